"""Write the reference's literal span-2 transfer tables as a JSON fixture.

Reads three numeric tables, as text, from the reference MATLAB sources and
stores them as data in ``transfer_tables.json``:

* ``inNODE``  27x8  trilinear weights (Operator4MultiGridRestrictionAndInterpolation.m:5-31)
* ``inDOF``   81x24 per-DOF weights  (Operator4MultiGridRestrictionAndInterpolation.m:33-113)
* ``subEleNodMat`` 8x8 child -> 27-node map (SubEleNodMat.m:4-11)

Run in a container that has the reference mounted:
    python tests/golden/extract_reference_tables.py /root/reference
"""
import json
import re
import sys
from pathlib import Path


def _matrix_after(text: str, start_pat: str, rows: int):
    i = text.index(start_pat)
    j = text.index("[", i)
    k = text.index("]", j)
    body = text[j + 1:k]
    out = []
    for line in body.splitlines():
        nums = re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", line)
        if nums:
            out.append([float(v) for v in nums])
    assert len(out) == rows, (start_pat, len(out))
    return out


def main(ref_root: str):
    root = Path(ref_root) / "src" / "kernel"
    op = (root / "Operator4MultiGridRestrictionAndInterpolation.m").read_text()
    sub = (root / "SubEleNodMat.m").read_text()
    span2 = op[op.index("case 2"):op.index("case 4")]
    in_node = _matrix_after(span2, "'inNODE'", 27)
    in_dof = _matrix_after(span2, "'inDOF'", 81)
    sub2 = sub[sub.index("case 2"):sub.index("case 4")]
    sub_ele = [[int(v) for v in row] for row in _matrix_after(sub2, "val", 8)]
    out = {
        "source": "Operator4MultiGridRestrictionAndInterpolation.m:5-113, SubEleNodMat.m:4-11",
        "inNODE": in_node,
        "inDOF": in_dof,
        "subEleNodMat": sub_ele,
    }
    dst = Path(__file__).with_name("transfer_tables.json")
    dst.write_text(json.dumps(out))
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
