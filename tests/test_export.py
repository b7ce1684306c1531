"""The on-disk formats against hand-built text on a 2x2x2 model whose first
voxel is void (so compact node ids differ from finest-grid ids by one):
common/ExportDataset_3DTSV.m:11-41 and common/WrapVoxelFEAmodel.m:7-25."""
import numpy as np


def model(oracle):
    vox = np.ones((2, 2, 2), np.int32)
    vox[0, 0, 0] = 0
    m = oracle.discretize(vox, np.array([[0, 0, 0], [2, 2, 2.0]]), num_levels=2)
    lev = m.levels[0]
    # the inputs the expected text below assumes
    assert np.array_equal(lev.eleMapBack, np.arange(2, 9))
    assert np.array_equal(lev.nodMapBack, np.arange(2, 28))
    return lev


FIXED = np.array([1, 2, 3])
LOADS = np.array([[26, 0.0, 0.0, -1.0], [5, 0.5, 0.0, 0.0]])


def stress():
    S = np.zeros((26, 6))
    S[0] = [1.5, -2.25e-3, 123456.789, 0.0, 1e-7, -4.0]
    S[25] = [-0.5, 0.0, 0.0, 0.0, 0.0, 7.0e12]
    return S


def test_carti_dataset_text(pkg, oracle):
    lev = model(oracle)
    text = pkg.export.dataset_3dtsv_text((2, 2, 2), np.array([[0, 0, 0], [2, 2, 2.0]]), lev.eleMapBack,
                                         lev.nodMapBack, FIXED, LOADS, stress())
    zero = "0.000000e+00 0.000000e+00 0.000000e+00 0.000000e+00 0.000000e+00 0.000000e+00\n"
    expected = ("DATASET CARTESIAN_GRID \n"
                "Resolution: 2 2 2\n"
                "LowerBound: 0.000000 0.000000 0.000000\n"
                "UpperBound: 2.000000 2.000000 2.000000\n"
                "ELEMENTS 7 int \n"
                "1\n2\n3\n4\n5\n6\n7\n"
                "Number of Stress Fields: 1\n"
                "Node Forces: 2\n"
                "26 0.000000 0.000000 -1.000000\n"
                "5 0.500000 0.000000 0.000000\n"
                "Fixed Nodes: 3\n"
                "1\n2\n3\n"
                "Cartesian Stress: 26\n"
                "1.500000e+00 -2.250000e-03 1.234568e+05 0.000000e+00 1.000000e-07 -4.000000e+00\n"
                + zero * 24 +
                "-5.000000e-01 0.000000e+00 0.000000e+00 0.000000e+00 0.000000e+00 7.000000e+12\n")
    assert text == expected


def test_wrapped_model_text(pkg, oracle, tmp_path):
    lev = model(oracle)
    p = pkg.export.wrap_voxel_fea_model(tmp_path / "wrapped.txt", (2, 2, 2), lev.eleMapBack, lev.nodMapBack,
                                        FIXED, LOADS)
    expected = ("domain type: 3D \n"
                "resolution: 2 2 2\n"
                "valid elements: 7\n"
                "2\n3\n4\n5\n6\n7\n8\n"
                "fixed position: 3\n"
                "2 1 1 1\n3 1 1 1\n4 1 1 1\n"
                "loading condition: 2\n"
                "27 0.000000 0.000000 -1.000000\n"
                "6 0.500000 0.000000 0.000000\n")
    assert p.read_text() == expected


def test_empty_conditions(pkg, oracle):
    lev = model(oracle)
    text = pkg.export.wrapped_model_text((2, 2, 2), lev.eleMapBack, lev.nodMapBack, np.zeros(0),
                                         np.zeros((0, 4)))
    assert text.endswith("fixed position: 0\nloading condition: 0\n")
