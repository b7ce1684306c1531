"""Pins for the CPU oracle (tests/../oracle/gmgs_oracle.py).

Each check is independent of the oracle's own formulas: closed-form
integrals, brute-force index construction from coordinates, the reference's
literal tables (tests/golden/transfer_tables.json), scipy sparse assembly and
direct solves, and invariants (R = P^T, Galerkin nesting, patch test).
"""
import itertools
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

GOLD = Path(__file__).resolve().parent / "golden"

# physical offsets (x, y, z) of the 8 hex8 corners: node 1 = (x-, y-, z-), ...
PHYS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


def exact_brick_stiffness(E, nu, h):
    """Ke of a brick hx*hy*hz from exact products of 1-D integrals."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))

    def i_ff(a, b, hh):      # int phi_a phi_b
        return hh / 3 if a == b else hh / 6

    def i_dd(a, b, hh):      # int phi_a' phi_b'
        return (1 if a == b else -1) / hh

    def i_df(a, b, hh):      # int phi_a' phi_b
        return (1 if a == 1 else -1) / 2.0

    def integral(i, j, di, dj):
        """int d_di N_i * d_dj N_j over the brick (di/dj in 0..2 or None)."""
        out = 1.0
        for ax in range(3):
            a, b = PHYS[i, ax], PHYS[j, ax]
            if di == ax and dj == ax:
                out *= i_dd(a, b, h[ax])
            elif di == ax:
                out *= i_df(a, b, h[ax])
            elif dj == ax:
                out *= i_df(b, a, h[ax])
            else:
                out *= i_ff(a, b, h[ax])
        return out

    K = np.zeros((24, 24))
    for i in range(8):
        for j in range(8):
            for c in range(3):
                for d in range(3):
                    v = lam * integral(i, j, c, d) + mu * integral(i, j, d, c)
                    if c == d:
                        v += mu * sum(integral(i, j, k, k) for k in range(3))
                    K[3 * i + c, 3 * j + d] = v
    return K


def probe_coords(h):
    return PHYS * np.asarray(h)[None, :]


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 2.0, 1.25)])
def test_ke0_closed_form(oracle, h):
    Ke = oracle.element_stiffness(1.0, 0.3, probe_coords(h))
    Kx = exact_brick_stiffness(1.0, 0.3, h)
    assert np.abs(Ke - Kx).max() < 1e-13 * np.abs(Kx).max()


def test_ke0_invariants(oracle):
    Ke = oracle.element_stiffness(1.0, 0.3, probe_coords((1, 1, 1)))
    lam, mu = 0.3 / (1.3 * 0.4), 1 / 2.6
    assert abs(Ke[0, 0] - (lam + 4 * mu) / 9) < 1e-14
    assert abs(Ke[0, 1] - (lam + mu) / 12) < 1e-14
    assert np.abs(Ke - Ke.T).max() < 1e-15
    X = PHYS.astype(float)
    modes = []
    for c in range(3):
        u = np.zeros((8, 3)); u[:, c] = 1; modes.append(u.ravel())
    for a, b in [(0, 1), (1, 2), (2, 0)]:
        u = np.zeros((8, 3)); u[:, a] = -X[:, b]; u[:, b] = X[:, a]; modes.append(u.ravel())
    R = np.stack(modes, axis=1)
    assert np.abs(Ke @ R).max() < 1e-14
    assert np.linalg.matrix_rank(Ke, tol=1e-10) == 18
    # Ke(h) = h Ke(1) for a cube
    K2 = oracle.element_stiffness(1.0, 0.3, probe_coords((2, 2, 2)))
    assert np.abs(K2 - 2 * Ke).max() < 1e-14


def test_transfer_tables_match_reference_literals(oracle):
    g = json.loads((GOLD / "transfer_tables.json").read_text())
    inNODE, inDOF, sub = oracle.transfer_tables()
    assert np.array_equal(inNODE, np.array(g["inNODE"]))
    assert np.array_equal(inDOF, np.array(g["inDOF"]))
    assert np.array_equal(sub, np.array(g["subEleNodMat"]))


def test_hooke_and_otp(oracle):
    D = oracle.hooke(2.0, 0.25)
    lam, mu = 2 * 0.25 / (1.25 * 0.5), 2 / 2.5
    ref = np.zeros((6, 6)); ref[:3, :3] = lam; ref[np.arange(3), np.arange(3)] = lam + 2 * mu
    ref[3, 3] = ref[4, 4] = ref[5, 5] = mu
    assert np.allclose(D, ref, rtol=1e-15, atol=0)
    OTP = oracle.outer_interpolation_matrix()
    s, t, p, _ = oracle.gauss_points()
    N = oracle.shape_function(s, t, p)
    # nodal values -> Gauss values -> back to nodes is the identity
    v = np.random.default_rng(0).standard_normal((8, 6))
    gp = N @ v
    back = (OTP @ gp.reshape(-1)).reshape(8, 6)
    assert np.abs(back - v).max() < 1e-12


# ------------------------------------------------------------------ indexing
def brute_force_hierarchy(vox, L):
    """Definition-level construction from coordinates (pure Python)."""
    vox = np.asarray(vox)
    ny, nx, nz = vox.shape
    out = []
    solid = {(y, x, z) for y in range(ny) for x in range(nx) for z in range(nz)
             if vox[y, x, z] == 1}
    fy = lambda y, x, z, ny_, nx_: y + ny_ * x + ny_ * nx_ * z + 1
    locl = [(1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 0), (1, 0, 1), (1, 1, 1), (0, 1, 1), (0, 0, 1)]
    prev = None
    dims = (ny, nx, nz)
    for lev in range(1, L + 1):
        ey, ex, ez = dims
        if lev > 1:
            pe = prev
            cells = {}
            for gz in range(ez):
                for gx in range(ex):
                    for gy in range(ey):
                        kids = []
                        for cz in range(2):
                            for cx in range(2):
                                for cy in range(2):
                                    c = (2 * gy + cy, 2 * gx + cx, 2 * gz + cz)
                                    kids.append(pe["ecompact"].get(c, 0))
                        if any(kids):
                            cells[(gy, gx, gz)] = kids
            solid = set(cells)
        emb = sorted(fy(y, x, z, ey, ex) for (y, x, z) in solid)
        inv = {}
        for (y, x, z) in solid:
            inv[fy(y, x, z, ey, ex)] = (y, x, z)
        ecompact = {inv[e]: i + 1 for i, e in enumerate(emb)}
        nodes = set()
        for (y, x, z) in solid:
            for (a, b, c) in locl:
                nodes.add(fy(y + a, x + b, z + c, ey + 1, ex + 1))
        nmb = sorted(nodes)
        ncompact = {n: i + 1 for i, n in enumerate(nmb)}
        enod = np.array([[ncompact[fy(inv[e][0] + a, inv[e][1] + b, inv[e][2] + c, ey + 1, ex + 1)]
                          for (a, b, c) in locl] for e in emb], dtype=np.int64).reshape(-1, 8)
        rec = dict(eleMapBack=np.array(emb), numNodes=len(nmb), eNodMat=enod, ecompact=ecompact)
        # forward maps, incident-element counts and boundary lists from
        # coordinates (DiscretizeDesignDomain.m:68-72,86-89,103-116)
        efwd = np.zeros(ey * ex * ez, dtype=np.int64)
        for i, e in enumerate(emb):
            efwd[e - 1] = i + 1
        nfwd = np.zeros((ey + 1) * (ex + 1) * (ez + 1), dtype=np.int64)
        for i, n in enumerate(nmb):
            nfwd[n - 1] = i + 1
        cnt = np.zeros(len(nmb), dtype=np.int64)
        for i, n in enumerate(nmb):
            q = n - 1
            y, x, z = q % (ey + 1), (q // (ey + 1)) % (ex + 1), q // ((ey + 1) * (ex + 1))
            cnt[i] = sum((y - a, x - b, z - c) in solid
                         for a in (0, 1) for b in (0, 1) for c in (0, 1))
        onb = {i + 1 for i in range(len(nmb)) if cnt[i] < 8}
        eob = [i + 1 for i, e in enumerate(emb)
               if any(ncompact[fy(inv[e][0] + a, inv[e][1] + b, inv[e][2] + c, ey + 1, ex + 1)] in onb
                      for (a, b, c) in locl)]
        rec.update(eleMapForward=efwd, nodMapForward=nfwd, numNod2Eles=cnt,
                   nodesOnBoundary=np.array(sorted(onb), dtype=np.int64),
                   elementsOnBoundary=np.array(eob, dtype=np.int64))
        s = 2 ** (lev - 1)
        Y1, X1 = dims[0] * s + 1, dims[1] * s + 1

        def to_finest(n):
            n -= 1
            y = n % (ey + 1); x = (n // (ey + 1)) % (ex + 1); z = n // ((ey + 1) * (ex + 1))
            return y * s + Y1 * (x * s) + Y1 * X1 * (z * s) + 1
        rec["nodMapBack"] = np.array(nmb) if lev == 1 else np.array([to_finest(n) for n in nmb])
        if lev > 1:
            up = np.array([cells[inv[e]] for e in emb], dtype=np.int64)
            rec["elementUpwardMap"] = up
            fyy, fxx = 2 * ey + 1, 2 * ex + 1
            inter = set()
            for e in emb:
                gy, gx, gz = inv[e]
                for c in range(3):
                    for b in range(3):
                        for a in range(3):
                            inter.add(fy(2 * gy + a, 2 * gx + b, 2 * gz + c, fyy, fxx))
            inter = sorted(inter)
            irank = {n: i + 1 for i, n in enumerate(inter)}
            tm = np.zeros((27, len(emb)), dtype=np.int64)
            for k, e in enumerate(emb):
                gy, gx, gz = inv[e]
                for c in range(3):
                    for b in range(3):
                        for a in range(3):
                            tm[a + 3 * b + 9 * c, k] = irank[fy(2 * gy + a, 2 * gx + b, 2 * gz + c, fyy, fxx)]
            rec["transferMat"] = tm
            rec["transferMatCoeffi"] = np.bincount(tm.ravel() - 1, minlength=len(inter)).astype(float)
            rec["solidNodeMapCoarser2Finer"] = np.array([irank[n] for n in prev["nodes_full"]])
        rec["nodes_full"] = nmb
        out.append(rec)
        prev = rec
        dims = (ey // 2, ex // 2, ez // 2)
    return out


@pytest.mark.parametrize("shape,L,seed", [((4, 4, 4), 2, 0), ((8, 6, 4), 3, 1),
                                           ((8, 8, 8), 4, 2), ((6, 10, 4), 2, 3)])
def test_hierarchy_brute_force(oracle, shape, L, seed):
    rng = np.random.default_rng(seed)
    vox = (rng.random(shape) < 0.6).astype(np.int32)
    vox[0, 0, 0] = 1
    m = oracle.discretize(vox, np.array([[0, 0, 0], [shape[1], shape[0], shape[2]]], float),
                          num_levels=L)
    oracle.build_hierarchy(m)
    bf = brute_force_hierarchy(oracle.pad_vox(vox, L) if hasattr(oracle, "pad_vox") else
                               m_pad(vox, L), L)
    for lev, ref in zip(m.levels, bf):
        assert np.array_equal(lev.eleMapBack, ref["eleMapBack"])
        assert lev.numNodes == ref["numNodes"]
        assert np.array_equal(lev.eNodMat, ref["eNodMat"])
        assert np.array_equal(lev.nodMapBack, ref["nodMapBack"])
        for key in ("eleMapForward", "nodMapForward", "numNod2Eles", "nodesOnBoundary",
                    "elementsOnBoundary"):
            assert np.array_equal(np.asarray(getattr(lev, key)).ravel(), ref[key]), key
        if "elementUpwardMap" in ref:
            assert np.array_equal(lev.elementUpwardMap, ref["elementUpwardMap"])
            assert np.array_equal(lev.transferMat, ref["transferMat"])
            assert np.array_equal(lev.transferMatCoeffi, ref["transferMatCoeffi"])
            assert np.array_equal(lev.solidNodeMapCoarser2Finer, ref["solidNodeMapCoarser2Finer"])
            assert set(np.unique(lev.transferMatCoeffi)) <= {1.0, 2.0, 4.0, 8.0}


def m_pad(vox, L):
    ny, nx, nz = vox.shape
    s = 2 ** (L - 1)
    out = np.zeros((-(-ny // s) * s, -(-nx // s) * s, -(-nz // s) * s), dtype=np.int32)
    out[:ny, :nx, :nz] = vox
    return out


def test_level_rule_and_box_counts(oracle, pkg):
    wl = pkg.workloads
    assert wl.level_rule(40, 40, 64, 1000.0) == 4
    assert wl.level_rule(169, 147, 256, 1e4) == 5
    assert wl.level_rule(337, 294, 512, 1e4) == 6
    assert wl.padded_dims(169, 147, 256, 5) == (176, 160, 256)
    m = oracle.discretize(np.ones((40, 40, 64), np.int32),
                          np.array([[0, 0, 0], [40, 40, 64.0]]), 1000.0)
    assert m.numLevels == 4
    oracle.build_hierarchy(m)
    for k, lev in enumerate(m.levels):
        s = 2 ** k
        ne = (40 // s) * (40 // s) * (64 // s)
        nn = (40 // s + 1) * (40 // s + 1) * (64 // s + 1)
        assert (lev.numElements, lev.numNodes) == (ne, nn)
    assert [l.numDOFs for l in m.levels] == [327795, 43659, 6171, 972]


def test_bc_last_row_wins(oracle):
    m = oracle.discretize(np.ones((2, 2, 2), np.int32), np.array([[0, 0, 0], [2, 2, 2.0]]),
                          num_levels=2)
    loads = np.array([[5, 1, 2, 3], [7, 0, 0, 1], [5, -1, -2, -3]], float)
    F = oracle.apply_bc(m, loads, np.array([1, 2]))
    assert np.array_equal(F[12:15], [-1, -2, -3]) and F[20] == 1
    assert np.array_equal(np.sort(m.levels[0].fixedDOFs), np.arange(1, 7))


# ------------------------------------------------------------------ operators
def small_model(oracle, pkg, exact=False, L=3, seed=1, shape=(16, 12, 16)):
    wl = pkg.workloads
    p = wl.ellipsoid_problem(shape[1], shape[0], shape[2], L, seed=seed, n_ell=4, n_loads=30,
                             tol=1e-10)
    m = oracle.build(p.vox, p.bbox, num_levels=L, omega=0.4, exact_galerkin=exact)
    oracle.apply_bc(m, p.loads, p.fixed)
    oracle.factor_coarsest(m)
    return m, p


def assemble_level(m, ii):
    lev = m.levels[ii - 1]
    ed = lev.eDofMat - 1
    if lev.storingState == 0:
        Ks = np.broadcast_to(lev.Ks, (lev.numElements, 24, 24))
    elif lev.storingState == 1:
        Ks = lev.Ks
    else:
        st = np.zeros(lev.numElements, dtype=np.int64)
        st[lev.uniqueElements - 1] = np.arange(1, 1 + lev.uniqueElements.size)
        Ks = lev.Ks[st]
    rows = np.repeat(ed, 24, axis=1).ravel()
    cols = np.tile(ed, (1, 24)).ravel()
    return sp.csr_matrix((Ks.reshape(-1), (rows, cols)), shape=(lev.numDOFs,) * 2)


@pytest.mark.parametrize("exact", [False, True])
def test_atx_all_levels_match_assembly(oracle, pkg, exact):
    m, _ = small_model(oracle, pkg, exact=exact, L=4, shape=(16, 16, 24))
    rng = np.random.default_rng(5)
    K1 = oracle.assemble_level1(m)
    for ii in range(1, m.numLevels + 1):
        lev = m.levels[ii - 1]
        x = rng.standard_normal(lev.numDOFs)
        K = K1 if ii == 1 else assemble_level(m, ii)
        y = K @ x
        y[lev.fixedDOFs - 1] = 0
        assert np.abs(oracle.atx(m, x, ii) - y).max() < 1e-12 * np.abs(y).max()
        d = K.diagonal()
        if ii < m.numLevels:
            assert np.abs(lev.diagK - d).max() < 1e-12 * d.max()


def structured_prolong(m, ii, xc):
    """Plain trilinear interpolation on the structured coarse grid."""
    lev, fine = m.levels[ii - 1], m.levels[ii - 2]
    ny, nx, nz = lev.resY, lev.resX, lev.resZ
    grid = np.zeros((ny + 1, nx + 1, nz + 1, 3))
    full = np.nonzero(lev.nodMapForward)[0]
    grid.reshape(-1, 3, order="F")
    g = grid.reshape((ny + 1) * (nx + 1) * (nz + 1), 3, order="F")
    g[lev.nodMapForward[full] .astype(int) * 0 + full] = xc.reshape(-1, 3)[lev.nodMapForward[full] - 1]
    grid = g.reshape(ny + 1, nx + 1, nz + 1, 3, order="F")
    fy, fx, fz = 2 * ny + 1, 2 * nx + 1, 2 * nz + 1
    # fine full ids of solid fine nodes
    fb = np.nonzero(fine.nodMapForward)[0]
    y = fb % fy; x = (fb // fy) % fx; z = fb // (fy * fx)
    out = np.zeros((fb.size, 3))
    for oy in (0, 1):
        for ox in (0, 1):
            for oz in (0, 1):
                wy = np.where(y % 2 == 0, 1.0 - oy, 0.5)
                wx = np.where(x % 2 == 0, 1.0 - ox, 0.5)
                wz = np.where(z % 2 == 0, 1.0 - oz, 0.5)
                cy = np.minimum(y // 2 + oy, ny); cx = np.minimum(x // 2 + ox, nx)
                cz = np.minimum(z // 2 + oz, nz)
                out += (wy * wx * wz)[:, None] * grid[cy, cx, cz]
    res = np.zeros((fine.numNodes, 3))
    res[fine.nodMapForward[fb] - 1] = out
    return res.reshape(-1)


def test_transfers(oracle, pkg):
    m, _ = small_model(oracle, pkg, L=3)
    rng = np.random.default_rng(2)
    for ii in range(2, m.numLevels + 1):
        lev, fine = m.levels[ii - 1], m.levels[ii - 2]
        rf = rng.standard_normal(fine.numDOFs)
        xc = rng.standard_normal(lev.numDOFs)
        Rr = oracle.restrict(m, rf, ii)
        Px = oracle.interpolate(m, xc, ii)
        assert abs(Rr @ xc - rf @ Px) < 1e-12 * abs(rf @ Px) + 1e-12   # R = P^T
        assert np.abs(Px - structured_prolong(m, ii, xc)).max() < 1e-14
    # linear fields are reproduced exactly on solid fine nodes
    L1 = m.levels[0]
    coords = m.nodeCoords
    lev = m.levels[1]
    cf = coords[lev.nodMapBack - 1]
    f = lambda c: np.stack([1 + 2 * c[:, 0] - c[:, 2], c[:, 1] * 0.5, c[:, 0] + c[:, 1] + c[:, 2]], 1)
    Px = oracle.interpolate(m, f(cf).reshape(-1), 2)
    assert np.abs(Px - f(coords[L1.nodMapBack - 1]).reshape(-1)).max() < 1e-12


def test_galerkin_nesting_and_quirk(oracle, pkg):
    Ke = oracle.element_stiffness(1.0, 0.3, probe_coords((1, 1, 1)))
    full = np.ones((1, 8))
    G = oracle.galerkin(np.broadcast_to(Ke, (1, 8, 24, 24)).copy(), full)
    assert np.abs(G[0] - 2 * Ke).max() < 1e-13
    # exact Galerkin: sum of coarse Ks equals P^T K_fine P on level 2
    m, _ = small_model(oracle, pkg, exact=True, L=2, shape=(8, 8, 8), seed=4)
    K1 = oracle.assemble_level1(m).toarray()
    n2 = m.levels[1].numDOFs
    P = np.stack([oracle.interpolate(m, np.eye(n2)[j], 2) for j in range(n2)], axis=1)
    K2 = assemble_level(m, 2).toarray()
    assert np.abs(P.T @ K1 @ P - K2).max() < 1e-12 * np.abs(K2).max()
    # reference placement (B12): mask on column 3 lands on geometric child 4
    pres = np.ones((1, 8)); pres[0, 2] = 0
    ref = oracle.galerkin(np.broadcast_to(Ke, (1, 8, 24, 24)).copy(), pres, exact=False)
    pres2 = np.ones((1, 8)); pres2[0, 3] = 0
    ex = oracle.galerkin(np.broadcast_to(Ke, (1, 8, 24, 24)).copy(), pres2, exact=True)
    assert np.abs(ref - ex).max() < 1e-14
    ex3 = oracle.galerkin(np.broadcast_to(Ke, (1, 8, 24, 24)).copy(), pres, exact=True)
    assert np.abs(ref - ex3).max() > 1e-3


def holey_brick(oracle, L=3, seed=7, shape=(10, 12, 16), omega=0.5):
    """Brick with random interior voids, bottom face fixed, random top loads."""
    rng = np.random.default_rng(seed)
    ny, nx, nz = shape
    vox = np.ones(shape, np.int32)
    vox[1:-1, 1:-1, 2:-2][rng.random((ny - 2, nx - 2, nz - 4)) < 0.15] = 0
    m = oracle.build(vox, np.array([[0, 0, 0], [nx, ny, nz]], float), num_levels=L, omega=omega)
    nyn, nxn = m.levels[0].resY + 1, m.levels[0].resX + 1
    full = m.levels[0].nodMapBack
    iz = (full - 1) // (nyn * nxn)
    fixed = np.nonzero(iz == 0)[0] + 1
    top = np.nonzero(iz == nz)[0] + 1
    loads = np.concatenate([top[:, None], rng.standard_normal((top.size, 3))], axis=1)
    oracle.apply_bc(m, loads, fixed)
    oracle.factor_coarsest(m)
    return m


def test_cg_matches_direct_solve(oracle):
    m = holey_brick(oracle)
    res = oracle.cg_solver(m, m.F, 1e-11, 500)
    assert res.converged
    K = oracle.assemble_level1(m)
    fr = m.levels[0].freeDOFs - 1
    u = np.zeros(m.levels[0].numDOFs)
    u[fr] = spla.spsolve(K[fr][:, fr].tocsc(), m.F[fr])
    assert np.linalg.norm(res.x - u) / np.linalg.norm(u) < 1e-9
    assert res.hist.shape == (res.its - 1, 3)
    # recursive residual agrees with the true one
    r = m.F - oracle.atx(m, res.x)
    assert np.linalg.norm(r) / np.linalg.norm(m.F) < 1e-10


def test_stress_patch(oracle, pkg):
    m, _ = small_model(oracle, pkg, L=2, shape=(8, 8, 8))
    L1 = m.levels[0]
    c = m.nodeCoords[L1.nodMapBack - 1]
    A = np.array([[1e-3, 2e-3, -1e-3], [0.5e-3, -2e-3, 3e-3], [1e-3, 0, 2e-3]])
    U = (c @ A.T).reshape(-1)
    S = oracle.cartesian_stress(m, U)
    eps = np.array([A[0, 0], A[1, 1], A[2, 2], A[1, 2] + A[2, 1], A[2, 0] + A[0, 2],
                    A[0, 1] + A[1, 0]])
    sig = oracle.hooke(1.0, 0.3) @ eps
    assert np.abs(S - sig[None, :]).max() < 1e-13


# ---- symmetric V(nu, nu) cycle (no reference counterpart) -------------------

def _free_random(m, rng):
    v = rng.standard_normal(m.levels[0].numDOFs)
    v[m.levels[0].fixedDOFs - 1] = 0.0
    return v


def _cheb_weights(oracle, m, nu, **kw):
    return {ii: oracle.chebyshev_weights(oracle.estimate_lmax(m, ii), nu, **kw) for ii in range(1, m.numLevels)}


@pytest.mark.parametrize("cheb", [False, True])
@pytest.mark.parametrize("nu", [1, 2, 3])
def test_symmetric_vcycle_is_symmetric_and_reference_is_not(oracle, nu, cheb):
    m = holey_brick(oracle)
    rng = np.random.default_rng(5)
    a, b = _free_random(m, rng), _free_random(m, rng)
    w = _cheb_weights(oracle, m, nu) if cheb else None
    Ma, Mb = oracle.vcycle_sym(m, a, nu, w), oracle.vcycle_sym(m, b, nu, w)
    assert abs(Ma @ b - a @ Mb) <= 1e-12 * abs(Ma @ b)
    assert a @ Ma > 0 and b @ Mb > 0
    # kernel/Vcycle.m:27-33 post-smooths with the stale residual: not symmetric
    Ra, Rb = oracle.vcycle(m, a), oracle.vcycle(m, b)
    assert abs(Ra @ b - a @ Rb) > 1e-8 * abs(Ra @ b)


@pytest.mark.parametrize("cheb", [False, True])
@pytest.mark.parametrize("nu", [1, 2, 3])
def test_symmetric_vcycle_two_level_error_propagation(oracle, nu, cheb):
    """I - M A = G (I - P A_c^-1 P^T A) G on the free DOFs, G = prod_s (I - w_s D^-1 A)
    (w_s = omega, or the Chebyshev weights), built from dense matrices of the
    level operators."""
    m = holey_brick(oracle, L=2, shape=(4, 4, 6), seed=3)
    L1, L2 = m.levels
    F = L1.freeDOFs - 1
    Fc = L2.freeDOFs - 1
    n, nc = L1.numDOFs, L2.numDOFs

    def unit(k, size):
        v = np.zeros(size)
        v[k] = 1.0
        return v

    A = np.stack([oracle.atx(m, unit(j, n), 1)[F] for j in F], axis=1)
    P = np.stack([oracle.interpolate(m, unit(k, nc), 2)[F] for k in Fc], axis=1)
    Ac_inv = np.stack([oracle.coarsest_solve(m, unit(k, nc))[Fc] for k in Fc], axis=1)
    Dinv = np.diag(1.0 / L1.diagK[F])
    I = np.eye(F.size)
    w = _cheb_weights(oracle, m, nu) if cheb else {1: [m.omega] * nu}
    Sm = I
    for ws in w[1]:
        Sm = Sm @ (I - ws * Dinv @ A)
    E = Sm @ (I - P @ Ac_inv @ P.T @ A) @ Sm
    M_formula = (I - E) @ np.linalg.inv(A)
    M = np.stack([oracle.vcycle_sym(m, unit(j, n), nu, w if cheb else None)[F] for j in F], axis=1)
    assert np.abs(M - M_formula).max() <= 1e-9 * np.abs(M_formula).max()


def test_symmetric_cycle_cg_converges_faster(oracle):
    m = holey_brick(oracle)
    ref = oracle.cg_solver(m, m.F, 1e-10, 500)
    sym = oracle.cg_solver(m, m.F, 1e-10, 500, jacobi_step=False,
                           cycle=lambda mm, r: oracle.vcycle_sym(mm, r, 2))
    assert sym.converged and sym.its < ref.its
    assert np.linalg.norm(sym.x - ref.x) / np.linalg.norm(ref.x) < 1e-8


# ---- Chebyshev sweep weights (no reference counterpart) ---------------------

@pytest.mark.parametrize("nu", [1, 2, 3])
def test_chebyshev_weights_give_the_shifted_chebyshev_polynomial(oracle, nu):
    """prod_i (1 - w_i l) = T_nu((theta - l)/delta) / T_nu(theta/delta), with T_nu
    from numpy's Chebyshev series, and |p| equioscillates at 1/T_nu(theta/delta)
    on [lower b, b]."""
    from numpy.polynomial import chebyshev as C
    lmax, lower, safety = 2.7, 0.2, 1.1
    w = oracle.chebyshev_weights(lmax, nu, lower, safety)
    b = safety * lmax
    a = lower * b
    th, de = 0.5 * (b + a), 0.5 * (b - a)
    Tn = [0.0] * nu + [1.0]
    lam = np.linspace(0.0, 1.3 * b, 101)
    p = np.prod([1.0 - wi * lam for wi in w], axis=0)
    ref = C.chebval((th - lam) / de, Tn) / C.chebval(th / de, Tn)
    assert np.abs(p - ref).max() < 1e-12
    inside = np.linspace(a, b, 2001)
    pin = np.abs(np.prod([1.0 - wi * inside for wi in w], axis=0))
    assert abs(pin.max() - 1.0 / C.chebval(th / de, Tn)) < 1e-9


def test_lanczos_estimate_is_a_lower_bound_close_to_lambda_max(oracle):
    """The top Ritz value never exceeds lambda_max(D^-1 A) on the free DOFs
    (dense eigenvalues of D^-1/2 A D^-1/2) and reaches it within 1% here."""
    m = holey_brick(oracle, L=2, shape=(4, 4, 6), seed=3)
    L1 = m.levels[0]
    F = L1.freeDOFs - 1
    n = L1.numDOFs
    A = np.stack([oracle.atx(m, np.eye(1, n, j).ravel(), 1)[F] for j in F], axis=1)
    s = 1.0 / np.sqrt(L1.diagK[F])
    lam = np.linalg.eigvalsh(s[:, None] * A * s[None, :])[-1]
    prev = 0.0
    for steps in (2, 5, 10, 20):
        est = oracle.estimate_lmax(m, 1, steps)
        assert est <= lam * (1 + 1e-12) and est >= prev - 1e-12
        prev = est
    assert prev >= 0.99 * lam


def test_chebyshev_cycle_cg_converges_faster(oracle):
    m = holey_brick(oracle)
    jac = oracle.cg_solver(m, m.F, 1e-10, 500, jacobi_step=False,
                           cycle=lambda mm, r: oracle.vcycle_sym(mm, r, 2))
    # the interval's lower end trades the smoothed band against the damping in
    # it: on this small brick [0.2 b, b] wins clearly, the default [0.05 b, b]
    # (chosen on the 57 M-DOF problem) at least ties
    for lower, better in ((0.2, True), (0.05, False)):
        w = _cheb_weights(oracle, m, 2, lower=lower)
        che = oracle.cg_solver(m, m.F, 1e-10, 500, jacobi_step=False,
                               cycle=lambda mm, r: oracle.vcycle_sym(mm, r, 2, w))
        assert che.converged and (che.its < jac.its if better else che.its <= jac.its)
        assert np.linalg.norm(che.x - jac.x) / np.linalg.norm(jac.x) < 1e-8


# ---- B3: the level-3 "unique element" test (SetupSolver.m:94-101) ----------

def level3_classes_by_coordinates(vox, L):
    """Level-3 elements of a padded mask (L > 3), classified from voxel
    coordinates: a level-3 cell covers 4x4x4 voxels and its 8 level-2 children
    2x2x2 each.  Returns (exact, reference) sets of compact ids: exact = some
    voxel of the cell is void (its Galerkin operator differs from 4 Ke0);
    reference = some *present* level-2 child holds a void voxel (children that
    are entirely void are skipped, SetupSolver.m:97-98)."""
    vox = np.asarray(vox)
    ny, nx, nz = (d // 4 for d in vox.shape)
    full_ids, exact, ref = [], set(), set()
    for gz in range(nz):
        for gx in range(nx):
            for gy in range(ny):
                cell = vox[4 * gy:4 * gy + 4, 4 * gx:4 * gx + 4, 4 * gz:4 * gz + 4]
                if cell.any():
                    full_ids.append((gy + ny * gx + ny * nx * gz + 1, cell))
    full_ids.sort(key=lambda t: t[0])
    for k, (_, cell) in enumerate(full_ids):
        kids = [cell[2 * a:2 * a + 2, 2 * b:2 * b + 2, 2 * c:2 * c + 2]
                for a in (0, 1) for b in (0, 1) for c in (0, 1)]
        if not cell.all():
            exact.add(k + 1)
        if any(kid.any() and not kid.all() for kid in kids):
            ref.add(k + 1)
    return exact, ref


def b3_mask():
    """16^3 solid with, inside level-3 cell (1,1,1), one level-2 child left
    entirely void (B3 fires: its present children are all full), and in cell
    (2,2,2) a single void voxel (unique either way)."""
    vox = np.ones((16, 16, 16), np.int32)
    vox[4:6, 4:6, 4:6] = 0
    vox[9, 10, 11] = 0
    return vox


@pytest.mark.parametrize("kind,seed", [("designed", 0), ("random", 1), ("random", 2)])
def test_level3_classification_b3_by_coordinates(oracle, kind, seed):
    if kind == "designed":
        vox = b3_mask()
    else:
        rng = np.random.default_rng(seed)
        vox = np.ones((16, 16, 24), np.int32)
        # void 2x2x2 blocks on the level-2 grid and scattered void voxels
        blocks = rng.random((8, 8, 12)) < 0.08
        vox[np.kron(blocks, np.ones((2, 2, 2), bool)).astype(bool)] = 0
        vox[rng.random(vox.shape) < 0.01] = 0
        vox[0, 0, 0] = 1
    bbox = np.array([[0, 0, 0], [vox.shape[1], vox.shape[0], vox.shape[2]]], float)
    exact_ids, ref_ids = level3_classes_by_coordinates(vox, 4)
    b3 = exact_ids - ref_ids
    assert ref_ids <= exact_ids
    if kind == "designed":
        assert b3 == {1 + 1 + 4 * 1 + 16 * 1}     # cell (1,1,1) of the 4^3 level-3 grid
    assert len(b3) >= 1
    for exact in (False, True):
        m = oracle.build(vox, bbox, num_levels=4, exact_galerkin=exact)
        lev3 = m.levels[2]
        assert lev3.storingState == 2
        got = set(int(v) for v in lev3.uniqueElements)
        assert got == (exact_ids if exact else ref_ids)
        # a B3 element is treated as universal: 4 Ke0 (SetupSolver.m:105) in
        # the reference, its Galerkin product (not 4 Ke0) when exact
        st = np.zeros(lev3.numElements + 1, dtype=np.int64)
        st[lev3.uniqueElements] = np.arange(1, lev3.uniqueElements.size + 1)
        Ke3 = 4.0 * m.Ke0
        assert np.abs(lev3.Ks[0] - Ke3).max() < 1e-13
        for e in b3:
            K = lev3.Ks[st[e]]
            if exact:
                assert st[e] > 0 and np.abs(K - Ke3).max() > 1e-3
            else:
                assert st[e] == 0


# ---- B2: error propagation of the reference V-cycle (Vcycle.m:1-36) ---------

@pytest.mark.parametrize("L,shape", [(2, (4, 4, 6)), (3, (8, 4, 8))])
def test_reference_vcycle_operator_identity(oracle, L, shape):
    """The reference cycle (x_l = S r; r_l = r - A S r; coarse correction;
    post-smoothing with the stale r_l, Vcycle.m:12-33 with JacobiSmoother.m)
    is the linear map M = Z B_1 with
        B_l = 2 S_l - S_l A_l S_l + P_l B_{l+1} R_l (I - A_l S_l),  B_L = C,
    S_l = omega D_l^-1, R_l = P_l^T, C = K_ff^-1 of the coarsest level (zero
    on fixed DOFs), A_l the assembled operator with fixed rows zeroed
    (AtX.m:57) and Z zeroing the level-1 fixed DOFs (Vcycle.m:35).  Every
    matrix here comes from an independent assembly (scipy sparse from the
    element matrices, a dense inverse of the coarsest K_ff, the diagonal of
    the assembly), not from the oracle's cycle code; a post-smoothing with
    the updated residual, a dropped pre-smoothed x or a wrong level's
    diagonal changes M by O(1)."""
    m = holey_brick(oracle, L=L, shape=shape, seed=3)
    n = [lev.numDOFs for lev in m.levels]

    def unit(k, size):
        v = np.zeros(size)
        v[k] = 1.0
        return v

    A, S = [], []
    for ii in range(1, L + 1):
        lev = m.levels[ii - 1]
        K = (oracle.assemble_level1(m) if ii == 1 else assemble_level(m, ii)).toarray()
        Kz = K.copy()
        Kz[lev.fixedDOFs - 1, :] = 0.0
        A.append(Kz)
        S.append(m.omega * np.diag(1.0 / K.diagonal()))
    P = [None] + [np.stack([oracle.interpolate(m, unit(k, n[ii - 1]), ii) for k in range(n[ii - 1])], axis=1)
                  for ii in range(2, L + 1)]
    Kc = assemble_level(m, L).toarray()
    fr = m.levels[-1].freeDOFs - 1
    C = np.zeros_like(Kc)
    C[np.ix_(fr, fr)] = np.linalg.inv(Kc[np.ix_(fr, fr)])
    B = C
    for ii in range(L - 1, 0, -1):
        Sl, Al, Pl = S[ii - 1], A[ii - 1], P[ii]
        I = np.eye(n[ii - 1])
        B = 2 * Sl - Sl @ Al @ Sl + Pl @ B @ Pl.T @ (I - Al @ Sl)
    Z = np.ones(n[0])
    Z[m.levels[0].fixedDOFs - 1] = 0.0
    M_formula = Z[:, None] * B
    M = np.stack([oracle.vcycle(m, unit(j, n[0])) for j in range(n[0])], axis=1)
    assert np.abs(M - M_formula).max() <= 1e-10 * np.abs(M_formula).max()
    # the transpose relation the formula uses: restriction is P^T
    rng = np.random.default_rng(0)
    for ii in range(2, L + 1):
        rf = rng.standard_normal(n[ii - 2])
        assert np.abs(oracle.restrict(m, rf, ii) - P[ii - 1].T @ rf).max() < 1e-13 * np.abs(rf).max() * 8
