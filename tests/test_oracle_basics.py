"""Pins the oracle's building blocks against closed forms, invariants, brute
force and values measured on the reference -- without re-stating the
oracle's own formulas."""
import itertools

import numpy as np
import pytest


def test_stencil_tables(oracle):
    c = np.stack([oracle.CX, oracle.CY, oracle.CZ], axis=1)
    # 18 distinct unit/diagonal links with |c|^2 in {1, 2}, rest direction last
    assert (c[18] == 0).all()
    assert len({tuple(v) for v in c[:18]}) == 18
    assert sorted((c[:18] ** 2).sum(1).tolist()) == [1] * 6 + [2] * 12
    # inverse maps c -> -c and is an involution
    for d in range(19):
        assert (c[oracle.INV[d]] == -c[d]).all()
        assert oracle.INV[oracle.INV[d]] == d
    # names of src/Kernel.h:121-140 spell the components: N/S = +-y, E/W = +-x, T/B = +-z
    names = "N S E W NE SE NW SW T TN TE TW TS B BS BN BW BE C".split()
    comp = {"N": (0, 1, 0), "S": (0, -1, 0), "E": (1, 0, 0), "W": (-1, 0, 0), "T": (0, 0, 1), "B": (0, 0, -1)}
    for d, n in enumerate(names[:18]):
        v = np.zeros(3, int)
        for ch in n:
            v += comp[ch]
        assert tuple(v) == tuple(c[d]), n


def test_weights_isotropy(oracle):
    w, c = oracle.W, np.stack([oracle.CX, oracle.CY, oracle.CZ], 1).astype(float)
    assert abs(w.sum() - 1.0) < 1e-15
    # second moment = c_s^2 delta_ij with c_s^2 = 1/3; odd moments vanish
    m2 = np.einsum("d,di,dj->ij", w, c, c)
    assert np.allclose(m2, np.eye(3) / 3, atol=1e-15)
    assert np.allclose(np.einsum("d,di->i", w, c), 0, atol=1e-16)
    # fourth moment isotropy: sum w cx^2 cy^2 = 1/9, sum w cx^4 = 1/3
    assert abs((w * c[:, 0] ** 2 * c[:, 1] ** 2).sum() - 1 / 9) < 1e-15
    assert abs((w * c[:, 0] ** 4).sum() - 1 / 3) < 1e-15


@pytest.mark.parametrize("n,expected", [(3584, 264208), (246016, 264208), (253952, 264208)])
def test_padding_measured_values(oracle, n, expected):
    # values measured on the reference build (its "# padding" report)
    assert oracle.pad_cells(n) == expected


def test_padding_closed_form(oracle):
    rng = np.random.default_rng(1)
    for n in list(rng.integers(1, 50_000_000, 300)) + [1, 2097151, 2097152, 262208, 264208]:
        p = oracle.pad_cells(int(n))
        assert p >= n
        # auto padding makes the plane size 128 mod 16 KiB and the node count
        # 262208 mod 2 Mi before the second entry adds its 2000 (= 16 - 64 mod 2048)
        assert (p * 8) % 16384 == 128
        assert p - n < 2097152 + 2048
    assert oracle.pad_cells(100, "no") == 100
    assert oracle.pad_cells(100, "off") == 100
    assert oracle.pad_cells(100, "64+8") == 105        # 100 -> 8 cells per modulus: (1 + 8 - 4) % 8 = 5
    assert oracle.pad_cells(100, "64+8,128+0") == 112


@pytest.mark.parametrize("s", ["", "x", "+8", "64+8+8", "64,", "8+8", "64 + 8", "64+8,"])
def test_padding_rejects(oracle, s):
    rc, _ = oracle.pad_parse(s)
    assert rc != 0


def _textbook_trt(f, omega, magic):
    """TRT from its definition (Ginzburg 2008): split f and the incompressible
    equilibrium into symmetric/antisymmetric parts, relax them with
    omega_e = omega and omega_o from Lambda = (1/omega_e - 1/2)(1/omega_o - 1/2)."""
    from lbm_checkers import NDIR
    import lbm_checkers as lc  # noqa: F401
    cx = np.array([0, 0, 1, -1, 1, 1, -1, -1, 0, 0, 1, -1, 0, 0, 0, 0, -1, 1, 0])
    cy = np.array([1, -1, 0, 0, 1, -1, 1, -1, 0, 1, 0, 0, -1, 0, -1, 1, 0, 0, 0])
    cz = np.array([0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, 0])
    inv = [1, 0, 3, 2, 7, 6, 5, 4, 13, 14, 16, 17, 15, 8, 9, 12, 10, 11, 18]
    w = np.array([1 / 18] * 4 + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 3])
    rho = f.sum()
    u = np.array([(f * cx).sum(), (f * cy).sum(), (f * cz).sum()])
    cu = cx * u[0] + cy * u[1] + cz * u[2]
    feq = w * (rho + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)
    oo = 1 / (0.5 + magic / (1 / omega - 0.5))
    fs, fa = 0.5 * (f + f[inv]), 0.5 * (f - f[inv])
    es, ea = 0.5 * (feq + feq[inv]), 0.5 * (feq - feq[inv])
    assert len(f) == NDIR
    return f - omega * (fs - es) - oo * (fa - ea)


@pytest.mark.parametrize("omega,magic", [(1.0, 1 / 12), (1.3, 1 / 12), (0.7, 3 / 16), (1.8, 0.25)])
def test_collision_matches_textbook_trt(oracle, omega, magic):
    rng = np.random.default_rng(7)
    w = oracle.W
    for tree in (0, 1):
        for _ in range(50):
            f = w * (1 + 0.05 * rng.standard_normal(19))
            post = oracle.collide(f, omega, magic, tree)
            want = _textbook_trt(f, omega, magic)
            assert np.allclose(post, want, rtol=0, atol=1e-15)


def test_collision_invariants(oracle):
    rng = np.random.default_rng(3)
    c = np.stack([oracle.CX, oracle.CY, oracle.CZ], 1)
    for _ in range(100):
        f = oracle.W * (1 + 0.1 * rng.standard_normal(19))
        q = oracle.collide(f, 1.4)
        assert abs(q.sum() - f.sum()) < 1e-15                    # mass
        assert np.allclose(q @ c, f @ c, atol=1e-16)              # momentum
    # rest equilibrium is a fixed point (to rounding)
    assert np.allclose(oracle.collide(oracle.W.copy(), 1.2), oracle.W, rtol=0, atol=1e-16)
    # the two association trees differ only by rounding
    f = oracle.W * (1 + 0.1 * rng.standard_normal(19))
    assert np.allclose(oracle.collide(f, 1.1, tree=0), oracle.collide(f, 1.1, tree=1), atol=1e-16)


def test_force_terms(oracle):
    F = 1e-5
    ft = np.array([oracle.force_term(d, F) for d in range(19)])
    c = np.stack([oracle.CX, oracle.CY, oracle.CZ], 1)
    assert abs(ft.sum()) < 1e-21                                   # mass neutral
    assert np.allclose(ft @ c, [F, 0, 0], atol=1e-21)              # momentum F in x
    assert (ft[oracle.CX == 0] == 0).all()


def _brute_adjacency(mask, dims, per, grid, n_cells):
    """Direct transcription of the rules in words: push neighbour of node n in
    direction d; solid or beyond a non-periodic face -> own slot inv(d); a
    periodic exit wraps every out-of-range coordinate."""
    X, Y, Z = dims
    m = mask.reshape(Z, Y, X)
    g = grid.reshape(Z, Y, X)
    cx = [0, 0, 1, -1, 1, 1, -1, -1, 0, 0, 1, -1, 0, 0, 0, 0, -1, 1]
    cy = [1, -1, 0, 0, 1, -1, 1, -1, 0, 1, 0, 0, -1, 0, -1, 1, 0, 0]
    cz = [0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1]
    inv = [1, 0, 3, 2, 7, 6, 5, 4, 13, 14, 16, 17, 15, 8, 9, 12, 10, 11]
    out = {}
    for z, y, x in itertools.product(range(Z), range(Y), range(X)):
        if not m[z, y, x]:
            continue
        n = int(g[z, y, x])
        for d in range(18):
            q = [x + cx[d], y + cy[d], z + cz[d]]
            outside = [not (0 <= q[k] < dims[k]) for k in range(3)]
            if any(o and p for o, p in zip(outside, per)):
                q = [(q[k] % dims[k]) if outside[k] else q[k] for k in range(3)]
            elif any(outside):
                out[(n, d)] = inv[d] * n_cells + n
                continue
            qx, qy, qz = q
            out[(n, d)] = inv[d] * n_cells + n if not m[qz, qy, qx] else d * n_cells + int(g[qz, qy, qx])
    return out


@pytest.mark.parametrize("gtype,dims,per,blk", [
    (0, (5, 4, 3), (1, 1, 0), (0, 0, 0)),
    (1, (6, 5, 4), (1, 0, 0), (2, 2, 2)),
    (3, (8, 8, 8), (0, 0, 0), (3, 0, 2)),
    (4, (3, 3, 3), (0, 0, 0), (0, 0, 0)),
    (2, (7, 9, 9), (1, 1, 1), (0, 4, 0)),
])
def test_list_build_brute_force(oracle, gtype, dims, per, blk):
    mask, _ = oracle.geometry(gtype, dims, per, block=2)
    b = oracle.list_build(mask, dims, per, blk, n_cells=int(np.count_nonzero(mask)) + 7)
    X, Y, Z = dims
    # numbering: bijection onto fluid cells, blocked traversal with z fastest
    bx, by, bz = (blk[0] or X), (blk[1] or Y), (blk[2] or Z)
    order = []
    for x0, y0, z0 in itertools.product(range(0, X, bx), range(0, Y, by), range(0, Z, bz)):
        for x, y, z in itertools.product(range(x0, min(x0 + bx, X)), range(y0, min(y0 + by, Y)), range(z0, min(z0 + bz, Z))):
            if mask[(z * Y + y) * X + x]:
                order.append((x, y, z))
    assert [tuple(v) for v in b.coords.tolist()] == order
    bf = _brute_adjacency(mask, dims, per, b.grid, b.n_cells)
    for (n, d), v in bf.items():
        assert b.adj[n, d] == v, (n, d)


def test_list_adjacency_is_a_bijection(oracle):
    # AA in-place invariant: (n, d) -> adj[n][d] hits each of the 18 nFluid
    # non-rest slots exactly once; bounce-back entries point at the own slot
    for gtype, dims, per in [(0, (9, 8, 7), (1, 1, 0)), (1, (10, 6, 6), (1, 0, 0)), (3, (12, 12, 12), (0, 0, 0)),
                             (4, (4, 5, 6), (1, 1, 1))]:
        mask, _ = oracle.geometry(gtype, dims, per, block=2)
        b = oracle.list_build(mask, dims, per)
        nf, nc = b.n_fluid, b.n_cells
        flat = b.adj.astype(np.int64).reshape(-1)
        assert len(np.unique(flat)) == flat.size
        d_of, n_of = flat // nc, flat % nc
        assert (n_of < nf).all() and (d_of < 18).all()
        inv = oracle.INV
        for n in range(nf):
            for d in range(18):
                v = int(b.adj[n, d])
                dd, m = divmod(v, nc)
                if dd == inv[d]:
                    assert m == n
                else:
                    assert dd == d
                    # symmetry: the neighbour's inverse link points back at us
                    assert b.adj[m, inv[d]] == inv[d] * nc + n


def test_runs_count(oracle):
    mask, _ = oracle.geometry(0, (16, 16, 16), (1, 1, 0))
    b = oracle.list_build(mask, (16, 16, 16), (1, 1, 0))
    # measured on the reference: 3584 nodes, 768 runs with 4 threads
    assert b.n_fluid == 3584
    assert oracle.list_runs(b.adj, b.n_fluid, 4) == 768
    r1 = oracle.list_runs(b.adj, b.n_fluid, 1)
    assert r1 <= 768


def test_duct_flow_converges_to_the_series(lbm, oracle):
    """The oracle's aa-soa (TRT at Lambda 3/16, half-way bounce-back, Luo
    force) in square ducts of side W = 16 and 30 (x periodic, walls on y and
    z) at steady state: the mean velocity (u = sum c_x f - F/2) approaches
    the textbook square-duct series F W^2 / (12 nu) (1 - 192 / pi^5
    sum_{n odd} tanh(n pi / 2) / n^5) with a second-order wall error (the
    relative error falls by (30 / 16)^2 from one size to the next).  A
    dropped force term, a wrong wall position or a wrong viscosity moves it
    by far more."""
    F, om = 1e-5, 1.0
    cx = np.array([0, 0, 1, -1, 1, 1, -1, -1, 0, 0, 1, -1, 0, 0, 0, 0, -1, 1, 0], dtype=float)
    s = sum(np.tanh(n * np.pi / 2) / n ** 5 for n in range(1, 400, 2))
    err = {}
    for W, steps in ((16, 6000), (30, 12000)):
        dims = (4, W + 2, W + 2)
        geo = lbm.Geometry.from_string("channel", dims, (1, 0, 0))
        d, _ = oracle.aa_run(geo.mask.reshape(-1), dims, geo.periodic, steps, om, F, magic=3 / 16)
        u = float((d @ cx - F / 2).mean())
        want = F * W * W / (12 * (1 / om - 0.5) / 3) * (1 - 192 / np.pi ** 5 * s)
        err[W] = u / want - 1
    assert 0 < err[16] <= 5e-3 and 0 < err[30] <= 1.5e-3
    assert abs(err[16] / err[30] / (30 / 16) ** 2 - 1) <= 0.1
