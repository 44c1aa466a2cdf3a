"""GPU tests of the plane passes (x and y transforms of a z-plane in one
kernel per direction, c64).

Bars: the c64 operator application with the plane passes agrees with the
separate X and Y passes to 1e-6 relative and with the FP64 application to
2e-6 (c64 rounding) for every plane class (L = 32, 48, 64, 96, 128), a box
whose rows are one shorter than L/2, a box that is not eligible, and a mixed
batch; refined solves with and without the plane passes both meet the 1e-5
field bound against a tight FP64 solve, with iteration counts within 15 %;
the scattered-geometry (box map) path through a refined compat solve meets
1e-5 against the tightly converged oracle.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def native():
    from dda_model import native
    yield native
    native.set_plane_passes(True)


def _problem(native, shape, seed, precision, lam=560.0):
    rng = np.random.default_rng(seed)
    req = native.Request(shape=shape, d=15.0, lam=lam, rho=rng.uniform(0, 1, int(np.prod(shape))))
    return native.Problem(req, precision=precision)


SHAPES = [(16, 16, 16), (24, 24, 24), (32, 32, 16), (48, 48, 24), (64, 64, 32), (31, 31, 5), (20, 24, 6)]


@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
def test_plane_apply_matches_separate_passes(native, shape):
    _log(f"[gpu] plane apply {shape}")
    p32 = _problem(native, shape, 1, "fp32_refined")
    p64 = _problem(native, shape, 1, "fp64")
    rng = np.random.default_rng(2)
    n = 3 * int(np.prod(shape))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y64 = native.apply([p64], [x])[0]
    native.set_plane_passes(False)
    y_sep = native.apply([p32], [x], c64=True)[0]
    native.set_plane_passes(True)
    y_pl = native.apply([p32], [x], c64=True)[0]
    e_pl, e_sep, d = _rel(y_pl, y64), _rel(y_sep, y64), _rel(y_pl, y_sep)
    _log(f"[gpu]   vs FP64: plane {e_pl:.2e} separate {e_sep:.2e}; plane vs separate {d:.2e}")
    assert e_pl < 2e-6 and e_sep < 2e-6 and d < 1e-6


def test_plane_apply_mixed_batch(native):
    _log("[gpu] plane apply, mixed batch")
    shapes = [(16, 16, 16), (48, 48, 24), (20, 24, 6), (64, 64, 32), (24, 24, 24), (32, 32, 16), (16, 16, 16)]
    p32 = [_problem(native, s, 10 + i, "fp32_refined", lam=500.0 + 7 * i) for i, s in enumerate(shapes)]
    p64 = [_problem(native, s, 10 + i, "fp64", lam=500.0 + 7 * i) for i, s in enumerate(shapes)]
    rng = np.random.default_rng(3)
    xs = [rng.standard_normal(3 * p.n) + 1j * rng.standard_normal(3 * p.n) for p in p32]
    y64 = native.apply(p64, xs)
    native.set_plane_passes(True)
    y_pl = native.apply(p32, xs, c64=True)
    alone = [native.apply([p], [x], c64=True)[0] for p, x in zip(p32, xs)]
    for i in range(len(shapes)):
        assert _rel(y_pl[i], y64[i]) < 2e-6, i
        assert np.array_equal(y_pl[i], alone[i]), i   # same arithmetic alone and in a batch


def test_plane_refined_solves(native):
    shapes = [(16, 16, 16), (24, 24, 24), (32, 32, 16), (48, 48, 24), (64, 64, 32)]
    _log("[gpu] refined solves with and without the plane passes")
    refs = []
    for i, s in enumerate(shapes):
        p = _problem(native, s, 20 + i, "fp64", lam=450.0 + 60 * i)
        info = p.solve(tol=1e-10)
        assert info["status"] == "converged"
        refs.append(p.field())
    res = {}
    for on in (False, True):
        native.set_plane_passes(on)
        ps = [_problem(native, s, 20 + i, "fp32_refined", lam=450.0 + 60 * i) for i, s in enumerate(shapes)]
        infos = native.solve(ps, tol=1e-5)
        errs = [_rel(p.field(), r) for p, r in zip(ps, refs)]
        res[on] = ([i["iterations"] for i in infos], errs)
        _log(f"[gpu]   plane={on}: iterations {res[on][0]}, field errors {['%.2e' % e for e in errs]}")
        for info, e in zip(infos, errs):
            assert info["status"] == "converged" and info["error"] <= 1e-5
            assert e <= 1e-5
    native.set_plane_passes(True)
    for a, b in zip(res[True][0], res[False][0]):
        assert abs(a - b) <= max(3, 0.15 * b)


def test_plane_scattered_geometry_refined(native, tables):
    """Box map (dipoles on a subset of whole z-columns, shuffled) through a
    refined compat solve: Lx = Ly = 32 is a plane class."""
    import dda_model
    import dda_oracle as O
    _log("[gpu] plane passes, scattered geometry, refined compat solve")
    shape = (16, 16, 6)
    full = O.generate_geometry(*shape).reshape(-1, 3)
    rng = np.random.default_rng(12)
    cols = rng.permutation(256)[:150]
    keep = np.isin(full[:, 0] + 16 * full[:, 1], cols)
    pick = rng.permutation(np.where(keep)[0])
    geo = full[pick].reshape(-1)
    n = pick.size
    diel = np.repeat(rng.uniform(0, 1, n), 3)
    a = O.wrapper_args(list(shape), 15.0, diel, [1.0 + 0j, 5.96 + 3.89j], [0, 0, 1], [1, 0, 0], 542.0,
                       symmetry_axes=[7.5, 7.5], symmetry_method="None", apc_max_m=0, apc_max_n=0,
                       apc_method_name="LDR", tables=tables)
    a[12] = geo
    a[17] = n
    o = O.OracleDDAModel(*a)
    o.solveElectricField(np.zeros(3 * n, complex), 100000, 1e-11)
    e_tight = o.getElectricField()
    g = dda_model.DDAModel(*a)
    g.setSolverPrecision("fp32_refined")
    g.solveElectricField(np.zeros(3 * n, complex), 100000, 1e-5)
    err = _rel(g.getElectricField(), e_tight)
    _log(f"[gpu]   {n} dipoles: {g.solveInfo(0)}, E err {err:.2e}")
    assert g.solveInfo(0)["status"] == "converged"
    assert err <= 1e-5
