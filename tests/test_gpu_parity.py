"""GPU parity of the compat DDAModel against the FP64 oracle (small sizes).

Tolerances: operator application in FP64 to 1e-11 relative; solves compared
against the oracle's solve with the same algorithm and tolerance, and against
a tightly converged oracle solution; gradients to the stated bounds below.
"""
import math
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_OBJECTIVE

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dm():
    import dda_model
    return dda_model


@pytest.fixture(scope="module")
def O():
    import dda_oracle
    return dda_oracle


def _args(O, tables, shape, method="LDR", symmetry="None", maxmn=0, seed=0, lam=542.0, d=15.0,
          mat=(1.0 + 0j, 5.96 + 3.89j), k=(0, 0, 1), e=(1, 0, 0), geometry=None, diel=None):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    if diel is None:
        diel = np.repeat(rng.uniform(0.0, 1.0, n), 3)
    a = O.wrapper_args(list(shape), d, diel, list(mat), list(k), list(e), lam,
                       symmetry_axes=[(nx - 1) / 2, (ny - 1) / 2], symmetry_method=symmetry,
                       apc_max_m=maxmn, apc_max_n=maxmn, apc_method_name=method, tables=tables)
    if geometry is not None:
        a[12] = geometry
        a[17] = geometry.size // 3
    return a


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


CASES = [
    ("ldr-5x4x3", dict(shape=(5, 4, 3), method="LDR")),
    ("ldr-16cube", dict(shape=(16, 16, 16), method="LDR")),
    ("fcd-periodic-6x5x4", dict(shape=(6, 5, 4), method="FCD", maxmn=3)),
    ("ldr-7x9x11-odd", dict(shape=(7, 9, 11), method="LDR", k=(0.6, 0, 0.8), e=(0.8, 0, -0.6))),
    ("ldr-1x1x1", dict(shape=(1, 1, 1), method="LDR")),
    ("ldr-24x3x2", dict(shape=(24, 3, 2), method="LDR")),
]


@pytest.mark.parametrize("name,case", CASES, ids=[c[0] for c in CASES])
def test_operator_apply_matches_oracle(dm, O, tables, name, case):
    _log(f"[gpu] operator apply {name}")
    a = _args(O, tables, **case)
    g = dm.DDAModel(*a)
    o = O.OracleDDAModel(*a)
    rng = np.random.default_rng(7)
    n = 3 * o.N
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = g.applyOperator(x)
    assert _rel(y, o.apply(x)) < 1e-11
    assert np.allclose(g.getInverseAlpha(), o.al[0::3], rtol=1e-13)


def test_operator_apply_scattered_geometry(dm, O, tables):
    """Dipoles on a subset of the box in shuffled order (the box map path)."""
    _log("[gpu] operator apply scattered geometry")
    shape = (6, 5, 4)
    full = O.generate_geometry(*shape).reshape(-1, 3)
    rng = np.random.default_rng(11)
    # whole z-columns of a random subset of (x, y), in shuffled order: the
    # reference's parameterisation needs NFpara = N / Nz columns
    cols = rng.permutation(30)[:17]
    keep = np.isin(full[:, 0] + 6 * full[:, 1], cols)
    pick = rng.permutation(np.where(keep)[0])
    geo = full[pick].reshape(-1)
    n = pick.size
    diel = np.repeat(rng.uniform(0, 1, n), 3)
    a = _args(O, tables, shape, geometry=geo, diel=diel)
    g = dm.DDAModel(*a)
    o = O.OracleDDAModel(*a)
    x = rng.standard_normal(3 * n) + 1j * rng.standard_normal(3 * n)
    assert _rel(g.applyOperator(x), o.apply(x)) < 1e-11
    # and a solve through the box map
    g.solveElectricField(np.zeros(3 * n, complex), 10000, 1e-11)
    o.solveElectricField(np.zeros(3 * n, complex), 10000, 1e-11)
    assert _rel(g.getElectricField(), o.getElectricField()) < 1e-8


def test_golden_wrapper_objective(dm, golden_inputs):
    """python_tests/test_DDAModelWrapper.py:32-40 on the GPU path."""
    _log("[gpu] golden wrapper objective (22x22x10 FCD +-50)")
    diel, _ = golden_inputs
    model = dm.DDAModelWrapper([22, 22, 10], 15.0, diel, dielectric_constants=[1.01 + 0j, 5.96282 + 3.80423e-7j],
                               light_direction=[0, 0, 1], light_polarization=[1, 0, 0], light_wavelength_nm=542,
                               symmetry_axes=[10.5, 10.5], verbose=False)
    obj = model.objective()
    _log(f"[gpu] golden objective {obj!r} info {model.solve_info()}")
    assert obj == pytest.approx(GOLDEN_OBJECTIVE, rel=1e-4)


def test_golden_field_matches_oracle(dm, O, tables, golden_inputs):
    """FP64 path (the reference's recursive-residual stopping rule) on the
    golden 22x22x10 FCD case against a tightly converged oracle (1e-11).

    Two solves that stop at a relative residual tol differ from the exact
    solution by about amp * tol in E, amp = 1.5 .. 3 for these systems (the
    field error over the residual, measured along whole trajectories).  So
    the field bound 1e-5 is asserted at tol = 2e-6, the reference's own rule
    at tol = 1e-5 is checked for its iteration window and its field error is
    logged, and at tol = 1e-11 both solve the same system to 1e-9."""
    _log("[gpu] golden field vs oracle")
    diel, _ = golden_inputs
    a = O.wrapper_args([22, 22, 10], 15.0, diel, [1.01 + 0j, 5.96282 + 3.80423e-7j], [0, 0, 1], [1, 0, 0], 542,
                       symmetry_axes=[10.5, 10.5], tables=tables)
    g = dm.DDAModel(*a)
    o = O.OracleDDAModel(*a, workers=4)
    p0 = np.zeros(3 * 4840, complex)
    o.solveElectricField(p0, 100000, 1e-11)
    e_tight = o.getElectricField()
    obj_tight = o.calculateObjective()
    g.solveElectricField(p0, 100000, 1e-5)
    info = g.solveInfo(0)
    err_ref_rule = _rel(g.getElectricField(), e_tight)
    o5 = O.OracleDDAModel(*a, workers=4)
    o5.solveElectricField(p0, 100000, 1e-5)
    _log(f"[gpu] tol 1e-5: iterations gpu {info['iterations']} oracle {o5.info.iterations}; "
         f"E err vs tight: gpu {err_ref_rule:.3e} oracle {_rel(o5.getElectricField(), e_tight):.3e}")
    # the oracle alone stops anywhere in 40..46 iterations here when only the
    # FFT grid (hence the rounding) changes; the recurrence is pinned step by
    # step in test_residual_history_matches_oracle
    assert info["status"] == "converged" and info["error"] <= 1e-5
    assert 36 <= info["iterations"] <= 50
    assert err_ref_rule < 5e-5
    # field within 1e-5 of the tight oracle at the tolerance that achieves it
    g.solveElectricField(p0, 100000, 2e-6)
    err = _rel(g.getElectricField(), e_tight)
    _log(f"[gpu] tol 2e-6: iterations {g.solveInfo(0)['iterations']}, E err vs tight {err:.3e}")
    assert err <= 1e-5
    # tightly converged: both solve the same system
    g.solveElectricField(p0, 100000, 1e-11)
    assert _rel(g.getElectricField(), e_tight) < 1e-9
    assert g.calculateObjective() == pytest.approx(obj_tight, rel=1e-9)


@pytest.mark.parametrize("name,case", [CASES[1], CASES[2], CASES[3]], ids=[CASES[i][0] for i in (1, 2, 3)])
def test_residual_history_matches_oracle(dm, O, tables, name, case):
    """Same BiCGSTAB recurrence and stopping rule (DDAModel.cpp:899-1009).

    The recurrence amplifies rounding about tenfold per step: the oracle on
    two FFT grids (the same operator, rounded differently) agrees to ~1e-11
    after 8 steps but only to ~1e-1 after 30, so the iteration count at which
    the residual first drops below tol moves by several steps under rounding
    alone (66..74 for the 7x9x11 case).  The recurrence itself is pinned by
    the residual after each of the first steps (maxit = k reports it), which
    a wrong scalar (a sign, a swapped omega, a stale rho) changes at O(1);
    the converged field and a loose count window check the rest."""
    _log(f"[gpu] residual history {name}")
    a = _args(O, tables, **case)
    o = O.OracleDDAModel(*a)
    hist = []
    O.bicgstab_reference(o.apply, o.E, np.zeros(3 * o.N, complex), 100000, 1e-6, history=hist)
    g = dm.DDAModel(*a)
    p0 = np.zeros(3 * o.N, complex)
    for k in (1, 2, 3, 4, 6, 8):
        g.solveElectricField(p0, k, 1e-30)
        info = g.solveInfo(0)
        assert info["status"] == "maxit" and info["iterations"] == k
        assert info["error"] == pytest.approx(hist[k - 1], rel=1e-7), k
    g.solveElectricField(p0, 100000, 1e-6)
    info = g.solveInfo(0)
    _log(f"[gpu] iterations gpu {info['iterations']} oracle {len(hist)}")
    assert info["status"] == "converged" and info["error"] <= 1e-6
    assert 0.8 * len(hist) <= info["iterations"] <= 1.25 * len(hist)
    o.solveElectricField(p0, 100000, 1e-6)
    assert _rel(g.getElectricField(), o.getElectricField()) < 1e-5


REFINED_CASES = [
    ("ldr-16cube", dict(shape=(16, 16, 16), method="LDR", seed=0)),
    ("ldr-6x5x32", dict(shape=(6, 5, 32), method="LDR", seed=0)),
    ("ldr-24x20x12-oblique", dict(shape=(24, 20, 12), method="LDR", seed=4, lam=450.0, k=(0.6, 0, 0.8),
                                  e=(0.8, 0, -0.6))),
    ("fcd-periodic-12x10x6", dict(shape=(12, 10, 6), method="FCD", maxmn=2, seed=1)),
]


@pytest.mark.parametrize("name,case", REFINED_CASES, ids=[c[0] for c in REFINED_CASES])
def test_refined_solve_accuracy(dm, O, tables, name, case):
    """complex64 transforms/Krylov with FP64 reliable updates at tol 1e-5:
    the FP64 true residual of the returned P meets 1e-5 and E is within 1e-5
    of the tightly converged oracle (1e-12).  The solver stops when its own
    estimate of the field error (amplification measured between reliable
    updates, times the true residual, times a safety factor 2) meets tol;
    the estimate is reported and checked against the actual error.  The z
    lengths 32 and 64 take the c64 Z splits 8 x 4 and 16 x 4 (the second
    with a partial last tile of pencils)."""
    _log(f"[gpu] refined solve {name}")
    a = _args(O, tables, **case)
    g = dm.DDAModel(*a)
    o = O.OracleDDAModel(*a)
    n = 3 * o.N
    p0 = np.zeros(n, complex)
    o.solveElectricField(p0, 100000, 1e-12)
    e_ref = o.getElectricField()
    g.setSolverPrecision("fp32_refined")
    g.solveElectricField(p0, 100000, 1e-5)
    info = g.solveInfo(0)
    err32 = _rel(g.getElectricField(), e_ref)
    r = np.linalg.norm(o.E - o.apply(g.getPolarization())) / np.linalg.norm(o.E)
    g.setSolverPrecision("fp64")
    g.solveElectricField(p0, 100000, 1e-5)
    info64 = g.solveInfo(0)
    err64 = _rel(g.getElectricField(), e_ref)
    _log(f"[gpu] refined {info}; E err refined {err32:.3e} (estimate {info['err_est']:.3e}), true residual "
         f"{r:.3e}; fp64 at tol 1e-5: {info64['iterations']} iterations, E err {err64:.3e}")
    assert info["status"] == "converged" and info["error"] <= 1e-5
    assert r <= 1.0001e-5
    assert err32 <= 1e-5
    assert 2.0 * info["err_est"] <= 1e-5
    # the estimate follows the actual error
    assert err32 / 3.0 <= info["err_est"] <= 3.0 * err32


def test_refined_maxit_returns_whole_solution(dm, O, tables):
    """A refined solve stopped by the iteration limit ends on a reliable
    update: the reported error is the FP64 true residual of the returned
    polarization, which therefore holds every step (the c64 increments
    accumulated since the last update included)."""
    _log("[gpu] refined maxit 16^3")
    a = _args(O, tables, (16, 16, 16), method="LDR", seed=3)
    g = dm.DDAModel(*a)
    o = O.OracleDDAModel(*a)
    p0 = np.zeros(3 * o.N, complex)
    g.setSolverPrecision("fp32_refined")
    for k in (5, 23):
        g.solveElectricField(p0, k, 1e-30)
        info = g.solveInfo(0)
        assert info["status"] == "maxit" and info["iterations"] == k
        P = g.getPolarization()
        r = np.linalg.norm(o.E - o.apply(P)) / np.linalg.norm(o.E)
        _log(f"[gpu] maxit {k}: reported {info['error']:.6e} oracle {r:.6e} updates {info.get('reliable_updates')}")
        assert info["error"] == pytest.approx(r, rel=1e-6)


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_gradients_match_oracle(dm, O, tables, mode):
    _log(f"[gpu] gradients {mode}")
    a = _args(O, tables, (8, 8, 3), method="FCD", maxmn=2, symmetry="4fold", seed=5)
    g = dm.DDAModel(*a)
    o = O.OracleDDAModel(*a)
    p0 = np.zeros(3 * o.N, complex)
    g.setGradientMode(mode)
    g.solveElectricField(p0, 100000, 1e-12)
    o.solveElectricField(p0, 100000, 1e-12)
    obj = g.calculateObjective()
    assert obj == pytest.approx(o.calculateObjective(), rel=1e-10)
    grad = g.calculateGradients(1e-3, obj, 100000, 1e-12)
    if mode == "reference":
        want = o.calculateGradients(1e-3, obj, 100000, 1e-12)
        assert _rel(grad, want) < 1e-6
    else:
        want = o.exact_gradient(tol=1e-12)
        assert _rel(grad, want) < 1e-8
    if mode == "reference":
        # side effect of the reference: P now holds the adjoint solution
        assert _rel(g.getPolarization(), o.P) < 1e-6


def test_parameters_round_trip(dm, O, tables):
    _log("[gpu] parameters round trip")
    a = _args(O, tables, (6, 6, 2), symmetry="4fold", seed=2)
    g = dm.DDAModel(*a)
    o = O.OracleDDAModel(*a)
    assert np.array_equal(g.getParameters(), o.getParameters())
    assert np.array_equal(g.getDielectrics(), o.getDielectrics())
    newp = np.linspace(0, 1, o.nfpara)
    g.setParameters(newp)
    o.setParameters(newp)
    assert np.array_equal(g.getDielectrics(), o.getDielectrics())
    assert np.allclose(g.getInverseAlpha(), o.al[0::3], rtol=1e-14)
    with pytest.raises(RuntimeError):
        g.setParameters(newp[:-1])


def test_nonconvergence_reports_status(dm, O, tables):
    _log("[gpu] maxit status")
    a = _args(O, tables, (8, 8, 8), seed=1)
    g = dm.DDAModel(*a)
    g.solveElectricField(np.zeros(3 * 512, complex), 2, 1e-12)
    info = g.solveInfo(0)
    assert info["status"] == "maxit" and info["iterations"] == 2


def test_against_committed_fixture(dm, O, tables):
    """Regression against tests/golden/small_cases.npz (written by
    tests/golden/make_golden.py from the oracle)."""
    path = os.path.join(GOLDEN, "small_cases.npz")
    data = np.load(path)
    _log("[gpu] committed fixtures")
    a = _args(O, tables, (6, 5, 4), method="FCD", maxmn=3, seed=21)
    g = dm.DDAModel(*a)
    assert _rel(g.applyOperator(data["x"]), data["y_fcd"]) < 1e-11
    g.solveElectricField(np.zeros(360, complex), 100000, 1e-11)
    assert _rel(g.getElectricField(), data["e_fcd"]) < 1e-8


def test_objective_region_quirk_on_gpu(dm, O, tables):
    """IntegratedE with integral bounds that make the region a proper subset:
    `xMin <= g <= xMax` parses as `(xMin <= g) <= xMax` (ObjDDAModel.cpp:135),
    so xMax = 0 keeps the dipoles with g_x < xMin, and xMax >= 1 keeps all of
    them.  Objective and both gradient modes against the oracle."""
    _log("[gpu] objective region quirk")
    shape = (8, 6, 3)
    n = int(np.prod(shape))
    rng = np.random.default_rng(9)
    diel = np.repeat(rng.uniform(0.0, 1.0, n), 3)
    a = O.wrapper_args(list(shape), 15.0, diel, [1.0 + 0j, 5.96 + 3.89j], [0, 0, 1], [1, 0, 0], 542.0,
                       symmetry_method="None", apc_max_m=0, apc_max_n=0, apc_method_name="LDR", tables=tables,
                       integral_xbounds=[3.0, 0.0], integral_ybounds=[0.0, 5.0], integral_zbounds=[2.0, 0.0])
    o = O.OracleDDAModel(*a)
    region = o.region
    gx = o.geometry.reshape(-1, 3)
    assert np.array_equal(region, (gx[:, 0] < 3) & (gx[:, 2] < 2))
    assert 0 < region.sum() < n
    g = dm.DDAModel(*a)
    p0 = np.zeros(3 * n, complex)
    g.solveElectricField(p0, 100000, 1e-12)
    o.solveElectricField(p0, 100000, 1e-12)
    obj = g.calculateObjective()
    assert obj == pytest.approx(o.calculateObjective(), rel=1e-10)
    # the whole-box objective differs: the region matters
    assert abs(obj - O.objective_integrated(o.al, o.P, np.ones(n, bool), o.pownum)) > 1e-3
    g.setGradientMode("exact")
    grad = g.calculateGradients(1e-3, obj, 100000, 1e-12)
    assert _rel(grad, o.exact_gradient(tol=1e-12)) < 1e-8


C64_SHAPES = [(40, 48, 16), (48, 64, 24), (64, 40, 32)]


def test_c64_apply_matches_fp64(dm, O, tables):
    """The c64 transform kernels (length-specialised X/Y passes for 80, 96
    and 128, the Z pass with its lane-exchanged octant records for 32, 48
    and 64) against the FP64 transforms of the same operators, all in one
    batched launch per pass; the FP64 apply itself is checked against the
    oracle for one of them.  Refined solves of the same shapes then take
    about as many iterations as FP64 ones: a wrong record in the c64 Z pass
    would show up as a large apply error or as extra iterations."""
    from dda_model import native
    _log("[gpu] c64 apply vs fp64 " + str(C64_SHAPES))
    rng = np.random.default_rng(12)
    probs, reqs = [], []
    for i, s in enumerate(C64_SHAPES):
        req = native.Request(shape=s, d=15.0, lam=480.0 + 70.0 * i, rho=rng.uniform(0, 1, int(np.prod(s))),
                             n_k=(0.0, 0.0, 1.0), n_e0=(1.0, 0.0, 0.0))
        reqs.append(req)
        probs.append(native.Problem(req, precision="fp32_refined"))
    xs = [rng.standard_normal(3 * p.n) + 1j * rng.standard_normal(3 * p.n) for p in probs]
    y64 = native.apply(probs, xs, c64=False)
    y32 = native.apply(probs, xs, c64=True)
    for s, a, b in zip(C64_SHAPES, y32, y64):
        e = _rel(a, b)
        _log(f"[gpu] {s}: c64 vs fp64 apply {e:.3e}")
        assert e < 2e-6
    box = O.OracleFullBox(C64_SHAPES[2], 15.0, reqs[2].lam, reqs[2].rho, workers=8)
    assert _rel(y64[2], box.apply(xs[2])) < 1e-11
    # iteration counts of refined vs FP64 solves of the same requests: the
    # refined solve stops later (on its field-error estimate, which needs a
    # true residual of about tol / (2 amp), amp ~ 2) and its c64 vectors
    # cost about a tenth more steps; a wrong octant record would cost far more
    its, infos = {}, {}
    for prec in ("fp64", "fp32_refined"):
        ps = [native.Problem(r, precision=prec) for r in reqs]
        infos[prec] = native.solve(ps, tol=1e-5)
        its[prec] = [i["iterations"] for i in infos[prec]]
    _log(f"[gpu] iterations fp64 {its['fp64']} refined {its['fp32_refined']}; refined true residual "
         f"{[round(i['error'], 8) for i in infos['fp32_refined']]}")
    for a, b in zip(its["fp32_refined"], its["fp64"]):
        assert 0.8 * b <= a <= 1.5 * b


# every class of the fused y/z pass that the serving mix uses, odd sizes in a
# class (ny = 47 -> Ly 96, nz = 31 -> Lz 64), and boxes outside it (Lx = 72
# has no specialised X pass; Ly = 80 is no multiple of 16) in the same batch
FUSED_SHAPES = [(16, 16, 16), (24, 24, 24), (32, 32, 16), (48, 48, 24), (64, 64, 32), (20, 47, 31),
                (40, 46, 16), (33, 31, 24), (20, 40, 16)]


def test_fused_yz_apply_matches_three_passes(dm, O, tables):
    """The fused c64 y/z pass (one workgroup per x frequency) against the
    three-pass c64 path and the FP64 apply on the same operators, in one
    mixed batch; then refined solves with and without it."""
    from dda_model import native
    # the plane passes take precedence over the fused y/z pass: off here
    native.set_plane_passes(False)
    try:
        _fused_yz_body(native, O)
    finally:
        native.set_plane_passes(True)


def _fused_yz_body(native, O):
    _log("[gpu] fused y/z apply " + str(FUSED_SHAPES))
    rng = np.random.default_rng(21)
    probs, reqs = [], []
    for i, s in enumerate(FUSED_SHAPES):
        req = native.Request(shape=s, d=15.0, lam=450.0 + 37.0 * i, rho=rng.uniform(0, 1, int(np.prod(s))),
                             n_k=(0.0, 0.0, 1.0), n_e0=(0.0, 1.0, 0.0))
        reqs.append(req)
        probs.append(native.Problem(req, precision="fp32_refined"))
    xs = [rng.standard_normal(3 * p.n) + 1j * rng.standard_normal(3 * p.n) for p in probs]
    y64 = native.apply(probs, xs, c64=False)
    try:
        native.set_fused_yz("off")
        y3 = native.apply(probs, xs, c64=True)
        native.set_fused_yz("all")
        yf = native.apply(probs, xs, c64=True)
    finally:
        native.set_fused_yz("schedule")
    for s, a, b, c in zip(FUSED_SHAPES, yf, y3, y64):
        ef, e3, d = _rel(a, c), _rel(b, c), _rel(a, b)
        _log(f"[gpu] {s}: fused vs fp64 {ef:.3e}, three-pass vs fp64 {e3:.3e}, fused vs three-pass {d:.3e}")
        assert ef < 2e-6 and e3 < 2e-6 and d < 2e-6
    box = O.OracleFullBox(FUSED_SHAPES[5], 15.0, reqs[5].lam, reqs[5].rho, workers=8)
    assert _rel(y64[5], box.apply(xs[5])) < 1e-11
    # refined solves through the fused pass reach the same fields as through
    # the three passes (each within its 1e-5 field-error target) in about as
    # many iterations
    out = {}
    for fused in (False, True):
        try:
            native.set_fused_yz("all" if fused else "off")
            ps = [native.Problem(r, precision="fp32_refined") for r in reqs[:6]]
            infos = native.solve(ps, tol=1e-5)
            out[fused] = ([i["iterations"] for i in infos], [p.field() for p in ps], infos)
        finally:
            native.set_fused_yz("schedule")
    _log(f"[gpu] refined iterations three-pass {out[False][0]} fused {out[True][0]}")
    for a, b, ea, eb, ia in zip(out[True][0], out[False][0], out[True][1], out[False][1], out[True][2]):
        assert ia["status"] == "converged" and ia["error"] <= 1e-5
        assert abs(a - b) <= max(2, 0.15 * b)
        assert _rel(ea, eb) < 2e-5
