"""Pins for the CPU oracle (oracle/dda_oracle.py, oracle/green_oracle.c).

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it: reference golden
values and fixtures, textbook closed forms, independent derivations and brute
force on tiny inputs.
"""
import math
import os

import numpy as np
import pytest
from scipy.special import sici

import dda_oracle as O
from conftest import GOLDEN, GOLDEN_OBJECTIVE


# ---------------------------------------------------------------------------
# reference fixtures
# ---------------------------------------------------------------------------
def test_geometry_fixture():
    """python_tests/test_DDAModelWrapper.py:42-48: x-fastest dipole order."""
    cfg = np.loadtxt(os.path.join(GOLDEN, "geometry.txt"), dtype=int)
    nx, ny, nz, nt = cfg[:4]
    assert (nx, ny, nz, nt) == (22, 22, 10, 4840)
    assert np.array_equal(O.generate_geometry(nx, ny, nz), cfg[4:])


def test_golden_objective(golden_inputs, tables):
    """python_tests/test_DDAModelWrapper.py:32-40: end to end (4-fold mapping,
    FCD lattice sum with the wrapper's Si/Ci tables, FFT matvec, BiCGSTAB,
    IntegratedE objective) reproduces 11.005338768207627 within 1e-4."""
    diel, _ = golden_inputs
    args = O.wrapper_args([22, 22, 10], 15.0, diel, [1.01 + 0j, 5.96282 + 3.80423e-7j],
                          [0, 0, 1], [1, 0, 0], 542, symmetry_axes=[10.5, 10.5], tables=tables)
    m = O.OracleDDAModel(*args, workers=4)
    assert m.nfpara == 121
    m.solveElectricField(np.zeros(3 * 4840, complex), 100000, 1e-5)
    assert m.info.converged
    obj = m.calculateObjective()
    assert obj == pytest.approx(GOLDEN_OBJECTIVE, rel=1e-4)


# ---------------------------------------------------------------------------
# Green's tensor
# ---------------------------------------------------------------------------
def _draine_flatau(K, r):
    """Interaction matrix of the coupled-dipole method (Draine & Flatau 1994,
    eq. 3): A = exp(ikr)/r [k^2 (rr - I) + (ikr - 1)/r^2 (3 rr - I)]."""
    rn = np.linalg.norm(r)
    u = np.outer(r, r) / rn ** 2
    I = np.eye(3)
    return np.exp(1j * K * rn) / rn * (K ** 2 * (u - I) + (1j * K * rn - 1) / rn ** 2 * (3 * u - I))


@pytest.mark.parametrize("r", [(15.0, 0.0, 0.0), (3.0, -7.0, 11.0), (-40.0, 25.0, -5.0), (300.0, 10.0, 7.0)])
def test_ldr_is_textbook_interaction(r):
    K = 2 * math.pi / 542.0
    got = O.green_point(K, 15.0, O.METHOD_LDR, *r)
    want = _draine_flatau(K, np.array(r))
    assert np.allclose(got, want, rtol=1e-12, atol=1e-18)


def test_green_zero_offset_is_zero():
    K = 2 * math.pi / 500.0
    assert np.all(O.green_point(K, 10.0, O.METHOD_LDR, 0, 0, 0) == 0)


def _fine_tables(delta, n):
    """Tables whose sample p sits exactly at (p+1)*delta, so the reference's
    lookup is plain linear interpolation of exact Si/Ci."""
    pos = (np.arange(n) + 1.0) * delta
    si, ci = sici(pos)
    return si, ci


def _fcd_from_definition(K, d, r):
    """Filtered coupled dipole tensor from its definition (Piller & Martin 1998):
    -[(K^2 g + g'/r + 4pi/3 h) I + (g'' - g'/r) rr], with the filtered scalar
    Green's function g(r) = [sin(Kr) (i pi + Ci((Kf-K)r) - Ci((Kf+K)r))
    + cos(Kr) (Si((Kf+K)r) + Si((Kf-K)r))]/(pi r) and the filtered delta
    h(r) = (sin(Kf r) - Kf r cos(Kf r))/(2 pi^2 r^3), Kf = pi/d.  The radial
    derivatives are taken numerically here (5-point stencil) with exact Si/Ci."""
    kf = math.pi / d

    def g(rr):
        sp, cp = sici((kf + K) * rr)
        sm, cm = sici((kf - K) * rr)
        return (np.sin(K * rr) * (1j * math.pi + cm - cp) + np.cos(K * rr) * (sp + sm)) / (math.pi * rr)

    rn = np.linalg.norm(r)
    hstep = 1e-3 * rn
    f = [g(rn + k * hstep) for k in (-2, -1, 0, 1, 2)]
    g0 = f[2]
    g1 = (f[0] - 8 * f[1] + 8 * f[3] - f[4]) / (12 * hstep)
    g2 = (-f[0] + 16 * f[1] - 30 * f[2] + 16 * f[3] - f[4]) / (12 * hstep ** 2)
    h = (math.sin(kf * rn) - kf * rn * math.cos(kf * rn)) / (2 * math.pi ** 2 * rn ** 3)
    u = np.outer(r, r) / rn ** 2
    return -((K ** 2 * g0 + g1 / rn + 4 * math.pi / 3 * h) * np.eye(3) + (g2 - g1 / rn) * u)


@pytest.mark.parametrize("r", [(15.0, 0.0, 0.0), (15.0, 30.0, -15.0), (-45.0, 60.0, 15.0)])
def test_fcd_matches_definition(r):
    K = 2 * math.pi / 542.0
    d = 15.0
    delta = 2e-5
    si, ci = _fine_tables(delta, 1_000_000)
    got = O.green_point(K, d, O.METHOD_FCD, *r, si=si, ci=ci, delta=delta)
    want = _fcd_from_definition(K, d, np.array(r))
    scale = np.abs(want).max()
    assert np.abs(got - want).max() / scale < 2e-6


def test_fcd_tends_to_ldr_far_away():
    """Kf r -> infinity: Si -> pi/2, Ci -> 0, the filtered function becomes the
    point-dipole one; the relative gap must shrink as r grows."""
    K = 2 * math.pi / 600.0
    d = 10.0
    si, ci = _fine_tables(0.1, 1_000_000)
    gaps = []
    # the filtered terms oscillate at Kf with relative size ~ Kf/(K^2 r), so
    # the limit is only reached for r >> Kf/K^2 (about 3 um here)
    for rr in (1e4, 3e4, 1e5):
        r = np.array([rr, 0.3 * rr, 0.2 * rr])
        f = O.green_point(K, d, O.METHOD_FCD, *r, si=si, ci=ci, delta=0.1)
        l = O.green_point(K, d, O.METHOD_LDR, *r)
        gaps.append(np.abs(f - l).max() / np.abs(l).max())
    assert gaps[2] < 0.02 and gaps[2] < gaps[0] / 3


def test_sici_lookup_semantics(tables):
    """SiCi.cpp:12-44: abscissae (p+1)*delta, past-the-end returns the last
    entry, exact grid points give 0/0."""
    si, ci, delta = tables
    x = 3.456
    p1 = math.floor(x / delta) - 1
    p2 = math.ceil(x / delta) - 1
    a1, a2 = (p1 + 1) * delta, (p2 + 1) * delta
    want = si[p1] + (si[p2] - si[p1]) / (a2 - a1) * (x - a1)
    assert O.sici_lookup(si, delta, x) == pytest.approx(want, rel=1e-15)
    assert O.sici_lookup(si, delta, 1e7) == si[-1]
    assert math.isnan(O.sici_lookup(si, delta, 0.5))   # 0.5/0.1 is exact: 0/0
    # the wrapper's tables are offset from the lookup's abscissae: the lookup is
    # not the exact Si, and the difference is visible at this level
    assert abs(O.sici_lookup(si, delta, 3.456) - sici(3.456)[0]) > 1e-7


def test_lattice_octant_equals_full(tables):
    si, ci, delta = tables
    K = 2 * math.pi / 542.0
    for method, mm in ((O.METHOD_FCD, 3), (O.METHOD_LDR, 2), (O.METHOD_LDR, 0)):
        a = O.green_offsets(3, 4, 2, 15.0, K, method, mm, mm, 3 * 15.0, 4 * 15.0, si, ci, delta, octant_only=False)
        b = O.green_offsets(3, 4, 2, 15.0, K, method, mm, mm, 3 * 15.0, 4 * 15.0, si, ci, delta, octant_only=True)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(a).max())


def test_lattice_sum_brute_force(tables):
    """Image sum with x shifted by m*Lm and y by n*Ln, each image damped by
    exp(-(0.01 K r)^4) (AProductCore.cpp:335-349)."""
    si, ci, delta = tables
    K = 2 * math.pi / 500.0
    d, lm, ln, mm = 10.0, 40.0, 30.0, 2
    cent = O.green_offsets(2, 2, 2, d, K, O.METHOD_FCD, mm, mm, lm, ln, si, ci, delta)
    for o in [(0, 0, 0), (1, -1, 1), (-1, 1, 0)]:
        acc = np.zeros((3, 3), complex)
        for m in range(-mm, mm + 1):
            for n in range(-mm, mm + 1):
                r = np.array([d * o[0] + m * lm, d * o[1] + n * ln, d * o[2]])
                w = math.exp(-(0.01 * K * np.linalg.norm(r)) ** 4)
                acc += w * O.green_point(K, d, O.METHOD_FCD, *r, si=si, ci=ci, delta=delta)
        got = cent[o[0] + 1, o[1] + 1, o[2] + 1]
        want = np.array([acc[i, j] for i, j in O.SYM_INDEX])
        assert np.allclose(got, want, rtol=1e-12, atol=1e-16)


# ---------------------------------------------------------------------------
# matvec, solver
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 4, 5)])
def test_fft_matvec_equals_dense(shape, tables):
    si, ci, delta = tables
    nx, ny, nz = shape
    K = 2 * math.pi / 542.0
    cent = O.green_offsets(nx, ny, nz, 15.0, K, O.METHOD_FCD, 1, 1, nx * 15.0, ny * 15.0, si, ci, delta)
    geo = O.generate_geometry(nx, ny, nz)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(3 * nx * ny * nz) + 1j * rng.standard_normal(3 * nx * ny * nz)
    dense = O.dense_interaction_matrix(cent, geo)
    assert np.allclose(dense, dense.T)   # complex symmetric
    y_ref = dense @ x
    y_fft = O.FFTOperator(cent, geo, nx, ny, nz).apply(x)
    y_2n = O.FFTOperator(cent, geo, nx, ny, nz, fft_shape=(2 * nx, 2 * ny, 2 * nz)).apply(x)
    scale = np.abs(y_ref).max()
    assert np.abs(y_fft - y_ref).max() < 1e-12 * scale
    assert np.abs(y_2n - y_ref).max() < 1e-12 * scale


def test_fft_matvec_scattered_geometry(tables):
    """Dipoles at a subset of the box, in a shuffled order."""
    si, ci, delta = tables
    K = 2 * math.pi / 542.0
    nx, ny, nz = 4, 3, 3
    cent = O.green_offsets(nx, ny, nz, 15.0, K, O.METHOD_LDR, 0, 0, 0.0, 0.0)
    full = O.generate_geometry(nx, ny, nz).reshape(-1, 3)
    rng = np.random.default_rng(1)
    pick = rng.permutation(len(full))[:20]
    geo = full[pick].reshape(-1)
    x = rng.standard_normal(60) + 1j * rng.standard_normal(60)
    y = O.FFTOperator(cent, geo, nx, ny, nz).apply(x)
    assert np.allclose(y, O.dense_interaction_matrix(cent, geo) @ x, rtol=1e-11, atol=1e-14)


def test_bicgstab_solves_dense_system(tables):
    si, ci, delta = tables
    K = 2 * math.pi / 542.0
    nx, ny, nz = 3, 3, 2
    cent = O.green_offsets(nx, ny, nz, 15.0, K, O.METHOD_LDR, 0, 0, 0.0, 0.0)
    geo = O.generate_geometry(nx, ny, nz)
    n = 3 * nx * ny * nz
    al = 1.0 / O.get_alpha(K, 15.0, np.full(n, 3.0 + 0.5j), [1, 0, 0], [0, 0, 1])
    A = O.dense_interaction_matrix(cent, geo) + np.diag(al)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, info = O.bicgstab_reference(lambda v: A @ v, b, np.zeros(n, complex), 1000, 1e-13)
    assert info.converged
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-14)
    # the stopping rule is on the recursive residual ||r||/||b|| after a full step
    _, info2 = O.bicgstab_reference(lambda v: A @ v, b, np.zeros(n, complex), 1000, 1e-3)
    assert info2.error <= 1e-3 and info2.iterations <= info.iterations


def _textbook_bicgstab(A, b, steps):
    """BiCGSTAB as published (van der Vorst 1992, Algorithm 3; inner product
    (u, w) = sum conj(u) w), written out independently of the oracle: returns
    the iterate and the relative residual after each of `steps` iterations."""
    x = np.zeros_like(b)
    r = b - A @ x
    rhat = r.copy()
    rho_old = alpha = omega = 1.0 + 0j
    v = np.zeros_like(b)
    p = np.zeros_like(b)
    xs, res = [], []
    for _ in range(steps):
        rho = np.vdot(rhat, r)
        beta = (rho / rho_old) * (alpha / omega)
        p = r + beta * (p - omega * v)
        v = A @ p
        alpha = rho / np.vdot(rhat, v)
        s = r - alpha * v
        t = A @ s
        omega = np.vdot(t, s) / np.vdot(t, t)
        x = x + alpha * p + omega * s
        r = s - omega * t
        rho_old = rho
        xs.append(x.copy())
        res.append(np.linalg.norm(r) / np.linalg.norm(b))
    return xs, res


def test_bicgstab_recurrence_is_textbook(tables):
    """The reference's BiCGSTAB (DDAModel.cpp:899-1009: a first step before
    the loop with p = r, then beta = (alpha/zeta) <r0,r>/<r0,rl>,
    p = r + beta (p - zeta v), P += alpha p + zeta r - alpha zeta v) is the
    textbook recurrence: same iterates and residuals step by step on a small
    dense DDA system (a dropped term, a swapped alpha/zeta or a conjugation
    on the wrong side changes them at O(1)).  The system is well conditioned
    (cond ~ 20) so that rounding, which the recurrence amplifies, stays below
    1e-9 over the ten steps compared."""
    K = 2 * math.pi / 800.0
    nx, ny, nz = 3, 3, 2
    cent = O.green_offsets(nx, ny, nz, 15.0, K, O.METHOD_LDR, 0, 0, 0.0, 0.0)
    geo = O.generate_geometry(nx, ny, nz)
    n = 3 * nx * ny * nz
    rng = np.random.default_rng(2)
    al = 1.0 / O.get_alpha(K, 15.0, 1.0 + rng.uniform(0, 1, n) * 0.4 * (4.96 + 3.89j), [1, 0, 0], [0, 0, 1])
    A = O.dense_interaction_matrix(cent, geo) + np.diag(al)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    steps = 10
    xs, res = _textbook_bicgstab(A, b, steps)
    for k in range(steps):
        # maxit = k loop steps after the first step = textbook iteration k + 1
        P, _ = O.bicgstab_reference(lambda v: A @ v, b, np.zeros(n, complex), k, 0.0)
        assert np.linalg.norm(P - xs[k]) <= 1e-8 * np.linalg.norm(xs[k]), k
    hist = []
    O.bicgstab_reference(lambda v: A @ v, b, np.zeros(n, complex), steps - 1, 0.0, history=hist)
    assert np.allclose(hist, res[1:steps], rtol=1e-8)
    assert res[-1] < 1e-4 * res[0]


def test_target_fields_local_field_identity():
    """At a dipole site the field of all the other dipoles plus the incident
    field is the local field that polarises it: E_inc,i - sum_{j != i}
    A(r_i - r_j) P_j = al_i P_i for a solved system (DDAModel.cpp:1153-1192
    evaluated at r_i; the interaction is zero at r = 0).  The solve's tensor
    carries the damping exp(-(0.01 K r)^4), which is 1 - 1e-8 over this box,
    and target_fields does not, hence the tolerance.  The vectorised LDR
    evaluation equals the loop version."""
    rng = np.random.default_rng(3)
    shape = (5, 4, 3)
    box = O.OracleFullBox(shape, 12.0, 542.0, rng.uniform(0, 1, 60), n_k=(0.0, 0.6, 0.8), n_e0=(1.0, 0.0, 0.0))
    box.solve(tol=1e-14)
    sites = box.geometry.reshape(-1, 3)[[0, 17, 33, 59]] * 12.0
    E = box.target_fields(sites)
    want = (box.al * box.P).reshape(-1, 3)[[0, 17, 33, 59]]
    assert np.allclose(E, want, rtol=1e-6, atol=1e-9 * np.abs(want).max())
    tg = np.array([[20.0, 18.0, 60.0], [0.0, 0.0, -40.0], [13.0, 7.0, 5.0]])
    assert np.allclose(box.target_fields_ldr(tg), box.target_fields(tg), rtol=1e-12, atol=1e-15)
    assert np.allclose(box.target_fields_ldr(sites), E, rtol=1e-12, atol=1e-15)


def test_bicgstab_maxit_zero_and_zero_rhs():
    A = np.eye(3) * (2 + 1j)
    b = np.ones(3, complex)
    _, info = O.bicgstab_reference(lambda v: A @ v, b, np.zeros(3, complex), 0, 1e-5)
    assert not info.converged and info.iterations == 0


# ---------------------------------------------------------------------------
# polarizability, objective, mapping
# ---------------------------------------------------------------------------
def test_alpha_clausius_mossotti_limit():
    d = 15.0
    eps = np.array([2.0 + 0.1j, 5.96 + 3.89j, 1.01 + 0j])
    cm = 3 * d ** 3 / (4 * math.pi) * (eps - 1) / (eps + 2)
    a = O.get_alpha(1e-9, d, eps, [1, 0, 0], [0, 0, 1])
    assert np.allclose(a, cm, rtol=1e-12)


def test_inverse_alpha_closed_form():
    """1/alpha = 1/a_CM + K^2 (b1 + eps (b2 + b3 S))/d - (2/3) i K^3, an
    algebraic rearrangement of Tools.cpp:353-366 (Draine & Goodman 1993 LDR);
    for real eps its imaginary part is the radiative reaction -2K^3/3."""
    K, d = 2 * math.pi / 542.0, 15.0
    ne, nk = np.array([0.6, 0.8, 0.0]), np.array([0.0, 0.6, 0.8])
    S = np.sum((ne * nk) ** 2)
    eps = np.array([1.5 + 0j, 4.0 + 0j, 5.96 + 3.89j])
    acm = 3 * d ** 3 / (4 * math.pi) * (eps - 1) / (eps + 2)
    want = 1 / acm + K ** 2 * (-1.891531 + eps * (0.1648469 - 1.7700004 * S)) / d - 2j / 3 * K ** 3
    got = 1 / O.get_alpha(K, d, eps, ne, nk)
    assert np.allclose(got, want, rtol=1e-12)
    assert np.allclose(got[:2].imag, -2 * K ** 3 / 3, rtol=1e-9)


def test_dal_deps_matches_difference():
    K, d = 2 * math.pi / 542.0, 15.0
    eps = 3.0 + 1.0j
    h = 1e-6
    fd = (1 / O.get_alpha(K, d, eps + h, [1, 0, 0], [0, 0, 1]) - 1 / O.get_alpha(K, d, eps - h, [1, 0, 0], [0, 0, 1])) / (2 * h)
    assert O.dal_deps(K, d, eps, [1, 0, 0], [0, 0, 1]) == pytest.approx(fd, rel=1e-8)


def test_material_nudge_and_interp():
    m = O.adjust_material([1.01 + 0j, 5.96282 + 3.80423e-7j], math.sqrt(1.01))
    assert m[0] == pytest.approx(1.01)          # 1.01/1.01 = 1 -> +0.01
    assert m[1] == pytest.approx((5.96282 + 3.80423e-7j) / 1.01)
    e = O.interp_eps([0.0, 0.25, 1.0, 1.5], m)
    assert e[0] == m[0] and e[2] == m[1] and e[3] == m[1]
    assert e[1] == pytest.approx(m[0] + 0.25 * (m[1] - m[0]))
    with pytest.raises(ValueError):
        O.interp_eps([-0.1], m)


def test_objective_region_predicate():
    """(xMin <= g) <= xMax: every dipole counts once the upper bounds are >= 1;
    with an upper bound of 0 only coordinates below the lower bound count."""
    geo = O.generate_geometry(3, 2, 2)
    all_in = O.objective_region(geo, [2, 1, 1, 1, 1, 1, 1])
    assert all_in.all()   # even though only x = y = z = 1 is "inside" [1,1]^3
    some = O.objective_region(geo, [2, 1, 0, 0, 5, 0, 5])
    g = geo.reshape(-1, 3)
    assert np.array_equal(some, g[:, 0] < 1)


def test_parameter_mapping_4fold():
    geo = O.generate_geometry(4, 4, 2)
    nf, gp, p2g = O.parameter_mapping(geo, 32, 2, "4fold", [1.5, 1.5])
    assert nf == 4
    assert sorted(len(c) for c in p2g) == [8, 8, 8, 8]
    g = geo.reshape(-1, 3)
    for i in range(32):
        x, y = g[i, :2]
        # mirror images share the parameter
        j = np.where((g[:, 0] == 3 - x) & (g[:, 1] == y) & (g[:, 2] == g[i, 2]))[0][0]
        assert gp[i] == gp[j]
    nf2, gp2, _ = O.parameter_mapping(geo, 32, 2, "None", [])
    assert nf2 == 16 and gp2[16] == gp2[0]   # z extrusion
    with pytest.raises(RuntimeError):
        O.parameter_mapping(geo, 32, 2, "8fold", [])


# ---------------------------------------------------------------------------
# gradients
# ---------------------------------------------------------------------------
def _small_model(tables, sym="None", method="LDR", seed=3, shape=(4, 4, 2), maxmn=0):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    n = nx * ny * nz
    diel = np.repeat(rng.uniform(0.05, 0.95, n), 3)
    args = O.wrapper_args(list(shape), 15.0, diel, [1.0 + 0j, 3.5 + 0.8j], [0, 0, 1], [1, 0, 0], 542,
                          symmetry_axes=[(nx - 1) / 2, (ny - 1) / 2], symmetry_method=sym,
                          apc_max_m=maxmn, apc_max_n=maxmn, apc_method_name=method, tables=tables)
    return O.OracleDDAModel(*args)


def test_exact_gradient_matches_central_fd(tables):
    m = _small_model(tables)
    m.solveElectricField(np.zeros(3 * m.N, complex), 10000, 1e-13)
    g = m.exact_gradient(tol=1e-13)
    p0 = m.getParameters()
    rng = np.random.default_rng(4)
    h = 1e-5
    for i in rng.choice(m.nfpara, 6, replace=False):
        vals = []
        for s in (+1, -1):
            p = p0.copy()
            p[i] += s * h
            m.setParameters(p)
            m.solveElectricField(np.zeros(3 * m.N, complex), 10000, 1e-13)
            vals.append(m.calculateObjective())
        m.setParameters(p0)
        fd = (vals[0] - vals[1]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_reference_gradient_vectorised_equals_sequential(tables):
    m = _small_model(tables, sym="4fold")
    m.solveElectricField(np.zeros(3 * m.N, complex), 10000, 1e-12)
    obj = m.calculateObjective()
    dx1, ad1, dp1 = m.fd_partials(1e-3, obj)
    dx2, ad2, dp2 = O.reference_gradient_sequential(m, 1e-3, obj)
    assert np.allclose(ad1, ad2, rtol=1e-9, atol=1e-12)
    # FD quotients amplify the rounding of the incremental sums: compare loosely
    assert np.allclose(dx1, dx2, rtol=1e-4, atol=1e-7 * np.abs(dx1).max())
    assert np.allclose(dp1, dp2, rtol=1e-3, atol=1e-4 * np.abs(dp1).max())


def test_reference_gradient_close_to_exact(tables):
    """One-sided FD with eps = 1e-3 differs from the exact gradient at O(eps)."""
    m = _small_model(tables, sym="4fold")
    m.solveElectricField(np.zeros(3 * m.N, complex), 10000, 1e-12)
    obj = m.calculateObjective()
    ge = m.exact_gradient(tol=1e-12)
    errs = []
    for eps in (1e-3, 1e-4):
        m.solveElectricField(np.zeros(3 * m.N, complex), 10000, 1e-12)
        gr = m.calculateGradients(eps, obj, 10000, 1e-12)
        errs.append(np.linalg.norm(gr - ge) / np.linalg.norm(ge))
    assert errs[0] < 0.1
    assert 5 < errs[0] / errs[1] < 20   # first-order convergence in eps


# ---------------------------------------------------------------------------
# full-box (per-voxel density) oracle used for the device problems
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["integrated", "targets"])
def test_full_box_gradient_matches_central_fd(kind):
    rng = np.random.default_rng(8)
    shape = (4, 3, 2)
    rho = rng.uniform(0.1, 0.9, 24)
    targets = [(15.0, 10.0, 60.0), (30.0, 20.0, 75.0)]
    box = O.OracleFullBox(shape, 12.0, 542.0, rho, n_k=(0.2, 0.0, math.sqrt(0.96)), n_e0=(0.0, 1.0, 0.0))
    box.solve(tol=1e-13)
    f = (lambda: box.objective_integrated(2)) if kind == "integrated" else (lambda: box.objective_targets(targets))
    g = box.gradient(kind=kind, targets=targets, tol=1e-13)
    h = 1e-5
    for i in (0, 7, 23):
        vals = []
        for s in (1, -1):
            r = rho.copy()
            r[i] += s * h
            box.set_density(r)
            box.solve(tol=1e-13)
            vals.append(f())
        box.set_density(rho)
        box.solve(tol=1e-13)
        assert g[i] == pytest.approx((vals[0] - vals[1]) / (2 * h), rel=1e-5, abs=1e-10)


def test_full_box_matches_model_when_layers_agree(tables):
    """With z-uniform densities the per-voxel box and the z-extruded model are
    the same system."""
    rng = np.random.default_rng(9)
    shape = (4, 4, 3)
    col = rng.uniform(0, 1, 16)
    rho = np.tile(col, 3)
    box = O.OracleFullBox(shape, 15.0, 542.0, rho, method="FCD", images=2, period=(4, 4), tables=tables)
    box.solve(tol=1e-12)
    a = O.wrapper_args(list(shape), 15.0, np.repeat(rho, 3), [1.0 + 0j, 5.96 + 3.89j], [0, 0, 1], [1, 0, 0], 542.0,
                       symmetry_method="None", apc_max_m=2, apc_max_n=2, apc_method_name="FCD", tables=tables)
    m = O.OracleDDAModel(*a)
    m.solveElectricField(np.zeros(3 * m.N, complex), 10000, 1e-12)
    assert np.allclose(box.field(), m.getElectricField(), rtol=1e-9, atol=1e-12)
