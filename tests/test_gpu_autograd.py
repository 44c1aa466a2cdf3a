"""GPU tests of the adjoint gradient and the torch.autograd.Function.

Bars: the per-voxel adjoint gradient of the IntegratedE objective equals the
oracle's exact adjoint gradient to 1e-7 relative (FP64, tight solves); both
objectives' gradients match central finite differences (h = 1e-4, FP64 solves
to 1e-12) to 1e-4 relative per voxel; autograd through the cone filter agrees
with finite differences of the filtered objective to 1e-4.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def native():
    from dda_model import native
    return native


def _req(native, shape, seed, method="LDR"):
    rng = np.random.default_rng(seed)
    return native.Request(shape=shape, d=12.0, lam=542.0, rho=rng.uniform(0.05, 0.95, int(np.prod(shape))),
                          n_k=(0.0, 0.0, 1.0), n_e0=(1.0, 0.0, 0.0), method=method)


def test_integrated_gradient_matches_oracle(native, tables):
    import dda_oracle as O
    _log("[gpu] integrated gradient vs oracle exact adjoint")
    shape = (8, 6, 4)
    req = _req(native, shape, seed=1)
    p = native.Problem(req, precision="fp64")
    assert p.solve(tol=1e-12)["status"] == "converged"
    g, val, info = p.gradient(kind="integrated", power=2, tol=1e-12)
    box = O.OracleFullBox(shape, 12.0, 542.0, req.rho)
    box.solve(tol=1e-12)
    assert val == pytest.approx(box.objective_integrated(2), rel=1e-11)
    ge = box.gradient(kind="integrated", tol=1e-12)
    assert np.linalg.norm(g - ge) / np.linalg.norm(ge) < 1e-7


def test_target_gradient_matches_oracle(native):
    import dda_oracle as O
    from dda_model.autograd import target_block
    _log("[gpu] target objective and gradient vs oracle")
    shape = (6, 5, 3)
    req = _req(native, shape, seed=6)
    tg = target_block(shape, 12.0, size=2, height_nm=30.0)
    p = native.Problem(req, precision="fp64")
    p.solve(tol=1e-12)
    g, val, _ = p.gradient(kind="targets", targets=tg, tol=1e-12)
    box = O.OracleFullBox(shape, 12.0, 542.0, req.rho)
    box.solve(tol=1e-12)
    assert val == pytest.approx(box.objective_targets(tg), rel=1e-10)
    ge = box.gradient(kind="targets", targets=tg, tol=1e-12)
    assert np.linalg.norm(g - ge) / np.linalg.norm(ge) < 1e-7


@pytest.mark.parametrize("kind", ["integrated", "targets"])
def test_gradient_matches_central_fd(native, kind):
    _log(f"[gpu] {kind} gradient vs central FD")
    from dda_model.autograd import target_block
    shape = (10, 10, 5)
    req = _req(native, shape, seed=2)
    targets = target_block(shape, 12.0, size=2, height_nm=36.0) if kind == "targets" else None
    p = native.Problem(req, precision="fp64")
    p.solve(tol=1e-12)
    g, val, _ = p.gradient(kind=kind, targets=targets, tol=1e-12)
    rng = np.random.default_rng(3)
    h = 1e-4
    for i in rng.choice(p.n, 6, replace=False):
        vals = []
        for s in (1, -1):
            r = req.rho.copy()
            r[i] += s * h
            p.set_density(r)
            p.solve(tol=1e-12)
            vals.append(p.objective(kind=kind, targets=targets))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(g).max())
    p.set_density(req.rho)


def test_autograd_function_and_filter(native):
    import torch
    from dda_model.autograd import cone_filter, dda_objective, target_block
    _log("[gpu] autograd through the cone filter")
    shape = (12, 12, 6)
    rng = np.random.default_rng(4)
    raw = torch.tensor(rng.uniform(0, 1, (6, 12, 12)), dtype=torch.float64, device="cuda", requires_grad=True)
    req = native.Request(shape=shape, d=12.0, lam=542.0, rho=np.full(int(np.prod(shape)), 0.5))
    p = native.Problem(req, precision="fp64")
    tg = target_block(shape, 12.0, size=2)

    def f(x):
        return dda_objective(cone_filter(x, 2.0).reshape(-1), p, kind="targets", targets=tg, tol=1e-12)

    val = f(raw)
    val.backward()
    grad = raw.grad.detach().cpu().numpy()
    h = 1e-4
    for idx in [(0, 0, 0), (3, 5, 7), (5, 11, 2)]:
        with torch.no_grad():
            xp = raw.detach().clone()
            xp[idx] += h
            xm = raw.detach().clone()
            xm[idx] -= h
            fd = (f(xp).item() - f(xm).item()) / (2 * h)
        assert grad[idx] == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(grad).max())


def test_refined_gradient_close_to_fp64(native):
    _log("[gpu] refined vs fp64 gradient")
    shape = (16, 16, 8)
    req = _req(native, shape, seed=5)
    grads = {}
    for prec in ("fp64", "fp32_refined"):
        p = native.Problem(req, precision=prec)
        p.solve(tol=1e-7)
        grads[prec], _, _ = p.gradient(kind="integrated", tol=1e-7)
    err = np.linalg.norm(grads["fp32_refined"] - grads["fp64"]) / np.linalg.norm(grads["fp64"])
    _log(f"[gpu] refined gradient rel err {err:.3e}")
    assert err < 1e-5


def test_gradient_outside_unit_interval(native):
    """Densities are clipped to [0, 1] before the permittivity mix, so the
    objective is flat in a voxel whose density lies outside: the gradient
    there is zero, and central differences agree at such voxels and inside."""
    _log("[gpu] gradient outside [0, 1]")
    shape = (8, 6, 4)
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.1, 0.9, int(np.prod(shape)))
    out_idx = {3: -0.1, 40: 1.2, 77: 1.0 + 1e-3}
    for i, v in out_idx.items():
        rho[i] = v
    req = native.Request(shape=shape, d=12.0, lam=542.0, rho=rho)
    p = native.Problem(req, precision="fp64")
    p.solve(tol=1e-13)
    g, val, _ = p.gradient(kind="integrated", tol=1e-13)
    h = 1e-4
    for i in list(out_idx) + [10, 55]:
        vals = []
        for s in (1, -1):
            r = rho.copy()
            r[i] += s * h
            p.set_density(r)
            p.solve(tol=1e-13)
            vals.append(p.objective(kind="integrated"))
        fd = (vals[0] - vals[1]) / (2 * h)
        if i in out_idx:
            assert g[i] == 0.0 and abs(fd) < 1e-9, (i, g[i], fd)
        else:
            assert g[i] == pytest.approx(fd, rel=1e-5), (i, g[i], fd)
    p.set_density(rho)
