"""GPU parity at the benchmark configurations' own sizes.

  periodic unit cell   32x32x12, d = 10 nm, FCD with +-50 images of a 320 nm
                       period, 600 nm, normal incidence, y-polarised, smooth
                       density (Gaussian-filtered noise, sigma 2, clipped)
  serving grids        48x48x24 and 64x64x32 (X/Y lengths 96 and 128, Z 48 and 64)
  adjoint gradient     48x48x24, d = 12 nm, cone-filtered density, mean |E|^2
                       over 4x4x4 points 60 nm above the structure
  multi-RHS            128x128x64, d = 8 nm (FFT lengths 256, 256, 128), smooth
                       density (sigma 3), 8 right-hand sides = 4 wavelengths x
                       2 polarisations in one batch

Bars: FP64 operator application against the oracle to 1e-11; refined (c64 +
FP64 reliable updates) fields within 1e-5 of a tightly converged FP64
reference, with the true residual recomputed by the oracle; FP64 gradient
against central differences (h = 1e-4); refined gradient within 1e-4 of the
FP64 one.  Where an oracle solve at full size would take minutes on the CPU,
the oracle supplies the operator (one FP64 matvec) and the GPU's FP64 solve
converged to 1e-10 is the reference field: its true residual is checked by
that oracle matvec.
"""
import math
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
    return native


@pytest.fixture(scope="module")
def O():
    import dda_oracle
    return dda_oracle


def smooth_density(shape, sigma, seed):
    """rho = clip(0.5 + 0.25 f / std f, 0, 1), f = Gaussian-filtered white noise
    (periodic in x and y, nearest-value edges in z), flattened x-fastest."""
    from scipy.ndimage import gaussian_filter
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal((nz, ny, nx)), sigma=sigma, mode=("nearest", "wrap", "wrap"))
    return np.clip(0.5 + 0.25 * f / f.std(), 0.0, 1.0).reshape(-1)


def _true_residual(box, P):
    return float(np.linalg.norm(box.Einc - box.apply(P)) / np.linalg.norm(box.Einc))


# ---------------------------------------------------------------------------
# periodic unit cell, FCD, +-50 images
# ---------------------------------------------------------------------------
def test_periodic_unit_cell_32x32x12(native, O, tables):
    shape = (32, 32, 12)
    rho = smooth_density(shape, 2.0, seed=1)
    _log("[gpu] periodic unit cell 32x32x12 FCD +-50: oracle lattice sum")
    box = O.OracleFullBox(shape, 10.0, 600.0, rho, n_k=(0.0, 0.0, 1.0), n_e0=(0.0, 1.0, 0.0), method="FCD",
                          images=50, period=(32, 32), tables=tables, workers=16)
    req = native.Request(shape=shape, d=10.0, lam=600.0, rho=rho, n_k=(0.0, 0.0, 1.0), n_e0=(0.0, 1.0, 0.0),
                         method="FCD", images=50, period=(32, 32))
    p64 = native.Problem(req, precision="fp64", tables=tables)
    p32 = native.Problem(req, precision="fp32_refined", tables=tables)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(3 * box.N) + 1j * rng.standard_normal(3 * box.N)
    assert _rel(native.apply([p64], [x])[0], box.apply(x)) < 1e-11
    _log("[gpu] periodic unit cell: oracle tight solve")
    box.solve(tol=1e-10)
    e_ref = box.field()
    i64 = p64.solve(tol=1e-10)
    assert i64["status"] == "converged"
    assert _rel(p64.field(), e_ref) < 1e-8
    i32 = p32.solve(tol=1e-5)
    err = _rel(p32.field(), e_ref)
    r = _true_residual(box, p32.polarization())
    i64b = p64.solve(tol=1e-5)
    err64 = _rel(p64.field(), e_ref)
    _log(f"[gpu] periodic unit cell: refined {i32['iterations']} its, true residual {r:.3e}, E err {err:.3e} "
         f"(estimate {i32['err_est']:.3e}); fp64 at tol 1e-5: {i64b['iterations']} its, E err {err64:.3e}")
    assert i32["status"] == "converged" and r <= 1.0001e-5
    assert err <= 1e-5


# ---------------------------------------------------------------------------
# the two largest serving grids
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(48, 48, 24), (64, 64, 32)], ids=["48x48x24", "64x64x32"])
def test_serving_grids(native, O, shape):
    rng = np.random.default_rng(sum(shape))
    th, ph, ps = 0.4, 1.1, 0.7
    k = (math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th))
    e1 = (math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th))
    e2 = (-math.sin(ph), math.cos(ph), 0.0)
    e = tuple(math.cos(ps) * a + math.sin(ps) * b for a, b in zip(e1, e2))
    rho = rng.uniform(0, 1, int(np.prod(shape)))
    lam = 610.0
    _log(f"[gpu] serving grid {shape}")
    box = O.OracleFullBox(shape, 15.0, lam, rho, n_k=k, n_e0=e, workers=16)
    req = native.Request(shape=shape, d=15.0, lam=lam, rho=rho, n_k=k, n_e0=e)
    p64 = native.Problem(req, precision="fp64")
    p32 = native.Problem(req, precision="fp32_refined")
    x = rng.standard_normal(3 * box.N) + 1j * rng.standard_normal(3 * box.N)
    assert _rel(native.apply([p64], [x])[0], box.apply(x)) < 1e-11
    assert _rel(native.apply([p32], [x], c64=True)[0], box.apply(x)) < 2e-6
    i64 = p64.solve(tol=1e-10)
    assert i64["status"] == "converged"
    assert _true_residual(box, p64.polarization()) < 1.1e-10
    e_ref = p64.field()
    i32 = p32.solve(tol=1e-5)
    err = _rel(p32.field(), e_ref)
    r = _true_residual(box, p32.polarization())
    _log(f"[gpu] {shape}: fp64 tight {i64['iterations']} its; refined {i32['iterations']} its, true residual "
         f"{r:.3e}, E err {err:.3e} (estimate {i32['err_est']:.3e})")
    assert i32["status"] == "converged" and r <= 1.0001e-5
    assert err <= 1e-5


# ---------------------------------------------------------------------------
# adjoint gradient of the target objective through the cone filter
# ---------------------------------------------------------------------------
def test_target_gradient_48x48x24(native, O):
    import torch
    from dda_model.autograd import cone_filter, dda_objective, target_block
    shape = (48, 48, 24)
    nx, ny, nz = shape
    rng = np.random.default_rng(3)
    raw = torch.tensor(rng.uniform(0, 1, (nz, ny, nx)), dtype=torch.float64, device="cuda", requires_grad=True)
    tg = np.asarray(target_block(shape, 12.0, size=4, height_nm=60.0))
    assert tg.shape == (64, 3)
    req = native.Request(shape=shape, d=12.0, lam=542.0, rho=np.full(nx * ny * nz, 0.5))
    probs = {p: native.Problem(req, precision=p) for p in ("fp64", "fp32_refined")}

    def f(x, prec="fp64", tol=1e-11):
        rho = cone_filter(x, 2.0).reshape(-1)
        return dda_objective(rho, probs[prec], kind="targets", targets=tg, tol=tol)

    _log("[gpu] target gradient 48x48x24: fp64")
    val = f(raw)
    val.backward()
    grad = raw.grad.detach().cpu().numpy().copy()
    # the objective on the device against the oracle's field evaluation of the same P
    p = probs["fp64"]
    box = O.OracleFullBox(shape, 12.0, 542.0, cone_filter(raw.detach(), 2.0).reshape(-1).cpu().numpy(), workers=16)
    box.P = p.polarization()
    want = float(np.mean(np.sum(np.abs(box.target_fields_ldr(tg)) ** 2, axis=1)))
    assert val.item() == pytest.approx(want, rel=1e-10)
    assert _true_residual(box, box.P) < 1e-10
    # central differences on 32 voxels: 24 in the top four layers under the
    # targets (where the objective is sensitive) and 8 anywhere.  A difference
    # quotient carries the objective's solve error over h: with solves to
    # 1e-11 that is about 1e-6 |f| absolute, the floor of the comparison.
    h = 1e-4
    zz, yy, xx = np.meshgrid(np.arange(nz - 4, nz), np.arange(ny // 2 - 8, ny // 2 + 8),
                             np.arange(nx // 2 - 8, nx // 2 + 8), indexing="ij")
    near = (zz * ny * nx + yy * nx + xx).reshape(-1)
    flat = np.concatenate([rng.choice(near, 24, replace=False), rng.choice(nx * ny * nz, 8, replace=False)])
    gmax = np.abs(grad).max()
    floor = max(1e-5 * gmax, 1e-6 * abs(val.item()))
    for i in flat:
        idx = np.unravel_index(i, (nz, ny, nx))
        with torch.no_grad():
            xp = raw.detach().clone()
            xp[idx] += h
            xm = raw.detach().clone()
            xm[idx] -= h
            fd = (f(xp).item() - f(xm).item()) / (2 * h)
        assert grad[idx] == pytest.approx(fd, rel=1e-4, abs=floor), (idx, grad[idx], fd)
    # the refined gradient (forward and adjoint solves within 1e-5 in field)
    _log("[gpu] target gradient 48x48x24: refined")
    raw.grad = None
    v32 = f(raw, "fp32_refined", 1e-5)
    v32.backward()
    g32 = raw.grad.detach().cpu().numpy()
    err = _rel(g32, grad)
    _log(f"[gpu] objective fp64 {val.item():.10e} refined {v32.item():.10e}; gradient rel err {err:.3e}")
    assert v32.item() == pytest.approx(val.item(), rel=5e-5)
    assert err < 1e-4


# ---------------------------------------------------------------------------
# multi-RHS at 128x128x64 (FFT lengths 256)
# ---------------------------------------------------------------------------
def test_multi_rhs_128x128x64(native, O):
    shape = (128, 128, 64)
    rho = smooth_density(shape, 3.0, seed=5)
    lams = (450.0, 550.0, 650.0, 750.0)
    pols = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    reqs = [native.Request(shape=shape, d=8.0, lam=l, rho=rho, n_k=(0.0, 0.0, 1.0), n_e0=e) for l in lams for e in pols]
    _log("[gpu] multi-RHS 128x128x64: oracle operator at 550 nm")
    box = O.OracleFullBox(shape, 8.0, 550.0, rho, n_k=(0.0, 0.0, 1.0), n_e0=(1.0, 0.0, 0.0), workers=16)
    p64 = native.Problem(reqs[2], precision="fp64")
    rng = np.random.default_rng(4)
    x = rng.standard_normal(3 * box.N) + 1j * rng.standard_normal(3 * box.N)
    assert _rel(native.apply([p64], [x])[0], box.apply(x)) < 1e-11
    del p64
    _log("[gpu] multi-RHS 128x128x64: batched refined solve of 8 right-hand sides")
    probs = [native.Problem(r, precision="fp32_refined") for r in reqs]
    infos = native.solve(probs, tol=1e-5)
    _log(f"[gpu] multi-RHS: {[(i['status'], i['iterations'], round(i['err_est'], 8)) for i in infos]}")
    for i in infos:
        assert i["status"] == "converged" and i["error"] <= 1e-5 and 2.0 * i["err_est"] <= 1e-5
    r = _true_residual(box, probs[2].polarization())
    _log(f"[gpu] multi-RHS: oracle true residual of the 550 nm x-polarised solve {r:.3e}")
    assert r <= 1.0001e-5
