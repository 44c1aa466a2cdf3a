"""GPU tests of the design side: the density parameterisation as torch ops
through the device solve, the design loop on the GPU model, and a request
with a dispersive material.

Bars: autograd through map -> cone filter -> beta projection -> adjoint
solve against central differences (h = 1e-5, FP64 solves to 1e-12) to 1e-4
per parameter; each op alone against central differences to 1e-8; the design
loop's objective history and parameters on the GPU model equal those of the
same loop on the oracle model to 1e-7; a SiO2 structure (Sellmeier
permittivity) solved on the device equals the oracle's field to 1e-8.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fd(f, x, i, h):
    xp = x.detach().clone()
    xm = x.detach().clone()
    xp[i] += h
    xm[i] -= h
    return (f(xp).item() - f(xm).item()) / (2 * h)


def test_density_ops_alone_match_central_differences():
    import torch
    import dda_oracle as O
    from dda_model import density
    _log("[gpu] density ops alone")
    nx, ny, nz = 10, 10, 2
    geo = O.generate_geometry(nx, ny, nz)
    pipe = density.DensityPipeline(geo, nz, "4fold", (4.5, 4.5), radius=2.2, periodic=True, lm=nx, ln=ny)
    rng = np.random.default_rng(5)
    x = torch.tensor(rng.uniform(0.1, 0.9, pipe.nfpara), dtype=torch.float64, device="cuda", requires_grad=True)
    wts = torch.tensor(rng.standard_normal(nx * ny * nz), dtype=torch.float64, device="cuda")
    ops = {
        "map": lambda v: (density.params_to_dipoles(v, pipe.gpara) * wts).sum(),
        "filter": lambda v: (density.cone_filter_params(v, pipe.weights) ** 3).sum(),
        "projection": lambda v: (density.smooth_density(v, 9.0, 0.5) * wts[: v.numel()]).sum(),
        "pipeline": lambda v: (pipe(v, beta=9.0) * wts).sum(),
    }
    for name, f in ops.items():
        x.grad = None
        f(x).backward()
        g = x.grad.detach().cpu().numpy()
        for i in (0, pipe.nfpara // 2, pipe.nfpara - 1):
            assert g[i] == pytest.approx(_fd(f, x, i, 1e-6), rel=1e-8, abs=1e-9), (name, i)


def test_pipeline_gradient_through_solve():
    import torch
    import dda_oracle as O
    from dda_model import density, native
    from dda_model.autograd import dda_objective, target_block
    _log("[gpu] pipeline gradient through the solve")
    shape = (12, 12, 4)
    nx, ny, nz = shape
    geo = O.generate_geometry(nx, ny, nz)
    pipe = density.DensityPipeline(geo, nz, "4fold", (5.5, 5.5), radius=2.0)
    rng = np.random.default_rng(6)
    x = torch.tensor(rng.uniform(0.2, 0.8, pipe.nfpara), dtype=torch.float64, device="cuda", requires_grad=True)
    req = native.Request(shape=shape, d=12.0, lam=542.0, rho=np.full(nx * ny * nz, 0.5))
    p = native.Problem(req, precision="fp64", precond="jacobi")
    tg = np.asarray(target_block(shape, 12.0, size=2, height_nm=36.0))

    def f(v):
        return dda_objective(pipe(v, beta=4.0), p, kind="targets", targets=tg, tol=1e-12)

    f(x).backward()
    g = x.grad.detach().cpu().numpy()
    gmax = np.abs(g).max()
    for i in rng.choice(pipe.nfpara, 6, replace=False):
        fd = _fd(f, x, i, 1e-5)
        assert g[i] == pytest.approx(fd, rel=1e-4, abs=1e-6 * gmax), (i, g[i], fd)


class _OracleWrapper:
    """DDAModelWrapper's interface over the oracle model (tight solves)."""

    def __init__(self, args, tol):
        import dda_oracle as O
        self.m = O.OracleDDAModel(*args)
        self.n = len(args[12]) // 3
        self.tol = tol

    def objective(self):
        self.m.solveElectricField(np.zeros(3 * self.n, complex), 100000, self.tol)
        return self.m.calculateObjective()

    def gradients(self, obj):
        return self.m.calculateGradients(1e-3, obj, 100000, self.tol)

    @property
    def parameters(self):
        return self.m.getParameters()

    @parameters.setter
    def parameters(self, v):
        self.m.setParameters(v)

    def allParameters(self):
        return self.m.getDielectrics()

    def getElectricField(self):
        return self.m.getElectricField()


class _GpuWrapper(_OracleWrapper):
    def __init__(self, args, tol):
        import dda_model
        self.m = dda_model.DDAModel(*args)
        self.n = len(args[12]) // 3
        self.tol = tol


def test_design_loop_on_gpu_matches_oracle(tmp_path, tables):
    import dda_oracle as O
    from dda_model import design
    _log("[gpu] design loop")
    shape = (8, 8, 3)
    n = int(np.prod(shape))
    rng = np.random.default_rng(7)
    diel = np.repeat(rng.uniform(0.2, 0.8, n), 3)
    args = O.wrapper_args(list(shape), 15.0, diel, [1.01 + 0j, 5.96282 + 3.80423e-7j], [0, 0, 1], [1, 0, 0], 542.0,
                          symmetry_axes=[3.5, 3.5], symmetry_method="4fold", apc_max_m=2, apc_max_n=2, tables=tables)
    cfg = {"slope": 1.0}
    ccfg = {"iter_start": 0, "iter_end": 10, "coeff_min": 0.0, "coeff_max": 0.5}
    hist = {}
    for name, cls in (("gpu", _GpuWrapper), ("oracle", _OracleWrapper)):
        m = cls(args, 1e-11)
        hist[name] = design.run_design(m, 3, 0.01, "triangular", cfg, "linear", ccfg, tmp_path / name)
        hist[name + "_p"] = np.loadtxt(tmp_path / name / "Parameters" / "Param2.txt")
    _log(f"[gpu] objectives gpu {hist['gpu']} oracle {hist['oracle']}")
    assert np.allclose(hist["gpu"], hist["oracle"], rtol=1e-7)
    assert np.allclose(hist["gpu_p"], hist["oracle_p"], rtol=1e-7, atol=1e-9)
    for f in ("Obj.txt", "Coeffs.txt", "Average_Penalties.txt", "stepsizes.txt", "Penalty_Shape.txt",
              "Structures/Structure2.txt", "E-Fields/E-Field2.txt"):
        assert (tmp_path / "gpu" / f).exists(), f


def test_dispersive_material_request(tables):
    import dda_oracle as O
    from dda_model import materials, native
    _log("[gpu] SiO2 structure at 633 nm")
    eps = materials.permittivity("SiO2", 633.0)
    rng = np.random.default_rng(8)
    shape = (10, 9, 6)
    req = native.Request(shape=shape, d=15.0, lam=633.0, rho=rng.uniform(0, 1, int(np.prod(shape))),
                         material=(1.0 + 0j, eps))
    p = native.Problem(req, precision="fp64")
    assert p.solve(tol=1e-12)["status"] == "converged"
    box = O.OracleFullBox(shape, 15.0, 633.0, req.rho, material=(1.0 + 0j, eps))
    box.solve(tol=1e-12)
    assert np.linalg.norm(p.field() - box.field()) / np.linalg.norm(box.field()) < 1e-8
