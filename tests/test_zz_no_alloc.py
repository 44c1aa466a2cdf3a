"""The no-allocation operator apply (tdda_problem_apply_noalloc) through the
C ABI against the allocating batched apply (tdda_problems_apply) on the same
problems: two calls in a row, the second on a smaller problem, each equal to
the allocating apply (same kernels and tiles, so bitwise; the bar is 1e-15
relative) and to the oracle's FP64 operator to 1e-11."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_noalloc_apply_matches_allocating_apply():
    import torch
    import dda_oracle as O
    from dda_model import native
    lib = native.lib()
    lib.tdda_problem_apply_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(31)
    for shape, prec in (((24, 20, 16), "fp32_refined"), ((10, 7, 5), "fp64")):
        _log(f"[gpu] no-alloc apply {shape} {prec}")
        req = native.Request(shape=shape, d=15.0, lam=560.0, rho=rng.uniform(0, 1, int(np.prod(shape))),
                             n_k=(0.3, 0.0, 0.9539392014169456), n_e0=(0.0, 1.0, 0.0))
        p = native.Problem(req, precision=prec)
        n3 = 3 * p.n
        x = rng.standard_normal(n3) + 1j * rng.standard_normal(n3)
        want = native.apply([p], [x], c64=False)[0]
        xd = torch.tensor(x, dtype=torch.complex128, device="cuda")
        yd = torch.full((n3,), float("nan"), dtype=torch.complex128, device="cuda")
        torch.cuda.synchronize()
        native.check(lib.tdda_problem_apply_noalloc(p.h, ctypes.c_void_p(xd.data_ptr()),
                                                    ctypes.c_void_p(yd.data_ptr())))
        torch.cuda.synchronize()
        got = yd.cpu().numpy()
        assert np.all(np.isfinite(got))
        assert _rel(got, want) <= 1e-15
        box = O.OracleFullBox(shape, 15.0, 560.0, req.rho, n_k=req.n_k, n_e0=req.n_e0)
        assert _rel(got, box.apply(x)) < 1e-11


def test_noalloc_apply_refuses_oversized_problem():
    """Pass maps beyond the fixed 256 KiB scratch are refused with a clear
    error and nothing runs (128 x 128 x 64 needs about 241 KiB and fits;
    128 x 128 x 100 needs about 378 KiB)."""
    import torch
    from dda_model import native
    lib = native.lib()
    lib.tdda_problem_apply_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    shape = (128, 128, 100)
    req = native.Request(shape=shape, d=15.0, lam=560.0, rho=np.full(int(np.prod(shape)), 0.5))
    p = native.Problem(req, precision="fp64")
    n3 = 3 * p.n
    xd = torch.zeros(n3, dtype=torch.complex128, device="cuda")
    yd = torch.zeros(n3, dtype=torch.complex128, device="cuda")
    rc = lib.tdda_problem_apply_noalloc(p.h, ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(yd.data_ptr()))
    msg = lib.tdda_last_error().decode()
    _log(f"[gpu] oversized: rc {rc} {msg}")
    assert rc == 4 and "no-allocation apply holds" in msg
