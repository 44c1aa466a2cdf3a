"""ctypes binding of the batched-problem and server entry points of
include/topodda.h (the parts with no reference counterpart).

Everything here calls the in-tree libtopodda.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_LIB = None
PREC = {"fp64": 0, "fp32_refined": 1}
PRECOND = {"none": 0, "jacobi": 1}
STATUS = {0: "active", 1: "converged", 2: "breakdown", 3: "maxit", 4: "idle"}
OBJ = {"integrated": 0, "targets": 1}


class ProblemDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("d", ctypes.c_double), ("lam", ctypes.c_double), ("nback", ctypes.c_double),
                ("method", ctypes.c_int), ("maxm", ctypes.c_int), ("maxn", ctypes.c_int),
                ("lm", ctypes.c_double), ("ln", ctypes.c_double),
                ("n_k", ctypes.c_double * 3), ("n_e0", ctypes.c_double * 3), ("e0", ctypes.c_double),
                ("material", ctypes.c_double * 4)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int), ("error", ctypes.c_double),
                ("reliable_updates", ctypes.c_int), ("ms", ctypes.c_double), ("err_est", ctypes.c_double)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("maxit", ctypes.c_int), ("tol", ctypes.c_double),
                ("ru_delta", ctypes.c_double), ("hbm_budget_gb", ctypes.c_double), ("max_slots", ctypes.c_int),
                ("chunk", ctypes.c_int), ("keep_fields", ctypes.c_int), ("err_tol", ctypes.c_double),
                ("precond", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_long), ("status", ctypes.c_int), ("iterations", ctypes.c_int),
                ("reliable_updates", ctypes.c_int), ("error", ctypes.c_double), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double), ("operator_cached", ctypes.c_int),
                ("err_est", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_long), ("completed", ctypes.c_long), ("chunks", ctypes.c_long),
                ("rebuilds", ctypes.c_long), ("launches", ctypes.c_long), ("max_active", ctypes.c_int),
                ("peak_bytes", ctypes.c_long), ("host_admit_ms", ctypes.c_double), ("host_wait_ms", ctypes.c_double),
                ("host_retire_ms", ctypes.c_double), ("host_build_ms", ctypes.c_double),
                ("host_enqueue_ms", ctypes.c_double), ("graph_chunks", ctypes.c_long)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtopodda.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        L.tdda_last_error.restype = ctypes.c_char_p
        L.tdda_problem_create.argtypes = [ctypes.POINTER(ProblemDesc), ctypes.c_int, dp, dp, ctypes.c_long,
                                          ctypes.c_double, ctypes.POINTER(vp)]
        L.tdda_problems_create.argtypes = [ctypes.POINTER(ProblemDesc), ctypes.c_int, ctypes.c_int, dp, dp,
                                           ctypes.c_long, ctypes.c_double, ctypes.POINTER(vp)]
        L.tdda_problem_destroy.argtypes = [vp]
        L.tdda_problem_set_density.argtypes = [vp, ctypes.c_void_p, ctypes.c_int]
        L.tdda_problems_solve.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.POINTER(SolveInfo)]
        L.tdda_problem_field.argtypes = [vp, ctypes.c_void_p, ctypes.c_int]
        L.tdda_problem_grid.argtypes = [vp, ctypes.POINTER(ctypes.c_int), dp, ctypes.POINTER(ctypes.c_int)]
        L.tdda_problem_set_preconditioner.argtypes = [vp, ctypes.c_int]
        L.tdda_problem_polarization.argtypes = [vp, ctypes.c_void_p, ctypes.c_int]
        L.tdda_problem_objective.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, dp]
        L.tdda_problem_gradient.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_double, dp, ctypes.c_int, dp, ctypes.POINTER(SolveInfo)]
        L.tdda_problems_time_passes.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.tdda_problems_time_iteration.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, dp]
        L.tdda_problems_apply.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, dp, dp]
        L.tdda_server_create.argtypes = [ctypes.POINTER(ServerOpts), dp, dp, ctypes.c_long, ctypes.c_double,
                                         ctypes.POINTER(vp)]
        L.tdda_server_destroy.argtypes = [vp]
        L.tdda_server_submit.argtypes = [vp, ctypes.c_long, ctypes.POINTER(ProblemDesc), dp, ctypes.c_long]
        L.tdda_server_poll.argtypes = [vp, ctypes.POINTER(Result), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.tdda_server_take_field.argtypes = [vp, ctypes.c_long, dp, ctypes.c_long]
        L.tdda_server_idle.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.tdda_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.tdda_set_graph_capture.argtypes = [ctypes.c_int]
        L.tdda_set_fused_yz.argtypes = [ctypes.c_int]
        L.tdda_set_plane_passes.argtypes = [ctypes.c_int]
        L.tdda_schedule_text.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
        L.tdda_schedule_parse.argtypes = [ctypes.c_char_p]
        L.tdda_schedule_load.argtypes = [ctypes.c_char_p]
        L.tdda_schedule_save.argtypes = [ctypes.c_char_p]
        L.tdda_tune.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
        L.tdda_device_bytes.restype = ctypes.c_long
        L.tdda_trim_caches.restype = ctypes.c_int
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        msg = lib().tdda_last_error().decode()
        if rc == 2:
            raise ValueError(msg)
        raise RuntimeError(msg or f"topodda error {rc}")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Request:
    """One DDA solve: a full box with a density per voxel (x fastest)."""
    shape: tuple
    d: float
    lam: float
    rho: np.ndarray
    n_k: tuple = (0.0, 0.0, 1.0)
    n_e0: tuple = (1.0, 0.0, 0.0)
    e0: float = 1.0
    material: tuple = (1.0 + 0j, 5.96 + 3.89j)
    nback: float = 1.0
    method: str = "LDR"
    images: int = 0
    period: tuple = (0.0, 0.0)

    def desc(self):
        nx, ny, nz = self.shape
        m = [complex(self.material[0]), complex(self.material[1])]
        return ProblemDesc(nx, ny, nz, self.d, self.lam, self.nback, 0 if self.method == "FCD" else 1,
                           self.images, self.images, self.period[0], self.period[1],
                           (ctypes.c_double * 3)(*self.n_k), (ctypes.c_double * 3)(*self.n_e0), self.e0,
                           (ctypes.c_double * 4)(m[0].real, m[0].imag, m[1].real, m[1].imag))


def _tables(tables):
    if tables is None:
        return None, None, 0, 0.1
    si, ci, delta = tables
    si = np.ascontiguousarray(si, dtype=np.float64)
    ci = np.ascontiguousarray(ci, dtype=np.float64)
    return si, ci, si.size, delta


class Problem:
    """A device-resident problem (operator from the device cache)."""

    def __init__(self, req: Request, precision="fp32_refined", tables=None, precond="none"):
        """precond: "none" (the reference's plain BiCGSTAB) or "jacobi" (right
        Jacobi preconditioning; the residual and the stopping rule are unchanged)."""
        si, ci, n, delta = _tables(tables)
        self._keep = (si, ci)
        h = ctypes.c_void_p()
        check(lib().tdda_problem_create(ctypes.byref(req.desc()), PREC[precision],
                                        _dp(si) if si is not None else None,
                                        _dp(ci) if ci is not None else None, n, delta, ctypes.byref(h)))
        self.h = h
        self.n = int(np.prod(req.shape))
        check(lib().tdda_problem_set_preconditioner(self.h, PRECOND[precond]))
        self.set_density(req.rho)

    @classmethod
    def create_many(cls, reqs, precision="fp32_refined", tables=None, precond="none"):
        """Several problems at once: the Green's set-ups of their new operators
        run together, one launch per step for all of them."""
        si, ci, n, delta = _tables(tables)
        k = len(reqs)
        descs = (ProblemDesc * max(k, 1))(*[r.desc() for r in reqs])
        hs = (ctypes.c_void_p * max(k, 1))()
        check(lib().tdda_problems_create(descs, k, PREC[precision], _dp(si) if si is not None else None,
                                         _dp(ci) if ci is not None else None, n, delta, hs))
        out = []
        for r, h in zip(reqs, hs[:k]):
            p = cls.__new__(cls)
            p._keep = (si, ci)
            p.h = ctypes.c_void_p(h)
            p.n = int(np.prod(r.shape))
            out.append(p)
        for r, p in zip(reqs, out):
            check(lib().tdda_problem_set_preconditioner(p.h, PRECOND[precond]))
            p.set_density(r.rho)
        return out

    def set_density(self, rho):
        """rho: numpy array (host) or a torch CUDA tensor of float64 (device)."""
        if hasattr(rho, "data_ptr"):
            check(lib().tdda_problem_set_density(self.h, ctypes.c_void_p(rho.data_ptr()), 1))
        else:
            r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
            assert r.size == self.n
            check(lib().tdda_problem_set_density(self.h, r.ctypes.data_as(ctypes.c_void_p), 0))

    def grid(self):
        """(FFT grid (Lx, Ly, Lz), operator set-up ms, operator taken from the cache)"""
        lxyz = (ctypes.c_int * 3)()
        ms = ctypes.c_double(0.0)
        cached = ctypes.c_int(0)
        check(lib().tdda_problem_grid(self.h, lxyz, ctypes.byref(ms), ctypes.byref(cached)))
        return tuple(lxyz), ms.value, bool(cached.value)

    def operator_setup_ms(self):
        return self.grid()[1]

    def field(self):
        out = np.empty(3 * self.n, dtype=np.complex128)
        check(lib().tdda_problem_field(self.h, out.ctypes.data_as(ctypes.c_void_p), 0))
        return out

    def polarization(self):
        out = np.empty(3 * self.n, dtype=np.complex128)
        check(lib().tdda_problem_polarization(self.h, out.ctypes.data_as(ctypes.c_void_p), 0))
        return out

    def solve(self, maxit=100000, tol=1e-5):
        return solve([self], maxit, tol)[0]

    @staticmethod
    def _targets(targets):
        if targets is None:
            return None, 0
        t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
        return t, t.shape[0]

    def objective(self, kind="integrated", power=2, targets=None, with_fields=False):
        """Objective of the current solution: "integrated" = log sum |al P|^power,
        "targets" = mean |E|^2 over external points (nm)."""
        t, nt = self._targets(targets)
        val = ctypes.c_double(0.0)
        fields = np.empty(3 * nt, dtype=np.complex128) if (with_fields and nt) else None
        check(lib().tdda_problem_objective(self.h, OBJ[kind], power, _dp(t) if t is not None else None, nt,
                                           ctypes.byref(val),
                                           _dp(fields.view(np.float64)) if fields is not None else None))
        return (val.value, fields) if with_fields else val.value

    def gradient(self, kind="integrated", power=2, targets=None, maxit=100000, tol=1e-5, out=None):
        """d objective / d rho per voxel by one adjoint solve.  out: numpy array
        (host) or torch CUDA float64 tensor (device); returns (grad, value, info)."""
        t, nt = self._targets(targets)
        val = ctypes.c_double(0.0)
        info = SolveInfo()
        if out is None:
            out = np.empty(self.n, dtype=np.float64)
        if hasattr(out, "data_ptr"):
            ptr, dev = ctypes.cast(ctypes.c_void_p(out.data_ptr()), ctypes.POINTER(ctypes.c_double)), 1
        else:
            ptr, dev = _dp(out), 0
        check(lib().tdda_problem_gradient(self.h, OBJ[kind], power, _dp(t) if t is not None else None, nt, maxit,
                                          tol, ptr, dev, ctypes.byref(val), ctypes.byref(info)))
        return out, val.value, dict(status=STATUS[info.status], iterations=info.iterations, error=info.error,
                                    err_est=info.err_est)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.tdda_problem_destroy(self.h)
            self.h = None


def solve(problems, maxit=100000, tol=1e-5):
    """Solve several problems in one batch; returns a list of info dicts."""
    arr = (ctypes.c_void_p * len(problems))(*[p.h for p in problems])
    infos = (SolveInfo * len(problems))()
    check(lib().tdda_problems_solve(arr, len(problems), maxit, tol, infos))
    return [dict(status=STATUS[i.status], iterations=i.iterations, error=i.error,
                 reliable_updates=i.reliable_updates, ms=i.ms, err_est=i.err_est) for i in infos]


def apply(problems, xs, c64=False):
    """y = A x + al x for every problem in one batched launch per pass, with
    c64 transforms (refined problems) or FP64 ones; xs: one complex vector of
    3 N entries per problem.  Returns the list of y."""
    arr = (ctypes.c_void_p * len(problems))(*[p.h for p in problems])
    x = np.ascontiguousarray(np.concatenate([np.asarray(v, np.complex128).reshape(-1) for v in xs]))
    y = np.empty_like(x)
    check(lib().tdda_problems_apply(arr, len(problems), int(c64), _dp(x.view(np.float64)), _dp(y.view(np.float64))))
    out, k = [], 0
    for p in problems:
        out.append(y[k:k + 3 * p.n].copy())
        k += 3 * p.n
    return out


def trim_caches():
    """Drop cached operators nobody holds and release cached device blocks."""
    check(lib().tdda_trim_caches())


def set_graph_capture(on: bool):
    """Record batched iterations as HIP graphs and replay them (default), or
    issue the launches one by one."""
    check(lib().tdda_set_graph_capture(int(bool(on))))


FUSED = {"off": 0, "schedule": 1, "all": 2}


def set_fused_yz(mode):
    """Which eligible boxes run the c64 y, z and inverse y passes as one
    kernel per x frequency: "off", "schedule" (as the kernel schedule says;
    the default) or "all"; applies to batches built afterwards."""
    check(lib().tdda_set_fused_yz(FUSED[mode]))


def set_plane_passes(on: bool):
    """Run the c64 x and y transforms of eligible boxes (Lx = Ly in {32, 48,
    64, 96, 128}) as one plane kernel per direction (default) or as the
    separate X and Y passes; applies to batches built afterwards."""
    check(lib().tdda_set_plane_passes(int(bool(on))))


def _text(call):
    n = ctypes.c_long(0)
    check(call(None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    check(call(buf, n.value + 1, ctypes.byref(n)))
    return buf.value.decode()


def schedule_text():
    """The kernel schedule, one entry per line."""
    return _text(lib().tdda_schedule_text)


def schedule_set(text: str):
    check(lib().tdda_schedule_parse(text.encode()))


def schedule_load(path):
    check(lib().tdda_schedule_load(str(path).encode()))


def schedule_save(path):
    check(lib().tdda_schedule_save(str(path).encode()))


def schedule_clear():
    check(lib().tdda_schedule_clear())


def tune(boxes, count=4, reps=10):
    """Autotune the kernel schedule for these (nx, ny, nz) boxes; returns the
    decisions as dicts (pass timings in ms summed over the boxes)."""
    b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 3))
    arr = b.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    n = ctypes.c_long(0)
    cap = 1 << 16
    buf = ctypes.create_string_buffer(cap)
    check(lib().tdda_tune(arr, b.shape[0], int(count), int(reps), buf, cap, ctypes.byref(n)))
    out = []
    for line in buf.value.decode().splitlines():
        f = line.split()
        if f[0] == "pass":
            out.append({"kind": "pass", "precision": f[1], "pass": f[2], "L": int(f[3]), "ms": float(f[5]),
                        "ms_alt": None if f[6] == "-" else float(f[6]), "chosen": int(f[8])})
        elif f[0] == "xy":
            out.append({"kind": "xy", "L": int(f[1]), "ms_off": float(f[3]), "ms_on": float(f[4]), "chosen": int(f[6])})
        else:
            out.append({"kind": "yz", "Ly": int(f[1]), "Lz": int(f[2]), "ms_three": float(f[4]),
                        "ms_fused": float(f[5]), "chosen": int(f[7])})
    return out


def time_passes(problems, c64=True, reps=20):
    """Average ms per batched launch of (X fwd, Y fwd, Z, Y inv, X inv)."""
    arr = (ctypes.c_void_p * len(problems))(*[p.h for p in problems])
    ms = np.zeros(5)
    check(lib().tdda_problems_time_passes(arr, len(problems), int(c64), reps, _dp(ms)))
    return ms


def time_iteration(problems, reps=20):
    """Average ms per batched launch of the ten passes of one refined
    BiCGSTAB step (p-update matvec, then t matvec); overwrites solver state."""
    arr = (ctypes.c_void_p * len(problems))(*[p.h for p in problems])
    ms = np.zeros(10)
    check(lib().tdda_problems_time_iteration(arr, len(problems), reps, _dp(ms)))
    return ms


class Server:
    """Continuous-batching request server (C++ loop thread)."""

    def __init__(self, precision="fp32_refined", maxit=100000, tol=1e-5, ru_delta=0.01, hbm_budget_gb=200.0,
                 max_slots=64, chunk=4, keep_fields=True, tables=None, err_tol=0.0, precond="none"):
        """err_tol: relative field-error target of refined solves (0: tol;
        negative: stop on the true residual alone).  precond: "none" or
        "jacobi" for every solve."""
        si, ci, n, delta = _tables(tables)
        self._keep = (si, ci)
        o = ServerOpts(PREC[precision], maxit, tol, ru_delta, hbm_budget_gb, max_slots, chunk, int(keep_fields),
                       err_tol, PRECOND[precond])
        h = ctypes.c_void_p()
        check(lib().tdda_server_create(ctypes.byref(o), _dp(si) if si is not None else None,
                                       _dp(ci) if ci is not None else None, n, delta, ctypes.byref(h)))
        self.h = h
        self.sizes = {}

    def submit(self, rid, req: Request):
        rho = np.ascontiguousarray(req.rho, dtype=np.float64).reshape(-1)
        self.sizes[rid] = rho.size
        check(lib().tdda_server_submit(self.h, rid, ctypes.byref(req.desc()), _dp(rho), rho.size))

    def poll(self, max_results=256):
        buf = (Result * max_results)()
        n = ctypes.c_int(0)
        check(lib().tdda_server_poll(self.h, buf, max_results, ctypes.byref(n)))
        return [dict(id=r.id, status=STATUS[r.status], iterations=r.iterations,
                     reliable_updates=r.reliable_updates, error=r.error, t_submit=r.t_submit,
                     t_admit=r.t_admit, t_done=r.t_done, operator_cached=bool(r.operator_cached),
                     err_est=r.err_est)
                for r in buf[:n.value]]

    def field(self, rid):
        out = np.empty(3 * self.sizes[rid], dtype=np.complex128)
        check(lib().tdda_server_take_field(self.h, rid, _dp(out.view(np.float64)), out.size))
        return out

    def idle(self):
        v = ctypes.c_int(0)
        check(lib().tdda_server_idle(self.h, ctypes.byref(v)))
        return bool(v.value)

    def stats(self):
        s = ServerStats()
        check(lib().tdda_server_get_stats(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            check(lib().tdda_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
