"""Benchmark of the DDA solve path on one MI355X; prints ONE JSON line.

Workload (headline): a stream of R DDA requests (default 512, seeds 1000..),
each independently drawn: box uniform over {16^3, 24^3, 32x32x16, 48x48x24,
64x64x32} with d = 15 nm, wavelength uniform in [400, 800] nm, polar angle
uniform in [0, 30 deg], azimuth and polarisation angle uniform, density
uniform in [0, 1) per voxel, LDR interaction, non-periodic.  Every request is
solved to a 1e-5 relative residual (true FP64 residual) by the continuous-
batching server in complex64 with FP64 reliable updates.

A step is one pass over R requests of its own (fresh seeds every step, so
every request pays its Green's set-up) with all of them queued at once (the
server's saturation throughput); `value` is solves/s over the K timed steps.
Another R requests then arrive as a Poisson stream at 0.8x that rate, which
gives the p50/p99 latency.  Each refined solve stops when its FP64 true
residual meets 1e-5 and its estimated field error meets 1e-5; the fields of
a subset are compared with a tight FP64 solve.  The plain FP64 single-solve
GPU path (one request at a time, FP64 transforms and Krylov vectors, the
reference's stopping rule) is timed on the same subset for the speedup, at
tol 1e-5 and at the tolerance that gives the same field accuracy.  Beside
them: the 128x128x64 eight-RHS solve, the FP64 lattice-sum rate, a device
copy, and the FP64 CPU oracle on a bounded sample (1 and 16 threads).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "topo-dda_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
SHAPES = [(16, 16, 16), (24, 24, 24), (32, 32, 16), (48, 48, 24), (64, 64, 32)]


def fft_len(n):
    L = max(2, 2 * n - 1)
    while True:
        m = L
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return L
        L += 1


def make_requests(native, count, seed0=1000):
    reqs = []
    for i in range(count):
        rng = np.random.default_rng(seed0 + i)
        shape = SHAPES[int(rng.integers(len(SHAPES)))]
        lam = float(rng.uniform(400.0, 800.0))
        th = float(rng.uniform(0.0, math.radians(30.0)))
        ph = float(rng.uniform(0.0, 2 * math.pi))
        ps = float(rng.uniform(0.0, 2 * math.pi))
        k = (math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th))
        e1 = (math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th))
        e2 = (-math.sin(ph), math.cos(ph), 0.0)
        e = tuple(math.cos(ps) * a + math.sin(ps) * b for a, b in zip(e1, e2))
        rho = rng.uniform(0.0, 1.0, int(np.prod(shape)))
        reqs.append(native.Request(shape=shape, d=15.0, lam=lam, rho=rho, n_k=k, n_e0=e, method="LDR"))
    return reqs


def percentile(v, q):
    s = sorted(v)
    if not s:
        return float("nan")
    k = (len(s) - 1) * q
    lo, hi = math.floor(k), math.ceil(k)
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


def run_stream(native, reqs, precision, max_slots, chunk, arrivals=None, fields=False, timeout=600.0, tol=1e-5,
               err_tol=0.0, precond="none", server=None, id0=0, ru_delta=0.01):
    """Submit every request (at its arrival time if given) and wait for all.
    A given server is used and kept open (request ids id0, id0 + 1, ...);
    otherwise one is made for the stream and closed.  Returns (wall seconds
    from first submission to last completion, results by request index,
    server statistics, fields or None)."""
    srv = server or native.Server(precision=precision, max_slots=max_slots, chunk=chunk, keep_fields=fields,
                                  tol=tol, err_tol=err_tol, precond=precond, ru_delta=ru_delta)
    results = {}
    t0 = time.perf_counter()
    nxt = 0
    deadline = t0 + timeout
    while len(results) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and (arrivals is None or arrivals[nxt] <= now):
            srv.submit(id0 + nxt, reqs[nxt])
            nxt += 1
        for r in srv.poll():
            results[r["id"] - id0] = r
        if time.perf_counter() > deadline:
            raise RuntimeError(f"stream did not finish in {timeout} s ({len(results)}/{len(reqs)})")
        time.sleep(0.0005)
    wall = time.perf_counter() - t0
    out_fields = {i: srv.field(id0 + i) for i in results} if fields else None
    stats = srv.stats()
    if server is None:
        srv.close()
    return wall, results, stats, out_fields


def host_cpu():
    """nproc and the lscpu model name of this host."""
    import subprocess
    info = {"nproc": os.cpu_count(), "model": None}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout.strip())
    except (OSError, ValueError, subprocess.CalledProcessError):
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.CalledProcessError):
        pass
    return info


def cpu_baseline(reqs, iterations, workers=1, per_shape_s=4.0):
    """FP64 oracle (numpy + scipy.fft with `workers` FFT threads) timed on a
    bounded sample.

    For the first request of every box shape in the stream, the oracle's set-up
    (Green's tensor + its FFT) and its BiCGSTAB iterations are timed, running
    iterations until about per_shape_s seconds. Each request of the stream then
    costs set-up + (its FP64 iteration count) x (per-iteration time of its
    shape); the CPU rate is the stream's length over the summed cost."""
    import dda_oracle as O
    per_shape = {}
    sample_s = 0.0
    for r in reqs:
        if r.shape in per_shape:
            continue
        a = O.wrapper_args(list(r.shape), r.d, np.repeat(r.rho, 3), list(r.material), list(r.n_k), list(r.n_e0),
                           r.lam, symmetry_method="None", apc_max_m=0, apc_max_n=0, apc_method_name="LDR",
                           tables=(np.zeros(2), np.zeros(2), 0.1))
        t0 = time.perf_counter()
        m = O.OracleDDAModel(*a, workers=workers)
        setup = time.perf_counter() - t0
        k, t_it = 1, 0.0
        while True:
            t1 = time.perf_counter()
            _, info = O.bicgstab_reference(m.apply, m.E, np.zeros(3 * m.N, complex), k, 0.0)
            dt = time.perf_counter() - t1
            # the first pass has three more matvecs than the k loop iterations
            t_it = dt * 2.0 / (info.matvecs + 1.0e-30)
            sample_s += dt
            if dt > per_shape_s / 4 or k >= 64:
                break
            k *= 4
        per_shape[r.shape] = (setup, t_it, info.matvecs)
        sample_s += setup
    total = sum(per_shape[r.shape][0] + iterations[i] * per_shape[r.shape][1] for i, r in enumerate(reqs))
    desc = ", ".join(f"{'x'.join(map(str, s))}: setup {v[0]:.2f} s, {v[1] * 1e3:.1f} ms/iteration ({v[2]} matvecs)"
                     for s, v in per_shape.items())
    return len(reqs) / total, sample_s, desc


def lattice_sum_rates(native, tables, threads):
    # tables: the package's Si/Ci tables for the device set-up; the CPU
    # oracle evaluates its own
    """FP64 Green's-tensor lattice sum of the periodic unit cell (32x32x12,
    d = 10 nm, 600 nm, FCD, period 32 cells), in Green evaluations per second
    = (2Nx-1)(2Ny-1)(2Nz-1)(2 MAXm+1)(2 MAXn+1) / time.  CPU: the oracle's C
    sum (OpenMP, `threads` threads, octant mode) at +-8 images, a bounded
    sample; GPU: the device operator set-up at the full +-50 images (lattice
    sum, DFTs and spectrum layout, timed by the library)."""
    import dda_oracle as O
    shape, d, lam = (32, 32, 12), 10.0, 600.0
    m_ref = (2 * shape[0] - 1) * (2 * shape[1] - 1) * (2 * shape[2] - 1)
    si, ci, delta = tables
    K = 2 * math.pi / lam
    cpu_images = 8
    osi, oci, odelta = O.sici_tables()
    t0 = time.perf_counter()
    O.green_offsets(*shape, d, K, O.METHOD_FCD, cpu_images, cpu_images, 32 * d, 32 * d, osi, oci, odelta,
                    octant_only=True)
    t_cpu = time.perf_counter() - t0
    cpu_rate = m_ref * (2 * cpu_images + 1) ** 2 / t_cpu
    req = native.Request(shape=shape, d=d, lam=lam, rho=np.full(int(np.prod(shape)), 0.5), n_k=(0.0, 0.0, 1.0),
                         n_e0=(0.0, 1.0, 0.0), method="FCD", images=50, period=(32, 32))
    native.check(native.lib().tdda_trim_caches())
    p = native.Problem(req, precision="fp32_refined", tables=tables)
    setup_ms = p.operator_setup_ms()
    del p
    gpu_rate = m_ref * 101 ** 2 / (setup_ms * 1e-3)
    return {"unit": "Green evaluations/s", "cpu": round(cpu_rate, 1), "cpu_threads": threads,
            "cpu_sample": f"oracle C lattice sum, +-{cpu_images} images, {t_cpu:.2f} s",
            "gpu": round(gpu_rate, 1), "gpu_setup_ms": round(setup_ms, 3), "gpu_images": 50}


PROBE_PER_SHAPE = 48
TRAFFIC_FILE = "matvec_probe240_pmc.md"
# launch groups of the batched c64 matvec (pass slots of time_passes): the
# x+y forward transform (plane passes; separate X and Y passes for boxes
# outside the plane classes), the z pass with the Green's multiply, and the
# inverse y+x transform with the al x epilogue
GROUPS = [("xy_fwd", (0, 1)), ("z", (2,)), ("yx_inv", (3, 4))]


def matvec_traffic():
    """HBM bytes per launch group (xy_fwd, z, yx_inv) of one batched matvec
    on the probe batch, from the committed rocprofv3 counter summary
    (counters cannot be read inside this process); None if absent."""
    try:
        with open(os.path.join(ROOT, "profiles", TRAFFIC_FILE)) as f:
            for line in f:
                if line.startswith("group_traffic_bytes:"):
                    return [int(x) for x in line.split(":")[1].split(",")]
    except OSError:
        pass
    return None


def probe_shapes():
    return [s for _ in range(PROBE_PER_SHAPE) for s in SHAPES]


def roofline_probe(native, reps, iteration=True):
    """One batched matvec (one launch per pass and class) over a serving-like
    batch (48 problems of every shape, about half of one step's 512 requests,
    distinct wavelengths so no two share a Green's spectrum, c64), optionally the ten
    launches of one refined BiCGSTAB step and the same matvec with separate
    x and y passes (returned with them), and a device copy of known size for
    calibrating byte counters."""
    import torch
    probs = [native.Problem(native.Request(shape=s, d=15.0, lam=500.0 + 1.0 * i, rho=np.full(int(np.prod(s)), 0.5)),
                            precision="fp32_refined") for i, s in enumerate(probe_shapes())]
    ms5 = native.time_passes(probs, c64=True, reps=reps)
    ms10 = native.time_iteration(probs, reps=reps) if iteration else None
    if iteration:   # the same matvec without the plane passes (separate x and y passes), for comparison
        native.set_plane_passes(False)
        try:
            ms5_sep = native.time_passes(probs, c64=True, reps=reps)
        finally:
            native.set_plane_passes(True)
        ms10 = (ms10, ms5_sep)
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")   # 256 MiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    return ms5, ms10, probs


def roofline_record(ms5):
    """Roofline of one batched matvec y = A x + al x over the probe batch.

    Algorithmic bytes per matvec and problem (the hot path's unit): with s = 8
    (c64), n = 3 Nx Ny Nz unknowns and M = Lx Ly Lz the padded grid,
    B_mv = s (2n + 12 M) + 6 s M: x in and y out, one read and one write of
    the 3-component grid per transform direction, and the 6-component Green's
    spectrum once.  `achieved` is the batch's sum of B_mv over the matvec's
    summed launch durations (HIP events on the launch stream).  The kernels
    move less than B_mv (pruned zero padding, the plane kept in LDS between
    the x and y transforms, octant spectrum), so the measured HBM bytes are
    reported beside it as `traffic` and `actual_GBs`."""
    Ms = [fft_len(s[0]) * fft_len(s[1]) * fft_len(s[2]) for s in probe_shapes()]
    ns = [3 * int(np.prod(s)) for s in probe_shapes()]
    b_mv = sum(8 * (2 * n + 12 * M) + 8 * 6 * M for n, M in zip(ns, Ms))
    t = float(sum(ms5)) * 1e-3
    pt = matvec_traffic()
    groups = {}
    for k, (nm, slots) in enumerate(GROUPS):
        ms = float(sum(ms5[i] for i in slots))
        e = {"ms": round(ms, 5)}
        if pt:
            e["bytes"] = pt[k]
            e["GBs"] = round(pt[k] / (ms * 1e-3) / 1e9, 1)
        groups[nm] = e
    traffic = sum(pt) if pt else None
    return {"bound": "hbm",
            "kernel": f"batched c64 matvec y = A x + al x over a {len(Ms)}-request batch ({PROBE_PER_SHAPE} per box "
                      f"shape): k_plane_fwd (x and y transforms of each z-plane), k_zpass_fast (z transform, Green's "
                      f"multiply, inverse z), k_plane_inv (inverse y and x, al x epilogue), one launch per length class",
            "achieved": round(b_mv / t / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(b_mv / t / 1e9 / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes": b_mv, "matvec_ms": round(t * 1e3, 5),
            "actual_GBs": round(traffic / t / 1e9, 1) if traffic else None,
            "actual_frac": round(traffic / t / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
            "traffic_source": f"profiles/{TRAFFIC_FILE} (rocprofv3 FETCH_SIZE and WRITE_SIZE, each scaled by "
                              f"its reading for a 256 MiB device copy, same batch)",
            "groups": groups}


C4_SHAPE = (128, 128, 64)
C4_LAMS = (450.0, 550.0, 650.0, 750.0)


def smooth_density(shape, sigma, seed):
    """rho = clip(0.5 + 0.25 f / std f, 0, 1), f = Gaussian-filtered white noise
    (periodic in x and y, nearest-value edges in z), flattened x-fastest."""
    from scipy.ndimage import gaussian_filter
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal((nz, ny, nx)), sigma=sigma, mode=("nearest", "wrap", "wrap"))
    return np.clip(0.5 + 0.25 * f / f.std(), 0.0, 1.0).reshape(-1)


def multi_rhs_probe(native, reps):
    """The 128x128x64 eight-RHS batch's c64 matvec (ms per pass), with its
    right-hand sides sharing 4 operators (4 wavelengths x 2 polarisations,
    R_op = 2: the Z tiles of systems that share an operator are interleaved so
    that each octant record is read from HBM about once per operator) and,
    for comparison, with 8 distinct operators (8 wavelengths)."""
    rho = smooth_density(C4_SHAPE, 3.0, seed=5)
    shared = [native.Request(shape=C4_SHAPE, d=8.0, lam=l, rho=rho, n_k=(0.0, 0.0, 1.0), n_e0=e)
              for l in C4_LAMS for e in ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0))]
    distinct = [native.Request(shape=C4_SHAPE, d=8.0, lam=450.0 + 40.0 * k, rho=rho, n_k=(0.0, 0.0, 1.0),
                               n_e0=(1.0, 0.0, 0.0)) for k in range(8)]
    out = {}
    for name, reqs in (("shared", shared), ("distinct", distinct)):
        probs = [native.Problem(r, precision="fp32_refined") for r in reqs]
        out[name] = [float(x) for x in native.time_passes(probs, c64=True, reps=reps)]
        del probs
    native.check(native.lib().tdda_trim_caches())
    return out


def multi_rhs_record(native, reps):
    """128x128x64 dipoles (FFT 256x256x128), d = 8 nm, smooth density (sigma
    3), 8 right-hand sides = 4 wavelengths x 2 polarisations, 2 per operator.
    The eight solves run as one batch to 1e-5, refined (field-error target
    1e-5) and FP64, each with and without the Jacobi preconditioner; the
    batched c64 matvec is timed per pass with B_mv = 8 (2n + 12 M) + 8 * 6 M /
    R_op per right-hand side, R_op = 2."""
    rho = smooth_density(C4_SHAPE, 3.0, seed=5)
    reqs = [native.Request(shape=C4_SHAPE, d=8.0, lam=l, rho=rho, n_k=(0.0, 0.0, 1.0), n_e0=e)
            for l in C4_LAMS for e in ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0))]
    out = {}
    for prec in ("fp32_refined", "fp64"):
        for pc in ("jacobi", "none"):
            probs = [native.Problem(r, precision=prec, precond=pc) for r in reqs]
            t0 = time.perf_counter()
            infos = native.solve(probs, tol=1e-5)
            dt = time.perf_counter() - t0
            rec = {"seconds": round(dt, 4), "solves_per_s": round(len(reqs) / dt, 3),
                   "iterations": [i["iterations"] for i in infos],
                   "converged": sum(1 for i in infos if i["status"] == "converged")}
            if prec == "fp32_refined":
                rec["err_est_max"] = max(i["err_est"] for i in infos)
            out[f"{prec}_{pc}"] = rec
            if prec == "fp32_refined" and pc == "jacobi":
                ms5 = native.time_passes(probs, c64=True, reps=reps)
                n = 3 * int(np.prod(C4_SHAPE))
                M = fft_len(C4_SHAPE[0]) * fft_len(C4_SHAPE[1]) * fft_len(C4_SHAPE[2])
                b_mv = len(reqs) * (8 * (2 * n + 12 * M) + 8 * 6 * M / 2)
                t = float(sum(ms5)) * 1e-3
                out["matvec"] = {"ms": round(t * 1e3, 4), "pass_ms": [round(float(x), 4) for x in ms5],
                                 "algorithmic_bytes": int(b_mv), "GBs": round(b_mv / t / 1e9, 1),
                                 "frac": round(b_mv / t / 1e9 / HBM_PEAK_GBS, 4)}
            del probs
    mp = multi_rhs_probe(native, reps)
    out["matvec"]["z_ms_shared_operators"] = round(mp["shared"][2], 4)
    out["matvec"]["z_ms_distinct_operators"] = round(mp["distinct"][2], 4)
    out["speedup_refined_vs_fp64_batched_jacobi"] = round(
        out["fp64_jacobi"]["seconds"] / out["fp32_refined_jacobi"]["seconds"], 3)
    native.check(native.lib().tdda_trim_caches())
    return out


def periodic_cell_record(native, tables):
    """BASELINE c1: 32x32x12 unit cell, d = 10 nm, periods 320 nm, FCD with
    +-50 images, 600 nm, normal incidence, y-polarised, smooth density
    (Gaussian-filtered noise, sigma 2, clipped).  Reports the FP64 operator
    set-up (lattice sum, DFTs), the refined (c64 + FP64 reliable updates) and
    FP64 solves to 1e-5 (solver ms, iterations) and their fields against a
    1e-10 FP64 solve."""
    shape = (32, 32, 12)
    rho = smooth_density(shape, 2.0, seed=1)
    req = native.Request(shape=shape, d=10.0, lam=600.0, rho=rho, n_k=(0.0, 0.0, 1.0), n_e0=(0.0, 1.0, 0.0),
                         method="FCD", images=50, period=(32, 32))
    native.check(native.lib().tdda_trim_caches())
    p64 = native.Problem(req, precision="fp64", tables=tables)
    setup_ms = p64.operator_setup_ms()
    p64.solve(tol=1e-10)
    e_ref = p64.field()
    rec = {"operator_setup_ms": round(setup_ms, 3)}
    for name, prec, pc in (("fp32_refined", "fp32_refined", "none"), ("fp64", "fp64", "none"),
                           ("fp32_refined_jacobi", "fp32_refined", "jacobi"), ("fp64_jacobi", "fp64", "jacobi")):
        p = native.Problem(req, precision=prec, tables=tables, precond=pc)
        p.solve(tol=1e-5)   # warm
        info = p.solve(tol=1e-5)
        rec[name] = {"solve_ms": round(info["ms"], 3), "iterations": info["iterations"],
                     "field_rel_err": rel_err(p.field(), e_ref)}
        del p
    del p64
    native.check(native.lib().tdda_trim_caches())
    return rec


def target_gradient_record(native):
    """BASELINE c3: 48x48x24, d = 12 nm, 542 nm, x-polarised, seeded uniform
    density through a linear cone filter of radius 2 voxels, mean |E|^2 over a
    4x4x4 target block 60 nm above the structure.  ms per objective + adjoint
    gradient through the torch.autograd.Function in FP32-refined (tol 1e-5)
    and FP64 (tol 1e-5 and 1e-10), and each gradient against the tight FP64
    one (the central-difference check on 32 voxels is a GPU test)."""
    import torch
    from dda_model.autograd import cone_filter, dda_objective, target_block
    shape = (48, 48, 24)
    nx, ny, nz = shape
    rng = np.random.default_rng(3)
    raw = torch.tensor(rng.uniform(0, 1, (nz, ny, nx)), dtype=torch.float64, device="cuda", requires_grad=True)
    tg = np.asarray(target_block(shape, 12.0, size=4, height_nm=60.0))
    req = native.Request(shape=shape, d=12.0, lam=542.0, rho=np.full(nx * ny * nz, 0.5))
    probs = {p: native.Problem(req, precision=p) for p in ("fp64", "fp32_refined")}
    out, grads = {}, {}
    for name, prec, tol in (("fp64_tight", "fp64", 1e-10), ("fp32_refined", "fp32_refined", 1e-5),
                            ("fp64", "fp64", 1e-5)):
        for rep in range(2):   # the first call is a warm-up
            raw.grad = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            val = dda_objective(cone_filter(raw, 2.0).reshape(-1), probs[prec], kind="targets", targets=tg, tol=tol)
            val.backward()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        grads[name] = raw.grad.detach().cpu().numpy().copy()
        out[name] = {"ms": round(dt * 1e3, 2), "objective": val.item()}
    for name in ("fp32_refined", "fp64"):
        out[name]["gradient_rel_err"] = rel_err(grads[name], grads["fp64_tight"])
    del probs
    native.check(native.lib().tdda_trim_caches())
    return out


def copy_bandwidth():
    """Device-to-device copy of 256 MiB timed with HIP events (read + write bytes / time)."""
    import torch
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    return 2 * (256 << 20) * 20 / (e0.elapsed_time(e1) * 1e-3) / 1e9


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--fp64-requests", type=int, default=96)
    ap.add_argument("--slots", type=int, default=512)
    ap.add_argument("--chunk", type=int, default=2)
    ap.add_argument("--load", type=float, default=0.8)
    ap.add_argument("--ru-delta", type=float, default=0.001,
                    help="reliable update when the recursive residual falls this far below its peak since the last "
                         "(a sweep gave 0.001 2.9 %% above 0.01 and 1.5 %% above 0.0001)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="skip the 128x128x64 multi-RHS record")
    ap.add_argument("--precond", default="jacobi", choices=["jacobi", "none"],
                    help="preconditioner of the refined server (the FP64 single-solve path is timed with both)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the batched pass timing and a 256 MiB copy (for counter runs)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default kernel schedule (no autotuning in the warm-up)")
    ap.add_argument("--multi-rhs-only", action="store_true",
                    help="only the 128x128x64 eight-RHS matvec, shared and distinct operators (for counter runs)")
    ap.add_argument("--saturation-only", action="store_true",
                    help="only warm-up and the timed saturation steps; prints solves/s (for kernel traces)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (no multi-GPU path is built)",
              file=sys.stderr)
        sys.exit(2)
    from dda_model import native

    if a.multi_rhs_only:
        if not a.no_tune:
            native.tune(SHAPES, count=16, reps=10)
        print(json.dumps({"multi_rhs_pass_ms": multi_rhs_probe(native, 30)}), flush=True)
        return

    if a.roofline_only:
        # the matvec probe alone (after the tuner, as in the full run), so that
        # a profile's last 30 launches of each kernel are the probe's
        if not a.no_tune:
            native.tune(SHAPES, count=16, reps=10)
        ms5, _, _ = roofline_probe(native, 30, iteration=False)
        print(json.dumps({"pass_ms": [float(x) for x in ms5], "copy_bytes": 2 * (256 << 20),
                          "roofline": roofline_record(ms5)}), flush=True)
        return

    # Every step serves its own requests (fresh seeds): wavelengths are
    # continuous, so no step finds its Green's spectra in the operator cache
    # and every request pays its set-up inside the timed region.
    R = a.requests

    # Kernel schedule: the autotuner times the split variants and the fused
    # y/z pass for the stream's boxes (untimed, part of the warm-up)
    tune_rec = {"tuned": False}
    if not a.no_tune:
        t0 = time.perf_counter()
        decisions = native.tune(SHAPES, count=16, reps=10)
        tune_rec = {"tuned": True, "seconds": round(time.perf_counter() - t0, 2),
                    "variant_1_chosen": sum(1 for d in decisions if d["kind"] == "pass" and d["chosen"] == 1),
                    "fused_yz_chosen": sum(1 for d in decisions if d["kind"] == "yz" and d["chosen"] == 1),
                    "plane_chosen": [d["L"] for d in decisions if d["kind"] == "xy" and d["chosen"] == 1],
                    "decisions": len(decisions)}
        print(f"# tuned schedule in {tune_rec['seconds']} s", file=sys.stderr, flush=True)

    # One long-lived server takes the warm-up and the timed steps, as a
    # service would; every step still submits R fresh requests and waits for
    # all of them (ramp-up and tail included).
    server = native.Server(precision="fp32_refined", max_slots=a.slots, chunk=a.chunk, keep_fields=False,
                           precond=a.precond, ru_delta=a.ru_delta)
    next_id = 0
    # warmup: requests outside the timed seeds (plans, twiddles, allocator,
    # system pool, code objects), as many as the server has slots so that its
    # per-shape system pool and the device allocator hold what a full batch
    # uses, as in a running service; every timed request still sets up its own
    # operator and system
    for w in range(a.warmup):
        wreqs = make_requests(native, max(128, min(a.slots, R)), seed0=900000 + 1024 * w)
        run_stream(native, wreqs, "fp32_refined", a.slots, a.chunk, server=server, id0=next_id)
        next_id += len(wreqs)
    print("# warmup done", file=sys.stderr, flush=True)

    # timed steps: saturation throughput of the refined batched server
    walls, hits, res_steps, stats_sat = [], 0, [], None
    reqs0 = None
    for k in range(a.steps):
        reqs = make_requests(native, R, seed0=1000 + R * k)
        if k == 0:
            reqs0 = reqs
        w, res, stats_sat, _ = run_stream(native, reqs, "fp32_refined", a.slots, a.chunk, server=server,
                                          id0=next_id)
        next_id += R
        walls.append(w)
        res_steps.append(res)
        hits += sum(1 for r in res.values() if r["operator_cached"])
        print(f"# step {k + 1}: {R / w:.1f} solves/s", file=sys.stderr, flush=True)
    total = sum(walls)
    value = a.steps * R / total
    res_all = [r for res in res_steps for r in res.values()]
    conv = sum(1 for r in res_all if r["status"] == "converged")
    if a.saturation_only:
        server.close()
        print(json.dumps({"saturation_solves_per_s": round(value, 3), "converged": conv, "operator_cache_hits": hits,
                          "server": stats_sat}), flush=True)
        return

    # Poisson arrivals at load x saturation -> latency (fresh requests again)
    rate = a.load * value
    rng = np.random.default_rng(2)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, R))
    arrivals -= arrivals[0]
    reqs_p = make_requests(native, R, seed0=1000 + R * a.steps)
    w_p, res_p, stats_p, _ = run_stream(native, reqs_p, "fp32_refined", a.slots, a.chunk, arrivals=arrivals,
                                        server=server, id0=next_id, ru_delta=a.ru_delta)
    next_id += R
    server.close()
    lat = [(r["t_done"] - r["t_submit"]) * 1e3 for r in res_p.values()]
    print(f"# poisson stream done: p50 {percentile(lat, 0.5):.2f} ms", file=sys.stderr, flush=True)

    # field accuracy and the FP64 single-solve path on the first requests of
    # step 1: reference fields from a tight FP64 solve (batched, tol 1e-10)
    sub = reqs0[:min(a.fp64_requests, R)]
    _, res_t, _, f_tight = run_stream(native, sub, "fp64", len(sub), 16, fields=True, tol=1e-10, precond="jacobi")
    _, res32s, _, f32 = run_stream(native, sub, "fp32_refined", a.slots, a.chunk, fields=True, precond=a.precond,
                                   ru_delta=a.ru_delta)
    err32 = [rel_err(f32[i], f_tight[i]) for i in f_tight]
    est = [res32s[i]["err_est"] for i in f_tight]
    # the FP64 single-solve path (one request at a time, FP64 transforms and
    # vectors, the reference's recursive-residual stopping rule), without and
    # with the preconditioner, at tol 1e-5 and at the tolerance that gives the
    # refined path's field accuracy (field error ~ 2 x residual here)
    tol_eq = 1e-5 / 3.0
    fp64 = {}
    for pc in ("none", "jacobi"):
        for tol in (1e-5, tol_eq):
            w64, res64, _, f64 = run_stream(native, sub, "fp64", 1, 8, fields=True, tol=tol, precond=pc)
            errs = [rel_err(f64[i], f_tight[i]) for i in f_tight]
            fp64[(pc, tol)] = {"solves_per_s": round(len(sub) / w64, 3), "tol": round(tol, 9),
                               "iterations_p50": percentile([r["iterations"] for r in res64.values()], 0.5),
                               "field_err_max": max(errs), "field_err_p50": percentile(errs, 0.5),
                               "_iters": {i: r["iterations"] for i, r in res64.items()}}
    fp64_rate = fp64[("none", 1e-5)]["solves_per_s"]
    print(f"# fp64 single-solve: {fp64_rate:.2f} solves/s (jacobi {fp64[('jacobi', 1e-5)]['solves_per_s']:.2f}); "
          f"refined field error max {max(err32):.3e}", file=sys.stderr, flush=True)
    # the refined server without preconditioning on one step's requests
    w_np, res_np, _, _ = run_stream(native, reqs0, "fp32_refined", a.slots, a.chunk, precond="none",
                                    ru_delta=a.ru_delta)

    # roofline of the batched matvec (one launch per pass over the probe batch)
    ms5, (ms10, ms5_sep), probs = roofline_probe(native, 30)
    roof = roofline_record(ms5)
    roof["iteration_pass_ms"] = [round(float(x), 5) for x in ms10]
    roof["separate_passes"] = {"matvec_ms": round(float(sum(ms5_sep)), 5),
                               "pass_ms": [round(float(x), 5) for x in ms5_sep],
                               "plane_passes_time_ratio": round(float(sum(ms5)) / float(sum(ms5_sep)), 4)}
    del probs
    native.check(native.lib().tdda_trim_caches())
    copy_gbs = copy_bandwidth()

    c4 = None if a.no_c4 else multi_rhs_record(native, 10)
    if c4:
        print(f"# multi-RHS 128x128x64: {c4['fp32_refined_jacobi']['solves_per_s']} solves/s refined", file=sys.stderr,
              flush=True)

    import dda_model
    pkg_tables = dda_model.sici_tables()   # the package's Si/Ci tables (FCD)
    c1 = None if a.no_c4 else periodic_cell_record(native, pkg_tables)
    c3 = None if a.no_c4 else target_gradient_record(native)

    cpu, lattice = None, None
    host = host_cpu()
    if not a.no_cpu:
        import dda_oracle as O
        # iteration counts: the FP64 GPU run where it covered the request
        # (same recurrence as the oracle), else the refined run's count
        its64 = fp64[("none", 1e-5)]["_iters"]
        its = [its64[i] if i in its64 else res_np[i]["iterations"] for i in range(R)]
        threads = min(16, host["nproc"] or 1)
        rate1, s1, desc1 = cpu_baseline(reqs0, its, workers=1)
        rate_n, s_n, desc_n = cpu_baseline(reqs0, its, workers=threads)
        cpu = {"value": round(rate_n, 5), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"FP64 oracle (numpy + scipy.fft with {threads} FFT threads) timed for {s_n:.1f} s on the "
                         f"first request of each box shape (set-up + BiCGSTAB iterations: {desc_n}); scaled to the "
                         f"{R}-request stream by each request's FP64 iteration count",
               "single_thread": {"value": round(rate1, 5), "cores": 1, "sample_s": round(s1, 1),
                                 "detail": desc1},
               "host": host}
        lattice = lattice_sum_rates(native, pkg_tables, int(os.environ.get("OMP_NUM_THREADS", threads)))

    iters = [r["iterations"] for r in res_all]
    out = {
        "metric": "DDA solves/sec at 1e-5 residual, FP32-refined batched vs FP64; p50/p99 ms",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(total / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64-refine",
        "data": "synthetic",
        "config": {"workload": f"serving stream of {R} DDA requests per step, fresh seeds every step (boxes "
                               "16^3..64x64x32, d=15nm, lambda U[400,800]nm, polar angle U[0,30deg], random "
                               "polarisation, density U[0,1), LDR, non-periodic), each solved to 1e-5 true residual "
                               "and 1e-5 estimated field error (right-Jacobi preconditioned BiCGSTAB unless precond "
                               "says otherwise); Green's set-up inside the timed region",
                   "slots": a.slots, "chunk": a.chunk, "ru_delta": a.ru_delta, "poisson_load": a.load,
                   "schedule": tune_rec},
        "operator_cache_hits": hits,
        "p50_ms": round(percentile(lat, 0.5), 3),
        "p99_ms": round(percentile(lat, 0.99), 3),
        "poisson_rate_per_s": round(rate, 2),
        "poisson_solves_per_s": round(R / w_p, 3),
        "converged": conv,
        "iterations_p50": percentile(iters, 0.5),
        "reliable_updates_p50": percentile([r["reliable_updates"] for r in res_all], 0.5),
        "reliable_updates_max": max(r["reliable_updates"] for r in res_all),
        "err_est_p50": percentile([r["err_est"] for r in res_all], 0.5),
        "err_est_max": max(r["err_est"] for r in res_all),
        "precond": a.precond,
        "field_accuracy": {
            "requests": len(sub), "reference": "FP64 solve to 1e-10 relative residual",
            "refined_max": max(err32), "refined_p50": percentile(err32, 0.5),
            "refined_within_1e-5": bool(max(err32) <= 1e-5),
            "refined_estimate_over_actual_min": min(e / a_ for e, a_ in zip(est, err32)),
            "refined_estimate_over_actual_max": max(e / a_ for e, a_ in zip(est, err32))},
        "fp64_single_solve": {f"{pc}_tol_{tol:.3g}": {k: v for k, v in rec.items() if k != "_iters"}
                              for (pc, tol), rec in fp64.items()},
        "speedup_vs_fp64_single": round(value / fp64_rate, 3),
        "speedup_vs_fp64_single_jacobi": round(value / fp64[("jacobi", 1e-5)]["solves_per_s"], 3),
        "speedup_vs_fp64_single_equal_accuracy": round(value / fp64[("none", tol_eq)]["solves_per_s"], 3),
        "speedup_vs_fp64_single_jacobi_equal_accuracy": round(value / fp64[("jacobi", tol_eq)]["solves_per_s"], 3),
        "refined_unpreconditioned": {"solves_per_s": round(R / w_np, 3), "requests": R,
                                     "iterations_p50": percentile([r["iterations"] for r in res_np.values()], 0.5)},
        "server": {k: stats_sat[k] for k in ("chunks", "rebuilds", "launches", "max_active", "peak_bytes")},
        "copy_GBs": round(copy_gbs, 1),
        "multi_rhs_128x128x64": c4,
        "periodic_cell_32x32x12": c1,
        "target_gradient_48x48x24": c3,
        "lattice_sum": lattice,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
