"""Benchmark of the DDA solve path on one MI355X; prints ONE JSON line.

Workload (headline): a stream of R DDA requests (default 512, seeds 1000..),
each independently drawn: box uniform over {16^3, 24^3, 32x32x16, 48x48x24,
64x64x32} with d = 15 nm, wavelength uniform in [400, 800] nm, polar angle
uniform in [0, 30 deg], azimuth and polarisation angle uniform, density
uniform in [0, 1) per voxel, LDR interaction, non-periodic.  Every request is
solved to a 1e-5 relative residual (true FP64 residual) by the continuous-
batching server in complex64 with FP64 reliable updates.

A step is one pass over the R requests with all of them queued at once (the
server's saturation throughput); `value` is solves/s over the K timed steps.
The same requests then arrive as a Poisson stream at 0.8x that rate, which
gives the p50/p99 latency.  The plain FP64 single-solve GPU path (one request
at a time, FP64 transforms and Krylov vectors) is timed on a subset of the
same requests for the speedup, and the FP64 CPU oracle on a bounded sample.
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


def run_stream(native, reqs, precision, max_slots, chunk, arrivals=None, fields=False, timeout=600.0):
    """Submit every request (at its arrival time if given) and wait for all.
    Returns (wall seconds from first submission to last completion, results)."""
    srv = native.Server(precision=precision, max_slots=max_slots, chunk=chunk, keep_fields=fields)
    results = {}
    t0 = time.perf_counter()
    nxt = 0
    deadline = t0 + timeout
    while len(results) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and (arrivals is None or arrivals[nxt] <= now):
            srv.submit(nxt, reqs[nxt])
            nxt += 1
        for r in srv.poll():
            results[r["id"]] = r
        if time.perf_counter() > deadline:
            raise RuntimeError(f"stream did not finish in {timeout} s ({len(results)}/{len(reqs)})")
        time.sleep(0.0005)
    wall = time.perf_counter() - t0
    out_fields = {i: srv.field(i) for i in results} if fields else None
    stats = srv.stats()
    srv.close()
    return wall, results, stats, out_fields


def cpu_baseline(reqs, iterations, per_shape_s=4.0):
    """FP64 oracle (numpy + scipy.fft, one worker) timed on a bounded sample.

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
        m = O.OracleDDAModel(*a, workers=1)
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


PROBE_PER_SHAPE = 8
TRAFFIC_FILE = "matvec_probe40_pmc.md"


def matvec_traffic():
    """HBM bytes per pass launch (X fwd, Y fwd, Z, Y inv, X inv) on the probe
    batch, from the committed rocprofv3 counter summary (counters cannot be
    read inside this process); None if absent."""
    try:
        with open(os.path.join(ROOT, "profiles", TRAFFIC_FILE)) as f:
            for line in f:
                if line.startswith("pass_traffic_bytes:"):
                    return [int(x) for x in line.split(":")[1].split(",")]
    except OSError:
        pass
    return None


def probe_shapes():
    return [s for _ in range(PROBE_PER_SHAPE) for s in SHAPES]


def roofline_probe(native, reps):
    """One batched launch per pass over a serving-like batch (8 problems of
    every shape, distinct wavelengths so no two share a Green's spectrum, c64),
    and a device copy of known size for calibrating byte counters."""
    import torch
    probs = [native.Problem(native.Request(shape=s, d=15.0, lam=500.0 + 3.0 * i, rho=np.full(int(np.prod(s)), 0.5)),
                            precision="fp32_refined") for i, s in enumerate(probe_shapes())]
    ms5 = native.time_passes(probs, c64=True, reps=reps)
    ms10 = native.time_iteration(probs, reps=reps)
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
    spectrum once.  `achieved` is the batch's sum of B_mv over the five pass
    launches' summed duration (HIP events on the launch stream).  The passes
    move less than B_mv (pruned zero padding, octant spectrum), so the
    measured HBM bytes are reported beside it as `traffic` and `actual_GBs`."""
    Ms = [fft_len(s[0]) * fft_len(s[1]) * fft_len(s[2]) for s in probe_shapes()]
    ns = [3 * int(np.prod(s)) for s in probe_shapes()]
    b_mv = sum(8 * (2 * n + 12 * M) + 8 * 6 * M for n, M in zip(ns, Ms))
    t = float(sum(ms5)) * 1e-3
    pt = matvec_traffic()
    names = ["x_fwd", "y_fwd", "z", "y_inv", "x_inv"]
    passes = {}
    for k, nm in enumerate(names):
        e = {"ms": round(float(ms5[k]), 5)}
        if pt:
            e["bytes"] = pt[k]
            e["GBs"] = round(pt[k] / (ms5[k] * 1e-3) / 1e9, 1)
        passes[nm] = e
    traffic = sum(pt) if pt else None
    return {"bound": "hbm",
            "kernel": f"batched matvec y = A x + al x, c64: k_xfwd_fast, k_ypass_fast<-1>, k_zpass_fast, "
                      f"k_ypass_fast<+1>, k_xinv_fast, one launch each over a {len(Ms)}-request batch "
                      f"({PROBE_PER_SHAPE} per box shape)",
            "achieved": round(b_mv / t / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(b_mv / t / 1e9 / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes": b_mv, "matvec_ms": round(t * 1e3, 5),
            "actual_GBs": round(traffic / t / 1e9, 1) if traffic else None,
            "actual_frac": round(traffic / t / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
            "traffic_source": f"profiles/{TRAFFIC_FILE} (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, same batch)",
            "passes": passes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--fp64-requests", type=int, default=96)
    ap.add_argument("--slots", type=int, default=128)
    ap.add_argument("--chunk", type=int, default=4)
    ap.add_argument("--load", type=float, default=0.8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the batched pass timing and a 256 MiB copy (for counter runs)")
    ap.add_argument("--saturation-only", action="store_true",
                    help="only warm-up and the timed saturation steps; prints solves/s (for kernel traces)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (no multi-GPU path is built)",
              file=sys.stderr)
        sys.exit(2)
    from dda_model import native

    if a.roofline_only:
        ms5, ms10, _ = roofline_probe(native, 30)
        print(json.dumps({"pass_ms": [float(x) for x in ms5], "iteration_pass_ms": [float(x) for x in ms10],
                          "copy_bytes": 2 * (256 << 20),
                          "roofline": roofline_record(ms5)}), flush=True)
        return

    reqs = make_requests(native, a.requests)
    print(f"# {len(reqs)} requests generated", file=sys.stderr, flush=True)

    # warmup: a short stream (plans, twiddles, allocator, code objects)
    for _ in range(a.warmup):
        run_stream(native, reqs[:min(48, len(reqs))], "fp32_refined", a.slots, a.chunk)
    print("# warmup done", file=sys.stderr, flush=True)

    # timed steps: saturation throughput of the refined batched server
    walls, res_sat, stats_sat = [], None, None
    for _ in range(a.steps):
        w, res_sat, stats_sat, _ = run_stream(native, reqs, "fp32_refined", a.slots, a.chunk)
        walls.append(w)
        print(f"# step {len(walls)}: {len(reqs) / w:.1f} solves/s", file=sys.stderr, flush=True)
    total = sum(walls)
    value = a.steps * len(reqs) / total
    conv = sum(1 for r in res_sat.values() if r["status"] == "converged")
    if a.saturation_only:
        print(json.dumps({"saturation_solves_per_s": round(value, 3), "converged": conv,
                          "server": stats_sat}), flush=True)
        return

    # Poisson arrivals at load x saturation -> latency
    rate = a.load * value
    rng = np.random.default_rng(2)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    arrivals -= arrivals[0]
    w_p, res_p, stats_p, _ = run_stream(native, reqs, "fp32_refined", a.slots, a.chunk, arrivals=arrivals)
    lat = [(r["t_done"] - r["t_submit"]) * 1e3 for r in res_p.values()]
    print(f"# poisson stream done: p50 {percentile(lat, 0.5):.2f} ms", file=sys.stderr, flush=True)

    # plain FP64 single-solve GPU path on a subset, plus field agreement
    sub = reqs[:min(a.fp64_requests, len(reqs))]
    w64, res64, _, f64 = run_stream(native, sub, "fp64", 1, 8, fields=True)
    _, res32s, _, f32 = run_stream(native, sub, "fp32_refined", a.slots, a.chunk, fields=True)
    fp64_rate = len(sub) / w64
    e_err = max(float(np.linalg.norm(f32[i] - f64[i]) / np.linalg.norm(f64[i])) for i in f64)
    print(f"# fp64 single-solve: {fp64_rate:.2f} solves/s", file=sys.stderr, flush=True)

    # roofline of the batched matvec (one launch per pass over the probe batch)
    ms5, ms10, probs = roofline_probe(native, 30)
    roof = roofline_record(ms5)
    roof["iteration_pass_ms"] = [round(float(x), 5) for x in ms10]
    del probs

    cpu = None
    if not a.no_cpu:
        # iteration counts: the FP64 GPU run where it covered the request
        # (same recurrence as the oracle), else the refined run's count
        its = [res64[i]["iterations"] if i in res64 else res_sat[i]["iterations"] for i in range(len(reqs))]
        rate_cpu, sample_s, desc = cpu_baseline(reqs, its)
        cpu = {"value": round(rate_cpu, 5), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"FP64 oracle (numpy+scipy.fft, workers=1) timed for {sample_s:.1f} s on the first "
                         f"request of each box shape (set-up + BiCGSTAB iterations: {desc}); scaled to the "
                         f"{len(reqs)}-request stream by each request's FP64 iteration count"}

    iters = [r["iterations"] for r in res_sat.values()]
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
        "config": {"workload": f"serving stream of {len(reqs)} DDA requests (boxes 16^3..64x64x32, d=15nm, "
                               "lambda U[400,800]nm, polar angle U[0,30deg], random polarisation, density "
                               "U[0,1), LDR, non-periodic), solved to 1e-5 true residual",
                   "slots": a.slots, "chunk": a.chunk, "poisson_load": a.load},
        "p50_ms": round(percentile(lat, 0.5), 3),
        "p99_ms": round(percentile(lat, 0.99), 3),
        "poisson_rate_per_s": round(rate, 2),
        "poisson_solves_per_s": round(len(reqs) / w_p, 3),
        "converged": conv,
        "iterations_p50": percentile(iters, 0.5),
        "fp64_single_solve": {"solves_per_s": round(fp64_rate, 3), "requests": len(sub),
                              "iterations_p50": percentile([r["iterations"] for r in res64.values()], 0.5)},
        "speedup_vs_fp64_single": round(value / fp64_rate, 3),
        "field_rel_err_refined_vs_fp64_max": e_err,
        "server": {k: stats_sat[k] for k in ("chunks", "rebuilds", "launches", "max_active", "peak_bytes")},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
