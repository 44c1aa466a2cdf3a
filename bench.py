"""Benchmark of the DDA solve path on one MI355X.

Prints ONE JSON line.  A step is one solve of the forward system to a 1e-5
relative residual, with the operator, 1/alpha and the right-hand side already
resident on the device.

Workload: 2-D periodic unit cell, 32x32x12 dipoles, d = 10 nm, periods 320 nm,
FCD interaction with a +-50 image lattice sum in FP64, lambda = 600 nm, normal
incidence, y polarisation, smooth seeded density (Gaussian filter radius 2,
clipped to [0, 1]).  `value` is the FP32-refined solver's solves/s; the plain
FP64 solver is timed on the same system in the same run.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "topo-dda_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def smooth_density(shape_xyz, sigma, seed):
    """rho = clip(0.5 + 0.25 f / std(f), 0, 1), f = Gaussian-filtered N(0,1)
    noise on (z, y, x), periodic in x and y, nearest at the z faces."""
    from scipy.ndimage import gaussian_filter
    nx, ny, nz = shape_xyz
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal((nz, ny, nx)), sigma=sigma, mode=("nearest", "wrap", "wrap"))
    rho = np.clip(0.5 + 0.25 * f / f.std(), 0.0, 1.0)
    return rho.reshape(-1)


def unit_cell_args(tables):
    """Positional DDAModel arguments of the 32x32x12 periodic unit cell."""
    nx, ny, nz = 32, 32, 12
    rho = smooth_density((nx, ny, nz), 2.0, seed=1)
    import dda_model
    geometry = dda_model._generate_geometry(nx, ny, nz)
    si, ci, delta = tables
    n = nx * ny * nz
    return [0.0, 50.0, 0.5, "piecewise", [100], [2.0], False,
            "None", [], True, "IntegratedE", [2, 0.0, nx - 1.0, 0.0, ny - 1.0, 0.0, nz - 1.0, 0.95, 50.0],
            geometry, np.repeat(rho, 3), nx, ny, nz, n,
            [0.0, 0.0, 1.0], 1.0, [0.0, 1.0, 0.0], 600.0, [1.0 + 0j, 5.96 + 3.89j], 1.0,
            50, 50, 32.0, 32.0, "FCD", 10.0, si, ci, delta, False]


def percentile(v, q):
    s = sorted(v)
    k = (len(s) - 1) * q
    lo, hi = math.floor(k), math.ceil(k)
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


def cpu_baseline(args, tol):
    """The FP64 oracle (numpy + scipy.fft, one worker) on the same system:
    Green's setup timed apart, then one full solve."""
    import dda_oracle as O
    t0 = time.perf_counter()
    m = O.OracleDDAModel(*args, workers=1)
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    m.solveElectricField(np.zeros(3 * m.N, complex), 100000, tol)
    solve = time.perf_counter() - t0
    return setup, solve, m


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (no multi-GPU path is built)",
              file=sys.stderr)
        sys.exit(2)

    import dda_model
    tables = dda_model.sici_tables()
    args = unit_cell_args(tables)
    t0 = time.perf_counter()
    model = dda_model.DDAModel(*args)
    build_s = time.perf_counter() - t0
    (lx, ly, lz), green_ms = model.fftGrid()
    n = 3 * 32 * 32 * 12
    M = lx * ly * lz

    def run(precision, steps, warmup):
        model.setSolverPrecision(precision)
        for _ in range(warmup):
            model.solveResident(100000, a.tol)
        times, infos = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            infos.append(model.solveResident(100000, a.tol))
            times.append((time.perf_counter() - s0) * 1e3)
        wall = time.perf_counter() - t0
        return wall, times, infos

    wall32, t32, i32 = run("fp32_refined", a.steps, a.warmup)
    wall64, t64, i64 = run("fp64", a.steps, a.warmup)

    # per-pass timing (HIP events, library stream)
    model.setSolverPrecision("fp32_refined")
    p32 = model.timePasses(50)
    model.setSolverPrecision("fp64")
    p64 = model.timePasses(50)
    s = 8
    b_z = s * (4 * M + 6 * M)                      # Z pass share of B_mv, R_op = 1
    b_mv = s * (2 * n + 12 * M) + s * 6 * M        # per matvec
    z_ms = p32[2]
    roof = {"bound": "hbm", "kernel": "k_zpass<float> (z FFT + Green multiply + inverse z FFT)",
            "achieved": round(b_z / (z_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(b_z / (z_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
            "matvec_achieved_GBs": round(b_mv / (sum(p32) * 1e-3) / 1e9, 1),
            "pass_ms_c64": [round(x, 4) for x in p32], "pass_ms_c128": [round(x, 4) for x in p64]}

    cpu = None
    if not a.no_cpu:
        setup, solve, om = cpu_baseline(args, a.tol)
        cpu = {"value": round(1.0 / solve, 4), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"one FP64 oracle solve of the same system (numpy+scipy.fft, workers=1, "
                         f"{om.info.iterations} iterations, {solve:.1f} s; Green's setup {setup:.1f} s apart)"}

    value = a.steps / wall32
    out = {
        "metric": "DDA solves/sec at 1e-5 residual, FP32-refined vs FP64; p50/p99 ms",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall32 / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64-refine",
        "data": "synthetic",
        "config": {"workload": "periodic unit cell 32x32x12, d=10nm, FCD +-50 images, 600nm y-pol, "
                               "smooth density seed 1; one solve per step to 1e-5",
                   "fft_grid": [lx, ly, lz], "green_setup_ms": round(green_ms, 2), "model_build_s": round(build_s, 3)},
        "p50_ms": round(percentile(t32, 0.5), 3),
        "p99_ms": round(percentile(t32, 0.99), 3),
        "fp64": {"solves_per_s": round(a.steps / wall64, 3), "p50_ms": round(percentile(t64, 0.5), 3),
                 "p99_ms": round(percentile(t64, 0.99), 3), "iterations": i64[-1]["iterations"]},
        "refined": {"iterations": i32[-1]["iterations"], "reliable_updates": i32[-1]["reliable_updates"],
                    "error": i32[-1]["error"]},
        "speedup_vs_fp64": round((a.steps / wall32) / (a.steps / wall64), 3),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
