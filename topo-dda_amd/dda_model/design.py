"""The design loop around the solve: optimizers, binarisation penalties,
their coefficient schedules, the on-disk outputs, and a driver.

Restates (same formulas, own code):
  * AdamOptimizer / AdaGradOptimizer of python_bindings/dda_model/optimizers.py:3-33
    (the bias corrections sit inside the moving averages there, and are kept);
  * the penalties parabolic / gaussian / triangular and their gradients, and
    the coefficient schedules piecewise / piecewise_absolute / exp / linear,
    of python_bindings/dda_model/binarizers.py:5-93;
  * the loop and output files of calculation_scripts/TopoOptCalculation.py:
    objective -> gradients -> penalty -> coefficient -> optimizer step ->
    clip to [0, 1] -> setParameters, saving Structures/StructureN.txt,
    E-Fields/E-FieldN.txt, Parameters/ParamN.txt, Penalties/PenaltyN.txt,
    Gradient_Penalties/GradientN.txt, Obj.txt, Coeffs.txt,
    Average_Penalties.txt, Penalty_Shape.txt, PenaltyGradient_Shape.txt and
    stepsizes.txt as numpy text files.
The model is any object with the DDAModelWrapper interface (objective,
gradients, parameters, allParameters, getElectricField), so the loop runs on
the GPU model unchanged.
"""
from __future__ import annotations

import math
import os
from pathlib import Path

import numpy as np


class AdamOptimizer:
    """optimizers.py:3-23: V <- b1 V + (1 - b1) g / (1 - b1^(t+1)),
    S <- b2 S + (1 - b2) g^2 / (1 - b2^(t+1)), step = V / (sqrt(S) + 1e-8)."""

    def __init__(self, beta1=0.9, beta2=0.99):
        self.beta1, self.beta2 = beta1, beta2
        self.V = 0.0
        self.S = 0.0
        self.timestep = 0

    def __call__(self, gradients):
        g = np.asarray(gradients, dtype=np.float64)
        t = self.timestep + 1
        self.V = self.beta1 * self.V + (1 - self.beta1) * g / (1 - self.beta1 ** t)
        self.S = self.beta2 * self.S + (1 - self.beta2) * g * g / (1 - self.beta2 ** t)
        self.timestep += 1
        return self.V / (np.sqrt(self.S) + 1e-8)


class AdaGradOptimizer:
    """optimizers.py:26-33: step = g / (sum of g^2 so far + 1)."""

    def __init__(self):
        self.squared = 0.0

    def __call__(self, gradients):
        g = np.asarray(gradients, dtype=np.float64)
        self.squared = self.squared + g * g
        return g / (self.squared + 1)


# ---------------------------------------------------------------------------
# penalties (binarizers.py:5-31)
# ---------------------------------------------------------------------------
def parabolic(p, cfg=None):
    return p - np.square(p)


def grad_parabolic(p, cfg=None):
    return 1 - 2 * p


def gaussian(p, cfg):
    s, mu = cfg["sigma"], cfg["mu"]
    return s * np.exp(-np.power((p - mu) / s, 2) / 2 + 0.5)


def grad_gaussian(p, cfg):
    s, mu = cfg["sigma"], cfg["mu"]
    return -(p - mu) * gaussian(p, cfg) / np.square(s)


def triangular(p, cfg):
    slope = cfg["slope"]
    out = np.array(p, dtype=np.float64) * slope
    hi = p > 0.5
    out[hi] = slope - out[hi]
    return out


def grad_triangular(p, cfg):
    out = np.ones_like(np.asarray(p, dtype=np.float64)) * cfg["slope"]
    out[p > 0.5] *= -1
    return out


PENALTIES = {"parabolic": (parabolic, grad_parabolic), "gaussian": (gaussian, grad_gaussian),
             "triangular": (triangular, grad_triangular)}


# ---------------------------------------------------------------------------
# coefficient schedules (binarizers.py:33-93)
# ---------------------------------------------------------------------------
def coeff_piecewise_absolute(it, c):
    if it <= c["iter1"]:
        return c["coeff1"]
    if it <= c["iter2"]:
        return c["coeff2"]
    if it <= c["iter3"]:
        return c["coeff3"]
    return c["coeff4"]


def coeff_piecewise(it, c):
    end = c["iter_end"]
    if it <= c["frac1"] * end:
        return c["coeff1"]
    if it <= c["frac2"] * end:
        return c["coeff2"]
    if it <= c["frac3"] * end:
        return c["coeff3"]
    return c["coeff4"]


def coeff_exp(it, c):
    if it < c["iter_start"]:
        return c["coeff_min"]
    if it > c["iter_end"]:
        return c["coeff_max"]
    frac = (it - c["iter_start"]) / (c["iter_end"] - c["iter_start"])
    return c["coeff_min"] + (c["coeff_max"] - c["coeff_min"]) * (math.pow(c["base"], frac) - 1) / (c["base"] - 1)


def coeff_linear(it, c):
    if it < c["iter_start"]:
        return c["coeff_min"]
    if it > c["iter_end"]:
        return c["coeff_max"]
    return c["coeff_min"] + (c["coeff_max"] - c["coeff_min"]) * it / c["iter_end"]


COEFFS = {"piecewise": coeff_piecewise, "piecewise_absolute": coeff_piecewise_absolute, "exp": coeff_exp,
          "linear": coeff_linear, "constant": lambda it, c: c["coeff_const"]}


# ---------------------------------------------------------------------------
# on-disk formats
# ---------------------------------------------------------------------------
DIRECTORIES = ("Structures", "Parameters", "Penalties", "Gradient_Penalties", "E-Fields")


def write_geometry(path, nx, ny, nz, geometry):
    """geometry.txt of python_tests/data: Nx, Ny, Nz, N, then the 3N dipole
    coordinates, one integer per line."""
    g = np.asarray(geometry, dtype=np.int64).reshape(-1)
    with open(path, "w") as f:
        f.write(f"{nx}\n{ny}\n{nz}\n{g.size // 3}\n")
        f.write("\n".join(str(int(v)) for v in g))
        f.write("\n")


def read_geometry(path):
    """Inverse of write_geometry: ((nx, ny, nz), geometry[3N])."""
    v = np.loadtxt(path, dtype=np.int64)
    nx, ny, nz, n = (int(x) for x in v[:4])
    g = v[4:]
    if g.size != 3 * n:
        raise ValueError(f"{path}: {g.size} coordinates for {n} dipoles")
    return (nx, ny, nz), g


def write_dielectrics(path, values):
    """dielectrics.txt: one density per line (3 per dipole)."""
    np.savetxt(path, np.asarray(values, dtype=np.float64), delimiter="\n")


def read_dielectrics(path):
    return np.loadtxt(path, dtype=np.float64)


def _save(path, values):
    np.savetxt(path, np.asarray(values), delimiter="\n")


def run_design(model, iterations, step_size, penalty_type, penalty_config, coeff_type, coeff_config, out_dir,
               optimizer=None, save_every=1):
    """The loop of calculation_scripts/TopoOptCalculation.py:165-215 on any
    DDAModelWrapper-like model.  Returns the objective history."""
    data = Path(out_dir)
    for d in DIRECTORIES:
        (data / d).mkdir(parents=True, exist_ok=True)
    pen, grad_pen = PENALTIES[penalty_type]
    coeff_of = COEFFS[coeff_type]
    opt = optimizer or AdamOptimizer()
    objs, coeffs, avg_pen, steps = [], [], [], []
    for it in range(iterations):
        obj = model.objective()
        objs.append(obj)
        g_obj = np.asarray(model.gradients(obj))
        params = np.asarray(model.parameters)
        g_pen = grad_pen(params, penalty_config)
        p = pen(params, penalty_config)
        if it % save_every == 0:
            _save(data / "Structures" / f"Structure{it}.txt", model.allParameters())
            _save(data / "E-Fields" / f"E-Field{it}.txt", model.getElectricField())
            _save(data / "Parameters" / f"Param{it}.txt", params)
            _save(data / "Gradient_Penalties" / f"Gradient{it}.txt", g_pen)
            _save(data / "Penalties" / f"Penalty{it}.txt", p)
        avg_pen.append(float(np.average(p)))
        c = coeff_of(it, coeff_config)
        coeffs.append(c)
        g = (1 - c) * g_obj - c * g_pen
        steps.append(step_size)
        model.parameters = np.clip(params + step_size * opt(g), 0, 1)
    _save(data / "Obj.txt", objs)
    _save(data / "Coeffs.txt", coeffs)
    _save(data / "Average_Penalties.txt", avg_pen)
    _save(data / "stepsizes.txt", steps)
    x = np.linspace(0, 1, 200)
    np.savetxt(data / "Penalty_Shape.txt", np.column_stack((x, pen(x, penalty_config))), delimiter=",")
    np.savetxt(data / "PenaltyGradient_Shape.txt", np.column_stack((x, grad_pen(x, penalty_config))),
               delimiter=",")
    return objs
