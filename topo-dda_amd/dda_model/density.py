"""Density parameterisation in front of the solve, as differentiable torch ops.

The reference keeps these inside the C++ model (DDAModel / FilterOption):

  * the free-parameter map with z extrusion and optional 4-fold mirror
    symmetry about two axes (src/DDAModel.cpp:37-64, 179-226);
  * the in-plane cone filter with weights r - |dxy| inside the circle
    |dxy|^2 <= r^2 - 0.01 of the same z layer, with the +-Lm/+-Ln images when
    periodic (src/DDAModel.cpp:66-77, 580-712), applied as
    filtered_i = sum_j w_ij origin_j / sum_j w_ij (src/DDAModel.cpp:802-811);
  * the beta projection SmoothDensity(x; beta, ita) and its derivative
    (src/FilterOption.cpp:48-62, src/DDAModel.cpp:340-353);
  * the beta schedules exp / piecewise / linear (src/Tools.cpp:473-516,
    src/FilterOption.cpp:23-46);
  * the filtered-gradient rule gradients_filtered (src/DDAModel.cpp:354-396).

Here every step is a torch function of float64 tensors on any device, so
autograd gives exact derivatives of the composed chain origin -> filtered ->
projected -> per-dipole density -> objective (dda_model.autograd).  The
reference's own gradient rule, which weights the gradient at parameter k by
sum_i w_ik / W_i and the projection slope at k (not the exact chain rule), is
kept as `gradients_filtered_reference` for like-for-like comparison.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _c_round(x: float) -> int:
    """C's round(): half away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def parameter_map(geometry, nz: int, symmetry: str = "None", symaxis=(0.0, 0.0)):
    """Free parameter of every dipole (src/DDAModel.cpp:37-64, 179-226).

    A dipole's (x, y) column is one parameter (z is extruded); with "4fold"
    the mirror images (2a-x, y), (2a-x, 2b-y), (x, 2b-y) of a column already
    seen share its parameter.  Returns (nfpara, gpara[N] int64)."""
    g = np.asarray(geometry, dtype=np.int64).reshape(-1, 3)
    if symmetry == "None":
        divide = 1
    elif symmetry == "4fold":
        divide = 4
    else:
        raise RuntimeError("SpacePara: symmetry must be None or 4fold")
    nfpara = int(g.shape[0] // nz // divide)   # geometry holds 3 N values (DDAModel.cpp:181)
    seen = {}
    pos = 0
    gpara = np.zeros(g.shape[0], dtype=np.int64)
    for i, (x, y, _) in enumerate(g):
        xy = (int(x), int(y))
        if xy not in seen:
            if symmetry == "None":
                seen[xy] = pos
                pos += 1
            else:
                mx = _c_round(2 * symaxis[0] - x)
                my = _c_round(2 * symaxis[1] - y)
                for s in ((mx, xy[1]), (mx, my), (xy[0], my)):
                    if s in seen:
                        seen[xy] = seen[s]
                        break
                else:
                    seen[xy] = pos
                    pos += 1
        gpara[i] = seen[xy]
    if pos > nfpara:
        raise RuntimeError("geometry has more free parameters than NFpara")
    return nfpara, gpara


def params_to_dipoles(params: torch.Tensor, gpara) -> torch.Tensor:
    """Per-dipole density from the free parameters (a gather; its adjoint sums
    the dipoles of a parameter, DDAModel.cpp:471-536)."""
    idx = torch.as_tensor(gpara, dtype=torch.long, device=params.device)
    return params.index_select(0, idx)


def filter_weights(geometry, gpara, nfpara: int, radius: float, periodic: bool = False, lm: int = 0,
                   ln: int = 0) -> np.ndarray:
    """Cone-filter weights between free parameters (src/DDAModel.cpp:580-712):
    for parameter i (its first dipole o) and every dipole r of parameter j in
    the same z layer, w = radius - |xy_o - xy_r| when |xy_o - xy_r|^2 <=
    radius^2 - 0.01; periodic structures also take the eight +-Lm, +-Ln
    images of r.  Entries accumulate per (i, j).  Returns a dense
    [nfpara, nfpara] array."""
    g = np.asarray(geometry, dtype=np.int64).reshape(-1, 3)
    gpara = np.asarray(gpara, dtype=np.int64)
    first = np.full(nfpara, -1, dtype=np.int64)
    for i in range(g.shape[0]):
        if first[gpara[i]] < 0:
            first[gpara[i]] = i
    shifts = [(0, 0)]
    if periodic:
        shifts += [(0, -ln), (-lm, -ln), (-lm, 0), (-lm, ln), (0, ln), (lm, ln), (lm, 0), (lm, -ln)]
    W = np.zeros((nfpara, nfpara))
    r2max = radius * radius - 0.01
    for i in range(nfpara):
        xo, yo, zo = g[first[i]]
        same = g[:, 2] == zo
        for sx, sy in shifts:
            dx = (g[:, 0] + sx - xo).astype(np.float64)
            dy = (g[:, 1] + sy - yo).astype(np.float64)
            d2 = dx * dx + dy * dy
            hit = same & (d2 <= r2max)
            np.add.at(W[i], gpara[hit], radius - np.sqrt(d2[hit]))
    return W


def cone_filter_params(origin: torch.Tensor, weights) -> torch.Tensor:
    """filtered_i = sum_j w_ij origin_j / sum_j w_ij (src/DDAModel.cpp:802-811)."""
    W = torch.as_tensor(weights, dtype=origin.dtype, device=origin.device)
    return (W @ origin) / W.sum(dim=1)


def smooth_density(x: torch.Tensor, beta: float, ita: float) -> torch.Tensor:
    """Beta projection of src/FilterOption.cpp:48-62 on [0, 1]:
    ita (exp(-beta (1 - x/ita)) - (1 - x/ita) exp(-beta)) for x <= ita and
    (1 - ita)(1 - exp(-beta (x - ita)/(1 - ita)) + (x - ita)/(1 - ita) exp(-beta)) + ita
    above; inputs outside [0, 1] are an error, as in the reference."""
    if bool(((x < 0) | (x > 1)).any()):
        raise ValueError("SmoothDensity: input out of range [0, 1]")
    eb = math.exp(-beta)
    lo = ita * (torch.exp(-beta * (1 - x / ita)) - (1 - x / ita) * eb)
    u = (x - ita) / (1 - ita)
    hi = (1 - ita) * (1 - torch.exp(-beta * u) + u * eb) + ita
    return torch.where(x <= ita, lo, hi)


def smooth_density_derivative(x, beta: float, ita: float):
    """PtoFderivative of src/DDAModel.cpp:340-353 (the derivative of
    smooth_density): beta exp(-beta (1 - x/ita)) + exp(-beta) below ita,
    beta exp(-beta (x - ita)/(1 - ita)) + exp(-beta) above."""
    x = torch.as_tensor(x, dtype=torch.float64)
    eb = math.exp(-beta)
    lo = beta * torch.exp(-beta * (1 - x / ita)) + eb
    hi = beta * torch.exp(-beta * (x - ita) / (1 - ita)) + eb
    return torch.where(x <= ita, lo, hi)


# ---------------------------------------------------------------------------
# beta schedules, src/Tools.cpp:473-516 (update_beta, FilterOption.cpp:23-46)
# ---------------------------------------------------------------------------
def beta_exp(it: float, it_max: float, lo: float, hi: float) -> float:
    """lo + (hi - lo)(100^(it/it_max) - 1)/99 for it <= 300, and 0.5 beyond
    (the reference's constant, kept)."""
    if it <= 300:
        return lo + (hi - lo) * (math.pow(100, it / it_max) - 1) / 99
    return 0.5


def beta_piecewise(it: float, it_max: float, lo: float, hi: float) -> float:
    """lo up to half the iterations, then lo + (hi-lo)/5 to 70 %, lo + (hi-lo)/2.5 to 80 %, hi."""
    if it <= 0.5 * it_max:
        return lo
    if it <= 0.7 * it_max:
        return lo + (hi - lo) / 5
    if it <= 0.8 * it_max:
        return lo + (hi - lo) / 2.5
    return hi


def beta_piecewise_absolute(it: float, it_max: float, lo: float, hi: float) -> float:
    """lo to iteration 100, lo + (hi-lo)/5 to 140, lo + (hi-lo)/2.5 to 160, then hi."""
    if it <= 100:
        return lo
    if it <= 140:
        return lo + (hi - lo) / 5
    if it <= 160:
        return lo + (hi - lo) / 2.5
    return hi


def beta_linear(it: float, it_max: float, lo: float, hi: float) -> float:
    return lo + (hi - lo) * it / it_max


BETA_SCHEDULES = {"exp": beta_exp, "piecewise": beta_piecewise, "linear": beta_linear}


def beta_at(kind: str, it: int, it_max: int, lo: float, hi: float) -> float:
    """update_beta (src/FilterOption.cpp:23-46); an unknown kind raises as the reference does."""
    if kind not in BETA_SCHEDULES:
        raise RuntimeError("ERROR: FilterOption::update_beta: beta_type not defined")
    return BETA_SCHEDULES[kind](it, it_max, lo, hi)


def gradients_filtered_reference(gradients, filtered, weights, beta: float, ita: float) -> np.ndarray:
    """The reference's gradient through filter and projection
    (src/DDAModel.cpp:354-396): result_k = (sum_i w_ik / W_i) *
    PtoF'(filtered_k) * gradients_k with W_i = sum_j w_ij."""
    W = np.asarray(weights, dtype=np.float64)
    col = (W / W.sum(axis=1, keepdims=True)).sum(axis=0)
    slope = smooth_density_derivative(torch.as_tensor(filtered), beta, ita).numpy()
    return col * slope * np.asarray(gradients, dtype=np.float64)


class DensityPipeline(torch.nn.Module):
    """origin parameters -> cone filter -> beta projection -> per-dipole
    density, differentiable end to end (the chain the reference applies in
    UpdateStr, src/DDAModel.cpp:774-837, with the exact derivative)."""

    def __init__(self, geometry, nz, symmetry="None", symaxis=(0.0, 0.0), radius=None, periodic=False, lm=0,
                 ln=0, ita=0.5):
        super().__init__()
        self.nfpara, self.gpara = parameter_map(geometry, nz, symmetry, symaxis)
        self.weights = None
        if radius is not None:
            self.weights = filter_weights(geometry, self.gpara, self.nfpara, radius, periodic, lm, ln)
        self.ita = ita

    def forward(self, origin: torch.Tensor, beta=None) -> torch.Tensor:
        x = origin
        if self.weights is not None:
            x = cone_filter_params(x, self.weights)
        if beta is not None:
            x = smooth_density(x, beta, self.ita)
        return params_to_dipoles(x, self.gpara)
