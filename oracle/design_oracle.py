"""CPU restatement of the reference's design-side arithmetic -- TEST
INFRASTRUCTURE ONLY (tests/ use it as the checker of dda_model.density,
dda_model.design and dda_model.materials; the product never imports it).

Literal loops over the reference's own data structures, one function per
reference routine:
  * filter weights (FreeWeight lists) ........ src/DDAModel.cpp:66-77, 580-712
  * filtered parameters ....................... src/DDAModel.cpp:802-811
  * SmoothDensity / PtoFderivative ............ src/FilterOption.cpp:48-62, src/DDAModel.cpp:340-353
  * gradients_filtered ........................ src/DDAModel.cpp:354-396
  * exp / piecewise / linear beta updates ..... src/Tools.cpp:473-516
  * Adam, AdaGrad ............................. python_bindings/dda_model/optimizers.py:3-33
  * penalties and coefficient schedules ....... python_bindings/dda_model/binarizers.py:5-93
  * Get_material table interpolation .......... src/Tools.cpp:249-321

Pins (tests/test_design_oracle.py): the derivative restated equals a central
difference of SmoothDensity; filter weights equal r - distance on hand-built
pairs; the schedules hit their closed-form end points; the Adam step of a
constant gradient is the textbook bias-corrected value at t = 1; the table
interpolation reproduces a linear function exactly.
"""
from __future__ import annotations

import math


def calweight(xo, yo, x, y, r):
    return r - math.sqrt(float(x - xo) * float(x - xo) + float(y - yo) * float(y - yo))


def circlerange(xo, yo, x, y, r):
    return float(x - xo) * float(x - xo) + float(y - yo) * float(y - yo) <= r * r - 0.01


def free_weights(geometry, para_to_geometry, rfilter, periodic, lm, ln):
    """FreeWeight[i] = [(weight, j), ...] exactly as assignFreeWeightsForFilter."""
    n = len(para_to_geometry)
    out = [[] for _ in range(n)]
    shifts = [(0, 0)]
    if periodic:
        shifts += [(0, -ln), (-lm, -ln), (-lm, 0), (-lm, ln), (0, ln), (lm, ln), (lm, 0), (lm, -ln)]
    for i in range(n):
        po = para_to_geometry[i][0]
        xo, yo, zo = geometry[3 * po], geometry[3 * po + 1], geometry[3 * po + 2]
        for j in range(n):
            for pr in para_to_geometry[j]:
                xr, yr, zr = geometry[3 * pr], geometry[3 * pr + 1], geometry[3 * pr + 2]
                for sx, sy in shifts:
                    if zo == zr and circlerange(xo, yo, xr + sx, yr + sy, rfilter):
                        out[i].append((calweight(xo, yo, xr + sx, yr + sy, rfilter), j))
    return out


def filtered_parameters(free_weight, origin):
    res = []
    for row in free_weight:
        num = sum(w * origin[j] for w, j in row)
        den = sum(w for w, _ in row)
        res.append(num / den)
    return res


def smooth_density(x, beta, ita):
    if 0.0 <= x <= ita:
        return ita * (math.exp(-beta * (1 - x / ita)) - (1 - x / ita) * math.exp(-beta))
    if ita < x <= 1.0:
        return (1 - ita) * (1 - math.exp(-beta * (x - ita) / (1 - ita)) + (x - ita) / (1 - ita) * math.exp(-beta)) + ita
    raise RuntimeError("SmoothDensity: input out of range")


def ptof_derivative(x, beta, ita):
    if 0.0 <= x <= ita:
        return beta * math.exp(-beta * (1 - x / ita)) + math.exp(-beta)
    if ita < x <= 1.0:
        return beta * math.exp(-beta * (x - ita) / (1 - ita)) + math.exp(-beta)
    raise RuntimeError("PtoFderivative: input out of range")


def gradients_filtered(gradients, para_filtered, free_weight, beta, ita):
    n = len(gradients)
    result = [0.0] * n
    for i in range(n):
        s = sum(w for w, _ in free_weight[i])
        for w, pos in free_weight[i]:
            if pos < n:
                result[pos] += w / s
    for i in range(n):
        result[i] = result[i] * ptof_derivative(para_filtered[i], beta, ita) * gradients[i]
    return result


def exp_update(x, x_max, lo, hi):
    if x <= 300:
        return lo + (hi - lo) * (math.pow(100, x / x_max) - 1) / (100 - 1)
    return 0.5


def piecewise_update(x, x_max, lo, hi):
    if x <= 0.5 * x_max:
        return lo
    elif 0.5 * x_max < x <= 0.7 * x_max:
        return lo + (hi - lo) / 5
    elif 0.7 * x_max < x <= 0.8 * x_max:
        return lo + (hi - lo) / 2.5
    return hi


def linear_update(x, x_max, lo, hi):
    return lo + (hi - lo) * x / x_max


class Adam:
    def __init__(self, beta1=0.9, beta2=0.99):
        self.b1, self.b2, self.V, self.S, self.t = beta1, beta2, None, None, 0

    def __call__(self, g):
        n = len(g)
        if self.V is None:
            self.V, self.S = [0.0] * n, [0.0] * n
        out = []
        for i in range(n):
            self.V[i] = self.b1 * self.V[i] + (1 - self.b1) * g[i] / (1 - self.b1 ** (self.t + 1))
            self.S[i] = self.b2 * self.S[i] + (1 - self.b2) * g[i] ** 2 / (1 - self.b2 ** (self.t + 1))
            out.append(self.V[i] / (math.sqrt(self.S[i]) + 1e-8))
        self.t += 1
        return out


def penalty(kind, p, cfg):
    out = []
    for x in p:
        if kind == "parabolic":
            out.append(x - x * x)
        elif kind == "gaussian":
            out.append(cfg["sigma"] * math.exp(-((x - cfg["mu"]) / cfg["sigma"]) ** 2 / 2 + 0.5))
        else:
            out.append(cfg["slope"] - x * cfg["slope"] if x > 0.5 else x * cfg["slope"])
    return out


def penalty_gradient(kind, p, cfg):
    out = []
    for x in p:
        if kind == "parabolic":
            out.append(1 - 2 * x)
        elif kind == "gaussian":
            g = cfg["sigma"] * math.exp(-((x - cfg["mu"]) / cfg["sigma"]) ** 2 / 2 + 0.5)
            out.append(-(x - cfg["mu"]) * g / cfg["sigma"] ** 2)
        else:
            out.append(-cfg["slope"] if x > 0.5 else cfg["slope"])
    return out


def get_material_from_rows(re_rows, im_rows, wl):
    """Get_material's scan over one table (Tools.cpp:286-309), for both parts."""
    def scan(rows):
        up, down, up_v, down_v = 1.0, 0.0, None, None
        for a, b in rows:
            if a <= wl and a > down:
                down, down_v = a, b
            elif a > wl and a < up:
                up, up_v = a, b
        if up == 1.0 or down == 0.0:
            raise ValueError("wavelength out of range")
        return down_v + (wl - down) * (up_v - down_v) / (up - down)
    return complex(scan(re_rows), scan(im_rows))
