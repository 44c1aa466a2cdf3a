"""Wavelength-dependent permittivities for the solve (Get_material /
Get_2_material / Get_X_material, src/Tools.cpp:249-351).

The reference reads two tables per material, "<name>Re_eps.txt" and
"<name>Im_eps.txt", with lines "wavelength_in_metres value", and interpolates
linearly between the largest tabulated wavelength <= wl and the smallest one
> wl; a wavelength outside the table is an error.  Its table files are not
part of the source, so this module keeps that reader (`TableMaterial`, same
file format and rule, with the error raised instead of printed) and adds
closed-form models where a published formula fixes the permittivity:
  * "Air": eps = 1;
  * "SiO2": Sellmeier fit of fused silica, Malitson, J. Opt. Soc. Am. 55,
    1205 (1965), n^2 = 1 + sum_i B_i l^2 / (l^2 - C_i^2), l in um;
  * "Al2O3": Sellmeier fit of sapphire (ordinary ray), Malitson & Dodge,
    J. Opt. Soc. Am. 62, 1405 (1972);
  * constant dielectrics given by their permittivity.
Use `permittivity(name, wl, unit)` for one material and `materials(names,
wl, unit)` for the (background, structure) pair a request takes.
"""
from __future__ import annotations

import os

import numpy as np

UNIT = {"nm": 1.0e9, "um": 1.0e6, "m": 1.0}   # units per metre (Tools.cpp:251-254)

# Sellmeier coefficients (B_i, C_i in um)
_SELLMEIER = {
    "SiO2": ((0.6961663, 0.0684043), (0.4079426, 0.1162414), (0.8974794, 9.896161)),
    "Al2O3": ((1.4313493, 0.0726631), (0.65054713, 0.1193242), (5.3414021, 18.028251)),
}


def sellmeier_eps(name: str, wl_m: float) -> complex:
    lu = wl_m * 1.0e6
    n2 = 1.0 + sum(b * lu * lu / (lu * lu - c * c) for b, c in _SELLMEIER[name])
    return complex(n2, 0.0)


def _bracket(table: np.ndarray, wl: float) -> float:
    """Tools.cpp:286-309: the nearest sample at or below wl and the nearest
    one above it, interpolated linearly; both must exist."""
    a, b = table[:, 0], table[:, 1]
    below = np.where(a <= wl)[0]
    above = np.where(a > wl)[0]
    if below.size == 0 or above.size == 0:
        raise ValueError(f"wavelength {wl} m out of the range of the permittivity table")
    i = below[np.argmax(a[below])]
    j = above[np.argmin(a[above])]
    return float(b[i] + (wl - a[i]) * (b[j] - b[i]) / (a[j] - a[i]))


class TableMaterial:
    """A material given by Re/Im permittivity tables in the reference's format."""

    def __init__(self, prefix: str):
        self.re = np.loadtxt(prefix + "Re_eps.txt", ndmin=2)
        self.im = np.loadtxt(prefix + "Im_eps.txt", ndmin=2)

    def __call__(self, wl_m: float) -> complex:
        return complex(_bracket(self.re, wl_m), _bracket(self.im, wl_m))


_TABLES: dict = {}


def register_table(name: str, prefix: str):
    """Make `name` resolve to the tables prefix + "Re_eps.txt" / "Im_eps.txt"."""
    _TABLES[name] = TableMaterial(prefix)


def permittivity(name, wl: float, unit: str = "nm") -> complex:
    """Permittivity of `name` at wavelength wl (Get_material, Tools.cpp:249-321).
    A number is a constant permittivity."""
    if unit not in UNIT:
        raise ValueError(f"unit must be one of {sorted(UNIT)}")
    wl_m = wl / UNIT[unit]
    if isinstance(name, (int, float, complex)):
        return complex(name)
    if name in _TABLES:
        return _TABLES[name](wl_m)
    if name == "Air":
        return complex(1.0, 0.0)
    if name in _SELLMEIER:
        return sellmeier_eps(name, wl_m)
    env = os.environ.get("TDDA_DIEL_DIR")
    if env and os.path.exists(os.path.join(env, name + "Re_eps.txt")):
        register_table(name, os.path.join(env, name))
        return _TABLES[name](wl_m)
    raise ValueError(f"no permittivity data for material {name!r}")


def materials(names, wl: float, unit: str = "nm"):
    """Get_X_material (Tools.cpp:331-351): the permittivity of each name."""
    return [permittivity(n, wl, unit) for n in names]
