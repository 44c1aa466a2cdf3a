"""dda_model: drop-in for the reference's Python package on the gfx950 solve path.

``DDAModel`` is the pybind11 class built in-tree (``_dda_model``), a thin layer
over the C ABI of ``include/topodda.h`` whose kernels run on the GPU.  There is
no CPU fallback: importing this package without the built extension fails.

Mirrors python_bindings/dda_model/__init__.py:1-203 (geometry order,
DDAModelWrapper defaults, Si/Ci tables, objective/gradients/parameters).
"""
from __future__ import annotations

import itertools

import numpy as np
from scipy.special import sici

try:
    # torch ships its own copy of the HIP runtime under the same soname; loading
    # it before libtopodda.so keeps one HIP runtime in the process (otherwise the
    # second copy to initialise finds no GPU).
    import torch  # noqa: F401
except ImportError:
    torch = None

from ._dda_model import DDAModel, device_bytes  # noqa: F401  (fails loudly if not built)

__all__ = ["DDAModel", "DDAModelWrapper", "_generate_geometry", "sici_tables"]


def _generate_geometry(num_x: int, num_y: int, num_z: int):
    """Dipole coordinates [x y z] flattened, x varying fastest
    (python_bindings/dda_model/__init__.py:8-20)."""
    cells = itertools.product(range(num_z), range(num_y), range(num_x))
    zyx = np.array(list(cells), dtype=int).reshape(-1, 3)
    return zyx[:, ::-1].reshape(-1).astype(int)


def sici_tables(delta: float = 0.1, count: int = 1_000_000):
    """Si/Ci samples as the reference wrapper builds them
    (python_bindings/dda_model/__init__.py:136-140): count-1 points from delta
    to delta*count; the lookup assumes abscissae (p+1)*delta."""
    positions = np.linspace(delta, delta * count, count - 1)
    si, ci = sici(positions)
    return si, ci, delta


_TABLES = None


def _cached_tables():
    global _TABLES
    if _TABLES is None:
        _TABLES = sici_tables()
    return _TABLES


class DDAModelWrapper:
    """Friendly constructor for DDAModel (python_bindings/dda_model/__init__.py:23-160)."""

    def __init__(
        self,
        num_pixels_xyz: list[int],
        pixel_size_nm: float,
        initial_parameter_values,
        dielectric_constants: list[complex],
        light_direction: list[float],
        light_polarization: list[float],
        light_wavelength_nm: float,
        light_amplitude: float = 1.0,
        symmetry_axes: list[float] = [],
        symmetry_method: str = "4fold",
        symmetry_is_periodic: bool = True,
        integral_power: int = 2,
        integral_xbounds: list[float] | None = None,
        integral_ybounds: list[float] | None = None,
        integral_zbounds: list[float] | None = None,
        background_refractive_index: float | None = None,
        objective_name: str = "IntegratedE",
        apc_max_m: int = 50,
        apc_max_n: int = 50,
        apc_l_m: int | None = None,
        apc_l_n: int | None = None,
        apc_method_name: str = "FCD",
        verbose: bool = False,
    ):
        if len(num_pixels_xyz) != 3:
            raise ValueError("num_pixels_xyz must contain the number of pixels "
                             f"in [x, y, z] but instead found {num_pixels_xyz}.")
        if len(symmetry_axes) != 2 and symmetry_method == "4fold":
            raise ValueError("If symmetry_method is 4fold, symmetry_axes must "
                             f"be 2-dimensional. Instead found {symmetry_axes}.")
        self._pixel_dimensions = num_pixels_xyz
        num_x, num_y, num_z = num_pixels_xyz
        num_total = int(np.prod(num_pixels_xyz))
        geometry = _generate_geometry(num_x, num_y, num_z)
        xb = integral_xbounds or [0.0, float(num_x) - 1]
        yb = integral_ybounds or [0.0, float(num_y) - 1]
        zb = integral_zbounds or [0.0, float(num_z) - 1]
        # power, bounds, then two filter values the objective never reads
        objective_config = [integral_power] + list(xb) + list(yb) + list(zb) + [0.95, 50.0]
        # image periods default to the box (x images step by apc_l_m, y by apc_l_n)
        apc_l_n = apc_l_n or num_x
        apc_l_m = apc_l_m or num_y
        si, ci, delta = _cached_tables()
        if not background_refractive_index:
            background_refractive_index = float(np.sqrt(np.real(dielectric_constants[0])))
        self._model = DDAModel(
            0.0, 50.0, 0.5, "piecewise", [100], [2.0], False,
            symmetry_method, list(symmetry_axes), symmetry_is_periodic,
            objective_name, objective_config,
            geometry, initial_parameter_values,
            num_x, num_y, num_z, num_total,
            light_direction, light_amplitude, light_polarization,
            light_wavelength_nm, dielectric_constants,
            background_refractive_index,
            apc_max_m, apc_max_n, apc_l_m, apc_l_n,
            apc_method_name, pixel_size_nm,
            si, ci, delta, verbose,
        )

    def objective(self, bgs_max_iter: int = 100_000, bgs_max_error: float = 1e-5):
        """Solve from P = 0, then the IntegratedE objective (__init__.py:162-174)."""
        n = int(np.prod(self._pixel_dimensions))
        self._model.solveElectricField(np.zeros(3 * n, dtype=np.complex128), bgs_max_iter, bgs_max_error)
        return self._model.calculateObjective()

    def gradients(self, current_objective_value: float, finite_difference_epsilon: float = 0.001,
                  bgs_max_iter: int = 100_000, bgs_max_error: float = 1e-5):
        """d objective / d parameters (__init__.py:176-188)."""
        return self._model.calculateGradients(finite_difference_epsilon, current_objective_value,
                                              bgs_max_iter, bgs_max_error)

    @property
    def parameters(self):
        return self._model.getParameters()

    @parameters.setter
    def parameters(self, value):
        self._model.setParameters(np.asarray(value, dtype=np.float64))

    def allParameters(self):
        return self._model.getDielectrics()

    def getElectricField(self):
        return self._model.getElectricField()

    # additions: solver precision and diagnostics
    def set_precision(self, precision: str):
        self._model.setSolverPrecision(precision)

    def solve_info(self, which: int = 0):
        return self._model.solveInfo(which)
