"""CPU checks of the C ABI library and of the host-side argument handling."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "topodda.h")
LIB = os.path.join(PKG, "dda_model", "libtopodda.so")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|long|const char\*)\s+(tdda_\w+)\s*\(", text, re.M)))


def test_header_declares_model_entry_points():
    names = declared_functions()
    for want in ["tdda_model_create", "tdda_model_destroy", "tdda_model_solve_electric_field",
                 "tdda_model_calculate_objective", "tdda_model_calculate_gradients",
                 "tdda_model_set_parameters", "tdda_model_get_parameters",
                 "tdda_model_get_dielectrics", "tdda_model_get_electric_field", "tdda_last_error"]:
        assert want in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_codes_without_device():
    """Argument errors raised before any device work surface as codes and a
    message (reference: `throw 1` -> RuntimeError, invalid method -> ValueError)."""
    lib = ctypes.CDLL(LIB)
    lib.tdda_last_error.restype = ctypes.c_char_p
    lib.tdda_model_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    out = ctypes.c_void_p()
    assert lib.tdda_model_create(None, ctypes.byref(out)) == 1
    assert b"null" in lib.tdda_last_error()


def _args(method="LDR", symmetry="None", **kw):
    import dda_oracle as O
    nx, ny, nz = 3, 3, 2
    diel = np.full(3 * nx * ny * nz, 0.5)
    a = O.wrapper_args([nx, ny, nz], 15.0, diel, [1.0 + 0j, 3.0 + 0.1j], [0, 0, 1], [1, 0, 0], 542.0,
                       symmetry_axes=[1.0, 1.0], symmetry_method=symmetry, apc_max_m=0, apc_max_n=0,
                       apc_method_name=method, tables=(np.zeros(4), np.zeros(4), 0.1))
    return a


def test_pybind_constructor_errors_on_cpu():
    import dda_model
    with pytest.raises(ValueError, match="not a valid polarizabiliy"):
        dda_model.DDAModel(*_args(method="XYZ"))
    with pytest.raises(RuntimeError):
        dda_model.DDAModel(*_args(symmetry="8fold"))
    a = _args()
    a[13] = a[13][:-3]   # inputDielectrics shorter than geometry
    with pytest.raises(RuntimeError):
        dda_model.DDAModel(*a)


def test_constructor_keywords_match_reference():
    """Keyword names and order of python_bindings/PythonBindings.cpp:47-58."""
    import dda_model
    doc = dda_model.DDAModel.__init__.__doc__
    names = ["betaMin", "betaMax", "ita", "betaType", "filterIterations", "filterRadii", "useFilter",
             "symmetryType", "symmetryAxis", "isPeriodic", "objectiveName", "objPara", "geometry",
             "inputDielectrics", "num_x", "num_y", "num_z", "num_total", "n_K", "E0", "n_E0", "lam",
             "material", "nback", "MAXm", "MAXn", "Lm", "Ln", "matrixAMethodName", "d",
             "sineIntegralValues", "cosineIntegralValues", "integralDelta", "verbose"]
    pos = [doc.index(n + ":") for n in names]
    assert pos == sorted(pos)
    for m in ["calculateObjective", "calculateGradients", "solveElectricField", "setParameters",
              "getParameters", "getDielectrics", "getElectricField"]:
        assert hasattr(dda_model.DDAModel, m)


def test_generate_geometry_matches_oracle():
    import dda_model
    import dda_oracle as O
    assert np.array_equal(dda_model._generate_geometry(5, 3, 2), O.generate_geometry(5, 3, 2))
