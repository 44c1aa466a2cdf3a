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


def test_schedule_text_format_and_variants(tmp_path):
    """The kernel schedule through the C ABI (host only): parse, text, save,
    load, clear; variants exist only where the two split tables differ
    (Z at 40, 48, 64; X and Y at 80, 96, 128); a bad line leaves the schedule
    as it was."""
    from dda_model import native
    native.schedule_clear()
    try:
        native.schedule_set("# tuned\nvar c64 z 64 1\nvar c128 xfwd 128 1  # trailing comment\nyz 128 64 1\n"
                            "xy 96 0\n")
        t = native.schedule_text()
        assert "var c64 z 64 1" in t and "var c128 xfwd 128 1" in t and "yz 128 64 1" in t and "xy 96 0" in t
        path = tmp_path / "schedule.txt"
        native.schedule_save(path)
        native.schedule_clear()
        assert "var " not in native.schedule_text()
        native.schedule_load(path)
        assert native.schedule_text() == t
        for bad in ["var c64 z 32 1",          # 32: both tables hold 8 x 4
                    "var c64 z 64 2",          # no variant 2
                    "var c64 y 100 1",         # no split at all
                    "var c32 z 64 1", "var c64 w 64 1", "yz 128 64 3", "yz 128", "tune 1 2 3",
                    "xy 96 2", "xy 96", "xy 0 1"]:
            with pytest.raises(RuntimeError):
                native.schedule_set(bad)
            assert native.schedule_text() == t, bad
        for ok in ["var c64 z 40 1", "var c64 z 48 1", "var c128 xinv 80 1", "var c64 y 96 1", "var c64 xfwd 128 0"]:
            native.schedule_set(ok)
        with pytest.raises(RuntimeError):
            native.schedule_load(tmp_path / "missing.txt")
    finally:
        native.schedule_clear()
