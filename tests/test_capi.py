"""The C-ABI library loads and exports every function declared in include/lcsx.h (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lcsx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lcsx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_all_declared_symbols():
    from luisacomputesimulator_amd import LIB_PATH, _SIGS
    lib = ctypes.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_SIGS), set(names) ^ set(_SIGS)


def test_error_paths_without_gpu():
    from luisacomputesimulator_amd import NewtonSolver, LcsxError, scenes
    s = NewtonSolver(scenes.Params())
    with pytest.raises(LcsxError):
        s.physics_step_gpu(1)  # step before init -> LCSX_ERR_STATE
    with pytest.raises(LcsxError):
        s.register_mesh("bad", [[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 5]])  # index out of range
    with pytest.raises(LcsxError):
        s.init_solver()  # no meshes


def test_precision_parameter_is_validated():
    import ctypes as C
    from luisacomputesimulator_amd import LIB_PATH, lcsx_params, _SIGS
    from luisacomputesimulator_amd import NewtonSolver, scenes
    lib = C.CDLL(LIB_PATH)
    for n, (res, args) in _SIGS.items():
        getattr(lib, n).restype, getattr(lib, n).argtypes = res, args
    p = lcsx_params()
    lib.lcsx_params_default(C.byref(p))
    assert p.precision == 0
    p.precision = 3
    h = C.c_void_p()
    assert lib.lcsx_solver_create(C.byref(p), 0, C.byref(h)) != 0
    assert b"precision" in lib.lcsx_last_error()
    NewtonSolver(scenes.Params(precision="mixed"))  # accepted
    NewtonSolver(scenes.Params(precision="mixed_bf16"))
    with pytest.raises(KeyError):
        NewtonSolver(scenes.Params(precision="fp16"))
    lib.lcsx_params_default(C.byref(p))
    assert p.contact_energy_type == 1  # the reference's default: the log barrier
    p.contact_energy_type = 2
    assert lib.lcsx_solver_create(C.byref(p), 0, C.byref(h)) != 0
    assert b"contact_energy_type" in lib.lcsx_last_error()
    NewtonSolver(scenes.Params(contact_energy_type=0))  # accepted
