"""The C ABI library loads and exports every symbol include/mlh.h declares
(no compute calls: no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "mlh.h")) as fh:
        txt = fh.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mlh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_all_declared_symbols():
    from meshlesshydro_amd import lib
    if not os.path.exists(lib.LIB_PATH):
        pytest.fail("libmlh_hip.so is not built (python -m meshlesshydro_amd.build)")
    so = ctypes.CDLL(lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) > 20
    missing = [s for s in syms if not hasattr(so, s)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert sorted(lib.SIGNATURES) == syms


def test_default_params_match_python_mirror():
    from meshlesshydro_amd import lib
    from meshlesshydro_amd._abi import default_params
    c = lib.default_params_c()
    py = default_params()
    for name, _ in c._fields_:
        a, b = getattr(c, name), getattr(py, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name
    assert lib.lib().mlh_version() == 2


def test_status_codes_match_reference_exit_codes():
    from meshlesshydro_amd import _abi
    # Particles.cpp:351 exit(1), :2252 exit(3), :1824-1827/1878-1881 exit(6),
    # MeshlessScheme.cpp:95-97 exit(9)
    assert _abi.MLH_ERR_NEIGHBOUR_OVERFLOW == 1
    assert _abi.MLH_ERR_GHOST_OVERFLOW == 3
    assert _abi.MLH_ERR_NEGATIVE_FACE_PRESSURE == 6
    assert _abi.MLH_ERR_DUMP_OVERRUN == 9


def test_create_without_device_fails_cleanly():
    """On a machine without a GPU mlh_create reports MLH_ERR_DEVICE (no CPU
    fallback); with a GPU it succeeds."""
    from meshlesshydro_amd import lib, _abi, ics
    s = ics.kelvin_helmholtz(16)
    p = _abi.default_params(kernel_size=0.2)
    try:
        sim = lib.Simulation(p, s)
    except lib.MLHError as e:
        assert e.status == _abi.MLH_ERR_DEVICE
    else:
        sim.close()


def test_create_rejects_bad_arguments_before_any_device_call():
    """mlh_create validates its arguments before it touches the device, so
    these run anywhere: each bad argument gives MLH_ERR_INVALID_ARGUMENT
    with its own message and leaves the context pointer null."""
    from meshlesshydro_amd import lib, _abi
    L = lib.lib()
    cases = [
        (dict(), 0, "n <= 0"),
        (dict(), -5, "n <= 0"),
        (dict(dim=4), 100, "dim must be 2 or 3"),
        (dict(), 1 << 27, "n exceeds 2^27-1"),
        (dict(max_interactions=0), 100, "bad interaction capacities"),
        (dict(max_ghost_interactions=-1), 100, "bad interaction capacities"),
        (dict(precision=7), 100, "bad precision"),
        (dict(work_order=9), 100, "bad work_order"),
    ]
    for kw, n, msg in cases:
        p = _abi.default_params(kernel_size=0.1, **kw)
        h = ctypes.c_void_p()
        st = L.mlh_create(ctypes.byref(p), n, ctypes.byref(h))
        assert st == _abi.MLH_ERR_INVALID_ARGUMENT, (kw, n, st)
        assert msg in L.mlh_last_error().decode(), (kw, n)
        assert not h.value
