"""The C ABI library loads and exports every entry point include/sph_hip.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sph_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_trait_entry_points():
    names = declared_functions()
    for required in ("sph_neighbors_find", "sph_pressure_solve", "sph_boundary_add_volume_map",
                     "sph_boundary_find_samples", "sph_volume", "sph_nonpressure_accel",
                     "sph_step", "sph_ctx_create", "sph_ctx_clone", "sph_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from sph_amd import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    names = declared_functions()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _native.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_last_error_is_callable_without_a_device():
    from sph_amd import _native
    lib = _native.lib()
    assert lib.sph_last_error() is not None
