"""CPU checks of the C-ABI boundary: libubg.so loads, exports every entry
point include/ubg_c.h declares, and the ctypes binding covers them.  No
compute call is made here (no GPU in this container)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ubg_c.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"UBG_API\s+[^;(]*?\b(ubg_\w+)\s*\(", src)))


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_every_declared_symbol():
    from ubootgl_amd import _lib
    decl = declared_symbols()
    assert len(decl) >= 40
    exp = exported_symbols(_lib.LIB_PATH)
    missing = [s for s in decl if s not in exp]
    assert not missing, f"declared but not exported: {missing}"
    # nothing else with the ubg_ prefix leaks out
    extra = sorted(s for s in exp if s.startswith("ubg_") and s not in decl)
    assert not extra, f"exported but not declared: {extra}"


def test_ctypes_binding_matches_header():
    from ubootgl_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_symbols()
    L = _lib.lib()
    for name in _lib.SIGNATURES:
        assert getattr(L, name) is not None


def test_no_device_is_reported_not_faked():
    from ubootgl_amd import _lib
    import ubootgl_amd as U
    L = _lib.lib()
    assert L.ubg_version() >= 100
    if _lib.device_count() == 0:
        h = ctypes.c_void_p()
        rc = L.ubg_mg_create(16, 16, ctypes.byref(h))
        assert rc == _lib.E_NO_DEVICE
        with pytest.raises(U.UbgError):
            U.MG(16, 16)


def test_level_sizes_entry_point_matches_oracle():
    from ubootgl_amd import level_sizes
    from oracle import pyoracle as O
    for W, H in ((4096, 2048), (1090, 436), (8, 8), (7, 100), (513, 257), (8192, 4096)):
        assert level_sizes(W, H) == O.level_sizes(W, H)


def test_params_defaults_are_the_reference_values():
    from ubootgl_amd import _lib
    p = _lib.default_params()
    # pressure_solver.cpp:204,211,243 and simulation.cpp:189-190
    assert (p.nu_pre, p.nu_post, p.nu_coarse, p.vcycles_per_step, p.alpha) == (3, 3, 5, 2, 1.0)
