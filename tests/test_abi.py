"""The C ABI library loads on the CPU host and exports every symbol declared
in include/*.h (no compute calls: there may be no GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("ostr_hip.h", "ostr_cnn.h")]
LIB = os.path.join(ROOT, "newrecon_amd", "libostr_hip.so")


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(ostr_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_surface():
    names = declared_functions()
    for required in ("ostr_create", "ostr_destroy", "ostr_last_error", "ostr_fproj",
                     "ostr_gamma", "ostr_bproj_sbp", "ostr_prior", "ostr_update_est",
                     "ostr_fproj_h", "ostr_update_fused", "ostr_precompute_d",
                     "ostr_batch_fproj_h", "ostr_batch_update_fused", "ostr_cnn_conv3d",
                     "ostr_cnn_wgrad", "ostr_cnn_adam"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libostr_hip.so not built")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and nothing else leaks from the C++ side
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True)
    exported = {l.split()[-1] for l in out.stdout.splitlines() if " T " in l}
    extra = sorted(e for e in exported if e.startswith("ostr_") and e not in declared_functions())
    assert not extra, extra


def test_error_path_without_device():
    """Argument validation happens before any device call."""
    if not os.path.exists(LIB):
        pytest.skip("libostr_hip.so not built")
    lib = ctypes.CDLL(LIB)
    lib.ostr_create.restype = ctypes.c_int
    lib.ostr_last_error.restype = ctypes.c_char_p
    lib.ostr_last_error.argtypes = [ctypes.c_void_p]
    h = ctypes.c_void_p()
    rc = lib.ostr_create(None, None, None, 0, ctypes.byref(h))
    assert rc == -1 and h.value is None
    assert b"geometry" in lib.ostr_last_error(None)


def test_volume_limit_is_checked_before_any_device_call():
    """nx ny nz + 2 nx ny + 2 nx + 2 must stay <= 2^30 (32-bit byte offsets
    with a two-plane margin for the walk's out-of-volume indices)."""
    if not os.path.exists(LIB):
        pytest.skip("libostr_hip.so not built")
    import numpy as np
    from newrecon_amd.geometry import Geometry
    lib = ctypes.CDLL(LIB)
    lib.ostr_create.restype = ctypes.c_int
    lib.ostr_last_error.restype = ctypes.c_char_p
    lib.ostr_last_error.argtypes = [ctypes.c_void_p]
    angles = np.zeros(1, np.float32)
    # 32767 x 32767 x 1 voxels: below 2^30, but not with the two-plane margin
    g = Geometry.half_cone(32767, 32767, 1, 0.1, 0.1, 1.0, 256, 128, 0.5, 0.5)
    h = ctypes.c_void_p()
    rc = lib.ostr_create(ctypes.byref(g.to_c()), angles.ctypes.data_as(ctypes.c_void_p), None,
                         1, ctypes.byref(h))
    assert rc == -1 and h.value is None
    assert b"volume too large" in lib.ostr_last_error(None)
