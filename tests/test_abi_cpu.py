"""C ABI checks that need no GPU: the library loads, exports every entry point
include/*.h declares, refuses to create a context without a device (with a
message), and the host-only helpers behave."""
import ctypes
import glob
import os
import re

import numpy as np

from lpmgpu import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\**\s*(lpm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("lpm_bve_sphere_velocity", "lpm_plane_velocity", "lpm_pse_plane_laplacian",
                 "lpm_bve_sphere_stream", "lpm_plane_stream", "lpm_bve_solver_timestep"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_every_export_is_bound_with_a_signature():
    unbound = [n for n in sorted(declared_functions()) if n not in L._SIGS]
    assert not unbound, unbound


def test_version():
    assert L.lib().lpm_version() >= 10000


def test_create_without_device_reports_error():
    lib = L.lib()
    h = ctypes.c_void_p()
    rc = lib.lpm_create(ctypes.byref(h), 0)
    if rc == L.LPM_OK:  # a GPU is present: nothing to check here
        lib.lpm_destroy(h)
        return
    assert rc in (L.LPM_ERR_HIP, L.LPM_ERR_ARG)
    assert h.value is None
    assert len(lib.lpm_last_error(None)) > 0


def test_null_context_is_rejected():
    lib = L.lib()
    x = np.zeros(4)
    act = np.ones(4, np.int32)
    rc = lib.lpm_bve_sphere_velocity(None, L.pd(x), L.pd(x), L.pd(x), L.pd(x), L.pd(x), L.pd(x), L.pd(x), L.pd(x),
                                     1.0, 0.0, L.pi(act), 4, 0, 4, 1)
    assert rc == L.LPM_ERR_ARG


def test_helpers_reject_bad_arguments():
    lib = L.lib()
    out = ctypes.c_double()
    assert lib.lpm_pse_eps(-1.0, 0.75, ctypes.byref(out)) == L.LPM_ERR_ARG
    s = np.zeros(2, np.int32)
    assert lib.lpm_load_balance(10, 0, L.pi(s), L.pi(s), L.pi(s)) == L.LPM_ERR_ARG
