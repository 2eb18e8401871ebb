"""The Fortran ISO_C_BINDING layer (lpm-v2_amd/fortran) end to end on the GPU:
lpm_BVESphereVelocity over two MPISetup-style blocks, then one RK4 step of the
device-resident GPUBVESolver, against the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from helpers import cfg0_inputs, rel_err_max

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lpm-v2_amd", "fortran",
                   "lpm_fortran_smoke")


def test_fortran_drop_in(oracle, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"Fortran smoke program not built: {EXE}")
    m, zeta = cfg0_inputs(2)
    n = m.n
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([n], np.int32).tofile(f)
        for a in (m.x, m.y, m.z, zeta, m.area):
            np.ascontiguousarray(a, np.float64).tofile(f)
        np.ascontiguousarray(m.active, np.int32).tofile(f)
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ERROR" not in r.stdout
    raw = np.fromfile(out, np.float64)
    assert raw.size == 18 * n + 10
    res = raw[:7 * n].reshape(7, n)
    ref = oracle.bve_sphere_velocity(m.x, m.y, m.z, zeta, m.area, m.active)
    assert rel_err_max(tuple(res[:3]), ref) < 1e-11
    st = dict(x=m.x, y=m.y, z=m.z, relvort=zeta, u=ref[0], v=ref[1], w=ref[2])
    o = oracle.bve_timestep(st, m.area, m.active, 1.0, 2 * np.pi, 0.01)
    assert rel_err_max(tuple(res[3:6]), (o["x"], o["y"], o["z"])) < 1e-11
    assert rel_err_max(res[6], o["relvort"]) < 1e-11

    # PSE / SWE entry points through their bind(C) interfaces
    ops = raw[7 * n: 14 * n].reshape(7, n)
    g = oracle.pse_sphere_gradient(m.x, m.y, m.z, zeta, m.area, m.active, 0.3)
    assert rel_err_max(tuple(ops[:3]), g) < 1e-11
    swe = oracle.swe_plane_rhs(m.x, m.y, zeta, 0.1 * zeta, 1.0 + 0.1 * m.z, 0.05 * m.x, m.area, m.active, 0.3)
    for k in range(4):
        assert rel_err_max(ops[3 + k], swe[k]) < 1e-11, k
    pint = raw[14 * n: 14 * n + 10]
    ip = oracle.pse_sphere_interpolate(m.x, m.y, m.z, zeta, m.area, m.active, 0.3, m.x[:10], m.y[:10], m.z[:10])
    assert rel_err_max(pint, ip) < 1e-11
    # the reloaded solver's step equals the first solver's, bit for bit
    reload = raw[14 * n + 10:].reshape(4, n)
    assert np.array_equal(reload, res[3:7])
