"""The Fortran ISO_C_BINDING layer (lpm-v2_amd/fortran) end to end on the GPU:
lpm_BVESphereVelocity over two MPISetup-style blocks, one RK4 step of the
device-resident GPUBVESolver, the PSE and SWE entry points, a solver reload,
the planar solver, a batch, the learned far-field object and the helpers,
against the FP64 oracle (and, for the far-field object, the same call made
from Python)."""
import os
import subprocess

import numpy as np
import pytest

from helpers import icos, rel_err_max
from lpmgpu import ic
from lpmgpu.farfield import EXPANSION, LEARNED, FarFieldBVE

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lpm-v2_amd", "fortran",
                   "lpm_fortran_smoke")


def test_fortran_drop_in(ctx, oracle, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"Fortran smoke program not built: {EXE}")
    m = icos(2, 0)  # 20 coarse cells for the far-field object
    zeta = ic.gaussian_vortex(m.x, m.y, m.z, ic.lat_lon_point(0.3, 0.0), 4.0 * np.pi, 16.0)
    n = m.n
    cell = FarFieldBVE.particle_cells(m)
    ncell = int(np.asarray(m.face_cell).max()) + 1
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([n], np.int32).tofile(f)
        for a in (m.x, m.y, m.z, zeta, m.area):
            np.ascontiguousarray(a, np.float64).tofile(f)
        np.ascontiguousarray(m.active, np.int32).tofile(f)
        np.array([ncell], np.int32).tofile(f)
        np.ascontiguousarray(cell, np.int32).tofile(f)
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ERROR" not in r.stdout
    raw = np.fromfile(out, np.float64)
    assert raw.size == 29 * n + 27
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
    reload = raw[14 * n + 10: 18 * n + 10].reshape(4, n)
    assert np.array_equal(reload, res[3:7])

    # planar solver: one FP64 RK4 step on (x + 0.1 z, y + 0.37 z) and its diagnostics
    k = 18 * n + 10
    plane = raw[k: k + 4 * n].reshape(4, n)
    pdiag = raw[k + 4 * n: k + 4 * n + 3]
    px, py = m.x + 0.1 * m.z, m.y + 0.37 * m.z
    pu, pv = oracle.plane_velocity(px, py, zeta, m.area, m.active)
    po = oracle.plane_timestep(dict(x=px, y=py, u=pu, v=pv), zeta, m.area, m.active, 0.01)
    assert rel_err_max(tuple(plane[:2]), (po["x"], po["y"])) < 1e-11
    assert rel_err_max(tuple(plane[2:]), (po["u"], po["v"])) < 1e-11
    want = oracle.diagnostics(po["u"], po["v"], None, zeta, m.area, m.active)
    assert np.allclose(pdiag, want, rtol=1e-11, atol=1e-14), (pdiag, want)

    # a batch holding the sphere problem steps like the solver
    k += 4 * n + 3
    batch = raw[k: k + 4 * n].reshape(4, n)
    assert rel_err_max(tuple(batch[:3]), tuple(res[3:6])) < 1e-13
    assert rel_err_max(batch[3], res[6]) < 1e-13

    # the learned far-field object in expansion mode, as from Python
    k += 4 * n
    ffv = raw[k: k + 3 * n].reshape(3, n)
    f = FarFieldBVE(ctx, m, zeta, rotation_rate=2 * np.pi, near_radius=0.3, band_radius=1.5, order=4)
    want = f.velocity(EXPANSION)
    f.delete()
    assert rel_err_max(tuple(ffv), want) < 1e-12
    assert rel_err_max(tuple(ffv), ref) < 1e-2  # the expansion approximates the direct sum

    # its training and judged learned-mode steps, driven from Fortran
    k += 3 * n
    total, loss, mode, last_error, judged = raw[k: k + 5]
    assert total == 20000 and np.isfinite(loss) and loss > 0
    assert mode == LEARNED and judged == 2 and 0 <= last_error < 1.0

    # helpers
    k += 5
    assert raw[k] == oracle.pse_eps(0.3, 0.75)
    s, e, _ = oracle.load_balance(n, 3)
    assert np.array_equal(raw[k + 1: k + 4], s) and np.array_equal(raw[k + 4: k + 7], e)
    # the batch's pairs per evaluation (self-pairs skipped) and the far-field
    # object's particle cells, read back through their bind(C) interfaces: the
    # object rebins moved particles at every evaluation, so after two learned
    # steps some boundary particles sit in a neighbouring cell (46 of 482 here)
    na = int(np.count_nonzero(m.active))
    assert raw[k + 7] == n * na - na
    assert 0 <= raw[k + 8] < n // 4
