"""Sphere BVE with a learned far field (lpm_bve_ff_*)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from ._lib import f64, i32, pd, pi

_pf = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int

L.extra_signatures({
    "lpm_bve_ff_new": ([_vp, ctypes.POINTER(_vp), _i, L._pd, L._pd, L._pd, L._pd, L._pd, L._pi, L._pi, _i, _d, _d, _d,
                        _d, _i], _i),
    "lpm_bve_ff_velocity": ([_vp, _i, L._pd, L._pd, L._pd], _i),
    "lpm_bve_ff_collect": ([_vp, _i, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64)], _i),
    "lpm_bve_ff_train": ([_vp, _i, _i, _d, ctypes.c_uint64, L._pd], _i),
    "lpm_bve_ff_set_online": ([_vp, _i, _i, _i, _d], _i),
    "lpm_bve_ff_set_online_from": ([_vp, _d], _i),
    "lpm_bve_ff_samples": ([_vp, ctypes.c_int64, ctypes.c_int64, _pf, _pf, L._pi], _i),
    "lpm_bve_ff_set_params": ([_vp, _pf, ctypes.c_int64], _i),
    "lpm_bve_ff_get_params": ([_vp, _pf, ctypes.c_int64], _i),
    "lpm_bve_ff_steps": ([_vp, _d, _i, _i, _i, _d, ctypes.POINTER(_i), L._pd, ctypes.POINTER(_i)], _i),
    "lpm_bve_ff_get_state": ([_vp, L._pd, L._pd, L._pd, L._pd, L._pd, L._pd, L._pd], _i),
    "lpm_bve_ff_set_state": ([_vp, L._pd, L._pd, L._pd, L._pd], _i),
    "lpm_bve_ff_set_particles": ([_vp, _i, L._pd, L._pd, L._pd, L._pd, L._pd, L._pi], _i),
    "lpm_bve_ff_cells": ([_vp, L._pi, L._pi], _i),
    "lpm_bve_ff_set_parents": ([_vp, _d], _i),
    "lpm_bve_ff_far_counts": ([_vp, L._pi], _i),
    "lpm_bve_ff_particle_cells": ([_vp, L._pi], _i),
    "lpm_bve_ff_delete": ([_vp], _i),
})

PLAIN, EXPANSION, LEARNED = 0, 1, 2
N_PARAMS = 13 * 128 + 128 + 2 * (128 * 128 + 128) + 3 * 128 + 3


class FarFieldBVE:
    """Sphere BVE whose far cells use an expansion of order `order` (2..4) in
    their moments, with a network correcting the remainder in the band.  The
    defaults are the measured icosahedral L8 configuration: cells nearer than
    0.1 rad summed directly, the network over the far cells within 0.2 rad,
    training samples taking the nearest of 4 band cells."""

    def __init__(self, ctx, mesh, relvort, rotation_rate=0.0, near_radius=0.1, band_radius=0.2, radius=1.0,
                 order=4, sample_draws=4, parent_radius=0.0):
        if mesh.face_cell is None or mesh.cell_level < 0:
            raise ValueError("mesh needs coarse cells (cell_level >= 0)")
        self.ctx = ctx
        self.order = order
        self.n = mesh.n
        self.ncell = int(mesh.face_cell.max()) + 1
        self.cell = self.particle_cells(mesh)
        arr = [f64(a) for a in (mesh.x, mesh.y, mesh.z, relvort, mesh.area)]
        act = i32(mesh.active)
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.lpm_bve_ff_new(ctx.handle, ctypes.byref(h), self.n, *[pd(a) for a in arr], pi(act),
                                          pi(self.cell), self.ncell, radius, rotation_rate, near_radius, band_radius,
                                          order | ((sample_draws & 7) << 3)))
        self.handle = h
        ctx.adopt(self)
        if parent_radius > 0.0:
            self.set_parents(parent_radius)

    def set_parents(self, parent_radius):
        """Coarse level of the expansion part (lpm_bve_ff_set_parents): beyond
        `parent_radius` four sibling far cells are one expansion; 0 turns it off.
        Needs the cells of a refined mesh (cells 4p .. 4p+3 share parent p)."""
        self.ctx.check(self.ctx._lib.lpm_bve_ff_set_parents(self.handle, float(parent_radius)))

    def far_counts(self):
        """Expansion records per target cell in the last evaluation."""
        a = np.zeros(self.ncell, np.int32)
        self.ctx.check(self.ctx._lib.lpm_bve_ff_far_counts(self.handle, pi(a)))
        return a

    @staticmethod
    def particle_cells(mesh):
        """Coarse cell of every particle: a face centre belongs to its face's cell,
        a vertex to the cell of the first leaf face that has it as a corner."""
        cell = np.full(mesh.n, -1, np.int64)
        fc = np.asarray(mesh.face_cell)
        for k in range(mesh.faces.shape[1] - 1, -1, -1):
            cell[mesh.faces[:, k]] = fc
        cell[mesh.face_centre] = fc
        return i32(cell)

    def velocity(self, mode):
        u, v, w = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        self.ctx.check(self.ctx._lib.lpm_bve_ff_velocity(self.handle, mode, pd(u), pd(v), pd(w)))
        return u, v, w

    def collect(self, n_samples, seed=0):
        tot = ctypes.c_int64()
        self.ctx.check(self.ctx._lib.lpm_bve_ff_collect(self.handle, n_samples, seed, ctypes.byref(tot)))
        return tot.value

    def samples(self, offset, count):
        """Inputs [count][13], targets [count][3] and ids [count][2] = (particle, cell)."""
        x = np.zeros((count, 13), np.float32)
        t = np.zeros((count, 3), np.float32)
        ids = np.zeros((count, 2), np.int32)
        self.ctx.check(self.ctx._lib.lpm_bve_ff_samples(self.handle, offset, count, x.ctypes.data_as(_pf),
                                                        t.ctypes.data_as(_pf), pi(ids)))
        return x, t, ids

    def train(self, epochs=10, batch=4096, lr=1e-3, seed=0):
        loss = ctypes.c_double()
        self.ctx.check(self.ctx._lib.lpm_bve_ff_train(self.handle, epochs, batch, lr, seed, ctypes.byref(loss)))
        return loss.value

    def set_online(self, n_samples, epochs=1, batch=4096, lr=2e-4, above=0.0):
        """Fine-tune on n_samples direct-sum samples at every judge that keeps
        the learned mode (0 turns it off) and whose error is above `above` x
        tol (lpm_bve_ff_set_online_from; 0: at every such judge)."""
        self.ctx.check(self.ctx._lib.lpm_bve_ff_set_online(self.handle, n_samples, epochs, batch, lr))
        self.ctx.check(self.ctx._lib.lpm_bve_ff_set_online_from(self.handle, float(above)))

    def params(self):
        p = np.zeros(N_PARAMS, np.float32)
        self.ctx.check(self.ctx._lib.lpm_bve_ff_get_params(self.handle, p.ctypes.data_as(_pf), N_PARAMS))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self.ctx.check(self.ctx._lib.lpm_bve_ff_set_params(self.handle, p.ctypes.data_as(_pf), p.shape[0]))

    def steps(self, dt, nsteps, mode, judge_every=10, tol=1e-4):
        m = ctypes.c_int()
        e = ctypes.c_double()
        j = ctypes.c_int()
        self.ctx.check(self.ctx._lib.lpm_bve_ff_steps(self.handle, dt, nsteps, mode, judge_every, tol, ctypes.byref(m),
                                                      ctypes.byref(e), ctypes.byref(j)))
        return dict(mode=m.value, last_error=e.value, judged=j.value)

    def state(self):
        out = [np.zeros(self.n) for _ in range(7)]
        self.ctx.check(self.ctx._lib.lpm_bve_ff_get_state(self.handle, *[pd(o) for o in out]))
        return dict(zip(["x", "y", "z", "relvort", "u", "v", "w"], out))

    def set_state(self, x, y, z, relvort):
        arr = [f64(a) for a in (x, y, z, relvort)]
        self.ctx.check(self.ctx._lib.lpm_bve_ff_set_state(self.handle, *[pd(a) for a in arr]))

    def set_particles(self, x, y, z, relvort, area, active):
        """Remesh re-upload (lpm_bve_ff_set_particles): a new particle set of any size; cells,
        the trained network and the collected samples are kept."""
        arr = [f64(a) for a in (x, y, z, relvort, area)]
        act = np.ascontiguousarray(active, dtype=np.int32)
        n = arr[0].shape[0]
        self.ctx.check(self.ctx._lib.lpm_bve_ff_set_particles(self.handle, n, *[pd(a) for a in arr],
                                                              act.ctypes.data_as(L._pi)))
        self.n = n

    def pretrain(self, dt, states=5, every=5, samples=2_000_000, epochs=10, batch=4096, lr=1e-3, seed=1):
        """Train on direct-sum samples of the method's own run: collect at
        `states` states `every` expansion-mode steps apart, return to the
        starting state, then train.  Returns the final training loss."""
        st = self.state()
        per = max(1, samples // states)
        for k in range(states):
            self.collect(per, seed=seed + k)
            if k + 1 < states:
                self.steps(dt, every, EXPANSION, judge_every=0)
        self.set_state(st["x"], st["y"], st["z"], st["relvort"])
        return self.train(epochs=epochs, batch=batch, lr=lr, seed=seed)

    def binning(self):
        """Cell of every particle as binned by the last evaluation."""
        c = np.zeros(self.n, np.int32)
        self.ctx.check(self.ctx._lib.lpm_bve_ff_particle_cells(self.handle, pi(c)))
        return c

    def cells(self):
        a = np.zeros(self.ncell, np.int32)
        b = np.zeros(self.ncell, np.int32)
        self.ctx.check(self.ctx._lib.lpm_bve_ff_cells(self.handle, pi(a), pi(b)))
        return a, b

    def delete(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.lpm_bve_ff_delete(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.delete()
        except Exception:
            pass
