"""Mesh generator checks (CPU): counts (PolyMesh2d.f90:717-766), areas,
seed tables and the reference's creation order."""
import numpy as np
import pytest

from lpmgpu import mesh


@pytest.mark.parametrize("level", range(0, 5))
def test_icosahedral_counts_and_area(level):
    m = mesh.icosahedral_mesh(level)
    assert m.n_vertices == 10 * 4**level + 2
    assert m.n_faces == 20 * 4**level
    assert m.n_active == m.n_faces
    assert abs(m.area.sum() - 4 * np.pi) < 1e-10  # OOMeshTest area check, ZERO_TOL
    r = np.sqrt(m.x**2 + m.y**2 + m.z**2)
    assert np.abs(r - 1.0).max() < 1e-15
    assert np.all(m.area[m.active == 0] == 0.0) and np.all(m.area[m.active == 1] > 0.0)


@pytest.mark.parametrize("level", range(0, 4))
def test_cubed_sphere_counts_and_area(level):
    m = mesh.cubed_sphere_mesh(level)
    assert m.n_vertices == 6 * 4**level + 2
    assert m.n_faces == 6 * 4**level
    assert abs(m.area.sum() - 4 * np.pi) < 1e-10


def test_icosahedron_nest0_has_30_edges():
    """IcosTriMeshTester's nest-0 check: 30 edges, every vertex of degree 5."""
    m = mesh.icosahedral_mesh(0)
    e = set()
    for f in m.faces:
        for k in range(3):
            a, b = int(f[k]), int(f[(k + 1) % 3])
            e.add((min(a, b), max(a, b)))
    assert len(e) == 30
    deg = np.bincount(np.array(list(e)).ravel())
    assert np.all(deg == 5)


def test_seed_particles_come_first():
    """Creation order: 12 seed vertices, the 20 seed centres (SphereTriCenter of the
    seed faces, PolyMesh2d.f90:875-879), then level-1 particles."""
    m0 = mesh.icosahedral_mesh(0)
    m2 = mesh.icosahedral_mesh(2)
    assert np.array_equal(m2.x[:32], m0.x) and np.array_equal(m2.z[:32], m0.z)
    # the seed centres are still active centres (middle children keep them)
    assert np.all(m2.active[12:32] == 1)


def test_quad_rect_is_the_uniform_grid():
    m = mesh.quad_rect_mesh(3)
    act = m.active == 1
    h = 2.0 / 16
    assert np.allclose(m.area[act], h * h, rtol=1e-14)
    xs = np.sort(np.unique(np.round(m.x[act], 12)))
    assert np.allclose(xs, -1 + h * (np.arange(16) + 0.5))


def test_grid_centres_jitter_is_seeded():
    a = mesh.grid_centres(16, jitter=1e-3, seed=3)
    b = mesh.grid_centres(16, jitter=1e-3, seed=3)
    assert np.array_equal(a[0], b[0])
    assert np.abs(a[0] - mesh.grid_centres(16)[0]).max() <= 1e-3


def test_cell_ids_partition_faces():
    m = mesh.icosahedral_mesh(4, cell_level=2)
    counts = np.bincount(m.face_cell)
    assert counts.shape[0] == 320 and np.all(counts == 16)
