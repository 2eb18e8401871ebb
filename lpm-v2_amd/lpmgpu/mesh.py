"""Particle meshes for the direct-sum path (input side of the hot path).

Restates lpm-v2's seed-plus-uniform-division mesh construction
(PolyMesh2d.f90:135-195, 795-939) with vectorised numpy, level by level:

* icosahedral triangles (`DivideTriFace`, Faces.f90:701-858): each parent
  triangle (a, b, c) gets edge midpoints projected to the sphere
  (`SphereMidpoint`, SphereGeometry.f90:280-289), children
  (a, m_ab, m_ca), (m_ab, b, m_bc), (m_ca, m_bc, c) get new centre particles
  at `SphereTriCenter` (SphereGeometry.f90:302-311), and the middle child
  (m_bc, m_ca, m_ab) keeps the parent's centre particle where it is
  (Faces.f90:846-850);
* cubed-sphere quadrilaterals (`DivideQuadFace`, Faces.f90:529-689): the
  parent centre becomes a passive vertex and all four children get new
  centres at `SphereQuadCenter` (SphereGeometry.f90:325-334);
* planar quadRect quadrilaterals (quadRectSeed.dat spans [-1, 1]^2): the
  same quad division with plain midpoints and centroids.

Face areas are the reference's `TriFaceArea` / `QuadFaceArea`
(Faces.f90:917-975): the sum over polygon edges of the spherical (or planar)
triangle (v_k, centre, v_k+1), with L'Huilier's formula
(`SphereTriAreaVector`, SphereGeometry.f90:347-366).

Particle order is [vertices..., face centres...]; vertices are passive
(area 0) and face centres are active, as in the reference
(Faces.f90:840-850; Particles.f90).  The reference interleaves vertices and
centres as they are created; the sums do not depend on the order.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class ParticleMesh:
    """SoA particle set: x, y, z, area, active (int32 0/1), plus bookkeeping."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    area: np.ndarray
    active: np.ndarray
    n_vertices: int
    n_faces: int
    kind: str
    level: int
    radius: float = 1.0
    faces: np.ndarray | None = None  # leaf-face vertex indices
    face_cell: np.ndarray | None = None  # ancestor face id at `cell_level`
    cell_level: int = -1
    max_edge_length: float = 0.0

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_active(self) -> int:
        return int(self.active.sum())

    def xyz(self) -> np.ndarray:
        return np.stack([self.x, self.y, self.z], axis=1)


def _normalize(v: np.ndarray, radius: float) -> np.ndarray:
    return v / np.sqrt(np.sum(v * v, axis=1))[:, None] * radius


def _arc(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    c = np.cross(a, b)
    return np.arctan2(np.sqrt(np.sum(c * c, axis=1)), np.sum(a * b, axis=1))


def sphere_tri_area(a: np.ndarray, b: np.ndarray, c: np.ndarray, radius: float = 1.0) -> np.ndarray:
    """L'Huilier's formula on the sphere (SphereGeometry.f90:347-366)."""
    s1 = _arc(a, b)
    s2 = _arc(b, c)
    s3 = _arc(c, a)
    hp = 0.5 * (s1 + s2 + s3)
    zz = np.tan(0.5 * hp) * np.tan(0.5 * (hp - s1)) * np.tan(0.5 * (hp - s2)) * np.tan(0.5 * (hp - s3))
    return 4.0 * np.arctan2(np.sqrt(np.maximum(zz, 0.0)), 1.0) * radius * radius


def _icosahedron():
    """Regular icosahedron: pole, upper ring at lon 36 + 72k, lower ring at 72k, pole
    (the vertex layout of icosTriSeed.dat)."""
    zr = 1.0 / np.sqrt(5.0)
    rr = 2.0 / np.sqrt(5.0)
    v = [(0.0, 0.0, 1.0)]
    for k in range(5):
        lon = np.deg2rad(36.0 + 72.0 * k)
        v.append((rr * np.cos(lon), rr * np.sin(lon), zr))
    for k in range(5):
        lon = np.deg2rad(72.0 * k)
        v.append((rr * np.cos(lon), rr * np.sin(lon), -zr))
    v.append((0.0, 0.0, -1.0))
    v = np.array(v, dtype=np.float64)
    up = [1, 2, 3, 4, 5]
    lo = [6, 7, 8, 9, 10]
    f = []
    for k in range(5):
        f.append((0, up[k], up[(k + 1) % 5]))
    for k in range(5):
        # up[k] (lon 36+72k) sits above the gap between lo[k] and lo[k+1]
        f.append((up[k], lo[k], lo[(k + 1) % 5]))
        f.append((lo[(k + 1) % 5], up[(k + 1) % 5], up[k]))
    for k in range(5):
        f.append((11, lo[(k + 1) % 5], lo[k]))
    f = np.array(f, dtype=np.int64)
    # orient every face counter-clockwise seen from outside
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    flip = np.sum(np.cross(b - a, c - a) * (a + b + c), axis=1) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def _edge_midpoints(verts, faces, nvert_per_face, sphere, radius):
    nf = faces.shape[0]
    k = nvert_per_face
    a = faces
    b = np.roll(faces, -1, axis=1)
    lo = np.minimum(a, b).ravel()
    hi = np.maximum(a, b).ravel()
    nv = verts.shape[0]
    key = lo.astype(np.int64) * nv + hi
    uniq, inv = np.unique(key, return_inverse=True)
    ea = uniq // nv
    eb = uniq % nv
    mid = 0.5 * (verts[ea] + verts[eb])
    if sphere:
        mid = _normalize(mid, radius)
    mid_idx = nv + inv.reshape(nf, k)
    return np.concatenate([verts, mid], axis=0), mid_idx


def icosahedral_mesh(level: int, radius: float = 1.0, cell_level: int = -1) -> ParticleMesh:
    """Icosahedral triangle mesh at uniform nest `level`: V = 10*4^L + 2 passive
    vertices, F = 20*4^L active centres (PolyMesh2d.f90:737-738, 761-762)."""
    verts, faces = _icosahedron()
    verts = verts * radius
    centres = _normalize(verts[faces].sum(axis=1) / 3.0, radius)
    cell = np.arange(faces.shape[0]) if cell_level == 0 else np.full(faces.shape[0], -1)
    for lev in range(level):
        verts, m = _edge_midpoints(verts, faces, 3, True, radius)
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        mab, mbc, mca = m[:, 0], m[:, 1], m[:, 2]
        c1 = np.stack([a, mab, mca], 1)
        c2 = np.stack([mab, b, mbc], 1)
        c3 = np.stack([mca, mbc, c], 1)
        c4 = np.stack([mbc, mca, mab], 1)
        newc = [
            _normalize(verts[cc].sum(axis=1) / 3.0, radius) for cc in (c1, c2, c3)
        ] + [centres]
        faces = np.stack([c1, c2, c3, c4], axis=1).reshape(-1, 3)
        centres = np.stack(newc, axis=1).reshape(-1, 3)
        if lev + 1 == cell_level:
            cell = np.arange(faces.shape[0])
        else:
            cell = np.repeat(cell, 4)
    area = np.zeros(faces.shape[0])
    for k in range(3):
        area += sphere_tri_area(verts[faces[:, k]], centres, verts[faces[:, (k + 1) % 3]], radius)
    return _assemble(verts, centres, area, faces, "icos", level, radius, cell, cell_level, sphere=True)


def _cube_seed():
    s = 1.0 / np.sqrt(3.0)
    corners = np.array(
        [[s, -s, s], [s, -s, -s], [s, s, -s], [s, s, s], [-s, s, -s], [-s, s, s], [-s, -s, -s], [-s, -s, s]]
    )
    quads = np.array([[0, 1, 2, 3], [3, 2, 4, 5], [5, 4, 6, 7], [7, 6, 1, 0], [7, 0, 3, 5], [1, 6, 4, 2]])
    return corners, quads


def _quad_divide(verts, faces, centres, sphere, radius):
    """One uniform quad division; parent centres become vertices."""
    nv0 = verts.shape[0]
    verts, m = _edge_midpoints(verts, faces, 4, sphere, radius)
    nf = faces.shape[0]
    cidx = verts.shape[0] + np.arange(nf)
    verts = np.concatenate([verts, centres], axis=0)
    v0, v1, v2, v3 = (faces[:, k] for k in range(4))
    m0, m1, m2, m3 = (m[:, k] for k in range(4))  # m_k on edge (v_k, v_k+1)
    ch = [
        np.stack([v0, m0, cidx, m3], 1),
        np.stack([m0, v1, m1, cidx], 1),
        np.stack([cidx, m1, v2, m2], 1),
        np.stack([m3, cidx, m2, v3], 1),
    ]
    faces = np.stack(ch, axis=1).reshape(-1, 4)
    cen = verts[faces].sum(axis=1) / 4.0
    if sphere:
        cen = _normalize(cen, radius)
    del nv0
    return verts, faces, cen


def cubed_sphere_mesh(level: int, radius: float = 1.0, cell_level: int = -1) -> ParticleMesh:
    """Cubed-sphere quad mesh at uniform nest `level`: V = 6*4^L + 2 passive
    vertices, F = 6*4^L active centres (PolyMesh2d.f90:739-740, 763-764)."""
    verts, faces = _cube_seed()
    verts = verts * radius
    centres = _normalize(verts[faces].sum(axis=1) / 4.0, radius)
    cell = np.arange(6) if cell_level == 0 else np.full(6, -1)
    for lev in range(level):
        verts, faces, centres = _quad_divide(verts, faces, centres, True, radius)
        cell = np.arange(faces.shape[0]) if lev + 1 == cell_level else np.repeat(cell, 4)
    area = np.zeros(faces.shape[0])
    for k in range(4):
        area += sphere_tri_area(verts[faces[:, k]], centres, verts[faces[:, (k + 1) % 4]], radius)
    return _assemble(verts, centres, area, faces, "cubed", level, radius, cell, cell_level, sphere=True)


def quad_rect_mesh(level: int, half_width: float = 1.0) -> ParticleMesh:
    """Planar quadRect mesh on [-w, w]^2 at nest `level` (4*4^L cells of side
    h = w / 2^L, centres at cell centroids, areas h^2 from the triangle sums)."""
    s = half_width
    verts = np.array([[-s, s], [-s, 0.0], [-s, -s], [0.0, -s], [s, -s], [s, 0.0], [s, s], [0.0, s], [0.0, 0.0]])
    faces = np.array([[0, 1, 8, 7], [1, 2, 3, 8], [8, 3, 4, 5], [7, 8, 5, 6]])
    verts3 = np.concatenate([verts, np.zeros((verts.shape[0], 1))], axis=1)
    centres = verts3[faces].sum(axis=1) / 4.0
    for _ in range(level):
        verts3, faces, centres = _quad_divide(verts3, faces, centres, False, 1.0)
    area = np.zeros(faces.shape[0])
    for k in range(4):
        a = verts3[faces[:, k]]
        b = verts3[faces[:, (k + 1) % 4]]
        # TriArea(v1, centre, v2), PlaneGeometry.f90:109-113
        area += 0.5 * np.abs((a[:, 0] - centres[:, 0]) * (b[:, 1] - centres[:, 1]) - (b[:, 0] - centres[:, 0]) * (a[:, 1] - centres[:, 1]))
    cell = np.full(faces.shape[0], -1)
    return _assemble(verts3, centres, area, faces, "quadrect", level, 1.0, cell, -1, sphere=False)


def _assemble(verts, centres, area, faces, kind, level, radius, cell, cell_level, sphere):
    nv = verts.shape[0]
    nf = centres.shape[0]
    xyz = np.concatenate([verts, centres], axis=0)
    ar = np.concatenate([np.zeros(nv), area])
    act = np.concatenate([np.zeros(nv, np.int32), np.ones(nf, np.int32)])
    # maximum leaf-edge length (Edges.f90:260-275): arc on the sphere, chord on the plane
    k = faces.shape[1]
    a = verts[faces].reshape(-1, 3)
    b = verts[np.roll(faces, -1, axis=1)].reshape(-1, 3)
    if sphere:
        maxlen = float(np.max(_arc(a, b)) * radius)
    else:
        maxlen = float(np.max(np.sqrt(np.sum((a - b) ** 2, axis=1))))
    del k
    return ParticleMesh(
        x=np.ascontiguousarray(xyz[:, 0]),
        y=np.ascontiguousarray(xyz[:, 1]),
        z=np.ascontiguousarray(xyz[:, 2]),
        area=ar,
        active=act,
        n_vertices=nv,
        n_faces=nf,
        kind=kind,
        level=level,
        radius=radius,
        faces=faces,
        face_cell=cell,
        cell_level=cell_level,
        max_edge_length=maxlen,
    )


def grid_centres(n_side: int, half_width: float = 1.0, jitter: float = 0.0, seed: int = 0):
    """Cell centres of a uniform n_side^2 grid on [-w, w]^2 (the active
    particles of quadRect at nest log2(n_side/2)), area h^2, optional seeded
    uniform +-jitter on each coordinate."""
    h = 2.0 * half_width / n_side
    c = -half_width + h * (np.arange(n_side) + 0.5)
    X, Y = np.meshgrid(c, c, indexing="ij")
    x = X.ravel().copy()
    y = Y.ravel().copy()
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        x += rng.uniform(-jitter, jitter, x.shape)
        y += rng.uniform(-jitter, jitter, y.shape)
    area = np.full(x.shape, h * h)
    return x, y, area, h
