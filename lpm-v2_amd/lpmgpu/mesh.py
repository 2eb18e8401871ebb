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

Particles are numbered in the reference's creation order (seed vertices,
seed centres, then per level and face: new edge midpoints, new child
centres); leaf-face centres are active, every other particle is a passive
vertex with area 0 (Faces.f90:840-850; Particles.f90).
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
    face_centre: np.ndarray | None = None  # centre particle of each leaf face
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


def _divide_level(pos, faces, centre, K, sphere, radius):
    """One uniform division of every leaf face, numbering new particles in the
    reference's creation order: faces are divided in order (PolyMesh2d.f90:179-189);
    for each face its edges are visited in slot order k = (v_k, v_k+1) and an
    edge's midpoint is inserted the first time it is met (Faces.f90:727-758,
    divideLinearEdge Edges.f90:567-617); then the new child centres are
    inserted, children 1..3 for triangles (child 4 keeps the parent's centre
    particle, Faces.f90:836-850) and children 1..4 for quadrilaterals (the
    parent centre becomes a passive vertex, Faces.f90:640-655)."""
    nf = faces.shape[0]
    n0 = pos.shape[0]
    a = faces
    b = np.roll(faces, -1, axis=1)
    lo = np.minimum(a, b).ravel().astype(np.int64)
    hi = np.maximum(a, b).ravel().astype(np.int64)
    key = lo * (n0 + 1) + hi
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    n_mid = uniq.shape[0]
    n_cen = 3 if K == 3 else 4
    # events: midpoint of edge e at (face, slot) of its first encounter, then centres
    ev_face = np.concatenate([first // K, np.repeat(np.arange(nf), n_cen)])
    ev_sub = np.concatenate([first % K, np.tile(np.arange(n_cen), nf) + K])
    order = np.lexsort((ev_sub, ev_face))
    new_index = np.empty(order.shape[0], np.int64)
    new_index[order] = n0 + np.arange(order.shape[0])
    mid_idx = new_index[:n_mid][inv].reshape(nf, K)  # midpoint of edge slot k of face f
    cen_idx = new_index[n_mid:].reshape(nf, n_cen)
    ea = uniq // (n0 + 1)
    eb = uniq % (n0 + 1)
    mid = 0.5 * (pos[ea] + pos[eb])
    if sphere:
        mid = _normalize(mid, radius)
    newpos = np.empty((order.shape[0], 3))
    newpos[new_index[:n_mid] - n0] = mid
    pos = np.concatenate([pos, newpos], axis=0)
    if K == 3:
        v1, v2, v3 = faces[:, 0], faces[:, 1], faces[:, 2]
        m1, m2, m3 = mid_idx[:, 0], mid_idx[:, 1], mid_idx[:, 2]
        ch = [np.stack([v1, m1, m3], 1), np.stack([m1, v2, m2], 1), np.stack([m3, m2, v3], 1),
              np.stack([m2, m3, m1], 1)]
        for c in range(3):
            cc = _normalize(pos[ch[c]].sum(axis=1) / 3.0, radius) if sphere else pos[ch[c]].sum(axis=1) / 3.0
            pos[cen_idx[:, c]] = cc
        newcentre = np.stack([cen_idx[:, 0], cen_idx[:, 1], cen_idx[:, 2], centre], 1).reshape(-1)
    else:
        v1, v2, v3, v4 = (faces[:, k] for k in range(4))
        m1, m2, m3, m4 = (mid_idx[:, k] for k in range(4))
        C = centre
        ch = [np.stack([v1, m1, C, m4], 1), np.stack([m1, v2, m2, C], 1), np.stack([C, m2, v3, m3], 1),
              np.stack([m4, C, m3, v4], 1)]
        for c in range(4):
            cc = pos[ch[c]].sum(axis=1) / 4.0
            pos[cen_idx[:, c]] = _normalize(cc, radius) if sphere else cc
        newcentre = cen_idx.reshape(-1)
    faces = np.stack(ch, axis=1).reshape(-1, K)
    return pos, faces, newcentre


def _build(seed_pos, seed_faces, seed_centre_pos, level, K, sphere, radius, cell_level):
    nv = seed_pos.shape[0]
    pos = np.concatenate([seed_pos, seed_centre_pos], axis=0)
    faces = seed_faces.astype(np.int64)
    centre = nv + np.arange(faces.shape[0])
    cell = np.arange(faces.shape[0]) if cell_level == 0 else np.full(faces.shape[0], -1)
    for lev in range(level):
        pos, faces, centre = _divide_level(pos, faces, centre, K, sphere, radius)
        cell = np.arange(faces.shape[0]) if lev + 1 == cell_level else np.repeat(cell, 4)
    cpos = pos[centre]
    area = np.zeros(faces.shape[0])
    for k in range(K):
        va = pos[faces[:, k]]
        vb = pos[faces[:, (k + 1) % K]]
        if sphere:
            area += sphere_tri_area(va, cpos, vb, radius)
        else:
            # TriArea(v1, centre, v2), PlaneGeometry.f90:109-113
            area += 0.5 * np.abs((va[:, 0] - cpos[:, 0]) * (vb[:, 1] - cpos[:, 1])
                                 - (vb[:, 0] - cpos[:, 0]) * (va[:, 1] - cpos[:, 1]))
    return pos, faces, centre, area, cell


def icosahedral_mesh(level: int, radius: float = 1.0, cell_level: int = -1) -> ParticleMesh:
    """Icosahedral triangle mesh at uniform nest `level`: V = 10*4^L + 2 passive
    vertices, F = 20*4^L active centres (PolyMesh2d.f90:737-738, 761-762).  The
    seed's 20 centres are SphereTriCenter of their vertices (PolyMesh2d.f90:875-879)."""
    verts, faces = _icosahedron()
    verts = verts * radius
    centres = _normalize(verts[faces].sum(axis=1) / 3.0, radius)
    pos, f, c, area, cell = _build(verts, faces, centres, level, 3, True, radius, cell_level)
    return _assemble(pos, f, c, area, "icos", level, radius, cell, cell_level, sphere=True)


def _cube_seed():
    s = 1.0 / np.sqrt(3.0)
    corners = np.array(
        [[s, -s, s], [s, -s, -s], [s, s, -s], [s, s, s], [-s, s, -s], [-s, s, s], [-s, -s, -s], [-s, -s, s]]
    )
    quads = np.array([[0, 1, 2, 3], [3, 2, 4, 5], [5, 4, 6, 7], [7, 6, 1, 0], [7, 0, 3, 5], [1, 6, 4, 2]])
    return corners, quads


def cubed_sphere_mesh(level: int, radius: float = 1.0, cell_level: int = -1) -> ParticleMesh:
    """Cubed-sphere quad mesh at uniform nest `level`: V = 6*4^L + 2 passive
    vertices, F = 6*4^L active centres (PolyMesh2d.f90:739-740, 763-764)."""
    verts, faces = _cube_seed()
    verts = verts * radius
    centres = _normalize(verts[faces].sum(axis=1) / 4.0, radius)
    pos, f, c, area, cell = _build(verts, faces, centres, level, 4, True, radius, cell_level)
    return _assemble(pos, f, c, area, "cubed", level, radius, cell, cell_level, sphere=True)


def quad_rect_mesh(level: int, half_width: float = 1.0) -> ParticleMesh:
    """Planar quadRect mesh on [-w, w]^2 at nest `level` (4*4^L cells of side
    h = w / 2^L, centres at cell centroids, areas h^2 from the triangle sums)."""
    s = half_width
    verts = np.array([[-s, s], [-s, 0.0], [-s, -s], [0.0, -s], [s, -s], [s, 0.0], [s, s], [0.0, s], [0.0, 0.0]])
    faces = np.array([[0, 1, 8, 7], [1, 2, 3, 8], [8, 3, 4, 5], [7, 8, 5, 6]])
    verts3 = np.concatenate([verts, np.zeros((verts.shape[0], 1))], axis=1)
    centres = verts3[faces].sum(axis=1) / 4.0
    pos, f, c, area, cell = _build(verts3, faces, centres, level, 4, False, 1.0, -1)
    return _assemble(pos, f, c, area, "quadrect", level, 1.0, cell, -1, sphere=False)


def _assemble(pos, faces, centre, area, kind, level, radius, cell, cell_level, sphere):
    n = pos.shape[0]
    ar = np.zeros(n)
    ar[centre] = area
    act = np.zeros(n, np.int32)
    act[centre] = 1
    # maximum leaf-edge length (Edges.f90:260-275): arc on the sphere, chord on the plane
    a = pos[faces].reshape(-1, 3)
    b = pos[np.roll(faces, -1, axis=1)].reshape(-1, 3)
    if sphere:
        maxlen = float(np.max(_arc(a, b)) * radius)
    else:
        maxlen = float(np.max(np.sqrt(np.sum((a - b) ** 2, axis=1))))
    return ParticleMesh(
        x=np.ascontiguousarray(pos[:, 0]),
        y=np.ascontiguousarray(pos[:, 1]),
        z=np.ascontiguousarray(pos[:, 2]),
        area=ar,
        active=act,
        n_vertices=n - centre.shape[0],
        n_faces=centre.shape[0],
        kind=kind,
        level=level,
        radius=radius,
        faces=faces,
        face_centre=centre,
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
