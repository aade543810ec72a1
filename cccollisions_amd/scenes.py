"""Synthetic cloth scenes used by the benchmark, the smoke check and the parity tests.

Each scene is a vertex/face mesh (numpy, fp64 vertices, int64 faces); the collision
entry points take triangle soups ``vertices[faces]`` of shape ``[N, 3, 3]``.  All
randomness comes from ``numpy.random.default_rng(seed)``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class ClothScene:
    """One simulation state: positions at the start (x0) and end (x1) of a step."""
    faces: np.ndarray      # [F, 3] int64
    x0: np.ndarray         # [V, 3] fp64
    x1: np.ndarray         # [V, 3] fp64
    vmass: np.ndarray      # [V] fp64 vertex masses

    @property
    def num_triangles(self) -> int:
        return int(self.faces.shape[0])

    def soups(self):
        return self.x0[self.faces], self.x1[self.faces]

    def corner_masses(self):
        return self.vmass[self.faces]


def grid_faces(nx: int, ny: int) -> np.ndarray:
    """Faces of an nx-by-ny vertex grid, two triangles per quad, row-major order."""
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="ij")
    v00 = (i * ny + j).ravel()
    v10 = ((i + 1) * ny + j).ravel()
    v01 = (i * ny + j + 1).ravel()
    v11 = ((i + 1) * ny + j + 1).ravel()
    f = np.empty((v00.size * 2, 3), np.int64)
    f[0::2] = np.stack([v00, v10, v11], 1)
    f[1::2] = np.stack([v00, v11, v01], 1)
    return f


def grid_vertices(nx: int, ny: int) -> np.ndarray:
    x, y = np.meshgrid(np.linspace(0.0, 1.0, nx), np.linspace(0.0, 1.0, ny), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), np.zeros(nx * ny)], 1)


def proximity_cloth(n: int = 64, seed: int = 1, jitter: float = 1e-3) -> tuple[np.ndarray, np.ndarray]:
    """Static cloth: n-by-n vertex grid on the unit square,
    z = 0.05 sin(6x) cos(6y) + N(0, jitter).  Returns (vertices, faces)."""
    rng = np.random.default_rng(seed)
    v = grid_vertices(n, n)
    v[:, 2] = 0.05 * np.sin(6 * v[:, 0]) * np.cos(6 * v[:, 1]) + rng.normal(0.0, jitter, v.shape[0])
    return v, grid_faces(n, n)


def folded_cloth(n: int = 256, seed: int = 2, gap: float = 1e-3, sigma: float = 0.5,
                 dt: float = 1.0 / 30.0) -> ClothScene:
    """Cloth folded over itself: n-by-n vertex grid on the unit square, vertices with
    x > 0.5 mirrored to (1 - x, y, gap); t1 = t0 + dt * v with v ~ N(0, sigma) per
    vertex component.  Vertex masses are 1."""
    rng = np.random.default_rng(seed)
    v = grid_vertices(n, n)
    right = v[:, 0] > 0.5
    v[right, 0] = 1.0 - v[right, 0]
    v[right, 2] = gap
    vel = rng.normal(0.0, sigma, v.shape)
    return ClothScene(grid_faces(n, n), v, v + dt * vel, np.ones(v.shape[0]))


def grid_size_for_triangles(num_tris: int) -> int:
    """Vertex count per side of the square grid with about num_tris triangles."""
    return max(2, int(round(math.sqrt(num_tris / 2.0))) + 1)


def batched_cloths(count: int = 64, lo: int = 2000, hi: int = 200000, seed: int = 4,
                   **kw) -> list[ClothScene]:
    """Independent folded cloths whose triangle counts are log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), count))
    return [folded_cloth(grid_size_for_triangles(int(s)), seed=seed * 1000 + i, **kw)
            for i, s in enumerate(sizes)]


def _frame(axis: np.ndarray):
    """Two unit vectors orthogonal to `axis` (unit) and to each other."""
    t = np.array([1.0, 0.0, 0.0]) if abs(axis[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(axis, t)
    e1 /= np.linalg.norm(e1)
    return e1, np.cross(axis, e1)


def capsule_mesh(a, b, radius: float, nu: int, nv: int) -> tuple[np.ndarray, np.ndarray]:
    """UV mesh of a capsule around segment a-b: nu rings along the axis (the first and last
    quarter of them on the hemispherical caps), nv vertices per ring, closed in v (the seam
    column is duplicated, so the mesh is a grid of nu x (nv + 1) vertices)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    axis = b - a
    length = float(np.linalg.norm(axis))
    axis /= length
    e1, e2 = _frame(axis)
    # arc-length parameter over cap + cylinder + cap, mapped to (axial offset, ring radius)
    total = math.pi * radius + length
    s = np.linspace(0.0, total, nu)
    cap = math.pi * radius / 2
    off = np.where(s < cap, -radius * np.cos(s / radius),
                   np.where(s > cap + length, length + radius * np.sin((s - cap - length) / radius), s - cap))
    rr = np.where(s < cap, radius * np.sin(s / radius),
                  np.where(s > cap + length, radius * np.cos((s - cap - length) / radius), radius))
    rr = np.maximum(rr, 1e-3 * radius)       # keep the pole rings non-degenerate
    th = np.linspace(0.0, 2 * math.pi, nv + 1)
    ring = np.cos(th)[:, None] * e1 + np.sin(th)[:, None] * e2
    v = a + off[:, None, None] * axis + rr[:, None, None] * ring[None]
    return v.reshape(-1, 3), grid_faces(nu, nv + 1)


# (start, end, radius) of the 12 capsules of the procedural body, metres, z up
BODY_CAPSULES = [
    ((0.0, 0.0, 0.95), (0.0, 0.0, 1.40), 0.15),      # torso
    ((0.0, 0.0, 1.52), (0.0, 0.0, 1.68), 0.09),      # head
    ((0.0, 0.0, 1.42), (0.0, 0.0, 1.50), 0.05),      # neck
    ((-0.10, 0.0, 0.80), (0.10, 0.0, 0.80), 0.14),   # hips
    ((0.22, 0.0, 1.36), (0.48, 0.0, 1.30), 0.055),   # upper arm r
    ((0.48, 0.0, 1.30), (0.72, 0.0, 1.26), 0.045),   # forearm r
    ((-0.22, 0.0, 1.36), (-0.48, 0.0, 1.30), 0.055), # upper arm l
    ((-0.48, 0.0, 1.30), (-0.72, 0.0, 1.26), 0.045), # forearm l
    ((0.10, 0.0, 0.72), (0.12, 0.0, 0.40), 0.075),   # thigh r
    ((0.12, 0.0, 0.40), (0.13, 0.0, 0.06), 0.055),   # shin r
    ((-0.10, 0.0, 0.72), (-0.12, 0.0, 0.40), 0.075), # thigh l
    ((-0.12, 0.0, 0.40), (-0.13, 0.0, 0.06), 0.055), # shin l
]


def garment_on_body(body_grid: int = 32, sleeve_rings: int = 224, sleeve_segments: int = 223, seed: int = 3,
                    gap: float = 2e-3, sigma: float = 0.3, dt: float = 1.0 / 30.0) -> ClothScene:
    """Static procedural body (12 capsule UV meshes of body_grid x body_grid vertices,
    ~24k triangles at 32) and a cylindrical sleeve of sleeve_rings x (sleeve_segments + 1)
    vertices (~100k triangles at 224 x 223) around the right arm at radius + gap, with
    inward radial velocity |N(0, sigma)| per vertex (seeded).  One soup: body vertices have
    t1 = t0, so body-body cubics are linear and contribute no roots.  Vertex masses 1."""
    rng = np.random.default_rng(seed)
    verts, faces, off = [], [], 0
    for a, b, r in BODY_CAPSULES:
        v, f = capsule_mesh(a, b, r, body_grid, body_grid - 1)
        verts.append(v)
        faces.append(f + off)
        off += v.shape[0]
    nbody = off
    # sleeve around the right arm (upper arm + forearm axis), open cylinder
    a, b = np.array(BODY_CAPSULES[4][0]), np.array(BODY_CAPSULES[5][1])
    axis = b - a
    length = float(np.linalg.norm(axis))
    axis /= length
    e1, e2 = _frame(axis)
    radius = BODY_CAPSULES[4][2] + gap
    t = np.linspace(0.0, length, sleeve_rings)
    th = np.linspace(0.0, 2 * math.pi, sleeve_segments + 1)
    ring = np.cos(th)[:, None] * e1 + np.sin(th)[:, None] * e2
    sv = (a + t[:, None, None] * axis + radius * ring[None]).reshape(-1, 3)
    radial = np.broadcast_to(ring[None], (sleeve_rings, sleeve_segments + 1, 3)).reshape(-1, 3)
    verts.append(sv)
    faces.append(grid_faces(sleeve_rings, sleeve_segments + 1) + off)
    x0 = np.concatenate(verts)
    x1 = x0.copy()
    speed = np.abs(rng.normal(0.0, sigma, sv.shape[0]))
    x1[nbody:] = sv - dt * speed[:, None] * radial
    return ClothScene(np.concatenate(faces), x0, x1, np.ones(x0.shape[0]))


def apply_corrections(faces: np.ndarray, x1: np.ndarray, dx: np.ndarray, cnt: np.ndarray) -> np.ndarray:
    """Caller-side update after one impulse pass: every vertex moves by the mean of the
    position corrections its triangle corners received.  A corner whose summed correction
    is not finite (the reference accepts collisions with NaN parameters, whose normal is
    NaN) contributes neither its correction nor its count.  Returns the new end-of-step
    positions."""
    ok = np.isfinite(dx.reshape(-1, 3)).all(axis=1)
    vs = np.zeros_like(x1)
    vc = np.zeros(x1.shape[0])
    np.add.at(vs, faces.reshape(-1)[ok], dx.reshape(-1, 3)[ok])
    np.add.at(vc, faces.reshape(-1)[ok], cnt.reshape(-1)[ok].astype(np.float64))
    out = x1.copy()
    hit = vc > 0
    out[hit] += vs[hit] / vc[hit, None]
    return out
