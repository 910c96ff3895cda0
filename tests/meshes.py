"""Closed triangle meshes for the mesh -> SDF tests (outward-facing triangles)."""
import numpy as np


def box(a=1.0, center=(0.0, 0.0, 0.0)):
    """Axis-aligned cube [-a, a]^3 + center: 8 vertices, 12 triangles."""
    c = np.asarray(center, dtype=np.float64)
    V = np.array([[x, y, z] for z in (-a, a) for y in (-a, a) for x in (-a, a)], dtype=np.float64) + c
    # vertex index = x + 2y + 4z (bits); faces as quads split in two
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    F = []
    for q in quads:
        F.append((q[0], q[1], q[2]))
        F.append((q[0], q[2], q[3]))
    return V, np.array(F, dtype=np.int64)


def icosphere(r=1.0, levels=2, center=(0.0, 0.0, 0.0)):
    """Subdivided icosahedron projected to radius r."""
    t = (1.0 + 5 ** 0.5) / 2
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(levels):
        mid = {}
        NF = []

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                mid[key] = len(V) - 1
            return mid[key]
        for a, b, c in F:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            NF += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = NF
    return np.array(V) * r + np.asarray(center), np.array(F, dtype=np.int64)


def torus(R=1.0, r=0.35, nu=24, nv=12):
    """Torus around the z axis: tube radius r at distance R."""
    V, F = [], []
    for i in range(nu):
        u = 2 * np.pi * i / nu
        for j in range(nv):
            v = 2 * np.pi * j / nv
            V.append(((R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u), r * np.sin(v)))
    for i in range(nu):
        for j in range(nv):
            a, b = i * nv + j, ((i + 1) % nu) * nv + j
            c, d = ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv
            F.append((a, b, c))
            F.append((a, c, d))
    return np.array(V), np.array(F, dtype=np.int64)


def box_exact_sdf(p, a=1.0):
    """Signed distance to the cube [-a, a]^3 (negative inside)."""
    q = np.abs(p) - a
    out = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inn = np.minimum(np.max(q, axis=-1), 0.0)
    return out + inn


def torus_exact_sdf(p, R=1.0, r=0.35):
    xy = np.linalg.norm(p[..., :2], axis=-1)
    return np.sqrt((xy - R) ** 2 + p[..., 2] ** 2) - r
