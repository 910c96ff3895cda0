"""Closed triangle meshes for the mesh -> SDF tests (outward-facing triangles) and
their exact signed distances."""
import numpy as np

from atac.workloads import box_mesh as box, icosphere, torus  # noqa: F401


def box_exact_sdf(p, a=1.0):
    """Signed distance to the cube [-a, a]^3 (negative inside)."""
    q = np.abs(p) - a
    out = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inn = np.minimum(np.max(q, axis=-1), 0.0)
    return out + inn


def torus_exact_sdf(p, R=1.0, r=0.35):
    xy = np.linalg.norm(p[..., :2], axis=-1)
    return np.sqrt((xy - R) ** 2 + p[..., 2] ** 2) - r
