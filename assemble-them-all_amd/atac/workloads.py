"""Procedural SDF parts, contact-point sets and poses for the contact workloads.

All grids are FP32 with x-fastest index order a[i + ni*(j + nj*k)] (the order of
sdfgen's Array3, simulation/externals/sdfgen/array3.h:70-79) and
max_box = min_box + (n-1)*dx exactly. Values are stored as numpy arrays of shape
(nk, nj, ni) so that C order equals that index order.

Workloads (seeds are 1000 + workload number):
  1 peg_in_hole    box 2x2x2 with a z-cylinder hole r=0.52 on a 64^3 grid (dx 0.05),
                   cylinder peg r=0.5 h=2 with 5,000 area-uniform surface points,
                   256 poses (tilt ~N(0,2 deg) about a random horizontal axis,
                   translation ~U(+-0.05)^3), twist phi1 ~N(0,0.1), phi2 = 0.
  2 blobs          10 superquadric pairs on 128^3 grids, 200k points, 64 poses per
                   pair placed by bisection so the max penetration is U[0,3%] of the
                   SDF body's bbox diagonal.
  3 batched        4,096 instances, N ~ log-uniform[1k, 50k], 16 distinct 128^3 grids
                   (cylindrical-hole boxes, square-hole boxes, blobs), pose = nominal
                   assembled pose + s*u, s ~U[0.01,0.1], u one of 26 unit directions.
  4 differentiated 1,024 instances (512 peg/hole + 512 blob), 20k points each.
  5 multi_part     a scene: a fixed base plate and 19 free parts (superquadrics and
                   square-hole blocks) resting on it, ~100k contact points and ~256^3
                   grid nodes each, every ordered pair a contact (multi_part_scene).
Force parameters for all of them: kn=1e4, damping=10, mu=0.3, kt=1e3, col_th=0.01.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

KN, DAMPING, MU, KT, COL_TH = 1.0e4, 10.0, 0.3, 1.0e3, 0.01


@dataclass
class SDFGrid:
    values: np.ndarray          # float32 (nk, nj, ni)
    min_box: np.ndarray         # float64 (3,)
    max_box: np.ndarray         # float64 (3,)
    dxyz: np.ndarray            # float64 (3,)

    @property
    def n(self):
        nk, nj, ni = self.values.shape
        return ni, nj, nk

    def node_coords(self):
        ni, nj, nk = self.n
        xs = self.min_box[0] + self.dxyz[0] * np.arange(ni)
        ys = self.min_box[1] + self.dxyz[1] * np.arange(nj)
        zs = self.min_box[2] + self.dxyz[2] * np.arange(nk)
        return xs, ys, zs


def make_grid(fn, n, min_box, dxyz) -> SDFGrid:
    """Sample fn(x, y, z) (broadcast FP64) on an n-node grid starting at min_box."""
    n = tuple(int(v) for v in n)
    min_box = np.asarray(min_box, dtype=np.float64)
    dxyz = np.asarray(dxyz, dtype=np.float64)
    max_box = min_box + (np.asarray(n) - 1) * dxyz
    xs = min_box[0] + dxyz[0] * np.arange(n[0])
    ys = min_box[1] + dxyz[1] * np.arange(n[1])
    zs = min_box[2] + dxyz[2] * np.arange(n[2])
    vals = np.empty((n[2], n[1], n[0]), dtype=np.float32)
    Y, X = np.meshgrid(ys, xs, indexing="ij")
    for k, z in enumerate(zs):
        vals[k] = fn(X, Y, np.full_like(X, z)).astype(np.float32)
    return SDFGrid(vals, min_box, max_box, dxyz)


# ------------------------------------------------------------------ analytic SDFs
def sdf_box(x, y, z, hx, hy, hz):
    qx, qy, qz = np.abs(x) - hx, np.abs(y) - hy, np.abs(z) - hz
    outside = np.sqrt(np.maximum(qx, 0) ** 2 + np.maximum(qy, 0) ** 2 + np.maximum(qz, 0) ** 2)
    inside = np.minimum(np.maximum(np.maximum(qx, qy), qz), 0.0)
    return outside + inside


def sdf_box_cyl_hole(x, y, z, half=1.0, r=0.52):
    return np.maximum(sdf_box(x, y, z, half, half, half), -(np.sqrt(x * x + y * y) - r))


def sdf_box_square_hole(x, y, z, half=1.0, w=0.52):
    hole = sdf_box(x, y, 0.0 * z, w, w, 1e9)
    return np.maximum(sdf_box(x, y, z, half, half, half), -hole)


def superquadric_G(x, y, z, axes, eps):
    a, b, c = axes
    e1, e2 = eps
    ax, ay, az = np.abs(x / a), np.abs(y / b), np.abs(z / c)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        t = (ax ** (2.0 / e2) + ay ** (2.0 / e2)) ** (e2 / e1) + az ** (2.0 / e1)
        return t ** (e1 / 2.0)


def sdf_superquadric(x, y, z, axes, eps):
    """Radial approximate SDF |x| (1 - 1/G(x)); G is the degree-1 homogeneous
    inside-outside function, so x / G(x) is the surface point on the ray."""
    r = np.sqrt(x * x + y * y + z * z)
    G = superquadric_G(x, y, z, axes, eps)
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = r * (1.0 - 1.0 / G)
    # at the origin use the radius of the surface along +x
    phi = np.where(r > 1e-12, phi, -axes[0])
    return phi


# ------------------------------------------------------------------ point sets
def sample_cylinder_surface(rng, n, r=0.5, h=2.0):
    side = 2 * np.pi * r * h
    caps = 2 * np.pi * r * r
    on_side = rng.random(n) < side / (side + caps)
    th = rng.random(n) * 2 * np.pi
    pts = np.empty((n, 3))
    ns = int(on_side.sum())
    pts[on_side, 0] = r * np.cos(th[on_side])
    pts[on_side, 1] = r * np.sin(th[on_side])
    pts[on_side, 2] = (rng.random(ns) - 0.5) * h
    nc = n - ns
    rr = r * np.sqrt(rng.random(nc))
    pts[~on_side, 0] = rr * np.cos(th[~on_side])
    pts[~on_side, 1] = rr * np.sin(th[~on_side])
    pts[~on_side, 2] = np.where(rng.random(nc) < 0.5, -0.5 * h, 0.5 * h)
    return pts


def sample_square_prism_surface(rng, n, w=0.5, h=2.0):
    side_area = 4 * (2 * w) * h
    cap_area = 2 * (2 * w) ** 2
    on_side = rng.random(n) < side_area / (side_area + cap_area)
    pts = np.empty((n, 3))
    ns = int(on_side.sum())
    face = rng.integers(0, 4, ns)
    u = (rng.random(ns) * 2 - 1) * w
    zz = (rng.random(ns) - 0.5) * h
    sx = np.select([face == 0, face == 1, face == 2, face == 3], [np.full(ns, w), np.full(ns, -w), u, u])
    sy = np.select([face == 0, face == 1, face == 2, face == 3], [u, u, np.full(ns, w), np.full(ns, -w)])
    pts[on_side] = np.stack([sx, sy, zz], axis=1)
    nc = n - ns
    pts[~on_side, 0] = (rng.random(nc) * 2 - 1) * w
    pts[~on_side, 1] = (rng.random(nc) * 2 - 1) * w
    pts[~on_side, 2] = np.where(rng.random(nc) < 0.5, -0.5 * h, 0.5 * h)
    return pts


def sample_superquadric_surface(rng, n, axes, eps):
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    G = superquadric_G(u[:, 0], u[:, 1], u[:, 2], axes, eps)
    return u / G[:, None]


def random_superquadric(rng):
    axes = tuple(rng.uniform(0.6, 1.0, 3))
    eps = tuple(rng.uniform(0.3, 3.0, 2))
    return axes, eps


def superquadric_grid(axes, eps, n=128, pad=0.10) -> SDFGrid:
    ext = np.asarray(axes) * (1.0 + pad)
    dxyz = 2 * ext / (n - 1)
    return make_grid(lambda x, y, z: sdf_superquadric(x, y, z, axes, eps), (n, n, n), -ext, dxyz)


# ------------------------------------------------------------------ poses
def rotation_exp(w):
    w = np.asarray(w, dtype=np.float64)
    th = np.linalg.norm(w)
    if th < 1e-15:
        return np.eye(3)
    k = w / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def pose(R, t):
    E = np.eye(4)
    E[:3, :3] = R
    E[:3, 3] = t
    return E


def cube_directions():
    """The 26 unit directions to the faces, edges and corners of a cube."""
    d = []
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                if a or b or c:
                    v = np.array([a, b, c], dtype=np.float64)
                    d.append(v / np.linalg.norm(v))
    return np.array(d)


def trilinear_np(grid: SDFGrid, x):
    """FP64 trilinear lookup (FLT_MAX outside the box), numpy, for pose placement."""
    x = np.asarray(x, dtype=np.float64)
    ni, nj, nk = grid.n
    c = (x - grid.min_box) / grid.dxyz
    inside = np.all((x >= grid.min_box) & (x <= grid.max_box), axis=1)
    i0 = np.clip(np.floor(c).astype(np.int64), 0, np.array([ni, nj, nk]) - 1)
    f = c - i0
    i1 = np.minimum(i0 + 1, np.array([ni, nj, nk]) - 1)
    v = grid.values.reshape(-1).astype(np.float64)

    def at(ii, jj, kk):
        return v[ii + ni * (jj + nj * kk)]

    fx, fy, fz = f[:, 0], f[:, 1], f[:, 2]
    a0, a1 = i0.T, i1.T
    c00 = (1 - fx) * at(a0[0], a0[1], a0[2]) + fx * at(a1[0], a0[1], a0[2])
    c10 = (1 - fx) * at(a0[0], a1[1], a0[2]) + fx * at(a1[0], a1[1], a0[2])
    c01 = (1 - fx) * at(a0[0], a0[1], a1[2]) + fx * at(a1[0], a0[1], a1[2])
    c11 = (1 - fx) * at(a0[0], a1[1], a1[2]) + fx * at(a1[0], a1[1], a1[2])
    c0 = (1 - fy) * c00 + fy * c10
    c1 = (1 - fy) * c01 + fy * c11
    out = (1 - fz) * c0 + fz * c1
    return np.where(inside, out, np.finfo(np.float32).max)


def place_by_bisection(grid: SDFGrid, pts, R, u, target_pen, s_hi=4.0, iters=40):
    """Translation t = s*u (SDF body at the origin) with max penetration
    max(-phi(R p + t)) equal to target_pen, found by bisection on s."""
    def pen(s):
        x = pts @ R.T + s * u
        return float(np.max(-trilinear_np(grid, x)))

    lo, hi = 0.0, s_hi
    if pen(hi) > target_pen:
        return hi * u
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        if pen(mid) > target_pen:
            lo = mid
        else:
            hi = mid
    return hi * u


def place_by_bisection_many(grid: SDFGrid, pts, Rs, us, targets, s_hi=4.0, iters=40):
    """place_by_bisection for many (R, u, target) at once: the same bisection per
    pose, run in lockstep over one trilinear lookup per step (the same values)."""
    m = len(Rs)
    X = np.stack([pts @ R.T for R in Rs])                       # (m, n, 3), as place_by_bisection forms them
    us = np.asarray(us, dtype=np.float64).reshape(m, 3)
    targets = np.asarray(targets, dtype=np.float64)

    def pen(s):
        x = X + s[:, None, None] * us[:, None, :]
        return np.max(-trilinear_np(grid, x.reshape(-1, 3)).reshape(m, -1), axis=1)

    lo, hi = np.zeros(m), np.full(m, s_hi)
    done = pen(hi) > targets                                    # already past the target at s_hi
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        over = pen(mid) > targets
        lo = np.where(over, mid, lo)
        hi = np.where(over, hi, mid)
    hi = np.where(done, s_hi, hi)
    return hi[:, None] * us


# ------------------------------------------------------------------ workloads
@dataclass
class Instance:
    grid: int
    points: int                 # index into Workload.point_sets
    E1: np.ndarray
    phi1: np.ndarray
    E2: np.ndarray
    phi2: np.ndarray
    col_th: float = COL_TH
    kn: float = KN
    kt: float = KT
    mu: float = MU
    damping: float = DAMPING
    scale: float = 1.0
    contact_scale: float = 1.0


@dataclass
class Workload:
    name: str
    grids: list
    point_sets: list            # list of (n_i, 3) FP64 arrays (body frame of the contact body)
    instances: list
    meta: dict = field(default_factory=dict)

    @property
    def n_queries(self):
        return int(sum(len(self.point_sets[i.points]) for i in self.instances))


def peg_in_hole(seed=1001, n_points=5000, n_poses=256, kt=KT) -> Workload:
    rng = np.random.default_rng(seed)
    grid = make_grid(lambda x, y, z: sdf_box_cyl_hole(x, y, z, 1.0, 0.52), (64, 64, 64), (-1.575,) * 3, (0.05,) * 3)
    pts = sample_cylinder_surface(rng, n_points, 0.5, 2.0)
    inst = []
    for _ in range(n_poses):
        ang = np.deg2rad(rng.normal(0.0, 2.0))
        al = rng.uniform(0, 2 * np.pi)
        R = rotation_exp(ang * np.array([np.cos(al), np.sin(al), 0.0]))
        t = rng.uniform(-0.05, 0.05, 3)
        inst.append(Instance(0, 0, pose(R, t), rng.normal(0, 0.1, 6), np.eye(4), np.zeros(6), kt=kt))
    return Workload("peg_in_hole_64^3_5k_x256", [grid], [pts], inst,
                    {"seed": seed, "kn": KN, "kt": kt, "mu": MU, "damping": DAMPING, "col_th": COL_TH})


def blobs(seed=1002, n_pairs=10, n_points=200_000, n_poses=64, grid_n=128, place_sample=4000,
          progress=None) -> Workload:
    """progress: optional callable, called with the pair index once a pair is placed."""
    rng = np.random.default_rng(seed)
    grids, psets, inst = [], [], []
    for p in range(n_pairs):
        axB, epB = random_superquadric(rng)
        axA, epA = random_superquadric(rng)
        g = superquadric_grid(axB, epB, grid_n)
        grids.append(g)
        pts = sample_superquadric_surface(rng, n_points, axA, epA)
        psets.append(pts)
        diag = float(np.linalg.norm(g.max_box - g.min_box))
        sub = pts[rng.choice(len(pts), min(place_sample, len(pts)), replace=False)]
        Rs, us, targets, phis = [], [], [], []
        for _ in range(n_poses):   # the random draws in their order; the placement uses none
            Rs.append(random_rotation(rng))
            u = rng.normal(size=3)
            us.append(u / np.linalg.norm(u))
            targets.append(rng.uniform(0.0, 0.03) * diag)
            phis.append(rng.normal(0, 0.1, 6))
        ts = place_by_bisection_many(g, sub, Rs, us, targets)
        for R, t, phi in zip(Rs, ts, phis):
            inst.append(Instance(p, p, pose(R, t), phi, np.eye(4), np.zeros(6)))
        if progress is not None:
            progress(p)
    return Workload(f"blobs_{grid_n}^3_{n_points}_x{n_pairs * n_poses}", grids, psets, inst, {"seed": seed})


def _part_library(rng, grid_n=128):
    """16 SDF parts with the contact body that assembles into each and its
    nominal assembled pose: 6 cylindrical-hole boxes, 4 square-hole boxes,
    6 superquadric blobs (nominal = just touching along a random direction)."""
    lib = []
    ext = 1.575
    dx = 2 * ext / (grid_n - 1)
    for r in (0.52, 0.53, 0.54, 0.55, 0.56, 0.58):
        g = make_grid(lambda x, y, z, r=r: sdf_box_cyl_hole(x, y, z, 1.0, r), (grid_n,) * 3, (-ext,) * 3, (dx,) * 3)
        lib.append((g, ("cyl", 0.5, 2.0), np.eye(4)))
    for w in (0.52, 0.53, 0.54, 0.55):
        g = make_grid(lambda x, y, z, w=w: sdf_box_square_hole(x, y, z, 1.0, w), (grid_n,) * 3, (-ext,) * 3, (dx,) * 3)
        lib.append((g, ("sq", 0.5, 2.0), np.eye(4)))
    for _ in range(6):
        axB, epB = random_superquadric(rng)
        axA, epA = random_superquadric(rng)
        g = superquadric_grid(axB, epB, grid_n)
        sub = sample_superquadric_surface(rng, 4000, axA, epA)
        R = random_rotation(rng)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        t = place_by_bisection(g, sub, R, u, 0.0)
        lib.append((g, ("sq_blob", axA, epA), pose(R, t)))
    return lib


def _sample_part(rng, kind, n):
    if kind[0] == "cyl":
        return sample_cylinder_surface(rng, n, kind[1], kind[2])
    if kind[0] == "sq":
        return sample_square_prism_surface(rng, n, kind[1], kind[2])
    return sample_superquadric_surface(rng, n, kind[1], kind[2])


def batched(seed=1003, n_inst=4096, n_min=1000, n_max=50_000, grid_n=128, fixed_n=None, kinds=None) -> Workload:
    rng = np.random.default_rng(seed)
    lib = _part_library(rng, grid_n)
    dirs = cube_directions()
    grids = [g for g, _, _ in lib]
    psets, inst = [], []
    for i in range(n_inst):
        gi = int(rng.integers(0, len(lib))) if kinds is None else int(kinds[i % len(kinds)])
        g, kind, E_nom = lib[gi]
        n = int(fixed_n) if fixed_n else int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        psets.append(_sample_part(rng, kind, n))
        s = rng.uniform(0.01, 0.1)
        u = dirs[rng.integers(0, len(dirs))]
        E1 = E_nom.copy()
        E1[:3, 3] += s * u
        inst.append(Instance(gi, len(psets) - 1, E1, rng.normal(0, 0.1, 6), np.eye(4), np.zeros(6)))
    return Workload(f"batched_{n_inst}_inst_{grid_n}^3", grids, psets, inst, {"seed": seed})


def differentiated(seed=1004, n_inst=1024, n_points=20_000, grid_n=128) -> Workload:
    """512 peg/hole (library parts 0-9) and 512 blob (parts 10-15) instances."""
    kinds = [i % 10 if (i % 2 == 0) else 10 + (i // 2) % 6 for i in range(n_inst)]
    w = batched(seed=seed, n_inst=n_inst, fixed_n=n_points, grid_n=grid_n, kinds=kinds)
    w.name = f"differentiated_{n_inst}_inst_{n_points}_pts_{grid_n}^3"
    return w


def round_points_f32(pts):
    """The GPU path stores contact points in FP32; the oracle is fed the same
    rounded values (promoted back to FP64)."""
    return np.asarray(pts, dtype=np.float32).astype(np.float64)


# ------------------------------------------------------------------ closed triangle meshes
# (outward-facing triangles) for the mesh -> SDF builder

def box_mesh(a=1.0, center=(0.0, 0.0, 0.0)):
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


# ------------------------------------------------------------------ multi-part scene (workload 5)

def superquadric_mesh(axes, eps, n_lat=64, n_lon=128):
    """Closed mesh of the superquadric |x/a|^(2/e2) + |y/b|^(2/e2))^(e2/e1) + |z/c|^(2/e1) = 1:
    rings of the (eta, omega) parametrisation plus two pole vertices, outward triangles.
    (n_lat - 1) * n_lon + 2 vertices."""
    a, b, c = axes
    e1, e2 = eps

    def spow(v, e):
        return np.sign(v) * np.abs(v) ** e

    eta = -np.pi / 2 + np.pi * np.arange(1, n_lat) / n_lat
    om = -np.pi + 2 * np.pi * np.arange(n_lon) / n_lon
    E, W = np.meshgrid(eta, om, indexing="ij")
    x = a * spow(np.cos(E), e1) * spow(np.cos(W), e2)
    y = b * spow(np.cos(E), e1) * spow(np.sin(W), e2)
    z = c * spow(np.sin(E), e1)
    V = np.concatenate([np.stack([x, y, z], axis=-1).reshape(-1, 3), [[0, 0, -c], [0, 0, c]]])
    south, north = len(V) - 2, len(V) - 1
    F = []
    ring = lambda i, j: i * n_lon + (j % n_lon)
    for i in range(n_lat - 2):
        for j in range(n_lon):
            p, q, r, s = ring(i, j), ring(i, j + 1), ring(i + 1, j + 1), ring(i + 1, j)
            F += [[p, q, r], [p, r, s]]
    for j in range(n_lon):
        F.append([south, ring(0, j + 1), ring(0, j)])
        F.append([north, ring(n_lat - 2, j), ring(n_lat - 2, j + 1)])
    return V.astype(np.float64), np.array(F, dtype=np.int64)


def frame_mesh(outer, hole, height, step):
    """Closed mesh of a square block with a square through-hole along z: |x|, |y| <= outer,
    not both |x|, |y| < hole, |z| <= height/2. Every face is split on one lattice of
    spacing `step` (outer, hole and height/2 multiples of it), so faces meet vertex to
    vertex and vertices are shared; outward triangles."""
    na, nb, nh = int(round(outer / step)), int(round(hole / step)), int(round(height / 2 / step))
    index = {}
    V = []

    def vid(p):
        if p not in index:
            index[p] = len(V)
            V.append(p)
        return index[p]

    F = []

    def rect(axis, level, u0, u1, v0, v1, outward_positive):
        """Lattice rectangle on the plane coord[axis] = level (lattice units), spanning
        [u0, u1] x [v0, v1] in the two other axes (cyclic order), normal +axis or -axis."""
        u, v = (axis + 1) % 3, (axis + 2) % 3
        for i in range(u0, u1):
            for j in range(v0, v1):
                def c(di, dj):
                    p = [0, 0, 0]
                    p[axis], p[u], p[v] = level, i + di, j + dj
                    return vid(tuple(p))
                q = [c(0, 0), c(1, 0), c(1, 1), c(0, 1)]
                if outward_positive:
                    F.extend([[q[0], q[1], q[2]], [q[0], q[2], q[3]]])
                else:
                    F.extend([[q[0], q[2], q[1]], [q[0], q[3], q[2]]])

    # outer walls x = +-outer (over y, z) and y = +-outer (over z, x)
    rect(0, na, -na, na, -nh, nh, True)
    rect(0, -na, -na, na, -nh, nh, False)
    rect(1, na, -nh, nh, -na, na, True)
    rect(1, -na, -nh, nh, -na, na, False)
    # hole walls, facing the hole
    rect(0, nb, -nb, nb, -nh, nh, False)
    rect(0, -nb, -nb, nb, -nh, nh, True)
    rect(1, nb, -nh, nh, -nb, nb, False)
    rect(1, -nb, -nh, nh, -nb, nb, True)
    # top and bottom annuli as four rectangles each (axis 2: u = x, v = y)
    for level, up in ((nh, True), (-nh, False)):
        rect(2, level, -na, na, nb, na, up)
        rect(2, level, -na, na, -na, -nb, up)
        rect(2, level, nb, na, -nb, nb, up)
        rect(2, level, -na, -nb, -nb, nb, up)
    return np.array(V, dtype=np.float64) * step, np.array(F, dtype=np.int64)


def multi_part_scene(folder, seed=1005, n_parts=20, points_per_part=100_000, grid_res=250, kn=1e4, kt=1e3, mu=0.3,
                     damping=10.0, col_th=0.01, timestep=1e-3, gravity=9.8):
    """Workload 5: a fixed base plate and n_parts - 1 free parts (free3d-exp joints) resting
    on it in a grid, superquadrics and square-hole blocks alternating; each part's mesh
    has about points_per_part vertices (its contact points) and its SDF grid about
    grid_res + 3 nodes per axis. Every ordered pair of parts gets a general_SDF_contact
    (the AABB broad phase keeps the ones that can touch). Writes part<k>.obj into
    `folder` and returns (xml string, part names)."""
    import os

    rng = np.random.default_rng(seed)
    names = [f"part{k}" for k in range(n_parts)]
    n_free = n_parts - 1
    cols = int(np.ceil(np.sqrt(n_free * 1.25)))
    rows = int(np.ceil(n_free / cols))
    pitch = 2.6
    plate = (cols * pitch + 1.0, rows * pitch + 1.0, 0.5)
    # base plate: a box with the requested vertex count
    side = np.sqrt(points_per_part / (2 * (plate[0] * plate[1] + plate[0] * plate[2] + plate[1] * plate[2])))
    step = 1.0 / max(side, 1e-9)
    nx, ny, nz = [max(1, int(round(L / step))) for L in plate]
    from itertools import product
    index, V, F = {}, [], []

    def vid(p):
        if p not in index:
            index[p] = len(V)
            V.append(p)
        return index[p]

    n = (nx, ny, nz)
    for axis in range(3):
        u, v = (axis + 1) % 3, (axis + 2) % 3
        for side_, pos in ((0, False), (n[axis], True)):
            for i, j in product(range(n[u]), range(n[v])):
                def c(di, dj):
                    p = [0, 0, 0]
                    p[axis], p[u], p[v] = side_, i + di, j + dj
                    return vid(tuple(p))
                q = [c(0, 0), c(1, 0), c(1, 1), c(0, 1)]
                F.extend([[q[0], q[1], q[2]], [q[0], q[2], q[3]]] if pos else [[q[0], q[2], q[1]], [q[0], q[3], q[2]]])
    Vp = (np.array(V, dtype=np.float64) / np.array(n) - 0.5) * np.array(plate)
    _write_obj(os.path.join(folder, "part0.obj"), Vp, np.array(F))
    links = [_scene_link("part0", "fixed", "part0.obj", (0.0, 0.0, -plate[2] / 2), col_th, grid_res)]
    top = 0.0
    for k in range(1, n_parts):
        r, c = divmod(k - 1, cols)
        cx = (c - (cols - 1) / 2) * pitch
        cy = (r - (rows - 1) / 2) * pitch
        if k % 2 == 1:
            axes = rng.uniform(0.55, 1.0, 3)
            eps = rng.uniform(0.3, 1.5, 2)
            n_lat = int(np.sqrt(points_per_part / 2))
            Vk, Fk = superquadric_mesh(axes, eps, n_lat=n_lat, n_lon=2 * n_lat)
        else:
            outer, hole, height = 1.0, 0.5, 0.8
            area = 8 * outer * height + 8 * hole * height + 2 * (4 * outer * outer - 4 * hole * hole)
            s = np.sqrt(area / points_per_part)
            s = 0.1 / max(1, int(round(0.1 / s)))         # outer, hole, height/2 multiples of s
            Vk, Fk = frame_mesh(outer, hole, height, s)
        _write_obj(os.path.join(folder, f"part{k}.obj"), Vk, Fk)
        # resting on the plate, its lowest vertex 0.2 mm past the contact threshold col_th
        z = top - Vk[:, 2].min() - (col_th + 2e-4)
        links.append(_scene_link(f"part{k}", "free3d-exp", f"part{k}.obj", (cx, cy, z), col_th, grid_res))
    contacts = "".join(f'<general_SDF_contact general_body="{a}" SDF_body="{b}"/>'
                       for a in names for b in names if a != b)
    xml = (f'<redmax model="multi_part"><option integrator="BDF1" timestep="{timestep}" gravity="0 0 {-gravity}"/>'
           f'<default><general_SDF_contact kn="{kn}" kt="{kt}" mu="{mu}" damping="{damping}"/></default>'
           + "".join(links) + f"<contact>{contacts}</contact></redmax>")
    return xml, names


def _scene_link(name, joint, filename, pos, col_th, grid_res):
    p = " ".join("%.17g" % float(v) for v in pos)
    return (f'<robot><link name="{name}"><joint name="{name}" type="{joint}" pos="{p}" '
            f'quat="1 0 0 0" frame="WORLD" damping="0"/><body name="{name}" type="SDF" filename="{filename}" '
            f'pos="0 0 0" quat="1 0 0 0" scale="1 1 1" transform_type="OBJ_TO_JOINT" density="1" dx="1000" '
            f'res="{grid_res}" col_th="{col_th}"/></link></robot>')


def _write_obj(path, V, F):
    with open(path, "w") as f:
        f.write("".join("v %.17g %.17g %.17g\n" % tuple(v) for v in V))
        f.write("".join("f %d %d %d\n" % tuple(t) for t in (np.asarray(F) + 1)))


def subdivided_box_mesh(size, n, center=(0.0, 0.0, 0.0)):
    """A box of the given edge lengths, each face split into n x n squares of two outward
    triangles, vertices shared along edges: 6 n^2 + 2 vertices."""
    size = np.asarray(size, dtype=np.float64)
    index, V, F = {}, [], []

    def vid(p):
        if p not in index:
            index[p] = len(V)
            V.append(p)
        return index[p]

    for axis in range(3):
        u, v = (axis + 1) % 3, (axis + 2) % 3
        for side, up in ((0, False), (n, True)):
            for i in range(n):
                for j in range(n):
                    def c(di, dj):
                        p = [0, 0, 0]
                        p[axis], p[u], p[v] = side, i + di, j + dj
                        return vid(tuple(p))
                    q = [c(0, 0), c(1, 0), c(1, 1), c(0, 1)]
                    F.extend([[q[0], q[1], q[2]], [q[0], q[2], q[3]]] if up else [[q[0], q[2], q[1]], [q[0], q[3], q[2]]])
    V = (np.array(V, dtype=np.float64) / n - 0.5) * size + np.asarray(center)
    return V, np.array(F, dtype=np.int64)


def peg_frame_scene(folder, peg_points=5000, kn=1e4, kt=1e3, mu=0.3, damping=10.0, col_th=0.01, rotation=True):
    """The disassembly planner's two-part scene (examples/run_joint_plan.py:140-180): a square
    peg (0.9 x 0.9 x 1.6, about peg_points mesh vertices) in the 1 x 1 hole of a fixed block
    (2 x 2 x 0.8, square-hole frame), the peg on a free3d-exp joint (rotation=True) or a
    translational one, gravity 1e-12 as the planner sets it, contacts both ways. Writes the
    OBJ files into `folder`; returns the XML string."""
    import os

    n = max(1, int(round(np.sqrt((peg_points - 2) / 6.0))))
    Vp, Fp = subdivided_box_mesh((0.9, 0.9, 1.6), n)
    _write_obj(os.path.join(folder, "peg.obj"), Vp, Fp)
    Vf, Ff = frame_mesh(1.0, 0.5, 0.8, 0.05)
    _write_obj(os.path.join(folder, "block.obj"), Vf, Ff)
    joint = "free3d-exp" if rotation else "translational"
    body = ('<body name="{n}" type="SDF" filename="{n}.obj" pos="0 0 0" quat="1 0 0 0" scale="1 1 1" '
            f'transform_type="OBJ_TO_JOINT" density="1" dx="0.05" col_th="{col_th}" mu="0"/>')
    return (f'<redmax model="peg_frame"><option integrator="BDF1" timestep="1e-3" gravity="0. 0. 1e-12"/>'
            f'<default><general_SDF_contact kn="{kn}" kt="{kt}" mu="{mu}" damping="{damping}"/></default>'
            f'<robot><link name="peg"><joint name="peg" type="{joint}" axis="0. 0. 0." pos="0 0 0" quat="1 0 0 0" '
            f'frame="WORLD" damping="0"/>{body.format(n="peg")}</link></robot>'
            f'<robot><link name="block"><joint name="block" type="fixed" axis="0. 0. 0." pos="0 0 0" quat="1 0 0 0" '
            f'frame="WORLD" damping="0"/>{body.format(n="block")}</link></robot>'
            '<contact><general_SDF_contact general_body="peg" SDF_body="block"/>'
            '<general_SDF_contact general_body="block" SDF_body="peg"/></contact></redmax>')


# the planner's 12 actions for rotation (examples/run_joint_plan.py:593-606): [rotation; translation]
ROTATION_ACTIONS = np.array([[0, 0, 0, 0, 0, 1], [0, 0, 0, 0, 0, -1], [0, 0, 0, 0, 1, 0], [0, 0, 0, 0, -1, 0],
                             [0, 0, 0, 1, 0, 0], [0, 0, 0, -1, 0, 0], [0, 0, 1, 0, 0, 0], [0, 0, -1, 0, 0, 0],
                             [0, 1, 0, 0, 0, 0], [0, -1, 0, 0, 0, 0], [1, 0, 0, 0, 0, 0], [-1, 0, 0, 0, 0, 0]],
                            dtype=np.float64)


def action_force(action, force_mag):
    """PhysicsPlanner.apply_action (run_joint_plan.py:263-272): the unit action scaled by
    force_mag, the rotational part three times as strong, as [torque; force]."""
    a = np.asarray(action, dtype=np.float64)
    a = a / np.linalg.norm(a) * force_mag
    return np.concatenate([a[:3] * 3, a[3:]]) if len(a) == 6 else np.concatenate([np.zeros(3), a])
