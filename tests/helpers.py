"""Shared test helpers: error metrics and oracle runs of workloads."""
import numpy as np

from oracle import oracle as O
from atac import workloads as W


def rel_err(a, b, scale):
    """||a - b||_2 / scale (scale > 0)."""
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b))) / max(float(scale), 1e-300)


def oracle_workload(w, flags=O.FLAG_DERIV | O.FLAG_P6, instances=None, mode=O.MODE_REFERENCE):
    """Oracle results for every instance of a workload, on the FP32-rounded points."""
    grids = [O.GridRef.from_grid(g) for g in w.grids]
    pts = [W.round_points_f32(p) for p in w.point_sets]
    out = []
    for ins in (instances if instances is not None else w.instances):
        oi = O.make_instance(ins.grid, pts[ins.points], ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt,
                             ins.mu, ins.damping, ins.scale, ins.contact_scale, mode=mode)
        out.append(O.eval_instance(grids, oi, flags))
    return out


def band_mask(w, ins, d_band=1e-6):
    """Points whose oracle detection distance lies within the |d| < d_band band."""
    g = O.GridRef.from_grid(w.grids[ins.grid])
    pts = W.round_points_f32(w.point_sets[ins.points])
    xw = pts @ ins.E1[:3, :3].T + ins.E1[:3, 3]
    d = O.distance(g, ins.E2, ins.col_th, xw)
    return np.abs(d) < d_band


def friction_terms(w, ins, ids):
    """(mu*|fc1|, kt*|tdot|) of each contact from the oracle's BodySDFObj::collision
    values: the friction is static where the first is >= the second - 1e-8
    (ForceGeneralSDFContact.cpp:202), dynamic otherwise."""
    g = O.GridRef.from_grid(w.grids[ins.grid])
    xi = W.round_points_f32(w.point_sets[ins.points])[ids]
    R1, p1 = ins.E1[:3, :3], ins.E1[:3, 3]
    xw = xi @ R1.T + p1
    xw_dot = (np.cross(ins.phi1[:3], xi) + ins.phi1[3:]) @ R1.T
    c = O.collision(g, ins.E2, ins.phi2, ins.col_th, xw, xw_dot)
    kn = ins.kn * ins.scale * ins.contact_scale
    kt = ins.kt * ins.scale * ins.contact_scale
    lam = -kn * c["d"] + ins.damping * c["ddot"] * c["d"]
    n1 = c["n"] @ R1
    fcn = np.abs(lam) * np.sqrt(np.sum(np.cross(xi, n1) ** 2, axis=1) + np.sum(n1 ** 2, axis=1))
    return ins.mu * fcn, kt * np.linalg.norm(c["tdot"], axis=1)


def friction_margin(w, ins, ids):
    """Relative distance of each contact from the friction branch switch: the Jacobian
    jumps there, so where FP32 rounding (~1e-7) can move a contact across it the
    FP32 and FP64 arithmetic may differ in Km/Dm while the wrench agrees."""
    a, b = friction_terms(w, ins, ids)
    return np.abs(a - b + 1e-8) / np.maximum(np.maximum(a, b), 1e-300)


def branch_stable_contacts(w, ins, ids, rel_margin=0.05, h=1e-4):
    """Active contacts whose friction branch (static if mu*|fc1| >= kt*|tdot| - 1e-8,
    ForceGeneralSDFContact.cpp:202) is not within rel_margin of switching and whose depth
    exceeds what a pose step h can move (|fc1| = |lam|*|ni1| has a kink at d = 0), so a
    finite-difference step crosses neither. Uses the oracle's BodySDFObj::collision values."""
    g = O.GridRef.from_grid(w.grids[ins.grid])
    xi = W.round_points_f32(w.point_sets[ins.points])[ids]
    R1, p1 = ins.E1[:3, :3], ins.E1[:3, 3]
    w1, v1 = ins.phi1[:3], ins.phi1[3:]
    xw = xi @ R1.T + p1
    xw_dot = (np.cross(w1, xi) + v1) @ R1.T
    c = O.collision(g, ins.E2, ins.phi2, ins.col_th, xw, xw_dot)
    kn = ins.kn * ins.scale * ins.contact_scale
    kt = ins.kt * ins.scale * ins.contact_scale
    lam = -kn * c["d"] + ins.damping * c["ddot"] * c["d"]
    n1 = c["n"] @ R1            # R1^T n per row
    ni1 = np.sqrt(np.sum(np.cross(xi, n1) ** 2, axis=1) + np.sum(n1 ** 2, axis=1))
    fcn = np.abs(lam) * ni1
    tn = np.linalg.norm(c["tdot"], axis=1)
    a, b = ins.mu * fcn, kt * tn
    lever = np.linalg.norm(xi, axis=1) + 1.0
    keep = (np.abs(a - b) > rel_margin * np.maximum(a, b)) & (np.abs(c["d"]) > 50.0 * h * lever)
    return ids[keep]


def cell_stable_contacts(w, ins, ids, h, eps=1e-5):
    """Contacts whose point stays in its grid cell, with its FD-stencil reach eps to
    spare, under every right-multiplied pose step of size h of either body
    (E <- E exp(+-h e_k)): the wrench is smooth in the pose over such a step (the
    trilinear interpolant has slope jumps at cell faces)."""
    g = w.grids[ins.grid]
    xi = W.round_points_f32(w.point_sets[ins.points])[ids].astype(np.float64)

    def cells(E1, E2):
        x = (xi @ E1[:3, :3].T + E1[:3, 3] - E2[:3, 3]) @ E2[:3, :3]
        c = (x - g.min_box) / g.dxyz
        return np.floor(c), c - np.floor(c)

    base, _ = cells(ins.E1, ins.E2)
    keep = np.ones(len(ids), dtype=bool)
    reach = eps / g.dxyz
    for k in range(6):
        for sgn in (1.0, -1.0):
            dq = np.zeros(6)
            dq[k] = sgn * h
            for E1, E2 in ((ins.E1 @ O.se3_exp(dq), ins.E2), (ins.E1, ins.E2 @ O.se3_exp(dq))):
                c, f = cells(E1, E2)
                keep &= np.all(c == base, axis=1) & np.all((f > reach) & (f < 1.0 - reach), axis=1)
    return ids[keep]
