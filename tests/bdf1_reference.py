"""FP64 CPU reference of the batched BDF1 stepping (test infrastructure): the
reference's Simulation::integration_BDF1 / newton (Simulation.cpp:999-1098,
1248-1270) for one part on a translational joint, with the contact forces and their
derivatives from the FP64 oracle (ForceGeneralSDFContact::computeForceWithDerivative)."""
import numpy as np

from atac import workloads as W
from oracle import oracle as O


class ReferenceBDF1:
    def __init__(self, parts, contacts, kn, kt, mu, damping, col_th, h, gravity):
        """parts: [(points (n,3) FP64 or empty, grid (SDFGrid) or None, E0 4x4)]."""
        self.parts = parts
        self.grids = [O.GridRef.from_grid(g) if g is not None else None for _, g, _ in parts]
        self.pts = [W.round_points_f32(np.asarray(p, dtype=np.float64).reshape(-1, 3)) for p, _, _ in parts]
        self.contacts = contacts
        self.kn, self.kt, self.mu, self.damping, self.col_th = kn, kt, mu, damping, col_th
        self.h = h
        self.g = np.asarray(gravity, dtype=np.float64)

    def joint_force(self, move, mass, u, x, xdot, deriv):
        E0 = self.parts[move][2]
        R = E0[:3, :3]
        f = np.asarray(u, dtype=np.float64) + mass * self.g
        K = np.zeros((3, 3))
        D = np.zeros((3, 3))
        for a, b in self.contacts:
            if move not in (a, b):
                continue
            E = [self.parts[a][2].copy(), self.parts[b][2].copy()]
            phi = [np.zeros(6), np.zeros(6)]
            s = 0 if a == move else 1
            E[s][:3, 3] = E0[:3, 3] + x
            phi[s][3:] = R.T @ xdot
            gi = [g for g in self.grids]
            inst = O.make_instance(0, self.pts[a], E[0], phi[0], E[1], phi[1], self.col_th, self.kn, self.kt, self.mu,
                                   self.damping)
            r = O.eval_instance([gi[b]], inst, O.FLAG_DERIV if deriv else 0)
            o = 3 if s == 0 else 9
            f = f + R @ r["fm"][o:o + 3]
            if deriv:
                K += R @ r["Km"][o:o + 3, o:o + 3] @ R.T
                D += R @ r["Dm"][o:o + 3, o:o + 3] @ R.T
        return f, K, D

    def step(self, move, mass, q0, qd0, u, n=1):
        h = self.h
        q0 = np.asarray(q0, dtype=np.float64).copy()
        qd0 = np.asarray(qd0, dtype=np.float64).copy()
        info = []
        for _ in range(n):
            x = q0 + h * qd0
            fail, conv, iters = 0, False, 0
            for _it in range(max(20 * 3, 100)):
                f, K, D = self.joint_force(move, mass, u, x, (x - q0) / h, True)
                g = mass * (x - q0 - h * qd0) - h * h * f
                H = mass * np.eye(3) - h * h * K - h * D
                dx = np.linalg.solve(H, -g)
                gn = np.linalg.norm(g)
                alpha, ok, gnew = 1.0, False, None
                for _t in range(20):
                    xt = x + alpha * dx
                    ft, _, _ = self.joint_force(move, mass, u, xt, (xt - q0) / h, False)
                    gnew = mass * (xt - q0 - h * qd0) - h * h * ft
                    if np.linalg.norm(gnew) < gn:
                        ok = True
                        break
                    alpha *= 0.5
                iters += 1
                if ok:
                    fail = 0
                else:
                    fail += 1
                    if fail >= 3:
                        break
                x = x + alpha * dx
                if np.linalg.norm(gnew) < 1e-8:
                    conv = True
                    break
            qd0 = (x - q0) / h
            q0 = x
            info.append((conv, iters))
        return q0, qd0, info
