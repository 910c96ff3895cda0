"""FP64 reference integrator for scenes of atac._redmax.Simulation (test infrastructure).

The same model as the product, derived independently: each body sits on a joint to
the world (fixed, translational, or free3d-exp with q = [p; r], Q = [exp([r]x) p]); its
world pose is E_pj0 Q(q) E_ji. The body twists and the joint Jacobian J come from
complex-step derivatives of that pose map (not from the product's closed forms), the
bias term Jdot qdot from a central difference of J along qdot, and the reduced
dynamics are M = J^T Mm J, f = J^T (fm - Mm Jdot qdot) + fr (Simulation.cpp:379-419)
with the body forces restated (Body.cpp:250-273: Coriolis ad(phi)^T (I phi), gravity,
the inertia-scaled external force) and joint damping -Dr qdot (Joint.cpp:226-231). The
contact forces are the oracle's FP64 ForceGeneralSDFContact over the contacts the GPU's
FP32 detection decides (the host build of that decision, which the detection tests pin
against the oracle), so a GPU run with FP64 contact arithmetic computes the same thing.
BDF1 (Simulation.cpp:1081-1091, 1248-1262) is solved by Newton's method with a
finite-difference Jacobian to |g| < 1e-11.

The mass properties, grids and contact points are taken from the scene: the tests pin
them separately against closed forms.
"""
import numpy as np

from atac.batch import make_instance
from hostcheck import hostcheck as H
from oracle import oracle as O


def skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]], dtype=v.dtype)


def rodrigues(r):
    """exp([r]x), analytic in r (complex inputs allowed, for complex-step derivatives)."""
    s = r @ r
    K = skew(r)
    if abs(s) < 1e-12:
        return np.eye(3, dtype=r.dtype) + K + 0.5 * K @ K
    t = np.sqrt(s)
    return np.eye(3, dtype=r.dtype) + (np.sin(t) / t) * K + ((1 - np.cos(t)) / s) * K @ K


class RefGrid:
    def __init__(self, d):
        self.values = np.ascontiguousarray(d["values"], dtype=np.float32)
        self.min_box = np.asarray(d["min_box"], dtype=np.float64).ravel()
        self.max_box = np.asarray(d["max_box"], dtype=np.float64).ravel()
        self.dxyz = np.asarray(d["dxyz"], dtype=np.float64).ravel()
        self.col_th = float(d["col_th"])
        self.oracle = O.GridRef(self.values, self.min_box, self.max_box, self.dxyz)


class RefBody:
    def __init__(self, sim, name, joint_type, damping=0.0):
        self.name = name
        self.type = joint_type
        self.n = {"fixed": 0, "translational": 3, "free3d-exp": 6}[joint_type]
        self.Mm = np.asarray(sim.get_body_inertia(name), dtype=np.float64).ravel()
        self.E_ji = np.asarray(sim.get_body_Eji(name), dtype=np.float64)
        self.damping = damping
        self.points = np.asarray(sim.get_body_contact_points(name), dtype=np.float64)
        self.points32 = self.points.astype(np.float32).astype(np.float64)
        try:
            self.grid = RefGrid(sim.get_body_sdf(name))
        except Exception:
            self.grid = None
        self.f_ext = np.zeros(6)
        self.E_pj0 = None   # set by RefScene from the pose at q = 0

    def Q(self, q):
        Q = np.eye(4, dtype=q.dtype if self.n else np.float64)
        if self.type == "translational":
            Q[:3, 3] = q
        elif self.type == "free3d-exp":
            Q[:3, :3] = rodrigues(q[3:])
            Q[:3, 3] = q[:3]
        return Q

    def pose(self, q):
        return self.E_pj0 @ self.Q(q) @ self.E_ji


def vee(X):
    return np.array([X[2, 1], X[0, 2], X[1, 0], X[0, 3], X[1, 3], X[2, 3]])


class RefScene:
    def __init__(self, sim, joints, contacts, gravity, h):
        """joints: [(body name, joint type, damping)] in the scene's order; contacts:
        [(general body, SDF body, kn, kt, mu, damping)]."""
        self.bodies = [RefBody(sim, n, t, dmp) for n, t, dmp in joints]
        self.offsets = np.cumsum([0] + [b.n for b in self.bodies])
        self.n = int(self.offsets[-1])
        self.contacts = contacts
        self.g = np.asarray(gravity, dtype=np.float64)
        self.h = float(h)
        self.scale = 1.0

    def set_E_pj0(self, sim, q0):
        """Joint frames at q = 0 from the scene's poses at q0."""
        for b, o in zip(self.bodies, self.offsets):
            E = np.asarray(sim.get_body_E0i(b.name), dtype=np.float64)
            # E = E_pj0 Q(q0) E_ji  ->  E_pj0 = E (Q(q0) E_ji)^-1
            b.E_pj0 = E @ np.linalg.inv(b.Q(q0[o:o + b.n].astype(np.float64)) @ b.E_ji)

    def _q(self, q, b, o):
        return q[o:o + b.n]

    def jacobian(self, b, qb):
        """6 x n body Jacobian by complex step of the pose map."""
        E = b.pose(qb)
        Ei = np.linalg.inv(E)
        J = np.zeros((6, b.n))
        hs = 1e-30
        for k in range(b.n):
            dq = np.zeros(b.n, dtype=complex)
            dq[k] = 1j * hs
            dE = np.imag(b.pose(qb.astype(complex) + dq)) / hs
            J[:, k] = vee(Ei @ dE)
        return J

    def forces_max(self, q, qd):
        """fm (6 per body): body forces and the contact forces."""
        nb = len(self.bodies)
        fm = np.zeros(6 * nb)
        poses, twists = [], []
        for i, (b, o) in enumerate(zip(self.bodies, self.offsets)):
            qb, qdb = self._q(q, b, o), self._q(qd, b, o)
            E = b.pose(qb)
            phi = self.jacobian(b, qb) @ qdb if b.n else np.zeros(6)
            poses.append(E)
            twists.append(phi)
            R = E[:3, :3]
            w, v = phi[:3], phi[3:]
            Iw, mv = b.Mm[:3] * w, b.Mm[3:] * v
            fm[6 * i:6 * i + 3] += np.cross(Iw, w) + np.cross(mv, v) + b.Mm[:3] * (R.T @ b.f_ext[:3])
            fm[6 * i + 3:6 * i + 6] += np.cross(mv, w) + R.T @ (b.Mm[3] * self.g) + b.Mm[3:] * (R.T @ b.f_ext[3:])
        for gi, si, kn, kt, mu, dmp in self.contacts:
            gb, sb = self.bodies[gi], self.bodies[si]
            if len(gb.points) == 0:
                continue
            inst = make_instance(0, 0, 0, poses[gi], twists[gi], poses[si], twists[si], sb.grid.col_th, kn, kt, mu, dmp,
                                 self.scale, 1.0)
            act = H.detect_values(sb.grid, gb.points32, inst) < 0
            if not act.any():
                continue
            r = O.force_fixed_contacts(sb.grid.oracle, gb.points32[act], poses[gi], twists[gi], poses[si], twists[si],
                                       sb.grid.col_th, kn, kt, mu, dmp, self.scale, 1.0)
            fm[6 * gi:6 * gi + 6] += r["fm"][:6]
            fm[6 * si:6 * si + 6] += r["fm"][6:]
        return fm

    def matrices(self, q, qd):
        """M, f of M qdd = f in reduced coordinates."""
        fm = self.forces_max(q, qd)
        M = np.zeros((self.n, self.n))
        f = np.zeros(self.n)
        for i, (b, o) in enumerate(zip(self.bodies, self.offsets)):
            if b.n == 0:
                continue
            qb, qdb = self._q(q, b, o), self._q(qd, b, o)
            J = self.jacobian(b, qb)
            dt = 1e-5
            Jdqd = (self.jacobian(b, qb + dt * qdb) - self.jacobian(b, qb - dt * qdb)) @ qdb / (2 * dt)
            M[o:o + b.n, o:o + b.n] = J.T @ (b.Mm[:, None] * J)
            f[o:o + b.n] = J.T @ (fm[6 * i:6 * i + 6] - b.Mm * Jdqd) - b.damping * qdb
        return M, f

    def residual(self, q1, q0, qd0):
        qd1 = (q1 - q0) / self.h
        M, f = self.matrices(q1, qd1)
        return M @ (q1 - q0 - self.h * qd0) - self.h * self.h * f

    def step(self, q0, qd0, tol=1e-11, max_iter=50):
        q = q0 + self.h * qd0
        g = self.residual(q, q0, qd0)
        for _ in range(max_iter):
            if np.linalg.norm(g) < tol:
                break
            Jg = np.zeros((self.n, self.n))
            eps = 1e-7
            for k in range(self.n):
                e = np.zeros(self.n)
                e[k] = eps
                Jg[:, k] = (self.residual(q + e, q0, qd0) - self.residual(q - e, q0, qd0)) / (2 * eps)
            dx = np.linalg.solve(Jg, -g)
            alpha = 1.0
            for _ in range(30):
                gn = self.residual(q + alpha * dx, q0, qd0)
                if np.linalg.norm(gn) < np.linalg.norm(g):
                    break
                alpha *= 0.5
            q = q + alpha * dx
            g = gn
        return q, (q - q0) / self.h, float(np.linalg.norm(g))
