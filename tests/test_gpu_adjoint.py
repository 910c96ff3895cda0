"""GPU adjoint (VJP) checks: against the oracle's Km/Dm (the reference's analytic
dfdq), against FP64 central finite differences (right-multiplied exp, active set
held fixed as in the reference FD tests, ForceGeneralSDFContact.cpp:1048-1110) and
d/d(kn) against a kn central difference."""
import numpy as np
import pytest
import torch

from atac import workloads as W
from atac.batch import batch_from_workload
from helpers import branch_stable_contacts, friction_terms, oracle_workload, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def setup(gpu):
    w = W.peg_in_hole(n_poses=24)
    ref = oracle_workload(w)
    b = batch_from_workload(w)
    return w, ref, b


def test_vjp_matches_oracle_jacobians(setup):
    w, ref, b = setup
    rng = np.random.default_rng(11)
    g = rng.normal(size=(b.n, 12))
    b.eval(checkpoint=True)
    grads = b.vjp(torch.tensor(g, device="cuda"))
    torch.cuda.synchronize()
    gq1, gq2 = grads["g_q1"].cpu().numpy(), grads["g_q2"].cpu().numpy()
    gp1, gp2 = grads["g_phi1"].cpu().numpy(), grads["g_phi2"].cpu().numpy()
    for i, r in enumerate(ref):
        if not r["n_active"]:
            assert np.all(gq1[i] == 0)
            continue
        Kc = g[i] @ r["Km"]
        Dc = g[i] @ r["Dm"]
        s = np.linalg.norm(g[i]) * r["scale_Km"]
        assert rel_err(np.concatenate([gq1[i], gq2[i]]), Kc, s) <= TOL
        sd = np.linalg.norm(g[i]) * max(r["scale_Dm"], 1e-30)
        assert rel_err(np.concatenate([gp1[i], gp2[i]]), Dc, sd) <= TOL


def _fixed_contact_fm(w, i, ids, E1=None, E2=None, kn=None):
    ins = w.instances[i]
    G = O.GridRef.from_grid(w.grids[ins.grid])
    pts = W.round_points_f32(w.point_sets[ins.points])[ids]
    return O.force_fixed_contacts(G, pts, ins.E1 if E1 is None else E1, ins.phi1, ins.E2 if E2 is None else E2,
                                  ins.phi2, ins.col_th, ins.kn if kn is None else kn, ins.kt, ins.mu, ins.damping)["fm"]


def multilinear_workload(n_poses=8, n_points=4000, seed=21, mu=W.MU, kt=W.KT):
    """A ground body whose field f = z + 0.05 x y + 0.03 y z - 0.02 x y z is multilinear, so
    trilinear interpolation reproduces it in every cell (no kinks at cell faces), and a
    point patch pressed into it with random tilt, offset and twists."""
    rng = np.random.default_rng(seed)
    grid = W.make_grid(lambda x, y, z: z + 0.05 * x * y + 0.03 * y * z - 0.02 * x * y * z, (41, 41, 41),
                       (-1.0, -1.0, -1.0), (0.05, 0.05, 0.05))
    pts = np.stack([rng.uniform(-0.4, 0.4, n_points), rng.uniform(-0.4, 0.4, n_points),
                    rng.uniform(-0.02, 0.0, n_points)], axis=1)
    inst = []
    for _ in range(n_poses):
        R = W.rotation_exp(rng.normal(0, 0.05, 3))
        t = np.array([rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1), rng.uniform(-0.01, 0.01)])
        E2 = W.pose(W.rotation_exp(rng.normal(0, 0.02, 3)), rng.normal(0, 0.01, 3))
        inst.append(W.Instance(0, 0, W.pose(R, t), rng.normal(0, 0.1, 6), E2, rng.normal(0, 0.1, 6), mu=mu, kt=kt))
    return W.Workload("multilinear_ground", [grid], [pts], inst)


@pytest.mark.parametrize("branch", ["static", "dynamic"])
def test_fp32_friction_branch_parity(gpu, branch):
    """The FP32 contact arithmetic of each friction branch against the oracle. On the GPU
    the FP32 quotients multiply by the hardware reciprocal (v_rcp_f32), which the host
    build of the same code does not run, so this pins that path. The twists are set so
    that nearly every contact is in one branch: static (sliding scaled by 1e-4) or
    dynamic (a fast tangential slide, where the friction scale mu*|fc|/(kt*|tdot|) and its
    derivative terms enter Km and Dm). Wrench, Km and Dm (the latter two through the
    evaluation's Km/Dm outputs and through the adjoint of random cotangents) within 1e-5
    of the per-contact scale."""
    w = multilinear_workload(n_poses=8, seed=31)
    for ins in w.instances:
        if branch == "static":
            ins.phi1 = ins.phi1 * 1e-4
            ins.phi2 = ins.phi2 * 1e-4
        else:
            ins.phi1 = ins.phi1 + np.array([0.0, 0.0, 0.0, 2.0, -1.5, 0.0])
    ref = oracle_workload(w)
    n_static = n_all = 0
    for ins, r in zip(w.instances, ref):
        a, bb = friction_terms(w, ins, np.nonzero(r["active"])[0])
        n_static += int(np.sum(a >= bb - 1e-8))
        n_all += len(a)
    frac = n_static / max(n_all, 1)
    print(f"  {branch}: {n_all} contacts, static fraction {frac:.4f}", flush=True)
    assert n_all > 1000
    assert frac >= 0.95 if branch == "static" else frac <= 0.05
    b = batch_from_workload(w)
    g = np.random.default_rng(5).normal(size=(b.n, 12))
    out = b.eval(deriv=True)
    b.eval(checkpoint=True)
    grads = b.vjp(torch.tensor(g, device="cuda"))
    torch.cuda.synchronize()
    o = {k: v.cpu().numpy() for k, v in out.items()}
    gr = {k: v.cpu().numpy() for k, v in grads.items()}
    worst = 0.0
    for i, r in enumerate(ref):
        assert int(o["n_active"][i]) == r["n_active"]
        sk, sd, gn = r["scale_Km"], max(r["scale_Dm"], 1e-300), np.linalg.norm(g[i])
        e = [rel_err(o["fm"][i], r["fm"], r["scale_fm"]), rel_err(o["Km"][i], r["Km"], sk),
             rel_err(o["Dm"][i], r["Dm"], sd),
             rel_err(np.concatenate([gr["g_q1"][i], gr["g_q2"][i]]), g[i] @ r["Km"], gn * sk),
             rel_err(np.concatenate([gr["g_phi1"][i], gr["g_phi2"][i]]), g[i] @ r["Dm"], gn * sd)]
        worst = max(worst, max(e))
    print(f"  {branch}: worst relative error {worst:.3g} (fm, Km, Dm, adjoint)", flush=True)
    assert worst <= TOL


def _fd_check(w, ref, h, tol, n_check=4, seed=12):
    """For up to n_check instances: the GPU adjoint over the contacts a step h cannot
    switch (friction branch, d = 0) against FP64 central FD of the oracle over the same
    fixed contacts, and d/d(kn) against a kn central difference."""
    rng = np.random.default_rng(seed)
    checked = 0
    for i, r in enumerate(ref):
        if not r["n_active"] or checked >= n_check:
            continue
        ins = w.instances[i]
        ids = branch_stable_contacts(w, ins, np.nonzero(r["active"])[0], h=h)
        if len(ids) < 10:
            continue
        print(f"[case] fd h={h} instance {i} with {len(ids)} contacts", flush=True)
        pts = W.round_points_f32(w.point_sets[ins.points])[ids]
        sub = W.Workload("fd_subset", [w.grids[ins.grid]], [pts],
                         [W.Instance(0, 0, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt, ins.mu,
                                     ins.damping, ins.scale, ins.contact_scale)])
        g = rng.normal(size=(1, 12))
        b = batch_from_workload(sub)
        out = b.eval(checkpoint=True)
        grads = b.vjp(torch.tensor(g, device="cuda"))
        torch.cuda.synchronize()
        assert int(out["n_active"][0].item()) == len(ids)
        an = np.concatenate([grads["g_q1"][0].cpu().numpy(), grads["g_q2"][0].cpu().numpy()])
        gkn = float(grads["g_kn"][0].item())
        all_ids = np.arange(len(ids))
        fd = np.zeros(12)
        for k in range(6):
            dq = np.zeros(6)
            dq[k] = h
            fp = _fixed_contact_fm(sub, 0, all_ids, E1=ins.E1 @ O.se3_exp(dq))
            fmn = _fixed_contact_fm(sub, 0, all_ids, E1=ins.E1 @ O.se3_exp(-dq))
            fd[k] = g[0] @ (fp - fmn) / (2 * h)
            fp = _fixed_contact_fm(sub, 0, all_ids, E2=ins.E2 @ O.se3_exp(dq))
            fmn = _fixed_contact_fm(sub, 0, all_ids, E2=ins.E2 @ O.se3_exp(-dq))
            fd[6 + k] = g[0] @ (fp - fmn) / (2 * h)
        G = O.GridRef.from_grid(w.grids[ins.grid])
        full = O.force_fixed_contacts(G, pts, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt, ins.mu,
                                      ins.damping, derivs=True)
        scale = np.linalg.norm(g[0]) * np.abs(full["Km"]).sum()
        assert rel_err(an, fd, scale) <= tol, (i, rel_err(an, fd, scale))
        dk = ins.kn * 1e-6
        fkp = _fixed_contact_fm(sub, 0, all_ids, kn=ins.kn + dk)
        fkm = _fixed_contact_fm(sub, 0, all_ids, kn=ins.kn - dk)
        fdk = g[0] @ (fkp - fkm) / (2 * dk)
        assert abs(gkn - fdk) <= tol * np.linalg.norm(g[0]) * np.abs(full["fm"]).sum() / ins.kn, (i, gkn, fdk)
        checked += 1
    assert checked > 0


def test_vjp_central_fd_h1e4_multilinear_field(gpu):
    """Central FD, h = 1e-4, FP64, on a field with no interpolation kinks."""
    w = multilinear_workload()
    _fd_check(w, oracle_workload(w), 1e-4, 1e-5)


def test_vjp_central_fd_peg_in_hole(setup):
    """Central FD on the peg-in-hole grid; h = 1e-6 keeps the FD stencil of almost every
    contact inside its trilinear cell (the interpolant has slope jumps at cell faces)."""
    w, ref, b = setup
    _fd_check(w, ref, 1e-6, 1e-4)


def test_vjp_requires_an_evaluation(setup):
    """The adjoint differentiates the last evaluation's records: a batch never
    evaluated (or re-planned since) has none."""
    import atac
    w, ref, _ = setup
    b = batch_from_workload(w)
    with pytest.raises(atac.AtacError):
        b.vjp(torch.zeros((b.n, 12), dtype=torch.float64, device="cuda"))
    b.eval()
    b.retune({"tile_points": 2048})
    with pytest.raises(atac.AtacError):
        b.vjp(torch.zeros((b.n, 12), dtype=torch.float64, device="cuda"))


def test_graph_with_adjoint(setup):
    w, ref, b = setup
    g = torch.tensor(np.random.default_rng(3).normal(size=(b.n, 12)), device="cuda")
    b.eval(checkpoint=True)
    eager = b.vjp(g)
    torch.cuda.synchronize()
    eager = {k: v.cpu().numpy() for k, v in eager.items()}
    out = b.alloc_outputs()
    grads = b.alloc_gradients()
    b.graph_capture(out, checkpoint=True, g_fm=g, grads=grads)
    for rep in range(3):
        for v in grads.values():
            v.zero_()
        b.graph_launch()
        torch.cuda.synchronize()
        for k in eager:
            assert np.array_equal(grads[k].cpu().numpy(), eager[k]), (rep, k)
        torch.ones(100, device="cuda").sum().item()


def test_autograd_function(setup):
    """contact_wrench (torch.autograd.Function over the hand-written adjoint): the
    forward equals the batch evaluation; the backward's phi and kn gradients and the
    pairing of the pose gradients with tangent directions E xi^ equal the oracle's
    Km / Dm contractions; evaluating again between forward and backward raises."""
    from atac.autograd import contact_wrench
    w, ref, b = setup
    n = b.n
    dev = "cuda"
    E1 = torch.tensor(np.stack([ins.E1 for ins in w.instances]), device=dev, requires_grad=True)
    E2 = torch.tensor(np.stack([ins.E2 for ins in w.instances]), device=dev, requires_grad=True)
    phi1 = torch.tensor(np.stack([ins.phi1 for ins in w.instances]), device=dev, requires_grad=True)
    phi2 = torch.tensor(np.stack([ins.phi2 for ins in w.instances]), device=dev, requires_grad=True)
    kn = torch.tensor([ins.kn for ins in w.instances], device=dev, dtype=torch.float64, requires_grad=True)
    kt = torch.tensor([ins.kt for ins in w.instances], device=dev, dtype=torch.float64, requires_grad=True)
    fm = contact_wrench(b, E1, E2, phi1, phi2, kn, kt)
    direct = b.eval()["fm"]
    torch.cuda.synchronize()
    assert torch.equal(fm.detach(), direct)
    # the direct evaluation above replaced the forward's records
    with pytest.raises(RuntimeError):
        fm.sum().backward()
    fm = contact_wrench(b, E1, E2, phi1, phi2, kn, kt)
    rng = np.random.default_rng(17)
    g = rng.normal(size=(n, 12))
    (fm * torch.tensor(g, device=dev)).sum().backward()
    gE1, gE2 = E1.grad.cpu().numpy(), E2.grad.cpu().numpy()
    gp1, gp2 = phi1.grad.cpu().numpy(), phi2.grad.cpu().numpy()
    for i, r in enumerate(ref):
        if not r["n_active"]:
            continue
        Kc, Dc = g[i] @ r["Km"], g[i] @ r["Dm"]
        s = np.linalg.norm(g[i]) * r["scale_Km"]
        for _ in range(3):
            xi = rng.normal(size=6)
            hat = np.zeros((4, 4))
            hat[:3, :3] = [[0, -xi[2], xi[1]], [xi[2], 0, -xi[0]], [-xi[1], xi[0], 0]]
            hat[:3, 3] = xi[3:]
            p1 = np.sum(gE1[i] * (w.instances[i].E1 @ hat))
            p2 = np.sum(gE2[i] * (w.instances[i].E2 @ hat))
            assert abs(p1 - Kc[:6] @ xi) <= TOL * s * np.linalg.norm(xi)
            assert abs(p2 - Kc[6:] @ xi) <= TOL * s * np.linalg.norm(xi)
        sd = np.linalg.norm(g[i]) * max(r["scale_Dm"], 1e-30)
        assert rel_err(np.concatenate([gp1[i], gp2[i]]), Dc, sd) <= TOL
    assert torch.isfinite(kn.grad).all() and torch.isfinite(kt.grad).all()


def test_plain_path_adjoint_matches_batched(setup):
    """The plain per-instance baseline (one launch per force object, FP64 atomics)
    gives the batched path's wrench and adjoint (same contact arithmetic, another
    summation order)."""
    w, ref, b = setup
    g = torch.tensor(np.random.default_rng(4).normal(size=(b.n, 12)), device="cuda")
    out = b.eval()
    grads = b.vjp(g)
    pout, pgrads = b.eval_plain_vjp(g)
    torch.cuda.synchronize()
    assert torch.equal(out["n_active"], pout["n_active"])
    gn = g.cpu().numpy()
    for i, r in enumerate(ref):
        if not r["n_active"]:
            continue
        assert rel_err(pout["fm"][i].cpu().numpy(), out["fm"][i].cpu().numpy(), r["scale_fm"]) <= 1e-6
        s = np.linalg.norm(gn[i]) * r["scale_Km"]
        for k in ("g_q1", "g_q2"):
            assert rel_err(pgrads[k][i].cpu().numpy(), grads[k][i].cpu().numpy(), s) <= 1e-6, k
