"""Full-size parity runs of two benchmark configurations against the FP64 oracle,
with the error of every result reported.

* Differentiated path: 1,024 force objects (512 peg/hole, 512 superquadric blob)
  x 20,000 points. Forward wrench (+ contact-scale column) and the hand-written
  adjoint for a random cotangent against the oracle's g.Km and g.Dm (the
  reference's analytic dfdq) on every instance; d/d(kn) against a central
  difference in kn; FP64 central finite differences (h = 1e-4, right-multiplied
  exp) on blob instances over contacts such a step moves across no cell face,
  friction branch or sign of d.
* Dense single-pair instances: 10 superquadric pairs x 200,000 points x 64 poses
  (640 results): wrench, Km, Dm and the contact-scale column of every result.

Tolerance 1e-5 relative to the per-contact scale (the sum over contacts of each
contact's contribution norm), as in test_gpu_parity.py. Where the GPU's active
set differs from the oracle's (points within |d| < 1e-6 of the threshold), the
oracle is evaluated over the GPU's own contacts (force_fixed_contacts). The
friction branch (static or dynamic) has a band of its own: Km and Dm jump at the
switch, and a contact within 1e-6 (relative) of it can take either branch under
FP32 rounding while the wrench agrees; a result holding such a contact is
compared through the GPU's FP64 contact arithmetic (ATAC_EVAL_PRECISE), which
takes the oracle's branch."""
import os
import time

import numpy as np
import pytest
import torch

from atac import workloads as W
from atac.batch import batch_from_workload
from helpers import branch_stable_contacts, cell_stable_contacts, friction_margin, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
THREADS = min(16, os.cpu_count() or 1)


def _oracle_instances(w):
    pts = [W.round_points_f32(p) for p in w.point_sets]
    return [O.make_instance(i.grid, pts[i.points], i.E1, i.phi1, i.E2, i.phi2, i.col_th, i.kn, i.kt, i.mu, i.damping,
                            i.scale, i.contact_scale) for i in w.instances], pts


def _fixed(w, pts, i, ids, **over):
    ins = w.instances[i]
    G = O.GridRef.from_grid(w.grids[ins.grid])
    a = dict(E1=ins.E1, phi1=ins.phi1, E2=ins.E2, phi2=ins.phi2, kn=ins.kn)
    a.update(over)
    return O.force_fixed_contacts(G, pts[ins.points][ids], a["E1"], a["phi1"], a["E2"], a["phi2"], ins.col_th, a["kn"],
                                  ins.kt, ins.mu, ins.damping, ins.scale, ins.contact_scale, derivs=True)


def _summary(name, errs):
    e = np.asarray(errs)
    print(f"  {name}: {len(e)} results, max {e.max():.2e}, median {np.median(e):.2e}, 99th pct "
          f"{np.percentile(e, 99):.2e}", flush=True)
    return float(e.max())


@pytest.fixture(scope="module")
def cfg4(gpu):
    t0 = time.time()
    w = W.differentiated()
    print(f"[case] differentiated: {len(w.instances)} instances x 20000 points ({time.time() - t0:.1f} s to build)",
          flush=True)
    b = batch_from_workload(w)
    rng = np.random.default_rng(1004)
    g = rng.normal(size=(b.n, 12))
    out = b.eval(p6=True)
    grads = b.vjp(torch.tensor(g, device="cuda"))
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in out.items()}
    grads = {k: v.cpu().numpy() for k, v in grads.items()}
    insts, pts = _oracle_instances(w)
    t0 = time.time()
    ref = O.eval_batch([O.GridRef.from_grid(gr) for gr in w.grids], insts, O.FLAG_DERIV | O.FLAG_P6, threads=THREADS)
    print(f"  oracle over {len(insts)} instances on {THREADS} threads: {time.time() - t0:.1f} s", flush=True)
    return w, b, g, out, grads, ref, pts


def _near_branch_switch(w, b, i):
    return bool(np.any(friction_margin(w, w.instances[i], b.active_ids(i)) < 1e-6))


def _precise_one(w, i, g=None):
    """Instance i alone, FP64 contact arithmetic: wrench, Km, Dm, p6 (and the adjoint of g)."""
    ins = w.instances[i]
    sub = W.Workload("one", [w.grids[ins.grid]], [w.point_sets[ins.points]],
                     [W.Instance(0, 0, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt, ins.mu,
                                 ins.damping, ins.scale, ins.contact_scale)])
    b1 = batch_from_workload(sub)
    o = b1.eval(deriv=True, precise=True)
    gr = b1.vjp(torch.tensor(g[i:i + 1], device="cuda")) if g is not None else {}
    torch.cuda.synchronize()
    res = {k: v[0].cpu().numpy() for k, v in list(o.items()) + list(gr.items())}
    b1.close()
    return res


def _reference_of(w, b, pts, ref, out, i):
    """The oracle's result for instance i over the same contacts as the GPU."""
    if int(out["n_active"][i]) == int(ref["n_active"][i]):
        return {"fm": ref["fm"][i], "Km": ref["Km"][i], "Dm": ref["Dm"][i], "dfm_dp6": ref["dfm_dp6"][i]}
    return _fixed(w, pts, i, b.active_ids(i))


def test_differentiated_forward_and_adjoint_every_instance(cfg4):
    w, b, g, out, grads, ref, pts = cfg4
    errs = {k: {"peg": [], "blob": []} for k in ("fm", "p6", "g_q", "g_phi")}
    differ = near = 0
    for i in range(b.n):
        if int(ref["n_active"][i]) == 0 and int(out["n_active"][i]) == 0:
            assert np.all(out["fm"][i] == 0) and np.all(grads["g_q1"][i] == 0)
            continue
        differ += int(out["n_active"][i]) != int(ref["n_active"][i])
        r = _reference_of(w, b, pts, ref, out, i)
        o, gq = out, grads
        if _near_branch_switch(w, b, i):
            near += 1
            p = _precise_one(w, i, g)
            o = {k: {i: p[k]} for k in ("fm", "dfm_dp6")}
            gq = {k: {i: p[k]} for k in ("g_q1", "g_q2", "g_phi1", "g_phi2")}
        fam = "peg" if i % 2 == 0 else "blob"
        sf, sk, sd = ref["scale_fm"][i], ref["scale_Km"][i], ref["scale_Dm"][i]
        gn = np.linalg.norm(g[i])
        errs["fm"][fam].append(rel_err(o["fm"][i], r["fm"], sf))
        errs["p6"][fam].append(rel_err(o["dfm_dp6"][i], r["dfm_dp6"], sf))
        errs["g_q"][fam].append(rel_err(np.concatenate([gq["g_q1"][i], gq["g_q2"][i]]), g[i] @ r["Km"], gn * sk))
        errs["g_phi"][fam].append(rel_err(np.concatenate([gq["g_phi1"][i], gq["g_phi2"][i]]), g[i] @ r["Dm"],
                                          gn * max(sd, 1e-300)))
    print(f"  instances whose active set differs within the |d| < 1e-6 band: {differ}; with a contact within "
          f"1e-6 of the friction switch (compared in FP64): {near}", flush=True)
    worst = 0.0
    for k, fams in errs.items():
        for fam, e in fams.items():
            worst = max(worst, _summary(f"{k} ({fam})", e))
    assert worst <= TOL, worst


def test_differentiated_kn_derivative(cfg4):
    """d/d(kn) of g.fm against a central difference in kn (active set held fixed)."""
    w, b, g, out, grads, ref, pts = cfg4
    errs = []
    for i in range(0, b.n, 8):
        if not int(out["n_active"][i]):
            continue
        ids = b.active_ids(i)
        kn = w.instances[i].kn
        dk = kn * 1e-6
        fp = _fixed(w, pts, i, ids, kn=kn + dk)["fm"]
        fmn = _fixed(w, pts, i, ids, kn=kn - dk)["fm"]
        fd = g[i] @ (fp - fmn) / (2 * dk)
        scale = np.linalg.norm(g[i]) * ref["scale_fm"][i] / kn
        errs.append(abs(float(grads["g_kn"][i]) - fd) / scale)
    assert len(errs) >= 32
    assert _summary("g_kn vs central difference", errs) <= TOL


def test_differentiated_central_fd_blobs(cfg4):
    """FP64 central FD (h = 1e-4) on blob instances, over the contacts the step
    moves across no cell face, friction branch or sign of d."""
    w, b, g, out, grads, ref, pts = cfg4
    h = 1e-4
    checked, errs = 0, []
    for i in range(1, b.n, 2):   # odd instances are blobs
        if checked >= 6 or not int(out["n_active"][i]):
            continue
        ins = w.instances[i]
        ids = b.active_ids(i)
        ids = cell_stable_contacts(w, ins, branch_stable_contacts(w, ins, ids, h=h), h)
        if len(ids) < 20:
            continue
        sub = W.Workload("fd_subset", [w.grids[ins.grid]], [pts[ins.points][ids]],
                         [W.Instance(0, 0, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt, ins.mu,
                                     ins.damping, ins.scale, ins.contact_scale)])
        bs = batch_from_workload(sub)
        o = bs.eval()
        gr = bs.vjp(torch.tensor(g[i:i + 1], device="cuda"))
        torch.cuda.synchronize()
        assert int(o["n_active"][0].item()) == len(ids)
        an = np.concatenate([gr["g_q1"][0].cpu().numpy(), gr["g_q2"][0].cpu().numpy()])
        spts = [sub.point_sets[0]]
        all_ids = np.arange(len(ids))
        fd = np.zeros(12)
        for k in range(6):
            dq = np.zeros(6)
            dq[k] = h
            for body in (0, 1):
                key = "E1" if body == 0 else "E2"
                E = ins.E1 if body == 0 else ins.E2
                fp = _fixed(sub, spts, 0, all_ids, **{key: E @ O.se3_exp(dq)})["fm"]
                fmn = _fixed(sub, spts, 0, all_ids, **{key: E @ O.se3_exp(-dq)})["fm"]
                fd[6 * body + k] = g[i] @ (fp - fmn) / (2 * h)
        full = _fixed(sub, spts, 0, all_ids)
        e = rel_err(an, fd, np.linalg.norm(g[i]) * np.abs(full["Km"]).sum())
        print(f"  blob instance {i}: {len(ids)} contacts, FD error {e:.2e}", flush=True)
        errs.append(e)
        checked += 1
    assert checked >= 4
    assert max(errs) <= TOL, errs


@pytest.mark.timeout(600)
def test_blob_pairs_full_size_every_result(gpu):
    t0 = time.time()
    w = W.blobs(progress=lambda p: print(f"  pair {p} placed ({time.time() - t0:.0f} s)", flush=True))
    n_pts = sum(len(p) for p in w.point_sets)
    print(f"[case] blobs: {len(w.grids)} pairs x {n_pts // len(w.grids)} points x "
          f"{len(w.instances) // len(w.grids)} poses", flush=True)
    b = batch_from_workload(w)
    out = b.eval(deriv=True)
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in out.items()}
    print(f"  GPU: {int(out['n_active'].sum())} contacts over {len(w.instances)} results; record pool "
          f"{b.workspace()}", flush=True)
    insts, pts = _oracle_instances(w)
    t0 = time.time()
    ref = O.eval_batch([O.GridRef.from_grid(gr) for gr in w.grids], insts, O.FLAG_DERIV | O.FLAG_P6, threads=THREADS)
    print(f"  oracle: {time.time() - t0:.1f} s", flush=True)
    errs = {"fm": [], "Km": [], "Dm": [], "p6": []}
    differ = near = 0
    for i in range(b.n):
        if not int(ref["n_active"][i]) and not int(out["n_active"][i]):
            assert np.all(out["fm"][i] == 0)
            continue
        differ += int(out["n_active"][i]) != int(ref["n_active"][i])
        r = _reference_of(w, b, pts, ref, out, i)
        o = {k: v[i] for k, v in out.items()}
        if _near_branch_switch(w, b, i):
            near += 1
            o = _precise_one(w, i)
        errs["fm"].append(rel_err(o["fm"], r["fm"], ref["scale_fm"][i]))
        errs["Km"].append(rel_err(o["Km"], r["Km"], ref["scale_Km"][i]))
        errs["Dm"].append(rel_err(o["Dm"], r["Dm"], max(ref["scale_Dm"][i], 1e-300)))
        errs["p6"].append(rel_err(o["dfm_dp6"], r["dfm_dp6"], ref["scale_fm"][i]))
    print(f"  results whose active set differs within the |d| < 1e-6 band: {differ}; with a contact within 1e-6 "
          f"of the friction switch (compared in FP64): {near}", flush=True)
    assert len(errs["fm"]) > 500
    worst = max(_summary(k, e) for k, e in errs.items())
    assert worst <= TOL, worst
