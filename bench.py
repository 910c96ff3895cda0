"""Contact-point SDF query throughput (force + adjoint) on one MI355X.

A step is one pass of the batched contact path over the whole workload with the
inputs resident in HBM: detection over every (instance, point) query, the contact
wrench of the active points, per-instance FP64 reduction, then the hand-written
adjoint (VJP) for a fixed synthetic cotangent g_fm. The step is replayed from a
captured HIP graph. value = queries per second, a query being one
(instance, contact point) pair.

Default workload: 4,096 force objects, N ~ log-uniform[1k, 50k] contact points each
(about 51M queries), 16 distinct 128^3 FP32 SDF grids (cylindrical-hole boxes,
square-hole boxes, superquadric blobs), pose = nominal assembled pose + s*u with
s ~ U[0.01, 0.1] along one of 26 cube directions; kn=1e4, damping=10, mu=0.3, kt=1e3.

Printed as one JSON line, with the dominant kernel's HBM roofline (algorithmic bytes
per launch over its HIP-event-timed launch duration), the FP64 CPU oracle on a
bounded sample of the same workload, the error against FP64 on that sample and the
plain per-instance GPU path for comparison.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "assemble-them-all_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "contact-point SDF queries/s (force+adjoint) at 1 GPU; % HBM roofline; err vs FP64"
# Algorithmic bytes of one detection launch (see roofline in main):
DETECT_INST_BYTES = 72 + 24 + 24 + 8 + 24 + 8 + 4   # instance fields detection reads: A, b, cmax, col_th,
                                                    # point pointers, point count, grid index
TILE_BYTES = 32 + 4 + 4 + 4                         # TileDev + queue entry + tile_active + tile_chunk0 writes
ACT_BYTES = 4                                       # active-list entry (point offset) per active point
CHUNK_BYTES = 16                                    # work-item entry, one per kChunk active points
RECORD_BYTES = 48                                   # FP32 active-contact record
K_CHUNK = 256
GROUP = 256                                         # points per culling group
GROUP_BYTES = 32 + 8 * 4                            # its box, and the 8 summed-area entries of its test


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name):
    from atac import workloads as W
    if name == "batched":
        return W.batched()
    if name == "differentiated":
        return W.differentiated()
    if name == "peg_in_hole":
        return W.peg_in_hole()
    raise SystemExit(f"unknown workload {name}")


def traffic_from_profiles(kernel="detect_kernel"):
    """HBM bytes per launch of the kernel from the committed rocprofv3 PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "detect_kernel_batched_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def host_cpu():
    """The host's CPU model (lscpu) and the cores this process may use."""
    model = None
    try:
        import subprocess
        for ln in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except Exception:
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0))}


def cpu_baseline(w, gpu, g, budget_s, threads, budget_1t_s):
    """FP64 oracle (reference-mode restatement, computeForceWithDerivative per force
    object: the reference's way to the gradient is Km/Dm, then the contraction) on the
    first instances of the workload, until about budget_s of CPU time on `threads`
    threads and budget_1t_s on one thread. Errors of the GPU step's outputs against
    it on every instance evaluated: the wrench and the adjoint (g.Km, g.Dm), each
    relative to the per-contact scale (the sum over contacts of each contact's
    contribution norm) and the wrench also to its own norm."""
    from oracle import oracle as O
    from atac import workloads as W
    grids = [O.GridRef.from_grid(g_) for g_ in w.grids]

    def oracle_insts(lo, hi):
        out = []
        for ins in w.instances[lo:hi]:
            pts = W.round_points_f32(w.point_sets[ins.points])
            out.append(O.make_instance(ins.grid, pts, ins.E1, ins.phi1, ins.E2, ins.phi2, ins.col_th, ins.kn, ins.kt,
                                       ins.mu, ins.damping, ins.scale, ins.contact_scale))
        return out

    t_total, q_total, n_done = 0.0, 0, 0
    errs = {"fm_sum": [], "fm_net": [], "vjp_q": [], "vjp_phi": []}
    band = 0
    chunk = 32
    while n_done < len(w.instances) and t_total < budget_s:
        oi = oracle_insts(n_done, n_done + chunk)
        t0 = time.perf_counter()
        r = O.eval_batch(grids, oi, O.FLAG_DERIV, threads=threads)
        t_total += time.perf_counter() - t0
        for k in range(len(oi)):
            i = n_done + k
            q_total += oi[k].n_points
            if int(gpu["n_active"][i]) != int(r["n_active"][k]):
                band += 1          # the active sets differ within the |d| < 1e-6 band: not compared
                continue
            if not r["n_active"][k]:
                continue
            ref = r["fm"][k]
            errs["fm_sum"].append(np.linalg.norm(gpu["fm"][i] - ref) / r["scale_fm"][k])
            if np.linalg.norm(ref) > 0:
                errs["fm_net"].append(np.linalg.norm(gpu["fm"][i] - ref) / np.linalg.norm(ref))
            gn = np.linalg.norm(g[i])
            gq = np.concatenate([gpu["g_q1"][i], gpu["g_q2"][i]])
            gp = np.concatenate([gpu["g_phi1"][i], gpu["g_phi2"][i]])
            errs["vjp_q"].append(np.linalg.norm(gq - g[i] @ r["Km"][k]) / (gn * r["scale_Km"][k]))
            errs["vjp_phi"].append(np.linalg.norm(gp - g[i] @ r["Dm"][k]) / (gn * max(r["scale_Dm"][k], 1e-300)))
        n_done += len(oi)
    # one thread, the reference's own setting (the planners pin OMP_NUM_THREADS=1)
    t1, q1, n1 = 0.0, 0, 0
    while n1 < len(w.instances) and t1 < budget_1t_s:
        oi = oracle_insts(n1, n1 + 4)
        t0 = time.perf_counter()
        O.eval_batch(grids, oi, O.FLAG_DERIV, threads=1)
        t1 += time.perf_counter() - t0
        q1 += sum(x.n_points for x in oi)
        n1 += len(oi)
    mx = lambda e: float(max(e)) if e else None
    dist = lambda e: ({"max": float(np.max(e)), "p99_9": float(np.percentile(e, 99.9)),
                       "median": float(np.median(e)), "above_1e-5": int(np.sum(np.asarray(e) > 1e-5))} if e else None)
    return {
        "value": q_total / t_total, "unit": "queries/s", "cores": threads, "kind": "port",
        "sample": f"first {n_done} of {len(w.instances)} force objects ({q_total} queries), FP64 oracle, "
                  f"detection + computeForceWithDerivative (FD stencils), OpenMP over force objects",
        "one_thread": {"value": q1 / t1, "unit": "queries/s", "cores": 1,
                       "sample": f"first {n1} force objects ({q1} queries)"},
        "host": host_cpu(),
    }, {
        "max_rel_err_wrench_vs_sum_of_contact_norms": mx(errs["fm_sum"]),
        "max_rel_err_wrench_vs_net": mx(errs["fm_net"]),
        "median_rel_err_wrench_vs_net": float(np.median(errs["fm_net"])) if errs["fm_net"] else None,
        "max_rel_err_vjp_pose_vs_g_Km": mx(errs["vjp_q"]),
        "max_rel_err_vjp_twist_vs_g_Dm": mx(errs["vjp_phi"]),
        # per-instance distributions (the Jacobian jumps at the friction branch switch: an
        # instance with a contact within FP32 rounding of it can differ there while the
        # wrench agrees)
        "wrench_vs_sum_of_contact_norms": dist(errs["fm_sum"]),
        "vjp_pose_vs_g_Km": dist(errs["vjp_q"]),
        "vjp_twist_vs_g_Dm": dist(errs["vjp_phi"]),
        "instances_checked": n_done - band, "instances_with_band_points_skipped": band,
    }


def sdf_build_measurement(cpu):
    """Mesh -> SDF grid (the level-set construction): GPU time for a 5,120-triangle
    sphere into 64^3 and 128^3 grids, and the CPU restatement at 64^3, with the
    number of nodes whose values differ (0 expected)."""
    import torch
    from atac import workloads as W
    from atac.sdfmesh import level_set
    V, F = W.icosphere(1.0, 4)
    res = {}
    for n in (64, 128):
        d = 2.6 / (n - 1)
        origin, dd, nn = (-1.3, -1.3, -1.3), (d, d, d), (n, n, n)
        level_set(V, F, origin, dd, nn)          # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        phi = level_set(V, F, origin, dd, nn)
        gpu_ms = (time.perf_counter() - t0) * 1e3
        e = {"grid": f"{n}^3", "triangles": int(len(F)), "gpu_ms": round(gpu_ms, 2)}
        if cpu and n == 64:
            from oracle import oracle as O
            t0 = time.perf_counter()
            ref = O.make_level_set3(V.astype(np.float32), F, origin, dd, nn)
            e["cpu_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
            e["cpu_kind"] = "port, 1 thread"
            e["nodes_differing"] = int((phi.view(np.uint32) != ref.view(np.uint32)).sum())
        res[f"{n}"] = e
    return res


def bdf1_measurement(cpu):
    """Implicit (BDF1) stepping of a peg on a translational joint in the hole, pushed
    along each of the 26 cube directions: 26 candidates stepped together, every Newton
    iteration and line-search trial of all of them one batched contact evaluation
    (FP32 contact arithmetic, FP64 Newton on the host)."""
    from atac import _redmax as R
    from atac import workloads as W
    w = W.peg_in_hole(n_points=5000, n_poses=8)
    g = w.grids[0]
    hole = R.BodySDFObj("hole", g.values.astype(np.float64), g.min_box, g.max_box, g.dxyz, 0.01, np.zeros((1, 3)), 0)
    E_peg = w.instances[3].E1.copy()
    parts = [(w.point_sets[0], None, E_peg), (np.zeros((0, 3)), hole, np.eye(4))]
    sim = R.BatchedBDF1(parts, [(0, 1)], kn=W.KN, kt=W.KT, mu=W.MU, damping=W.DAMPING, col_th=0.01, h=1e-3)
    dirs = W.cube_directions()
    for u in dirs:
        sim.add_candidate(0, 1.5, np.zeros(3), np.zeros(3), 500.0 * np.asarray(u))
    sim.step(1)                                  # builds the batch
    e0 = sim.n_evaluations
    n = 20
    t0 = time.perf_counter()
    sim.step(n)
    ms = (time.perf_counter() - t0) * 1e3 / n
    conv = sum(bool(sim.converged(c)) for c in range(len(dirs)))
    res = {"candidates": len(dirs), "contact_points": int(len(w.point_sets[0])), "ms_per_step": round(ms, 3),
           "batched_evaluations_per_step": round((sim.n_evaluations - e0) / n, 2),
           "converged_last_step": conv}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from bdf1_reference import ReferenceBDF1
        ref = ReferenceBDF1([(w.point_sets[0], None, E_peg), (np.zeros((0, 3)), g, np.eye(4))], [(0, 1)], W.KN, W.KT,
                            W.MU, W.DAMPING, 0.01, 1e-3, np.zeros(3))
        t0 = time.perf_counter()
        ref.step(0, 1.5, np.zeros(3), np.zeros(3), 500.0 * np.asarray(dirs[0]), 3)
        res["cpu_ms_per_step_per_candidate"] = round((time.perf_counter() - t0) * 1e3 / 3, 1)
        res["cpu_kind"] = "port (FP64 oracle forces, Python Newton), 1 thread"
    return res


def multipart_measurement(cpu, steps):
    """The multi-part step loop (workload 5): a fixed plate and 19 free parts (superquadrics
    and square-hole blocks on free3d-exp joints) resting on it under gravity, ~100k
    contact points and ~253^3 grid nodes each, all 380 ordered pairs declared and the
    AABB broad phase choosing the ones each evaluation runs; BDF1 at dt = 1e-3 with the
    FP64 Newton on the host (FP32 contact arithmetic). `steps` timed steps after 5
    untimed ones, then `steps` more each followed by the adjoint of its last contact
    evaluation; the 2,000-step loop's time is extrapolated from them."""
    import tempfile
    from atac import _redmax as R
    from atac import workloads as W
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        xml, names = W.multi_part_scene(d)
        t_gen = time.perf_counter() - t0
        t0 = time.perf_counter()
        sim = R.Simulation(xml, d)
        t_build = time.perf_counter() - t0
    n_points = int(sum(len(sim.get_body_contact_points(m)) for m in names))
    sim.reset()
    sim.forward(5)
    sim.reset_stats()
    t0 = time.perf_counter()
    sim.forward(steps)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    st = sim.stats()
    ev = max(st["evaluations"], 1)
    # the same loop with the adjoint every step: after each step, the VJP of its last
    # contact evaluation for the loss 0.5 |fm|^2 over the evaluated pairs (cotangent fm)
    t_adj = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.forward(1)
        t1 = time.perf_counter()
        sim.contact_vjp(sim.last_contact_eval()["f"])
        t_adj += time.perf_counter() - t1
    ms_adj = (time.perf_counter() - t0) * 1e3 / steps
    res = {"parts": len(names), "contact_points": n_points, "pairs_declared": len(names) * (len(names) - 1),
           "grid": "x".join(str(v) for v in sim.get_body_sdf(names[1])["values"].shape[::-1]) + " (part1)",
           "steps_timed": steps, "ms_per_step": round(ms, 3),
           "newton_iterations_per_step": round(st["newton_iterations"] / steps, 2),
           "evaluations_per_step": round(st["evaluations"] / steps, 2),
           "pairs_evaluated_per_evaluation": round(st["pairs_evaluated"] / ev, 1),
           "pairs_culled_per_evaluation": round(st["pairs_culled"] / ev, 1),
           "contact_ms_per_step": round(st["eval_ms"] / steps, 3),
           "converged_last_step": bool(sim.is_converged()),
           "loop_2000_steps_s": round(ms * 2000 / 1e3, 1),
           "ms_per_step_with_adjoint": round(ms_adj, 3),
           "adjoint_ms_per_step": round(t_adj * 1e3 / steps, 3),
           "loop_2000_steps_with_adjoint_s": round(ms_adj * 2000 / 1e3, 1),
           "scene_generation_s": round(t_gen, 1), "scene_build_s": round(t_build, 1)}
    if cpu:
        # the FP64 oracle over the pairs in contact (each free part with the plate, both
        # ways) at the final state, on all host threads, per evaluation with and without
        # derivatives; a step's CPU time counts the evaluations the GPU step made
        from oracle import oracle as O
        grids, gidx = [], {}
        insts = []
        q = {m: W.round_points_f32(np.asarray(sim.get_body_contact_points(m))) for m in names}
        for a_, b_ in [(m, names[0]) for m in names[1:]] + [(names[0], m) for m in names[1:]]:
            if b_ not in gidx:
                gd = sim.get_body_sdf(b_)
                gidx[b_] = len(grids)
                grids.append(O.GridRef(gd["values"], gd["min_box"].ravel(), gd["max_box"].ravel(), gd["dxyz"]))
            insts.append(O.make_instance(gidx[b_], q[a_], sim.get_body_E0i(a_), np.zeros(6), sim.get_body_E0i(b_),
                                         np.zeros(6), 0.01, W.KN, W.KT, W.MU, W.DAMPING))
        threads = host_cpu()["nproc"]
        t0 = time.perf_counter()
        O.eval_batch(grids, insts, O.FLAG_DERIV, threads)
        t_d = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.eval_batch(grids, insts, 0, threads)
        t_f = time.perf_counter() - t0
        cpu_step = (st["derivative_evaluations"] * t_d + (st["evaluations"] - st["derivative_evaluations"]) * t_f) \
            / steps * 1e3
        res["cpu_contact_ms_per_step"] = round(cpu_step, 1)
        res["cpu_kind"] = (f"port (FP64 oracle) over the {len(insts)} plate pairs, {threads} threads, per "
                           f"evaluation {t_d * 1e3:.0f} ms with derivatives, {t_f * 1e3:.0f} ms without")
    return res


def rotation_sweep_measurement(cpu, frame_skip=100):
    """One expansion of the disassembly planner's rotation search (plan_rot,
    examples/run_joint_plan.py:593-660): from a state, each of the 12 actions (a torque or a
    force along each axis, force_mag 1e3, torques three times as strong) is stepped
    frame_skip = 100 BDF1 steps of a peg on a free3d-exp joint in the square hole of a
    fixed block (the planner's contact parameters: kn 1e6, kt 1e3, mu 0, damping 0,
    dt 1e-3). Timed both ways: the 12 candidates in lockstep (forward_candidates, one
    batched contact evaluation per Newton round of all of them) and one at a time
    (reset, set the state and the force, forward), as the planner's loop does."""
    import tempfile
    from atac import _redmax as R
    from atac import workloads as W
    with tempfile.TemporaryDirectory() as d:
        xml = W.peg_frame_scene(d, kn=1e6, kt=1e3, mu=0.0, damping=0.0)
        sim = R.Simulation(xml, d)
    names = sim.body_names()
    K = len(W.ROTATION_ACTIONS)
    forces = np.zeros((K, len(names), 6))
    for c, a in enumerate(W.ROTATION_ACTIONS):
        forces[c, names.index("peg")] = W.action_force(a, 1e3)
    q0, qd0 = np.zeros((K, sim.ndof_r)), np.zeros((K, sim.ndof_r))
    sim.forward_candidates(q0, qd0, forces, 2)           # builds the batches
    sim.reset_stats()
    t0 = time.perf_counter()
    out = sim.forward_candidates(q0, qd0, forces, frame_skip)
    ms_lock = (time.perf_counter() - t0) * 1e3
    st_lock = sim.stats()
    sim.reset_stats()
    t0 = time.perf_counter()
    moved = []
    for c in range(K):
        sim.set_state_init(q0[c], qd0[c])
        sim.reset()
        sim.set_body_external_force("peg", forces[c, names.index("peg")])
        sim.forward(frame_skip)
        moved.append(float(np.abs(sim.get_q() - out["q"][c]).max()))
    ms_seq = (time.perf_counter() - t0) * 1e3
    st_seq = sim.stats()
    sim.set_body_external_force("peg", np.zeros(6))
    res = {"candidates": K, "steps_per_candidate": frame_skip,
           "contact_points": {m: len(sim.get_body_contact_points(m)) for m in names},
           "lockstep_ms": round(ms_lock, 1), "one_at_a_time_ms": round(ms_seq, 1),
           "speedup": round(ms_seq / ms_lock, 2),
           "lockstep_evaluations": st_lock["evaluations"], "one_at_a_time_evaluations": st_seq["evaluations"],
           "lockstep_contact_ms": round(st_lock["eval_ms"], 1), "one_at_a_time_contact_ms": round(st_seq["eval_ms"], 1),
           "max_abs_q_difference": max(moved),
           "converged": int(np.sum(out["converged"]))}
    if cpu:
        # the FP64 oracle for both pairs of one candidate at its final state, one thread,
        # with derivatives, times the evaluations the one-at-a-time loop made
        from oracle import oracle as O
        sim.set_state(out["q"][6], out["qdot"][6])
        grids, insts = [], []
        for i, (a_, b_) in enumerate([("peg", "block"), ("block", "peg")]):
            gd = sim.get_body_sdf(b_)
            grids.append(O.GridRef(gd["values"], gd["min_box"].ravel(), gd["max_box"].ravel(), gd["dxyz"]))
            insts.append(O.make_instance(i, W.round_points_f32(np.asarray(sim.get_body_contact_points(a_))),
                                         sim.get_body_E0i(a_), np.zeros(6), sim.get_body_E0i(b_), np.zeros(6), 0.01,
                                         1e6, 1e3, 0.0, 0.0))
        t0 = time.perf_counter()
        O.eval_batch(grids, insts, O.FLAG_DERIV, 1)
        t_d = time.perf_counter() - t0
        res["cpu_contact_ms_one_at_a_time"] = round(t_d * 1e3 * st_seq["evaluations"], 1)
        res["cpu_kind"] = f"port (FP64 oracle), 1 thread, {t_d * 1e3:.1f} ms per evaluation with derivatives"
    return res


def bvh_measurement():
    """BVH contact variant on a 4,096-triangle torus: exact signed distance +
    pseudonormal queries per second (100k points around the mesh, host arrays in and
    out, so the PCIe copies are included) and one force evaluation with Km / Dm and
    the p6 column for a 10,242-point ball pressed into the surface."""
    import torch
    from atac import _lib as L
    from atac.bvh import BVHMesh
    from atac import workloads as W
    V, F = W.torus(1.0, 0.35, 64, 32)
    m = BVHMesh(V, F)
    rng = np.random.default_rng(5)
    xw = rng.uniform(V.min(0) - 0.05, V.max(0) + 0.05, size=(100_000, 3))
    d_x = torch.tensor(xw, device="cuda")
    d_d = torch.empty(len(xw), dtype=torch.float64, device="cuda")
    d_n = torch.empty((len(xw), 3), dtype=torch.float64, device="cuda")
    E2 = np.eye(4)
    for _ in range(2):
        L.check(L.lib().atac_mesh_distance_normal(m.handle, L.dptr(E2), 0.0, d_x.data_ptr(), len(xw), d_d.data_ptr(),
                                                  d_n.data_ptr(), None, None))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 10
    for _ in range(reps):
        L.check(L.lib().atac_mesh_distance_normal(m.handle, L.dptr(E2), 0.0, d_x.data_ptr(), len(xw), d_d.data_ptr(),
                                                  d_n.data_ptr(), None, None))
    torch.cuda.synchronize()
    q_ms = (time.perf_counter() - t0) * 1e3 / reps
    P, _ = W.icosphere(0.3, levels=5)
    E1 = np.eye(4)
    E1[:3, 3] = V[0] * 1.05
    args = dict(col_th=0.005, kn=3e4, kt=1e3, mu=0.6, damping=5.0)
    m.force(P, E1, np.zeros(6), np.eye(4), np.zeros(6), deriv=True, p6=True, **args)
    t0 = time.perf_counter()
    for _ in range(reps):
        out = m.force(P, E1, np.zeros(6), np.eye(4), np.zeros(6), deriv=True, p6=True, **args)
    f_ms = (time.perf_counter() - t0) * 1e3 / reps
    return {"triangles": int(len(F)), "query_points": int(len(xw)), "query_ms": round(q_ms, 3),
            "queries_per_s": round(len(xw) / (q_ms * 1e-3), 1), "force_points": int(len(P)),
            "force_contacts": out["n_active"], "force_deriv_ms": round(f_ms, 3),
            "note": "device-resident queries; the force call includes its host copies"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="batched", choices=["batched", "differentiated", "peg_in_hole"])
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU oracle work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the mesh -> SDF build and BDF1 stepping measurements (profiling runs)")
    ap.add_argument("--tile-points", type=int, default=0)
    ap.add_argument("--detect-block", type=int, default=0)
    ap.add_argument("--order", type=int, default=0, help="tile queues: 0 default, 1 instance order, 2 grid-grouped, 3 per XCD by grid, "
                                                                 "4 per XCD by place in the instance")
    ap.add_argument("--points-per-lane", type=int, default=0)
    ap.add_argument("--group-cull", type=int, default=0, help="0 default (on), -1 detect every point")
    ap.add_argument("--multipart-steps", type=int, default=100, help="timed steps of the multi-part step loop")
    ap.add_argument("--sort-cell", type=float, default=0.0,
                    help="Morton sort granularity of the point sets (0: extent/1024; 1e30: generated order)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from atac.batch import batch_from_workload, device_init

    n_cu = device_init(0)
    t0 = time.perf_counter()
    w = make_workload(args.workload)
    log(f"workload {w.name}: {len(w.instances)} instances, {w.n_queries} queries, generated in "
        f"{time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    tuning = {"tile_points": args.tile_points, "detect_block": args.detect_block, "order": args.order,
              "points_per_lane": args.points_per_lane, "group_cull": args.group_cull}
    b = batch_from_workload(w, tuning=tuning, sort_cell=args.sort_cell)
    info = b.info()
    log(f"batch ready in {time.perf_counter() - t0:.1f}s: {info}")
    n_q = info["n_points"]

    g = torch.tensor(np.random.default_rng(7).normal(size=(b.n, 12)), device="cuda")
    out = b.alloc_outputs()
    grads = b.alloc_gradients()
    b.graph_capture(out, checkpoint=True, g_fm=g, grads=grads)

    for _ in range(args.warmup):
        b.graph_launch()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        b.graph_launch()
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    n_active = int(out["n_active"].sum().item())
    n_live, _ = b.live_points()
    log(f"step {ms:.3f} ms, {n_q / ms * 1e3:.3e} queries/s, active {n_active} ({n_active / n_q:.2%}), "
        f"points read after group culling {n_live} ({n_live / n_q:.2%})")

    # per-kernel durations with HIP events on the library's launch stream (eager launches)
    b.timing(True)
    for _ in range(args.steps):
        b.eval(checkpoint=True, out=out)
        b.vjp(g, out=grads)
    st = b.timing(False)
    torch.cuda.synchronize()
    stage_ms = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in st.items()}
    log(f"stage ms: {stage_ms}")

    # the same step with FP64 contact arithmetic and FP64 records (ATAC_EVAL_PRECISE), eager launches
    b.timing(True)
    for _ in range(args.steps):
        b.eval(checkpoint=True, precise=True, out=out)
        b.vjp(g, out=grads)
    st64 = b.timing(False)
    torch.cuda.synchronize()
    stage64 = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in st64.items()}
    ms64 = sum(stage64.values())
    log(f"FP64 contact arithmetic: stage ms {stage64}")

    # end to end from host memory: the instances' poses uploaded (atac_batch_set_instances
    # composes the detection maps on the host), evaluation and adjoint with d(loss)/d(wrench)
    # copied in, the wrench and the gradients copied back, every step
    inst_arr = b.instances
    g_host = g.cpu()
    e2e_steps = max(1, args.steps // 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(e2e_steps):
        b.set_instances(inst_arr)
        b.eval(checkpoint=True, out=out)
        g.copy_(g_host)
        b.vjp(g, out=grads)
        host_out = [out["fm"].cpu()] + [v.cpu() for v in grads.values()]
    torch.cuda.synchronize()
    e2e_ms = (time.perf_counter() - t0) * 1e3 / e2e_steps
    log(f"end to end from host memory: {e2e_ms:.3f} ms per step")

    touched = b.touched_values()
    # detection (the culling pass and the detection kernel): every group's box read
    # (32 B) and its summed-area counts (8 x 4 B) when culling, every point of the
    # kept groups read once (12 B), every grid value its stencils touch once (4 B),
    # the instance fields and tile entries read, and the active list (one offset per
    # active point) and its work items written
    n_groups = sum(-(-len(w.point_sets[ins.points]) // GROUP) for ins in w.instances)
    culled = n_live < n_q or args.group_cull >= 0
    n_chunks_lb = -(-n_active // K_CHUNK)
    det_bytes = (12 * n_live + (GROUP_BYTES * n_groups + 4 * info["n_tiles"] if culled else 0) + 4 * touched
                 + DETECT_INST_BYTES * b.n + TILE_BYTES * info["n_tiles"] + ACT_BYTES * n_active
                 + CHUNK_BYTES * n_chunks_lb)
    det_ms = stage_ms["detect"]
    achieved = det_bytes / (det_ms * 1e-3) / 1e9
    # the step: detection; the contact pass reads the active list and points and
    # writes one record per active point; the adjoint pass reads the records; both
    # write one partial row per work item; the reductions write fm, read g_fm and
    # write the 26 gradients per instance
    step_bytes = (det_bytes + (ACT_BYTES + 12 + RECORD_BYTES) * n_active + RECORD_BYTES * n_active
                  + 8 * (12 + 26) * n_chunks_lb + 8 * 12 * b.n * 2 + 8 * 26 * b.n)
    # a plain device-to-device copy of 1 GiB (read + write) on the same GPU: the bandwidth a
    # streaming kernel reaches here, reported beside the published peak
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(10):
        dst.copy_(src)
    c1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * src.numel() * 4 / (c0.elapsed_time(c1) / 10 * 1e-3) / 1e9
    del src, dst
    log(f"device copy: {copy_gbs:.0f} GB/s")
    # the whole step's algorithmic bytes: every point read once, every touched grid
    # value once, per instance its poses, twists and parameters (408 B) and g_fm (96 B)
    # read and its wrench (96 B) and gradients (208 B) written, and the adjoint's
    # checkpoint of d and grad phi (16 B) per active contact
    step_alg = (12 * n_live + (GROUP_BYTES * n_groups if culled else 0) + 4 * touched + (408 + 96 + 96 + 208) * b.n
                + 16 * n_active)
    step_gbs = step_alg / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                "kernel": "cull_kernel + detect_kernel" if culled else "detect_kernel",
                "algorithmic_bytes_per_launch": int(det_bytes),
                "points_read": n_live, "groups": n_groups,
                # the same launch priced as if every query read its point (12 B per query, the
                # per-query rate without culling): what a non-culling detection would have to stream
                "all_queries_point_bytes_GBps": round(12 * n_q / (det_ms * 1e-3) / 1e9, 1),
                "touched_grid_values": touched, "launch_ms": round(det_ms, 4),
                "measured_copy_GBps": round(copy_gbs, 1), "frac_of_measured_copy": round(achieved / copy_gbs, 4),
                "step_algorithmic_bytes": int(step_alg), "step_achieved": round(step_gbs, 1),
                "step_frac": round(step_gbs / HBM_PEAK_GBS, 4)}

    # the same outputs by the fused pass (atac_batch_eval_vjp): for a cotangent known when the
    # contacts are evaluated (a backward pass re-evaluating a step's contacts), one pass computes
    # the wrench and the adjoint without writing and re-reading the records; graph replay
    b.eval(checkpoint=True, out=out)
    b.vjp(g, out=grads)
    fout, fgr = b.alloc_outputs(), b.alloc_gradients()
    b.graph_capture(fout, g_fm=g, grads=fgr, fused=True)
    for _ in range(args.warmup):
        b.graph_launch()
    torch.cuda.synchronize()
    start.record()
    for _ in range(args.steps):
        b.graph_launch()
    stop.record()
    torch.cuda.synchronize()
    ms_f = start.elapsed_time(stop) / args.steps
    fdiff = 0.0
    for ref_t, got_t in [(out["fm"], fout["fm"])] + [(grads[k], fgr[k]) for k in grads]:
        r_, g_ = ref_t.cpu().numpy(), got_t.cpu().numpy()
        fdiff = max(fdiff, float(np.abs(g_ - r_).max() / max(np.abs(r_).max(), 1e-300)))
    b.timing(True)
    for _ in range(args.steps):
        b.eval_vjp(g, out=fout, grads=fgr)
    stf = b.timing(False)
    torch.cuda.synchronize()
    fused_alg = step_alg - 16 * n_active   # no (d, grad phi) checkpoint to write and read back
    fused = {"ms_per_step": round(ms_f, 4), "queries_per_s": n_q / ms_f * 1e3,
             "stage_ms": {k: round(v[0] / v[1], 4) for k, v in stf.items() if v[1]},
             "max_rel_diff_vs_separate_passes": fdiff,
             "step_algorithmic_bytes": int(fused_alg),
             "step_frac": round(fused_alg / (ms_f * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
             "what": "wrench + adjoint of a cotangent given up front, one fused pass over the contacts "
                     "(no records kept), HIP graph replay; the headline keeps the separate passes a "
                     "torch.autograd backward needs"}
    log(f"fused pass: {fused}")
    b.graph_capture(out, checkpoint=True, g_fm=g, grads=grads)   # the headline graph again

    plain = None
    if not args.no_plain:
        po, pg = b.eval_plain_vjp(g)
        torch.cuda.synchronize()
        reps = 3
        s2 = torch.cuda.Event(enable_timing=True)
        e2 = torch.cuda.Event(enable_timing=True)
        s2.record()
        for _ in range(reps):
            b.eval_plain_vjp(g, out=po, grads=pg)
        e2.record()
        torch.cuda.synchronize()
        pms = s2.elapsed_time(e2) / reps
        plain = {"ms_per_step": round(pms, 3), "queries_per_s": n_q / pms * 1e3,
                 "what": "the step's outputs (wrench and adjoint) by the plain path: one launch per force object, "
                         "one lane per point, the contact math and its adjoint inline, FP64 atomics",
                 "speedup_batched_vs_plain": round(pms / ms, 1)}
        log(f"plain path {pms:.2f} ms/step ({pms / ms:.1f}x slower)")

    cpu = None
    err = None
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        b.graph_launch()                      # the graph's outputs: the step's wrench and adjoint
        torch.cuda.synchronize()
        gpu = {k: v.cpu().numpy() for k, v in list(out.items()) + list(grads.items())}
        cpu, err = cpu_baseline(w, gpu, g.cpu().numpy(), args.cpu_budget, threads, args.cpu_budget / 3)
        log(f"cpu baseline {cpu['value']:.3e} queries/s on {threads} threads, {cpu['one_thread']['value']:.3e} "
            f"on 1; err {err}")

    sdf_build = bdf1 = bvh = multipart = sweep = None
    if not args.no_extras:
        multipart = multipart_measurement(not args.no_cpu, args.multipart_steps)
        log(f"multi-part step loop: {multipart}")
        sweep = rotation_sweep_measurement(not args.no_cpu)
        log(f"rotation sweep: {sweep}")
        sdf_build = sdf_build_measurement(not args.no_cpu)
        log(f"sdf build: {sdf_build}")
        bdf1 = bdf1_measurement(not args.no_cpu)
        log(f"bdf1: {bdf1}")
        bvh = bvh_measurement()
        log(f"bvh contact: {bvh}")

    line = {
        "metric": METRIC, "value": n_q / ms * 1e3, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "instances": b.n, "queries": n_q, "grids": f"{len(w.grids)} x "
                   f"{'x'.join(str(v) for v in w.grids[0].n)} FP32", "active_fraction": round(n_active / n_q, 5),
                   "step": "group culling + detect (cell classification, active list) + slow-path records + "
                           "contact wrench (records built) + reduce + adjoint (VJP) + reduce, HIP graph replay",
                   "group_cull": info["group_cull"], "points_read_fraction": round(n_live / n_q, 5),
                   "tile_points": info["tile_points"], "detect_block": info["detect_block"],
                   "points_per_lane": info["points_per_lane"],
                   "persistent_blocks": info["persistent_blocks"], "n_cu": n_cu, "parallelism": "none"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "err_vs_fp64": err,
        "plain_path": plain,
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "fused_force_adjoint": fused,
        "fp64_contact": {"kernel_ms_per_step": round(ms64, 4), "queries_per_s": n_q / ms64 * 1e3,
                         "stage_ms": {k: round(v, 4) for k, v in stage64.items()},
                         "what": "detection unchanged; records, wrench and adjoint in FP64 (sum of the stage times)"},
        "end_to_end_host": {"ms_per_step": round(e2e_ms, 4), "queries_per_s": n_q / e2e_ms * 1e3,
                            "what": "instances uploaded from host memory (detection maps composed on the host), "
                                    "evaluation + adjoint, wrench and gradients copied back, eager launches"},
        "step_effective_GBps": round(step_bytes / (ms * 1e-3) / 1e9, 1),
        "sdf_build": sdf_build,
        "bdf1_translational": bdf1,
        "bvh_contact": bvh,
        "multi_part_loop": multipart,
        "rotation_sweep": sweep,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
