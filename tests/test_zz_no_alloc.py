"""atac_bvh_force_noalloc (the BVH force evaluation over a fixed per-device
scratch array, for callers that cannot allocate) against atac_bvh_force_host,
through the C ABI: the same kernels in the same order, so the outputs agree bit
for bit. Two calls in a row, the second over fewer points (fewer workgroups:
rows of the first call stay in the scratch beyond what the second reads)."""
import ctypes as C

import numpy as np
import pytest

from atac import _lib as L
from atac.batch import make_instance
from atac.bvh import BVHMesh
from atac.workloads import icosphere, torus
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(seed, n_shells):
    V, F = torus(1.0, 0.35, 32, 16)
    rng = np.random.default_rng(seed)
    P, _ = icosphere(0.35, levels=4)
    xi = np.concatenate([P * (1.0 - 0.05 * k) for k in range(n_shells)])
    E1 = O.se3_exp(np.concatenate([rng.normal(size=3) * 0.3, np.zeros(3)]))
    E1[:3, 3] = V[rng.integers(len(V))] * 1.02
    inst = make_instance(0, 0, 0, E1, rng.normal(size=6) * 0.2, np.eye(4), rng.normal(size=6) * 0.1,
                         0.005, 3e4, 1e3, 0.6, 5.0, 1.1, 0.9)
    return V, F, np.ascontiguousarray(xi), inst


def _noalloc(m, inst, xi, flags):
    import torch
    x = torch.from_numpy(xi).cuda()
    out = torch.full((312,), float("nan"), dtype=torch.float64, device="cuda")
    na = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    L.check(L.lib().atac_bvh_force_noalloc(m.handle, C.byref(inst), C.c_void_p(x.data_ptr()), len(xi), flags,
                                           C.c_void_p(out.data_ptr()), C.c_void_p(na.data_ptr()), C.c_void_p(st)))
    torch.cuda.synchronize()
    return out.cpu().numpy(), int(na.cpu().item())


def _host(m, inst, xi, flags):
    fm, Km, Dm, p6 = np.zeros(12), np.zeros(144), np.zeros(144), np.zeros(12)
    na = C.c_int64(0)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    L.check(L.lib().atac_bvh_force_host(m.handle, C.byref(inst), vp(xi), len(xi), flags, vp(fm), vp(Km), vp(Dm),
                                        vp(p6), C.byref(na)))
    return np.concatenate([fm, p6, Km, Dm]), int(na.value)


def test_noalloc_matches_host_variant(gpu):
    flags = L.EVAL_DERIV | L.EVAL_P6
    for seed, shells in ((3, 4), (4, 1)):   # 10,248 then 2,562 points: 6 then 2 workgroups
        V, F, xi, inst = _case(seed, shells)
        m = BVHMesh(V, F)
        print("no-alloc BVH force:", len(xi), "points, seed", seed, flush=True)
        got, na = _noalloc(m, inst, xi, flags)
        ref, nref = _host(m, inst, xi, flags)
        assert na == nref and na > 20, (na, nref)
        assert np.array_equal(got, ref), np.abs(got - ref).max()


def test_noalloc_wrench_only_and_limit(gpu):
    """Without DERIV/P6 the p6, Km and Dm blocks come out zero; n beyond the
    scratch is refused before any launch."""
    V, F, xi, inst = _case(5, 2)
    m = BVHMesh(V, F)
    got, na = _noalloc(m, inst, xi, 0)
    ref, nref = _host(m, inst, xi, 0)
    assert na == nref > 0
    assert np.array_equal(got[:12], ref[:12])
    assert np.all(got[12:] == 0.0)
    rc = L.lib().atac_bvh_force_noalloc(m.handle, C.byref(inst), None, L.BVH_NOALLOC_MAX_POINTS + 1, 0, None, None,
                                        None)
    assert rc == L.ATAC_ERR_INVALID
    assert b"fixed scratch" in L.lib().atac_last_error()
