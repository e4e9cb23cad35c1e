"""sdfo_contact_evaluate_noalloc (static device scratch) against
sdfo_contact_evaluate through the C-ABI: two calls in a row, the second with
fewer pairs and fewer points, both with the broad phase on; the kernels are
the same, so results are equal bit for bit.  Sizes past the static maximum
are rejected."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _states(n_pairs, seed):
    from sdfo import scenes
    Rs, ts = scenes.ring_poses(n_pairs, seed=seed)
    rng = np.random.default_rng(seed)
    st = np.zeros((n_pairs, 36))
    for b in range(n_pairs):
        RB = scenes.rot_axis(rng.normal(size=3), rng.uniform(0.2, np.pi))
        rB = rng.normal(0, 0.3, 3)
        RA, rA = RB @ Rs[b], rB + RB @ ts[b]
        vA, wA, vB, wB = rng.normal(0, 0.2, (4, 3))
        st[b] = np.concatenate([rA, RA.ravel(), vA, wA, rB, RB.ravel(), vB, wB])
    return st


def test_noalloc_evaluate_equals_allocating_evaluate(sdfo_gpu):
    from sdfo import _lib, scenes
    from sdfo import contact as SC
    lib = _lib.lib
    v, t = scenes.cube_mesh()
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=64, max_resolution=64)
    p = _lib.ContactParams(1e11, 1e4, 0.3, 1e-6, 0.07, 2, 3)
    stream = torch.cuda.current_stream().cuda_stream
    ptr = lambda a: C.c_void_p(a.data_ptr())  # noqa: E731
    for call, (nmaj, nmin, n_pairs, seed) in enumerate([(120, 32, 300, 1), (60, 16, 37, 2)]):
        print(f"[gpu] noalloc evaluate call {call}: {n_pairs} pairs", flush=True)
        X, w = scenes.ring_quadrature(nmaj, nmin)
        bvh = SC.PointBvh(X, leaf_size=16)
        Xt, wt = torch.from_numpy(X).cuda(), torch.from_numpy(w).cuda()
        St = torch.from_numpy(_states(n_pairs, seed)).cuda()
        outs = []
        for fn in (lib.sdfo_contact_evaluate, lib.sdfo_contact_evaluate_noalloc):
            out = torch.full((n_pairs, 12), np.nan, dtype=torch.float64, device="cuda")
            act = torch.full((n_pairs,), -1, dtype=torch.int32, device="cuda")
            rc = fn(g.handle, bvh.handle, ptr(Xt), ptr(wt), len(X), ptr(St), None, n_pairs,
                    C.byref(p), None, ptr(out), ptr(act), stream)
            assert rc == 0, lib.sdfo_last_error()
            outs.append((out.cpu().numpy(), act.cpu().numpy()))
        (o1, a1), (o2, a2) = outs
        assert (a1 > 0).sum() > 10
        assert np.array_equal(a1, a2)
        assert np.array_equal(o1, o2)
        bvh.close()
    # past the static maximum: rejected before any launch
    big = torch.zeros((4097, 36), dtype=torch.float64, device="cuda")
    out = torch.zeros((4097, 12), dtype=torch.float64, device="cuda")
    rc = lib.sdfo_contact_evaluate_noalloc(g.handle, None, ptr(Xt), ptr(wt), len(Xt), ptr(big),
                                           None, 4097, C.byref(p), None, ptr(out), None, stream)
    assert rc == -1 and b"4096" in lib.sdfo_last_error()
    torch.cuda.synchronize()
    g.close()
