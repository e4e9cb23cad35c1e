"""Threading and stream contract: a grid handle is immutable after its build,
so queries and contact calls on different streams at once give the same
bits as the same calls made one after another; a stream-ordered release
returns the arrays after the work already on that stream."""
import numpy as np
import pytest
import torch

from sdfo import contact as SC
from sdfo import scenes

pytestmark = pytest.mark.gpu


def _work(g, pts32, pts64, X, w, poses, states, stream):
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    law_e = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, narrow_margin=0.07,
                          order=2)
    with torch.cuda.stream(stream):
        out = []
        for _ in range(3):
            out += list(g.query(pts32, "Tricubic", stream=stream))
            out += list(g.query(pts64, "ContactAware", stream=stream))
            out.append(SC.contact_forward(g, X, w, poses, law, stream=stream))
            out.append(SC.contact_evaluate(g, X, w, states, law_e, stream=stream))
    return out


def test_two_streams_on_one_handle(sdfo_gpu):
    v, t = scenes.cube_mesh()
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=128, max_resolution=128)
    try:
        cuda = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        X, w = scenes.ring_quadrature(120, 32)
        inputs = []
        for seed in (1, 2):
            R, tt = scenes.ring_poses(200, seed=seed)
            v0, Om = scenes.pose_velocities(200, seed=seed + 10)
            st = np.concatenate([tt, R.reshape(-1, 9), v0, Om, np.zeros((200, 3)),
                                 np.tile(np.eye(3).ravel(), (200, 1)), np.zeros((200, 6))], 1)
            inputs.append((cuda(scenes.valid_region_points(g.spec, 400_000, seed=seed)),
                           cuda(scenes.valid_region_points(g.spec, 100_000, seed=seed + 5,
                                                           dtype=np.float64)),
                           cuda(X), cuda(w), cuda(scenes.pose_array(R, tt, v0=v0, Om=Om)),
                           cuda(st)))
        main = torch.cuda.current_stream()
        serial = [_work(g, *inp, main) for inp in inputs]
        torch.cuda.synchronize()
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        s1.wait_stream(main)
        s2.wait_stream(main)
        conc = [None, None]
        conc[0] = _work(g, *inputs[0], s1)
        conc[1] = _work(g, *inputs[1], s2)
        torch.cuda.synchronize()
        for a_list, b_list in zip(serial, conc):
            assert len(a_list) == len(b_list)
            for a, b in zip(a_list, b_list):
                assert torch.equal(a, b)
    finally:
        g.close()


def test_stream_ordered_release(sdfo_gpu):
    """A grid released on the stream that used it: the release waits for that
    stream's work (the queries see valid data), with no device-wide sync;
    grids of one batched build released on two streams free their shared
    block after both."""
    v, t = scenes.cube_mesh()
    s = torch.cuda.Stream()
    ref_g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=128, max_resolution=128)
    pts = torch.from_numpy(scenes.valid_region_points(ref_g.spec, 1_000_000, seed=4)).cuda()
    want = ref_g.query(pts, "Tricubic")[0]
    torch.cuda.synchronize()
    for _ in range(3):
        g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=128, max_resolution=128)
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            got = g.query(pts, "Tricubic", stream=s)[0]
            g.close(stream=s)
            got2 = ref_g.query(pts, "Tricubic", stream=s)[0]
        torch.cuda.synchronize()
        assert torch.equal(got, want) and torch.equal(got2, want)
    grids = sdfo_gpu.SdfGrid.build_dense_batched([(v, t), (v, t)],
                                                 [dict(target_resolution=64, max_resolution=64)] * 2)
    s2 = torch.cuda.Stream()
    for g, st in zip(grids, (s, s2)):
        st.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(st):
            p = torch.from_numpy(scenes.valid_region_points(g.spec, 200_000, seed=5)).cuda()
            out = g.query(p, "Tricubic", stream=st)[0]
            g.close(stream=st)
        torch.cuda.synchronize()
        assert torch.isfinite(out).all()
    ref_g.close()


def test_fp64_cell_records_made_on_one_stream_used_on_another(sdfo_gpu):
    """A dense grid's FP64 cell records are made by its first FP64 Tricubic
    query, on that query's stream; a query issued at once on a second stream
    waits for them.  Both give the bits of the same grid built without cell
    records (FP64 corners gathered from phi and n)."""
    from sdfo import _lib

    v, t = scenes.cube_mesh()
    cuda = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    _lib.debug_skip_cell_records(True)
    try:
        plain = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=96, max_resolution=96)
    finally:
        _lib.debug_skip_cell_records(False)
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=96, max_resolution=96)
    try:
        pts = [cuda(scenes.valid_region_points(g.spec, 300_000, seed=s, dtype=np.float64))
               for s in (4, 5)]
        want = [plain.query(p, mode) for p in pts for mode in ("Tricubic", "Trilinear")]
        torch.cuda.synchronize()
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        s1.wait_stream(torch.cuda.current_stream())
        s2.wait_stream(torch.cuda.current_stream())
        print("[gpu] FP64 records: first query on s1, second on s2 at once", flush=True)
        with torch.cuda.stream(s1):
            a = [g.query(pts[0], mode, stream=s1) for mode in ("Tricubic", "Trilinear")]
        with torch.cuda.stream(s2):
            b = [g.query(pts[1], mode, stream=s2) for mode in ("Tricubic", "Trilinear")]
        torch.cuda.synchronize()
        for got, exp in zip(a + b, want):
            for x, y in zip(got, exp):
                assert torch.equal(x, y)
    finally:
        g.close()
        plain.close()
