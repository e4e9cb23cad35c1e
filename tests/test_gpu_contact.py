"""GPU contact reduction (contacts_sdf.py:329-385) against the FP64 oracle.

Both evaluate the law in FP64 on the same FP32-rounded field data; they
differ only by FMA contraction and summation order, so forward sums agree to
1e-9 relative and the analytic gradients to 1e-7 relative.
"""
import numpy as np
import pytest
import torch

from conftest import f32, ograd
from sdfo import contact as SC
from sdfo import meshes, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cube(O, sdfo_gpu):
    v, t = scenes.cube_mesh()
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=64, max_resolution=64)
    ref = g.download_dense()
    yield g, ref
    g.close()


def _scene(n_poses, seed, nmaj=60, nmin=16):
    X, w = scenes.ring_quadrature(nmaj, nmin)
    R, t = scenes.ring_poses(n_poses, seed=seed)
    v0, Om = scenes.pose_velocities(n_poses, seed=seed + 1)
    return X, w, scenes.pose_array(R, t, v0=v0, Om=Om)


def _cuda(*arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


@pytest.mark.parametrize("mode,mu,order", [("Tricubic", 0.0, 2), ("Tricubic", 0.3, 2),
                                           ("Trilinear", 0.3, 1), ("Tricubic", 0.5, 1)])
def test_contact_forward_matches_oracle(O, cube, mode, mu, order):
    g, ref = cube
    print(f"[gpu] contact forward {mode} mu={mu} order={order}", flush=True)
    X, w, poses = _scene(64, seed=3)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=mu, regularizer=1e-6, order=order, mode=mode)
    out, act = SC.contact_forward(g, *_cuda(X, w, poses), law, active=True)
    o_ref, a_ref = O.contact_forward(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X, w, poses,
                                     kn=1e11, c=1e4, mu=mu, reg=1e-6, order=order, mode=mode)
    assert np.array_equal(act.cpu().numpy(), a_ref)
    assert (a_ref > 0).sum() > 8
    out = out.cpu().numpy()
    scale = np.abs(o_ref).max(axis=0) + 1e-300
    assert np.all(np.abs(out - o_ref) <= 1e-9 * scale)


def test_contact_forward_deterministic(cube):
    g, _ = cube
    X, w, poses = _scene(32, seed=4)
    law = SC.ContactLaw(mu=0.3)
    a = SC.contact_forward(g, *_cuda(X, w, poses), law)
    b = SC.contact_forward(g, *_cuda(X, w, poses), law)
    assert torch.equal(a, b)


@pytest.mark.parametrize("mode,mu,order", [("Tricubic", 0.3, 2), ("Trilinear", 0.3, 1)])
def test_contact_backward_matches_oracle(O, cube, mode, mu, order):
    g, ref = cube
    print(f"[gpu] contact backward {mode}", flush=True)
    X, w, poses = _scene(24, seed=7)
    rng = np.random.default_rng(0)
    gout = rng.normal(size=(len(poses), 6))
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=mu, regularizer=1e-6, order=order, mode=mode)
    gt, gp, gx, gw = SC.contact_backward(g, *_cuda(X, w, poses), torch.from_numpy(gout).cuda(), law)
    rt, rp, rx, rw = O.contact_backward(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X, w,
                                        poses, gout, kn=1e11, c=1e4, mu=mu, reg=1e-6, order=order,
                                        mode=mode)
    for got, exp in ((gt, rt), (gp, rp), (gx, rx), (gw, rw)):
        got = got.cpu().numpy()
        scale = np.abs(exp).max() + 1e-300
        assert np.abs(got - exp).max() <= 1e-7 * scale


def test_contact_autograd_function(O, cube):
    g, ref = cube
    X, w, poses = _scene(6, seed=9)
    R0 = torch.from_numpy(poses[:, :9].reshape(-1, 3, 3).copy()).cuda()
    t0 = torch.from_numpy(poses[:, 9:12].copy()).cuda()
    v0 = torch.from_numpy(poses[:, 24:27].copy()).cuda()
    Om = torch.from_numpy(poses[:, 27:30].copy()).cuda()
    Xt, wt = _cuda(X, w)
    Xt.requires_grad_(True)
    twist = torch.zeros((6, 6), dtype=torch.float64, device="cuda", requires_grad=True)
    kn = torch.tensor(1e11, dtype=torch.float64, requires_grad=True)
    out = SC.contact(g, twist, kn, 1e4, 0.3, Xt, wt, R0, t0, v0, Om)
    loss = (out[:, 2] + 10.0 * out[:, 5]).sum()
    loss.backward()
    assert twist.grad is not None and torch.isfinite(twist.grad).all()
    assert kn.grad is not None and torch.isfinite(kn.grad)
    assert Xt.grad is not None and Xt.grad.shape == Xt.shape
    # kn enters linearly in the penalty and friction terms (damping is kn free)
    law0 = SC.ContactLaw(kn=1e11, damping=0.0, mu=0.3)
    base = SC.contact_forward(g, Xt.detach(), wt, SC.pose_records(R0, t0, v0=v0, Om=Om), law0)
    assert torch.allclose(kn.grad * 1e11, (base[:, 2] + 10.0 * base[:, 5]).sum(), rtol=1e-9)


# ------------------------------------------------------ linearisation cache
@pytest.mark.parametrize("margin", [0.0, 0.065])
def test_linearize_fill_matches_oracle(O, cube, margin):
    g, ref = cube
    print(f"[gpu] linearize fill margin={margin}", flush=True)
    X, w, poses = _scene(4, seed=21)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, narrow_margin=margin)
    Xd, wd = _cuda(X, w)
    for b in range(len(poses)):
        cache = SC.LinearizationCache(g, Xd, wd, law).fill(_cuda(poses[b])[0])
        g0, grad0, valid0 = O.contact_linearize(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]),
                                                X, poses[b], kn=1e11, c=1e4, mu=0.3, reg=1e-6,
                                                margin=margin)
        assert np.array_equal(cache.valid0.cpu().numpy(), valid0)
        assert valid0.sum() > 8
        assert np.abs(cache.g0.cpu().numpy() - g0).max() <= 1e-12
        assert np.abs(cache.grad0.cpu().numpy() - grad0).max() <= 1e-10


def test_linearized_forward_at_reference_pose_is_bitwise_forward(cube):
    g, _ = cube
    X, w, poses = _scene(8, seed=22)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    Xd, wd, Pd = _cuda(X, w, poses)
    direct, da = SC.contact_forward(g, Xd, wd, Pd, law, active=True)
    for b in range(len(poses)):
        cache = SC.LinearizationCache(g, Xd, wd, law).fill(Pd[b])
        lin, la = cache.forward(Pd[b:b + 1].contiguous(), active=True)
        assert torch.equal(la[0], da[b])
        assert torch.equal(lin[0], direct[b])


@pytest.mark.parametrize("mu,order,margin", [(0.3, 2, 0.0), (0.0, 1, 0.0), (0.5, 2, 0.065)])
def test_linearized_forward_matches_oracle(O, cube, mu, order, margin):
    """Perturbed poses (translation and rotation) around one reference pose,
    the cache filled on the GPU and handed to the oracle's linearised forward."""
    g, _ = cube
    print(f"[gpu] linearized forward mu={mu} order={order} margin={margin}", flush=True)
    X, w, base = _scene(1, seed=23)
    rng = np.random.default_rng(5)
    from sdfo.scenes import rot_axis
    poses = np.repeat(base, 64, axis=0)
    for b in range(1, 64):
        Rr = rot_axis(rng.normal(size=3), rng.uniform(0, 0.01))
        poses[b, :9] = (Rr @ base[0, :9].reshape(3, 3)).ravel()
        poses[b, 12:21] = poses[b, :9]
        poses[b, 9:12] += rng.uniform(-2e-3, 2e-3, 3)
        poses[b, 11] -= 4e-3                          # deeper than the reference pose
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=mu, regularizer=1e-6, narrow_margin=margin,
                        order=order)
    Xd, wd, Pd = _cuda(X, w, poses)
    cache = SC.LinearizationCache(g, Xd, wd, law)
    out, act = cache.forward(Pd, active=True)        # fills at the first pose
    assert cache.filled
    o_ref, a_ref = O.contact_forward_linearized(
        cache.g0.cpu().numpy(), cache.grad0.cpu().numpy(), cache.valid0.cpu().numpy(),
        poses[0, 9:12], X, w, poses, kn=1e11, c=1e4, mu=mu, reg=1e-6, margin=margin, order=order)
    assert np.array_equal(act.cpu().numpy(), a_ref)
    assert (a_ref > 0).sum() > 40 and a_ref.max() > 100
    out = out.cpu().numpy()
    scale = np.abs(o_ref).max(axis=0) + 1e-300
    assert np.all(np.abs(out - o_ref) <= 1e-9 * scale)
    cache.clear()
    assert not cache.filled
    again = cache.forward(Pd[5:].contiguous())      # refills at pose 5
    assert torch.allclose(again[0], SC.contact_forward(g, Xd, wd, Pd[5:6].contiguous(), law)[0],
                          rtol=0, atol=0)


# --------------------------------------------------------------- broad phase
def _cull_scene(g, n_poses=48, seed=31):
    X, w, poses = _scene(n_poses, seed=seed)
    rng = np.random.default_rng(seed)
    poses[:, 9:11] += rng.uniform(-0.09, 0.09, (n_poses, 2))   # ring partly off the SDF box
    # no yaw on even poses: their corner-mapped boxes stay upright
    poses[::2, :9] = np.eye(3).ravel()
    poses[::2, 12:21] = np.eye(3).ravel()
    poses[-1, 9] = 1.0                                          # no leaf overlaps
    return X, w, poses


@pytest.mark.parametrize("margin", [0.0, 0.01])
def test_cull_mask_matches_oracle(O, cube, margin):
    g, _ = cube
    print(f"[gpu] broad-phase cull margin={margin}", flush=True)
    X, _, poses = _cull_scene(g)
    bvh = SC.PointBvh(X, leaf_size=16)
    nodes = O.point_bvh_build(X, leaf_size=16)
    Pd = _cuda(poses)[0]
    boxes = SC.broad_phase_boxes(Pd, g.spec.bmin, g.spec.bmax, 0.02)
    ob = O.broad_phase_boxes(poses, g.spec.bmin, g.spec.bmax, 0.02)
    assert np.abs(boxes.cpu().numpy() - ob).max() <= 1e-15
    mask = bvh.cull(boxes, margin).cpu().numpy()
    B = boxes.cpu().numpy()
    sizes = []
    for b in range(len(poses)):
        want = np.zeros(len(X), np.uint8)
        want[O.point_bvh_query(nodes, B[b, :3], B[b, 3:], margin)] = 1
        assert np.array_equal(mask[b], want), b
        sizes.append(int(want.sum()))
    assert min(sizes) == 0 and max(sizes) == len(X) and len(set(sizes)) > 10


def test_culled_forward_matches_oracle(O, cube):
    g, ref = cube
    print("[gpu] culled contact forward", flush=True)
    X, w, poses = _cull_scene(g)
    nodes = O.point_bvh_build(X, leaf_size=16)
    bvh = SC.PointBvh(X)
    Xd, wd, Pd = _cuda(X, w, poses)
    boxes = SC.broad_phase_boxes(Pd, g.spec.bmin, g.spec.bmax, 0.0)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    out, act = SC.contact_forward_culled(g, bvh, boxes, Xd, wd, Pd, law, active=True)
    full, fact = SC.contact_forward(g, Xd, wd, Pd, law, active=True)
    B = boxes.cpu().numpy()
    out, act = out.cpu().numpy(), act.cpu().numpy()
    for b in range(len(poses)):
        idx = O.culled_points(nodes, B[b], len(X))
        o, a = O.contact_forward(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X[idx], w[idx],
                                 poses[b:b + 1], kn=1e11, c=1e4, mu=0.3, reg=1e-6, order=2)
        assert act[b] == a[0], b
        scale = np.abs(o[0]).max() + 1e-300
        assert np.abs(out[b] - o[0]).max() <= 1e-9 * scale, b
    # contacts lie inside the SDF box, so upright boxes cull none of them
    assert np.array_equal(act[::2], fact.cpu().numpy()[::2]) and act.max() > 100
    assert torch.equal(torch.from_numpy(out[-1]), full[-1].cpu())   # empty cull -> all points

    # the linearised forward behind the same cull
    cache = SC.LinearizationCache(g, Xd, wd, law).fill(Pd[0])
    lo, la = cache.forward(Pd, active=True, bvh=bvh, boxes=boxes)
    g0, gr, v0 = (t.cpu().numpy() for t in (cache.g0, cache.grad0, cache.valid0))
    for b in range(len(poses)):
        idx = O.culled_points(nodes, B[b], len(X))
        o, a = O.contact_forward_linearized(g0[idx], gr[idx], v0[idx], poses[0, 9:12], X[idx],
                                            w[idx], poses[b:b + 1], kn=1e11, c=1e4, mu=0.3,
                                            reg=1e-6, order=2)
        assert la[b].item() == a[0], b
        scale = np.abs(o[0]).max() + 1e-300
        assert np.abs(lo[b].cpu().numpy() - o[0]).max() <= 1e-9 * scale, b


@pytest.mark.parametrize("n_radial", [20, 40])
def test_torsion_on_sdf_face_matches_closed_form(sdfo_gpu, n_radial):
    """The annular-patch torsion check (ring_on_cube_torsion.py:55-60) through the
    GPU path: an annulus (Ri 0.1, Ro 0.2, midpoint rings x 128 sectors) pressed
    0.01 into the top face of a cube SDF and spinning about the face normal,
    mu = 0.3, linear penalty.  Normal force = kn * depth * area and torque =
    -mu N (2/3)(Ro^3 - Ri^3)/(Ro^2 - Ri^2) times the quadrature's exact relative
    error, both to 1e-6; the symmetric patch's tangential residual to 1e-15 mu N."""
    v, t = meshes.cube(1.0)
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=64, max_resolution=64)
    try:
        Ri, Ro, mu, kn, pen = 0.1, 0.2, 0.3, 1e6, 0.01
        X, w = meshes.annular_quadrature(Ri, Ro, n_radial, 128)
        P = scenes.pose_array(np.eye(3)[None], np.array([[0.0, 0.0, 0.5 - pen]]),
                              Om=np.array([[0.0, 0.0, 1.0]]))
        law = SC.ContactLaw(kn=kn, damping=0.0, mu=mu, regularizer=0.0, order=1, mode="Tricubic")
        out, act = SC.contact_forward(g, *_cuda(X, w, P), law, active=True)
        out = out.cpu().numpy()[0]
        N = kn * pen * np.pi * (Ro ** 2 - Ri ** 2)
        Tz = -mu * N * (2.0 / 3.0) * (Ro ** 3 - Ri ** 3) / (Ro ** 2 - Ri ** 2)
        quad = -(Ro - Ri) ** 3 / (4 * n_radial ** 2 * (Ro ** 3 - Ri ** 3))
        assert int(act.cpu()[0]) == len(X)
        assert np.isclose(out[2], N, rtol=1e-6)
        assert np.isclose(out[5], Tz * (1 + quad), rtol=1e-6)
        resid = np.hypot(out[0], out[1]) / (mu * N)
        print(f"[gpu] torsion tangential residual / (mu N) = {resid:.3e}", flush=True)
        assert resid <= 1e-15
    finally:
        g.close()


def test_contact_gradients_match_central_differences(cube):
    """The hand-written backward through `sdfo.contact.contact` against FP64
    central differences of the same GPU forward (h = 1e-7 m / rad), friction
    on: pose twist, kn and a few quadrature points, to 1e-4 relative."""
    g, _ = cube
    X, w, poses = _scene(4, seed=21)
    R0 = torch.from_numpy(poses[:, :9].reshape(-1, 3, 3).copy()).cuda()
    t0 = torch.from_numpy(poses[:, 9:12].copy()).cuda()
    v0 = torch.from_numpy(poses[:, 24:27].copy()).cuda()
    Om = torch.from_numpy(poses[:, 27:30].copy()).cuda()
    Xt, wt = _cuda(X, w)
    wsum = torch.from_numpy(np.random.default_rng(4).normal(size=(4, 6))).cuda()

    def loss(twist, kn, Xq):
        out = SC.contact(g, twist, kn, 1e4, 0.3, Xq, wt, R0, t0, v0, Om)
        return (out * wsum).sum()

    twist = torch.zeros((4, 6), dtype=torch.float64, device="cuda", requires_grad=True)
    kn = torch.tensor(1e11, dtype=torch.float64, requires_grad=True)
    Xq = Xt.clone().requires_grad_(True)
    L = loss(twist, kn, Xq)
    L.backward()
    assert float(L.detach().abs()) > 0
    h = 1e-7
    with torch.no_grad():
        fd = torch.zeros_like(twist)
        for b in range(4):
            for k in range(6):
                tp, tm = twist.detach().clone(), twist.detach().clone()
                tp[b, k] += h
                tm[b, k] -= h
                fd[b, k] = (loss(tp, kn.detach(), Xt) - loss(tm, kn.detach(), Xt)) / (2 * h)
        gt = twist.grad
        # per component: 1e-4 relative, plus a floor at the difference quotient's
        # rounding (|L| eps / h ~ 1e-9 of the largest component)
        floor = 1e-9 * fd.abs().max()
        assert ((gt - fd).abs() <= 1e-4 * fd.abs() + floor).all()
        dk = 1e11 * 1e-6
        fdk = (loss(twist.detach(), kn.detach() + dk, Xt) - loss(twist.detach(), kn.detach() - dk, Xt)) / (2 * dk)
        assert abs(float(kn.grad) - float(fdk)) <= 1e-4 * abs(float(fdk))
        gx = Xq.grad
        rows = torch.argsort(gx.abs().sum(1), descending=True)[:5]
        floor_x = 1e-9 * gx.abs().max()
        for r in rows.tolist():
            for k in range(3):
                xp, xm = Xt.clone(), Xt.clone()
                xp[r, k] += h
                xm[r, k] -= h
                fdx = (loss(twist.detach(), kn.detach(), xp) - loss(twist.detach(), kn.detach(), xm)) / (2 * h)
                assert abs(float(gx[r, k]) - float(fdx)) <= 1e-4 * abs(float(fdx)) + float(floor_x)


@pytest.mark.parametrize("mode,mu,order,margin", [("Tricubic", 0.3, 2, 0.0),
                                                  ("Trilinear", 0.3, 1, 0.0),
                                                  ("Tricubic", 0.0, 1, 0.058),
                                                  ("Tricubic", 0.2, 0, 0.0)])
def test_saved_forward_and_backward(O, cube, mode, mu, order, margin):
    """The forward that keeps its pairs in contact returns the plain forward's
    bits; the backward on that record matches the oracle and the screening
    backward (up to the order of the sums) and is run-to-run reproducible."""
    g, ref = cube
    print(f"[gpu] saved forward/backward {mode} mu={mu} order={order} margin={margin}",
          flush=True)
    X, w, poses = _scene(40, seed=21)
    Xd, wd, Pd = _cuda(X, w, poses)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=mu, regularizer=1e-6, order=order, mode=mode,
                        narrow_margin=margin)
    plain, act0 = SC.contact_forward(g, Xd, wd, Pd, law, active=True)
    out, act, rec = SC.contact_forward(g, Xd, wd, Pd, law, active=True, save=True)
    assert torch.equal(out, plain) and torch.equal(act, act0)
    assert int(act.sum()) > 0
    gout = torch.from_numpy(np.random.default_rng(3).normal(size=(len(poses), 6))).cuda()
    a = SC.contact_backward(g, Xd, wd, Pd, gout, law, saved=rec)
    b = SC.contact_backward(g, Xd, wd, Pd, gout, law, saved=rec)
    s = SC.contact_backward(g, Xd, wd, Pd, gout, law)
    exp = O.contact_backward(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X, w, poses,
                             gout.cpu().numpy(), kn=1e11, c=1e4, mu=mu, reg=1e-6, order=order,
                             mode=mode, margin=margin)
    for x, y, z, e in zip(a, b, s, exp):
        assert torch.equal(x, y)
        x, z = x.cpu().numpy(), z.cpu().numpy()
        scale = np.abs(e).max() + 1e-300
        assert np.abs(x - e).max() <= 1e-7 * scale
        assert np.abs(x - z).max() <= 1e-12 * scale


def test_saved_record_size_checked(cube):
    from sdfo import _lib
    g, _ = cube
    X, w, poses = _scene(4, seed=2)
    Xd, wd, Pd = _cuda(X, w, poses)
    law = SC.ContactLaw()
    _, rec = SC.contact_forward(g, Xd, wd, Pd, law, save=True)
    gout = torch.zeros((4, 6), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        SC.contact_backward(g, Xd, wd, Pd, gout, law, saved=rec[:10])
    assert _lib.lib.sdfo_contact_saved_bytes(-1, 4) == -1
    assert _lib.lib.sdfo_contact_saved_bytes(len(X), 0) == 0


def test_saved_path_edge_sizes(O, cube):
    """One pose (30 chunks of one pose), and no points at all."""
    g, ref = cube
    X, w, poses = _scene(1, seed=31)
    Xd, wd, Pd = _cuda(X, w, poses)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    out, rec = SC.contact_forward(g, Xd, wd, Pd, law, save=True)
    assert torch.equal(out, SC.contact_forward(g, Xd, wd, Pd, law))
    gout = torch.ones((1, 6), dtype=torch.float64, device="cuda")
    a = SC.contact_backward(g, Xd, wd, Pd, gout, law, saved=rec)
    exp = O.contact_backward(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X, w, poses,
                             gout.cpu().numpy(), kn=1e11, c=1e4, mu=0.3, reg=1e-6, order=2)
    for x, e in zip(a, exp):
        assert np.abs(x.cpu().numpy() - e).max() <= 1e-7 * (np.abs(e).max() + 1e-300)
    X0 = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    w0 = torch.zeros(0, dtype=torch.float64, device="cuda")
    out0, rec0 = SC.contact_forward(g, X0, w0, Pd, law, save=True)
    assert float(out0.abs().max()) == 0.0
    gt, gp, gx, gw = SC.contact_backward(g, X0, w0, Pd, gout, law, saved=rec0)
    assert float(gt.abs().max()) == 0.0 and float(gp.abs().max()) == 0.0
    assert gx.shape == (0, 3) and gw.shape == (0,)


def test_tuner_variants_give_the_same_bits(cube):
    """Every launch variant the autotuner chooses among (forward occupancy,
    recording-forward occupancy, backward persistent grid, query points per
    lane x block size) gives the default variant's bits: a tuning run, then
    each variant pinned in turn."""
    import sdfo

    g, _ = cube
    X, w, poses = _scene(40, seed=23)
    Xd, wd, Pd = _cuda(X, w, poses)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    gout = torch.from_numpy(np.random.default_rng(5).normal(size=(len(poses), 6))).cuda()
    pts = torch.from_numpy(scenes.valid_region_points(g.spec, 20000, seed=9)).cuda()

    def run():
        f = SC.contact_forward(g, Xd, wd, Pd, law)
        fs, rec = SC.contact_forward(g, Xd, wd, Pd, law, save=True)
        b = SC.contact_backward(g, Xd, wd, Pd, gout, law, saved=rec)
        q = g.query(pts, "Tricubic")
        torch.cuda.synchronize()
        return [f, fs, *[x for x in b if x is not None], *q]

    sdfo.tuner_reset()
    ref = run()
    sdfo.tuner_set_mode(1)
    try:
        print("[gpu] tuner: contact and query variants", flush=True)
        tuned = run()
        rep = sdfo.tuner_report()
    finally:
        sdfo.tuner_set_mode(0)
    print(rep, flush=True)
    for a, b in zip(tuned, ref):
        assert torch.equal(a, b)
    keys = {}
    for line in rep.strip().splitlines():
        key, rest = line.split(": ", 1)
        keys[key] = rest.count("ms")
    for prefix in ("contact_forward tricubic points", "contact_forward_save tricubic",
                   "contact_backward_pieces",
                   "query dense tricubic f32"):
        assert any(k.startswith(prefix) for k in keys), prefix
    for key, nvar in keys.items():
        if not key.startswith(("contact_", "query dense tricubic f32")):
            continue
        for v in range(nvar):
            print(f"[gpu] pinned {key} -> {v}", flush=True)
            sdfo._lib.debug_tuner_pin(key, v)
            for a, b in zip(run(), ref):
                assert torch.equal(a, b), (key, v)
    sdfo.tuner_reset()
