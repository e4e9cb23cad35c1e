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
from sdfo import scenes

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
