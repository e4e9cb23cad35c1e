"""World-frame SDFContactEngine.evaluate (contacts_sdf.py:225-385) on the GPU
against the oracle's step-for-step restatement.

Body B (the SDF) is rotated and translated, so the rotation of the raw
gradient into the world (rg = R_B grad) and the F_B, tau_B terms are all
exercised; both sides evaluate the law in FP64 on the same FP32-rounded field
and differ by FMA contraction and summation order: 1e-9 relative.
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


def _states(O, n_pairs, seed):
    """Ring poses relative to the cube, carried by a randomly placed body B."""
    Rs, ts = scenes.ring_poses(n_pairs, seed=seed)
    rng = np.random.default_rng(seed)
    st = np.zeros((n_pairs, 36))
    for b in range(n_pairs):
        RB = scenes.rot_axis(rng.normal(size=3), rng.uniform(0.2, np.pi))
        rB = rng.normal(0, 0.3, 3)
        RA, rA = RB @ Rs[b], rB + RB @ ts[b]
        vA, wA, vB, wB = rng.normal(0, 0.2, (4, 3)) * [[1], [5], [1], [5]]
        st[b] = O.pair_states(rA[None], RA[None], vA[None], wA[None], rB[None], RB[None],
                              vB[None], wB[None])[0]
    return st


def _cmp(got, exp, tol=1e-9):
    scale = np.abs(exp).max(axis=0) + 1e-300
    assert np.all(np.abs(got - exp) <= tol * scale), np.abs(got - exp).max()


@pytest.mark.parametrize("mode,mu,order,margin", [("Tricubic", 0.3, 2, 0.07),
                                                  ("Tricubic", 0.0, 1, 0.01),
                                                  ("Trilinear", 0.5, 2, 0.07)])
def test_evaluate_matches_oracle_rotated_b(O, cube, mode, mu, order, margin):
    g, ref = cube
    print(f"[gpu] evaluate {mode} mu={mu} order={order} margin={margin}", flush=True)
    X, w = scenes.ring_quadrature(60, 16)
    st = _states(O, 48, seed=31)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=mu, regularizer=1e-6, narrow_margin=margin,
                        order=order, mode=mode)
    Xt, wt, St = (torch.from_numpy(a).cuda() for a in (X, w, st))
    out, act = SC.contact_evaluate(g, Xt, wt, St, law, active=True)
    o_ref, a_ref = O.contact_evaluate(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X, w, st,
                                      kn=1e11, c=1e4, mu=mu, reg=1e-6, margin=margin, order=order,
                                      mode=mode)
    assert np.array_equal(act.cpu().numpy(), a_ref)
    if margin > 0.05:
        assert (a_ref > 0).sum() > 8
    _cmp(out.cpu().numpy(), o_ref)


def test_evaluate_frames_broad_phase_and_body_check(O, cube):
    g, ref = cube
    print("[gpu] evaluate frames + bvh + body check", flush=True)
    X, w = scenes.ring_quadrature(60, 16)
    st = _states(O, 40, seed=7)
    rng = np.random.default_rng(2)
    # marker frames: the body frames with the action marker offset inside A
    fr = np.zeros((len(st), 24))
    for b in range(len(st)):
        rA, RA, _, _, rB, RB, _, _ = O._state_fields(st[b])
        d = rng.normal(0, 1e-3, 3)
        rAm = rA + RA @ d
        fr[b] = np.concatenate([(RB.T @ RA).ravel(), RB.T @ (rAm - rB), RA.ravel(), rAm])
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, narrow_margin=0.07,
                        order=2)
    kw = dict(kn=1e11, c=1e4, mu=0.3, reg=1e-6, margin=0.07, order=2, mode="Tricubic")
    og = ograd(O, g.spec)
    phi, n3 = f32(ref["phi"]), f32(ref["n"])
    Xt, wt, St, Ft = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (X, w, st, fr))
    # frames
    out, act = SC.contact_evaluate(g, Xt, wt, St, law, frames=Ft, active=True)
    o_ref, a_ref = O.contact_evaluate(og, phi, n3, X, w, st, frames=fr, **kw)
    assert np.array_equal(act.cpu().numpy(), a_ref) and (a_ref > 0).sum() > 8
    _cmp(out.cpu().numpy(), o_ref)
    # broad phase: the ring's points under the SDF box (a thin box keeps part of them)
    bvh = SC.PointBvh(X, leaf_size=16)
    nodes = O.point_bvh_build(X, leaf_size=16)
    keep = O.evaluate_keep_mask(nodes, st, None, g.spec.bmin, g.spec.bmax, 0.07, len(X))
    out, act = SC.contact_evaluate(g, Xt, wt, St, law, bvh=bvh, active=True)
    o_ref, a_ref = O.contact_evaluate(og, phi, n3, X, w, st, keep=keep, **kw)
    assert np.array_equal(act.cpu().numpy(), a_ref)
    _cmp(out.cpu().numpy(), o_ref)
    # body-level check: B's half extents small enough to reject some pairs
    lo, hi = X.min(axis=0), X.max(axis=0)
    ca, ha = (lo + hi) / 2, (hi - lo) / 2
    dists = []
    for b in range(len(st)):
        rA, RA, _, _, rB, _, _, _ = O._state_fields(st[b])
        dists.append(np.abs(rA + RA @ ca - rB) - np.abs(RA) @ ha)
    hb = np.full(3, np.median(np.max(dists, axis=1)))
    on = O.evaluate_body_check(st, ca, ha, hb)
    assert 0 < on.sum() < len(st)
    out, act = SC.contact_evaluate(g, Xt, wt, St, law, bvh=bvh, body_boxes=(ca, ha, hb),
                                   active=True)
    o_ref, a_ref = O.contact_evaluate(og, phi, n3, X, w, st, keep=keep, pair_on=on, **kw)
    assert np.array_equal(act.cpu().numpy(), a_ref)
    _cmp(out.cpu().numpy(), o_ref)
    assert not out[torch.from_numpy(on == 0).cuda()].any()


def test_engine_mirror_matches_evaluate(O, cube):
    """SdfContactEngine.register_pair + evaluate: weights normalised, tree
    and narrow margin of the registration, split into four (B,3) results."""
    g, ref = cube
    X, w = scenes.ring_quadrature(60, 16)
    st = _states(O, 16, seed=12)
    eng = SC.SdfContactEngine()
    eng.register_pair(7, X, 3.0 * w / w.sum(), g, narrow_margin=0.07)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    FA, TA, FB, TB = eng.evaluate(7, torch.from_numpy(st).cuda(), law)
    wn = w / w.sum()
    nodes = O.point_bvh_build(X, leaf_size=16)
    keep = O.evaluate_keep_mask(nodes, st, None, g.spec.bmin, g.spec.bmax, 0.07, len(X))
    o_ref, _ = O.contact_evaluate(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X, wn, st,
                                  keep=keep, kn=1e11, c=1e4, mu=0.3, reg=1e-6, margin=0.07,
                                  order=2)
    _cmp(torch.cat([FA, TA, FB, TB], 1).cpu().numpy(), o_ref)
    z = eng.evaluate(99, torch.from_numpy(st).cuda(), law)
    assert all(not t.any() for t in z)
