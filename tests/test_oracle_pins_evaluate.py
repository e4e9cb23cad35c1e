"""Pins the oracle's contact law (contacts_sdf.py:300-385) where the plane
checks elsewhere leave it open: a scaled gradient (|grad phi| != 1, which
separates p from p_hat = p |grad phi|), a relative velocity with normal and
tangential parts (which exercises the projection u_t = u - (u.n)n), and the
world-frame evaluate (rg = R_B grad, F_B, tau_B) with a rotated body B.

Expectations are worked out by hand for a linear field phi = s (z - z0),
which the tricubic interpolant reproduces exactly, plus invariance of
evaluate under a rigid motion of the whole scene.
"""
import numpy as np
import pytest

from sdfo import meshes, scenes

S = 2.5
Z0 = 0.1
LAW = dict(kn=1e8, c=1e3, mu=0.3, reg=1e-3, order=2)


def _plane_grid(O):
    g = O.Grid(np.array([-0.3, -0.2, -0.25]), None, np.array([0.05, 0.04, 0.06]), (16, 17, 18))
    g.bmax = g.bmin + g.voxel * np.array(g.shape)
    c = g.centers()
    return g, dict(phi=S * (c[:, 2] - Z0), n=np.tile([0.0, 0.0, S], (len(c), 1)))


def _hand_force(depth_z, u, kn, c, mu, reg, order):
    """Per-point force of the law in the plane's frame for a point at
    z = Z0 - depth_z moving with velocity u."""
    g = -S * depth_z
    p = kn * (-g) ** order
    G = np.array([0.0, 0.0, S])
    fp = p * G
    fd = c * (-(u @ G)) * G
    ut = np.array([u[0], u[1], 0.0])
    ft = -mu * (p * S) * ut / np.sqrt(ut @ ut + reg * reg)
    return fp + fd + ft


def test_friction_projection_and_phat_closed_form(O):
    g, d = _plane_grid(O)
    X, w = meshes.annular_quadrature(0.05, 0.15, 6, 24)
    w = w / w.sum()
    R = scenes.rot_z(0.3)[None]
    t = np.array([[0.1, 0.14, Z0 - 0.002]])
    m = np.array([[0.01, 0.02, 0.03]])
    u = np.array([0.3, -0.2, -0.05])
    P = scenes.pose_array(R, t, m=m, v0=u[None])
    out, act = O.contact_forward(g, d["phi"], d["n"], X, w, P, mode="Tricubic", **LAW)
    assert act[0] == len(X)
    fq = _hand_force(0.002, u, **LAW)
    off = X @ R[0].T + m[0]
    assert np.allclose(out[0, :3], fq, rtol=1e-11, atol=0)
    assert np.allclose(out[0, 3:], np.cross(w @ off, fq), rtol=1e-10, atol=1e-12 * np.abs(fq).max())
    # the friction term is a third of the tangential force budget here, so
    # dropping -(u.n)n (u_t gains u_z) or using p for p_hat both show up
    wrong_phat = fq.copy()
    wrong_phat[:2] /= S
    assert not np.allclose(out[0, :3], wrong_phat, rtol=1e-6)


def _states_for(O, RB, rB, Rs, ts, vA, wA, vB, wB):
    RA = RB @ Rs
    rA = rB + RB @ ts
    return O.pair_states(rA[None], RA[None], vA[None], wA[None], rB[None], RB[None], vB[None],
                         wB[None]), RA, rA


def test_evaluate_world_frame_closed_form_rotated_b(O):
    g, d = _plane_grid(O)
    X, w = meshes.annular_quadrature(0.05, 0.15, 6, 24)
    w = w / w.sum()
    RB = scenes.rot_axis(np.array([1.0, 1.0, 0.3]), np.radians(50.0))
    rB = np.array([0.3, -0.1, 0.2])
    Rs, ts = scenes.rot_z(0.3), np.array([0.1, 0.14, Z0 - 0.002])
    vA = np.array([0.2, 0.1, -0.3])
    st, RA, rA = _states_for(O, RB, rB, Rs, ts, vA, np.zeros(3), np.zeros(3), np.zeros(3))
    out, act = O.contact_evaluate(g, d["phi"], d["n"], X, w, st, margin=0.2, mode="Tricubic", **LAW)
    assert act[0] == len(X)
    fq_local = _hand_force(0.002, RB.T @ vA, **LAW)
    FA = RB @ fq_local
    off = X @ RA.T                               # x_w - r_A
    TA = np.cross(w @ off, FA)
    scale = np.abs(FA).max()
    assert np.allclose(out[0, 0:3], FA, rtol=1e-11, atol=0)
    assert np.allclose(out[0, 3:6], TA, rtol=0, atol=1e-11 * scale)
    assert np.array_equal(out[0, 6:9], -out[0, 0:3])
    assert np.allclose(out[0, 9:12], -(TA + np.cross(rA - rB, FA)), rtol=0, atol=1e-11 * scale)
    # the B-frame force differs from the world one by R_B: a missing rotation fails
    assert np.abs(fq_local - FA).max() > 0.1 * scale
    # the narrow filter |Y_z| < margin applies in B's frame (Y_z = 0.098 here)
    out0, act0 = O.contact_evaluate(g, d["phi"], d["n"], X, w, st, margin=0.05, mode="Tricubic",
                                    **LAW)
    assert act0[0] == 0 and not out0.any()


def _scene(O, n_pairs=4, seed=3):
    v, t = scenes.cube_mesh()
    g = O.setup_grid(v, 0.1, 0.0, 32, 32)
    d = O.build_dense(v, t, g)
    X, w = scenes.ring_quadrature(30, 8)
    Rs, ts = scenes.ring_poses(n_pairs, seed=seed, z_range=(0.057, 0.060))
    rng = np.random.default_rng(seed)
    states = []
    for b in range(n_pairs):
        RB = scenes.rot_axis(rng.normal(size=3), rng.uniform(0, np.pi))
        rB = rng.normal(0, 0.2, 3)
        st, _, _ = _states_for(O, RB, rB, Rs[b], ts[b], *rng.normal(0, 0.3, (4, 3)))
        states.append(st[0])
    return g, d, X, w, np.array(states), Rs, ts


def test_evaluate_invariant_under_rigid_motion_of_the_scene(O):
    g, d, X, w, st, _, _ = _scene(O)
    law = dict(LAW, kn=1e11, c=1e4, reg=1e-6)
    out, act = O.contact_evaluate(g, d["phi"], d["n"], X, w, st, margin=0.1, **law)
    assert (act > 0).all()
    Q = scenes.rot_axis(np.array([0.2, -1.0, 0.5]), 1.1)
    s = np.array([1.5, -0.7, 2.0])
    st2 = st.copy()
    for b in range(len(st)):
        rA, RA, vA, wA, rB, RB, vB, wB = O._state_fields(st[b])
        st2[b] = O.pair_states((Q @ rA + s)[None], (Q @ RA)[None], (Q @ vA)[None],
                               (Q @ wA)[None], (Q @ rB + s)[None], (Q @ RB)[None],
                               (Q @ vB)[None], (Q @ wB)[None])[0]
    out2, act2 = O.contact_evaluate(g, d["phi"], d["n"], X, w, st2, margin=0.1, **law)
    assert np.array_equal(act, act2)
    scale = np.abs(out).max()
    for k in range(4):
        assert np.allclose(out2[:, 3 * k:3 * k + 3], out[:, 3 * k:3 * k + 3] @ Q.T, rtol=0,
                           atol=1e-9 * scale)


def test_evaluate_equals_b_frame_forward_when_b_is_the_world(O):
    """With R_B = I, r_B = 0 and B at rest, evaluate's sums are the B-frame
    forward with R = M = R_A, t = r_A, m = 0, v0 = v_A, Om = w_A."""
    g, d, X, w, st, Rs, ts = _scene(O, n_pairs=3, seed=9)
    law = dict(LAW, kn=1e11, c=1e4, reg=1e-6)
    for b in range(len(st)):
        _, _, vA, wA, _, _, _, _ = O._state_fields(st[b])
        st[b] = O.pair_states(ts[b][None], Rs[b][None], vA[None], wA[None], np.zeros((1, 3)),
                              np.eye(3)[None], np.zeros((1, 3)), np.zeros((1, 3)))[0]
    out, act = O.contact_evaluate(g, d["phi"], d["n"], X, w, st, margin=1.0, **law)
    P = np.stack([scenes.pose_array(O._state_fields(s_)[1][None], O._state_fields(s_)[0][None],
                                    v0=O._state_fields(s_)[2][None],
                                    Om=O._state_fields(s_)[3][None])[0] for s_ in st])
    ref, ra = O.contact_forward(g, d["phi"], d["n"], X, w, P, margin=1.0, **law)
    assert np.array_equal(act, ra) and (ra > 0).all()
    assert np.allclose(out[:, :6], ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_evaluate_frames_keep_mask_and_pair_gate(O):
    g, d, X, w, st, _, _ = _scene(O, n_pairs=3, seed=5)
    law = dict(LAW, kn=1e11, c=1e4, reg=1e-6)
    out, act = O.contact_evaluate(g, d["phi"], d["n"], X, w, st, margin=0.1, **law)
    # marker frames equal to the body frames reproduce the frameless result
    fr = np.zeros((len(st), 24))
    for b in range(len(st)):
        rA, RA, _, _, rB, RB, _, _ = O._state_fields(st[b])
        fr[b] = np.concatenate([(RB.T @ RA).ravel(), RB.T @ (rA - rB), RA.ravel(), rA])
    out_f, act_f = O.contact_evaluate(g, d["phi"], d["n"], X, w, st, frames=fr, margin=0.1, **law)
    assert np.array_equal(act, act_f)
    assert np.allclose(out_f, out, rtol=0, atol=1e-10 * np.abs(out).max())
    # a broad-phase mask equals evaluating the kept subset of the points
    rng = np.random.default_rng(1)
    keep = (rng.random((len(st), len(X))) < 0.6).astype(np.uint8)
    out_k, _ = O.contact_evaluate(g, d["phi"], d["n"], X, w, st, keep=keep, margin=0.1, **law)
    for b in range(len(st)):
        sel = keep[b].astype(bool)
        sub, _ = O.contact_evaluate(g, d["phi"], d["n"], X[sel], w[sel], st[b:b + 1], margin=0.1,
                                    **law)
        assert np.array_equal(out_k[b], sub[0])
    # a pair rejected by the body-level check returns zeros
    out_g, act_g = O.contact_evaluate(g, d["phi"], d["n"], X, w, st, pair_on=[1, 0, 1], margin=0.1,
                                      **law)
    assert not out_g[1].any() and act_g[1] == 0 and np.array_equal(out_g[[0, 2]], out[[0, 2]])


def test_evaluate_broad_phase_known_answers(O):
    """evaluate's cull box (contacts_sdf.py:243-271) and body check (:68-89)."""
    X = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0], [3.0, 0, 0]])
    nodes = O.point_bvh_build(X, leaf_size=1)
    ident = O.pair_states(np.zeros((1, 3)), np.eye(3)[None], np.zeros((1, 3)), np.zeros((1, 3)),
                          np.zeros((1, 3)), np.eye(3)[None], np.zeros((1, 3)), np.zeros((1, 3)))
    # SDF box x in [0.9, 2.2] grown by max(0.01, 0.05): points 1 and 2
    k = O.evaluate_keep_mask(nodes, ident, None, [0.9, -1, -1], [2.2, 1, 1], 0.01, 4)
    assert list(k[0]) == [0, 1, 1, 0]
    # A shifted by +1 in x: the box maps to x in [-0.15, 1.25] in A's frame
    shifted = ident.copy()
    shifted[0, 0] = 1.0
    k = O.evaluate_keep_mask(nodes, shifted, None, [0.9, -1, -1], [2.2, 1, 1], 0.01, 4)
    assert list(k[0]) == [1, 1, 0, 0]
    # nothing overlaps: every point takes part
    k = O.evaluate_keep_mask(nodes, ident, None, [10, 10, 10], [11, 11, 11], 0.01, 4)
    assert k.all()
    # body check: A's box (centre 1.5, half 1.5 in x) rotated 90 deg about z
    st = ident.copy()
    st[0, 3:12] = scenes.rot_z(np.pi / 2).ravel()
    st[0, 18:21] = [0.0, 3.5, 0.0]                 # B at y = 3.5 with half 0.4
    on = O.evaluate_body_check(st, [1.5, 0, 0], [1.5, 0.1, 0.1], [0.4, 0.4, 0.4])
    # rotated A spans y in [0, 3] around centre (0, 1.5): |3.5 - 1.5| = 2.0 > 1.5 + 0.4
    assert on[0] == 0
    st[0, 19] = 3.3
    assert O.evaluate_body_check(st, [1.5, 0, 0], [1.5, 0.1, 0.1], [0.4, 0.4, 0.4])[0] == 1
