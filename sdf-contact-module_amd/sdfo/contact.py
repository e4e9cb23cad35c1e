"""Per-pose penalty / damping / friction contact reduction on the GPU.

Mirrors the per-point law and sums of ``SDFContactEngine.evaluate``
(prototype/src/dynamics/contacts_sdf.py:300-385) evaluated in the SDF body's
frame, batched over poses, with a hand-written backward
(``ContactFunction``, a ``torch.autograd.Function``).

Pose record (30 float64 per pose, see include/sdfo.h):
``R(9) t(3)``: quadrature point X (body A frame) -> SDF frame, ``Y = R X + t``;
``M(9) m(3)``: offset from A's centre of mass, ``off = M X + m``;
``v0(3) Om(3)``: relative velocity ``u = v0 + Om x off``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import MODES, check, lib
from .grid import SdfGrid, _stream_handle


@dataclass
class ContactLaw:
    """ContactPair law parameters (ir/contact.py:1-18, contacts_sdf.py:300-305)."""

    kn: float = 1e11
    damping: float = 1e4
    mu: float = 0.0
    regularizer: float = 1e-6
    narrow_margin: float = 0.0
    order: int = 2
    mode: str = "Tricubic"

    def ctypes(self) -> _lib.ContactParams:
        key = (self.kn, self.damping, self.mu, self.regularizer, self.narrow_margin, self.order,
               self.mode)
        c = self.__dict__.get("_ctypes")
        if c is None or c[0] != key:   # built once per set of values
            c = (key, _lib.ContactParams(float(self.kn), float(self.damping), float(self.mu),
                                         float(self.regularizer), float(self.narrow_margin),
                                         int(self.order), int(MODES.get(self.mode, self.mode))))
            self.__dict__["_ctypes"] = c
        return c[1]


def pose_records(R, t, M=None, m=None, v0=None, Om=None):
    """Packs (B,3,3)/(B,3) tensors into (B,30) float64 pose records.

    Defaults: ``M = R`` and ``m = 0`` (torque about the frame origin of A),
    zero relative velocity."""
    R = torch.as_tensor(R, dtype=torch.float64)
    t = torch.as_tensor(t, dtype=torch.float64, device=R.device)
    B = R.shape[0]
    M = R if M is None else torch.as_tensor(M, dtype=torch.float64, device=R.device)
    z = torch.zeros((B, 3), dtype=torch.float64, device=R.device)
    m = z if m is None else torch.as_tensor(m, dtype=torch.float64, device=R.device)
    v0 = z if v0 is None else torch.as_tensor(v0, dtype=torch.float64, device=R.device)
    Om = z if Om is None else torch.as_tensor(Om, dtype=torch.float64, device=R.device)
    return torch.cat([R.reshape(B, 9), t.reshape(B, 3), M.reshape(B, 9), m.reshape(B, 3),
                      v0.reshape(B, 3), Om.reshape(B, 3)], dim=1).contiguous()


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _check_inputs(X, w, poses):
    for name, t in (("X", X), ("w", w), ("poses", poses)):
        if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous float64 GPU tensor")
    if X.dim() != 2 or X.shape[1] != 3 or w.shape != (X.shape[0],):
        raise ValueError("X must be (N,3) and w (N,)")
    if poses.dim() != 2 or poses.shape[1] != 30:
        raise ValueError("poses must be (B,30)")


def contact_forward(grid: SdfGrid, X, w, poses, law: ContactLaw = ContactLaw(), stream=None,
                    active=False, save=False):
    """(F_A, tau_A) per pose as a (B,6) float64 tensor (SDF frame).

    With ``save``, also returns the record of the (pose, point) pairs in
    contact (a uint8 GPU tensor) for ``contact_backward(..., saved=...)`` on
    the same inputs: ``out[, act], saved``."""
    _check_inputs(X, w, poses)
    B, N = poses.shape[0], X.shape[0]
    out = torch.empty((B, 6), dtype=torch.float64, device=X.device)
    act = torch.empty(B, dtype=torch.int32, device=X.device) if active else None
    p = law.ctypes()
    if save:
        nbytes = int(lib.sdfo_contact_saved_bytes(N, B))
        rec = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=X.device)
        check(lib.sdfo_contact_forward_save(grid.handle, _ptr(X), _ptr(w), N, _ptr(poses), B,
                                            C.byref(p), _ptr(out), _ptr(act), _ptr(rec), nbytes,
                                            _stream_handle(stream)))
        return (out, act, rec) if active else (out, rec)
    check(lib.sdfo_contact_forward(grid.handle, _ptr(X), _ptr(w), N, _ptr(poses), B,
                                   C.byref(p), _ptr(out), _ptr(act), _stream_handle(stream)))
    return (out, act) if active else out


def pair_states(rA, RA, vA, wA, rB, RB, vB, wB):
    """Packs per-pair body states (contacts_sdf.py:231-240) into the (B,36)
    float64 records of ``contact_evaluate``: rA RA vA wA rB RB vB wB (world
    frame, rotations row major).  Each argument is (B,3) or (B,3,3)."""
    parts = [rA, RA, vA, wA, rB, RB, vB, wB]
    ref = next(p for p in parts if torch.is_tensor(p)) if any(torch.is_tensor(p) for p in parts) \
        else None
    dev = ref.device if ref is not None else "cpu"
    ts = [torch.as_tensor(p, dtype=torch.float64, device=dev) for p in parts]
    B = ts[0].reshape(-1, 3).shape[0]
    return torch.cat([t.reshape(B, -1) for t in ts], dim=1).contiguous()


def pair_frames(R_ab, t_ab, RAm, rAm):
    """(B,24) marker-frame records (contacts_sdf.py:243-249, 275-279): the
    action marker relative to the base marker (``Y = R_ab X + t_ab``) and the
    action marker's world pose (``x_w = rAm + RAm X``)."""
    parts = [R_ab, t_ab, RAm, rAm]
    dev = next((p.device for p in parts if torch.is_tensor(p)), "cpu")
    ts = [torch.as_tensor(p, dtype=torch.float64, device=dev) for p in parts]
    B = ts[1].reshape(-1, 3).shape[0]
    return torch.cat([t.reshape(B, -1) for t in ts], dim=1).contiguous()


def _check_evaluate(X, w, states, frames):
    _check_inputs(X, w, torch.zeros((0, 30), dtype=torch.float64, device=X.device))
    if not states.is_cuda or states.dtype != torch.float64 or not states.is_contiguous() \
            or states.dim() != 2 or states.shape[1] != 36:
        raise ValueError("states must be a contiguous (B,36) float64 GPU tensor")
    if frames is not None and (not frames.is_cuda or frames.dtype != torch.float64
                               or not frames.is_contiguous()
                               or tuple(frames.shape) != (states.shape[0], 24)):
        raise ValueError("frames must be a contiguous (B,24) float64 GPU tensor")


def contact_evaluate(grid: SdfGrid, X, w, states, law: "ContactLaw" = None, frames=None,
                     bvh: "PointBvh" = None, body_boxes=None, stream=None, active=False):
    """``SDFContactEngine.evaluate`` (contacts_sdf.py:225-385) for B pairs at
    once: (B,12) float64 = F_A, tau_A, F_B, tau_B in the world frame.

    ``states``: (B,36) from ``pair_states``; ``frames``: optional (B,24) from
    ``pair_frames``; ``law.narrow_margin`` is the engine's narrow filter
    (default 0.01 here, as ``register_pair``); ``bvh``: the broad-phase tree
    over X (``PointBvh(X, 16)``) or None; ``body_boxes``: (A centre, A half,
    B half) for ``_check_broad_phase`` or None."""
    law = ContactLaw(narrow_margin=0.01, order=1) if law is None else law
    _check_evaluate(X, w, states, frames)
    B = states.shape[0]
    out = torch.empty((B, 12), dtype=torch.float64, device=X.device)
    act = torch.empty(B, dtype=torch.int32, device=X.device) if active else None
    p = law.ctypes()
    bb = None
    if body_boxes is not None:
        import numpy as np
        bb = np.ascontiguousarray(np.concatenate([np.ravel(b) for b in body_boxes]), np.float64)
        if bb.shape != (9,):
            raise ValueError("body_boxes = (A centre 3, A half 3, B half 3)")
    check(lib.sdfo_contact_evaluate(grid.handle, None if bvh is None else bvh.handle, _ptr(X),
                                    _ptr(w), X.shape[0], _ptr(states), _ptr(frames), B, C.byref(p),
                                    None if bb is None else bb.ctypes.data_as(C.c_void_p),
                                    _ptr(out), _ptr(act), _stream_handle(stream)))
    return (out, act) if active else out


class PointBvh:
    """Broad-phase AABB tree over a body's quadrature points (AABB_BVH,
    prototype/src/geometry/bvh.py:16-75), built on the host.

    ``query`` is the reference's host query (indices in its order); ``cull``
    evaluates the same leaf test for many boxes on the device."""

    def __init__(self, points, leaf_size=16):
        import numpy as np
        P = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        h = C.c_void_p()
        check(lib.sdfo_point_bvh_build(P.ctypes.data_as(C.c_void_p), len(P), int(leaf_size),
                                       C.byref(h)))
        self.handle = h
        self.n = len(P)
        nn, nl = C.c_int64(), C.c_int64()
        check(lib.sdfo_point_bvh_get_info(h, None, C.byref(nn), C.byref(nl)))
        self.n_nodes, self.n_leaves = nn.value, nl.value

    def close(self):
        if self.handle:
            lib.sdfo_point_bvh_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001  (interpreter shutdown)
            pass

    def query(self, query_min, query_max, margin=0.0):
        """AABB_BVH.query: int32 indices of the points of every overlapping leaf."""
        import numpy as np
        qmin = np.ascontiguousarray(query_min, np.float64).reshape(3)
        qmax = np.ascontiguousarray(query_max, np.float64).reshape(3)
        out = np.empty(self.n, np.int32)
        cnt = C.c_int64()
        check(lib.sdfo_point_bvh_query(self.handle, qmin.ctypes.data_as(C.c_void_p),
                                       qmax.ctypes.data_as(C.c_void_p), float(margin),
                                       out.ctypes.data_as(C.c_void_p), C.byref(cnt)))
        return out[:cnt.value].copy()

    def cull(self, boxes, margin=0.0, stream=None):
        """(B,N) uint8 device mask for (B,6) float64 device boxes (qmin, qmax)."""
        if not boxes.is_cuda or boxes.dtype != torch.float64 or boxes.dim() != 2 \
                or boxes.shape[1] != 6:
            raise ValueError("boxes must be a (B,6) float64 GPU tensor")
        boxes = boxes.contiguous()
        mask = torch.empty((boxes.shape[0], self.n), dtype=torch.uint8, device=boxes.device)
        check(lib.sdfo_point_bvh_cull(self.handle, _ptr(boxes), boxes.shape[0], float(margin),
                                      _ptr(mask), _stream_handle(stream)))
        return mask


def broad_phase_boxes(poses, bmin, bmax, margin):
    """Cull boxes of SDFContactEngine.evaluate (contacts_sdf.py:247-258): the
    SDF box corners ``bmin``, ``bmax`` mapped into A's frame by the inverse of
    ``Y = R X + t``, grown by ``margin`` (the engine passes
    ``max(narrow_margin, 0.05)``).  (B,30) poses -> (B,6) (qmin, qmax)."""
    R = poses[:, :9].reshape(-1, 3, 3)
    t = poses[:, 9:12]
    Rt = R.transpose(1, 2)
    t_ba = -(Rt @ t[:, :, None])[..., 0]
    lo = torch.as_tensor(bmin, dtype=torch.float64, device=poses.device).reshape(1, 3, 1)
    hi = torch.as_tensor(bmax, dtype=torch.float64, device=poses.device).reshape(1, 3, 1)
    qmin = (Rt @ lo)[..., 0] + t_ba - margin
    qmax = (Rt @ hi)[..., 0] + t_ba + margin
    return torch.cat([qmin, qmax], dim=1).contiguous()


def contact_forward_culled(grid: SdfGrid, bvh: PointBvh, boxes, X, w, poses,
                           law: ContactLaw = ContactLaw(), stream=None, active=False):
    """contact_forward behind the broad phase (contacts_sdf.py:245-272): for
    pose b only the points of ``bvh``'s leaves that overlap ``boxes[b]`` take
    part, or all points when none does."""
    _check_inputs(X, w, poses)
    boxes = boxes.contiguous()
    if boxes.shape != (poses.shape[0], 6) or boxes.dtype != torch.float64 or not boxes.is_cuda:
        raise ValueError("boxes must be a (B,6) float64 GPU tensor")
    B = poses.shape[0]
    out = torch.empty((B, 6), dtype=torch.float64, device=X.device)
    act = torch.empty(B, dtype=torch.int32, device=X.device) if active else None
    p = law.ctypes()
    check(lib.sdfo_contact_forward_culled(grid.handle, bvh.handle, _ptr(boxes), _ptr(X), _ptr(w),
                                          X.shape[0], _ptr(poses), B, C.byref(p), _ptr(out),
                                          _ptr(act), _stream_handle(stream)))
    return (out, act) if active else out


class LinearizationCache:
    """SDF linearisation cache for Newton iterations (contacts_sdf.py:40-50,
    307-327; enabled around each implicit step by implicit_be.py:474-500).

    ``fill(pose)`` queries the field once at a reference pose and keeps the
    gaps, raw gradients and validity on the device; ``forward(poses)`` then
    evaluates the contact law on ``g0 + grad0 . (t - t_ref)`` for every pose
    without touching the SDF.  Entries are kept per quadrature point; a point
    takes part only where it was valid at the reference pose and passes the
    narrow filter at the current one.  ``clear()`` drops the cache (the
    reference's ``_cache_cleared``); ``forward`` on an empty cache fills it at
    the first pose, as the reference's first evaluation of a step does."""

    def __init__(self, grid: SdfGrid, X, w, law: ContactLaw = ContactLaw()):
        _check_inputs(X, w, torch.zeros((0, 30), dtype=torch.float64, device=X.device))
        self.grid, self.X, self.w, self.law = grid, X, w, law
        n = X.shape[0]
        self.g0 = torch.empty(n, dtype=torch.float64, device=X.device)
        self.grad0 = torch.empty((n, 3), dtype=torch.float64, device=X.device)
        self.valid0 = torch.empty(n, dtype=torch.uint8, device=X.device)
        self.t_ref = None

    @property
    def filled(self) -> bool:
        return self.t_ref is not None

    def clear(self):
        self.t_ref = None

    def fill(self, pose, stream=None):
        """Caches the field at one (30,) or (1,30) pose record."""
        pose = pose.reshape(30).contiguous()
        if not pose.is_cuda or pose.dtype != torch.float64:
            raise ValueError("pose must be a float64 GPU tensor")
        p = self.law.ctypes()
        check(lib.sdfo_contact_linearize(self.grid.handle, _ptr(self.X), self.X.shape[0],
                                         _ptr(pose), C.byref(p), _ptr(self.g0), _ptr(self.grad0),
                                         _ptr(self.valid0), _stream_handle(stream)))
        self.t_ref = (C.c_double * 3)(*pose[9:12].tolist())
        return self

    def forward(self, poses, stream=None, active=False, bvh: PointBvh = None, boxes=None):
        """(F_A, tau_A) per pose as a (B,6) float64 tensor, on the cache;
        with ``bvh`` and ``boxes``, behind the broad phase of
        ``contact_forward_culled``."""
        _check_inputs(self.X, self.w, poses)
        if (bvh is None) != (boxes is None):
            raise ValueError("bvh and boxes go together")
        if boxes is not None:
            boxes = boxes.contiguous()
            if boxes.shape != (poses.shape[0], 6) or boxes.dtype != torch.float64:
                raise ValueError("boxes must be a (B,6) float64 GPU tensor")
        if not self.filled:
            self.fill(poses[0], stream)
        B = poses.shape[0]
        out = torch.empty((B, 6), dtype=torch.float64, device=self.X.device)
        act = torch.empty(B, dtype=torch.int32, device=self.X.device) if active else None
        p = self.law.ctypes()
        check(lib.sdfo_contact_forward_linearized(
            _ptr(self.g0), _ptr(self.grad0), _ptr(self.valid0), self.t_ref, _ptr(self.X),
            _ptr(self.w), self.X.shape[0], _ptr(poses), B, C.byref(p),
            None if bvh is None else bvh.handle, _ptr(boxes), _ptr(out), _ptr(act),
            _stream_handle(stream)))
        return (out, act) if active else out


def contact_backward(grid: SdfGrid, X, w, poses, g_out, law: ContactLaw = ContactLaw(),
                     want_twist=True, want_params=True, want_X=True, want_w=True, stream=None,
                     saved=None):
    """Gradients of sum(g_out * forward): (g_twist (B,6) [omega, dt] left twist,
    g_params (3,) [kn, damping, mu], g_X (N,3), g_w (N,)).

    ``saved``: the record of ``contact_forward(..., save=True)`` on the same
    grid, X, w, poses and law; the backward then reads which pairs are in
    contact instead of screening every pair again."""
    _check_inputs(X, w, poses)
    g_out = g_out.contiguous().to(torch.float64)
    B, N = poses.shape[0], X.shape[0]
    if g_out.shape != (B, 6) or not g_out.is_cuda:
        raise ValueError("g_out must be a (B,6) GPU tensor")
    # one allocation for all outputs, sliced into views
    o1 = 6 * B if want_twist else 0
    o2 = o1 + (3 if want_params else 0)
    o3 = o2 + (3 * N if want_X else 0)
    buf = torch.empty(o3 + (N if want_w else 0), dtype=torch.float64, device=X.device)
    gt = buf[:o1].view(B, 6) if want_twist else None
    gp = buf[o1:o2] if want_params else None
    gx = buf[o2:o3].view(N, 3) if want_X else None
    gw = buf[o3:] if want_w else None
    p = law.ctypes()
    if saved is not None:
        nbytes = int(lib.sdfo_contact_saved_bytes(N, B))
        if saved.dtype != torch.uint8 or not saved.is_cuda or saved.numel() < nbytes:
            raise ValueError("saved must be the record of contact_forward(..., save=True) "
                             "on the same inputs")
        check(lib.sdfo_contact_backward_saved(grid.handle, _ptr(X), _ptr(w), N, _ptr(poses), B,
                                              C.byref(p), _ptr(g_out), _ptr(saved), nbytes,
                                              _ptr(gt), _ptr(gp), _ptr(gx), _ptr(gw),
                                              _stream_handle(stream)))
        return gt, gp, gx, gw
    check(lib.sdfo_contact_backward(grid.handle, _ptr(X), _ptr(w), N, _ptr(poses), B, C.byref(p),
                                    _ptr(g_out), _ptr(gt), _ptr(gp), _ptr(gx), _ptr(gw),
                                    _stream_handle(stream)))
    return gt, gp, gx, gw


def _hat(v):
    z = torch.zeros_like(v[..., 0])
    return torch.stack([torch.stack([z, -v[..., 2], v[..., 1]], -1),
                        torch.stack([v[..., 2], z, -v[..., 0]], -1),
                        torch.stack([-v[..., 1], v[..., 0], z], -1)], -2)


def so3_exp(omega):
    """Rodrigues: exp([omega]x) for (B,3) float64."""
    th = torch.linalg.vector_norm(omega, dim=-1, keepdim=True)[..., None]
    K = _hat(omega)
    small = th < 1e-8
    ths = torch.where(small, torch.ones_like(th), th)
    a = torch.where(small, 1.0 - th ** 2 / 6.0, torch.sin(ths) / ths)
    b = torch.where(small, 0.5 - th ** 2 / 24.0, (1.0 - torch.cos(ths)) / ths ** 2)
    eye = torch.eye(3, dtype=omega.dtype, device=omega.device).expand_as(K)
    return eye + a * K + b * (K @ K)


def so3_left_jacobian(omega):
    """J_l(omega) = I + (1 - cos t)/t^2 [w] + (t - sin t)/t^3 [w]^2."""
    th = torch.linalg.vector_norm(omega, dim=-1, keepdim=True)[..., None]
    K = _hat(omega)
    small = th < 1e-6
    ths = torch.where(small, torch.ones_like(th), th)
    a = torch.where(small, 0.5 - th ** 2 / 24.0, (1.0 - torch.cos(ths)) / ths ** 2)
    b = torch.where(small, 1.0 / 6.0 - th ** 2 / 120.0, (ths - torch.sin(ths)) / ths ** 3)
    eye = torch.eye(3, dtype=omega.dtype, device=omega.device).expand_as(K)
    return eye + a * K + b * (K @ K)


class ContactFunction(torch.autograd.Function):
    """F, tau (B,6) of the contact reduction as a function of a per-pose twist
    ``(omega, dt)`` applied to base poses ``R = exp([omega]) R0``, ``t = t0 + dt``
    (torque about A's origin, ``off = R X``), of ``kn``, ``damping``, ``mu``, of
    the quadrature points ``X`` and weights ``w``."""

    @staticmethod
    def forward(ctx, twist, kn, damping, mu, X, w, grid, R0, t0, v0, Om, regularizer, order,
                mode):
        R = so3_exp(twist[:, :3]) @ R0
        t = t0 + twist[:, 3:]
        poses = pose_records(R, t, v0=v0, Om=Om)
        law = ContactLaw(float(kn), float(damping), float(mu), float(regularizer), 0.0, int(order),
                         mode)
        Xc = X.detach().contiguous()
        wc = w.detach().contiguous()
        out, rec = contact_forward(grid, Xc, wc, poses, law, save=True)
        ctx.save_for_backward(twist, Xc, wc, poses, rec)
        ctx.grid = grid
        ctx.law = law
        return out

    @staticmethod
    def backward(ctx, g_out):
        twist, X, w, poses, rec = ctx.saved_tensors
        need = ctx.needs_input_grad
        gt, gp, gx, gw = contact_backward(ctx.grid, X, w, poses, g_out, ctx.law,
                                          want_twist=need[0], want_params=any(need[1:4]),
                                          want_X=need[4], want_w=need[5], saved=rec)
        g_twist = None
        if need[0]:
            Jl = so3_left_jacobian(twist[:, :3])
            g_omega = (Jl.transpose(1, 2) @ gt[:, :3, None])[..., 0]
            g_twist = torch.cat([g_omega, gt[:, 3:]], dim=1)
        gkn = gp[0] if need[1] else None
        gc = gp[1] if need[2] else None
        gmu = gp[2] if need[3] else None
        return (g_twist, gkn, gc, gmu, gx if need[4] else None, gw if need[5] else None, None,
                None, None, None, None, None, None, None)


def contact(grid, twist, kn, damping, mu, X, w, R0, t0, v0=None, Om=None, regularizer=1e-6,
            order=2, mode="Tricubic"):
    """Differentiable batched contact: returns (B,6) = (F_A, tau_A) per pose."""
    B = R0.shape[0]
    z = torch.zeros((B, 3), dtype=torch.float64, device=R0.device)
    v0 = z if v0 is None else v0
    Om = z if Om is None else Om
    as_t = lambda v: v if torch.is_tensor(v) else torch.tensor(float(v), dtype=torch.float64)  # noqa: E731
    return ContactFunction.apply(twist, as_t(kn), as_t(damping), as_t(mu), X, w, grid, R0, t0, v0,
                                 Om, regularizer, order, mode)


class SdfContactEngine:
    """Pair registry and ``evaluate`` of ``SDFContactEngine``
    (contacts_sdf.py:33-66, 68-89, 225-385) on the device, batched over the
    body states of many evaluations of one pair.

    ``register_pair`` normalises the weights to sum 1, keeps A's point AABB
    and B's half extents for the body-level check and builds the broad-phase
    tree (leaf size 16); like the reference, the narrow margin is one value
    for the engine, the one given at the last registration."""

    def __init__(self):
        self._pairs = {}
        self._narrow_margin = 0.01

    def register_pair(self, cp_id, X_q, w_q, grid: SdfGrid, aabb_b_half=None,
                      narrow_margin=0.01, device=None):
        import numpy as np
        X = np.ascontiguousarray(X_q, np.float64).reshape(-1, 3)
        w = np.array(w_q, np.float64).reshape(-1)
        total = w.sum()
        if total > 1e-30:
            w = w / total
        if len(X):
            lo, hi = X.min(axis=0), X.max(axis=0)
            box_a = ((lo + hi) / 2, (hi - lo) / 2)
        else:
            box_a = (np.zeros(3), np.zeros(3))
        dev = torch.device("cuda") if device is None else torch.device(device)
        self._pairs[cp_id] = dict(
            grid=grid, X=torch.from_numpy(X).to(dev), w=torch.from_numpy(w).to(dev),
            bvh=PointBvh(X, leaf_size=16) if len(X) else None, box_a=box_a,
            half_b=None if aabb_b_half is None else np.asarray(aabb_b_half, np.float64).reshape(3))
        self._narrow_margin = float(narrow_margin)

    def evaluate(self, cp_id, states, law: ContactLaw = ContactLaw(order=1), frames=None,
                 stream=None):
        """(F_A, tau_A, F_B, tau_B), each (B,3) float64, world frame."""
        B = states.shape[0]
        pair = self._pairs.get(cp_id)
        if pair is None or pair["X"].shape[0] == 0:
            z = torch.zeros((B, 3), dtype=torch.float64, device=states.device)
            return z, z.clone(), z.clone(), z.clone()
        lw = ContactLaw(law.kn, law.damping, law.mu, law.regularizer, self._narrow_margin,
                        law.order, law.mode)
        boxes = None if pair["half_b"] is None else (*pair["box_a"], pair["half_b"])
        out = contact_evaluate(pair["grid"], pair["X"], pair["w"], states, lw, frames=frames,
                               bvh=pair["bvh"], body_boxes=boxes, stream=stream)
        return out[:, 0:3], out[:, 3:6], out[:, 6:9], out[:, 9:12]
