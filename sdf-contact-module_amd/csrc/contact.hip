// Per-point penalty / damping / regularised Coulomb friction contact
// reduction (contacts_sdf.py:329-385) and its hand-written backward.
//
// All arithmetic after the field gather is FP64 and every sum has a fixed
// order (forward: per-lane strided partials, LDS tree, then chunk partials in
// chunk order; backward: a fixed butterfly per piece of 64 listed pairs, then
// the pieces in list order), so results are bit-reproducible run to run.  The
// backward runs only on the (pose, point) pairs in contact, listed by a
// screen or by the forward that preceded it, and recomputes their forward
// terms instead of storing them.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>

#include "interp_device.hpp"
#include "sdfo_internal.hpp"

namespace sdfo {

namespace dev {
QGrid makeQGrid(const sdfo_grid& g);
}

namespace {

using dev::Cell;
using dev::Interp;
using dev::QGrid;

constexpr int kBlock = 256;

struct V
{
    double x, y, z;
};
__device__ __forceinline__ V vadd(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V vsub(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V vmul(double s, V a) { return V{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V vcross(V a, V b)
{
    return V{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct Pose
{
    double R[9], t[3], M[9], m[3], v0[3], Om[3];
};

__device__ __forceinline__ Pose loadPose(const double* __restrict__ p)
{
    Pose P;
#pragma unroll
    for (int i = 0; i < 9; ++i) P.R[i] = p[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) P.t[i] = p[9 + i];
#pragma unroll
    for (int i = 0; i < 9; ++i) P.M[i] = p[12 + i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        P.m[i] = p[21 + i];
        P.v0[i] = p[24 + i];
        P.Om[i] = p[27 + i];
    }
    return P;
}

__device__ __forceinline__ V matvec(const double* A, V x)
{
    return V{A[0] * x.x + A[1] * x.y + A[2] * x.z, A[3] * x.x + A[4] * x.y + A[5] * x.z,
             A[6] * x.x + A[7] * x.y + A[8] * x.z};
}
__device__ __forceinline__ V matTvec(const double* A, V x)
{
    return V{A[0] * x.x + A[3] * x.y + A[6] * x.z, A[1] * x.x + A[4] * x.y + A[7] * x.z,
             A[2] * x.x + A[5] * x.y + A[8] * x.z};
}

struct Law
{
    double kn, c, mu, reg, margin;
    int order, mode;
    // 1/kn, 1/c, 1/mu (0 where the parameter is 0), for the backward's
    // parameter terms
    double ikn, ic, imu;
};

// The penalty, damping and friction forces are kept as their scalar factors
// (f_p = p G, f_d = cd G, f_t = s u_t) and rebuilt where the backward needs
// them, with the same operations as the forward, to keep the live state small.
struct Terms
{
    bool active, damp, fric;
    V Y, off, u, G, fq, ut;
    double g, p, dp, gn, D, cd, s;
    double hess[6];
    __device__ __forceinline__ V fp() const { return vmul(p, G); }
    __device__ __forceinline__ V fd() const { return damp ? vmul(cd, G) : V{0, 0, 0}; }
    __device__ __forceinline__ V ft() const { return fric ? vmul(s, ut) : V{0, 0, 0}; }
};

__device__ __forceinline__ double powi(double x, int k)
{
    double r = 1.0;
    for (int i = 0; i < k; ++i) r *= x;
    return r;
}

// Gap and raw gradient predicted from a cached query at a reference
// translation (contacts_sdf.py:307-327): g = g0 + grad0 . (t - t_ref).
struct LinCache
{
    const double* g0;
    const double* grad0;
    const uint8_t* valid0;
    double tref[3];
};

// How the contact kernels sample the field: kFieldAny handles every grid and
// both modes at run time; kFieldDenseCubic is Tricubic on a dense grid with
// cell records (one record read, no other path compiled in, which leaves the
// kernels fewer registers).  Both give the same bits for the same grid.
constexpr int kFieldAny = 0;
constexpr int kFieldDenseCubic = 1;

__host__ __device__ __forceinline__ int fieldKind(const QGrid& g, const Law& L)
{
    return (!g.sparse && g.cellRec != nullptr && L.mode == SDFO_TRICUBIC) ? kFieldDenseCubic
                                                                          : kFieldAny;
}

// The field at Y = R X + t after the narrow filter (contacts_sdf.py:91-98,
// 319): false where the point is filtered out or outside the field.
template <bool WithHess, int Kind = kFieldAny>
__device__ __forceinline__ bool sampleField(const QGrid& g, const Law& L, V Y, Interp<double>& it)
{
    if (L.margin > 0.0 && !(fabs(Y.z) < L.margin)) return false;
    const Cell c = dev::locateCell(g, Y.x, Y.y, Y.z);
    if (!c.valid) return false;
    if constexpr (Kind == kFieldDenseCubic)
    {
        const float4* r = g.cellRec + 8 * (int64_t(c.i0) + int64_t(g.n[0] - 1) *
                                                               (c.j0 + int64_t(g.n[1] - 1) * c.k0));
        float4 f[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) f[q] = r[q];
        dev::tricubicCorners<double, WithHess>(g, c, f, it);
        return true;
    }
    else
        return L.mode == SDFO_TRILINEAR ? dev::trilinear<double, WithHess>(g, c, it)
                                        : dev::tricubic<double, WithHess>(g, c, it);
}

// Forward terms of one (pose, point) pair (contacts_sdf.py:329-379); with
// Cached, the field comes from the linearisation cache instead of the grid.
template <bool WithHess>
__device__ __forceinline__ void lawTerms(Terms& t, const Pose& P, const Law& L,
                                         const Interp<double>& it);

template <bool WithHess, bool Cached = false, int Kind = kFieldAny>
__device__ __forceinline__ Terms pointTerms(const QGrid& g, const Pose& P, V X, const Law& L,
                                            const LinCache* lc = nullptr, int64_t i = 0)
{
    Terms t;
    t.active = false;
    t.Y = vadd(matvec(P.R, X), V{P.t[0], P.t[1], P.t[2]});
    t.off = vadd(matvec(P.M, X), V{P.m[0], P.m[1], P.m[2]});
    Interp<double> it;
    if (Cached)
    {
        if (L.margin > 0.0 && !(fabs(t.Y.z) < L.margin)) return t;
        if (!lc->valid0[i]) return t;
        it.grad = dev::V3T<double>{lc->grad0[3 * i], lc->grad0[3 * i + 1], lc->grad0[3 * i + 2]};
        const double dt[3] = {P.t[0] - lc->tref[0], P.t[1] - lc->tref[1], P.t[2] - lc->tref[2]};
        it.gap = lc->g0[i] + (it.grad.x * dt[0] + it.grad.y * dt[1] + it.grad.z * dt[2]);
#pragma unroll
        for (int e = 0; e < 6; ++e) it.hess[e] = 0.0;
    }
    else if (!sampleField<WithHess, Kind>(g, L, t.Y, it))
        return t;
    lawTerms<WithHess>(t, P, L, it);
    return t;
}

// The law of one point once the field (gap, raw gradient, Hessian) is known.
template <bool WithHess>
__device__ __forceinline__ void lawTerms(Terms& t, const Pose& P, const Law& L,
                                         const Interp<double>& it)
{
    t.g = it.gap;
    const double delta = t.g < 0.0 ? -t.g : 0.0;
    t.p = L.kn * powi(delta, L.order);
    if (!(t.p > 1e-30)) return; // contacts_sdf.py:329
    t.active = true;
    t.dp = (L.order >= 1 && t.g < 0.0) ? -L.kn * double(L.order) * powi(delta, L.order - 1) : 0.0;
    if (WithHess)
    {
#pragma unroll
        for (int e = 0; e < 6; ++e) t.hess[e] = it.hess[e];
    }
    t.G = V{it.grad.x, it.grad.y, it.grad.z};
    t.gn = sqrt(vdot(t.G, t.G));
    t.u = vadd(V{P.v0[0], P.v0[1], P.v0[2]}, vcross(V{P.Om[0], P.Om[1], P.Om[2]}, t.off));
    t.damp = L.c > 0.0 && t.g < 0.0; // contacts_sdf.py:355-362
    t.cd = t.damp ? L.c * (-vdot(t.u, t.G)) : 0.0;
    const double gnc = t.gn > 1e-30 ? t.gn : 1e-30; // contacts_sdf.py:369
    const double ig = 1.0 / gnc;
    const V un = V{t.G.x * ig, t.G.y * ig, t.G.z * ig};
    t.ut = vsub(t.u, vmul(vdot(t.u, un), un));
    t.D = sqrt(vdot(t.ut, t.ut) + L.reg * L.reg);
    const double phat = t.p * t.gn;
    t.fric = L.mu * phat > 0.0; // contacts_sdf.py:375-377
    t.s = t.fric ? -L.mu * phat / t.D : 0.0;
    t.fq = vadd(vadd(t.fp(), t.fd()), t.ft());
}

// Fixed-order LDS tree reduction of NV doubles per lane.
template <int NV>
__device__ __forceinline__ void blockReduce(double (&v)[NV], double* sh)
{
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k * kBlock + threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1)
    {
        if (int(threadIdx.x) < s)
        {
#pragma unroll
            for (int k = 0; k < NV; ++k)
                sh[k * kBlock + threadIdx.x] += sh[k * kBlock + threadIdx.x + s];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = sh[k * kBlock];
}

// Cache fill at one reference pose (contacts_sdf.py:319-327): the same
// point transform, narrow filter and interpolation as the forward.
__global__ void __launch_bounds__(kBlock)
contactLinearizeKernel(QGrid g, Law L, const double* __restrict__ X, int64_t n,
                       const double* __restrict__ pose, double* __restrict__ g0,
                       double* __restrict__ grad0, uint8_t* __restrict__ valid0)
{
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const Pose P = loadPose(pose);
    const V x = V{X[3 * i], X[3 * i + 1], X[3 * i + 2]};
    const V Y = vadd(matvec(P.R, x), V{P.t[0], P.t[1], P.t[2]});
    Interp<double> it;
    const bool ok = sampleField<false>(g, L, Y, it);
    g0[i] = ok ? it.gap : 0.0;
    grad0[3 * i] = ok ? it.grad.x : 0.0;
    grad0[3 * i + 1] = ok ? it.grad.y : 0.0;
    grad0[3 * i + 2] = ok ? it.grad.z : 0.0;
    valid0[i] = ok ? 1 : 0;
}

// The overlap test of AABB_BVH.query (bvh.py:66) on one leaf box.
__device__ __forceinline__ bool leafOverlaps(const double* __restrict__ lb, const double (&q)[6],
                                             double margin)
{
    bool hit = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) hit = hit && !(lb[3 + k] + margin < q[k]) && !(lb[k] - margin > q[3 + k]);
    return hit;
}

__device__ __forceinline__ void loadBox(const double* __restrict__ b, double (&q)[6])
{
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = b[k];
}

// grid = poses.  anyLeaf[pose] = 1 when some leaf overlaps the pose's box.
__global__ void __launch_bounds__(kBlock) cullAnyKernel(CullArgs cu, uint8_t* __restrict__ anyLeaf)
{
    double q[6];
    loadBox(cu.boxes + 6 * int64_t(blockIdx.x), q);
    int hit = 0;
    for (int64_t l = threadIdx.x; l < cu.nLeaves && !hit; l += kBlock)
        hit = leafOverlaps(cu.leafBox + 6 * l, q, cu.margin) ? 1 : 0;
    hit = __syncthreads_or(hit);
    if (threadIdx.x == 0) anyLeaf[blockIdx.x] = uint8_t(hit ? 1 : 0);
}

// grid = (point blocks, boxes).  mask[box * n + i]: point i's leaf overlaps.
__global__ void __launch_bounds__(kBlock)
cullMaskKernel(CullArgs cu, int64_t n, uint8_t* __restrict__ mask)
{
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    double q[6];
    loadBox(cu.boxes + 6 * int64_t(blockIdx.y), q);
    mask[int64_t(blockIdx.y) * n + i] = leafOverlaps(cu.leafBox + 6 * int64_t(cu.leafOf[i]), q, cu.margin) ? 1 : 0;
}

// The (pose, point) pairs in contact, by (chunk, pose) slot: slot
// pose * chunks + chunk lists its pairs' point indices at
// list[slot * chunkLen ...] and their number in count[slot]; act holds one
// flag per (pose, point) (pose-major, n per pose).  Written by the backward's
// screen, or by a forward that keeps them for its backward.
struct PairList
{
    int32_t* list = nullptr;
    int32_t* count = nullptr;
    uint8_t* act = nullptr;
};

// Ordered compaction of one pass of a block over kBlock points: the point
// indices of the hit lanes, in (wave, lane) order, appended at out[base...];
// returns the new base.  Every thread of the block calls it.
__device__ __forceinline__ int appendHits(bool hit, int64_t i, int32_t* __restrict__ out, int base,
                                          int* waveCount)
{
    const int wave = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
    const unsigned long long m = __ballot(hit);
    if (lane == 0) waveCount[wave] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int q = 0; q < wave; ++q) off += waveCount[q];
    if (hit) out[off + __popcll(m & ((1ull << lane) - 1ull))] = int32_t(i);
    for (int q = 0; q < kBlock / 64; ++q) base += waveCount[q];
    __syncthreads();
    return base;
}

// grid = (chunks, poses).  partial[(pose*chunks + chunk)*7 + k]: F, tau, active.
// With Culled, points outside the pose's broad-phase leaves are skipped
// unless no leaf overlaps (then every point takes part, contacts_sdf.py:270-271).
// With Save, the pose's pairs in contact are recorded in save too.
template <bool Cached, bool Culled, bool Save, int Kind = kFieldAny>
__device__ __forceinline__ void forwardBlock(const QGrid& g, const Law& L, const double* __restrict__ X,
                                             const double* __restrict__ w, int64_t n,
                                             const double* __restrict__ poses, int64_t chunkLen,
                                             double* __restrict__ partial, const LinCache& lc,
                                             const CullArgs& cu, const uint8_t* __restrict__ anyLeaf,
                                             const PairList& save, double* sh, int* waveCount)
{
    const int64_t pose = blockIdx.y;
    const int chunk = blockIdx.x;
    const Pose P = loadPose(poses + 30 * pose);
    const int64_t beg = int64_t(chunk) * chunkLen;
    const int64_t end = beg + chunkLen < n ? beg + chunkLen : n;
    double q[6];
    bool cull = false;
    if (Culled)
    {
        loadBox(cu.boxes + 6 * pose, q);
        cull = anyLeaf[pose] != 0;
    }
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    auto add = [&](const Terms& t, int64_t i) {
        const double wi = w[i];
        const V tq = vcross(t.off, t.fq);
        acc[0] += wi * t.fq.x;
        acc[1] += wi * t.fq.y;
        acc[2] += wi * t.fq.z;
        acc[3] += wi * tq.x;
        acc[4] += wi * tq.y;
        acc[5] += wi * tq.z;
        acc[6] += 1.0;
    };
    // one point per step: at 128-136 VGPRs (three or four waves per SIMD)
    // this beat requesting two points' records before interpolating either
    // (218 VGPRs, two waves), since a pose's points hit L2-resident records
    // the same points in the same order per thread; with Save each pair's
    // contact flag is recorded too (pairListKernel lists them)
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        if (Culled && cull && !leafOverlaps(cu.leafBox + 6 * int64_t(cu.leafOf[i]), q, cu.margin))
            continue;
        const V x = V{X[3 * i], X[3 * i + 1], X[3 * i + 2]};
        const Terms t = pointTerms<false, Cached, Kind>(g, P, x, L, &lc, i);
        if (Save) save.act[pose * n + i] = t.active ? 1 : 0;
        if (!t.active) continue;
        add(t, i);
    }
    blockReduce<7>(acc, sh);
    if (threadIdx.x == 0)
    {
        double* o = partial + (pose * gridDim.x + chunk) * 7;
#pragma unroll
        for (int k = 0; k < 7; ++k) o[k] = acc[k];
    }
}

// The forward kernels: W waves per SIMD, four by default (128 registers;
// +10% over three waves for the 1000-pose forward), three as the tuner's
// alternative (no spill); the sums are the same either way.
template <bool Cached, bool Culled, int Kind = kFieldAny, int W = 4>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W)))
contactForwardKernel(QGrid g, Law L, const double* __restrict__ X, const double* __restrict__ w,
                     int64_t n, const double* __restrict__ poses, int64_t chunkLen,
                     double* __restrict__ partial, LinCache lc, CullArgs cu,
                     const uint8_t* __restrict__ anyLeaf)
{
    __shared__ double sh[7 * kBlock];
    forwardBlock<Cached, Culled, false, Kind>(g, L, X, w, n, poses, chunkLen, partial, lc, cu,
                                              anyLeaf, PairList{}, sh, nullptr);
}

template <int Kind, int W = 4>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W)))
contactForwardSaveKernel(QGrid g, Law L, const double* __restrict__ X, const double* __restrict__ w,
                         int64_t n, const double* __restrict__ poses, int64_t chunkLen,
                         double* __restrict__ partial, PairList save)
{
    __shared__ double sh[7 * kBlock];
    forwardBlock<false, false, true, Kind>(g, L, X, w, n, poses, chunkLen, partial, LinCache{},
                                           CullArgs{}, nullptr, save, sh, nullptr);
}

__device__ __forceinline__ void finalizePose(int64_t pose, int chunks,
                                             const double* __restrict__ partial,
                                             double* __restrict__ out, int32_t* __restrict__ active);

// After a recording forward, one launch: blocks [0, chunks * nb) compact
// slot b's contact flags, in point order, into list[b * chunkLen ...] and
// count[b] (slot b = pose * chunks + chunk); the rest finalise the forces of
// kBlock poses each, as contactForwardFinalize.
__global__ void __launch_bounds__(kBlock)
pairListKernel(int64_t n, int64_t nb, int chunks, int64_t chunkLen, PairList pl,
               const double* __restrict__ partial, double* __restrict__ outF,
               int32_t* __restrict__ active)
{
    __shared__ int waveCount[kBlock / 64];
    const int64_t slot = blockIdx.x;
    if (slot >= nb * chunks)
    {
        const int64_t pose = (slot - nb * chunks) * kBlock + threadIdx.x;
        if (pose < nb) finalizePose(pose, chunks, partial, outF, active);
        return;
    }
    const int64_t pose = slot / chunks;
    const int64_t beg = int64_t(slot - pose * chunks) * chunkLen;
    const int64_t end = beg + chunkLen < n ? beg + chunkLen : n;
    int32_t* out = pl.list + slot * chunkLen;
    int base = 0;
    for (int64_t i0 = beg; i0 < end; i0 += kBlock)
    {
        const int64_t i = i0 + threadIdx.x;
        const bool hit = i < end && pl.act[pose * n + i] != 0;
        base = appendHits(hit, i, out, base, waveCount);
    }
    if (threadIdx.x == 0) pl.count[slot] = base;
}

// Pose `pose`'s forces: its chunk partials summed in chunk order.
__device__ __forceinline__ void finalizePose(int64_t pose, int chunks,
                                             const double* __restrict__ partial,
                                             double* __restrict__ out, int32_t* __restrict__ active)
{
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < chunks; ++c)
    {
        const double* p = partial + (pose * chunks + c) * 7;
#pragma unroll
        for (int k = 0; k < 7; ++k) s[k] += p[k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) out[6 * pose + k] = s[k];
    if (active) active[pose] = int32_t(s[6]);
}

__global__ void contactForwardFinalize(int64_t nposes, int chunks, const double* __restrict__ partial,
                                       double* __restrict__ out, int32_t* __restrict__ active)
{
    const int64_t pose = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (pose < nposes) finalizePose(pose, chunks, partial, out, active);
}

// Backward per (pose, point), given a = w (gF + gT x off) -- the derivation
// is stated next to the CPU restatement; the same formulas are used here.
struct Back
{
    V gY, gOff;
    double skn, sc, smu;
};

__device__ __forceinline__ Back pointBackward(const Terms& t, const Law& L, V a, V gT, double wi,
                                              V Om)
{
    Back b;
    const V G = t.G;
    const double gn = t.gn;
    const double aG = vdot(a, G);
    double sg = t.dp * aG;
    if (t.fric && t.p != 0.0) sg += (t.dp / t.p) * vdot(a, t.ft());
    const double ig = 1.0 / gn, iD = 1.0 / t.D;
    V sG = vmul(t.p, a);
    V su = V{0, 0, 0};
    const double uG = vdot(t.u, G);
    if (t.damp)
    {
        sG = vadd(sG, vmul(-L.c, vadd(vmul(aG, t.u), vmul(uG, a))));
        su = vadd(su, vmul(-L.c * aG, G));
    }
    b.smu = 0.0;
    if (t.fric)
    {
        const double k = -L.mu * t.p;
        const V h = vmul(gn, t.ut);
        const V ahG = vadd(vsub(vsub(vmul(vdot(a, t.u) * ig, G), vmul(aG * ig, t.u)), vmul(uG * ig, a)),
                           vmul(uG * aG * (ig * ig * ig), G));
        const V dD = vmul(-uG * (ig * ig * iD), t.ut);
        sG = vadd(sG, vmul(k, vsub(vmul(iD, ahG), vmul(vdot(a, h) * (iD * iD), dD))));
        const V Pa = vsub(a, vmul(aG * (ig * ig), G));
        su = vadd(su, vmul(k * gn, vsub(vmul(iD, Pa), vmul(vdot(a, t.ut) * (iD * iD * iD), t.ut))));
        b.smu = vdot(a, t.ft()) * L.imu;
    }
    const double* h6 = t.hess;
    const V HsG = V{h6[0] * sG.x + h6[1] * sG.y + h6[2] * sG.z,
                    h6[1] * sG.x + h6[3] * sG.y + h6[4] * sG.z,
                    h6[2] * sG.x + h6[4] * sG.y + h6[5] * sG.z};
    b.gY = vadd(vmul(sg, G), HsG);
    b.gOff = vadd(vmul(wi, vcross(t.fq, gT)), vcross(su, Om));
    b.skn = L.kn != 0.0 ? vdot(a, vadd(t.fp(), t.ft())) * L.ikn : 0.0;
    b.sc = t.damp ? vdot(a, t.fd()) * L.ic : 0.0;
    return b;
}

// Backward, first pass: which (pose, point) pairs can be in contact.  Only a
// small share of the pairs is (about 9% for a ring resting on a cube), and
// in point order they are scattered over most waves, so the backward proper
// runs on compacted lists instead of every pair.  A pair is listed when its
// field sample exists and its gap, computed with the value part of the same
// interpolation after the same transform, narrow filter and cell, is below
// kScreenGap (or for every sampled pair when the law's order is 0, where
// p = kn for any gap); the second pass re-evaluates each listed pair exactly
// and drops the ones whose law is not active, so the screen's tolerance only
// has to admit every active pair.  Block (chunk, pose) writes its slot of
// the group's PairList (slots and flags counted from pose poseBase).
constexpr double kScreenGap = 1e-9;

template <bool Trilinear, int Kind = kFieldAny>
__global__ void __launch_bounds__(kBlock)
contactScreenKernel(QGrid g, Law L, const double* __restrict__ X, int64_t n,
                    const double* __restrict__ poses, int64_t chunkLen, int64_t poseBase,
                    PairList pl)
{
    __shared__ int waveCount[kBlock / 64];
    const int64_t lp = blockIdx.y;
    const double* pp = poses + 30 * (poseBase + lp);
    double R[9], tv[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = pp[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) tv[k] = pp[9 + k];
    const int64_t slot = lp * gridDim.x + blockIdx.x;
    const int64_t beg = int64_t(blockIdx.x) * chunkLen;
    const int64_t end = beg + chunkLen < n ? beg + chunkLen : n;
    int32_t* out = pl.list + slot * chunkLen;
    int base = 0;
    for (int64_t i0 = beg; i0 < end; i0 += kBlock)
    {
        const int64_t i = i0 + threadIdx.x;
        bool hit = false;
        if (i < end)
        {
            const V x = V{X[3 * i], X[3 * i + 1], X[3 * i + 2]};
            const V Y = vadd(matvec(R, x), V{tv[0], tv[1], tv[2]});
            if (!(L.margin > 0.0 && !(fabs(Y.z) < L.margin)))
            {
                const Cell c = dev::locateCell(g, Y.x, Y.y, Y.z);
                if (c.valid)
                {
                    double gap = 0.0;
                    bool ok;
                    if (Trilinear)
                    {
                        Interp<double> it;
                        ok = dev::trilinear<double, false>(g, c, it);
                        gap = it.gap;
                    }
                    else if constexpr (Kind == kFieldDenseCubic)
                    {
                        const float4* r =
                            g.cellRec + 8 * (int64_t(c.i0) +
                                             int64_t(g.n[0] - 1) * (c.j0 + int64_t(g.n[1] - 1) * c.k0));
                        float4 f[8];
#pragma unroll
                        for (int q = 0; q < 8; ++q) f[q] = r[q];
                        ok = true;
                        gap = dev::tricubicValueCorners<double>(g, c, f);
                    }
                    else
                    {
                        float4 f[8];
                        ok = dev::cellCorners(g, c, f);
                        if (ok) gap = dev::tricubicValueCorners<double>(g, c, f);
                    }
                    hit = ok && (L.order <= 0 ? L.kn > 1e-30 : gap < kScreenGap);
                }
            }
            if (pl.act) pl.act[lp * n + i] = hit ? 1 : 0;
        }
        base = appendHits(hit, i, out, base, waveCount);
    }
    if (threadIdx.x == 0) pl.count[slot] = base;
}

// Backward, second pass, in pieces of kPiece listed pairs: piece p of slot
// s covers entries [p kPiece, (p+1) kPiece) of the slot's list, and item code
// s * maxPieces + p names it.  pieceTableKernel (one block) lays the items of
// all slots out in slot order, so the pairs in contact are spread evenly over
// the waves of the next kernel whatever their distribution over poses, and
// the items of pose pl of the group are [poseStart[pl], poseStart[pl + 1]).
constexpr int kPiece = 64;
constexpr int kTableBlock = 1024;

__global__ void __launch_bounds__(kTableBlock)
pieceTableKernel(int64_t slots, int chunks, int maxPieces, const int32_t* __restrict__ count,
                 int32_t* __restrict__ items, int32_t* __restrict__ poseStart,
                 int32_t* __restrict__ total)
{
    __shared__ int sh[kTableBlock];
    const int tid = int(threadIdx.x);
    const int64_t per = (slots + kTableBlock - 1) / kTableBlock;
    const int64_t s0 = tid * per < slots ? tid * per : slots;
    const int64_t s1 = s0 + per < slots ? s0 + per : slots;
    int mine = 0;
    for (int64_t sl = s0; sl < s1; ++sl) mine += (count[sl] + kPiece - 1) / kPiece;
    sh[tid] = mine;
    __syncthreads();
    // inclusive scan of the per-thread piece counts (Hillis-Steele)
    for (int d = 1; d < kTableBlock; d <<= 1)
    {
        const int v = tid >= d ? sh[tid - d] : 0;
        __syncthreads();
        sh[tid] += v;
        __syncthreads();
    }
    int off = sh[tid] - mine;
    for (int64_t sl = s0; sl < s1; ++sl)
    {
        if (sl % chunks == 0) poseStart[sl / chunks] = off;
        const int np = (count[sl] + kPiece - 1) / kPiece;
        for (int p = 0; p < np; ++p) items[off++] = int32_t(sl * maxPieces + p);
    }
    if (tid == kTableBlock - 1)
    {
        *total = sh[tid];
        poseStart[slots / chunks] = sh[tid];
    }
}

// One wave per item, the waves looping over the items: lane l takes entry
// p kPiece + l of the piece, with the same per-pair arithmetic as the
// forward's terms and the derivation stated next to the CPU restatement.  A
// listed pair whose law is inactive writes a zero point partial.  The
// piece's nine sums are added by a fixed butterfly into pieceSum[k * 9 + q]
// for item k, so results are bit-reproducible run to run.
template <int Kind>
__global__ void __launch_bounds__(64)
contactBackwardPieceKernel(QGrid g, Law L, const double* __restrict__ X,
                           const double* __restrict__ w, int64_t n,
                           const double* __restrict__ poses, const double* __restrict__ gOut,
                           int chunks, int64_t chunkLen, int maxPieces, int64_t poseBase,
                           const int32_t* __restrict__ list, const int32_t* __restrict__ count,
                           const int32_t* __restrict__ items, const int32_t* __restrict__ total,
                           double* __restrict__ pieceSum, double4* __restrict__ pointPartial)
{
    const int nItems = *total;
    for (int k = blockIdx.x; k < nItems; k += gridDim.x)
    {
        const int code = items[k];
        const int64_t slot = code / maxPieces;
        const int piece = code - int(slot) * maxPieces;
        const int64_t pl = slot / chunks;
        const int64_t pose = poseBase + pl;
        const int j = piece * kPiece + int(threadIdx.x);
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (j < count[slot])
        {
            const int64_t i = list[slot * chunkLen + j];
            const Pose P = loadPose(poses + 30 * pose);
            const V gF = V{gOut[6 * pose], gOut[6 * pose + 1], gOut[6 * pose + 2]};
            const V gT = V{gOut[6 * pose + 3], gOut[6 * pose + 4], gOut[6 * pose + 5]};
            double4* pp = pointPartial ? pointPartial + pl * n : nullptr;
            const V x = V{X[3 * i], X[3 * i + 1], X[3 * i + 2]};
            const Terms t = pointTerms<true, false, Kind>(g, P, x, L);
            if (!t.active)
            {
                if (pp) pp[i] = make_double4(0.0, 0.0, 0.0, 0.0);
            }
            else
            {
                const double wi = w[i];
                const V a = vmul(wi, vadd(gF, vcross(gT, t.off)));
                const Back b = pointBackward(t, L, a, gT, wi, V{P.Om[0], P.Om[1], P.Om[2]});
                const V gw = vadd(vcross(vsub(t.Y, V{P.t[0], P.t[1], P.t[2]}), b.gY),
                                  vcross(t.off, b.gOff));
                acc[0] = gw.x;
                acc[1] = gw.y;
                acc[2] = gw.z;
                acc[3] = b.gY.x;
                acc[4] = b.gY.y;
                acc[5] = b.gY.z;
                acc[6] = b.skn;
                acc[7] = b.sc;
                acc[8] = b.smu;
                if (pp)
                {
                    const V gx = vadd(matTvec(P.R, b.gY), matTvec(P.M, b.gOff));
                    pp[i] = make_double4(gx.x, gx.y, gx.z,
                                         vdot(gF, t.fq) + vdot(gT, vcross(t.off, t.fq)));
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 9; ++q)
        {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
        }
        if (threadIdx.x == 0)
        {
            double* o = pieceSum + int64_t(k) * 9;
#pragma unroll
            for (int q = 0; q < 9; ++q) o[q] = acc[q];
        }
    }
}

// Point-gradient reduction over the poses of one pose group, in a fixed
// order: a block holds kReducePoints points x kReduceSlices pose slices; each
// thread sums its slice's poses for its point in pose order, then the slices
// are added in slice order through LDS and the result is added to (or, for
// the first pose group, stored in) gX / gW.  Only the pairs the screen
// listed (flag act) have a partial.  Consecutive lanes read consecutive
// points of one pose, so every load is coalesced.
constexpr int kReducePoints = 16;
constexpr int kReduceSlices = 16;
constexpr int kReduceBatch = 8;

__device__ __forceinline__ void pointReduceBlock(int64_t blk, int64_t n, int poses,
                                                 const double4* __restrict__ pointPartial,
                                                 const uint8_t* __restrict__ act, int accumulate,
                                                 double* __restrict__ gX, double* __restrict__ gW,
                                                 double4 (&sh)[kReduceSlices][kReducePoints])
{
    const int lane = int(threadIdx.x) % kReducePoints;
    const int slice = int(threadIdx.x) / kReducePoints;
    const int64_t i = blk * kReducePoints + lane;
    const int per = (poses + kReduceSlices - 1) / kReduceSlices;
    const int p0 = slice * per;
    const int p1 = p0 + per < poses ? p0 + per : poses;
    double sx = 0.0, sy = 0.0, sz = 0.0, sw = 0.0;
    if (i < n)
    {
        // kReduceBatch poses at a time: every flag is loaded, then every
        // listed partial (pairs the screen did not list have none), then
        // they are added in pose order
        for (int pb = p0; pb < p1; pb += kReduceBatch)
        {
            bool on[kReduceBatch];
#pragma unroll
            for (int q = 0; q < kReduceBatch; ++q)
                on[q] = pb + q < p1 && act[int64_t(pb + q) * n + i] != 0;
            double4 v[kReduceBatch];
#pragma unroll
            for (int q = 0; q < kReduceBatch; ++q)
                v[q] = on[q] ? pointPartial[int64_t(pb + q) * n + i] : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
            for (int q = 0; q < kReduceBatch; ++q)
            {
                if (!on[q]) continue;
                sx += v[q].x;
                sy += v[q].y;
                sz += v[q].z;
                sw += v[q].w;
            }
        }
    }
    sh[slice][lane] = make_double4(sx, sy, sz, sw);
    __syncthreads();
    if (slice != 0 || i >= n) return;
    sx = 0.0, sy = 0.0, sz = 0.0, sw = 0.0;
#pragma unroll
    for (int q = 0; q < kReduceSlices; ++q)
    {
        const double4 v = sh[q][lane];
        sx += v.x;
        sy += v.y;
        sz += v.z;
        sw += v.w;
    }
    if (gX)
    {
        if (accumulate)
        {
            sx = gX[3 * i] + sx;
            sy = gX[3 * i + 1] + sy;
            sz = gX[3 * i + 2] + sz;
        }
        gX[3 * i] = sx;
        gX[3 * i + 1] = sy;
        gX[3 * i + 2] = sz;
    }
    if (gW) gW[i] = accumulate ? gW[i] + sw : sw;
}

// One wave finalises pose pl of a group: its items' nine sums, read seven
// items (63 consecutive doubles) at a time and handed to lanes 0-8 by
// shuffles, are added in item order into d/d(omega, dt) (gTwist) and the
// (kn, c, mu) terms (perPose3).  Every lane of the wave calls it.
__device__ __forceinline__ void poseFinalizeWave(int64_t pl, int64_t poseBase,
                                                 const int32_t* __restrict__ poseStart,
                                                 const double* __restrict__ pieceSum,
                                                 double* __restrict__ gTwist,
                                                 double* __restrict__ perPose3)
{
    const int lane = int(threadIdx.x) & 63;
    const int k0 = poseStart[pl], k1 = poseStart[pl + 1];
    const int q = lane < 9 ? lane : 0;
    double s = 0.0;
    for (int kb = k0; kb < k1; kb += 7)
    {
        const double v = lane < 63 && kb + lane / 9 < k1 ? pieceSum[int64_t(kb) * 9 + lane] : 0.0;
        const int m = k1 - kb < 7 ? k1 - kb : 7;
#pragma unroll
        for (int j = 0; j < 7; ++j)
        {
            const double x = __shfl(v, j * 9 + q);
            if (lane < 9 && j < m) s += x;
        }
    }
    const int64_t pose = poseBase + pl;
    if (lane < 6)
    {
        if (gTwist) gTwist[6 * pose + lane] = s;
    }
    else if (lane < 9 && perPose3)
        perPose3[3 * pose + lane - 6] = s;
}

// The (kn, c, mu) gradient of a single-group call straight from the items:
// thread t adds items t, t + blockDim, ... in order, then the block's fixed
// LDS tree.
__device__ __forceinline__ void paramsFromItems(const int32_t* __restrict__ total,
                                                const double* __restrict__ pieceSum,
                                                double* __restrict__ gParams, double* sh)
{
    double acc[3] = {0, 0, 0};
    const int nItems = *total;
    for (int k = threadIdx.x; k < nItems; k += kBlock)
    {
        acc[0] += pieceSum[int64_t(k) * 9 + 6];
        acc[1] += pieceSum[int64_t(k) * 9 + 7];
        acc[2] += pieceSum[int64_t(k) * 9 + 8];
    }
    blockReduce<3>(acc, sh);
    if (threadIdx.x == 0)
    {
        gParams[0] = acc[0];
        gParams[1] = acc[1];
        gParams[2] = acc[2];
    }
}

// The backward's tail in one launch, all on one pose group: blocks
// [0, nReduce) reduce the point gradients, the next ceil(cnt / 4) blocks
// finalise four poses each (one per wave), and with gParams the last block
// sums the parameter terms.
static_assert(kReducePoints * kReduceSlices == kBlock, "tail blocks share one size");

__global__ void __launch_bounds__(kBlock)
backwardTailKernel(int64_t nReduce, int64_t n, int cnt, const double4* __restrict__ pointPartial,
                   const uint8_t* __restrict__ act, int accumulate, double* __restrict__ gX,
                   double* __restrict__ gW, int64_t poseBase, const int32_t* __restrict__ poseStart,
                   const double* __restrict__ pieceSum, const int32_t* __restrict__ total,
                   double* __restrict__ gTwist, double* __restrict__ perPose3,
                   double* __restrict__ gParams)
{
    __shared__ double4 shR[kReduceSlices][kReducePoints];
    __shared__ double shS[3 * kBlock];
    const int64_t b = blockIdx.x;
    if (b < nReduce)
    {
        pointReduceBlock(b, n, cnt, pointPartial, act, accumulate, gX, gW, shR);
        return;
    }
    const int64_t poseBlocks = (cnt + kBlock / 64 - 1) / (kBlock / 64);
    if (b < nReduce + poseBlocks)
    {
        const int wave = int(threadIdx.x) >> 6;
        const int64_t pl = (b - nReduce) * (kBlock / 64) + wave;
        if (pl < cnt) poseFinalizeWave(pl, poseBase, poseStart, pieceSum, gTwist, perPose3);
        return;
    }
    paramsFromItems(total, pieceSum, gParams, shS);
}

// Sum of the per-pose parameter terms in pose order (one block).
__global__ void paramReduceKernel(int64_t nposes, const double* __restrict__ perPose3,
                                  double* __restrict__ gParams)
{
    __shared__ double sh[3 * kBlock];
    double acc[3] = {0, 0, 0};
    for (int64_t p = threadIdx.x; p < nposes; p += kBlock)
    {
        acc[0] += perPose3[3 * p];
        acc[1] += perPose3[3 * p + 1];
        acc[2] += perPose3[3 * p + 2];
    }
    blockReduce<3>(acc, sh);
    if (threadIdx.x == 0)
    {
        gParams[0] = acc[0];
        gParams[1] = acc[1];
        gParams[2] = acc[2];
    }
}

// ---- SDFContactEngine.evaluate in the world frame (contacts_sdf.py:225-385)
// Per-pair record laid out by evaluatePrepKernel:
//   R_sdf(9) t_sdf(3)  quadrature point X -> SDF frame, Y = R_sdf X + t_sdf
//   RAm(9) rAm(3)      quadrature frame in the world, x_w = rAm + RAm X
//   rA vA wA (3 each), rB(3) RB(9) vB wB (3 each), on (1: the pair passed
//   the body-level check of _check_broad_phase)
constexpr int kEvalRec = 52;

struct EvalPrep
{
    const double* states; // pairs x 36: rA RA vA wA rB RB vB wB
    const double* frames; // pairs x 24: R_ab t_ab RAm rAm, or null
    double bmin[3], bmax[3];
    double cullMargin;    // max(narrow margin, 0.05), contacts_sdf.py:243
    int bodyCheck;
    double centerA[3], halfA[3], halfB[3];
};

// One thread per pair: the transforms of contacts_sdf.py:275-283, the cull
// box of :247-258 and the body-level check of :68-89.
__global__ void evaluatePrepKernel(EvalPrep in, int64_t nb, double* __restrict__ rec,
                                   double* __restrict__ boxes)
{
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const double* s = in.states + 36 * b;
    double* r = rec + kEvalRec * b;
    const V rA = V{s[0], s[1], s[2]};
    const double* RA = s + 3;
    const V rB = V{s[18], s[19], s[20]};
    const double* RB = s + 21;
    double lo[3], hi[3];
    if (in.frames)
    {
        const double* f = in.frames + 24 * b;
#pragma unroll
        for (int k = 0; k < 24; ++k) r[k] = f[k];
        // R_ba = R_ab^T, t_ba = -R_ab^T t_ab
        const V tba = vmul(-1.0, matTvec(f, V{f[9], f[10], f[11]}));
        const V l = vadd(matTvec(f, V{in.bmin[0], in.bmin[1], in.bmin[2]}), tba);
        const V h = vadd(matTvec(f, V{in.bmax[0], in.bmax[1], in.bmax[2]}), tba);
        lo[0] = l.x; lo[1] = l.y; lo[2] = l.z;
        hi[0] = h.x; hi[1] = h.y; hi[2] = h.z;
    }
    else
    {
        // R_sdf = RB^T RA, t_sdf = RB^T (rA - rB), marker = body frame
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                r[3 * i + j] = RB[i] * RA[j] + RB[3 + i] * RA[3 + j] + RB[6 + i] * RA[6 + j];
        const V ts = matTvec(RB, vsub(rA, rB));
        r[9] = ts.x; r[10] = ts.y; r[11] = ts.z;
#pragma unroll
        for (int k = 0; k < 9; ++k) r[12 + k] = RA[k];
        r[21] = rA.x; r[22] = rA.y; r[23] = rA.z;
        const V l = matTvec(RA, vsub(V{in.bmin[0], in.bmin[1], in.bmin[2]}, rA));
        const V h = matTvec(RA, vsub(V{in.bmax[0], in.bmax[1], in.bmax[2]}, rA));
        lo[0] = l.x; lo[1] = l.y; lo[2] = l.z;
        hi[0] = h.x; hi[1] = h.y; hi[2] = h.z;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        boxes[6 * b + k] = lo[k] - in.cullMargin;
        boxes[6 * b + 3 + k] = hi[k] + in.cullMargin;
    }
    // rA vA wA
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        r[24 + k] = s[k];
        r[27 + k] = s[12 + k];
        r[30 + k] = s[15 + k];
        r[33 + k] = s[18 + k];
        r[45 + k] = s[30 + k];
        r[48 + k] = s[33 + k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) r[36 + k] = RB[k];
    int on = 1;
    if (in.bodyCheck)
    {
        const V c = vadd(rA, matvec(RA, V{in.centerA[0], in.centerA[1], in.centerA[2]}));
        const double d[3] = {c.x - rB.x, c.y - rB.y, c.z - rB.z};
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            const double h = fabs(RA[3 * i]) * in.halfA[0] + fabs(RA[3 * i + 1]) * in.halfA[1] +
                             fabs(RA[3 * i + 2]) * in.halfA[2];
            if (!(fabs(d[i]) <= h + in.halfB[i])) on = 0;
        }
    }
    r[51] = double(on);
}

// grid = (chunks, pairs).  partial[(pair*chunks + chunk)*7 + k]: F_A, tau_A
// (world frame) and the active count.
template <bool Culled, int Kind = kFieldAny, int W = 0>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W > 0 ? W : 1, W > 0 ? W : 8)))
contactEvaluateKernel(QGrid g, Law L, const double* __restrict__ X, const double* __restrict__ w,
                      int64_t n, const double* __restrict__ rec, int64_t chunkLen,
                      double* __restrict__ partial, CullArgs cu, const uint8_t* __restrict__ anyLeaf)
{
    __shared__ double sh[7 * kBlock];
    const int64_t pair = blockIdx.y;
    const int chunk = blockIdx.x;
    const double* r = rec + kEvalRec * pair;
    double Rs[9], RAm[9], RB[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
    {
        Rs[k] = r[k];
        RAm[k] = r[12 + k];
        RB[k] = r[36 + k];
    }
    const V ts = V{r[9], r[10], r[11]}, rAm = V{r[21], r[22], r[23]};
    const V rA = V{r[24], r[25], r[26]}, vA = V{r[27], r[28], r[29]}, wA = V{r[30], r[31], r[32]};
    const V rB = V{r[33], r[34], r[35]}, vB = V{r[45], r[46], r[47]}, wB = V{r[48], r[49], r[50]};
    const bool on = r[51] != 0.0;
    const int64_t beg = int64_t(chunk) * chunkLen;
    const int64_t end = beg + chunkLen < n ? beg + chunkLen : n;
    double q[6];
    bool cull = false;
    if (Culled)
    {
        loadBox(cu.boxes + 6 * pair, q);
        cull = anyLeaf[pair] != 0;
    }
    Law Lq = L;
    Lq.margin = 0.0; // the narrow filter is applied below, always (RB is given)
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = beg + threadIdx.x; on && i < end; i += kBlock)
    {
        if (Culled && cull && !leafOverlaps(cu.leafBox + 6 * int64_t(cu.leafOf[i]), q, cu.margin))
            continue;
        const V x = V{X[3 * i], X[3 * i + 1], X[3 * i + 2]};
        const V Y = vadd(matvec(Rs, x), ts);
        if (!(fabs(Y.z) < L.margin)) continue; // _narrow_filter, contacts_sdf.py:91-98
        Interp<double> it;
        if (!sampleField<false, Kind>(g, Lq, Y, it)) continue;
        const double p = L.kn * powi(it.gap < 0.0 ? -it.gap : 0.0, L.order);
        if (!(p > 1e-30)) continue; // contacts_sdf.py:329
        const V gl = V{it.grad.x, it.grad.y, it.grad.z};
        const V G = matvec(RB, gl);           // rg = RB rg_local (:342)
        const double gn = sqrt(vdot(gl, gl)); // grad_norm of the local gradient
        const V xw = vadd(rAm, matvec(RAm, x));
        const V off = vsub(xw, rA);
        const V u = vsub(vadd(vA, vcross(wA, off)), vadd(vB, vcross(wB, vsub(xw, rB))));
        V fq = vmul(p, G);
        if (L.c > 0.0 && it.gap < 0.0) fq = vadd(fq, vmul(L.c * (-vdot(u, G)), G)); // :355-362
        const double gnc = gn > 1e-30 ? gn : 1e-30;
        const double ig = 1.0 / gnc;
        const V un = V{G.x * ig, G.y * ig, G.z * ig};
        const V ut = vsub(u, vmul(vdot(u, un), un));
        const double utn = sqrt(vdot(ut, ut));
        const double D = sqrt(utn * utn + L.reg * L.reg);
        const double phat = p * gn;
        if (L.mu * phat > 0.0) // :374-377
        {
            const double k = -L.mu * phat * (1.0 / D);
            fq = vadd(fq, V{k * ut.x, k * ut.y, k * ut.z});
        }
        const double wi = w[i];
        const V tq = vcross(off, fq);
        acc[0] += wi * fq.x;
        acc[1] += wi * fq.y;
        acc[2] += wi * fq.z;
        acc[3] += wi * tq.x;
        acc[4] += wi * tq.y;
        acc[5] += wi * tq.z;
        acc[6] += 1.0;
    }
    blockReduce<7>(acc, sh);
    if (threadIdx.x == 0)
    {
        double* o = partial + (pair * gridDim.x + chunk) * 7;
#pragma unroll
        for (int k = 0; k < 7; ++k) o[k] = acc[k];
    }
}

// F_A, tau_A from the chunk partials in chunk order; F_B = -F_A,
// tau_B = -(tau_A + (rA - rB) x F_A) (contacts_sdf.py:383-385).
__global__ void contactEvaluateFinalize(int64_t npairs, int chunks, const double* __restrict__ partial,
                                        const double* __restrict__ rec, double* __restrict__ out,
                                        int32_t* __restrict__ active)
{
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= npairs) return;
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < chunks; ++c)
    {
        const double* p = partial + (b * chunks + c) * 7;
#pragma unroll
        for (int k = 0; k < 7; ++k) s[k] += p[k];
    }
    const double* r = rec + kEvalRec * b;
    const bool on = r[51] != 0.0;
    const V F = V{s[0], s[1], s[2]}, T = V{s[3], s[4], s[5]};
    const V dr = V{r[24] - r[33], r[25] - r[34], r[26] - r[35]};
    const V TB = vadd(T, vcross(dr, F));
    const double o[12] = {F.x, F.y, F.z, T.x, T.y, T.z, -F.x, -F.y, -F.z, -TB.x, -TB.y, -TB.z};
#pragma unroll
    for (int k = 0; k < 12; ++k) out[12 * b + k] = on ? o[k] : 0.0;
    if (active) active[b] = on ? int32_t(s[6]) : 0;
}

// Compute units of the current device, looked up once per device.
int computeUnits()
{
    static std::atomic<int> cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int c = cached[dev].load(std::memory_order_relaxed);
    if (c > 0) return c;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c < 1)
        c = 256;
    cached[dev].store(c, std::memory_order_relaxed);
    return c;
}

Law makeLaw(const sdfo_contact_params& p)
{
    return Law{p.kn, p.damping, p.mu, p.regularizer, p.narrow_margin, p.order, p.mode,
               p.kn != 0.0 ? 1.0 / p.kn : 0.0, p.damping != 0.0 ? 1.0 / p.damping : 0.0,
               p.mu != 0.0 ? 1.0 / p.mu : 0.0};
}

// Split each pose's points into chunks so that the launch holds at least
// ~2048 workgroups (8 per CU) when there are few poses.
int chunksFor(int64_t n, int64_t nposes)
{
    int64_t want = (2048 + nposes - 1) / nposes;
    const int64_t maxc = (n + kBlock - 1) / kBlock;
    if (want > maxc) want = maxc;
    if (want < 1) want = 1;
    if (want > 65535) want = 65535;
    return int(want);
}

struct Scratch
{
    void* p = nullptr;
    hipStream_t s;
    explicit Scratch(size_t bytes, hipStream_t st) : s(st)
    {
        if (bytes > 0) SDFO_HIP(hipMallocAsync(&p, bytes, s));
    }
    ~Scratch()
    {
        if (p) (void)hipFreeAsync(p, s);
    }
};

} // namespace

namespace {

// The backward's buffers for n points and nb poses.  Point gradients need
// one partial per (pose, point) of a pose group; groups are bounded so that
// the per-(pose, point) arrays stay under groupBytes.  lists / counts / flags
// cover the group (screening backward) or every pose (a forward's record);
// the piece table, piece sums, pose starts and per-pose terms cover one
// group, the forward partials every pose.
struct ContactLayout
{
    int chunks = 1, maxPieces = 1;
    int64_t chunkLen = 0, group = 0;
    size_t oCount = 0, oList = 0, oAct = 0, oPart = 0, oItems = 0, oStart = 0, oTotal = 0,
           oSum = 0, oPer = 0, oPts = 0, bytes = 0;
};

ContactLayout contactLayout(int64_t n, int64_t nb, bool wantPoint, bool record)
{
    ContactLayout o;
    o.chunks = n > 0 ? chunksFor(n, nb) : 1;
    o.chunkLen = n > 0 ? (n + o.chunks - 1) / o.chunks : 0;
    o.maxPieces = int(std::max<int64_t>(1, (o.chunkLen + kPiece - 1) / kPiece));
    // a record keeps its point partials until the backward: 64 MiB of them
    // per group (its lists and flags cover every pose anyway); a screening
    // backward allocates lists, flags and partials per group for the call:
    // 256 MiB
    const int64_t groupBytes = int64_t(record ? 64 : 256) << 20;
    o.group = nb;
    if (n > 0)
    {
        const int64_t perPair =
            record ? int64_t(sizeof(double4))
                   : int64_t(sizeof(int32_t)) + (wantPoint ? int64_t(sizeof(double4)) + 1 : 0);
        const int64_t cap = groupBytes / (n * perPair);
        o.group = cap < 1 ? 1 : (cap < nb ? cap : nb);
    }
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const int64_t listPoses = record ? nb : o.group;
    const size_t listSlots = size_t(listPoses) * size_t(o.chunks);
    const size_t pieces = size_t(o.group) * size_t(o.chunks) * size_t(o.maxPieces);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t r = off;
        off += al(bytes);
        return r;
    };
    o.oCount = take(sizeof(int32_t) * listSlots);
    o.oList = take(sizeof(int32_t) * listSlots * size_t(o.chunkLen));
    o.oAct = take(wantPoint ? size_t(listPoses) * size_t(n > 0 ? n : 0) : 0);
    o.oPart = take(record ? sizeof(double) * 7 * size_t(nb) * size_t(o.chunks) : 0);
    o.oItems = take(sizeof(int32_t) * pieces);
    o.oStart = take(sizeof(int32_t) * (size_t(o.group) + 1));
    o.oTotal = take(sizeof(int32_t));
    o.oSum = take(sizeof(double) * 9 * pieces);
    o.oPer = take(sizeof(double) * 3 * size_t(nb));
    o.oPts = take(wantPoint ? sizeof(double4) * size_t(o.group) * size_t(n) : 0);
    o.bytes = off;
    return o;
}

// The record of contactForwardSaveKernel for its backward: the pairs in
// contact of every pose plus every buffer the backward uses, so that neither
// pass allocates.
ContactLayout savedLayout(int64_t n, int64_t nb) { return contactLayout(n, nb, n > 0, true); }

PairList savedPairs(void* saved, const ContactLayout& o)
{
    char* b = static_cast<char*>(saved);
    return PairList{reinterpret_cast<int32_t*>(b + o.oList), reinterpret_cast<int32_t*>(b + o.oCount),
                    reinterpret_cast<uint8_t*>(b + o.oAct)};
}

// Occupancy variants of the forward kernels the tuner chooses from, the
// default first: four waves per SIMD for the generic field read, three for
// the dense Tricubic specialisation (whose register use spills at four).
constexpr int kFwdVariants = 2;
const char* const kFwdVariantNames[kFwdVariants] = {"4waves", "3waves"};
const char* const kFwdVariantNamesDC[kFwdVariants] = {"3waves", "4waves"};

std::string forwardKey(const char* what, int64_t n, int64_t nb)
{
    return std::string(what) + " points~2^" + std::to_string(log2Bucket(n)) + " poses~2^" +
           std::to_string(log2Bucket(nb));
}

// Occupancy variants of the world-frame evaluate on the dense Tricubic path
// (the default first: the compiler's own register budget).
constexpr int kEvalVariants = 3;
const char* const kEvalVariantNames[kEvalVariants] = {"auto", "3waves", "4waves"};

// Persistent grid of the backward's piece kernel the tuner chooses from, in
// blocks per CU (the default first); the pieces and their sums are fixed by
// the piece table, so every variant gives the same bits.
constexpr int kPieceGridCount = 4;
constexpr int kPieceGrids[kPieceGridCount] = {8, 4, 16, 32};
const char* const kPieceGridNames[kPieceGridCount] = {"8/CU", "4/CU", "16/CU", "32/CU"};

// Forward launch shared by the plain, linearised and culled entry points;
// with saved (plain forward only), the pair record is written too.
template <bool Cached>
void forwardLaunch(const QGrid& q, const Law& L, const double* X, const double* w, int64_t n,
                   const double* poses, int64_t nb, const LinCache& lc, const CullArgs* cull,
                   double* out, int32_t* active, hipStream_t s, void* saved = nullptr)
{
    const int chunks = n > 0 ? chunksFor(n, nb) : 1;
    const int64_t chunkLen = n > 0 ? (n + chunks - 1) / chunks : 0;
    // a recording forward keeps its chunk partials in the record (no allocation)
    const bool rec = saved && !Cached && !cull;
    const ContactLayout lay = rec ? savedLayout(n, nb) : ContactLayout{};
    Scratch part(rec ? 0 : sizeof(double) * 7 * size_t(nb) * chunks + (cull ? size_t(nb) : 0), s);
    double* partial = rec ? reinterpret_cast<double*>(static_cast<char*>(saved) + lay.oPart)
                          : static_cast<double*>(part.p);
    uint8_t* anyLeaf = reinterpret_cast<uint8_t*>(partial + 7 * size_t(nb) * chunks);
    const dim3 grid(static_cast<unsigned>(chunks), static_cast<unsigned>(nb));
    if (cull)
    {
        hipLaunchKernelGGL(cullAnyKernel, dim3(unsigned(nb)), dim3(kBlock), 0, s, *cull, anyLeaf);
        SDFO_HIP(hipGetLastError());
        hipLaunchKernelGGL((contactForwardKernel<Cached, true>), grid, dim3(kBlock), 0, s, q, L, X,
                           w, n, poses, chunkLen, partial, lc, *cull, anyLeaf);
    }
    else if (rec)
    {
        const bool dc = fieldKind(q, L) == kFieldDenseCubic;
        auto launchSave = [&](int var) {
            auto go = [&](auto kindTag, auto wTag) {
                hipLaunchKernelGGL((contactForwardSaveKernel<decltype(kindTag)::value, decltype(wTag)::value>),
                                   grid, dim3(kBlock), 0, s, q, L, X, w, n, poses, chunkLen, partial,
                                   savedPairs(saved, lay));
            };
            using DC = std::integral_constant<int, kFieldDenseCubic>;
            using AnyK = std::integral_constant<int, kFieldAny>;
            using W3 = std::integral_constant<int, 3>;
            using W4 = std::integral_constant<int, 4>;
            if (dc)
                var == 1 ? go(DC{}, W4{}) : go(DC{}, W3{});
            else
                var == 1 ? go(AnyK{}, W3{}) : go(AnyK{}, W4{});
            SDFO_HIP(hipGetLastError());
        };
        launchSave(tunedVariant(forwardKey(dc ? "contact_forward_save tricubic" : "contact_forward_save",
                                           n, nb),
                                s, kFwdVariants, dc ? kFwdVariantNamesDC : kFwdVariantNames, launchSave));
        // the lists and the force finalize in one launch
        const int64_t blocks = nb * chunks + (nb + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(pairListKernel, dim3(unsigned(blocks)), dim3(kBlock), 0, s, n, nb, chunks,
                           chunkLen, savedPairs(saved, lay), static_cast<const double*>(partial),
                           out, active);
        SDFO_HIP(hipGetLastError());
        return;
    }
    else
    {
        // the direct forward on a dense Tricubic grid with cell records takes
        // the specialised field read (one record read, no other path)
        const bool dc = !Cached && fieldKind(q, L) == kFieldDenseCubic;
        auto launchFwd = [&](int var) {
            auto go = [&](auto kindTag, auto wTag) {
                hipLaunchKernelGGL((contactForwardKernel<Cached, false, decltype(kindTag)::value,
                                                         decltype(wTag)::value>),
                                   grid, dim3(kBlock), 0, s, q, L, X, w, n, poses, chunkLen, partial, lc,
                                   CullArgs{}, nullptr);
            };
            using DC = std::integral_constant<int, kFieldDenseCubic>;
            using AnyK = std::integral_constant<int, kFieldAny>;
            using W3 = std::integral_constant<int, 3>;
            using W4 = std::integral_constant<int, 4>;
            if (dc)
                var == 1 ? go(DC{}, W4{}) : go(DC{}, W3{});
            else
                var == 1 ? go(AnyK{}, W3{}) : go(AnyK{}, W4{});
            SDFO_HIP(hipGetLastError());
        };
        launchFwd(tunedVariant(forwardKey(Cached ? "contact_forward_linearized"
                                                 : (dc ? "contact_forward tricubic" : "contact_forward"),
                                          n, nb),
                               s, kFwdVariants, dc ? kFwdVariantNamesDC : kFwdVariantNames, launchFwd));
    }
    SDFO_HIP(hipGetLastError());
    hipLaunchKernelGGL(contactForwardFinalize, dim3(unsigned((nb + 255) / 256)), dim3(256), 0, s,
                       nb, chunks, static_cast<const double*>(partial), out, active);
    SDFO_HIP(hipGetLastError());
}

} // namespace

int64_t contactSavedBytes(int64_t n, int64_t nb)
{
    return nb > 0 ? int64_t(savedLayout(n, nb).bytes) : 0;
}

void launchContactForward(const sdfo_grid& g, const double* X, const double* w, int64_t n,
                          const double* poses, int64_t nb, const sdfo_contact_params& p,
                          double* out, int32_t* active, hipStream_t s, const CullArgs* cull,
                          void* saved)
{
    if (nb <= 0) return;
    forwardLaunch<false>(dev::makeQGrid(g), makeLaw(p), X, w, n, poses, nb, LinCache{}, cull, out,
                         active, s, saved);
}

namespace {

// Bytes of device scratch one evaluate call over nb pairs uses: the pair
// records, the cull boxes, the chunk partials and the per-pair leaf flags.
size_t evaluateScratchBytes(int64_t n, int64_t nb, int* chunksOut)
{
    const int chunks = n > 0 ? chunksFor(n, nb) : 1;
    if (chunksOut) *chunksOut = chunks;
    return sizeof(double) * (size_t(kEvalRec + 6 + 7 * chunks) * size_t(nb)) + size_t(nb);
}

void evaluateLaunch(const sdfo_grid& g, const double* X, const double* w, int64_t n,
                    const double* states, const double* frames, int64_t nb,
                    const sdfo_contact_params& p, const CullArgs* cull, const double* bodyBoxes9,
                    double* out, int32_t* active, hipStream_t s, int chunks, void* scratchBase)
{
    const QGrid q = dev::makeQGrid(g);
    const Law L = makeLaw(p);
    EvalPrep prep{};
    prep.states = states;
    prep.frames = frames;
    for (int k = 0; k < 3; ++k)
    {
        prep.bmin[k] = g.spec.bmin[k];
        prep.bmax[k] = g.spec.bmax[k];
    }
    prep.cullMargin = p.narrow_margin > 0.05 ? p.narrow_margin : 0.05;
    prep.bodyCheck = bodyBoxes9 ? 1 : 0;
    if (bodyBoxes9)
        for (int k = 0; k < 3; ++k)
        {
            prep.centerA[k] = bodyBoxes9[k];
            prep.halfA[k] = bodyBoxes9[3 + k];
            prep.halfB[k] = bodyBoxes9[6 + k];
        }
    const int64_t chunkLen = n > 0 ? (n + chunks - 1) / chunks : 0;
    double* rec = static_cast<double*>(scratchBase);
    double* boxes = rec + size_t(kEvalRec) * nb;
    double* partial = boxes + 6 * size_t(nb);
    uint8_t* anyLeaf = reinterpret_cast<uint8_t*>(partial + 7 * size_t(nb) * chunks);
    hipLaunchKernelGGL(evaluatePrepKernel, dim3(unsigned((nb + 255) / 256)), dim3(256), 0, s, prep,
                       nb, rec, boxes);
    SDFO_HIP(hipGetLastError());
    if (cull)
    {
        CullArgs cu = *cull;
        cu.boxes = boxes;
        hipLaunchKernelGGL(cullAnyKernel, dim3(unsigned(nb)), dim3(kBlock), 0, s, cu, anyLeaf);
        SDFO_HIP(hipGetLastError());
        if (fieldKind(q, L) == kFieldDenseCubic)
            hipLaunchKernelGGL((contactEvaluateKernel<true, kFieldDenseCubic>),
                               dim3(chunks, unsigned(nb)), dim3(kBlock), 0, s, q, L, X, w, n,
                               static_cast<const double*>(rec), chunkLen, partial, cu,
                               static_cast<const uint8_t*>(anyLeaf));
        else
            hipLaunchKernelGGL((contactEvaluateKernel<true>), dim3(chunks, unsigned(nb)),
                               dim3(kBlock), 0, s, q, L, X, w, n, static_cast<const double*>(rec),
                               chunkLen, partial, cu, static_cast<const uint8_t*>(anyLeaf));
    }
    else if (fieldKind(q, L) == kFieldDenseCubic)
    {
        auto launchEval = [&](int var) {
            auto go = [&](auto wTag) {
                hipLaunchKernelGGL((contactEvaluateKernel<false, kFieldDenseCubic, decltype(wTag)::value>),
                                   dim3(chunks, unsigned(nb)), dim3(kBlock), 0, s, q, L, X, w, n,
                                   static_cast<const double*>(rec), chunkLen, partial, CullArgs{},
                                   static_cast<const uint8_t*>(nullptr));
            };
            if (var == 1)
                go(std::integral_constant<int, 3>{});
            else if (var == 2)
                go(std::integral_constant<int, 4>{});
            else
                go(std::integral_constant<int, 0>{});
            SDFO_HIP(hipGetLastError());
        };
        launchEval(tunedVariant(forwardKey("contact_evaluate tricubic", n, nb), s, kEvalVariants,
                                kEvalVariantNames, launchEval));
    }
    else
        hipLaunchKernelGGL((contactEvaluateKernel<false>), dim3(chunks, unsigned(nb)), dim3(kBlock),
                           0, s, q, L, X, w, n, static_cast<const double*>(rec), chunkLen, partial,
                           CullArgs{}, static_cast<const uint8_t*>(nullptr));
    SDFO_HIP(hipGetLastError());
    hipLaunchKernelGGL(contactEvaluateFinalize, dim3(unsigned((nb + 255) / 256)), dim3(256), 0, s,
                       nb, chunks, static_cast<const double*>(partial),
                       static_cast<const double*>(rec), out, active);
    SDFO_HIP(hipGetLastError());
}

// Scratch of the allocation-free evaluate: one call of at most
// kNoAllocPairs pairs (chunks x pairs <= 2048 + pairs - 1 by chunksFor).
constexpr int64_t kNoAllocPairs = 4096;
constexpr size_t kNoAllocScratchBytes =
    sizeof(double) * size_t((kEvalRec + 6) * kNoAllocPairs + 7 * (2048 + kNoAllocPairs)) +
    size_t(kNoAllocPairs);
// A host-side name only: kernels get the per-device address that
// hipGetSymbolAddress returns for it.
static __device__ __align__(256) unsigned char scratch[(kNoAllocScratchBytes + 255) / 256 * 256];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
void* noAllocScratch(int device)
{
    static std::mutex m;
    static void* ptrs[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices)
        throw Error(SDFO_ERR_DEVICE, "device index " + std::to_string(device) + " out of range");
    std::lock_guard<std::mutex> lk(m);
    if (!ptrs[device])
    {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error(SDFO_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                             hipGetErrorString(e));
        ptrs[device] = ptr;
    }
    return ptrs[device];
}

} // namespace

void launchContactEvaluate(const sdfo_grid& g, const double* X, const double* w, int64_t n,
                           const double* states, const double* frames, int64_t nb,
                           const sdfo_contact_params& p, const CullArgs* cull,
                           const double* bodyBoxes9, double* out, int32_t* active, hipStream_t s)
{
    if (nb <= 0) return;
    int chunks = 1;
    const size_t bytes = evaluateScratchBytes(n, nb, &chunks);
    Scratch part(bytes, s);
    evaluateLaunch(g, X, w, n, states, frames, nb, p, cull, bodyBoxes9, out, active, s, chunks,
                   part.p);
}

void launchContactEvaluateNoAlloc(const sdfo_grid& g, const double* X, const double* w, int64_t n,
                                  const double* states, const double* frames, int64_t nb,
                                  const sdfo_contact_params& p, const CullArgs* cull,
                                  const double* bodyBoxes9, double* out, int32_t* active,
                                  hipStream_t s)
{
    if (nb > kNoAllocPairs)
        throw Error(SDFO_ERR_INVALID_ARGUMENT,
                    "sdfo_contact_evaluate_noalloc takes at most " + std::to_string(kNoAllocPairs) +
                        " pairs per call (got " + std::to_string(nb) + ")");
    if (nb <= 0) return;
    int chunks = 1;
    const size_t bytes = evaluateScratchBytes(n, nb, &chunks);
    if (bytes > sizeof(scratch))
        throw Error(SDFO_ERR_INVALID_ARGUMENT, "evaluate scratch of " + std::to_string(bytes) +
                                                   " bytes exceeds the static " +
                                                   std::to_string(sizeof(scratch)));
    evaluateLaunch(g, X, w, n, states, frames, nb, p, cull, bodyBoxes9, out, active, s, chunks,
                   noAllocScratch(g.device));
}

void launchCullMask(const CullArgs& cu, int64_t n, int64_t nBoxes, uint8_t* mask, hipStream_t s)
{
    if (n <= 0 || nBoxes <= 0) return;
    hipLaunchKernelGGL(cullMaskKernel, dim3(unsigned((n + kBlock - 1) / kBlock), unsigned(nBoxes)),
                       dim3(kBlock), 0, s, cu, n, mask);
    SDFO_HIP(hipGetLastError());
}

void launchContactLinearize(const sdfo_grid& g, const double* X, int64_t n, const double* pose,
                            const sdfo_contact_params& p, double* g0, double* grad0,
                            uint8_t* valid0, hipStream_t s)
{
    if (n <= 0) return;
    const QGrid q = dev::makeQGrid(g);
    const Law L = makeLaw(p);
    hipLaunchKernelGGL(contactLinearizeKernel, dim3(unsigned((n + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, q, L, X, n, pose, g0, grad0, valid0);
    SDFO_HIP(hipGetLastError());
}

void launchContactForwardLinearized(const double* g0, const double* grad0, const uint8_t* valid0,
                                    const double* tref, const double* X, const double* w, int64_t n,
                                    const double* poses, int64_t nb, const sdfo_contact_params& p,
                                    double* out, int32_t* active, hipStream_t s,
                                    const CullArgs* cull)
{
    if (nb <= 0) return;
    const LinCache lc{g0, grad0, valid0, {tref[0], tref[1], tref[2]}};
    forwardLaunch<true>(QGrid{}, makeLaw(p), X, w, n, poses, nb, lc, cull, out, active, s);
}

void launchContactBackward(const sdfo_grid& g, const double* X, const double* w, int64_t n,
                           const double* poses, int64_t nb, const sdfo_contact_params& p,
                           const double* gOut, double* gTwist, double* gParams, double* gX,
                           double* gW, hipStream_t s, const void* saved)
{
    if (nb <= 0) return;
    const QGrid q = dev::makeQGrid(g);
    const Law L = makeLaw(p);
    const bool wantPoint = (gX || gW) && n > 0;
    // with a forward's record every buffer is in it (the lists, counts and
    // flags of all poses included) and nothing is allocated; otherwise one
    // scratch block holds them for the call
    const ContactLayout lay = saved ? savedLayout(n, nb) : contactLayout(n, nb, wantPoint, false);
    const int chunks = lay.chunks, maxPieces = lay.maxPieces;
    const int64_t chunkLen = lay.chunkLen, group = lay.group;
    Scratch scr(saved ? 0 : lay.bytes, s);
    char* base8 = saved ? static_cast<char*>(const_cast<void*>(saved)) : static_cast<char*>(scr.p);
    int32_t* list = reinterpret_cast<int32_t*>(base8 + lay.oList);
    int32_t* count = reinterpret_cast<int32_t*>(base8 + lay.oCount);
    int32_t* items = reinterpret_cast<int32_t*>(base8 + lay.oItems);
    int32_t* poseStart = reinterpret_cast<int32_t*>(base8 + lay.oStart);
    int32_t* total = reinterpret_cast<int32_t*>(base8 + lay.oTotal);
    double* pieceSum = reinterpret_cast<double*>(base8 + lay.oSum);
    double* perPose = reinterpret_cast<double*>(base8 + lay.oPer);
    uint8_t* act = wantPoint ? reinterpret_cast<uint8_t*>(base8 + lay.oAct) : nullptr;
    double4* pointPartial = wantPoint ? reinterpret_cast<double4*>(base8 + lay.oPts) : nullptr;
    const bool fused = gParams && group == nb;
    const int cus = computeUnits();
    const bool dc = fieldKind(q, L) == kFieldDenseCubic;
    for (int64_t base = 0; base < nb; base += group)
    {
        const int64_t cnt = base + group < nb ? group : nb - base;
        const dim3 grid(static_cast<unsigned>(chunks), static_cast<unsigned>(cnt));
        // this group's slots and flags (a record holds every pose's)
        PairList gl{list, count, act};
        if (saved)
        {
            gl.list = list + base * chunks * chunkLen;
            gl.count = count + base * chunks;
            gl.act = wantPoint ? act + base * n : nullptr;
        }
        else if (L.mode == SDFO_TRILINEAR)
            hipLaunchKernelGGL(contactScreenKernel<true>, grid, dim3(kBlock), 0, s, q, L, X, n,
                               poses, chunkLen, base, gl);
        else if (dc)
            hipLaunchKernelGGL((contactScreenKernel<false, kFieldDenseCubic>), grid, dim3(kBlock), 0,
                               s, q, L, X, n, poses, chunkLen, base, gl);
        else
            hipLaunchKernelGGL(contactScreenKernel<false>, grid, dim3(kBlock), 0, s, q, L, X, n,
                               poses, chunkLen, base, gl);
        SDFO_HIP(hipGetLastError());
        hipLaunchKernelGGL(pieceTableKernel, dim3(1), dim3(kTableBlock), 0, s,
                           int64_t(cnt) * chunks, chunks, maxPieces,
                           static_cast<const int32_t*>(gl.count), items, poseStart, total);
        SDFO_HIP(hipGetLastError());
        // persistent one-wave blocks: by default eight per CU (two waves per
        // SIMD, the kernel's register budget, four SIMDs per CU)
        auto pieces = [&](int var) {
            if (var < 0 || var >= kPieceGridCount) var = 0;
            auto go = [&](auto kindTag) {
                hipLaunchKernelGGL(contactBackwardPieceKernel<decltype(kindTag)::value>,
                                   dim3(unsigned(kPieceGrids[var] * cus)), dim3(64), 0, s, q, L, X, w, n,
                                   poses, gOut, chunks, chunkLen, maxPieces, base,
                                   static_cast<const int32_t*>(gl.list),
                                   static_cast<const int32_t*>(gl.count),
                                   static_cast<const int32_t*>(items),
                                   static_cast<const int32_t*>(total), pieceSum, pointPartial);
            };
            // the generic field read: the dense Tricubic specialisation
            // compiles to 256 VGPRs (one wave per SIMD) here and is slower
            go(std::integral_constant<int, kFieldAny>{});
            SDFO_HIP(hipGetLastError());
        };
        pieces(tunedVariant("contact_backward_pieces pairs~2^" +
                                std::to_string(log2Bucket(int64_t(cnt) * n)),
                            s, kPieceGridCount, kPieceGridNames, pieces));
        // the tail: point reduction, pose finalize and, for a single group,
        // the parameter sum, as the blocks of one launch
        const int64_t nReduce = wantPoint ? (n + kReducePoints - 1) / kReducePoints : 0;
        const int64_t poseBlocks = (gTwist || gParams) ? (cnt + kBlock / 64 - 1) / (kBlock / 64) : 0;
        const int64_t tailBlocks = nReduce + poseBlocks + (fused ? 1 : 0);
        if (tailBlocks > 0)
        {
            hipLaunchKernelGGL(backwardTailKernel, dim3(unsigned(tailBlocks)), dim3(kBlock), 0, s,
                               nReduce, n, int(cnt), static_cast<const double4*>(pointPartial),
                               static_cast<const uint8_t*>(gl.act), base > 0 ? 1 : 0, gX, gW, base,
                               static_cast<const int32_t*>(poseStart),
                               static_cast<const double*>(pieceSum),
                               static_cast<const int32_t*>(total), gTwist,
                               fused ? nullptr : perPose, fused ? gParams : nullptr);
            SDFO_HIP(hipGetLastError());
        }
    }
    if (gParams && !fused)
    {
        hipLaunchKernelGGL(paramReduceKernel, dim3(1), dim3(kBlock), 0, s, nb,
                           static_cast<const double*>(perPose), gParams);
        SDFO_HIP(hipGetLastError());
    }
}

} // namespace sdfo
