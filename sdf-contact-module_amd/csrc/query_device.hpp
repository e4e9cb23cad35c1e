// The per-point query modes (SdfOracle.cpp:176-432, 448-1092) shared by the
// FP32 query kernels (query.hip) and the FP64 ones on a dense grid's FP64
// fields (query_f64.hip, built without FMA contraction so the results are the
// CPU path's bit for bit).  F64 selects the FP64 field arrays.
#pragma once

#include <hip/hip_runtime.h>

#include "interp_device.hpp"
#include "sdfo_internal.hpp"

namespace sdfo {

using dev::Cell;
using dev::Interp;
using dev::QGrid;
using dev::V3T;

namespace {

template <typename T>
struct QOut
{
    T gap;
    V3T<T> grad;
    V3T<T> normal;
    bool valid;
    int branch;
};

template <typename T>
__device__ __forceinline__ T tsqrt(T x);
template <>
__device__ __forceinline__ float tsqrt<float>(float x)
{
    return sqrtf(x);
}
template <>
__device__ __forceinline__ double tsqrt<double>(double x)
{
    return sqrt(x);
}

// Unit normal of trilinear / tricubic (SdfOracle.cpp:585-590, 1083-1087).
template <typename T>
__device__ __forceinline__ void finishInterp(const Interp<T>& it, QOut<T>& o, bool withNormal = true)
{
    o.gap = it.gap;
    o.grad = it.grad;
    o.valid = true;
    if (!withNormal) return; // the caller writes no normal and no witness
    const T len = tsqrt<T>(it.grad.x * it.grad.x + it.grad.y * it.grad.y + it.grad.z * it.grad.z);
    if (len > T(1e-15))
        o.normal = V3T<T>{it.grad.x / len, it.grad.y / len, it.grad.z / len};
    else
        o.normal = V3T<T>{T(1), T(0), T(0)};
    o.valid = true;
}

// First / second order Taylor at the nearest voxel (SdfOracle.cpp:448-520,
// 600-654).  Second order without a Hessian degrades to first order.
template <typename T, bool F64 = false>
__device__ __forceinline__ void taylorQuery(const QGrid& g, double px, double py, double pz,
                                            bool second, QOut<T>& o)
{
    int i, j, k;
    if (!dev::nearestVoxel(g, px, py, pz, i, j, k)) return;
    const int64_t rec = dev::recordOf(g, i, j, k);
    if (rec < 0) return;
    const auto f = dev::fieldAt<F64>(g, rec);
    const T dx = T(px - (g.bmin[0] + g.voxel[0] * (i + 0.5)));
    const T dy = T(py - (g.bmin[1] + g.voxel[1] * (j + 0.5)));
    const T dz = T(pz - (g.bmin[2] + g.voxel[2] * (k + 0.5)));
    const T nx = T(f.y), ny = T(f.z), nz = T(f.w);
    const T first = nx * dx + ny * dy + nz * dz;
    std::conditional_t<F64, double, float> h[6];
    if (!second || !g.hasHessian || !dev::hessianAt<F64>(g, rec, h))
    {
        o.gap = T(f.x) + first;
        o.grad = V3T<T>{nx, ny, nz};
        o.normal = o.grad;
        o.valid = true;
        return;
    }
    const T hxx = T(h[0]), hxy = T(h[1]), hxz = T(h[2]), hyy = T(h[3]), hyz = T(h[4]),
            hzz = T(h[5]);
    const T Hx = hxx * dx + hxy * dy + hxz * dz;
    const T Hy = hxy * dx + hyy * dy + hyz * dz;
    const T Hz = hxz * dx + hyz * dy + hzz * dz;
    o.gap = T(f.x) + first + T(0.5) * (Hx * dx + Hy * dy + Hz * dz);
    o.grad = V3T<T>{nx + Hx, ny + Hy, nz + Hz};
    const T len = tsqrt<T>(o.grad.x * o.grad.x + o.grad.y * o.grad.y + o.grad.z * o.grad.z);
    o.normal = len > T(1e-15) ? V3T<T>{o.grad.x / len, o.grad.y / len, o.grad.z / len}
                              : V3T<T>{nx, ny, nz};
    o.valid = true;
}

// ContactAware candidate selection (SdfOracle.cpp:269-432): no candidate
// arrays -> tricubic; zero candidates at the voxel -> first order.
template <typename T, bool F64 = false>
__device__ __forceinline__ void contactAwareQuery(const QGrid& g, double px, double py,
                                                  double pz, QOut<T>& o, bool wantBranch = true)
{
    int i, j, k;
    if (!dev::nearestVoxel(g, px, py, pz, i, j, k)) return;
    const int64_t rec = dev::recordOf(g, i, j, k);
    if (rec < 0) return;
    const uint8_t* cnt = g.sparse ? g.sccount : g.ccount;
    const int32_t* cb = g.sparse ? g.scbranch : g.cbranch;
    int count = int(cnt[rec]);
    if (count > g.stride) count = g.stride;
    if (count <= 0)
    {
        taylorQuery<T, F64>(g, px, py, pz, false, o);
        return;
    }
    if (count > 8) count = 8;
    const T dx = T(px - (g.bmin[0] + g.voxel[0] * (i + 0.5)));
    const T dy = T(py - (g.bmin[1] + g.voxel[1] * (j + 0.5)));
    const T dz = T(pz - (g.bmin[2] + g.voxel[2] * (k + 0.5)));
    T best = T(0);
    int bestC = 0;
    bool have = false;
    for (int c = 0; c < count; ++c)
    {
        const auto f = dev::candidateAt<F64>(g, rec * g.stride + c);
        T nx = T(f.y), ny = T(f.z), nz = T(f.w);
        const T nn = tsqrt<T>(nx * nx + ny * ny + nz * nz);
        if (nn > T(1e-15))
        {
            nx /= nn;
            ny /= nn;
            nz /= nn;
        }
        else
        {
            nx = T(1);
            ny = T(0);
            nz = T(0);
        }
        const T gap = T(f.x) + (nx * dx + ny * dy + nz * dz);
        const T ag = gap < T(0) ? -gap : gap;
        if (!have || ag < best)
        {
            have = true;
            best = ag;
            o.gap = gap;
            o.normal = V3T<T>{nx, ny, nz};
            bestC = c;
        }
    }
    if (wantBranch) o.branch = cb[rec * g.stride + bestC];
    o.grad = o.normal;
    o.valid = true;
}

// ContactAware on a grid known to hold candidate arrays (no tricubic branch to
// carry); a kernel specialisation, not a query mode of the interface.
constexpr int kContactAwareCand = 100;

template <typename T, int Mode, bool F64 = false>
__device__ __forceinline__ QOut<T> queryPoint(const QGrid& g, double px, double py, double pz,
                                              bool wantBranch = true)
{
    constexpr int mode = Mode == kContactAwareCand ? SDFO_CONTACT_AWARE : Mode;
    QOut<T> o;
    o.gap = T(0);
    o.grad = V3T<T>{T(0), T(0), T(0)};
    o.normal = V3T<T>{T(1), T(0), T(0)};
    o.valid = false;
    o.branch = -1;
    const bool cubic = mode == SDFO_TRICUBIC ||
                       (Mode == SDFO_CONTACT_AWARE && mode == SDFO_CONTACT_AWARE && g.stride <= 0);
    if (mode == SDFO_TRILINEAR || cubic)
    {
        const Cell c = dev::locateCell(g, px, py, pz);
        if (!c.valid) return o;
        Interp<T> it;
        const bool ok = cubic ? dev::tricubic<T, false, F64>(g, c, it)
                              : dev::trilinear<T, false, F64>(g, c, it);
        if (ok) finishInterp<T>(it, o);
        return o;
    }
    if (mode == SDFO_CONTACT_AWARE)
    {
        contactAwareQuery<T, F64>(g, px, py, pz, o, wantBranch);
        return o;
    }
    taylorQuery<T, F64>(g, px, py, pz, mode != SDFO_FIRST_ORDER, o);
    return o;
}

// 3-vector stored contiguously (12 or 24 bytes per lane, one vector memory
// instruction per lane and coalesced across the wave).
template <typename TP>
struct P3
{
    TP x, y, z;
};

template <typename TP, typename T>
__device__ __forceinline__ void writeOut(int64_t i, const QOut<T>& o, double px, double py,
                                         double pz, TP* __restrict__ gap,
                                         P3<TP>* __restrict__ grad3, uint8_t* __restrict__ valid,
                                         P3<TP>* __restrict__ normal3,
                                         P3<TP>* __restrict__ witness3, int32_t* __restrict__ branch)
{
    gap[i] = TP(o.gap);
    grad3[i] = P3<TP>{TP(o.grad.x), TP(o.grad.y), TP(o.grad.z)};
    valid[i] = o.valid ? 1 : 0;
    if (normal3) normal3[i] = P3<TP>{TP(o.normal.x), TP(o.normal.y), TP(o.normal.z)};
    if (witness3)
    {
        // witness = p - gap * normal (SdfOracle.cpp:593); zero when invalid.
        witness3[i] = o.valid ? P3<TP>{TP(T(px) - o.gap * o.normal.x), TP(T(py) - o.gap * o.normal.y),
                                       TP(T(pz) - o.gap * o.normal.z)}
                              : P3<TP>{TP(0), TP(0), TP(0)};
    }
    if (branch) branch[i] = o.branch;
}

template <typename TP, typename T, int Mode, bool F64 = false>
__global__ void __launch_bounds__(256)
queryKernel(QGrid g, const P3<TP>* __restrict__ pts, int64_t n, TP* __restrict__ gap,
            P3<TP>* __restrict__ grad3, uint8_t* __restrict__ valid, P3<TP>* __restrict__ normal3,
            P3<TP>* __restrict__ witness3, int32_t* __restrict__ branch)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const P3<TP> p = pts[i];
    const double px = double(p.x), py = double(p.y), pz = double(p.z);
    writeOut<TP, T>(i, queryPoint<T, Mode, F64>(g, px, py, pz, branch != nullptr), px, py, pz, gap,
                    grad3, valid, normal3, witness3, branch);
}


} // namespace

} // namespace sdfo
