// Point-query kernels: the five query modes over dense or sparse grids
// (SdfOracle.cpp:176-432, 448-1092).  One lane per point; the point stream
// and the outputs are read/written once and coalesced, the field corners
// are gathered from the FP32 mirror (float4 = phi + normal per voxel).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <memory>
#include <vector>

#include "interp_device.hpp"
#include "sdfo_internal.hpp"

namespace sdfo {

using dev::Cell;
using dev::Interp;
using dev::QGrid;
using dev::V3T;

namespace dev {

QGrid makeQGrid(const sdfo_grid& g)
{
    QGrid q{};
    for (int a = 0; a < 3; ++a)
    {
        q.bmin[a] = g.spec.bmin[a];
        q.voxel[a] = g.spec.voxel[a];
        q.n[a] = g.spec.n[a];
    }
    q.sparse = g.sparse ? 1 : 0;
    q.hasHessian = g.hasHessian ? 1 : 0;
    q.stride = g.stride;
    q.pn = g.pn;
    q.cellRec = g.sparse ? nullptr : g.cellRec;
    q.h6f = g.h6f;
    q.ccount = g.ccount;
    q.cpn = g.cpn;
    q.cbranch = g.cbranch;
    q.bd = g.blockDim;
    q.nbx = g.nbx;
    q.nby = g.nby;
    q.nbz = g.nbz;
    q.table = g.table;
    q.spn = g.spn;
    q.sh6 = g.sh6;
    q.sccount = g.sccount;
    q.scpn = g.scpn;
    q.scbranch = g.scbranch;
    q.scellRec = g.sparse ? g.scellRec : nullptr;
    q.scellMask = g.scellMask;
    q.maskWords = int((int64_t(g.blockDim) * g.blockDim * g.blockDim + 63) / 64);
    return q;
}

} // namespace dev

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
template <typename T>
__device__ __forceinline__ void taylorQuery(const QGrid& g, double px, double py, double pz,
                                            bool second, QOut<T>& o)
{
    int i, j, k;
    if (!dev::nearestVoxel(g, px, py, pz, i, j, k)) return;
    const int64_t rec = dev::recordOf(g, i, j, k);
    if (rec < 0) return;
    const float4 f = dev::fieldAt(g, rec);
    const T dx = T(px - (g.bmin[0] + g.voxel[0] * (i + 0.5)));
    const T dy = T(py - (g.bmin[1] + g.voxel[1] * (j + 0.5)));
    const T dz = T(pz - (g.bmin[2] + g.voxel[2] * (k + 0.5)));
    const T nx = T(f.y), ny = T(f.z), nz = T(f.w);
    const T first = nx * dx + ny * dy + nz * dz;
    const float* hp = g.sparse ? g.sh6 : g.h6f;
    if (!second || hp == nullptr || !g.hasHessian)
    {
        o.gap = T(f.x) + first;
        o.grad = V3T<T>{nx, ny, nz};
        o.normal = o.grad;
        o.valid = true;
        return;
    }
    const float* h = hp + 6 * rec;
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
template <typename T>
__device__ __forceinline__ void contactAwareQuery(const QGrid& g, double px, double py,
                                                  double pz, QOut<T>& o, bool wantBranch = true)
{
    int i, j, k;
    if (!dev::nearestVoxel(g, px, py, pz, i, j, k)) return;
    const int64_t rec = dev::recordOf(g, i, j, k);
    if (rec < 0) return;
    const uint8_t* cnt = g.sparse ? g.sccount : g.ccount;
    const float4* cp = g.sparse ? g.scpn : g.cpn;
    const int32_t* cb = g.sparse ? g.scbranch : g.cbranch;
    int count = int(cnt[rec]);
    if (count > g.stride) count = g.stride;
    if (count <= 0)
    {
        taylorQuery<T>(g, px, py, pz, false, o);
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
        const float4 f = cp[rec * g.stride + c];
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

template <typename T, int Mode>
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
        const bool ok = cubic ? dev::tricubic<T, false>(g, c, it) : dev::trilinear<T, false>(g, c, it);
        if (ok) finishInterp<T>(it, o);
        return o;
    }
    if (mode == SDFO_CONTACT_AWARE)
    {
        contactAwareQuery<T>(g, px, py, pz, o, wantBranch);
        return o;
    }
    taylorQuery<T>(g, px, py, pz, mode != SDFO_FIRST_ORDER, o);
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

template <typename TP, typename T, int Mode>
__global__ void __launch_bounds__(256)
queryKernel(QGrid g, const P3<TP>* __restrict__ pts, int64_t n, TP* __restrict__ gap,
            P3<TP>* __restrict__ grad3, uint8_t* __restrict__ valid, P3<TP>* __restrict__ normal3,
            P3<TP>* __restrict__ witness3, int32_t* __restrict__ branch)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const P3<TP> p = pts[i];
    const double px = double(p.x), py = double(p.y), pz = double(p.z);
    writeOut<TP, T>(i, queryPoint<T, Mode>(g, px, py, pz, branch != nullptr), px, py, pz, gap, grad3,
                    valid, normal3, witness3, branch);
}

// Trilinear / Tricubic on a dense grid with cell records, two points per
// thread: both points' 128-byte records are requested before either is
// interpolated, doubling the loads in flight per lane (the gather is
// latency-bound at one point per lane).  Same arithmetic as queryKernel.
constexpr int kPairBlock = 128;

// Points ia and ib (each processed when below its end) of one grid.
template <typename TP, typename T, int Mode, bool Sparse>
__device__ __forceinline__ void pairQuery(const QGrid& g, const P3<TP>* __restrict__ pts,
                                          int64_t ia, int64_t ib, int64_t end,
                                          TP* __restrict__ gap, P3<TP>* __restrict__ grad3,
                                          uint8_t* __restrict__ valid, P3<TP>* __restrict__ normal3,
                                          P3<TP>* __restrict__ witness3,
                                          int32_t* __restrict__ branch)
{
    double pc[2][3];
    dev::Cell c[2];
    float4 f[2][8];
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
        const int64_t i = h ? ib : ia;
        P3<TP> p{TP(0), TP(0), TP(0)};
        if (i < end) p = pts[i];
        pc[h][0] = double(p.x);
        pc[h][1] = double(p.y);
        pc[h][2] = double(p.z);
        c[h] = dev::locateCell(g, pc[h][0], pc[h][1], pc[h][2]);
        c[h].valid = c[h].valid && i < end;
        if (c[h].valid)
        {
            if (!Sparse)
            {
                const float4* r = g.cellRec + 8 * (int64_t(c[h].i0) +
                                                   int64_t(g.n[0] - 1) * (c[h].j0 + int64_t(g.n[1] - 1) * c[h].k0));
#pragma unroll
                for (int q = 0; q < 8; ++q) f[h][q] = r[q];
            }
            else if (const float4* r = dev::cellRecord(g, c[h].i0, c[h].j0, c[h].k0))
            {
#pragma unroll
                for (int q = 0; q < 8; ++q) f[h][q] = r[q];
            }
            else
                c[h].valid = false; // a missing corner sample: an invalid result
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
        const int64_t i = h ? ib : ia;
        if (i >= end) continue;
        QOut<T> o;
        o.gap = T(0);
        o.grad = V3T<T>{T(0), T(0), T(0)};
        o.normal = V3T<T>{T(1), T(0), T(0)};
        o.valid = false;
        o.branch = -1;
        if (c[h].valid)
        {
            Interp<T> it;
            if (Mode == SDFO_TRICUBIC)
                dev::tricubicCorners<T, false>(g, c[h], f[h], it);
            else
                dev::trilinearCorners<T, false>(g, c[h], f[h], it);
            finishInterp<T>(it, o, normal3 != nullptr || witness3 != nullptr);
        }
        writeOut<TP, T>(i, o, pc[h][0], pc[h][1], pc[h][2], gap, grad3, valid, normal3, witness3,
                        branch);
    }
}

// Trilinear / Tricubic on a dense grid with cell records, two points per
// thread: both points' 128-byte records are requested before either is
// interpolated, doubling the loads in flight per lane (the gather is
// latency-bound at one point per lane).  Same arithmetic as queryKernel.
template <typename TP, typename T, int Mode, bool Sparse>
__global__ void __launch_bounds__(kPairBlock)
queryPairKernel(QGrid g, const P3<TP>* __restrict__ pts, int64_t n, TP* __restrict__ gap,
                P3<TP>* __restrict__ grad3, uint8_t* __restrict__ valid,
                P3<TP>* __restrict__ normal3, P3<TP>* __restrict__ witness3,
                int32_t* __restrict__ branch)
{
    const int64_t i0 = int64_t(blockIdx.x) * (2 * kPairBlock) + threadIdx.x;
    pairQuery<TP, T, Mode, Sparse>(g, pts, i0, i0 + kPairBlock, n, gap, grad3, valid, normal3, witness3,
                           branch);
}

// One point stream of a batched query: its grid, its points and its range of
// 128-point work units (one wave each, two points per lane) in the launch.
constexpr int kUnitPoints = 128;
struct StreamDesc
{
    QGrid g;
    int64_t pointBegin, pointEnd;
    int64_t unitBegin;
};

template <int Mode, int Block>
__global__ void __launch_bounds__(Block)
queryBatchedKernel(const StreamDesc* __restrict__ sd, int ns, int64_t units,
                   const P3<float>* __restrict__ pts, float* __restrict__ gap,
                   P3<float>* __restrict__ grad3, uint8_t* __restrict__ valid)
{
    const int64_t u = int64_t(blockIdx.x) * (Block / 64) + (threadIdx.x >> 6);
    if (u >= units) return;
    // the stream owning this unit (wave-uniform binary search)
    int lo = 0, hi = ns - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (sd[mid].unitBegin <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    const StreamDesc& D = sd[lo];
    const int64_t i = D.pointBegin + (u - D.unitBegin) * kUnitPoints + (threadIdx.x & 63);
    const bool cubic = Mode == SDFO_TRICUBIC || (Mode == SDFO_CONTACT_AWARE && D.g.stride <= 0);
    if ((D.g.sparse ? D.g.scellRec != nullptr : D.g.cellRec != nullptr) &&
        (cubic || Mode == SDFO_TRILINEAR))
    {
        if (cubic)
            pairQuery<float, float, SDFO_TRICUBIC, true>(D.g, pts, i, i + 64, D.pointEnd, gap, grad3,
                                                         valid, nullptr, nullptr, nullptr);
        else
            pairQuery<float, float, SDFO_TRILINEAR, true>(D.g, pts, i, i + 64, D.pointEnd, gap,
                                                          grad3, valid, nullptr, nullptr, nullptr);
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
        const int64_t k = i + 64 * h;
        if (k >= D.pointEnd) continue;
        const P3<float> p = pts[k];
        const QOut<float> o = queryPoint<float, Mode>(D.g, double(p.x), double(p.y), double(p.z));
        gap[k] = o.gap;
        grad3[k] = P3<float>{o.grad.x, o.grad.y, o.grad.z};
        valid[k] = o.valid ? 1 : 0;
    }
}

template <int Block>
void launchBatchedBlock(int mode, const StreamDesc* sd, int ns, int64_t units, const float* pts,
                        float* gap, float* grad3, uint8_t* valid, hipStream_t s)
{
    using V = P3<float>;
    const V* p = reinterpret_cast<const V*>(pts);
    V* gr = reinterpret_cast<V*>(grad3);
    constexpr int per = Block / 64;
    const dim3 grid(unsigned((units + per - 1) / per)), block(Block);
    switch (mode)
    {
    case SDFO_FIRST_ORDER:
        hipLaunchKernelGGL((queryBatchedKernel<SDFO_FIRST_ORDER, Block>), grid, block, 0, s, sd, ns,
                           units, p, gap, gr, valid);
        break;
    case SDFO_SECOND_ORDER:
        hipLaunchKernelGGL((queryBatchedKernel<SDFO_SECOND_ORDER, Block>), grid, block, 0, s, sd, ns,
                           units, p, gap, gr, valid);
        break;
    case SDFO_TRILINEAR:
        hipLaunchKernelGGL((queryBatchedKernel<SDFO_TRILINEAR, Block>), grid, block, 0, s, sd, ns,
                           units, p, gap, gr, valid);
        break;
    case SDFO_TRICUBIC:
        hipLaunchKernelGGL((queryBatchedKernel<SDFO_TRICUBIC, Block>), grid, block, 0, s, sd, ns,
                           units, p, gap, gr, valid);
        break;
    default:
        hipLaunchKernelGGL((queryBatchedKernel<SDFO_CONTACT_AWARE, Block>), grid, block, 0, s, sd, ns,
                           units, p, gap, gr, valid);
        break;
    }
}

template <typename TP, typename T>
void launchQuery(const sdfo_grid& g, int mode, const TP* pts, int64_t n, TP* gap, TP* grad3,
                 uint8_t* valid, TP* normal3, TP* witness3, int32_t* branch, hipStream_t s)
{
    if (n <= 0) return;
    const QGrid q = dev::makeQGrid(g);
    using V = P3<TP>;
    const dim3 grid(unsigned((n + 255) / 256)), block(256);
    const V* p = reinterpret_cast<const V*>(pts);
    V* gr = reinterpret_cast<V*>(grad3);
    V* nm = reinterpret_cast<V*>(normal3);
    V* wt = reinterpret_cast<V*>(witness3);
    const bool cubic = mode == SDFO_TRICUBIC || (mode == SDFO_CONTACT_AWARE && q.stride <= 0);
    if ((q.sparse ? q.scellRec != nullptr : q.cellRec != nullptr) && (cubic || mode == SDFO_TRILINEAR))
    {
        const dim3 pg(unsigned((n + 2 * kPairBlock - 1) / (2 * kPairBlock))), pb(kPairBlock);
        auto launch = [&](auto sparseTag) {
            constexpr bool Sp = decltype(sparseTag)::value;
            if (cubic)
                hipLaunchKernelGGL((queryPairKernel<TP, T, SDFO_TRICUBIC, Sp>), pg, pb, 0, s, q, p, n,
                                   gap, gr, valid, nm, wt, branch);
            else
                hipLaunchKernelGGL((queryPairKernel<TP, T, SDFO_TRILINEAR, Sp>), pg, pb, 0, s, q, p, n,
                                   gap, gr, valid, nm, wt, branch);
        };
        if (q.sparse)
            launch(std::true_type{});
        else
            launch(std::false_type{});
        SDFO_HIP(hipGetLastError());
        return;
    }
    switch (mode)
    {
    case SDFO_FIRST_ORDER:
        hipLaunchKernelGGL((queryKernel<TP, T, SDFO_FIRST_ORDER>), grid, block, 0, s, q, p, n, gap,
                           gr, valid, nm, wt, branch);
        break;
    case SDFO_SECOND_ORDER:
        hipLaunchKernelGGL((queryKernel<TP, T, SDFO_SECOND_ORDER>), grid, block, 0, s, q, p, n, gap,
                           gr, valid, nm, wt, branch);
        break;
    case SDFO_TRILINEAR:
        hipLaunchKernelGGL((queryKernel<TP, T, SDFO_TRILINEAR>), grid, block, 0, s, q, p, n, gap,
                           gr, valid, nm, wt, branch);
        break;
    case SDFO_TRICUBIC:
        hipLaunchKernelGGL((queryKernel<TP, T, SDFO_TRICUBIC>), grid, block, 0, s, q, p, n, gap, gr,
                           valid, nm, wt, branch);
        break;
    default:
        if (q.stride > 0)
            hipLaunchKernelGGL((queryKernel<TP, T, kContactAwareCand>), grid, block, 0, s, q, p, n,
                               gap, gr, valid, nm, wt, branch);
        else
            hipLaunchKernelGGL((queryKernel<TP, T, SDFO_CONTACT_AWARE>), grid, block, 0, s, q, p, n,
                               gap, gr, valid, nm, wt, branch);
        break;
    }
    SDFO_HIP(hipGetLastError());
}

} // namespace

void launchQueryF32(const sdfo_grid& g, int mode, const float* pts, int64_t n, float* gap,
                    float* grad3, uint8_t* valid, float* normal3, float* witness3,
                    int32_t* branch, hipStream_t s)
{
    launchQuery<float, float>(g, mode, pts, n, gap, grad3, valid, normal3, witness3, branch, s);
}

void launchQueryF64(const sdfo_grid& g, int mode, const double* pts, int64_t n, double* gap,
                    double* grad3, uint8_t* valid, double* normal3, double* witness3,
                    int32_t* branch, hipStream_t s)
{
    launchQuery<double, double>(g, mode, pts, n, gap, grad3, valid, normal3, witness3, branch, s);
}

// ---- batched query plans ------------------------------------------------
QueryPlan::~QueryPlan()
{
    if (table) (void)hipFree(table);
}

std::unique_ptr<QueryPlan> makeQueryPlan(int ns, const sdfo_grid* const* grids,
                                         const int64_t* offsets)
{
    auto plan = std::make_unique<QueryPlan>();
    plan->streams = ns;
    std::vector<StreamDesc> t(static_cast<size_t>(ns));
    int64_t units = 0;
    for (int q = 0; q < ns; ++q)
    {
        t[size_t(q)].g = dev::makeQGrid(*grids[q]);
        t[size_t(q)].pointBegin = offsets[q];
        t[size_t(q)].pointEnd = offsets[q + 1];
        t[size_t(q)].unitBegin = units;
        units += (offsets[q + 1] - offsets[q] + kUnitPoints - 1) / kUnitPoints;
    }
    plan->points = ns > 0 ? offsets[ns] : 0;
    plan->units = units;
    SDFO_HIP(hipMalloc(&plan->table, sizeof(StreamDesc) * std::max<size_t>(1, t.size())));
    if (!t.empty())
        SDFO_HIP(hipMemcpy(plan->table, t.data(), sizeof(StreamDesc) * t.size(),
                           hipMemcpyHostToDevice));
    return plan;
}

void runQueryPlan(const QueryPlan& plan, int mode, int block, const float* pts, float* gap,
                  float* grad3, uint8_t* valid, hipStream_t s)
{
    if (plan.streams <= 0 || plan.units <= 0) return;
    const StreamDesc* sd = static_cast<const StreamDesc*>(plan.table);
    switch (block)
    {
    case 64:
        launchBatchedBlock<64>(mode, sd, plan.streams, plan.units, pts, gap, grad3, valid, s);
        break;
    case 128:
        launchBatchedBlock<128>(mode, sd, plan.streams, plan.units, pts, gap, grad3, valid, s);
        break;
    case 512:
        launchBatchedBlock<512>(mode, sd, plan.streams, plan.units, pts, gap, grad3, valid, s);
        break;
    default:
        launchBatchedBlock<256>(mode, sd, plan.streams, plan.units, pts, gap, grad3, valid, s);
        break;
    }
    SDFO_HIP(hipGetLastError());
}

} // namespace sdfo
