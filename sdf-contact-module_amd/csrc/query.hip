// Point-query kernels: the five query modes over dense or sparse grids
// (SdfOracle.cpp:176-432, 448-1092).  One lane per point; the point stream
// and the outputs are read/written once and coalesced, the field corners
// are gathered from the FP32 mirror (float4 = phi + normal per voxel).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "interp_device.hpp"
#include "query_device.hpp"
#include "sdfo_internal.hpp"

namespace sdfo {

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
    if (!g.sparse)
    {
        q.phi64 = g.phi;
        q.n64 = g.n3;
        q.h64 = g.h6;
        q.cphi64 = g.cphi;
        q.cn64 = g.cn3;
    }
    return q;
}

} // namespace dev

namespace {
// Default block size of the two-points-per-lane query (queryPairKernel).
constexpr int kPairBlock = 128;
// Points per lane of the single-grid kernel: two (three took 32.4 us against
// 29.8 us for 1e6 Tricubic points on a 128^3 grid).
constexpr int kLanePoints = 2;

// Where pairQuery takes a cell's 8 corner samples from: the dense record
// array, any grid's records (dense or through the brick table), or gathered
// corner by corner from the FP32 mirror of a grid without records (dense:
// eight addresses from the cell index; sparse: eight brick-table lookups).
enum CornerSrc
{
    kDenseRec = 0,
    kAnyRec = 1,
    kDenseGather = 2,
    kSparseGather = 3
};

// The 8 corners of a cell from the mirror; false if a sample is missing.  All
// addresses are formed before any load, so the eight loads are in flight
// together (a missing sparse sample reads record 0 and invalidates the cell).
template <int Src>
__device__ __forceinline__ bool gatherCorners(const QGrid& g, const dev::Cell& c, float4 (&f)[8])
{
    int64_t rec[8];
    bool ok = true;
    if (Src == kDenseGather)
    {
        const int64_t nx = g.n[0], nxy = int64_t(g.n[0]) * g.n[1];
        const int64_t b = c.i0 + nx * c.j0 + nxy * c.k0;
#pragma unroll
        for (int q = 0; q < 8; ++q) rec[q] = b + (q & 1) + nx * ((q >> 1) & 1) + nxy * (q >> 2);
    }
    else
    {
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            const int64_t r = dev::recordOf(g, c.i0 + (q & 1), c.j0 + ((q >> 1) & 1), c.k0 + (q >> 2));
            ok = ok && r >= 0;
            rec[q] = r < 0 ? 0 : r;
        }
    }
    const float4* pn = Src == kDenseGather ? g.pn : g.spn;
#pragma unroll
    for (int q = 0; q < 8; ++q) f[q] = pn[rec[q]];
    return ok;
}

// Points i0, i0 + stride, ... (Pts of them, each processed when below its
// end) of one grid: every point's corner loads are issued before any point
// is interpolated.
template <typename TP, typename T, int Mode, int Src, int Pts = 2>
__device__ __forceinline__ void pairQuery(const QGrid& g, const P3<TP>* __restrict__ pts,
                                          int64_t i0, int64_t stride, int64_t end,
                                          TP* __restrict__ gap, P3<TP>* __restrict__ grad3,
                                          uint8_t* __restrict__ valid, P3<TP>* __restrict__ normal3,
                                          P3<TP>* __restrict__ witness3,
                                          int32_t* __restrict__ branch)
{
    double pc[Pts][3];
    dev::Cell c[Pts];
    float4 f[Pts][8];
#pragma unroll
    for (int h = 0; h < Pts; ++h)
    {
        const int64_t i = i0 + h * stride;
        P3<TP> p{TP(0), TP(0), TP(0)};
        if (i < end) p = pts[i];
        pc[h][0] = double(p.x);
        pc[h][1] = double(p.y);
        pc[h][2] = double(p.z);
        c[h] = dev::locateCell(g, pc[h][0], pc[h][1], pc[h][2]);
        c[h].valid = c[h].valid && i < end;
        if (c[h].valid)
        {
            if (Src == kDenseRec)
            {
                const float4* r = g.cellRec + 8 * (int64_t(c[h].i0) +
                                                   int64_t(g.n[0] - 1) * (c[h].j0 + int64_t(g.n[1] - 1) * c[h].k0));
#pragma unroll
                for (int q = 0; q < 8; ++q) f[h][q] = r[q];
            }
            else if (Src == kAnyRec)
            {
                if (const float4* r = dev::cellRecord(g, c[h].i0, c[h].j0, c[h].k0))
                {
#pragma unroll
                    for (int q = 0; q < 8; ++q) f[h][q] = r[q];
                }
                else
                    c[h].valid = false; // a missing corner sample: an invalid result
            }
            else if (!gatherCorners<Src>(g, c[h], f[h]))
                c[h].valid = false;
        }
    }
#pragma unroll
    for (int h = 0; h < Pts; ++h)
    {
        const int64_t i = i0 + h * stride;
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

// The chip's 8 XCDs (each with its own L2) take a launch's workgroups in
// turn: workgroup b runs on XCD b % 8.  Renumbering so that XCD x takes a
// contiguous eighth of the work (x * q + min(x, r) + b / 8, a bijection of
// [0, G) for G = 8 q + r) keeps a spatially coherent point stream's grid
// reads within one XCD's L2 instead of every XCD reading the whole grid.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcdBlock(int64_t b, int64_t G)
{
    const int64_t q = G / kXcds, r = G % kXcds, x = b % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}

// Trilinear / Tricubic, Pts points per thread (two by default): every
// point's corner samples (one 128-byte record each, or eight mirror loads)
// are requested before any is interpolated, multiplying the loads in flight
// per lane (the gather is latency-bound at one point per lane).  Same
// arithmetic as queryKernel.
template <typename TP, typename T, int Mode, int Src, int Block = kPairBlock, int Pts = kLanePoints>
__global__ void __launch_bounds__(Block)
queryPairKernel(QGrid g, const P3<TP>* __restrict__ pts, int64_t n, TP* __restrict__ gap,
                P3<TP>* __restrict__ grad3, uint8_t* __restrict__ valid,
                P3<TP>* __restrict__ normal3, P3<TP>* __restrict__ witness3,
                int32_t* __restrict__ branch)
{
    const int64_t i0 = xcdBlock(blockIdx.x, gridDim.x) * (Pts * Block) + threadIdx.x;
    pairQuery<TP, T, Mode, Src, Pts>(g, pts, i0, Block, n, gap, grad3, valid, normal3, witness3,
                                     branch);
}

// Launch shapes of the record-path query the tuner chooses from (the default
// first): points per lane x block size.  Every variant computes the same
// outputs.
constexpr int kPairBlockCount = 4;
constexpr int kPairBlocks[kPairBlockCount] = {128, 64, 256, 512};
constexpr int kPairShapeCount = 2 * kPairBlockCount;
// On a grid with cell records the tuner also tries the per-corner gather from
// the FP32 mirror (the records' eighth of the bytes, which neighbouring
// points share through the caches when the stream is spatially coherent).
constexpr int kPairVariantCount = 2 * kPairShapeCount;
const char* const kPairVariantNames[kPairVariantCount] = {
    "rec/2pts/block128",    "rec/2pts/block64",    "rec/2pts/block256",    "rec/2pts/block512",
    "rec/1pt/block128",     "rec/1pt/block64",     "rec/1pt/block256",     "rec/1pt/block512",
    "gather/2pts/block128", "gather/2pts/block64", "gather/2pts/block256", "gather/2pts/block512",
    "gather/1pt/block128",  "gather/1pt/block64",  "gather/1pt/block256",  "gather/1pt/block512"};

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
    const int64_t u = xcdBlock(blockIdx.x, gridDim.x) * (Block / 64) + (threadIdx.x >> 6);
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
    if (cubic || Mode == SDFO_TRILINEAR)
    {
        const bool rec = D.g.sparse ? D.g.scellRec != nullptr : D.g.cellRec != nullptr;
        auto run = [&](auto modeTag) {
            constexpr int M = decltype(modeTag)::value;
            if (rec)
                pairQuery<float, float, M, kAnyRec>(D.g, pts, i, 64, D.pointEnd, gap, grad3, valid,
                                                    nullptr, nullptr, nullptr);
            else if (D.g.sparse)
                pairQuery<float, float, M, kSparseGather>(D.g, pts, i, 64, D.pointEnd, gap, grad3,
                                                          valid, nullptr, nullptr, nullptr);
            else
                pairQuery<float, float, M, kDenseGather>(D.g, pts, i, 64, D.pointEnd, gap, grad3,
                                                         valid, nullptr, nullptr, nullptr);
        };
        if (cubic)
            run(std::integral_constant<int, SDFO_TRICUBIC>{});
        else
            run(std::integral_constant<int, SDFO_TRILINEAR>{});
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
    if (cubic || mode == SDFO_TRILINEAR)
    {
        const bool rec = q.sparse ? q.scellRec != nullptr : q.cellRec != nullptr;
        const int src = rec ? (q.sparse ? kAnyRec : kDenseRec) : (q.sparse ? kSparseGather : kDenseGather);
        auto launchBlock = [&](auto srcTag, auto blockTag, auto ptsTag) {
            constexpr int Src = decltype(srcTag)::value;
            constexpr int B = decltype(blockTag)::value;
            constexpr int Pts = decltype(ptsTag)::value;
            const dim3 pg(unsigned((n + Pts * B - 1) / (Pts * B))), pb(B);
            if (cubic)
                hipLaunchKernelGGL((queryPairKernel<TP, T, SDFO_TRICUBIC, Src, B, Pts>), pg, pb, 0, s,
                                   q, p, n, gap, gr, valid, nm, wt, branch);
            else
                hipLaunchKernelGGL((queryPairKernel<TP, T, SDFO_TRILINEAR, Src, B, Pts>), pg, pb, 0, s,
                                   q, p, n, gap, gr, valid, nm, wt, branch);
        };
        const int nVariants = rec ? kPairVariantCount : kPairShapeCount;
        auto launch = [&](int variant) {
            if (variant < 0 || variant >= nVariants) variant = 0;
            const int from = variant >= kPairShapeCount ? (q.sparse ? kSparseGather : kDenseGather) : src;
            variant %= kPairShapeCount;
            auto withPts = [&](auto srcTag, auto ptsTag) {
                switch (variant % kPairBlockCount)
                {
                case 1: launchBlock(srcTag, std::integral_constant<int, kPairBlocks[1]>{}, ptsTag); break;
                case 2: launchBlock(srcTag, std::integral_constant<int, kPairBlocks[2]>{}, ptsTag); break;
                case 3: launchBlock(srcTag, std::integral_constant<int, kPairBlocks[3]>{}, ptsTag); break;
                default: launchBlock(srcTag, std::integral_constant<int, kPairBlocks[0]>{}, ptsTag); break;
                }
            };
            auto withSrc = [&](auto srcTag) {
                if (variant >= kPairBlockCount)
                    withPts(srcTag, std::integral_constant<int, 1>{});
                else
                    withPts(srcTag, std::integral_constant<int, kLanePoints>{});
            };
            switch (from)
            {
            case kDenseRec: withSrc(std::integral_constant<int, kDenseRec>{}); break;
            case kAnyRec: withSrc(std::integral_constant<int, kAnyRec>{}); break;
            case kDenseGather: withSrc(std::integral_constant<int, kDenseGather>{}); break;
            default: withSrc(std::integral_constant<int, kSparseGather>{}); break;
            }
            SDFO_HIP(hipGetLastError());
        };
        // the tuner times every launch shape on this call's own points (each
        // run writes the same outputs) the first time a size class is seen
        const std::string key = std::string("query ") + (q.sparse ? "sparse " : "dense ") +
                                (rec ? "" : "gather ") + (cubic ? "tricubic" : "trilinear") +
                                (sizeof(TP) == 8 ? " f64" : " f32") + " points~2^" +
                                std::to_string(log2Bucket(n));
        int variant = 0;
        if (!tunerLookup(key, variant) && tunerEnabled() && !streamCapturing(s))
        {
            launch(0); // warm the first launch out of the timings
            const std::vector<std::pair<int, double>> t =
                timeVariants(s, nVariants, [&](int v) { launch(v); });
            tunerRecord(key, t, rec ? kPairVariantNames : kPairVariantNames + kPairShapeCount);
            tunerLookup(key, variant);
        }
        launch(variant);
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

namespace {
// The memory system's floor for the record-path query: every lane reads two
// uniformly random records of F4 float4 (the query's 128-byte cell record:
// F4 = 8) from an array, in 128-lane blocks as queryPairKernel, and keeps
// only a sum of their words.
__device__ __forceinline__ uint32_t mixBits(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

template <int F4>
__global__ void __launch_bounds__(kPairBlock)
randomRecordReadKernel(const float4* __restrict__ a, uint32_t nrec, uint32_t seed, int64_t n,
                       float* __restrict__ out)
{
    const int64_t i0 = int64_t(blockIdx.x) * (2 * kPairBlock) + threadIdx.x;
    float4 v[2][F4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
        const uint32_t r = mixBits(uint32_t(i0 + h * kPairBlock) * 2654435761u + seed) % nrec;
#pragma unroll
        for (int q = 0; q < F4; ++q) v[h][q] = a[size_t(r) * F4 + q];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
        float sum = 0.f;
#pragma unroll
        for (int q = 0; q < F4; ++q) sum += (v[h][q].x + v[h][q].y) + (v[h][q].z + v[h][q].w);
        const int64_t i = i0 + h * kPairBlock;
        if (i < n) out[i] = sum;
    }
}
} // namespace

double randomRecordReadUs(int recordBytes, int64_t nReads, int64_t arrayBytes, int reps)
{
    if ((recordBytes != 64 && recordBytes != 128) || nReads <= 0 || reps <= 0 ||
        arrayBytes < recordBytes || arrayBytes / recordBytes > int64_t(0xffffffffu))
        throw Error(SDFO_ERR_INVALID_ARGUMENT, "random read floor: bad sizes");
    const uint32_t nrec = uint32_t(arrayBytes / recordBytes);
    float4* a = nullptr;
    float* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double best = 1e30;
    try
    {
        SDFO_HIP(hipMalloc(&a, size_t(nrec) * size_t(recordBytes)));
        SDFO_HIP(hipMalloc(&out, sizeof(float) * size_t(nReads)));
        SDFO_HIP(hipMemset(a, 0, size_t(nrec) * size_t(recordBytes)));
        SDFO_HIP(hipEventCreate(&e0));
        SDFO_HIP(hipEventCreate(&e1));
        const dim3 grid(unsigned((nReads + 2 * kPairBlock - 1) / (2 * kPairBlock))), block(kPairBlock);
        for (int r = 0; r <= reps; ++r) // the first launch warms up
        {
            SDFO_HIP(hipEventRecord(e0, nullptr));
            if (recordBytes == 128)
                hipLaunchKernelGGL(randomRecordReadKernel<8>, grid, block, 0, nullptr, a, nrec,
                                   1234u + 7919u * uint32_t(r), nReads, out);
            else
                hipLaunchKernelGGL(randomRecordReadKernel<4>, grid, block, 0, nullptr, a, nrec,
                                   1234u + 7919u * uint32_t(r), nReads, out);
            SDFO_HIP(hipGetLastError());
            SDFO_HIP(hipEventRecord(e1, nullptr));
            SDFO_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            SDFO_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0 && ms * 1e3 < best) best = ms * 1e3;
        }
    }
    catch (...)
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (a) (void)hipFree(a);
        if (out) (void)hipFree(out);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(out);
    return best;
}

void launchQueryF32(const sdfo_grid& g, int mode, const float* pts, int64_t n, float* gap,
                    float* grad3, uint8_t* valid, float* normal3, float* witness3,
                    int32_t* branch, hipStream_t s)
{
    launchQuery<float, float>(g, mode, pts, n, gap, grad3, valid, normal3, witness3, branch, s);
}

const double4* cellRecords64(const sdfo_grid& g, hipStream_t s)
{
    sdfo_grid::Records64& R = g.rec64;
    if (g.sparse || !g.cellRec) return nullptr;
    if (R.done.load(std::memory_order_acquire)) return R.rec;
    std::lock_guard<std::mutex> lk(R.lock);
    if (!R.tried)
    {
        if (streamCapturing(s)) return nullptr; // no allocation inside a capture
        R.tried = true;
        const int64_t cells = int64_t(g.spec.n[0] - 1) * (g.spec.n[1] - 1) * (g.spec.n[2] - 1);
        const int64_t bytes = cells * 8 * int64_t(sizeof(double4));
        if (cells <= 0 || bytes > cellRecordLimit()) return nullptr;
        double4* rec = nullptr;
        if (hipMallocAsync(reinterpret_cast<void**>(&rec), size_t(bytes), s) != hipSuccess)
        {
            (void)hipGetLastError();
            return nullptr; // out of memory: the gather path
        }
        launchCellPack64(g.spec, g.phi, g.n3, rec, s);
        SDFO_HIP(hipEventCreateWithFlags(&R.ready, hipEventDisableTiming));
        SDFO_HIP(hipEventRecord(R.ready, s));
        R.rec = rec;
        R.bytes = bytes;
        return rec; // ordered on s
    }
    if (!R.rec) return nullptr;
    const hipError_t q = hipEventQuery(R.ready);
    if (q == hipSuccess)
    {
        R.done.store(true, std::memory_order_release);
        return R.rec;
    }
    if (q != hipErrorNotReady) SDFO_HIP(q);
    if (streamCapturing(s)) return nullptr; // still packing: gather, wait for nothing
    SDFO_HIP(hipStreamWaitEvent(s, R.ready, 0));
    return R.rec;
}

void launchQueryF64(const sdfo_grid& g, int mode, const double* pts, int64_t n, double* gap,
                    double* grad3, uint8_t* valid, double* normal3, double* witness3,
                    int32_t* branch, hipStream_t s)
{
    // FP64 points run the exact kernels: a dense grid's FP64 fields, a sparse
    // grid's FP32 fields widened, no FMA contraction
    launchQueryExactF64(g, mode, pts, n, gap, grad3, valid, normal3, witness3, branch, s);
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
