// Contact-aware SDF build (SdfBuilder.cpp:28-359, 416-821) for gfx950.
//
// Per voxel: scan the triangles of the voxel's tile (or all triangles), keep
// those within min distance + tolerance, merge candidates whose blended
// corner normals are within the merge angle (keeping the smaller |phi|),
// order by |phi| and keep the first maxCandidates.  FP64 without FMA
// contraction, so every candidate is the CPU path's bit for bit.
//
// The merged list is held per lane with room for kMergeCap entries; a voxel
// that needs more (smooth corner normals over many small triangles) is listed
// and re-run by a second pass whose lists live in device memory, sized to the
// longest triangle list, so no voxel is ever truncated.  The |phi| ordering repeats the steps of the
// std::sort the CPU path uses (GNU libstdc++ introsort: median-of-three
// partitions above 16 elements, heap sort when the depth budget runs out,
// then a final insertion sort), so candidates with equal |phi| come out in
// the same order as on the CPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <vector>

#include "absphi_sort.hpp"
#include "geom_device.hpp"

namespace sdfo {

using dev::D3;

void finishDense(sdfo_grid& g, hipStream_t s);
void finishSparse(sdfo_grid& g, const std::vector<int32_t>& bijk, hipStream_t s);

namespace {

constexpr int kBlock = 128;

// Triangle record: a, b, c, corner normals n0, n1, n2, face normal (21 doubles).
struct CTri
{
    D3 a, b, c, n0, n1, n2, face;
};

__device__ __forceinline__ CTri loadCTri(const double* __restrict__ t21, int t)
{
    const double* p = t21 + 21 * size_t(t);
    CTri r;
    r.a = dev::d3(p[0], p[1], p[2]);
    r.b = dev::d3(p[3], p[4], p[5]);
    r.c = dev::d3(p[6], p[7], p[8]);
    r.n0 = dev::d3(p[9], p[10], p[11]);
    r.n1 = dev::d3(p[12], p[13], p[14]);
    r.n2 = dev::d3(p[15], p[16], p[17]);
    r.face = dev::d3(p[18], p[19], p[20]);
    return r;
}

// SdfBuilder.cpp:28-43
__host__ __device__ inline void normalizedOrDefault(double vx, double vy, double vz, double fx,
                                                    double fy, double fz, double& ox, double& oy,
                                                    double& oz)
{
    const double n = sqrt(vx * vx + vy * vy + vz * vz);
    if (n <= 1.0e-15)
    {
        const double fn = sqrt(fx * fx + fy * fy + fz * fz);
        if (fn > 1.0e-15)
        {
            ox = fx / fn;
            oy = fy / fn;
            oz = fz / fn;
            return;
        }
        ox = 1.0;
        oy = 0.0;
        oz = 0.0;
        return;
    }
    ox = vx / n;
    oy = vy / n;
    oz = vz / n;
}

// SdfBuilder.cpp:45-67
__device__ __forceinline__ D3 barycentric(D3 p, D3 a, D3 b, D3 c)
{
    const D3 v0 = dev::sub(b, a), v1 = dev::sub(c, a), v2 = dev::sub(p, a);
    const double d00 = dev::dot(v0, v0), d01 = dev::dot(v0, v1), d11 = dev::dot(v1, v1);
    const double d20 = dev::dot(v2, v0), d21 = dev::dot(v2, v1);
    const double denom = d00 * d11 - d01 * d01;
    if (fabs(denom) <= 1.0e-15) return dev::d3(1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0);
    const double v = (d11 * d20 - d01 * d21) / denom;
    const double w = (d00 * d21 - d01 * d20) / denom;
    return dev::d3(1.0 - v - w, v, w);
}

// SdfBuilder.cpp:116-134: FNV-1a over the normal quantised to 1/4095.
__device__ __forceinline__ int32_t branchId(D3 n0)
{
    double nx, ny, nz;
    normalizedOrDefault(n0.x, n0.y, n0.z, 1.0, 0.0, 0.0, nx, ny, nz);
    const double q[3] = {nx, ny, nz};
    uint32_t hash = 2166136261u;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        double v = q[k];
        v = (1.0 < v) ? 1.0 : v;  // std::min(1.0, v)
        v = (v < -1.0) ? -1.0 : v; // std::max(-1.0, .)
        const int32_t qi = int32_t(llround(v * 4095.0));
        hash ^= uint32_t(qi);
        hash *= 16777619u;
    }
    return int32_t(1u + (hash & 0x3fffffffu));
}

struct Cand
{
    double phi, dist;
    D3 normal;
    D3 witness;
};

// makeContactCandidate (SdfBuilder.cpp:260-279).
__device__ __forceinline__ Cand makeCandidate(D3 p, const CTri& t)
{
    const D3 cp = dev::closestPoint(p, t.a, t.b, t.c);
    const D3 bc = barycentric(cp, t.a, t.b, t.c);
    const D3 blend = dev::add(dev::add(dev::scale(bc.x, t.n0), dev::scale(bc.y, t.n1)),
                              dev::scale(bc.z, t.n2));
    Cand c;
    normalizedOrDefault(blend.x, blend.y, blend.z, t.face.x, t.face.y, t.face.z, c.normal.x,
                        c.normal.y, c.normal.z);
    const D3 d = dev::sub(p, cp);
    c.phi = dev::dot(d, c.normal);
    c.dist = sqrt(dev::sqnorm(d));
    c.witness = cp;
    return c;
}

// Nearest triangle only (KdTree::findNearest, KdTree.h:123-133, 266-328).
__device__ inline int nearestSlot(D3 p, const BvhNode* __restrict__ nodes,
                                  const double* __restrict__ tri9, dev::LaneStack st, double& dist)
{
    double bestSq = DBL_MAX;
    int bestSlot = -1;
    int sp = 0;
    st.set(sp++, 0);
    while (sp > 0)
    {
        const int ni = st.get(--sp);
        const BvhNode nd = nodes[ni];
        if (!(dev::boxSq(p, nd) <= bestSq)) continue;
        if (nd.triCount > 0)
        {
            for (int s = nd.triStart; s < nd.triStart + nd.triCount; ++s)
            {
                const dev::Tri t = dev::loadTri(tri9, s);
                const double sq = dev::sqnorm(dev::sub(p, dev::closestPoint(p, t.a, t.b, t.c)));
                if (sq < bestSq)
                {
                    bestSq = sq;
                    bestSlot = s;
                }
            }
            continue;
        }
        const double dl = dev::boxSq(p, nodes[nd.left]);
        const double dr = dev::boxSq(p, nodes[nd.right]);
        if (dr < dl)
        {
            st.set(sp++, nd.left);
            st.set(sp++, nd.right);
        }
        else
        {
            st.set(sp++, nd.right);
            st.set(sp++, nd.left);
        }
    }
    dist = bestSlot >= 0 ? sqrt(bestSq) : DBL_MAX;
    return bestSlot;
}

struct CAParams
{
    double tol, cosMerge;
    int stride;
    int useTiles;
    int tileDim, tnx, tny, tnz;
    int ntris;
};

// Merged candidate list: entry e is (phi[e], n[e], tri[e]); order[i] is the
// entry of the i-th candidate after sorting.  Per-lane private arrays on the
// fast path, a device-memory slice sized to the longest triangle list on the
// overflow path.
struct MergeView
{
    double *phi, *nx, *ny, *nz;
    int *tri, *order;
    int cap;
    int count;
};

struct LocalMerge
{
    double phi[kMergeCap], nx[kMergeCap], ny[kMergeCap], nz[kMergeCap];
    int tri[kMergeCap], order[kMergeCap];
    __device__ __forceinline__ MergeView view() { return MergeView{phi, nx, ny, nz, tri, order, kMergeCap, 0}; }
};

struct MergePool
{
    double *phi, *nx, *ny, *nz;
    int *tri, *order;
    int cap;
    __device__ __forceinline__ MergeView view(int64_t k) const
    {
        const int64_t o = k * cap;
        return MergeView{phi + o, nx + o, ny + o, nz + o, tri + o, order + o, cap, 0};
    }
};

__device__ inline void sortByAbsPhi(MergeView& m)
{
    for (int i = 0; i < m.count; ++i) m.order[i] = i;
    sortPerm(m.order, m.count, AbsPhiOrder{m.phi});
}

// buildContactCandidatesForPoint (SdfBuilder.cpp:281-359).  Returns the
// merged count (> m.cap on overflow); candidates sorted by |phi|.
__device__ inline int candidatesForPoint(D3 p, const CAParams& P, const double* __restrict__ t21,
                                         const int32_t* __restrict__ list, int listLen,
                                         const BvhNode* __restrict__ nodes,
                                         const double* __restrict__ tri9,
                                         const int32_t* __restrict__ slotTri, dev::LaneStack st,
                                         MergeView& m)
{
    double minDist = DBL_MAX;
    for (int q = 0; q < listLen; ++q)
    {
        const int t = P.useTiles ? list[q] : q;
        const Cand c = makeCandidate(p, loadCTri(t21, t));
        minDist = (c.dist < minDist) ? c.dist : minDist;
    }
    // DBL_MAX is finite, so the reference's isfinite() fallback never fires for
    // finite input; an empty list reaches the nearest-triangle fallback below.
    const double allowed = minDist + P.tol;
    m.count = 0;
    auto tryTri = [&](int t) -> bool {
        const Cand c = makeCandidate(p, loadCTri(t21, t));
        if (!(c.dist <= allowed)) return true;
        for (int e = 0; e < m.count; ++e)
        {
            if (m.nx[e] * c.normal.x + m.ny[e] * c.normal.y + m.nz[e] * c.normal.z >= P.cosMerge)
            {
                if (fabs(c.phi) < fabs(m.phi[e]))
                {
                    m.phi[e] = c.phi;
                    m.nx[e] = c.normal.x;
                    m.ny[e] = c.normal.y;
                    m.nz[e] = c.normal.z;
                    m.tri[e] = t;
                }
                return true;
            }
        }
        if (m.count >= m.cap)
        {
            m.count = m.cap + 1;
            return false;
        }
        const int e = m.count++;
        m.phi[e] = c.phi;
        m.nx[e] = c.normal.x;
        m.ny[e] = c.normal.y;
        m.nz[e] = c.normal.z;
        m.tri[e] = t;
        return true;
    };
    for (int q = 0; q < listLen; ++q)
        if (!tryTri(P.useTiles ? list[q] : q)) return m.count;
    if (m.count == 0)
    {
        double d;
        const int slot = nearestSlot(p, nodes, tri9, st, d);
        if (slot >= 0) tryTri(slotTri[slot]);
    }
    sortByAbsPhi(m);
    return m.count;
}

struct CAData
{
    const double* t21;
    const int32_t* tileStart;
    const int32_t* tileTris;
    const BvhNode* nodes;
    const double* tri9;
    const int32_t* slotTri;
};

__device__ __forceinline__ void tileList(const CAParams& P, const CAData& D, int i, int j, int k,
                                         const int32_t*& list, int& len)
{
    if (!P.useTiles)
    {
        list = nullptr;
        len = P.ntris;
        return;
    }
    // ContactTileIndex::candidatesForVoxel (SdfBuilder.cpp:102-113)
    int ti = i / P.tileDim, tj = j / P.tileDim, tk = k / P.tileDim;
    ti = ti < 0 ? 0 : (ti > P.tnx - 1 ? P.tnx - 1 : ti);
    tj = tj < 0 ? 0 : (tj > P.tny - 1 ? P.tny - 1 : tj);
    tk = tk < 0 ? 0 : (tk > P.tnz - 1 ? P.tnz - 1 : tk);
    const int64_t tile = ti + int64_t(P.tnx) * (tj + int64_t(P.tny) * tk);
    list = D.tileTris + D.tileStart[tile];
    len = D.tileStart[tile + 1] - D.tileStart[tile];
}

struct DenseOut
{
    double *phi, *n3, *wit3;
    uint8_t* ccount;
    double *cphi, *cn3;
    int32_t* cbranch;
};

struct SparseOut
{
    const int32_t* bijk;
    float *sphi, *sn3;
    uint8_t* sccount;
    float *scphi, *scn3;
    int32_t* scbranch;
};

struct Overflow
{
    int* count;   // voxels whose merged list exceeded the private capacity
    int64_t* ids; // their ids (dense linear index or brick-pool slot)
};

// Writes one dense voxel (SdfBuilder.cpp:498-519; defaults SdfData.h:52-74).
__device__ inline void writeDense(const DenseOut& o, int64_t idx, D3 c, const MergeView& m,
                                  const CAParams& P, const double* __restrict__ t21)
{
    const int keep = m.count < P.stride ? m.count : P.stride;
    o.ccount[idx] = uint8_t(keep);
    double ph0 = 0.0;
    D3 n0 = dev::d3(1.0, 0.0, 0.0), w0 = dev::d3(0.0, 0.0, 0.0);
    if (keep > 0)
    {
        const int e = m.order[0];
        ph0 = m.phi[e];
        n0 = dev::d3(m.nx[e], m.ny[e], m.nz[e]);
        const CTri tr = loadCTri(t21, m.tri[e]);
        w0 = dev::closestPoint(c, tr.a, tr.b, tr.c);
    }
    o.phi[idx] = ph0;
    o.n3[3 * idx] = n0.x;
    o.n3[3 * idx + 1] = n0.y;
    o.n3[3 * idx + 2] = n0.z;
    o.wit3[3 * idx] = w0.x;
    o.wit3[3 * idx + 1] = w0.y;
    o.wit3[3 * idx + 2] = w0.z;
    for (int q = 0; q < P.stride; ++q)
    {
        const int64_t k = idx * P.stride + q;
        if (q < keep)
        {
            const int e = m.order[q];
            o.cphi[k] = m.phi[e];
            o.cn3[3 * k] = m.nx[e];
            o.cn3[3 * k + 1] = m.ny[e];
            o.cn3[3 * k + 2] = m.nz[e];
            o.cbranch[k] = branchId(dev::d3(m.nx[e], m.ny[e], m.nz[e]));
        }
        else
        {
            o.cphi[k] = 0.0;
            o.cn3[3 * k] = 1.0;
            o.cn3[3 * k + 1] = 0.0;
            o.cn3[3 * k + 2] = 0.0;
            o.cbranch[k] = -1;
        }
    }
}

// Writes one in-grid sparse voxel (SdfBuilder.cpp:764-801).
__device__ inline void writeSparse(const SparseOut& o, int64_t t, const MergeView& m,
                                   const CAParams& P)
{
    const int keep = m.count < P.stride ? m.count : P.stride;
    o.sccount[t] = uint8_t(keep);
    if (keep > 0)
    {
        const int e = m.order[0];
        o.sphi[t] = float(m.phi[e]);
        o.sn3[3 * t] = float(m.nx[e]);
        o.sn3[3 * t + 1] = float(m.ny[e]);
        o.sn3[3 * t + 2] = float(m.nz[e]);
    }
    else
    {
        o.sphi[t] = 0.0f;
        o.sn3[3 * t] = 1.0f;
        o.sn3[3 * t + 1] = 0.0f;
        o.sn3[3 * t + 2] = 0.0f;
    }
    for (int q = 0; q < keep; ++q)
    {
        const int64_t k = t * P.stride + q;
        const int e = m.order[q];
        o.scphi[k] = float(m.phi[e]);
        o.scn3[3 * k] = float(m.nx[e]);
        o.scn3[3 * k + 1] = float(m.ny[e]);
        o.scn3[3 * k + 2] = float(m.nz[e]);
        o.scbranch[k] = branchId(dev::d3(m.nx[e], m.ny[e], m.nz[e]));
    }
}

__device__ __forceinline__ D3 voxelCentre(const GridSpec& g, int i, int j, int k)
{
    return dev::d3(g.bmin[0] + g.voxel[0] * (i + 0.5), g.bmin[1] + g.voxel[1] * (j + 0.5),
                   g.bmin[2] + g.voxel[2] * (k + 0.5));
}

// Dense candidate fill (SdfBuilder.cpp:477-534): one lane per voxel, a 4x4x4
// voxel brick per wave (all lanes of a wave share one tile's triangle list).
// Voxels whose merged list outgrows the private arrays go to the overflow list.
__global__ void __launch_bounds__(kBlock)
denseCandidateKernel(GridSpec g, CAParams P, CAData D, int bx, int by, int bz, DenseOut o,
                     Overflow ov)
{
    __shared__ int stack[kStackDepth * kBlock];
    const int lane = threadIdx.x & 63;
    const int64_t brick = int64_t(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
    if (brick >= int64_t(bx) * by * bz) return;
    const int i = int(brick % bx) * 4 + (lane & 3);
    const int j = int((brick / bx) % by) * 4 + ((lane >> 2) & 3);
    const int k = int(brick / (int64_t(bx) * by)) * 4 + (lane >> 4);
    if (i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) return;
    const D3 c = voxelCentre(g, i, j, k);
    const int32_t* list;
    int len;
    tileList(P, D, i, j, k, list, len);
    LocalMerge lm;
    MergeView m = lm.view();
    dev::LaneStack st{stack + threadIdx.x, kBlock};
    const int cnt = candidatesForPoint(c, P, D.t21, list, len, D.nodes, D.tri9, D.slotTri, st, m);
    const int64_t idx = i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k);
    if (cnt > m.cap)
    {
        ov.ids[atomicAdd(ov.count, 1)] = idx;
        return;
    }
    writeDense(o, idx, c, m, P, D.t21);
}

// Overflow voxels of the dense fill, one lane each, merged lists in device memory.
__global__ void __launch_bounds__(kBlock)
denseOverflowKernel(GridSpec g, CAParams P, CAData D, const int64_t* __restrict__ ids, int n,
                    MergePool pool, DenseOut o)
{
    __shared__ int stack[kStackDepth * kBlock];
    const int q = int(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= n) return;
    const int64_t idx = ids[q];
    const int i = int(idx % g.n[0]), j = int((idx / g.n[0]) % g.n[1]),
              k = int(idx / (int64_t(g.n[0]) * g.n[1]));
    const D3 c = voxelCentre(g, i, j, k);
    const int32_t* list;
    int len;
    tileList(P, D, i, j, k, list, len);
    MergeView m = pool.view(q);
    dev::LaneStack st{stack + threadIdx.x, kBlock};
    candidatesForPoint(c, P, D.t21, list, len, D.nodes, D.tri9, D.slotTri, st, m);
    writeDense(o, idx, c, m, P, D.t21);
}

// Sparse candidate fill (SdfBuilder.cpp:731-803): one lane per brick voxel.
__global__ void __launch_bounds__(kBlock)
sparseCandidateKernel(GridSpec g, CAParams P, CAData D, int bd, int64_t total, SparseOut o,
                      Overflow ov)
{
    __shared__ int stack[kStackDepth * kBlock];
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int bvc = bd * bd * bd;
    const int64_t b = t / bvc;
    const int local = int(t % bvc);
    const int i = o.bijk[3 * b] * bd + local % bd;
    const int j = o.bijk[3 * b + 1] * bd + (local / bd) % bd;
    const int k = o.bijk[3 * b + 2] * bd + local / (bd * bd);
    // block allocation defaults (SdfBuilder.cpp:697-714)
    for (int q = 0; q < P.stride; ++q)
    {
        const int64_t e = t * P.stride + q;
        o.scphi[e] = 0.0f;
        o.scn3[3 * e] = 0.0f;
        o.scn3[3 * e + 1] = 0.0f;
        o.scn3[3 * e + 2] = 0.0f;
        o.scbranch[e] = -1;
    }
    o.sccount[t] = 0;
    if (i < 0 || i >= g.n[0] || j < 0 || j >= g.n[1] || k < 0 || k >= g.n[2])
    {
        o.sphi[t] = 0.0f;
        o.sn3[3 * t] = 1.0f;
        o.sn3[3 * t + 1] = 0.0f;
        o.sn3[3 * t + 2] = 0.0f;
        return;
    }
    const int32_t* list;
    int len;
    tileList(P, D, i, j, k, list, len);
    LocalMerge lm;
    MergeView m = lm.view();
    dev::LaneStack st{stack + threadIdx.x, kBlock};
    const int cnt = candidatesForPoint(voxelCentre(g, i, j, k), P, D.t21, list, len, D.nodes,
                                       D.tri9, D.slotTri, st, m);
    if (cnt > m.cap)
    {
        ov.ids[atomicAdd(ov.count, 1)] = t;
        return;
    }
    writeSparse(o, t, m, P);
}

__global__ void __launch_bounds__(kBlock)
sparseOverflowKernel(GridSpec g, CAParams P, CAData D, int bd, const int64_t* __restrict__ ids,
                     int n, MergePool pool, SparseOut o)
{
    __shared__ int stack[kStackDepth * kBlock];
    const int q = int(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= n) return;
    const int64_t t = ids[q];
    const int bvc = bd * bd * bd;
    const int64_t b = t / bvc;
    const int local = int(t % bvc);
    const int i = o.bijk[3 * b] * bd + local % bd;
    const int j = o.bijk[3 * b + 1] * bd + (local / bd) % bd;
    const int k = o.bijk[3 * b + 2] * bd + local / (bd * bd);
    const int32_t* list;
    int len;
    tileList(P, D, i, j, k, list, len);
    MergeView m = pool.view(q);
    dev::LaneStack st{stack + threadIdx.x, kBlock};
    candidatesForPoint(voxelCentre(g, i, j, k), P, D.t21, list, len, D.nodes, D.tri9, D.slotTri, st,
                       m);
    writeSparse(o, t, m, P);
}

// Sparse activation without tiles: unsigned nearest distance at the brick
// centre within activation + brick radius (SdfBuilder.cpp:665-682).
__global__ void __launch_bounds__(kBlock)
sparseCAActivateKernel(GridSpec g, const BvhNode* __restrict__ nodes,
                       const double* __restrict__ tri9, int bd, int nbx, int nby, int nbz,
                       double limit, uint8_t* __restrict__ mask)
{
    __shared__ int stack[kStackDepth * kBlock];
    const int64_t bl = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (bl >= int64_t(nbx) * nby * nbz) return;
    const int bi = int(bl % nbx), bj = int((bl / nbx) % nby), bk = int(bl / (int64_t(nbx) * nby));
    const int i0 = bi * bd, j0 = bj * bd, k0 = bk * bd;
    const int i1 = min(i0 + bd, g.n[0]), j1 = min(j0 + bd, g.n[1]), k1 = min(k0 + bd, g.n[2]);
    const D3 c = dev::d3(g.bmin[0] + g.voxel[0] * (0.5 * double(i0 + i1)),
                         g.bmin[1] + g.voxel[1] * (0.5 * double(j0 + j1)),
                         g.bmin[2] + g.voxel[2] * (0.5 * double(k0 + k1)));
    dev::LaneStack st{stack + threadIdx.x, kBlock};
    double d;
    nearestSlot(c, nodes, tri9, st, d);
    mask[bl] = d <= limit ? 1 : 0;
}

inline unsigned blocksFor(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }

// Host side -------------------------------------------------------------------

struct HostCA
{
    std::vector<double> t21;
    std::vector<double> bmin, bmax; // 3 per triangle
};

// makeContactTriangleInfos (SdfBuilder.cpp:155-184)
HostCA makeTriangleInfos(const double* verts, int64_t nv, const int32_t* tris, int64_t nt,
                         const double* cn)
{
    HostCA h;
    h.t21.resize(size_t(nt) * 21);
    h.bmin.resize(size_t(nt) * 3);
    h.bmax.resize(size_t(nt) * 3);
    for (int64_t t = 0; t < nt; ++t)
    {
        const int32_t* ix = tris + 3 * t;
        for (int c = 0; c < 3; ++c)
            if (ix[c] < 0 || ix[c] >= nv)
                throw Error(SDFO_ERR_INVALID_ARGUMENT,
                            "Contact-aware mesh triangle references invalid vertex");
        double* o = &h.t21[21 * size_t(t)];
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) o[3 * c + k] = verts[3 * size_t(ix[c]) + k];
        for (int q = 0; q < 9; ++q) o[9 + q] = cn[9 * size_t(t) + q];
        const double e1[3] = {o[3] - o[0], o[4] - o[1], o[5] - o[2]};
        const double e2[3] = {o[6] - o[0], o[7] - o[1], o[8] - o[2]};
        const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                              e1[0] * e2[1] - e1[1] * e2[0]};
        normalizedOrDefault(cr[0], cr[1], cr[2], 1.0, 0.0, 0.0, o[18], o[19], o[20]);
        for (int k = 0; k < 3; ++k)
        {
            const double a = o[k], b = o[3 + k], c = o[6 + k];
            double lo = (b < a) ? b : a;
            lo = (c < lo) ? c : lo;
            double hi = (a < b) ? b : a;
            hi = (hi < c) ? c : hi;
            h.bmin[3 * size_t(t) + k] = lo;
            h.bmax[3 * size_t(t) + k] = hi;
        }
    }
    return h;
}

struct HostTiles
{
    int tileDim = 1, nx = 1, ny = 1, nz = 1;
    std::vector<int32_t> start, tris;
    bool empty(int64_t tile) const { return start[tile + 1] == start[tile]; }
};

// buildContactTileIndex (SdfBuilder.cpp:202-258), as CSR lists in triangle order.
HostTiles buildTiles(const GridSpec& g, const HostCA& ca, int64_t nt, int reqTileDim, double infl)
{
    HostTiles T;
    T.tileDim = std::max(1, reqTileDim);
    T.nx = (g.n[0] + T.tileDim - 1) / T.tileDim;
    T.ny = (g.n[1] + T.tileDim - 1) / T.tileDim;
    T.nz = (g.n[2] + T.tileDim - 1) / T.tileDim;
    const int64_t ntiles = int64_t(T.nx) * T.ny * T.nz;
    const double h = g.hmax();
    const double tileRadius = 0.5 * std::sqrt(3.0) * double(T.tileDim) * h;
    const double expansion = (0.0 < infl ? infl : 0.0) + tileRadius + h;
    std::vector<int> ranges(size_t(nt) * 6);
    std::vector<int32_t> count(size_t(ntiles) + 1, 0);
    auto clampi = [](int v, int hi) { return std::max(0, std::min(hi, v)); };
    for (int64_t t = 0; t < nt; ++t)
    {
        int r[6];
        for (int a = 0; a < 3; ++a)
        {
            const double lo = ca.bmin[3 * size_t(t) + a] - expansion;
            const double hi = ca.bmax[3 * size_t(t) + a] + expansion;
            const int v0 = int(std::floor((lo - g.bmin[a]) / g.voxel[a] - 0.5));
            const int v1 = int(std::ceil((hi - g.bmin[a]) / g.voxel[a] - 0.5));
            r[a] = clampi(v0, g.n[a] - 1) / T.tileDim;
            r[3 + a] = clampi(v1, g.n[a] - 1) / T.tileDim;
        }
        for (int q = 0; q < 6; ++q) ranges[6 * size_t(t) + q] = r[q];
        for (int tk = r[2]; tk <= r[5]; ++tk)
            for (int tj = r[1]; tj <= r[4]; ++tj)
                for (int ti = r[0]; ti <= r[3]; ++ti) ++count[ti + int64_t(T.nx) * (tj + int64_t(T.ny) * tk)];
    }
    T.start.assign(size_t(ntiles) + 1, 0);
    for (int64_t q = 0; q < ntiles; ++q) T.start[q + 1] = T.start[q] + count[q];
    T.tris.resize(size_t(T.start[ntiles]));
    std::vector<int32_t> fill(T.start.begin(), T.start.end() - 1);
    for (int64_t t = 0; t < nt; ++t)
    {
        const int* r = &ranges[6 * size_t(t)];
        for (int tk = r[2]; tk <= r[5]; ++tk)
            for (int tj = r[1]; tj <= r[4]; ++tj)
                for (int ti = r[0]; ti <= r[3]; ++ti)
                    T.tris[fill[ti + int64_t(T.nx) * (tj + int64_t(T.ny) * tk)]++] = int32_t(t);
    }
    return T;
}

struct DevBuf
{
    std::vector<void*> ptrs;
    template <typename T>
    T* up(const std::vector<T>& h)
    {
        void* p = nullptr;
        SDFO_HIP(hipMalloc(&p, std::max<size_t>(1, h.size()) * sizeof(T)));
        ptrs.push_back(p);
        if (!h.empty()) SDFO_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return static_cast<T*>(p);
    }
    void* raw(size_t bytes)
    {
        void* p = nullptr;
        SDFO_HIP(hipMalloc(&p, std::max<size_t>(1, bytes)));
        ptrs.push_back(p);
        return p;
    }
    ~DevBuf()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

} // namespace

sdfo_grid* buildContactAware(const double* verts, int64_t nv, const int32_t* tris, int64_t nt,
                             const double* cornerN, const sdfo_build_params& p, bool sparse,
                             hipStream_t s)
{
    if (nv <= 0 || nt <= 0 || !verts || !tris)
        throw Error(SDFO_ERR_EMPTY_MESH, sparse ? "Empty mesh provided to sparse contact-aware SdfBuilder"
                                                : "Empty mesh provided to contact-aware SdfBuilder");
    if (!cornerN)
        throw Error(SDFO_ERR_INVALID_ARGUMENT,
                    "Contact-aware SDF requires one corner-normal triplet per triangle");
    if (p.max_candidates <= 0)
        throw Error(SDFO_ERR_INVALID_ARGUMENT,
                    sparse ? "Sparse contact-aware SDF maxCandidates must be positive"
                           : "Contact-aware SDF maxCandidates must be positive");
    if (sparse && p.block_dim <= 0)
        throw Error(SDFO_ERR_INVALID_ARGUMENT, "Sparse contact-aware SDF blockDim must be positive");
    auto g = std::make_unique<sdfo_grid>();
    SDFO_HIP(hipGetDevice(&g->device));
    g->spec = setupGrid(verts, nv, p);
    const GridSpec& G = g->spec;
    const int stride = std::max(0, std::min(int(p.max_candidates), 8));
    g->stride = stride;
    const double h = G.hmax();
    const double tol = p.feature_distance_tolerance > 0.0 ? p.feature_distance_tolerance : 1.5 * h;
    const double cosMerge = std::cos(p.normal_merge_angle_degrees * std::acos(-1.0) / 180.0);
    const HostCA ca = makeTriangleInfos(verts, nv, tris, nt, cornerN);
    const HostBvh hb = buildHostBvh(verts, nv, tris, nt);
    double band = 0.0, activation = 0.0;
    if (sparse)
    {
        band = p.band_distance > 0.0 ? p.band_distance : double(p.band_voxels) * h;
        activation = band + double(p.halo_voxels) * h;
    }
    const double infl = sparse ? std::max(activation, tol + 2.0 * h) : tol + 2.0 * h;
    HostTiles tiles;
    if (p.use_tile_acceleration) tiles = buildTiles(G, ca, nt, p.tile_dim, infl);

    DevBuf buf;
    DeviceBvh db;
    CAData D{};
    CAParams P{};
    P.tol = tol;
    P.cosMerge = cosMerge;
    P.stride = stride;
    P.useTiles = p.use_tile_acceleration ? 1 : 0;
    P.tileDim = tiles.tileDim;
    P.tnx = tiles.nx;
    P.tny = tiles.ny;
    P.tnz = tiles.nz;
    P.ntris = int(nt);
    // The merged list of a voxel holds at most one entry per triangle it tests
    // (its tile list, or every triangle without tiles), plus the nearest-triangle
    // fallback when that list yields nothing.
    int maxList = int(nt);
    if (P.useTiles)
    {
        maxList = 0;
        for (size_t q = 0; q + 1 < tiles.start.size(); ++q)
            maxList = std::max(maxList, tiles.start[q + 1] - tiles.start[q]);
    }
    const int poolCap = std::max(maxList, 1);
    Overflow ov{nullptr, nullptr};
    // Re-runs the voxels the first pass listed, with merged lists in device
    // memory, in batches of at most ~1 GiB of list storage.
    auto resolveOverflow = [&](auto launch) {
        int cnt = 0;
        SDFO_HIP(hipMemcpyAsync(&cnt, ov.count, sizeof(int), hipMemcpyDeviceToHost, s));
        SDFO_HIP(hipStreamSynchronize(s));
        if (cnt <= 0) return;
        const size_t entry = 4 * sizeof(double) + 2 * sizeof(int);
        const int64_t per = std::max<int64_t>(1, int64_t((size_t(1) << 30) / (entry * size_t(poolCap))));
        const int batch = int(std::min<int64_t>(cnt, per));
        DevBuf pb;
        const size_t e = size_t(batch) * size_t(poolCap);
        MergePool pool{};
        pool.cap = poolCap;
        pool.phi = static_cast<double*>(pb.raw(e * sizeof(double)));
        pool.nx = static_cast<double*>(pb.raw(e * sizeof(double)));
        pool.ny = static_cast<double*>(pb.raw(e * sizeof(double)));
        pool.nz = static_cast<double*>(pb.raw(e * sizeof(double)));
        pool.tri = static_cast<int*>(pb.raw(e * sizeof(int)));
        pool.order = static_cast<int*>(pb.raw(e * sizeof(int)));
        for (int off = 0; off < cnt; off += batch)
        {
            launch(ov.ids + off, std::min(batch, cnt - off), pool);
            SDFO_HIP(hipGetLastError());
        }
        SDFO_HIP(hipStreamSynchronize(s));
    };
    try
    {
        keepStreamPoolMemory(g->device);
        db.upload(hb, s);
        D.t21 = buf.up(ca.t21);
        if (P.useTiles)
        {
            D.tileStart = buf.up(tiles.start);
            D.tileTris = buf.up(tiles.tris);
        }
        D.nodes = db.nodes;
        D.tri9 = db.tri9;
        D.slotTri = db.triIdx;
        ov.count = buf.up(std::vector<int>{0});
        if (!sparse)
        {
            const int64_t N = G.total();
            g->phi = g->alloc<double>(size_t(N));
            g->n3 = g->alloc<double>(size_t(N) * 3);
            g->h6 = g->alloc<double>(size_t(N) * 6);
            g->witness3 = g->alloc<double>(size_t(N) * 3);
            g->ccount = g->alloc<uint8_t>(size_t(N));
            g->cphi = g->alloc<double>(size_t(N) * stride);
            g->cn3 = g->alloc<double>(size_t(N) * stride * 3);
            g->cbranch = g->alloc<int32_t>(size_t(N) * stride);
            const int bx = (G.n[0] + 3) / 4, by = (G.n[1] + 3) / 4, bz = (G.n[2] + 3) / 4;
            const DenseOut o{g->phi, g->n3, g->witness3, g->ccount, g->cphi, g->cn3, g->cbranch};
            ov.ids = static_cast<int64_t*>(buf.raw(size_t(N) * sizeof(int64_t)));
            hipLaunchKernelGGL(denseCandidateKernel, dim3(blocksFor(int64_t(bx) * by * bz, kBlock / 64)),
                               dim3(kBlock), 0, s, G, P, D, bx, by, bz, o, ov);
            SDFO_HIP(hipGetLastError());
            resolveOverflow([&](const int64_t* ids, int n, const MergePool& pool) {
                hipLaunchKernelGGL(denseOverflowKernel, dim3(blocksFor(n, kBlock)), dim3(kBlock), 0, s,
                                   G, P, D, ids, n, pool, o);
            });
            if (p.store_hessian)
                launchDenseHessian(G, g->n3, g->h6, s);
            else
                SDFO_HIP(hipMemsetAsync(g->h6, 0, size_t(N) * 6 * sizeof(double), s));
            finishDense(*g, s);
        }
        else
        {
            g->sparse = true;
            g->blockDim = p.block_dim;
            g->halo = p.halo_voxels;
            g->band = band;
            g->activation = activation;
            g->hasHessian = p.store_hessian != 0;
            const int bd = p.block_dim;
            g->nbx = (G.n[0] + bd - 1) / bd;
            g->nby = (G.n[1] + bd - 1) / bd;
            g->nbz = (G.n[2] + bd - 1) / bd;
            const int64_t cells = int64_t(g->nbx) * g->nby * g->nbz;
            std::vector<uint8_t> mask(static_cast<size_t>(cells), 0);
            g->activeMask = g->alloc<uint8_t>(size_t(cells));
            if (P.useTiles)
            {
                // SdfBuilder.cpp:638-664: any non-empty tile over the brick's voxel range
                for (int64_t bl = 0; bl < cells; ++bl)
                {
                    const int bi = int(bl % g->nbx), bj = int((bl / g->nbx) % g->nby),
                              bk = int(bl / (int64_t(g->nbx) * g->nby));
                    const int i0 = bi * bd, j0 = bj * bd, k0 = bk * bd;
                    const int i1 = std::min(i0 + bd - 1, G.n[0] - 1);
                    const int j1 = std::min(j0 + bd - 1, G.n[1] - 1);
                    const int k1 = std::min(k0 + bd - 1, G.n[2] - 1);
                    bool act = false;
                    for (int tk = k0 / tiles.tileDim; tk <= k1 / tiles.tileDim && !act; ++tk)
                        for (int tj = j0 / tiles.tileDim; tj <= j1 / tiles.tileDim && !act; ++tj)
                            for (int ti = i0 / tiles.tileDim; ti <= i1 / tiles.tileDim && !act; ++ti)
                                act = !tiles.empty(ti + int64_t(tiles.nx) * (tj + int64_t(tiles.ny) * tk));
                    mask[size_t(bl)] = act ? 1 : 0;
                }
                SDFO_HIP(hipMemcpy(g->activeMask, mask.data(), mask.size(), hipMemcpyHostToDevice));
            }
            else
            {
                const double br = 0.5 * std::sqrt(3.0) * double(bd) * h;
                hipLaunchKernelGGL(sparseCAActivateKernel, dim3(blocksFor(cells, kBlock)), dim3(kBlock),
                                   0, s, G, db.nodes, db.tri9, bd, g->nbx, g->nby, g->nbz,
                                   activation + br, g->activeMask);
                SDFO_HIP(hipGetLastError());
                SDFO_HIP(hipMemcpyAsync(mask.data(), g->activeMask, mask.size(), hipMemcpyDeviceToHost, s));
                SDFO_HIP(hipStreamSynchronize(s));
            }
            std::vector<int32_t> bijk;
            for (int64_t c = 0; c < cells; ++c)
            {
                if (!mask[size_t(c)]) continue;
                bijk.push_back(int32_t(c % g->nbx));
                bijk.push_back(int32_t((c / g->nbx) % g->nby));
                bijk.push_back(int32_t(c / (int64_t(g->nbx) * g->nby)));
            }
            g->nb = int(bijk.size() / 3);
            const int64_t total = int64_t(g->nb) * bd * bd * bd;
            if (g->nb > 0)
            {
                g->bijk = g->alloc<int32_t>(bijk.size());
                SDFO_HIP(hipMemcpy(g->bijk, bijk.data(), bijk.size() * sizeof(int32_t),
                                   hipMemcpyHostToDevice));
                g->sphi = g->alloc<float>(size_t(total));
                g->sn3 = g->alloc<float>(size_t(total) * 3);
                g->sccount = g->alloc<uint8_t>(size_t(total));
                g->scphi = g->alloc<float>(size_t(total) * stride);
                g->scn3 = g->alloc<float>(size_t(total) * stride * 3);
                g->scbranch = g->alloc<int32_t>(size_t(total) * stride);
                const SparseOut o{g->bijk, g->sphi, g->sn3, g->sccount, g->scphi, g->scn3, g->scbranch};
                ov.ids = static_cast<int64_t*>(buf.raw(size_t(total) * sizeof(int64_t)));
                hipLaunchKernelGGL(sparseCandidateKernel, dim3(blocksFor(total, kBlock)), dim3(kBlock), 0,
                                   s, G, P, D, bd, total, o, ov);
                SDFO_HIP(hipGetLastError());
                resolveOverflow([&](const int64_t* ids, int n, const MergePool& pool) {
                    hipLaunchKernelGGL(sparseOverflowKernel, dim3(blocksFor(n, kBlock)), dim3(kBlock), 0,
                                       s, G, P, D, bd, ids, n, pool, o);
                });
            }
            finishSparse(*g, bijk, s);
            if (g->hasHessian && g->nb > 0)
            {
                g->sh6 = g->alloc<float>(size_t(total) * 6);
                SDFO_HIP(hipMemsetAsync(g->sh6, 0, size_t(total) * 6 * sizeof(float), s));
                launchSparseHessian(G, bd, g->nb, g->nbx, g->nby, g->nbz, g->bijk, g->table, g->sn3,
                                    g->sh6, s);
            }
        }
        SDFO_HIP(hipStreamSynchronize(s));
    }
    catch (...)
    {
        (void)hipStreamSynchronize(s);
        db.release();
        throw;
    }
    db.release();
    return g.release();
}

} // namespace sdfo
