// Contact-aware SDF build (SdfBuilder.cpp:28-359, 416-821) for gfx950.
//
// Per voxel: scan the triangles of the voxel's tile (or all triangles), keep
// those within min distance + tolerance, merge candidates whose blended
// corner normals are within the merge angle (keeping the smaller |phi|),
// order by |phi| and keep the first maxCandidates.  FP64 without FMA
// contraction, so every candidate is the CPU path's bit for bit.
//
// The first pass holds each voxel's merged list in registers, with room for a
// few entries (fully unrolled, so nothing goes to scratch memory); a voxel
// that needs more (smooth corner normals over many small triangles) is listed
// and re-run by a second pass whose lists live in device memory, sized to the
// longest triangle list, so no voxel is ever truncated.  The |phi| ordering
// gives what the std::sort of the CPU path gives: up to 16 entries that sort
// is GNU libstdc++'s insertion sort, a stable sort, which the register pass
// computes as ranks; the second pass repeats the steps of the whole introsort
// (median-of-three partitions above 16 elements, heap sort when the depth
// budget runs out, then a final insertion sort), so candidates with equal
// |phi| come out in the same order as on the CPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <exception>
#include <thread>
#include <vector>

#include "absphi_sort.hpp"
#include "geom_device.hpp"

namespace sdfo {

using dev::D3;

int& debugMergeCap()
{
    thread_local int cap = 0;
    return cap;
}

int& debugSecondCap()
{
    thread_local int cap = 0;
    return cap;
}

namespace {

constexpr int kBlock = 128;
constexpr int kSecondCap = 64; // merged entries per voxel in the second pass's first launch

// Triangle record: a, b, c, corner normals n0, n1, n2, face normal (21 doubles).
struct CTri
{
    D3 a, b, c, n0, n1, n2, face;
};

// Vertices only (the distance test needs no normals).
__device__ __forceinline__ void loadCTriVerts(const double* __restrict__ t21, int t, CTri& r)
{
    const double* p = t21 + 21 * size_t(t);
    r.a = dev::d3(p[0], p[1], p[2]);
    r.b = dev::d3(p[3], p[4], p[5]);
    r.c = dev::d3(p[6], p[7], p[8]);
}

__device__ __forceinline__ void loadCTriNormals(const double* __restrict__ t21, int t, CTri& r)
{
    const double* p = t21 + 21 * size_t(t);
    r.n0 = dev::d3(p[9], p[10], p[11]);
    r.n1 = dev::d3(p[12], p[13], p[14]);
    r.n2 = dev::d3(p[15], p[16], p[17]);
    r.face = dev::d3(p[18], p[19], p[20]);
}

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

// The rest of makeContactCandidate once the closest point cp is known.
__device__ __forceinline__ void finishCandidate(D3 p, const CTri& t, D3 cp, Cand& c)
{
    const D3 bc = barycentric(cp, t.a, t.b, t.c);
    const D3 blend = dev::add(dev::add(dev::scale(bc.x, t.n0), dev::scale(bc.y, t.n1)),
                              dev::scale(bc.z, t.n2));
    normalizedOrDefault(blend.x, blend.y, blend.z, t.face.x, t.face.y, t.face.z, c.normal.x,
                        c.normal.y, c.normal.z);
    c.phi = dev::dot(dev::sub(p, cp), c.normal);
    c.witness = cp;
}

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

// Merged candidate list of the second pass: entry e is (phi[e], n[e], tri[e]);
// order[i] is the entry of the i-th candidate after sorting.  The lists of a
// launch live in device memory interleaved across lanes (entry e of lane q at
// e * lanes + q), so a wave's accesses to its lanes' entry e coalesce.
struct MergeView
{
    Strided<double> phi, nx, ny, nz;
    Strided<int> tri, order;
    int cap;
    int count;
};

// Merged list of the first pass: K entries in registers (every index below is
// a compile-time constant after unrolling); rank[e] is entry e's place in the
// |phi| order.
template <int K>
struct RegMerge
{
    double phi[K], nx[K], ny[K], nz[K];
    int tri[K], rank[K];
    int count;
};

struct MergePool
{
    double *phi, *nx, *ny, *nz;
    int *tri, *order;
    int cap;       // entries per list
    int64_t lanes; // lists per launch
    __device__ __forceinline__ MergeView view(int64_t q) const
    {
        return MergeView{{phi + q, lanes}, {nx + q, lanes}, {ny + q, lanes}, {nz + q, lanes},
                         {tri + q, lanes}, {order + q, lanes}, cap, 0};
    }
};

__device__ inline void sortByAbsPhi(MergeView& m)
{
    for (int i = 0; i < m.count; ++i) m.order[i] = i;
    sortPerm(m.order, m.count, AbsPhiOrderT<Strided<double>>{m.phi});
}

// buildContactCandidatesForPoint (SdfBuilder.cpp:281-359) over the voxel's
// triangle list, visited by each(f) in list order (f returns false to stop).
// Returns the merged count (> m.cap on overflow); candidates sorted by |phi|.
// haveMin: minDistIn is already the list's minimum distance.  Lanes with
// !active take part in each() (which may be wave-collective) and nothing else.
template <class Each>
__device__ inline int candidatesFor(D3 p, const CAParams& P, const double* __restrict__ t21,
                                    const Each& each, bool haveMin, double minDistIn, bool active,
                                    const BvhNode* __restrict__ nodes,
                                    const double* __restrict__ tri9,
                                    const int32_t* __restrict__ slotTri, dev::LaneStack st,
                                    MergeView& m)
{
    double minDist = minDistIn;
    if (!haveMin)
    {
        minDist = DBL_MAX;
        each([&](int t) {
            const Cand c = makeCandidate(p, loadCTri(t21, t));
            minDist = (c.dist < minDist) ? c.dist : minDist;
            return true;
        });
    }
    // DBL_MAX is finite, so the reference's isfinite() fallback never fires for
    // finite input; an empty list reaches the nearest-triangle fallback below.
    const double allowed = minDist + P.tol;
    m.count = 0;
    auto tryTri = [&](int t) -> bool {
        // makeCandidate's distance first: the rest only for triangles in reach
        CTri tr;
        loadCTriVerts(t21, t, tr);
        const D3 cp = dev::closestPoint(p, tr.a, tr.b, tr.c);
        Cand c;
        c.dist = sqrt(dev::sqnorm(dev::sub(p, cp)));
        if (!(c.dist <= allowed)) return true;
        loadCTriNormals(t21, t, tr);
        finishCandidate(p, tr, cp, c);
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
    each(tryTri);
    if (m.count > m.cap || !active) return m.count;
    if (m.count == 0)
    {
        double d;
        const int slot = nearestSlot(p, nodes, tri9, st, d);
        if (slot >= 0) tryTri(slotTri[slot]);
    }
    sortByAbsPhi(m);
    return m.count;
}

// candidatesFor on a register list of K entries.  Returns the merged count, or
// K + 1 when the list would outgrow K entries or an |phi| is NaN (the ranks
// below assume an ordered |phi|): such voxels go to the second pass.
template <int K, class Each>
__device__ inline int fastCandidates(D3 p, const CAParams& P, const double* __restrict__ t21,
                                     const Each& each, double minDist, bool active,
                                     const BvhNode* __restrict__ nodes,
                                     const double* __restrict__ tri9,
                                     const int32_t* __restrict__ slotTri, dev::LaneStack st,
                                     RegMerge<K>& m)
{
    const double allowed = minDist + P.tol;
    m.count = 0;
    auto tryTri = [&](int t) -> bool {
        CTri tr;
        loadCTriVerts(t21, t, tr);
        const D3 cp = dev::closestPoint(p, tr.a, tr.b, tr.c);
        const double dist = sqrt(dev::sqnorm(dev::sub(p, cp)));
        if (!(dist <= allowed)) return true;
        loadCTriNormals(t21, t, tr);
        Cand c;
        finishCandidate(p, tr, cp, c);
        // mergeContactCandidate: the first entry within the merge angle takes
        // the candidate if its |phi| is smaller
        bool merged = false;
#pragma unroll
        for (int e = 0; e < K; ++e)
        {
            if (!merged && e < m.count &&
                m.nx[e] * c.normal.x + m.ny[e] * c.normal.y + m.nz[e] * c.normal.z >= P.cosMerge)
            {
                merged = true;
                if (fabs(c.phi) < fabs(m.phi[e]))
                {
                    m.phi[e] = c.phi;
                    m.nx[e] = c.normal.x;
                    m.ny[e] = c.normal.y;
                    m.nz[e] = c.normal.z;
                    m.tri[e] = t;
                }
            }
        }
        if (merged) return true;
        if (m.count >= K)
        {
            m.count = K + 1;
            return false;
        }
#pragma unroll
        for (int e = 0; e < K; ++e)
        {
            if (e == m.count)
            {
                m.phi[e] = c.phi;
                m.nx[e] = c.normal.x;
                m.ny[e] = c.normal.y;
                m.nz[e] = c.normal.z;
                m.tri[e] = t;
            }
        }
        ++m.count;
        return true;
    };
    each(tryTri);
    if (m.count > K || !active) return m.count;
    if (m.count == 0)
    {
        double d;
        const int slot = nearestSlot(p, nodes, tri9, st, d);
        if (slot >= 0) tryTri(slotTri[slot]);
    }
    bool nan = false;
#pragma unroll
    for (int e = 0; e < K; ++e) nan = nan || (e < m.count && m.phi[e] != m.phi[e]);
    if (nan) return K + 1;
    // place of entry e in a stable sort by |phi|
#pragma unroll
    for (int e = 0; e < K; ++e)
    {
        int r = 0;
#pragma unroll
        for (int f = 0; f < K; ++f)
        {
            if (f == e || f >= m.count) continue;
            const double af = fabs(m.phi[f]), ae = fabs(m.phi[e]);
            r += (af < ae || (af == ae && f < e)) ? 1 : 0;
        }
        m.rank[e] = e < m.count ? r : K;
    }
    return m.count;
}

struct CAData
{
    const double* t21;
    const double* tbox6;      // 6 per triangle: bmin3, bmax3
    const int64_t* tileStart; // with tiles: CSR lists, triangle order within each tile
    const int32_t* tileTris;
    const int32_t* triTile;   // with tiles: 6 per triangle, the tile range whose lists hold it
    const int32_t* nodeTile;  // with tiles: 6 per BVH node, the union of its triangles' ranges
    const BvhNode* nodes;
    const double* tri9;
    const double* slotBox6; // 6 per BVH slot: the slot triangle's box
    const int32_t* slotTri;
    double mu; // search margin, far above the rounding of box and triangle distances
};

// The tile of a voxel (ContactTileIndex::candidatesForVoxel, SdfBuilder.cpp:102-113).
__device__ __forceinline__ void voxelTile(const CAParams& P, int i, int j, int k, int& ti, int& tj,
                                          int& tk)
{
    ti = i / P.tileDim;
    tj = j / P.tileDim;
    tk = k / P.tileDim;
    ti = ti < 0 ? 0 : (ti > P.tnx - 1 ? P.tnx - 1 : ti);
    tj = tj < 0 ? 0 : (tj > P.tny - 1 ? P.tny - 1 : tj);
    tk = tk < 0 ? 0 : (tk > P.tnz - 1 ? P.tnz - 1 : tk);
}

__device__ __forceinline__ bool inRange(const int32_t* __restrict__ r, int ti, int tj, int tk)
{
    return r[0] <= ti && ti <= r[3] && r[1] <= tj && tj <= r[4] && r[2] <= tk && tk <= r[5];
}

// Triangle t is on the list of tile (ti, tj, tk) exactly when its tile range
// holds the tile (buildContactTileIndex, SdfBuilder.cpp:202-258); without
// tiles every triangle is on every list.  Lists are in triangle order.
__device__ __forceinline__ bool onList(const CAParams& P, const CAData& D, int t, int ti, int tj,
                                       int tk)
{
    return !P.useTiles || inRange(D.triTile + 6 * int64_t(t), ti, tj, tk);
}

// Ascending sort of distinct triangle ids: insertion sort for short rows,
// heap sort for long ones.
__device__ inline void sortIds(int* a, int n)
{
    if (n <= 32)
    {
        for (int i = 1; i < n; ++i)
        {
            const int v = a[i];
            int j = i - 1;
            while (j >= 0 && a[j] > v)
            {
                a[j + 1] = a[j];
                --j;
            }
            a[j + 1] = v;
        }
        return;
    }
    auto sift = [&](int r, int end) {
        while (2 * r + 1 < end)
        {
            int c = 2 * r + 1;
            if (c + 1 < end && a[c] < a[c + 1]) ++c;
            if (!(a[r] < a[c])) return;
            const int t = a[r];
            a[r] = a[c];
            a[c] = t;
            r = c;
        }
    };
    for (int r = n / 2 - 1; r >= 0; --r) sift(r, n);
    for (int end = n - 1; end > 0; --end)
    {
        const int t = a[0];
        a[0] = a[end];
        a[end] = t;
        sift(0, end);
    }
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
    int* count;       // voxels whose merged list outgrew the pass's capacity
    int64_t* ids;     // their ids (dense linear index or brick-pool slot)
    double* minDist;  // their lists' minimum distance (DBL_MAX: empty list)
};

__device__ __forceinline__ void pushOverflow(const Overflow& ov, int64_t id, double minDist)
{
    const int slot = atomicAdd(ov.count, 1);
    ov.ids[slot] = id;
    ov.minDist[slot] = minDist;
}

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

// writeDense for a register list: entry e goes to slot rank[e].
template <int K>
__device__ inline void writeDenseReg(const DenseOut& o, int64_t idx, D3 c, const RegMerge<K>& m,
                                     const CAParams& P, const double* __restrict__ t21)
{
    const int keep = m.count < P.stride ? m.count : P.stride;
    o.ccount[idx] = uint8_t(keep);
    double ph0 = 0.0;
    D3 n0 = dev::d3(1.0, 0.0, 0.0), w0 = dev::d3(0.0, 0.0, 0.0);
    int t0 = -1;
#pragma unroll
    for (int e = 0; e < K; ++e)
    {
        if (m.rank[e] == 0 && keep > 0)
        {
            ph0 = m.phi[e];
            n0 = dev::d3(m.nx[e], m.ny[e], m.nz[e]);
            t0 = m.tri[e];
        }
    }
    if (t0 >= 0)
    {
        const CTri tr = loadCTri(t21, t0);
        w0 = dev::closestPoint(c, tr.a, tr.b, tr.c);
    }
    o.phi[idx] = ph0;
    o.n3[3 * idx] = n0.x;
    o.n3[3 * idx + 1] = n0.y;
    o.n3[3 * idx + 2] = n0.z;
    o.wit3[3 * idx] = w0.x;
    o.wit3[3 * idx + 1] = w0.y;
    o.wit3[3 * idx + 2] = w0.z;
#pragma unroll
    for (int e = 0; e < K; ++e)
    {
        if (m.rank[e] >= keep) continue;
        const int64_t k = idx * P.stride + m.rank[e];
        o.cphi[k] = m.phi[e];
        o.cn3[3 * k] = m.nx[e];
        o.cn3[3 * k + 1] = m.ny[e];
        o.cn3[3 * k + 2] = m.nz[e];
        o.cbranch[k] = branchId(dev::d3(m.nx[e], m.ny[e], m.nz[e]));
    }
    for (int q = keep; q < P.stride; ++q)
    {
        const int64_t k = idx * P.stride + q;
        o.cphi[k] = 0.0;
        o.cn3[3 * k] = 1.0;
        o.cn3[3 * k + 1] = 0.0;
        o.cn3[3 * k + 2] = 0.0;
        o.cbranch[k] = -1;
    }
}

// writeSparse for a register list: entry e goes to slot rank[e].
template <int K>
__device__ inline void writeSparseReg(const SparseOut& o, int64_t t, const RegMerge<K>& m,
                                      const CAParams& P)
{
    const int keep = m.count < P.stride ? m.count : P.stride;
    o.sccount[t] = uint8_t(keep);
    float ph0 = 0.0f, nx0 = 1.0f, ny0 = 0.0f, nz0 = 0.0f;
#pragma unroll
    for (int e = 0; e < K; ++e)
    {
        if (m.rank[e] >= keep) continue;
        const int64_t k = t * P.stride + m.rank[e];
        const float ph = float(m.phi[e]), nx = float(m.nx[e]), ny = float(m.ny[e]),
                    nz = float(m.nz[e]);
        o.scphi[k] = ph;
        o.scn3[3 * k] = nx;
        o.scn3[3 * k + 1] = ny;
        o.scn3[3 * k + 2] = nz;
        o.scbranch[k] = branchId(dev::d3(m.nx[e], m.ny[e], m.nz[e]));
        if (m.rank[e] == 0)
        {
            ph0 = ph;
            nx0 = nx;
            ny0 = ny;
            nz0 = nz;
        }
    }
    o.sphi[t] = ph0;
    o.sn3[3 * t] = nx0;
    o.sn3[3 * t + 1] = ny0;
    o.sn3[3 * t + 2] = nz0;
}

__device__ __forceinline__ D3 voxelCentre(const GridSpec& g, int i, int j, int k)
{
    return dev::d3(g.bmin[0] + g.voxel[0] * (i + 0.5), g.bmin[1] + g.voxel[1] * (j + 0.5),
                   g.bmin[2] + g.voxel[2] * (k + 0.5));
}

// Distance^2 from p to a triangle's box (bmin3, bmax3), as boxSq.
__device__ __forceinline__ double triBoxSq(D3 p, const double* __restrict__ b)
{
    double sq = dev::addAxisGap(0.0, p.x, b[0], b[3]);
    sq = dev::addAxisGap(sq, p.y, b[1], b[4]);
    return dev::addAxisGap(sq, p.z, b[2], b[5]);
}

__device__ __forceinline__ bool waveAny(bool v) { return __ballot(v) != 0ull; }

// Smallest squared distance from p to a triangle on the lane's list
// (DBL_MAX when the list is empty): the minimum the list scan of
// buildContactCandidatesForPoint computes first (SdfBuilder.cpp:287-297),
// found by a wave-cooperative walk of the triangle BVH instead.  One shared
// stack (entry e in lane e), a node visited if any lane needs it; a lane
// prunes a node when its box lies farther than sqrt(best) + mu (mu is far
// above the rounding of box and triangle distances) or when its triangles'
// tile ranges miss the lane's tile.  Every lane of the wave must call it.
__device__ double waveListMinSq(D3 p, bool active, int ti, int tj, int tk, const CAParams& P,
                                const CAData& D)
{
    const int lane = threadIdx.x & 63;
    double bestSq = DBL_MAX, bound = DBL_MAX;
    int stk = 0;
    int sp = 1;
    while (sp > 0)
    {
        --sp;
        const int ni = __builtin_amdgcn_readlane(stk, sp);
        const BvhNode nd = D.nodes[ni];
        bool need = active && dev::boxSq(p, nd) <= bound;
        if (P.useTiles) need = need && inRange(D.nodeTile + 6 * int64_t(ni), ti, tj, tk);
        if (!waveAny(need)) continue;
        if (nd.triCount > 0)
        {
            for (int s = nd.triStart; s < nd.triStart + nd.triCount; ++s)
            {
                // the triangle's own box bounds its distance like a node's
                const bool mem = need && triBoxSq(p, D.slotBox6 + 6 * int64_t(s)) <= bound &&
                                 onList(P, D, D.slotTri[s], ti, tj, tk);
                if (!waveAny(mem)) continue;
                const dev::Tri tr = dev::loadTri(D.tri9, s);
                if (!mem) continue;
                const double sq = dev::sqnorm(dev::sub(p, dev::closestPoint(p, tr.a, tr.b, tr.c)));
                if (sq < bestSq)
                {
                    bestSq = sq;
                    const double r = sqrt(sq) + D.mu;
                    bound = r * r;
                }
            }
            continue;
        }
        const bool rf = need && dev::boxSq(p, D.nodes[nd.right]) < dev::boxSq(p, D.nodes[nd.left]);
        const bool rightFirst = 2 * __popcll(__ballot(rf)) > __popcll(__ballot(need));
        if (lane == sp) stk = rightFirst ? nd.left : nd.right;
        if (lane == sp + 1) stk = rightFirst ? nd.right : nd.left;
        sp += 2;
    }
    return bestSq;
}

// The voxel's triangle list (ContactTileIndex::candidatesForVoxel,
// SdfBuilder.cpp:102-113); every triangle without tiles.
struct VoxelList
{
    const int32_t* list;
    int64_t len;
};

__device__ __forceinline__ VoxelList tileList(const CAParams& P, const CAData& D, int64_t tile)
{
    if (!P.useTiles) return VoxelList{nullptr, P.ntris};
    return VoxelList{D.tileTris + D.tileStart[tile], D.tileStart[tile + 1] - D.tileStart[tile]};
}

__device__ __forceinline__ VoxelList voxelList(const CAParams& P, const CAData& D, int ti, int tj,
                                               int tk)
{
    return tileList(P, D, ti + int64_t(P.tnx) * (tj + int64_t(P.tny) * tk));
}

// Visits the list in order, skipping triangles whose box lies farther than
// reach (= min distance + tol + mu): those cannot be candidates, so the merge
// over the rest is the merge over the whole list.
struct ListScan
{
    VoxelList L;
    const double* tbox6;
    D3 p;
    double reach2;
    template <class F>
    __device__ __forceinline__ void operator()(F&& f) const
    {
        for (int64_t q = 0; q < L.len; ++q)
        {
            const int t = L.list ? L.list[q] : int(q);
            if (!(triBoxSq(p, tbox6 + 6 * int64_t(t)) <= reach2)) continue;
            if (!f(t)) return;
        }
    }
};

__device__ __forceinline__ double reachOf(double minDist, const CAParams& P, const CAData& D)
{
    return minDist + P.tol + D.mu;
}

constexpr int kScanBuf = 256; // LDS list entries per wave

// The same visit for a wave whose lanes share one list (every lane calls it):
// the wave tests 64 list entries at a time against the union of its lanes'
// reach boxes, appends the survivors in list order to an LDS buffer, and each
// lane then runs the per-lane test over the buffer.  Entries no lane can
// reach are the only ones dropped, so every lane sees what ListScan gives it.
struct WaveListScan
{
    VoxelList L;
    const double* tbox6;
    D3 p;
    double reach2;       // this lane's (negative: takes nothing)
    double lo[3], hi[3]; // union of the lanes' reach boxes
    int* buf;            // kScanBuf entries of LDS for this wave
    template <class F>
    __device__ void operator()(F&& f) const
    {
        const int lane = threadIdx.x & 63;
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        bool stopped = false;
        int n = 0;
        for (int64_t base = 0; base < L.len; base += 64)
        {
            const int64_t q = base + lane;
            int t = 0;
            bool pass = false;
            if (q < L.len)
            {
                t = L.list ? L.list[q] : int(q);
                const double* b = tbox6 + 6 * int64_t(t);
                pass = b[3] >= lo[0] && b[0] <= hi[0] && b[4] >= lo[1] && b[1] <= hi[1] &&
                       b[5] >= lo[2] && b[2] <= hi[2];
            }
            const uint64_t m = __ballot(pass);
            if (pass) buf[n + __popcll(m & below)] = t;
            n += __popcll(m);
            if (n > kScanBuf - 64 || base + 64 >= L.len)
            {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                for (int e = 0; e < n; ++e)
                {
                    // every lane reads the same entry: a wave-uniform index lets
                    // the box and vertex loads below go through the scalar cache
                    const int tt = __builtin_amdgcn_readfirstlane(buf[e]);
                    if (stopped || !(triBoxSq(p, tbox6 + 6 * int64_t(tt)) <= reach2)) continue;
                    if (!f(tt)) stopped = true;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                n = 0;
            }
        }
    }
};

__device__ __forceinline__ double waveMinD(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
        const double w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ double waveMaxD(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
        const double w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}

// Runs candidatesFor for one voxel per lane (every lane of the wave calls
// it): the wave-shared scan when all active lanes share a tile list, the
// per-lane scan otherwise.
template <class Run>
__device__ inline int waveCandidates(D3 c, bool active, int ti, int tj, int tk, double minDist,
                                     const CAParams& P, const CAData& D, int* buf, const Run& run)
{
    const double r = reachOf(minDist, P, D);
    const double r2 = active ? r * r : -1.0;
    const int64_t tile = ti + int64_t(P.tnx) * (tj + int64_t(P.tny) * tk);
    const uint64_t act = __ballot(active);
    if (act == 0ull) return 0;
    const int first = __ffsll((unsigned long long)act) - 1;
    const int64_t t0 = __shfl(tile, first);
    const bool shared = !P.useTiles || __ballot(active && tile != t0) == 0ull;
    if (shared)
    {
        // every lane walks the active lanes' common list
        WaveListScan each{tileList(P, D, t0), D.tbox6, c, r2, {0, 0, 0}, {0, 0, 0}, buf};
        each.lo[0] = waveMinD(active ? c.x - r : DBL_MAX);
        each.lo[1] = waveMinD(active ? c.y - r : DBL_MAX);
        each.lo[2] = waveMinD(active ? c.z - r : DBL_MAX);
        each.hi[0] = waveMaxD(active ? c.x + r : -DBL_MAX);
        each.hi[1] = waveMaxD(active ? c.y + r : -DBL_MAX);
        each.hi[2] = waveMaxD(active ? c.z + r : -DBL_MAX);
        return run(each);
    }
    const ListScan each{voxelList(P, D, ti, tj, tk), D.tbox6, c, r2};
    return run(each);
}

// Dense candidate fill (SdfBuilder.cpp:477-534): one lane per voxel, a 4x4x4
// voxel brick per wave.  Voxels whose merged list outgrows the K register
// entries go to the overflow list.
template <int K>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
denseCandidateKernel(GridSpec g, CAParams P, CAData D, int bx, int by, int bz, DenseOut o,
                     Overflow ov)
{
    __shared__ int stack[kStackDepth * kBlock];
    __shared__ int scanBuf[(kBlock / 64) * kScanBuf];
    const int lane = threadIdx.x & 63;
    const int64_t brick = int64_t(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
    if (brick >= int64_t(bx) * by * bz) return; // whole waves
    const int i = int(brick % bx) * 4 + (lane & 3);
    const int j = int((brick / bx) % by) * 4 + ((lane >> 2) & 3);
    const int k = int(brick / (int64_t(bx) * by)) * 4 + (lane >> 4);
    const bool inGrid = i < g.n[0] && j < g.n[1] && k < g.n[2];
    const D3 c = voxelCentre(g, i, j, k);
    int ti, tj, tk;
    voxelTile(P, i, j, k, ti, tj, tk);
    const double minSq = waveListMinSq(c, inGrid, ti, tj, tk, P, D);
    const double minDist = minSq == DBL_MAX ? DBL_MAX : sqrt(minSq);
    RegMerge<K> m;
    const int cnt = waveCandidates(c, inGrid, ti, tj, tk, minDist, P, D,
                                   scanBuf + (threadIdx.x >> 6) * kScanBuf, [&](const auto& each) {
                                       return fastCandidates<K>(c, P, D.t21, each, minDist, inGrid,
                                                                D.nodes, D.tri9, D.slotTri,
                                                                dev::LaneStack{stack + threadIdx.x, kBlock}, m);
                                   });
    if (!inGrid) return;
    const int64_t idx = i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k);
    if (cnt > K)
    {
        pushOverflow(ov, idx, minDist);
        return;
    }
    writeDenseReg(o, idx, c, m, P, D.t21);
}

// Voxels of the first pass with more merged candidates than it holds, one lane
// each, merged lists in device memory (pool.cap entries); the minimum distance
// comes from the first pass.  A voxel outgrowing pool.cap goes on to esc (a
// later launch with room for the longest list); esc.count is NULL there.
__global__ void __launch_bounds__(kBlock)
denseOverflowKernel(GridSpec g, CAParams P, CAData D, const int64_t* __restrict__ ids,
                    const double* __restrict__ minDists, int n, MergePool pool, DenseOut o,
                    Overflow esc)
{
    __shared__ int stack[kStackDepth * kBlock];
    const int q = int(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= n) return;
    const int64_t idx = ids[q];
    const int i = int(idx % g.n[0]), j = int((idx / g.n[0]) % g.n[1]),
              k = int(idx / (int64_t(g.n[0]) * g.n[1]));
    const D3 c = voxelCentre(g, i, j, k);
    int ti, tj, tk;
    voxelTile(P, i, j, k, ti, tj, tk);
    const VoxelList L = voxelList(P, D, ti, tj, tk);
    const double minDist = minDists[q];
    const double r = reachOf(minDist, P, D);
    const ListScan each{L, D.tbox6, c, r * r};
    MergeView m = pool.view(q);
    const int cnt = candidatesFor(c, P, D.t21, each, true, minDist, true, D.nodes, D.tri9,
                                  D.slotTri, dev::LaneStack{stack + threadIdx.x, kBlock}, m);
    if (cnt > m.cap) // never with room for the longest list (esc.count NULL)
    {
        if (esc.count) pushOverflow(esc, idx, minDist);
        return;
    }
    writeDense(o, idx, c, m, P, D.t21);
}

__device__ __forceinline__ void sparseVoxel(const SparseOut& o, int bd, int64_t t, int& i, int& j,
                                            int& k)
{
    const int bvc = bd * bd * bd;
    const int64_t b = t / bvc;
    const int local = int(t % bvc);
    i = o.bijk[3 * b] * bd + local % bd;
    j = o.bijk[3 * b + 1] * bd + (local / bd) % bd;
    k = o.bijk[3 * b + 2] * bd + local / (bd * bd);
}

// Sparse candidate fill (SdfBuilder.cpp:731-803): one lane per brick voxel,
// 64 brick voxels per wave laid out by brickUnitVoxel.
template <int K>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
sparseCandidateKernel(GridSpec g, CAParams P, CAData D, int bd, int64_t total, SparseOut o,
                      Overflow ov)
{
    __shared__ int stack[kStackDepth * kBlock];
    __shared__ int scanBuf[(kBlock / 64) * kScanBuf];
    const int64_t w = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if ((w & ~int64_t(63)) >= total) return; // whole waves only
    const int64_t t = dev::brickUnitVoxel(bd, w >> 6, int(w & 63));
    const bool live = t < total;
    int i = -1, j = -1, k = -1;
    if (live)
    {
        sparseVoxel(o, bd, t, i, j, k);
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
    }
    const bool inGrid = live && i >= 0 && i < g.n[0] && j >= 0 && j < g.n[1] && k >= 0 && k < g.n[2];
    if (live && !inGrid)
    {
        o.sphi[t] = 0.0f;
        o.sn3[3 * t] = 1.0f;
        o.sn3[3 * t + 1] = 0.0f;
        o.sn3[3 * t + 2] = 0.0f;
    }
    const D3 c = voxelCentre(g, i, j, k);
    int ti, tj, tk;
    voxelTile(P, i, j, k, ti, tj, tk);
    const double minSq = waveListMinSq(c, inGrid, ti, tj, tk, P, D);
    const double minDist = minSq == DBL_MAX ? DBL_MAX : sqrt(minSq);
    RegMerge<K> m;
    const int cnt = waveCandidates(c, inGrid, ti, tj, tk, minDist, P, D,
                                   scanBuf + (threadIdx.x >> 6) * kScanBuf, [&](const auto& each) {
                                       return fastCandidates<K>(c, P, D.t21, each, minDist, inGrid,
                                                                D.nodes, D.tri9, D.slotTri,
                                                                dev::LaneStack{stack + threadIdx.x, kBlock}, m);
                                   });
    if (!inGrid) return;
    if (cnt > K)
    {
        pushOverflow(ov, t, minDist);
        return;
    }
    writeSparseReg(o, t, m, P);
}

__global__ void __launch_bounds__(kBlock)
sparseOverflowKernel(GridSpec g, CAParams P, CAData D, int bd, const int64_t* __restrict__ ids,
                     const double* __restrict__ minDists, int n, MergePool pool, SparseOut o,
                     Overflow esc)
{
    __shared__ int stack[kStackDepth * kBlock];
    const int q = int(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= n) return;
    const int64_t t = ids[q];
    int i, j, k;
    sparseVoxel(o, bd, t, i, j, k);
    const D3 c = voxelCentre(g, i, j, k);
    int ti, tj, tk;
    voxelTile(P, i, j, k, ti, tj, tk);
    const VoxelList L = voxelList(P, D, ti, tj, tk);
    const double minDist = minDists[q];
    const double r = reachOf(minDist, P, D);
    const ListScan each{L, D.tbox6, c, r * r};
    MergeView m = pool.view(q);
    const int cnt = candidatesFor(c, P, D.t21, each, true, minDist, true, D.nodes, D.tri9,
                                  D.slotTri, dev::LaneStack{stack + threadIdx.x, kBlock}, m);
    if (cnt > m.cap) // never with room for the longest list (esc.count NULL)
    {
        if (esc.count) pushOverflow(esc, t, minDist);
        return;
    }
    writeSparse(o, t, m, P);
}

// Tile lists in triangle order (buildContactTileIndex, SdfBuilder.cpp:202-258),
// built on the device: per-tile counts, an exclusive scan, an unordered fill,
// then each list sorted.
__global__ void __launch_bounds__(256)
tileCountKernel(const int32_t* __restrict__ triTile, int64_t nt, int tnx, int tny,
                int32_t* __restrict__ count)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int32_t* r = triTile + 6 * t;
    for (int tk = r[2]; tk <= r[5]; ++tk)
        for (int tj = r[1]; tj <= r[4]; ++tj)
            for (int ti = r[0]; ti <= r[3]; ++ti)
                atomicAdd(&count[ti + int64_t(tnx) * (tj + int64_t(tny) * tk)], 1);
}

// One block: start[q] = sum of count[0..q), start[n] = the total; *maxCount = max count.
constexpr int kScanBlock = 1024;
__global__ void __launch_bounds__(kScanBlock)
tileScanKernel(const int32_t* __restrict__ count, int64_t n, int64_t* __restrict__ start,
               int32_t* __restrict__ maxCount)
{
    __shared__ int64_t part[kScanBlock];
    __shared__ int32_t mx[kScanBlock];
    const int tid = threadIdx.x;
    const int64_t chunk = (n + kScanBlock - 1) / kScanBlock;
    const int64_t b = tid * chunk, e = b + chunk < n ? b + chunk : n;
    int64_t s = 0;
    int32_t m = 0;
    for (int64_t q = b; q < e; ++q)
    {
        s += count[q];
        m = count[q] > m ? count[q] : m;
    }
    part[tid] = s;
    mx[tid] = m;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1)
    {
        const int64_t v = tid >= off ? part[tid - off] : 0;
        const int32_t w = tid >= off ? mx[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        mx[tid] = w > mx[tid] ? w : mx[tid];
        __syncthreads();
    }
    int64_t run = part[tid] - s; // exclusive prefix of this chunk
    for (int64_t q = b; q < e; ++q)
    {
        start[q] = run;
        run += count[q];
    }
    if (tid == kScanBlock - 1)
    {
        start[n] = part[tid];
        *maxCount = mx[tid];
    }
}

__global__ void __launch_bounds__(256)
tileFillKernel(const int32_t* __restrict__ triTile, int64_t nt, int tnx, int tny,
               const int64_t* __restrict__ start, int32_t* __restrict__ cursor,
               int32_t* __restrict__ tris)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int32_t* r = triTile + 6 * t;
    for (int tk = r[2]; tk <= r[5]; ++tk)
        for (int tj = r[1]; tj <= r[4]; ++tj)
            for (int ti = r[0]; ti <= r[3]; ++ti)
            {
                const int64_t tile = ti + int64_t(tnx) * (tj + int64_t(tny) * tk);
                tris[start[tile] + atomicAdd(&cursor[tile], 1)] = int32_t(t);
            }
}

// Sorts every tile's list ascending: a bitonic sort in LDS for lists of up
// to kSortMax entries, a heap sort by one lane for longer ones.
constexpr int kSortMax = 4096;
__global__ void __launch_bounds__(256)
tileSortKernel(const int64_t* __restrict__ start, int64_t ntiles, int32_t* __restrict__ tris)
{
    __shared__ int32_t v[kSortMax];
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
    {
        const int64_t b = start[tile];
        const int len = int(start[tile + 1] - b);
        if (len <= 1) continue;
        int32_t* a = tris + b;
        if (len > kSortMax)
        {
            if (threadIdx.x == 0) sortIds(a, len);
            __syncthreads();
            continue;
        }
        int n2 = 1;
        while (n2 < len) n2 <<= 1;
        for (int q = threadIdx.x; q < n2; q += blockDim.x) v[q] = q < len ? a[q] : INT32_MAX;
        __syncthreads();
        for (int k = 2; k <= n2; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1)
            {
                for (int q = threadIdx.x; q < n2; q += blockDim.x)
                {
                    const int r = q ^ j;
                    if (r > q)
                    {
                        const bool up = (q & k) == 0;
                        const int32_t x = v[q], y = v[r];
                        if ((x > y) == up)
                        {
                            v[q] = y;
                            v[r] = x;
                        }
                    }
                }
                __syncthreads();
            }
        for (int q = threadIdx.x; q < len; q += blockDim.x) a[q] = v[q];
        __syncthreads();
    }
}

// Sparse activation with tiles (SdfBuilder.cpp:638-664): a brick is active
// when any tile over its voxel range has a non-empty list.
__global__ void __launch_bounds__(256)
sparseTileActivateKernel(GridSpec g, int bd, int nbx, int nby, int nbz, int tileDim, int tnx,
                         int tny, const int32_t* __restrict__ count, uint8_t* __restrict__ mask)
{
    const int64_t bl = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (bl >= int64_t(nbx) * nby * nbz) return;
    const int bi = int(bl % nbx), bj = int((bl / nbx) % nby), bk = int(bl / (int64_t(nbx) * nby));
    const int i0 = bi * bd, j0 = bj * bd, k0 = bk * bd;
    const int i1 = min(i0 + bd - 1, g.n[0] - 1);
    const int j1 = min(j0 + bd - 1, g.n[1] - 1);
    const int k1 = min(k0 + bd - 1, g.n[2] - 1);
    bool act = false;
    for (int tk = k0 / tileDim; tk <= k1 / tileDim && !act; ++tk)
        for (int tj = j0 / tileDim; tj <= j1 / tileDim && !act; ++tj)
            for (int ti = i0 / tileDim; ti <= i1 / tileDim && !act; ++ti)
                act = count[ti + int64_t(tnx) * (tj + int64_t(tny) * tk)] != 0;
    mask[bl] = act ? 1 : 0;
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
    std::unique_ptr<double[]> t21;  // 21 per triangle (see CTri)
    std::unique_ptr<double[]> box6; // bmin3, bmax3 per triangle
};

// makeContactTriangleInfos (SdfBuilder.cpp:155-184), triangles in parallel;
// arrays left uninitialised until written (every element is).
HostCA makeTriangleInfos(const double* verts, int64_t nv, const int32_t* tris, int64_t nt,
                         const double* cn)
{
    HostCA h;
    h.t21.reset(new double[size_t(std::max<int64_t>(1, nt)) * 21]);
    h.box6.reset(new double[size_t(std::max<int64_t>(1, nt)) * 6]);
    std::atomic<bool> bad{false};
    parallelRange(nt, 8192, [&](int64_t t0, int64_t t1) {
        for (int64_t t = t0; t < t1; ++t)
        {
            const int32_t* ix = tris + 3 * t;
            if (ix[0] < 0 || ix[0] >= nv || ix[1] < 0 || ix[1] >= nv || ix[2] < 0 || ix[2] >= nv)
            {
                bad = true;
                continue;
            }
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
                h.box6[6 * size_t(t) + k] = lo;
                h.box6[6 * size_t(t) + 3 + k] = hi;
            }
        }
    });
    if (bad) throw Error(SDFO_ERR_INVALID_ARGUMENT, "Contact-aware mesh triangle references invalid vertex");
    return h;
}

struct HostTiles
{
    int tileDim = 1, nx = 1, ny = 1, nz = 1;
    std::unique_ptr<int32_t[]> triTile; // 6 per triangle: tile range (lo xyz, hi xyz)
    std::vector<int32_t> nodeTile;      // 6 per BVH node: union of its triangles' ranges
};

// buildContactTileIndex (SdfBuilder.cpp:202-258) without the lists: each
// triangle's tile range (a tile's list is its triangles in triangle order).
HostTiles buildTiles(const GridSpec& g, const HostCA& ca, int64_t nt, int reqTileDim, double infl)
{
    HostTiles T;
    T.tileDim = std::max(1, reqTileDim);
    T.nx = (g.n[0] + T.tileDim - 1) / T.tileDim;
    T.ny = (g.n[1] + T.tileDim - 1) / T.tileDim;
    T.nz = (g.n[2] + T.tileDim - 1) / T.tileDim;
    const double h = g.hmax();
    const double tileRadius = 0.5 * std::sqrt(3.0) * double(T.tileDim) * h;
    const double expansion = (0.0 < infl ? infl : 0.0) + tileRadius + h;
    auto clampi = [](int v, int hi) { return std::max(0, std::min(hi, v)); };
    T.triTile.reset(new int32_t[size_t(std::max<int64_t>(1, nt)) * 6]);
    parallelRange(nt, 16384, [&](int64_t t0, int64_t t1) {
        for (int64_t t = t0; t < t1; ++t)
            for (int a = 0; a < 3; ++a)
            {
                const double lo = ca.box6[6 * size_t(t) + a] - expansion;
                const double hi = ca.box6[6 * size_t(t) + 3 + a] + expansion;
                const int v0 = int(std::floor((lo - g.bmin[a]) / g.voxel[a] - 0.5));
                const int v1 = int(std::ceil((hi - g.bmin[a]) / g.voxel[a] - 0.5));
                T.triTile[6 * size_t(t) + a] = clampi(v0, g.n[a] - 1) / T.tileDim;
                T.triTile[6 * size_t(t) + 3 + a] = clampi(v1, g.n[a] - 1) / T.tileDim;
            }
    });
    return T;
}

// Per BVH node, the union of the tile ranges of the triangles below it.
void nodeTileUnion(HostTiles& T, const HostBvh& hb)
{
    // nodes are in pre-order: children after their parent
    const size_t nn = hb.nodes.size();
    T.nodeTile.assign(nn * 6, 0);
    for (size_t q = nn; q-- > 0;)
    {
        const BvhNode& nd = hb.nodes[q];
        int32_t* u = &T.nodeTile[6 * q];
        for (int a = 0; a < 3; ++a)
        {
            u[a] = INT32_MAX;
            u[3 + a] = -1;
        }
        auto merge = [&](const int32_t* r) {
            for (int a = 0; a < 3; ++a)
            {
                u[a] = std::min(u[a], r[a]);
                u[3 + a] = std::max(u[3 + a], r[3 + a]);
            }
        };
        if (nd.triCount > 0)
            for (int sl = nd.triStart; sl < nd.triStart + nd.triCount; ++sl)
                merge(&T.triTile[6 * size_t(hb.slotTri[size_t(sl)])]);
        else
        {
            merge(&T.nodeTile[6 * size_t(nd.left)]);
            merge(&T.nodeTile[6 * size_t(nd.right)]);
        }
    }
}

// Search margin of the device search: 2^-30 of the largest coordinate
// magnitude of the mesh and the grid, far above the rounding of box and
// triangle distances at that scale.
double searchMargin(const double* verts, int64_t nv, const GridSpec& g)
{
    double m = 0.0;
    for (int64_t q = 0; q < 3 * nv; ++q) m = std::max(m, std::fabs(verts[q]));
    for (int a = 0; a < 3; ++a) m = std::max({m, std::fabs(g.bmin[a]), std::fabs(g.bmax[a])});
    return std::ldexp(m, -30);
}

// Device buffers of one build: stream-ordered allocations on the build's
// stream, returned to the pool on that stream when the build ends.
struct DevBuf
{
    hipStream_t s;
    std::vector<void*> ptrs;
    explicit DevBuf(hipStream_t st) : s(st) {}
    void* raw(size_t bytes)
    {
        void* p = nullptr;
        SDFO_HIP(hipMallocAsync(&p, std::max<size_t>(1, bytes), s));
        ptrs.push_back(p);
        return p;
    }
    // a copy of n host elements; the host array must stay valid until the
    // stream has passed the copy
    template <typename T>
    T* up(const T* h, size_t n)
    {
        T* p = static_cast<T*>(raw(n * sizeof(T)));
        if (n) SDFO_HIP(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
        return p;
    }
    template <typename T>
    T* up(const std::vector<T>& h)
    {
        return up(h.data(), h.size());
    }
    int* zeroInt()
    {
        int* p = static_cast<int*>(raw(sizeof(int)));
        SDFO_HIP(hipMemsetAsync(p, 0, sizeof(int), s));
        return p;
    }
    ~DevBuf()
    {
        for (void* p : ptrs) (void)hipFreeAsync(p, s);
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
    g->allocStream = s;
    g->spec = setupGrid(verts, nv, p);
    const GridSpec& G = g->spec;
    const int stride = std::max(0, std::min(int(p.max_candidates), 8));
    g->stride = stride;
    const double h = G.hmax();
    const double tol = p.feature_distance_tolerance > 0.0 ? p.feature_distance_tolerance : 1.5 * h;
    const double cosMerge = std::cos(p.normal_merge_angle_degrees * std::acos(-1.0) / 180.0);
    const HostCA ca = makeTriangleInfos(verts, nv, tris, nt, cornerN); // validates indices
    // The triangle BVH is built on a host thread while the tile lists, the
    // activation and the allocations proceed; ensureBvh() joins it.
    HostBvh hb;
    std::exception_ptr bvhError;
    std::thread bvhThread([&] {
        try
        {
            hb = buildHostBvh(verts, nv, tris, nt);
        }
        catch (...)
        {
            bvhError = std::current_exception();
        }
    });
    struct Joiner
    {
        std::thread& t;
        ~Joiner()
        {
            if (t.joinable()) t.join();
        }
    } joiner{bvhThread};
    double band = 0.0, activation = 0.0;
    if (sparse)
    {
        band = p.band_distance > 0.0 ? p.band_distance : double(p.band_voxels) * h;
        activation = band + double(p.halo_voxels) * h;
    }
    const double infl = sparse ? std::max(activation, tol + 2.0 * h) : tol + 2.0 * h;
    HostTiles tiles;
    if (p.use_tile_acceleration) tiles = buildTiles(G, ca, nt, p.tile_dim, infl);

    DevBuf buf(s);
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
    Overflow ov{nullptr, nullptr, nullptr};
    // register entries of the first pass (4 unless a test asks for 1 or 8)
    const int mergeCap = debugMergeCap();
    // The merged list of a voxel holds at most one entry per triangle on its
    // list (the longest tile list, or every triangle without tiles), plus the
    // nearest-triangle fallback when that list is empty.
    int poolCap = int(nt);
    // Re-runs the voxels the first pass listed, with merged lists in device
    // memory: first kSecondCap entries per voxel, then, for the few voxels
    // that need more, room for the longest list.  Each launch holds at most
    // ~1 GiB of list storage.
    auto runLevel = [&](auto launch, const int64_t* ids, const double* md, int cnt, int cap,
                        const Overflow& esc) {
        const size_t entry = 4 * sizeof(double) + 2 * sizeof(int);
        const int64_t per = std::max<int64_t>(1, int64_t((size_t(1) << 30) / (entry * size_t(cap))));
        const int batch = int(std::min<int64_t>(cnt, per));
        DevBuf pb(s);
        const size_t e = size_t(batch) * size_t(cap);
        MergePool pool{};
        pool.cap = cap;
        pool.lanes = batch;
        pool.phi = static_cast<double*>(pb.raw(e * sizeof(double)));
        pool.nx = static_cast<double*>(pb.raw(e * sizeof(double)));
        pool.ny = static_cast<double*>(pb.raw(e * sizeof(double)));
        pool.nz = static_cast<double*>(pb.raw(e * sizeof(double)));
        pool.tri = static_cast<int*>(pb.raw(e * sizeof(int)));
        pool.order = static_cast<int*>(pb.raw(e * sizeof(int)));
        for (int off = 0; off < cnt; off += batch)
        {
            launch(ids + off, md + off, std::min(batch, cnt - off), pool, esc);
            SDFO_HIP(hipGetLastError());
        }
        SDFO_HIP(hipStreamSynchronize(s));
    };
    auto readCount = [&](const int* c) {
        int n = 0;
        SDFO_HIP(hipMemcpyAsync(&n, c, sizeof(int), hipMemcpyDeviceToHost, s));
        SDFO_HIP(hipStreamSynchronize(s));
        return n;
    };
    const int secondCap = debugSecondCap() > 0 ? debugSecondCap() : kSecondCap;
    auto resolveOverflow = [&](auto launch) {
        const int cnt = readCount(ov.count);
        if (cnt <= 0) return;
        if (poolCap <= secondCap)
        {
            runLevel(launch, ov.ids, ov.minDist, cnt, poolCap, Overflow{nullptr, nullptr, nullptr});
            return;
        }
        Overflow esc{buf.zeroInt(),
                     static_cast<int64_t*>(buf.raw(size_t(cnt) * sizeof(int64_t))),
                     static_cast<double*>(buf.raw(size_t(cnt) * sizeof(double)))};
        runLevel(launch, ov.ids, ov.minDist, cnt, secondCap, esc);
        const int cnt2 = readCount(esc.count);
        if (cnt2 > 0)
            runLevel(launch, esc.ids, esc.minDist, cnt2, poolCap, Overflow{nullptr, nullptr, nullptr});
    };
    bool bvhReady = false;
    auto ensureBvh = [&] {
        if (bvhReady) return;
        bvhThread.join();
        if (bvhError) std::rethrow_exception(bvhError);
        db.upload(hb, s);
        D.nodes = db.nodes;
        D.tri9 = db.tri9;
        D.slotBox6 = db.box6;
        D.slotTri = db.triIdx;
        if (P.useTiles)
        {
            nodeTileUnion(tiles, hb);
            D.nodeTile = buf.up(tiles.nodeTile);
        }
        bvhReady = true;
    };
    try
    {
        keepStreamPoolMemory(g->device);
        D.t21 = buf.up(ca.t21.get(), size_t(nt) * 21);
        D.tbox6 = buf.up(ca.box6.get(), size_t(nt) * 6);
        int32_t* tileCount = nullptr;
        if (P.useTiles)
        {
            D.triTile = buf.up(tiles.triTile.get(), size_t(nt) * 6);
            // tile lists on the device (count, scan, fill, sort)
            const int64_t ntiles = int64_t(tiles.nx) * tiles.ny * tiles.nz;
            tileCount = static_cast<int32_t*>(buf.raw(size_t(ntiles) * sizeof(int32_t)));
            int32_t* cursor = static_cast<int32_t*>(buf.raw(size_t(ntiles) * sizeof(int32_t)));
            int64_t* start = static_cast<int64_t*>(buf.raw(size_t(ntiles + 1) * sizeof(int64_t)));
            int32_t* maxCount = static_cast<int32_t*>(buf.raw(sizeof(int32_t)));
            SDFO_HIP(hipMemsetAsync(tileCount, 0, size_t(ntiles) * sizeof(int32_t), s));
            SDFO_HIP(hipMemsetAsync(cursor, 0, size_t(ntiles) * sizeof(int32_t), s));
            hipLaunchKernelGGL(tileCountKernel, dim3(blocksFor(nt, 256)), dim3(256), 0, s, D.triTile,
                               nt, tiles.nx, tiles.ny, tileCount);
            SDFO_HIP(hipGetLastError());
            hipLaunchKernelGGL(tileScanKernel, dim3(1), dim3(kScanBlock), 0, s, tileCount, ntiles,
                               start, maxCount);
            SDFO_HIP(hipGetLastError());
            int64_t total = 0;
            int32_t longest = 0;
            SDFO_HIP(hipMemcpyAsync(&total, start + ntiles, sizeof(int64_t), hipMemcpyDeviceToHost, s));
            SDFO_HIP(hipMemcpyAsync(&longest, maxCount, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            SDFO_HIP(hipStreamSynchronize(s));
            int32_t* tris = static_cast<int32_t*>(buf.raw(size_t(std::max<int64_t>(total, 1)) * sizeof(int32_t)));
            hipLaunchKernelGGL(tileFillKernel, dim3(blocksFor(nt, 256)), dim3(256), 0, s, D.triTile,
                               nt, tiles.nx, tiles.ny, start, cursor, tris);
            SDFO_HIP(hipGetLastError());
            hipLaunchKernelGGL(tileSortKernel, dim3(unsigned(std::min<int64_t>(ntiles, 65536))),
                               dim3(256), 0, s, start, ntiles, tris);
            SDFO_HIP(hipGetLastError());
            D.tileStart = start;
            D.tileTris = tris;
            poolCap = std::max(int(longest), 1);
        }
        D.mu = searchMargin(verts, nv, G);
        ov.count = buf.zeroInt();
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
            ov.minDist = static_cast<double*>(buf.raw(size_t(N) * sizeof(double)));
            ensureBvh();
            const dim3 grid(blocksFor(int64_t(bx) * by * bz, kBlock / 64));
            if (mergeCap == 1)
                hipLaunchKernelGGL(denseCandidateKernel<1>, grid, dim3(kBlock), 0, s, G, P, D, bx, by, bz, o, ov);
            else if (mergeCap == 8)
                hipLaunchKernelGGL(denseCandidateKernel<8>, grid, dim3(kBlock), 0, s, G, P, D, bx, by, bz, o, ov);
            else
                hipLaunchKernelGGL(denseCandidateKernel<4>, grid, dim3(kBlock), 0, s, G, P, D, bx, by, bz, o, ov);
            SDFO_HIP(hipGetLastError());
            resolveOverflow([&](const int64_t* ids, const double* md, int n, const MergePool& pool,
                                const Overflow& esc) {
                hipLaunchKernelGGL(denseOverflowKernel, dim3(blocksFor(n, kBlock)), dim3(kBlock), 0, s,
                                   G, P, D, ids, md, n, pool, o, esc);
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
                hipLaunchKernelGGL(sparseTileActivateKernel, dim3(blocksFor(cells, 256)), dim3(256),
                                   0, s, G, bd, g->nbx, g->nby, g->nbz, tiles.tileDim, tiles.nx,
                                   tiles.ny, tileCount, g->activeMask);
                SDFO_HIP(hipGetLastError());
                SDFO_HIP(hipMemcpyAsync(mask.data(), g->activeMask, mask.size(), hipMemcpyDeviceToHost, s));
                SDFO_HIP(hipStreamSynchronize(s));
            }
            else
            {
                ensureBvh();
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
                ov.minDist = static_cast<double*>(buf.raw(size_t(total) * sizeof(double)));
                ensureBvh();
                const dim3 grid(blocksFor(total, kBlock));
                if (mergeCap == 1)
                    hipLaunchKernelGGL(sparseCandidateKernel<1>, grid, dim3(kBlock), 0, s, G, P, D, bd, total, o, ov);
                else if (mergeCap == 8)
                    hipLaunchKernelGGL(sparseCandidateKernel<8>, grid, dim3(kBlock), 0, s, G, P, D, bd, total, o, ov);
                else
                    hipLaunchKernelGGL(sparseCandidateKernel<4>, grid, dim3(kBlock), 0, s, G, P, D, bd, total, o, ov);
                SDFO_HIP(hipGetLastError());
                resolveOverflow([&](const int64_t* ids, const double* md, int n, const MergePool& pool,
                                    const Overflow& esc) {
                    hipLaunchKernelGGL(sparseOverflowKernel, dim3(blocksFor(n, kBlock)), dim3(kBlock), 0,
                                       s, G, P, D, bd, ids, md, n, pool, o, esc);
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
        if (!bvhReady)
        {
            // no voxel needed the BVH; its errors still stand
            bvhThread.join();
            if (bvhError) std::rethrow_exception(bvhError);
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
