// Host-side geometry set-up: grid derivation and the BVH that the build
// kernels traverse.  Both run once per build and are O(N_t log N_t); the
// per-voxel work is on the GPU (build.hip).
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

#include "sdfo_internal.hpp"

namespace sdfo {

namespace {

// GridSpec.h:65-74: smallest ladder value >= n, capped at maxRes.
int snapResolution(int n, int maxRes)
{
    static const int ladder[] = {8, 16, 32, 64, 128, 256, 512, 1024, 2048};
    for (int level : ladder)
        if (level >= n || level >= maxRes) return level < maxRes ? level : maxRes;
    return maxRes;
}

inline double minOf(double a, double b) { return (b < a) ? b : a; } // std::min
inline double maxOf(double a, double b) { return (a < b) ? b : a; } // std::max

} // namespace

// SdfBuilder::setupGrid (SdfBuilder.cpp:994-1035).
GridSpec setupGrid(const double* verts, int64_t nv, const sdfo_build_params& p)
{
    double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
    for (int64_t v = 0; v < nv; ++v)
        for (int a = 0; a < 3; ++a)
        {
            lo[a] = minOf(lo[a], verts[3 * v + a]);
            hi[a] = maxOf(hi[a], verts[3 * v + a]);
        }
    const double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    const double diag = std::sqrt(ex * ex + ey * ey + ez * ez);
    const double pad = p.padding * diag;
    GridSpec g;
    double size[3];
    for (int a = 0; a < 3; ++a)
    {
        g.bmin[a] = lo[a] - pad;
        size[a] = (hi[a] + pad) - g.bmin[a];
    }
    double vs = p.voxel_size;
    if (vs <= 0.0) vs = maxOf(maxOf(size[0], size[1]), size[2]) / double(p.target_resolution);
    for (int a = 0; a < 3; ++a)
    {
        const int raw = std::max(1, int(std::ceil(size[a] / vs)));
        g.n[a] = snapResolution(raw, p.max_resolution);
        g.voxel[a] = vs;
        g.bmax[a] = g.bmin[a] + double(g.n[a]) * vs;
    }
    return g;
}

// KdTree::build / buildRecursive (KdTree.h:91-121, 188-242): object-median
// split on the largest centroid extent via std::nth_element, leaves of <= 8.
// Node order is the recursion's creation order so indices match the
// reference tree; the traversal order (near child first) is reproduced by
// the kernels.
HostBvh buildHostBvh(const double* verts, int64_t nv, const int32_t* tris, int64_t nt)
{
    HostBvh h;
    const int n = int(nt);
    std::vector<double> tb(size_t(n) * 9); // bmin3, bmax3, centroid3 per triangle
    for (int t = 0; t < n; ++t)
    {
        const int32_t* ix = tris + 3 * size_t(t);
        for (int c = 0; c < 3; ++c)
            if (ix[c] < 0 || ix[c] >= nv)
                throw Error(SDFO_ERR_INVALID_ARGUMENT, "triangle references invalid vertex");
        const double* a = verts + 3 * size_t(ix[0]);
        const double* b = verts + 3 * size_t(ix[1]);
        const double* c = verts + 3 * size_t(ix[2]);
        double* o = &tb[9 * size_t(t)];
        for (int k = 0; k < 3; ++k)
        {
            o[k] = minOf(minOf(a[k], b[k]), c[k]);
            o[3 + k] = maxOf(maxOf(a[k], b[k]), c[k]);
            o[6 + k] = ((a[k] + b[k]) + c[k]) / 3.0;
        }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    h.nodes.reserve(2 * size_t(n));

    struct Rec
    {
        HostBvh& h;
        std::vector<int>& idx;
        const std::vector<double>& tb;
        int build(int lo, int hi, int depth)
        {
            if (depth > h.depth) h.depth = depth;
            const int me = int(h.nodes.size());
            h.nodes.emplace_back();
            const double big = std::numeric_limits<double>::max();
            double bmn[3] = {big, big, big}, bmx[3] = {-big, -big, -big};
            double cmn[3] = {big, big, big}, cmx[3] = {-big, -big, -big};
            for (int i = lo; i < hi; ++i)
            {
                const double* r = &tb[9 * size_t(idx[i])];
                for (int k = 0; k < 3; ++k)
                {
                    bmn[k] = minOf(bmn[k], r[k]);
                    bmx[k] = maxOf(bmx[k], r[3 + k]);
                    cmn[k] = minOf(cmn[k], r[6 + k]);
                    cmx[k] = maxOf(cmx[k], r[6 + k]);
                }
            }
            BvhNode nd{};
            for (int k = 0; k < 3; ++k)
            {
                nd.bmin[k] = bmn[k];
                nd.bmax[k] = bmx[k];
            }
            nd.left = nd.right = -1;
            nd.triStart = -1;
            nd.triCount = 0;
            const int count = hi - lo;
            if (count <= 8)
            {
                nd.triStart = lo;
                nd.triCount = count;
                h.nodes[me] = nd;
                return me;
            }
            const double ext[3] = {cmx[0] - cmn[0], cmx[1] - cmn[1], cmx[2] - cmn[2]};
            int axis = 0;
            if (ext[1] > ext[0]) axis = 1;
            if (ext[2] > ext[axis]) axis = 2;
            const int mid = lo + count / 2;
            const std::vector<double>& tbr = tb;
            std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi,
                             [&tbr, axis](int a, int b) {
                                 return tbr[9 * size_t(a) + 6 + axis] < tbr[9 * size_t(b) + 6 + axis];
                             });
            nd.left = build(lo, mid, depth + 1);
            nd.right = build(mid, hi, depth + 1);
            h.nodes[me] = nd;
            return me;
        }
    } rec{h, idx, tb};
    if (n > 0) rec.build(0, n, 1);

    h.slotTri.assign(idx.begin(), idx.end());
    h.slotTri9.resize(size_t(n) * 9);
    for (int s = 0; s < n; ++s)
    {
        const int32_t* ix = tris + 3 * size_t(idx[s]);
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k)
                h.slotTri9[9 * size_t(s) + 3 * c + k] = verts[3 * size_t(ix[c]) + k];
    }
    // KdTree.h:330-333: fixed irrational ray direction, normalised once on the host.
    const double d0[3] = {1.0, 0.3713906763541037, 0.52999894000318};
    const double len = std::sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
    for (int k = 0; k < 3; ++k) h.rayDir[k] = d0[k] / len;
    return h;
}

void DeviceBvh::upload(const HostBvh& h, hipStream_t s)
{
    release();
    nnodes = int(h.nodes.size());
    depth = h.depth;
    for (int k = 0; k < 3; ++k) rayDir[k] = h.rayDir[k];
    if (depth + 2 > kStackDepth)
        throw Error(SDFO_ERR_UNSUPPORTED, "BVH depth " + std::to_string(depth) +
                                              " exceeds the traversal stack");
    SDFO_HIP(hipMalloc(&nodes, sizeof(BvhNode) * std::max<size_t>(1, h.nodes.size())));
    SDFO_HIP(hipMalloc(&tri9, sizeof(double) * std::max<size_t>(1, h.slotTri9.size())));
    SDFO_HIP(hipMalloc(&triIdx, sizeof(int32_t) * std::max<size_t>(1, h.slotTri.size())));
    SDFO_HIP(hipMemcpyAsync(nodes, h.nodes.data(), sizeof(BvhNode) * h.nodes.size(),
                            hipMemcpyHostToDevice, s));
    SDFO_HIP(hipMemcpyAsync(tri9, h.slotTri9.data(), sizeof(double) * h.slotTri9.size(),
                            hipMemcpyHostToDevice, s));
    SDFO_HIP(hipMemcpyAsync(triIdx, h.slotTri.data(), sizeof(int32_t) * h.slotTri.size(),
                            hipMemcpyHostToDevice, s));
}

void DeviceBvh::release()
{
    if (nodes) (void)hipFree(nodes);
    if (tri9) (void)hipFree(tri9);
    if (triIdx) (void)hipFree(triIdx);
    nodes = nullptr;
    tri9 = nullptr;
    triIdx = nullptr;
}

} // namespace sdfo
