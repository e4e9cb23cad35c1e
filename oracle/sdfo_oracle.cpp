// ============================================================================
// sdfo_oracle.cpp -- CPU FP64 restatement of the SdfOracle build/query path and
// of the prototype's per-point contact reduction.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
// library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load it.  The product library never links, calls or falls back to it.
//
// Every routine cites the reference file:line it restates.  Arithmetic is FP64,
// compiled with -ffp-contract=off so that every product/sum rounds separately
// (the reference is built without -march, i.e. without FMA contraction).
// Vector reductions (dot, squaredNorm) are evaluated left to right,
// ((x + y) + z); Eigen's internal reduction order is version dependent and the
// reference cannot be compiled here (Eigen3 is absent), so this order is the
// documented convention shared with the HIP kernels.
//
// Exported C ABI (ctypes): see oracle/oracle.py.
// ============================================================================

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3
{
    double x, y, z;
};

inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double sqnorm(V3 a) { return dot(a, a); }
inline double norm(V3 a) { return std::sqrt(sqnorm(a)); }
inline V3 cross(V3 a, V3 b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// std::min / std::max semantics (first argument wins on equality / NaN).
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }
inline V3 cmin(V3 a, V3 b) { return V3{smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)}; }
inline V3 cmax(V3 a, V3 b) { return V3{smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)}; }
inline double comp(V3 a, int axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }

inline V3 load3(const double* p, int64_t i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline void store3(double* p, int64_t i, V3 v)
{
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

// ---------------------------------------------------------------------------
// Grid spec -- GridSpec.h:17-74, SdfBuilder.cpp:994-1035
// ---------------------------------------------------------------------------
struct Grid
{
    V3 bmin, bmax, voxel;
    int n[3];

    int64_t total() const { return int64_t(n[0]) * n[1] * n[2]; }
    // GridSpec.h:30-34
    V3 center(int i, int j, int k) const
    {
        return V3{bmin.x + voxel.x * (i + 0.5), bmin.y + voxel.y * (j + 0.5),
                  bmin.z + voxel.z * (k + 0.5)};
    }
    // GridSpec.h:37-45 (std::round: half away from zero)
    void indexFromWorld(V3 p, int& i, int& j, int& k) const
    {
        const double fx = (p.x - bmin.x) / voxel.x - 0.5;
        const double fy = (p.y - bmin.y) / voxel.y - 0.5;
        const double fz = (p.z - bmin.z) / voxel.z - 0.5;
        i = static_cast<int>(std::round(fx));
        j = static_cast<int>(std::round(fy));
        k = static_cast<int>(std::round(fz));
    }
    // GridSpec.h:48-53
    bool inBounds(int i, int j, int k) const
    {
        return i >= 0 && i < n[0] && j >= 0 && j < n[1] && k >= 0 && k < n[2];
    }
    // GridSpec.h:56-59
    int64_t lin(int i, int j, int k) const { return i + int64_t(n[0]) * (j + int64_t(n[1]) * k); }
    double hmax() const { return smax(smax(voxel.x, voxel.y), voxel.z); }
};

Grid gridFromArrays(const double* g9, const int* shape)
{
    Grid g;
    g.bmin = v3(g9[0], g9[1], g9[2]);
    g.bmax = v3(g9[3], g9[4], g9[5]);
    g.voxel = v3(g9[6], g9[7], g9[8]);
    g.n[0] = shape[0];
    g.n[1] = shape[1];
    g.n[2] = shape[2];
    return g;
}

void gridToArrays(const Grid& g, double* g9, int* shape)
{
    g9[0] = g.bmin.x; g9[1] = g.bmin.y; g9[2] = g.bmin.z;
    g9[3] = g.bmax.x; g9[4] = g.bmax.y; g9[5] = g.bmax.z;
    g9[6] = g.voxel.x; g9[7] = g.voxel.y; g9[8] = g.voxel.z;
    shape[0] = g.n[0]; shape[1] = g.n[1]; shape[2] = g.n[2];
}

// GridSpec.h:65-74
int snapToStandardResolution(int n, int maxRes)
{
    static const int levels[] = {8, 16, 32, 64, 128, 256, 512, 1024, 2048};
    for (int v : levels)
    {
        if (v >= n || v >= maxRes)
            return std::min(v, maxRes);
    }
    return maxRes;
}

// SdfBuilder.cpp:994-1035
Grid setupGrid(const double* verts, int64_t nv, double padding, double voxelSize, int targetRes,
               int maxRes)
{
    V3 vmin = v3(1e30, 1e30, 1e30);
    V3 vmax = v3(-1e30, -1e30, -1e30);
    for (int64_t i = 0; i < nv; ++i)
    {
        const V3 v = load3(verts, i);
        vmin = cmin(vmin, v);
        vmax = cmax(vmax, v);
    }
    const double diag = norm(vmax - vmin);
    const double pad = padding * diag;
    const V3 bmin = v3(vmin.x - pad, vmin.y - pad, vmin.z - pad);
    const V3 bmax = v3(vmax.x + pad, vmax.y + pad, vmax.z + pad);
    const V3 size = bmax - bmin;
    double vs = voxelSize;
    if (vs <= 0.0)
    {
        const double longest = smax(smax(size.x, size.y), size.z);
        vs = longest / static_cast<double>(targetRes);
    }
    const int rx = std::max(1, static_cast<int>(std::ceil(size.x / vs)));
    const int ry = std::max(1, static_cast<int>(std::ceil(size.y / vs)));
    const int rz = std::max(1, static_cast<int>(std::ceil(size.z / vs)));
    Grid g;
    g.n[0] = snapToStandardResolution(rx, maxRes);
    g.n[1] = snapToStandardResolution(ry, maxRes);
    g.n[2] = snapToStandardResolution(rz, maxRes);
    g.bmin = bmin;
    g.voxel = v3(vs, vs, vs);
    g.bmax = v3(bmin.x + g.n[0] * vs, bmin.y + g.n[1] * vs, bmin.z + g.n[2] * vs);
    return g;
}

// ---------------------------------------------------------------------------
// Closest point on triangle -- KdTree.h:26-74 (Ericson, RTCD 5.1.5)
// ---------------------------------------------------------------------------
V3 closestPointOnTriangle(V3 p, V3 a, V3 b, V3 c)
{
    const V3 ab = b - a;
    const V3 ac = c - a;
    const V3 ap = p - a;
    const double d1 = dot(ab, ap);
    const double d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;

    const V3 bp = p - b;
    const double d3 = dot(ab, bp);
    const double d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;

    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0)
    {
        const double v = d1 / (d1 - d3);
        return a + v * ab;
    }

    const V3 cp = p - c;
    const double d5 = dot(ab, cp);
    const double d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;

    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0)
    {
        const double w = d2 / (d2 - d6);
        return a + w * ac;
    }

    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
    {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + w * (c - b);
    }

    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom;
    const double w = vc * denom;
    return (a + ab * v) + ac * w;
}

// ---------------------------------------------------------------------------
// BVH -- KdTree.h:76-442
// ---------------------------------------------------------------------------
struct Node
{
    V3 bmin, bmax;
    int left = -1, right = -1, triStart = -1, triCount = 0;
};

struct NearestResult
{
    V3 point{0, 0, 0};
    int tri = -1;
    double distance = 0.0;
    double bestSq = 0.0;
    V3 faceNormal{0, 0, 1};
    int64_t triTests = 0;
    int64_t boxTests = 0;
};

struct Bvh
{
    const double* verts = nullptr;
    const int* tris = nullptr;
    int nt = 0;
    std::vector<V3> tbmin, tbmax, tcent;
    std::vector<int> idx;
    std::vector<Node> nodes;
    V3 rayDir{0, 0, 0};

    V3 vert(int t, int c) const { return load3(verts, tris[3 * t + c]); }

    // KdTree.h:91-121
    void build(const double* v, const int* t, int n)
    {
        verts = v;
        tris = t;
        nt = n;
        idx.resize(n);
        std::iota(idx.begin(), idx.end(), 0);
        tbmin.resize(n);
        tbmax.resize(n);
        tcent.resize(n);
        for (int i = 0; i < n; ++i)
        {
            const V3 a = vert(i, 0), b = vert(i, 1), c = vert(i, 2);
            tbmin[i] = cmin(cmin(a, b), c);
            tbmax[i] = cmax(cmax(a, b), c);
            tcent[i] = ((a + b) + c) / 3.0;
        }
        nodes.clear();
        if (n > 0)
        {
            nodes.reserve(2 * size_t(n));
            buildRec(0, n);
        }
        // KdTree.h:330-333, Eigen normalized(): v / sqrt(squaredNorm)
        const V3 d0 = v3(1.0, 0.3713906763541037, 0.52999894000318);
        rayDir = d0 / std::sqrt(sqnorm(d0));
    }

    // KdTree.h:188-242
    int buildRec(int lo, int hi)
    {
        const int nodeIdx = static_cast<int>(nodes.size());
        nodes.emplace_back();
        const double big = std::numeric_limits<double>::max();
        V3 bmin = v3(big, big, big), bmax = v3(-big, -big, -big);
        V3 cmn = bmin, cmx = bmax;
        for (int i = lo; i < hi; ++i)
        {
            const int t = idx[i];
            bmin = cmin(bmin, tbmin[t]);
            bmax = cmax(bmax, tbmax[t]);
            cmn = cmin(cmn, tcent[t]);
            cmx = cmax(cmx, tcent[t]);
        }
        nodes[nodeIdx].bmin = bmin;
        nodes[nodeIdx].bmax = bmax;
        const int count = hi - lo;
        if (count <= 8)
        {
            nodes[nodeIdx].triStart = lo;
            nodes[nodeIdx].triCount = count;
            return nodeIdx;
        }
        const V3 ext = cmx - cmn;
        int axis = 0;
        if (ext.y > ext.x) axis = 1;
        if (ext.z > comp(ext, axis)) axis = 2;
        const int mid = lo + count / 2;
        std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi,
                         [&](int a, int b) { return comp(tcent[a], axis) < comp(tcent[b], axis); });
        const int l = buildRec(lo, mid);
        const int r = buildRec(mid, hi);
        nodes[nodeIdx].left = l;
        nodes[nodeIdx].right = r;
        return nodeIdx;
    }

    // KdTree.h:244-264
    static double sqDistAabb(V3 p, V3 bmin, V3 bmax)
    {
        double sq = 0.0;
        for (int axis = 0; axis < 3; ++axis)
        {
            const double pa = comp(p, axis), lo = comp(bmin, axis), hi = comp(bmax, axis);
            if (pa < lo)
            {
                const double d = lo - pa;
                sq += d * d;
            }
            else if (pa > hi)
            {
                const double d = pa - hi;
                sq += d * d;
            }
        }
        return sq;
    }

    // KdTree.h:266-328 (recursive near-first; strict < in leaves)
    void searchNearest(int ni, V3 p, NearestResult& best) const
    {
        const Node& node = nodes[ni];
        ++best.boxTests;
        if (sqDistAabb(p, node.bmin, node.bmax) > best.bestSq) return;
        if (node.triCount > 0)
        {
            for (int i = node.triStart; i < node.triStart + node.triCount; ++i)
            {
                const int t = idx[i];
                const V3 a = vert(t, 0), b = vert(t, 1), c = vert(t, 2);
                const V3 cp = closestPointOnTriangle(p, a, b, c);
                const double sq = sqnorm(p - cp);
                ++best.triTests;
                if (sq < best.bestSq)
                {
                    best.bestSq = sq;
                    best.point = cp;
                    best.tri = t;
                    best.distance = std::sqrt(sq);
                    const V3 n = cross(b - a, c - a);
                    const double nLen = norm(n);
                    best.faceNormal = nLen > 1.0e-15 ? n / nLen : v3(0.0, 0.0, 1.0);
                }
            }
            return;
        }
        const double leftSq = sqDistAabb(p, nodes[node.left].bmin, nodes[node.left].bmax);
        const double rightSq = sqDistAabb(p, nodes[node.right].bmin, nodes[node.right].bmax);
        best.boxTests += 2;
        int first = node.left, second = node.right;
        double firstSq = leftSq, secondSq = rightSq;
        if (rightSq < leftSq)
        {
            std::swap(first, second);
            std::swap(firstSq, secondSq);
        }
        if (firstSq <= best.bestSq) searchNearest(first, p, best);
        if (secondSq <= best.bestSq) searchNearest(second, p, best);
    }

    // KdTree.h:123-133
    NearestResult findNearest(V3 p) const
    {
        NearestResult best;
        best.distance = std::numeric_limits<double>::max();
        best.bestSq = std::numeric_limits<double>::max();
        if (!nodes.empty()) searchNearest(0, p, best);
        return best;
    }

    // KdTree.h:335-360
    static bool rayAabb(V3 o, V3 d, V3 bmin, V3 bmax)
    {
        double tmin = 0.0;
        double tmax = std::numeric_limits<double>::max();
        for (int axis = 0; axis < 3; ++axis)
        {
            const double invD = 1.0 / comp(d, axis);
            double t0 = (comp(bmin, axis) - comp(o, axis)) * invD;
            double t1 = (comp(bmax, axis) - comp(o, axis)) * invD;
            if (t0 > t1) std::swap(t0, t1);
            tmin = smax(tmin, t0);
            tmax = smin(tmax, t1);
            if (tmax < tmin) return false;
        }
        return true;
    }

    // KdTree.h:362-396 (Moller-Trumbore 1997, eps = 1e-12)
    static bool rayTriangle(V3 o, V3 d, V3 a, V3 b, V3 c)
    {
        const double eps = 1.0e-12;
        const V3 e1 = b - a;
        const V3 e2 = c - a;
        const V3 h = cross(d, e2);
        const double det = dot(e1, h);
        if (std::abs(det) < eps) return false;
        const double invDet = 1.0 / det;
        const V3 s = o - a;
        const double u = invDet * dot(s, h);
        if (u < -eps || u > 1.0 + eps) return false;
        const V3 q = cross(s, e1);
        const double v = invDet * dot(d, q);
        if (v < -eps || u + v > 1.0 + eps) return false;
        const double t = invDet * dot(e2, q);
        return t > eps;
    }

    // KdTree.h:398-441 (no early exit: every box the ray crosses is visited)
    int rayCountRec(int ni, V3 o) const
    {
        const Node& node = nodes[ni];
        if (!rayAabb(o, rayDir, node.bmin, node.bmax)) return 0;
        if (node.triCount > 0)
        {
            int hits = 0;
            for (int i = node.triStart; i < node.triStart + node.triCount; ++i)
            {
                const int t = idx[i];
                if (rayTriangle(o, rayDir, vert(t, 0), vert(t, 1), vert(t, 2))) ++hits;
            }
            return hits;
        }
        return rayCountRec(node.left, o) + rayCountRec(node.right, o);
    }
    int rayCount(V3 o) const { return nodes.empty() ? 0 : rayCountRec(0, o); }

    struct Signed
    {
        double sd;
        V3 cp;
        V3 normal;
        int tri;
        int hits;
        int64_t triTests, boxTests;
    };

    // KdTree.h:143-165
    Signed findSignedNearest(V3 p) const
    {
        const NearestResult cp = findNearest(p);
        Signed sr;
        sr.cp = cp.point;
        sr.tri = cp.tri;
        sr.normal = cp.faceNormal;
        sr.triTests = cp.triTests;
        sr.boxTests = cp.boxTests;
        const int hits = rayCount(p);
        sr.hits = hits;
        if (hits > 0)
        {
            const bool inside = (hits % 2) == 1;
            sr.sd = inside ? -cp.distance : cp.distance;
        }
        else
        {
            const V3 diff = p - cp.point;
            sr.sd = dot(diff, cp.faceNormal) >= 0.0 ? cp.distance : -cp.distance;
        }
        return sr;
    }
};

// SdfBuilder.cpp:1070-1085 / 946-958
V3 voxelGradient(V3 center, const Bvh::Signed& sr)
{
    const V3 diff = center - sr.cp;
    const double dist = norm(diff);
    const double sign = (sr.sd >= 0.0) ? 1.0 : -1.0;
    if (dist > 1e-12) return (sign * diff) / dist;
    return sign * sr.normal;
}

// SdfBuilder.cpp:1105-1152 (dense central differences, clamped, symmetrised)
void estimateHessian(const Grid& g, const double* n3, double* H6)
{
    const double vx = g.voxel.x, vy = g.voxel.y, vz = g.voxel.z;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    auto nrm = [&](int i, int j, int k) {
        i = std::max(0, std::min(i, nx - 1));
        j = std::max(0, std::min(j, ny - 1));
        k = std::max(0, std::min(k, nz - 1));
        return load3(n3, g.lin(i, j, k));
    };
    const int64_t total = g.total();
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < total; ++idx)
    {
        const int i = int(idx % nx);
        const int j = int((idx / nx) % ny);
        const int k = int(idx / (int64_t(nx) * ny));
        const V3 dx = (nrm(i + 1, j, k) - nrm(i - 1, j, k)) / (2.0 * vx);
        const V3 dy = (nrm(i, j + 1, k) - nrm(i, j - 1, k)) / (2.0 * vy);
        const V3 dz = (nrm(i, j, k + 1) - nrm(i, j, k - 1)) / (2.0 * vz);
        // H.col(0)=dx, col(1)=dy, col(2)=dz; H = 0.5 (H + H^T)
        double* h = H6 + 6 * idx;
        h[0] = 0.5 * (dx.x + dx.x);
        h[1] = 0.5 * (dy.x + dx.y);
        h[2] = 0.5 * (dz.x + dx.z);
        h[3] = 0.5 * (dy.y + dy.y);
        h[4] = 0.5 * (dz.y + dy.z);
        h[5] = 0.5 * (dz.z + dz.z);
    }
}

// ---------------------------------------------------------------------------
// Sparse container -- SdfData.h:120-303
// ---------------------------------------------------------------------------
struct SparseView
{
    Grid g;
    int blockDim;
    int nblocks;
    const int* bijk;       // 3 per block
    const float* phi;      // B^3 per block
    const float* nrm;      // 3 B^3 per block
    const float* hes;      // 6 B^3 per block or null
    int stride;            // candidate stride (0 = none)
    const uint8_t* ccount; // B^3 per block
    const float* cphi;     // B^3*stride per block
    const float* cnrm;     // 3 B^3 stride per block
    const int* cbranch;    // B^3 stride per block
    std::unordered_map<uint64_t, int> lookup;

    static uint64_t key(int bi, int bj, int bk)
    {
        return uint64_t(uint32_t(bi)) | (uint64_t(uint32_t(bj)) << 21) |
               (uint64_t(uint32_t(bk)) << 42);
    }
    int bvc() const { return blockDim * blockDim * blockDim; }
    void rebuild()
    {
        lookup.clear();
        lookup.reserve(nblocks);
        for (int b = 0; b < nblocks; ++b)
            lookup.emplace(key(bijk[3 * b], bijk[3 * b + 1], bijk[3 * b + 2]), b);
    }
    // SdfData.h:216-231
    int blockForVoxel(int i, int j, int k) const
    {
        if (!g.inBounds(i, j, k) || blockDim <= 0) return -1;
        auto it = lookup.find(key(i / blockDim, j / blockDim, k / blockDim));
        return it == lookup.end() ? -1 : it->second;
    }
    // SdfData.h:233-302
    bool sample(int i, int j, int k, double& phi0, V3& n0, double* h6 = nullptr) const
    {
        const int b = blockForVoxel(i, j, k);
        if (b < 0) return false;
        const int lx = i - bijk[3 * b] * blockDim;
        const int ly = j - bijk[3 * b + 1] * blockDim;
        const int lz = k - bijk[3 * b + 2] * blockDim;
        if (lx < 0 || lx >= blockDim || ly < 0 || ly >= blockDim || lz < 0 || lz >= blockDim)
            return false;
        const int64_t local = lx + blockDim * (ly + blockDim * lz);
        const int64_t base = int64_t(b) * bvc() + local;
        phi0 = double(phi[base]);
        n0 = v3(double(nrm[3 * base]), double(nrm[3 * base + 1]), double(nrm[3 * base + 2]));
        if (h6 != nullptr)
        {
            if (hes == nullptr) return false;
            for (int c = 0; c < 6; ++c) h6[c] = double(hes[6 * base + c]);
        }
        return true;
    }
    bool hasCandidates() const { return stride > 0 && nblocks > 0 && ccount != nullptr; }
};

// ---------------------------------------------------------------------------
// Query results
// ---------------------------------------------------------------------------
struct QueryOut
{
    double gap = 0.0;
    V3 grad{0, 0, 0};    // raw gradient (Python contract, sdf_grid.py:129-136)
    V3 normal{1, 0, 0};  // reference gapNormal
    V3 witness{0, 0, 0};
    bool valid = false;
    int branch = -1;
    double hess[6] = {0, 0, 0, 0, 0, 0}; // d2 phi (xx,xy,xz,yy,yz,zz) -- tricubic/trilinear only
};

// Corner record used by trilinear / tricubic.
struct Corners
{
    double phi[2][2][2];
    double gx[2][2][2], gy[2][2][2], gz[2][2][2];
};

// Hermite basis -- SdfOracle.cpp:976-979, 1009-1012; second derivatives are
// the analytic derivatives of the same cubics.
inline double h00(double t) { return t * t * (2.0 * t - 3.0) + 1.0; }
inline double h10(double t) { return t * (t * (t - 2.0) + 1.0); }
inline double h01(double t) { return t * t * (3.0 - 2.0 * t); }
inline double h11(double t) { return t * t * (t - 1.0); }
inline double dh00(double t) { return 6.0 * t * (t - 1.0); }
inline double dh10(double t) { return t * (3.0 * t - 4.0) + 1.0; }
inline double dh01(double t) { return 6.0 * t * (1.0 - t); }
inline double dh11(double t) { return t * (3.0 * t - 2.0); }
inline double ddh00(double t) { return 12.0 * t - 6.0; }
inline double ddh10(double t) { return 6.0 * t - 4.0; }
inline double ddh01(double t) { return 6.0 - 12.0 * t; }
inline double ddh11(double t) { return 6.0 * t - 2.0; }

// Tricubic Hermite -- SdfOracle.cpp:975-1091 (value, gradient) plus the
// analytic Hessian of the same interpolant (no counterpart in the reference).
void tricubicEval(const Corners& c, double u, double v, double w, V3 h, QueryOut& r, bool hess)
{
    const double a[4] = {h00(u), h10(u), h01(u), h11(u)};
    const double b[4] = {h00(v), h10(v), h01(v), h11(v)};
    const double cc[4] = {h00(w), h10(w), h01(w), h11(w)};
    const double da[4] = {dh00(u), dh10(u), dh01(u), dh11(u)};
    const double db[4] = {dh00(v), dh10(v), dh01(v), dh11(v)};
    const double dc[4] = {dh00(w), dh10(w), dh01(w), dh11(w)};
    double gap = 0.0, fu = 0.0, fv = 0.0, fw = 0.0;
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di)
            {
                const double p0 = a[2 * di], p1 = a[2 * di + 1];
                const double q0 = b[2 * dj], q1 = b[2 * dj + 1];
                const double r0 = cc[2 * dk], r1 = cc[2 * dk + 1];
                gap += (p0 * q0 * r0) * c.phi[di][dj][dk];
                gap += (p1 * q0 * r0) * c.gx[di][dj][dk];
                gap += (p0 * q1 * r0) * c.gy[di][dj][dk];
                gap += (p0 * q0 * r1) * c.gz[di][dj][dk];
            }
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di)
            {
                const double dp0 = da[2 * di], dp1 = da[2 * di + 1];
                const double q0 = b[2 * dj], q1 = b[2 * dj + 1];
                const double r0 = cc[2 * dk], r1 = cc[2 * dk + 1];
                fu += (dp0 * q0 * r0) * c.phi[di][dj][dk];
                fu += (dp1 * q0 * r0) * c.gx[di][dj][dk];
                fu += (dp0 * q1 * r0) * c.gy[di][dj][dk];
                fu += (dp0 * q0 * r1) * c.gz[di][dj][dk];
            }
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di)
            {
                const double p0 = a[2 * di], p1 = a[2 * di + 1];
                const double dq0 = db[2 * dj], dq1 = db[2 * dj + 1];
                const double r0 = cc[2 * dk], r1 = cc[2 * dk + 1];
                fv += (p0 * dq0 * r0) * c.phi[di][dj][dk];
                fv += (p1 * dq0 * r0) * c.gx[di][dj][dk];
                fv += (p0 * dq1 * r0) * c.gy[di][dj][dk];
                fv += (p0 * dq0 * r1) * c.gz[di][dj][dk];
            }
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di)
            {
                const double p0 = a[2 * di], p1 = a[2 * di + 1];
                const double q0 = b[2 * dj], q1 = b[2 * dj + 1];
                const double dr0 = dc[2 * dk], dr1 = dc[2 * dk + 1];
                fw += (p0 * q0 * dr0) * c.phi[di][dj][dk];
                fw += (p1 * q0 * dr0) * c.gx[di][dj][dk];
                fw += (p0 * q1 * dr0) * c.gy[di][dj][dk];
                fw += (p0 * q0 * dr1) * c.gz[di][dj][dk];
            }
    r.gap = gap;
    r.grad = v3(fu / h.x, fv / h.y, fw / h.z);
    if (hess)
    {
        const double ea[4] = {ddh00(u), ddh10(u), ddh01(u), ddh11(u)};
        const double eb[4] = {ddh00(v), ddh10(v), ddh01(v), ddh11(v)};
        const double ec[4] = {ddh00(w), ddh10(w), ddh01(w), ddh11(w)};
        // d2/(dx_s dx_t): basis factor per axis is value (0), first (1) or second (2) derivative.
        const int ord[6][3] = {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}};
        const double hs[6] = {h.x * h.x, h.x * h.y, h.x * h.z, h.y * h.y, h.y * h.z, h.z * h.z};
        for (int e = 0; e < 6; ++e)
        {
            const double* A = ord[e][0] == 0 ? a : (ord[e][0] == 1 ? da : ea);
            const double* B = ord[e][1] == 0 ? b : (ord[e][1] == 1 ? db : eb);
            const double* C = ord[e][2] == 0 ? cc : (ord[e][2] == 1 ? dc : ec);
            double s = 0.0;
            for (int dk = 0; dk < 2; ++dk)
                for (int dj = 0; dj < 2; ++dj)
                    for (int di = 0; di < 2; ++di)
                    {
                        s += (A[2 * di] * B[2 * dj] * C[2 * dk]) * c.phi[di][dj][dk];
                        s += (A[2 * di + 1] * B[2 * dj] * C[2 * dk]) * c.gx[di][dj][dk];
                        s += (A[2 * di] * B[2 * dj + 1] * C[2 * dk]) * c.gy[di][dj][dk];
                        s += (A[2 * di] * B[2 * dj] * C[2 * dk + 1]) * c.gz[di][dj][dk];
                    }
            r.hess[e] = s / hs[e];
        }
    }
}

// Trilinear -- SdfOracle.cpp:551-583 (value + analytic gradient of the
// interpolant) plus its mixed second derivatives (diagonal terms are zero).
void trilinearEval(const Corners& c, double u, double v, double w, V3 h, QueryOut& r, bool hess)
{
    const double c000 = c.phi[0][0][0], c100 = c.phi[1][0][0], c010 = c.phi[0][1][0],
                 c110 = c.phi[1][1][0], c001 = c.phi[0][0][1], c101 = c.phi[1][0][1],
                 c011 = c.phi[0][1][1], c111 = c.phi[1][1][1];
    const double c00 = c000 * (1 - u) + c100 * u;
    const double c10 = c010 * (1 - u) + c110 * u;
    const double c01 = c001 * (1 - u) + c101 * u;
    const double c11 = c011 * (1 - u) + c111 * u;
    const double c0 = c00 * (1 - v) + c10 * v;
    const double c1 = c01 * (1 - v) + c11 * v;
    r.gap = c0 * (1 - w) + c1 * w;
    const double dfdx = ((c100 - c000) * (1 - v) * (1 - w) + (c110 - c010) * v * (1 - w) +
                         (c101 - c001) * (1 - v) * w + (c111 - c011) * v * w) / h.x;
    const double dfdy = ((c010 - c000) * (1 - u) * (1 - w) + (c110 - c100) * u * (1 - w) +
                         (c011 - c001) * (1 - u) * w + (c111 - c101) * u * w) / h.y;
    const double dfdz = ((c001 - c000) * (1 - u) * (1 - v) + (c101 - c100) * u * (1 - v) +
                         (c011 - c010) * (1 - u) * v + (c111 - c110) * u * v) / h.z;
    r.grad = v3(dfdx, dfdy, dfdz);
    if (hess)
    {
        const double dxy = ((c110 - c010) - (c100 - c000)) * (1 - w) +
                           ((c111 - c011) - (c101 - c001)) * w;
        const double dxz = ((c101 - c001) - (c100 - c000)) * (1 - v) +
                           ((c111 - c011) - (c110 - c010)) * v;
        const double dyz = ((c011 - c001) - (c010 - c000)) * (1 - u) +
                           ((c111 - c101) - (c110 - c100)) * u;
        r.hess[0] = 0.0;
        r.hess[1] = dxy / (h.x * h.y);
        r.hess[2] = dxz / (h.x * h.z);
        r.hess[3] = 0.0;
        r.hess[4] = dyz / (h.y * h.z);
        r.hess[5] = 0.0;
    }
}

// Shared tail of trilinear/tricubic -- SdfOracle.cpp:585-595, 1082-1091
void finishInterp(V3 p, QueryOut& r)
{
    const double gLen = norm(r.grad);
    r.normal = gLen > 1e-15 ? r.grad / gLen : v3(1.0, 0.0, 0.0);
    r.witness = p - r.gap * r.normal;
    r.valid = true;
}

// Continuous index + validity -- SdfOracle.cpp:530-549, 937-956
bool cellOf(const Grid& g, V3 p, int& i0, int& j0, int& k0, double& u, double& v, double& w)
{
    const double fx = (p.x - g.bmin.x) / g.voxel.x - 0.5;
    const double fy = (p.y - g.bmin.y) / g.voxel.y - 0.5;
    const double fz = (p.z - g.bmin.z) / g.voxel.z - 0.5;
    i0 = static_cast<int>(std::floor(fx));
    j0 = static_cast<int>(std::floor(fy));
    k0 = static_cast<int>(std::floor(fz));
    u = fx - i0;
    v = fy - j0;
    w = fz - k0;
    return !(i0 < 0 || i0 + 1 >= g.n[0] || j0 < 0 || j0 + 1 >= g.n[1] || k0 < 0 ||
             k0 + 1 >= g.n[2]);
}

struct DenseView
{
    Grid g;
    const double* phi;
    const double* nrm;
    const double* hes; // 6 per voxel, may be null
    int stride;
    const uint8_t* ccount;
    const double* cphi;
    const double* cnrm;
    const int* cbranch;
    bool hasCandidates() const { return stride > 0 && ccount != nullptr; }
};

// SdfOracle.cpp:494-520
QueryOut denseFirstOrder(const DenseView& d, V3 p)
{
    QueryOut r;
    int i, j, k;
    d.g.indexFromWorld(p, i, j, k);
    if (!d.g.inBounds(i, j, k)) return r;
    const int64_t l = d.g.lin(i, j, k);
    const V3 dx = p - d.g.center(i, j, k);
    const V3 n0 = load3(d.nrm, l);
    r.gap = d.phi[l] + dot(n0, dx);
    r.grad = n0;
    r.normal = n0;
    r.witness = p - r.gap * r.normal;
    r.valid = true;
    return r;
}

// Second-order Taylor core -- SdfOracle.cpp:471-489, 647-653.
// 0.5 * dx^T H dx is evaluated as 0.5 * ((dx^T H) . dx), row vector first.
void secondOrderCore(double phi0, V3 n0, const double* h6, V3 p, V3 dx, QueryOut& r)
{
    const double H[3][3] = {{h6[0], h6[1], h6[2]}, {h6[1], h6[3], h6[4]}, {h6[2], h6[4], h6[5]}};
    const V3 rowv = v3(dx.x * H[0][0] + dx.y * H[1][0] + dx.z * H[2][0],
                       dx.x * H[0][1] + dx.y * H[1][1] + dx.z * H[2][1],
                       dx.x * H[0][2] + dx.y * H[1][2] + dx.z * H[2][2]);
    const double first = dot(n0, dx);
    const double second = 0.5 * dot(rowv, dx);
    r.gap = phi0 + first + second;
    const V3 Hdx = v3(H[0][0] * dx.x + H[0][1] * dx.y + H[0][2] * dx.z,
                      H[1][0] * dx.x + H[1][1] * dx.y + H[1][2] * dx.z,
                      H[2][0] * dx.x + H[2][1] * dx.y + H[2][2] * dx.z);
    const V3 upd = n0 + Hdx;
    const double nLen = norm(upd);
    r.grad = upd;
    r.normal = nLen > 1e-15 ? upd / nLen : n0;
    r.witness = p - r.gap * r.normal;
    r.valid = true;
}

// SdfOracle.cpp:448-490
QueryOut denseSecondOrder(const DenseView& d, V3 p)
{
    QueryOut r;
    int i, j, k;
    d.g.indexFromWorld(p, i, j, k);
    if (!d.g.inBounds(i, j, k)) return r;
    const int64_t l = d.g.lin(i, j, k);
    static const double zero6[6] = {0, 0, 0, 0, 0, 0};
    secondOrderCore(d.phi[l], load3(d.nrm, l), d.hes ? d.hes + 6 * l : zero6, p,
                    p - d.g.center(i, j, k), r);
    return r;
}

// SdfOracle.cpp:524-596 and 932-1092
QueryOut denseInterp(const DenseView& d, V3 p, bool cubic, bool hess)
{
    QueryOut r;
    int i0, j0, k0;
    double u, v, w;
    if (!cellOf(d.g, p, i0, j0, k0, u, v, w)) return r;
    Corners c;
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di)
            {
                const int64_t l = d.g.lin(i0 + di, j0 + dj, k0 + dk);
                c.phi[di][dj][dk] = d.phi[l];
                const V3 n = load3(d.nrm, l);
                c.gx[di][dj][dk] = n.x * d.g.voxel.x;
                c.gy[di][dj][dk] = n.y * d.g.voxel.y;
                c.gz[di][dj][dk] = n.z * d.g.voxel.z;
            }
    if (cubic)
        tricubicEval(c, u, v, w, d.g.voxel, r, hess);
    else
        trilinearEval(c, u, v, w, d.g.voxel, r, hess);
    finishInterp(p, r);
    return r;
}

// Candidate evaluation -- SdfOracle.cpp:387-431 (dense), 307-360 (sparse)
template <typename NF>
void candidateSelect(V3 p, V3 center, int count, NF fetch, QueryOut& r)
{
    const V3 dx = p - center;
    int selected = 0;
    double selAbs = std::numeric_limits<double>::max();
    double gaps[8];
    V3 nrms[8];
    int br[8];
    for (int c = 0; c < count; ++c)
    {
        double phi0;
        V3 n;
        fetch(c, phi0, n, br[c]);
        const double nn = norm(n);
        n = nn > 1.0e-15 ? n / nn : v3(1.0, 0.0, 0.0);
        gaps[c] = phi0 + dot(n, dx);
        nrms[c] = n;
        const double ag = std::abs(gaps[c]);
        if (ag < selAbs)
        {
            selAbs = ag;
            selected = c;
        }
    }
    r.gap = gaps[selected];
    r.normal = nrms[selected];
    r.grad = nrms[selected];
    r.witness = p - r.gap * r.normal;
    r.branch = br[selected];
    r.valid = true;
}

// SdfOracle.cpp:363-431 (dense queryFast ContactAware)
QueryOut denseContactAware(const DenseView& d, V3 p, bool hess)
{
    if (!d.hasCandidates()) return denseInterp(d, p, true, hess);
    QueryOut r;
    int i, j, k;
    d.g.indexFromWorld(p, i, j, k);
    if (!d.g.inBounds(i, j, k)) return r;
    const int64_t vox = d.g.lin(i, j, k);
    const int count = std::min(int(d.ccount[vox]), d.stride);
    if (count <= 0) return denseFirstOrder(d, p);
    candidateSelect(p, d.g.center(i, j, k), std::min(count, 8),
                    [&](int c, double& ph, V3& n, int& b) {
                        const int64_t idx = vox * d.stride + c;
                        ph = d.cphi[idx];
                        n = load3(d.cnrm, idx);
                        b = d.cbranch[idx];
                    },
                    r);
    return r;
}

QueryOut denseQuery(const DenseView& d, V3 p, int mode, bool hess)
{
    switch (mode)
    {
    case 0: return denseFirstOrder(d, p);
    case 1: return denseSecondOrder(d, p);
    case 2: return denseInterp(d, p, false, hess);
    case 3: return denseInterp(d, p, true, hess);
    case 4: return denseContactAware(d, p, hess);
    default: return denseSecondOrder(d, p);
    }
}

// SdfOracle.cpp:600-622
QueryOut sparseFirstOrder(const SparseView& s, V3 p)
{
    QueryOut r;
    int i, j, k;
    s.g.indexFromWorld(p, i, j, k);
    double phi0;
    V3 n0;
    if (!s.sample(i, j, k, phi0, n0, nullptr)) return r;
    const V3 dx = p - s.g.center(i, j, k);
    r.gap = phi0 + dot(n0, dx);
    r.grad = n0;
    r.normal = n0;
    r.witness = p - r.gap * r.normal;
    r.valid = true;
    return r;
}

// SdfOracle.cpp:624-654
QueryOut sparseSecondOrder(const SparseView& s, V3 p)
{
    if (s.hes == nullptr) return sparseFirstOrder(s, p);
    QueryOut r;
    int i, j, k;
    s.g.indexFromWorld(p, i, j, k);
    double phi0, h6[6];
    V3 n0;
    if (!s.sample(i, j, k, phi0, n0, h6)) return r;
    secondOrderCore(phi0, n0, h6, p, p - s.g.center(i, j, k), r);
    return r;
}

// SdfOracle.cpp:656-728, 730-851
QueryOut sparseInterp(const SparseView& s, V3 p, bool cubic, bool hess)
{
    QueryOut r;
    int i0, j0, k0;
    double u, v, w;
    if (!cellOf(s.g, p, i0, j0, k0, u, v, w)) return r;
    Corners c;
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di)
            {
                V3 n;
                if (!s.sample(i0 + di, j0 + dj, k0 + dk, c.phi[di][dj][dk], n, nullptr)) return r;
                c.gx[di][dj][dk] = n.x * s.g.voxel.x;
                c.gy[di][dj][dk] = n.y * s.g.voxel.y;
                c.gz[di][dj][dk] = n.z * s.g.voxel.z;
            }
    if (cubic)
        tricubicEval(c, u, v, w, s.g.voxel, r, hess);
    else
        trilinearEval(c, u, v, w, s.g.voxel, r, hess);
    finishInterp(p, r);
    return r;
}

// SdfOracle.cpp:269-361
QueryOut sparseContactAware(const SparseView& s, V3 p, bool hess)
{
    if (!s.hasCandidates()) return sparseInterp(s, p, true, hess);
    QueryOut r;
    int i, j, k;
    s.g.indexFromWorld(p, i, j, k);
    const int b = s.blockForVoxel(i, j, k);
    if (b < 0) return r;
    const int lx = i - s.bijk[3 * b] * s.blockDim;
    const int ly = j - s.bijk[3 * b + 1] * s.blockDim;
    const int lz = k - s.bijk[3 * b + 2] * s.blockDim;
    if (lx < 0 || lx >= s.blockDim || ly < 0 || ly >= s.blockDim || lz < 0 || lz >= s.blockDim)
        return r;
    const int64_t local = lx + s.blockDim * (ly + s.blockDim * lz);
    const int64_t base = int64_t(b) * s.bvc() + local;
    const int count = std::min(int(s.ccount[base]), s.stride);
    if (count <= 0) return sparseFirstOrder(s, p);
    candidateSelect(p, s.g.center(i, j, k), std::min(count, 8),
                    [&](int c, double& ph, V3& n, int& br) {
                        const int64_t idx = base * s.stride + c;
                        ph = double(s.cphi[idx]);
                        n = v3(double(s.cnrm[3 * idx]), double(s.cnrm[3 * idx + 1]),
                               double(s.cnrm[3 * idx + 2]));
                        br = s.cbranch[idx];
                    },
                    r);
    return r;
}

QueryOut sparseQuery(const SparseView& s, V3 p, int mode, bool hess)
{
    switch (mode)
    {
    case 0: return sparseFirstOrder(s, p);
    case 1: return sparseSecondOrder(s, p);
    case 2: return sparseInterp(s, p, false, hess);
    case 3: return sparseInterp(s, p, true, hess);
    case 4: return sparseContactAware(s, p, hess);
    default: return sparseSecondOrder(s, p);
    }
}

void writeQuery(const QueryOut& r, int64_t i, double* gap, double* grad, double* nrm, double* wit,
                uint8_t* valid, int* branch)
{
    if (gap) gap[i] = r.gap;
    if (grad) store3(grad, i, r.grad);
    if (nrm) store3(nrm, i, r.normal);
    if (wit) store3(wit, i, r.witness);
    if (valid) valid[i] = r.valid ? 1 : 0;
    if (branch) branch[i] = r.branch;
}

// ---------------------------------------------------------------------------
// Contact-aware build helpers -- SdfBuilder.cpp:28-359
// ---------------------------------------------------------------------------
// SdfBuilder.cpp:28-43
V3 normalizedOrDefault(V3 value, V3 fallback)
{
    const double n = norm(value);
    if (n <= 1.0e-15)
    {
        const double fn = norm(fallback);
        if (fn > 1.0e-15) return fallback / fn;
        return v3(1.0, 0.0, 0.0);
    }
    return value / n;
}

// SdfBuilder.cpp:45-67
V3 barycentricOnTriangle(V3 p, V3 a, V3 b, V3 c)
{
    const V3 v0 = b - a, v1 = c - a, v2 = p - a;
    const double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    const double d20 = dot(v2, v0), d21 = dot(v2, v1);
    const double denom = d00 * d11 - d01 * d01;
    if (std::abs(denom) <= 1.0e-15) return v3(1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0);
    const double v = (d11 * d20 - d01 * d21) / denom;
    const double w = (d00 * d21 - d01 * d20) / denom;
    return v3(1.0 - v - w, v, w);
}

// SdfBuilder.cpp:116-134 (FNV-1a over quantised unit normal)
int branchIdFromNormal(V3 raw)
{
    const V3 n = normalizedOrDefault(raw, v3(1.0, 0.0, 0.0));
    auto q = [](double value) {
        value = smax(-1.0, smin(1.0, value));
        return static_cast<int32_t>(std::lround(value * 4095.0));
    };
    uint32_t hash = 2166136261u;
    const int32_t qs[3] = {q(n.x), q(n.y), q(n.z)};
    for (int32_t value : qs)
    {
        hash ^= static_cast<uint32_t>(value);
        hash *= 16777619u;
    }
    return static_cast<int>(1u + (hash & 0x3fffffffu));
}

struct CTri
{
    V3 a, b, c;
    V3 cn[3];
    V3 face;
    V3 bmin, bmax;
};

struct CCand
{
    double phi, distance;
    V3 normal, witness;
    int branch;
};

// SdfBuilder.cpp:260-279
CCand makeCandidate(V3 p, const CTri& t)
{
    const V3 cp = closestPointOnTriangle(p, t.a, t.b, t.c);
    const V3 bc = barycentricOnTriangle(cp, t.a, t.b, t.c);
    const V3 n = normalizedOrDefault((bc.x * t.cn[0] + bc.y * t.cn[1]) + bc.z * t.cn[2], t.face);
    return CCand{dot(p - cp, n), norm(p - cp), n, cp, branchIdFromNormal(n)};
}

// SdfBuilder.cpp:136-153
void mergeCandidate(std::vector<CCand>& cands, const CCand& in, double cosMerge)
{
    for (CCand& e : cands)
    {
        if (dot(e.normal, in.normal) >= cosMerge)
        {
            if (std::abs(in.phi) < std::abs(e.phi)) e = in;
            return;
        }
    }
    cands.push_back(in);
}

struct TileIndex
{
    int tileDim = 8, nx = 1, ny = 1, nz = 1;
    std::vector<std::vector<int>> ids;
    int lin(int ti, int tj, int tk) const { return ti + nx * (tj + ny * tk); }
    // SdfBuilder.cpp:102-113
    const std::vector<int>& forVoxel(int i, int j, int k) const
    {
        static const std::vector<int> empty;
        if (ids.empty() || tileDim <= 0) return empty;
        const int ti = std::max(0, std::min(nx - 1, i / tileDim));
        const int tj = std::max(0, std::min(ny - 1, j / tileDim));
        const int tk = std::max(0, std::min(nz - 1, k / tileDim));
        return ids[size_t(lin(ti, tj, tk))];
    }
};

// SdfBuilder.cpp:202-258
TileIndex buildTileIndex(const Grid& g, const std::vector<CTri>& tris, int reqTileDim, double infl)
{
    TileIndex t;
    t.tileDim = std::max(1, reqTileDim);
    t.nx = (g.n[0] + t.tileDim - 1) / t.tileDim;
    t.ny = (g.n[1] + t.tileDim - 1) / t.tileDim;
    t.nz = (g.n[2] + t.tileDim - 1) / t.tileDim;
    t.ids.resize(size_t(t.nx) * t.ny * t.nz);
    const double h = g.hmax();
    const double tileRadius = 0.5 * std::sqrt(3.0) * static_cast<double>(t.tileDim) * h;
    const double expansion = smax(0.0, infl) + tileRadius + h;
    auto fl = [](double value, double bmin, double voxel) {
        return static_cast<int>(std::floor((value - bmin) / voxel - 0.5));
    };
    auto ce = [](double value, double bmin, double voxel) {
        return static_cast<int>(std::ceil((value - bmin) / voxel - 0.5));
    };
    auto clampi = [](int v, int hi) { return std::max(0, std::min(hi, v)); };
    for (int tri = 0; tri < int(tris.size()); ++tri)
    {
        const CTri& info = tris[tri];
        const V3 lo = v3(info.bmin.x - expansion, info.bmin.y - expansion, info.bmin.z - expansion);
        const V3 hi = v3(info.bmax.x + expansion, info.bmax.y + expansion, info.bmax.z + expansion);
        const int i0 = clampi(fl(lo.x, g.bmin.x, g.voxel.x), g.n[0] - 1);
        const int j0 = clampi(fl(lo.y, g.bmin.y, g.voxel.y), g.n[1] - 1);
        const int k0 = clampi(fl(lo.z, g.bmin.z, g.voxel.z), g.n[2] - 1);
        const int i1 = clampi(ce(hi.x, g.bmin.x, g.voxel.x), g.n[0] - 1);
        const int j1 = clampi(ce(hi.y, g.bmin.y, g.voxel.y), g.n[1] - 1);
        const int k1 = clampi(ce(hi.z, g.bmin.z, g.voxel.z), g.n[2] - 1);
        for (int tk = k0 / t.tileDim; tk <= k1 / t.tileDim; ++tk)
            for (int tj = j0 / t.tileDim; tj <= j1 / t.tileDim; ++tj)
                for (int ti = i0 / t.tileDim; ti <= i1 / t.tileDim; ++ti)
                    t.ids[size_t(t.lin(ti, tj, tk))].push_back(tri);
    }
    return t;
}

// SdfBuilder.cpp:281-359
std::vector<CCand> candidatesForPoint(V3 p, const std::vector<CTri>& tris,
                                      const std::vector<int>& cand, const Bvh& tree, double tol,
                                      double cosMerge, int maxC)
{
    double minDist = std::numeric_limits<double>::max();
    for (int t : cand)
    {
        if (t < 0 || t >= int(tris.size())) continue;
        minDist = smin(minDist, makeCandidate(p, tris[size_t(t)]).distance);
    }
    int fallback = -1;
    if (!std::isfinite(minDist))
    {
        const NearestResult nr = tree.findNearest(p);
        fallback = nr.tri;
        minDist = nr.distance;
    }
    std::vector<CCand> merged;
    merged.reserve(size_t(maxC));
    const double allowed = minDist + tol;
    auto tryTri = [&](int t) {
        if (t < 0 || t >= int(tris.size())) return;
        const CCand c = makeCandidate(p, tris[size_t(t)]);
        if (c.distance <= allowed) mergeCandidate(merged, c, cosMerge);
    };
    if (cand.empty())
        tryTri(fallback);
    else
        for (int t : cand) tryTri(t);
    if (merged.empty())
    {
        if (fallback < 0) fallback = tree.findNearest(p).tri;
        tryTri(fallback);
    }
    std::sort(merged.begin(), merged.end(),
              [](const CCand& a, const CCand& b) { return std::abs(a.phi) < std::abs(b.phi); });
    if (int(merged.size()) > maxC) merged.resize(size_t(maxC));
    return merged;
}

std::vector<CTri> makeContactTris(const double* verts, const int* tris, int nt, const double* cn)
{
    std::vector<CTri> out(nt);
    for (int t = 0; t < nt; ++t)
    {
        CTri& c = out[t];
        c.a = load3(verts, tris[3 * t]);
        c.b = load3(verts, tris[3 * t + 1]);
        c.c = load3(verts, tris[3 * t + 2]);
        for (int q = 0; q < 3; ++q) c.cn[q] = load3(cn, 3 * int64_t(t) + q);
        c.face = normalizedOrDefault(cross(c.b - c.a, c.c - c.a), v3(1.0, 0.0, 0.0));
        c.bmin = cmin(cmin(c.a, c.b), c.c);
        c.bmax = cmax(cmax(c.a, c.b), c.c);
    }
    return out;
}

// ---------------------------------------------------------------------------
// Contact reduction -- contacts_sdf.py:329-385, expressed in body B's SDF frame.
// Pose record (30 doubles): R(9, row major), t(3)   : Y   = R X + t
//                          M(9), m(3)               : off = M X + m (offset from A's COM)
//                          v0(3), Om(3)             : u_rel = v0 + Om x off
// ---------------------------------------------------------------------------
struct Pose
{
    double R[9], t[3], M[9], m[3], v0[3], Om[3];
};

inline V3 matvec(const double* A, V3 x)
{
    return v3(A[0] * x.x + A[1] * x.y + A[2] * x.z, A[3] * x.x + A[4] * x.y + A[5] * x.z,
              A[6] * x.x + A[7] * x.y + A[8] * x.z);
}
inline V3 matTvec(const double* A, V3 x)
{
    return v3(A[0] * x.x + A[3] * x.y + A[6] * x.z, A[1] * x.x + A[4] * x.y + A[7] * x.z,
              A[2] * x.x + A[5] * x.y + A[8] * x.z);
}

struct ContactParams
{
    double kn, c, mu, reg, margin;
    int order, mode;
};

struct PointTerms
{
    bool active;
    V3 Y, off, u, G, fq, fp, fd, ft;
    double g, p, dp, gn, hess[6];
    V3 ut;
    double D;
    bool damp, fric;
};

// max(-g,0)^k as numpy's ** with an integer exponent.
inline double powi(double x, int k)
{
    double r = 1.0;
    for (int i = 0; i < k; ++i) r *= x;
    return r;
}

template <typename QF>
PointTerms pointForward(const Pose& P, V3 X, const ContactParams& cp, QF query)
{
    PointTerms t{};
    t.active = false;
    t.Y = matvec(P.R, X) + v3(P.t[0], P.t[1], P.t[2]);
    t.off = matvec(P.M, X) + v3(P.m[0], P.m[1], P.m[2]);
    if (cp.margin > 0.0 && !(std::abs(t.Y.z) < cp.margin)) return t; // contacts_sdf.py:91-98
    const QueryOut q = query(t.Y);
    if (!q.valid) return t;
    t.g = q.gap;
    const double delta = smax(-q.gap, 0.0);
    t.p = cp.kn * powi(delta, cp.order);
    if (!(t.p > 1e-30)) return t; // contacts_sdf.py:329
    t.active = true;
    t.dp = (cp.order >= 1 && q.gap < 0.0) ? -cp.kn * cp.order * powi(delta, cp.order - 1) : 0.0;
    for (int e = 0; e < 6; ++e) t.hess[e] = q.hess[e];
    t.G = q.grad;
    t.gn = norm(t.G);
    t.u = v3(P.v0[0], P.v0[1], P.v0[2]) + cross(v3(P.Om[0], P.Om[1], P.Om[2]), t.off);
    t.fp = t.p * t.G;
    t.fd = v3(0, 0, 0);
    t.damp = cp.c > 0.0 && t.g < 0.0;
    if (t.damp)
    {
        const double vn = dot(t.u, t.G);
        t.fd = (cp.c * (-vn)) * t.G;
    }
    const V3 un = t.G / smax(t.gn, 1e-30);
    const double udn = dot(t.u, un);
    t.ut = t.u - udn * un;
    const double utn = norm(t.ut);
    t.D = std::sqrt(utn * utn + cp.reg * cp.reg);
    const double phat = t.p * t.gn;
    t.ft = v3(0, 0, 0);
    t.fric = cp.mu * phat > 0.0;
    if (t.fric) t.ft = (-cp.mu * phat) * t.ut / t.D;
    t.fq = (t.fp + t.fd) + t.ft;
    return t;
}

Pose loadPose(const double* p30)
{
    Pose P;
    std::memcpy(P.R, p30, 9 * sizeof(double));
    std::memcpy(P.t, p30 + 9, 3 * sizeof(double));
    std::memcpy(P.M, p30 + 12, 9 * sizeof(double));
    std::memcpy(P.m, p30 + 21, 3 * sizeof(double));
    std::memcpy(P.v0, p30 + 24, 3 * sizeof(double));
    std::memcpy(P.Om, p30 + 27, 3 * sizeof(double));
    return P;
}

template <typename QF>
void contactForward(const double* X, const double* w, int64_t n, const double* poses, int nb,
                    const ContactParams& cp, QF query, double* out6, int* activeCount)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b)
    {
        const Pose P = loadPose(poses + 30 * int64_t(b));
        double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
        int act = 0;
        for (int64_t i = 0; i < n; ++i)
        {
            const PointTerms t = pointForward(P, load3(X, i), cp, query);
            if (!t.active) continue;
            ++act;
            const V3 wf = w[i] * t.fq;
            const V3 tq = w[i] * cross(t.off, t.fq);
            F[0] += wf.x; F[1] += wf.y; F[2] += wf.z;
            T[0] += tq.x; T[1] += tq.y; T[2] += tq.z;
        }
        for (int c = 0; c < 3; ++c)
        {
            out6[6 * b + c] = F[c];
            out6[6 * b + 3 + c] = T[c];
        }
        if (activeCount) activeCount[b] = act;
    }
}

// SDFContactEngine.evaluate (contacts_sdf.py:225-385) restated in the world
// frame, step for step, for one contact pair per record.
// state (36 doubles): rA(3) RA(9) vA(3) wA(3) rB(3) RB(9) vB(3) wB(3)
// frames (24 doubles, optional): R_ab(9) t_ab(3) RAm(9) rAm(3) -- the action
//   marker relative to the base marker and the action marker's world pose
//   (contacts_sdf.py:243-249, 275-279); without them R_sdf = RB^T RA,
//   t_sdf = RB^T (rA - rB), rA_m = rA, RA_m = RA (:280-283).
// keep (nb x n, optional): point i of pair b survives the broad phase.
// pairOn (nb, optional): 0 where _check_broad_phase (:68-89) rejects the pair.
// The narrow filter |Y_z| < margin always applies (:91-98, RB is given).
// out12 = F_A, tau_A, F_B, tau_B (world frame).
struct PairState
{
    V3 rA, vA, wA, rB, vB, wB;
    double RA[9], RB[9];
};

inline PairState loadPairState(const double* s)
{
    PairState p;
    p.rA = load3(s, 0);
    std::memcpy(p.RA, s + 3, 9 * sizeof(double));
    p.vA = load3(s + 12, 0);
    p.wA = load3(s + 15, 0);
    p.rB = load3(s + 18, 0);
    std::memcpy(p.RB, s + 21, 9 * sizeof(double));
    p.vB = load3(s + 30, 0);
    p.wB = load3(s + 33, 0);
    return p;
}

// A^T B for row-major 3x3 (numpy's RB.T @ RA), sums in index order.
inline void matTmat(const double* A, const double* B, double* C)
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C[3 * r + c] = A[r] * B[c] + A[3 + r] * B[3 + c] + A[6 + r] * B[6 + c];
}

template <typename QF>
void contactEvaluate(const double* X, const double* w, int64_t n, const double* states36,
                     const double* frames24, int nb, const ContactParams& cp,
                     const uint8_t* keep, const uint8_t* pairOn, QF query, double* out12,
                     int* activeCount)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b)
    {
        const PairState S = loadPairState(states36 + 36 * int64_t(b));
        double Rs[9], RAm[9];
        V3 ts, rAm;
        if (frames24)
        {
            const double* f = frames24 + 24 * int64_t(b);
            std::memcpy(Rs, f, 9 * sizeof(double));
            ts = load3(f + 9, 0);
            std::memcpy(RAm, f + 12, 9 * sizeof(double));
            rAm = load3(f + 21, 0);
        }
        else
        {
            matTmat(S.RB, S.RA, Rs);
            ts = matTvec(S.RB, S.rA - S.rB);
            std::memcpy(RAm, S.RA, 9 * sizeof(double));
            rAm = S.rA;
        }
        V3 F = v3(0, 0, 0), T = v3(0, 0, 0);
        int act = 0;
        const bool on = pairOn == nullptr || pairOn[b] != 0;
        for (int64_t i = 0; on && i < n; ++i)
        {
            if (keep && !keep[int64_t(b) * n + i]) continue;
            const V3 x = load3(X, i);
            const V3 Y = matvec(Rs, x) + ts;
            if (!(std::abs(Y.z) < cp.margin)) continue; // _narrow_filter
            const QueryOut q = query(Y);
            if (!q.valid) continue;
            const double p = cp.kn * powi(smax(-q.gap, 0.0), cp.order);
            if (!(p > 1e-30)) continue; // valid_mask, :329
            ++act;
            const V3 rg = matvec(S.RB, q.grad); // :342
            const double gn = norm(q.grad);      // grad_norm of query_batch
            const V3 xw = rAm + matvec(RAm, x);  // :345
            const V3 offA = xw - S.rA;           // :348
            V3 fn = p * rg;                      // :351
            const double phat = p * gn;          // :352
            if (cp.c > 0.0)                      // :355-362
            {
                const V3 uAd = S.vA + cross(S.wA, offA);
                const V3 uBd = S.vB + cross(S.wB, xw - S.rB);
                const double vn = dot(uAd - uBd, rg);
                if (q.gap < 0.0) fn = fn + (cp.c * (-vn)) * rg;
            }
            const V3 uA = S.vA + cross(S.wA, offA); // :364-367
            const V3 uB = S.vB + cross(S.wB, xw - S.rB);
            const V3 urel = uA - uB;
            const V3 un = rg / smax(gn, 1e-30);     // :369
            const double udn = dot(urel, un);
            const V3 ut = urel - udn * un;
            const double utn = norm(ut);
            const double denom = std::sqrt(utn * utn + cp.reg * cp.reg);
            V3 ft = v3(0, 0, 0);
            if (cp.mu * phat > 0.0) ft = (-cp.mu * phat) * ut / denom; // :374-377
            const V3 fq = fn + ft;
            F = F + w[i] * fq;                      // :380-381
            T = T + w[i] * cross(offA, fq);
        }
        const V3 FB = -1.0 * F;                     // :383-384
        const V3 TB = -1.0 * (T + cross(S.rA - S.rB, F));
        const double o[12] = {F.x, F.y, F.z, T.x, T.y, T.z, FB.x, FB.y, FB.z, TB.x, TB.y, TB.z};
        for (int c = 0; c < 12; ++c) out12[12 * int64_t(b) + c] = on ? o[c] : 0.0;
        if (activeCount) activeCount[b] = act;
    }
}

// Analytic backward of contactForward (no counterpart in the reference; the
// derivation is in the comments).  Upstream gradient per pose gF, gT.
//   a = w (gF + gT x off)                    dL/dfq
//   dL/doff |torque = w (fq x gT)
//   dL/dY = s_g G + Hess s_G,  s_g = a.dfq/dg,  s_G = a^T dfq/dG,  s_u = a^T dfq/du
//   dL/doff += s_u x Om
//   left twist (omega, dt) on A:  dY = omega x (Y - t) + dt,  doff = omega x off
template <typename QF>
void contactBackward(const double* X, const double* w, int64_t n, const double* poses, int nb,
                     const ContactParams& cp, QF query, const double* gOut6, double* gTwist6,
                     double* gParams4, double* gX, double* gW)
{
    std::vector<double> gkn(nb, 0.0), gc(nb, 0.0), gmu(nb, 0.0);
    std::vector<double> gXb;
    const int nthreads =
#ifdef _OPENMP
        omp_get_max_threads();
#else
        1;
#endif
    if (gX) gXb.assign(size_t(nthreads) * size_t(n) * 3, 0.0);
    std::vector<double> gWb;
    if (gW) gWb.assign(size_t(nthreads) * size_t(n), 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        const Pose P = loadPose(poses + 30 * int64_t(b));
        const V3 gF = load3(gOut6 + 6 * b, 0), gT = load3(gOut6 + 6 * b + 3, 0);
        const V3 Om = v3(P.Om[0], P.Om[1], P.Om[2]);
        const V3 tvec = v3(P.t[0], P.t[1], P.t[2]);
        V3 gw = v3(0, 0, 0), gtr = v3(0, 0, 0);
        double skn = 0.0, sc = 0.0, smu = 0.0;
        for (int64_t i = 0; i < n; ++i)
        {
            const PointTerms t = pointForward(P, load3(X, i), cp, query);
            if (!t.active) continue;
            const V3 a = w[i] * (gF + cross(gT, t.off));
            V3 gOff = w[i] * cross(t.fq, gT);
            if (gW) gWb[size_t(tid) * n + i] += dot(gF, t.fq) + dot(gT, cross(t.off, t.fq));
            const V3 G = t.G;
            const double gn = t.gn;
            // d/dg
            double sg = t.dp * dot(a, G);
            if (t.fric && t.p != 0.0) sg += (t.dp / t.p) * dot(a, t.ft);
            // d/dG
            V3 sG = t.p * a;
            V3 su = v3(0, 0, 0);
            const double uG = dot(t.u, G);
            if (t.damp)
            {
                sG = sG + (-cp.c) * (dot(a, G) * t.u + uG * a);
                su = su + (-cp.c * dot(a, G)) * G;
            }
            if (t.fric)
            {
                const double k = -cp.mu * t.p;
                const V3 hvec = gn * t.ut; // gn u_t
                // a^T dh/dG
                const V3 ahG = (dot(a, t.u) / gn) * G - (dot(a, G) / gn) * t.u - (uG / gn) * a +
                               (uG * dot(a, G) / (gn * gn * gn)) * G;
                const V3 dD = (-uG / (gn * gn * t.D)) * t.ut;
                sG = sG + k * (ahG / t.D - (dot(a, hvec) / (t.D * t.D)) * dD);
                // a^T dft/du = k gn [P a / D - (a.ut) ut / D^3]
                const V3 Pa = a - (dot(a, G) / (gn * gn)) * G;
                su = su + (k * gn) * (Pa / t.D - (dot(a, t.ut) / (t.D * t.D * t.D)) * t.ut);
                smu += dot(a, t.ft) / cp.mu;
            }
            const double* h = t.hess;
            const V3 HsG = v3(h[0] * sG.x + h[1] * sG.y + h[2] * sG.z,
                              h[1] * sG.x + h[3] * sG.y + h[4] * sG.z,
                              h[2] * sG.x + h[4] * sG.y + h[5] * sG.z);
            const V3 gY = sg * G + HsG;
            gOff = gOff + cross(su, Om);
            gtr = gtr + gY;
            gw = gw + cross(t.Y - tvec, gY) + cross(t.off, gOff);
            if (cp.kn != 0.0) skn += dot(a, t.fp + t.ft) / cp.kn;
            if (t.damp) sc += dot(a, t.fd) / cp.c;
            if (gX)
            {
                const V3 gx = matTvec(P.R, gY) + matTvec(P.M, gOff);
                double* dst = &gXb[(size_t(tid) * n + i) * 3];
                dst[0] += gx.x; dst[1] += gx.y; dst[2] += gx.z;
            }
        }
        gTwist6[6 * b + 0] = gw.x; gTwist6[6 * b + 1] = gw.y; gTwist6[6 * b + 2] = gw.z;
        gTwist6[6 * b + 3] = gtr.x; gTwist6[6 * b + 4] = gtr.y; gTwist6[6 * b + 5] = gtr.z;
        gkn[b] = skn;
        gc[b] = sc;
        gmu[b] = smu;
    }
    if (gParams4)
    {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int b = 0; b < nb; ++b)
        {
            s0 += gkn[b];
            s1 += gc[b];
            s2 += gmu[b];
        }
        gParams4[0] = s0;
        gParams4[1] = s1;
        gParams4[2] = s2;
        gParams4[3] = 0.0;
    }
    if (gX)
    {
        for (int64_t i = 0; i < 3 * n; ++i)
        {
            double s = 0.0;
            for (int tid = 0; tid < nthreads; ++tid) s += gXb[size_t(tid) * 3 * n + i];
            gX[i] = s;
        }
    }
    if (gW)
    {
        for (int64_t i = 0; i < n; ++i)
        {
            double s = 0.0;
            for (int tid = 0; tid < nthreads; ++tid) s += gWb[size_t(tid) * n + i];
            gW[i] = s;
        }
    }
}

// SdfBuilder.cpp:1154-1225: central differences of the stored FP32 normals
// over the active bricks; a neighbour outside the grid or in a missing brick
// (SparseSdfData::sample returns false) falls back to the voxel's own normal.
void estimateSparseHessian(const Grid& g, int blockDim, int nb, const int* bijk, const float* nrm,
                           float* hes)
{
    const int bvc = blockDim * blockDim * blockDim;
    SparseView sv;
    sv.g = g;
    sv.blockDim = blockDim;
    sv.nblocks = nb;
    sv.bijk = bijk;
    sv.phi = nullptr;
    sv.nrm = nrm;
    sv.hes = nullptr;
    sv.stride = 0;
    sv.rebuild();
    const double vx = g.voxel.x, vy = g.voxel.y, vz = g.voxel.z;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b)
        for (int lz = 0; lz < blockDim; ++lz)
            for (int ly = 0; ly < blockDim; ++ly)
                for (int lx = 0; lx < blockDim; ++lx)
                {
                    const int i = bijk[3 * b] * blockDim + lx;
                    const int j = bijk[3 * b + 1] * blockDim + ly;
                    const int k = bijk[3 * b + 2] * blockDim + lz;
                    if (!g.inBounds(i, j, k)) continue;
                    const int64_t base = int64_t(b) * bvc + lx + blockDim * (ly + blockDim * lz);
                    const V3 cur = v3(double(nrm[3 * base]), double(nrm[3 * base + 1]),
                                      double(nrm[3 * base + 2]));
                    auto nf = [&](int ii, int jj, int kk) {
                        const int nbk = sv.blockForVoxel(ii, jj, kk);
                        if (nbk < 0) return cur;
                        const int64_t nl = int64_t(nbk) * bvc + (ii - bijk[3 * nbk] * blockDim) +
                                           blockDim * ((jj - bijk[3 * nbk + 1] * blockDim) +
                                                       blockDim * (kk - bijk[3 * nbk + 2] * blockDim));
                        return v3(double(nrm[3 * nl]), double(nrm[3 * nl + 1]), double(nrm[3 * nl + 2]));
                    };
                    const V3 dx = (nf(i + 1, j, k) - nf(i - 1, j, k)) / (2.0 * vx);
                    const V3 dy = (nf(i, j + 1, k) - nf(i, j - 1, k)) / (2.0 * vy);
                    const V3 dz = (nf(i, j, k + 1) - nf(i, j, k - 1)) / (2.0 * vz);
                    float* o = &hes[6 * base];
                    o[0] = static_cast<float>(0.5 * (dx.x + dx.x));
                    o[1] = static_cast<float>(0.5 * (dy.x + dx.y));
                    o[2] = static_cast<float>(0.5 * (dz.x + dx.z));
                    o[3] = static_cast<float>(0.5 * (dy.y + dy.y));
                    o[4] = static_cast<float>(0.5 * (dz.y + dy.z));
                    o[5] = static_cast<float>(0.5 * (dz.z + dz.z));
                }
}

// Sparse build state kept for the accessors.
struct SparseBuild
{
    Grid g;
    int blockDim, halo;
    double band, activation;
    int nbx, nby, nbz;
    std::vector<uint8_t> active;
    std::vector<int> bijk;
    std::vector<float> phi, nrm, hes;
    int stride = 0;
    std::vector<uint8_t> ccount;
    std::vector<float> cphi, cnrm;
    std::vector<int> cbranch;
};

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int oracle_num_threads()
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// SdfBuilder.cpp:994-1035.  grid9 = bmin3, bmax3, voxel3.
void oracle_setup_grid(const double* verts, int64_t nv, double padding, double voxelSize,
                       int targetRes, int maxRes, double* grid9, int* shape3)
{
    const Grid g = setupGrid(verts, nv, padding, voxelSize, targetRes, maxRes);
    gridToArrays(g, grid9, shape3);
}

int oracle_snap_resolution(int n, int maxRes) { return snapToStandardResolution(n, maxRes); }

void oracle_closest_point(const double* p, const double* a, const double* b, const double* c,
                          double* out)
{
    store3(out, 0, closestPointOnTriangle(load3(p, 0), load3(a, 0), load3(b, 0), load3(c, 0)));
}

int oracle_ray_triangle(const double* o, const double* a, const double* b, const double* c)
{
    const V3 d0 = v3(1.0, 0.3713906763541037, 0.52999894000318);
    const V3 dir = d0 / std::sqrt(sqnorm(d0));
    return Bvh::rayTriangle(load3(o, 0), dir, load3(a, 0), load3(b, 0), load3(c, 0)) ? 1 : 0;
}

// BVH export (layout check against the product's host BVH).
// nodes6: bmin3,bmax3 per node; nodeI4: left,right,triStart,triCount; perm: slot -> triangle.
int oracle_bvh_export(const double* verts, const int* tris, int nt, double* nodes6, int* nodeI4,
                      int* perm, int maxNodes)
{
    Bvh t;
    t.build(verts, tris, nt);
    const int nn = int(t.nodes.size());
    if (nn > maxNodes) return -nn;
    for (int i = 0; i < nn; ++i)
    {
        const Node& n = t.nodes[i];
        nodes6[6 * i + 0] = n.bmin.x; nodes6[6 * i + 1] = n.bmin.y; nodes6[6 * i + 2] = n.bmin.z;
        nodes6[6 * i + 3] = n.bmax.x; nodes6[6 * i + 4] = n.bmax.y; nodes6[6 * i + 5] = n.bmax.z;
        nodeI4[4 * i + 0] = n.left; nodeI4[4 * i + 1] = n.right;
        nodeI4[4 * i + 2] = n.triStart; nodeI4[4 * i + 3] = n.triCount;
    }
    for (int i = 0; i < nt; ++i) perm[i] = t.idx[i];
    return nn;
}

// KdTree::findSignedNearest per point (KdTree.h:143-165).
void oracle_signed_nearest(const double* verts, const int* tris, int nt, const double* pts,
                           int64_t n, double* sd, double* cp3, int* tri, double* face3, int* hits,
                           int64_t* triTests, int64_t* boxTests)
{
    Bvh t;
    t.build(verts, tris, nt);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i)
    {
        const Bvh::Signed s = t.findSignedNearest(load3(pts, i));
        if (sd) sd[i] = s.sd;
        if (cp3) store3(cp3, i, s.cp);
        if (tri) tri[i] = s.tri;
        if (face3) store3(face3, i, s.normal);
        if (hits) hits[i] = s.hits;
        if (triTests) triTests[i] = s.triTests;
        if (boxTests) boxTests[i] = s.boxTests;
    }
}

// O(N_t) scan: minimal squared distance and lowest index attaining it.
void oracle_brute_nearest(const double* verts, const int* tris, int nt, const double* pts,
                          int64_t n, double* sq, int* tri)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
    {
        const V3 p = load3(pts, i);
        double best = std::numeric_limits<double>::max();
        int bt = -1;
        for (int t = 0; t < nt; ++t)
        {
            const V3 a = load3(verts, tris[3 * t]), b = load3(verts, tris[3 * t + 1]),
                     c = load3(verts, tris[3 * t + 2]);
            const double s = sqnorm(p - closestPointOnTriangle(p, a, b, c));
            if (s < best)
            {
                best = s;
                bt = t;
            }
        }
        sq[i] = best;
        tri[i] = bt;
    }
}

// Generalised winding number (van Oosterom & Strackee solid angle / 4 pi).
void oracle_winding_number(const double* verts, const int* tris, int nt, const double* pts,
                           int64_t n, double* wn)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
    {
        const V3 p = load3(pts, i);
        double s = 0.0;
        for (int t = 0; t < nt; ++t)
        {
            const V3 a = load3(verts, tris[3 * t]) - p, b = load3(verts, tris[3 * t + 1]) - p,
                     c = load3(verts, tris[3 * t + 2]) - p;
            const double la = norm(a), lb = norm(b), lc = norm(c);
            const double num = dot(a, cross(b, c));
            const double den = la * lb * lc + dot(a, b) * lc + dot(b, c) * la + dot(c, a) * lb;
            s += 2.0 * std::atan2(num, den);
        }
        wn[i] = s / (4.0 * M_PI);
    }
}

// SdfBuilder::build (SdfBuilder.cpp:364-407, 1039-1152).  Outputs are dense,
// x fastest: phi[N], n[3N], H[6N] (Hxx,Hxy,Hxz,Hyy,Hyz,Hzz), witness[3N],
// tri[N] (nearest triangle index), optional per-voxel test counts.
int oracle_build_dense(const double* verts, int64_t nv, const int* tris, int nt,
                       const double* grid9, const int* shape3, double* phi, double* n3, double* H6,
                       double* witness3, int* tri, int64_t* triTests, int64_t* boxTests)
{
    if (nv <= 0 || nt <= 0) return -1;
    const Grid g = gridFromArrays(grid9, shape3);
    Bvh t;
    t.build(verts, tris, nt);
    const int64_t total = g.total();
    const int nx = g.n[0], ny = g.n[1];
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t idx = 0; idx < total; ++idx)
    {
        const int i = int(idx % nx);
        const int j = int((idx / nx) % ny);
        const int k = int(idx / (int64_t(nx) * ny));
        const V3 c = g.center(i, j, k);
        const Bvh::Signed s = t.findSignedNearest(c);
        phi[idx] = s.sd;
        if (witness3) store3(witness3, idx, s.cp);
        store3(n3, idx, voxelGradient(c, s));
        if (tri) tri[idx] = s.tri;
        if (triTests) triTests[idx] = s.triTests;
        if (boxTests) boxTests[idx] = s.boxTests;
    }
    if (H6) estimateHessian(g, n3, H6);
    return 0;
}

// SdfBuilder::buildSparse (SdfBuilder.cpp:823-983) and estimateSparseHessian
// (SdfBuilder.cpp:1154-1225).  Returns an opaque handle.
void* oracle_build_sparse(const double* verts, int64_t nv, const int* tris, int nt,
                          const double* grid9, const int* shape3, int blockDim, int halo,
                          int bandVoxels, double bandDistance, int storeHessian)
{
    if (nv <= 0 || nt <= 0 || blockDim <= 0) return nullptr;
    SparseBuild* sb = new SparseBuild;
    sb->g = gridFromArrays(grid9, shape3);
    const Grid& g = sb->g;
    sb->blockDim = blockDim;
    sb->halo = halo;
    const double h = g.hmax();
    sb->band = bandDistance > 0.0 ? bandDistance : static_cast<double>(bandVoxels) * h;
    const double blockRadius = 0.5 * std::sqrt(3.0) * static_cast<double>(blockDim) * h;
    sb->activation = sb->band + static_cast<double>(halo) * h + blockRadius;
    Bvh t;
    t.build(verts, tris, nt);
    sb->nbx = (g.n[0] + blockDim - 1) / blockDim;
    sb->nby = (g.n[1] + blockDim - 1) / blockDim;
    sb->nbz = (g.n[2] + blockDim - 1) / blockDim;
    const int nbx = sb->nbx, nby = sb->nby, nbz = sb->nbz;
    const int64_t totalBlocks = int64_t(nbx) * nby * nbz;
    sb->active.assign(size_t(totalBlocks), 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t bl = 0; bl < totalBlocks; ++bl)
    {
        const int bi = int(bl % nbx), bj = int((bl / nbx) % nby), bk = int(bl / (int64_t(nbx) * nby));
        const int i0 = bi * blockDim, j0 = bj * blockDim, k0 = bk * blockDim;
        const int i1 = std::min(i0 + blockDim, g.n[0]);
        const int j1 = std::min(j0 + blockDim, g.n[1]);
        const int k1 = std::min(k0 + blockDim, g.n[2]);
        const V3 bc = v3(g.bmin.x + g.voxel.x * (0.5 * static_cast<double>(i0 + i1)),
                         g.bmin.y + g.voxel.y * (0.5 * static_cast<double>(j0 + j1)),
                         g.bmin.z + g.voxel.z * (0.5 * static_cast<double>(k0 + k1)));
        const Bvh::Signed s = t.findSignedNearest(bc);
        sb->active[size_t(bl)] = std::abs(s.sd) <= sb->activation ? 1 : 0;
    }
    for (int64_t bl = 0; bl < totalBlocks; ++bl)
    {
        if (!sb->active[size_t(bl)]) continue;
        sb->bijk.push_back(int(bl % nbx));
        sb->bijk.push_back(int((bl / nbx) % nby));
        sb->bijk.push_back(int(bl / (int64_t(nbx) * nby)));
    }
    const int nb = int(sb->bijk.size() / 3);
    const int bvc = blockDim * blockDim * blockDim;
    sb->phi.assign(size_t(nb) * bvc, 0.0f);
    sb->nrm.assign(size_t(nb) * bvc * 3, 0.0f);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b)
    {
        for (int lz = 0; lz < blockDim; ++lz)
            for (int ly = 0; ly < blockDim; ++ly)
                for (int lx = 0; lx < blockDim; ++lx)
                {
                    const int i = sb->bijk[3 * b] * blockDim + lx;
                    const int j = sb->bijk[3 * b + 1] * blockDim + ly;
                    const int k = sb->bijk[3 * b + 2] * blockDim + lz;
                    const int64_t local = lx + blockDim * (ly + blockDim * lz);
                    const int64_t base = int64_t(b) * bvc + local;
                    if (!g.inBounds(i, j, k))
                    {
                        sb->phi[base] = 0.0f;
                        sb->nrm[3 * base] = 1.0f;
                        continue;
                    }
                    const V3 c = g.center(i, j, k);
                    const Bvh::Signed s = t.findSignedNearest(c);
                    const V3 gr = voxelGradient(c, s);
                    sb->phi[base] = static_cast<float>(s.sd);
                    sb->nrm[3 * base] = static_cast<float>(gr.x);
                    sb->nrm[3 * base + 1] = static_cast<float>(gr.y);
                    sb->nrm[3 * base + 2] = static_cast<float>(gr.z);
                }
    }
    if (storeHessian)
    {
        sb->hes.assign(size_t(nb) * bvc * 6, 0.0f);
        estimateSparseHessian(g, blockDim, nb, sb->bijk.data(), sb->nrm.data(), sb->hes.data());
    }
    return sb;
}

int oracle_sparse_info(void* h, int* nblocks, int* nbxyz3, double* bandActivation2)
{
    SparseBuild* sb = static_cast<SparseBuild*>(h);
    if (!sb) return -1;
    *nblocks = int(sb->bijk.size() / 3);
    nbxyz3[0] = sb->nbx;
    nbxyz3[1] = sb->nby;
    nbxyz3[2] = sb->nbz;
    bandActivation2[0] = sb->band;
    bandActivation2[1] = sb->activation;
    return sb->stride;
}

void oracle_sparse_copy(void* h, uint8_t* activeMask, int* bijk, float* phi, float* nrm,
                        float* hes, uint8_t* ccount, float* cphi, float* cnrm, int* cbranch)
{
    SparseBuild* sb = static_cast<SparseBuild*>(h);
    if (activeMask) std::memcpy(activeMask, sb->active.data(), sb->active.size());
    if (bijk) std::memcpy(bijk, sb->bijk.data(), sb->bijk.size() * sizeof(int));
    if (phi) std::memcpy(phi, sb->phi.data(), sb->phi.size() * sizeof(float));
    if (nrm) std::memcpy(nrm, sb->nrm.data(), sb->nrm.size() * sizeof(float));
    if (hes && !sb->hes.empty()) std::memcpy(hes, sb->hes.data(), sb->hes.size() * sizeof(float));
    if (ccount && !sb->ccount.empty()) std::memcpy(ccount, sb->ccount.data(), sb->ccount.size());
    if (cphi && !sb->cphi.empty()) std::memcpy(cphi, sb->cphi.data(), sb->cphi.size() * sizeof(float));
    if (cnrm && !sb->cnrm.empty()) std::memcpy(cnrm, sb->cnrm.data(), sb->cnrm.size() * sizeof(float));
    if (cbranch && !sb->cbranch.empty())
        std::memcpy(cbranch, sb->cbranch.data(), sb->cbranch.size() * sizeof(int));
}

void oracle_sparse_free(void* h) { delete static_cast<SparseBuild*>(h); }

// SdfBuilder::buildContactAware (SdfBuilder.cpp:416-549).  cornerNormals: 9 per triangle.
int oracle_build_contact_aware_dense(const double* verts, int64_t nv, const int* tris, int nt,
                                     const double* cornerN, const double* grid9,
                                     const int* shape3, int maxCandidates, double featureTol,
                                     double mergeDeg, int tileDim, int useTiles, int storeHessian,
                                     double* phi, double* n3, double* H6, double* witness3,
                                     uint8_t* ccount, double* cphi, double* cn3, int* cbranch)
{
    if (nv <= 0 || nt <= 0) return -1;
    if (maxCandidates <= 0) return -2;
    const Grid g = gridFromArrays(grid9, shape3);
    const int stride = std::max(0, std::min(maxCandidates, 8));
    const std::vector<CTri> ct = makeContactTris(verts, tris, nt, cornerN);
    Bvh tree;
    tree.build(verts, tris, nt);
    const double h = g.hmax();
    const double tol = featureTol > 0.0 ? featureTol : 1.5 * h;
    const double cosMerge = std::cos(mergeDeg * std::acos(-1.0) / 180.0);
    std::vector<int> all(nt);
    std::iota(all.begin(), all.end(), 0);
    TileIndex ti;
    if (useTiles) ti = buildTileIndex(g, ct, tileDim, tol + 2.0 * h);
    const int64_t total = g.total();
    const int nx = g.n[0], ny = g.n[1];
#pragma omp parallel for schedule(dynamic, 128)
    for (int64_t idx = 0; idx < total; ++idx)
    {
        const int i = int(idx % nx), j = int((idx / nx) % ny), k = int(idx / (int64_t(nx) * ny));
        const V3 c = g.center(i, j, k);
        const std::vector<int>& cand = useTiles ? ti.forVoxel(i, j, k) : all;
        const std::vector<CCand> cs = candidatesForPoint(c, ct, cand, tree, tol, cosMerge, stride);
        const int cnt = std::min(int(cs.size()), stride);
        ccount[idx] = uint8_t(cnt);
        // SdfData::resize defaults: phi 0, n UnitX, witness 0
        phi[idx] = 0.0;
        store3(n3, idx, v3(1, 0, 0));
        if (witness3) store3(witness3, idx, v3(0, 0, 0));
        for (int q = 0; q < stride; ++q)
        {
            cphi[idx * stride + q] = 0.0;
            store3(cn3, idx * stride + q, v3(1, 0, 0));
            cbranch[idx * stride + q] = -1;
        }
        if (cnt > 0)
        {
            phi[idx] = cs[0].phi;
            store3(n3, idx, cs[0].normal);
            if (witness3) store3(witness3, idx, cs[0].witness);
        }
        for (int q = 0; q < cnt; ++q)
        {
            cphi[idx * stride + q] = cs[q].phi;
            store3(cn3, idx * stride + q, cs[q].normal);
            cbranch[idx * stride + q] = cs[q].branch;
        }
    }
    if (H6)
    {
        if (storeHessian)
            estimateHessian(g, n3, H6);
        else
            std::memset(H6, 0, sizeof(double) * 6 * size_t(total));
    }
    return stride;
}

// SdfBuilder::buildSparseContactAware (SdfBuilder.cpp:551-821).
void* oracle_build_contact_aware_sparse(const double* verts, int64_t nv, const int* tris, int nt,
                                        const double* cornerN, const double* grid9,
                                        const int* shape3, int maxCandidates, double featureTol,
                                        double mergeDeg, int tileDim, int useTiles,
                                        int storeHessian, int blockDim, int halo, int bandVoxels,
                                        double bandDistance)
{
    if (nv <= 0 || nt <= 0 || maxCandidates <= 0 || blockDim <= 0) return nullptr;
    SparseBuild* sb = new SparseBuild;
    sb->g = gridFromArrays(grid9, shape3);
    const Grid& g = sb->g;
    sb->blockDim = blockDim;
    sb->halo = halo;
    const double h = g.hmax();
    sb->band = bandDistance > 0.0 ? bandDistance : static_cast<double>(bandVoxels) * h;
    sb->activation = sb->band + static_cast<double>(halo) * h;
    sb->stride = std::max(0, std::min(maxCandidates, 8));
    const std::vector<CTri> ct = makeContactTris(verts, tris, nt, cornerN);
    Bvh tree;
    tree.build(verts, tris, nt);
    const double tol = featureTol > 0.0 ? featureTol : 1.5 * h;
    TileIndex ti;
    if (useTiles) ti = buildTileIndex(g, ct, tileDim, smax(sb->activation, tol + 2.0 * h));
    const double cosMerge = std::cos(mergeDeg * std::acos(-1.0) / 180.0);
    std::vector<int> all(nt);
    std::iota(all.begin(), all.end(), 0);
    sb->nbx = (g.n[0] + blockDim - 1) / blockDim;
    sb->nby = (g.n[1] + blockDim - 1) / blockDim;
    sb->nbz = (g.n[2] + blockDim - 1) / blockDim;
    const int nbx = sb->nbx, nby = sb->nby, nbz = sb->nbz;
    const int64_t totalBlocks = int64_t(nbx) * nby * nbz;
    sb->active.assign(size_t(totalBlocks), 0);
    const bool tilesUsable = useTiles && !ti.ids.empty();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t bl = 0; bl < totalBlocks; ++bl)
    {
        const int bi = int(bl % nbx), bj = int((bl / nbx) % nby), bk = int(bl / (int64_t(nbx) * nby));
        bool act = false;
        if (tilesUsable)
        {
            const int i0 = bi * blockDim, j0 = bj * blockDim, k0 = bk * blockDim;
            const int i1 = std::min(i0 + blockDim - 1, g.n[0] - 1);
            const int j1 = std::min(j0 + blockDim - 1, g.n[1] - 1);
            const int k1 = std::min(k0 + blockDim - 1, g.n[2] - 1);
            for (int tk = k0 / ti.tileDim; tk <= k1 / ti.tileDim && !act; ++tk)
                for (int tj = j0 / ti.tileDim; tj <= j1 / ti.tileDim && !act; ++tj)
                    for (int tii = i0 / ti.tileDim; tii <= i1 / ti.tileDim; ++tii)
                        if (!ti.ids[size_t(ti.lin(tii, tj, tk))].empty())
                        {
                            act = true;
                            break;
                        }
        }
        else
        {
            const int i0 = bi * blockDim, j0 = bj * blockDim, k0 = bk * blockDim;
            const int i1 = std::min(i0 + blockDim, g.n[0]);
            const int j1 = std::min(j0 + blockDim, g.n[1]);
            const int k1 = std::min(k0 + blockDim, g.n[2]);
            const V3 bc = v3(g.bmin.x + g.voxel.x * (0.5 * static_cast<double>(i0 + i1)),
                             g.bmin.y + g.voxel.y * (0.5 * static_cast<double>(j0 + j1)),
                             g.bmin.z + g.voxel.z * (0.5 * static_cast<double>(k0 + k1)));
            const NearestResult nr = tree.findNearest(bc);
            const double br = 0.5 * std::sqrt(3.0) * static_cast<double>(blockDim) * h;
            act = nr.distance <= sb->activation + br;
        }
        sb->active[size_t(bl)] = act ? 1 : 0;
    }
    for (int64_t bl = 0; bl < totalBlocks; ++bl)
    {
        if (!sb->active[size_t(bl)]) continue;
        sb->bijk.push_back(int(bl % nbx));
        sb->bijk.push_back(int((bl / nbx) % nby));
        sb->bijk.push_back(int(bl / (int64_t(nbx) * nby)));
    }
    const int nb = int(sb->bijk.size() / 3);
    const int bvc = blockDim * blockDim * blockDim;
    const int S = sb->stride;
    sb->phi.assign(size_t(nb) * bvc, 0.0f);
    sb->nrm.assign(size_t(nb) * bvc * 3, 0.0f);
    sb->ccount.assign(size_t(nb) * bvc, 0);
    sb->cphi.assign(size_t(nb) * bvc * S, 0.0f);
    sb->cnrm.assign(size_t(nb) * bvc * S * 3, 0.0f);
    sb->cbranch.assign(size_t(nb) * bvc * S, -1);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b)
    {
        for (int lz = 0; lz < blockDim; ++lz)
            for (int ly = 0; ly < blockDim; ++ly)
                for (int lx = 0; lx < blockDim; ++lx)
                {
                    const int i = sb->bijk[3 * b] * blockDim + lx;
                    const int j = sb->bijk[3 * b + 1] * blockDim + ly;
                    const int k = sb->bijk[3 * b + 2] * blockDim + lz;
                    const int64_t local = lx + blockDim * (ly + blockDim * lz);
                    const int64_t base = int64_t(b) * bvc + local;
                    if (!g.inBounds(i, j, k))
                    {
                        sb->phi[base] = 0.0f;
                        sb->nrm[3 * base] = 1.0f;
                        continue;
                    }
                    const V3 c = g.center(i, j, k);
                    const std::vector<int>& cand = useTiles ? ti.forVoxel(i, j, k) : all;
                    const std::vector<CCand> cs = candidatesForPoint(c, ct, cand, tree, tol, cosMerge, S);
                    const int cnt = std::min(int(cs.size()), S);
                    sb->ccount[base] = uint8_t(cnt);
                    if (cnt > 0)
                    {
                        sb->phi[base] = static_cast<float>(cs[0].phi);
                        sb->nrm[3 * base] = static_cast<float>(cs[0].normal.x);
                        sb->nrm[3 * base + 1] = static_cast<float>(cs[0].normal.y);
                        sb->nrm[3 * base + 2] = static_cast<float>(cs[0].normal.z);
                    }
                    else
                    {
                        sb->phi[base] = 0.0f;
                        sb->nrm[3 * base] = 1.0f;
                        sb->nrm[3 * base + 1] = 0.0f;
                        sb->nrm[3 * base + 2] = 0.0f;
                    }
                    for (int q = 0; q < cnt; ++q)
                    {
                        const int64_t o = base * S + q;
                        sb->cphi[o] = static_cast<float>(cs[q].phi);
                        sb->cnrm[3 * o] = static_cast<float>(cs[q].normal.x);
                        sb->cnrm[3 * o + 1] = static_cast<float>(cs[q].normal.y);
                        sb->cnrm[3 * o + 2] = static_cast<float>(cs[q].normal.z);
                        sb->cbranch[o] = cs[q].branch;
                    }
                }
    }
    if (storeHessian)
    {
        sb->hes.assign(size_t(nb) * bvc * 6, 0.0f);
        estimateSparseHessian(g, blockDim, nb, sb->bijk.data(), sb->nrm.data(), sb->hes.data());
    }
    return sb;
}

// Hessian stencil alone (SdfBuilder.cpp:1105-1152) on a given normal field.
void oracle_dense_hessian(const double* grid9, const int* shape3, const double* n3, double* H6)
{
    estimateHessian(gridFromArrays(grid9, shape3), n3, H6);
}

// Dense queries, 5 modes (SdfOracle.cpp:176-432, 448-1092).
// mode: 0 FirstOrder, 1 SecondOrder, 2 Trilinear, 3 Tricubic, 4 ContactAware.
// hess6 (optional): analytic second derivatives of the trilinear/tricubic interpolant.
void oracle_query_dense(const double* grid9, const int* shape3, const double* phi,
                        const double* n3, const double* H6, int stride, const uint8_t* ccount,
                        const double* cphi, const double* cn3, const int* cbranch, int mode,
                        const double* pts, int64_t n, double* gap, double* grad3, double* normal3,
                        double* witness3, uint8_t* valid, int* branch, double* hess6)
{
    DenseView d;
    d.g = gridFromArrays(grid9, shape3);
    d.phi = phi;
    d.nrm = n3;
    d.hes = H6;
    d.stride = stride;
    d.ccount = ccount;
    d.cphi = cphi;
    d.cnrm = cn3;
    d.cbranch = cbranch;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
    {
        const QueryOut r = denseQuery(d, load3(pts, i), mode, hess6 != nullptr);
        writeQuery(r, i, gap, grad3, normal3, witness3, valid, branch);
        if (hess6)
            for (int e = 0; e < 6; ++e) hess6[6 * i + e] = r.hess[e];
    }
}

// Sparse queries (SdfOracle.cpp:600-851, 269-361) over a block list.
void oracle_query_sparse(const double* grid9, const int* shape3, int blockDim, int nblocks,
                         const int* bijk, const float* phi, const float* nrm, const float* hes,
                         int stride, const uint8_t* ccount, const float* cphi, const float* cnrm,
                         const int* cbranch, int mode, const double* pts, int64_t n, double* gap,
                         double* grad3, double* normal3, double* witness3, uint8_t* valid,
                         int* branch)
{
    SparseView s;
    s.g = gridFromArrays(grid9, shape3);
    s.blockDim = blockDim;
    s.nblocks = nblocks;
    s.bijk = bijk;
    s.phi = phi;
    s.nrm = nrm;
    s.hes = hes;
    s.stride = stride;
    s.ccount = ccount;
    s.cphi = cphi;
    s.cnrm = cnrm;
    s.cbranch = cbranch;
    s.rebuild();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        writeQuery(sparseQuery(s, load3(pts, i), mode, false), i, gap, grad3, normal3, witness3,
                   valid, branch);
}

// Contact forward on a dense grid (contacts_sdf.py:300-385).  params6 =
// kn, c, mu, reg, narrowMargin(<=0 off), unused; iparams2 = order, queryMode.
void oracle_contact_forward(const double* grid9, const int* shape3, const double* phi,
                            const double* n3, const double* X, const double* w, int64_t n,
                            const double* poses30, int nb, const double* params6,
                            const int* iparams2, double* out6, int* activeCount)
{
    DenseView d{};
    d.g = gridFromArrays(grid9, shape3);
    d.phi = phi;
    d.nrm = n3;
    ContactParams cp{params6[0], params6[1], params6[2], params6[3], params6[4], iparams2[0],
                     iparams2[1]};
    auto q = [&](V3 y) { return denseQuery(d, y, cp.mode, true); };
    contactForward(X, w, n, poses30, nb, cp, q, out6, activeCount);
}

void oracle_contact_backward(const double* grid9, const int* shape3, const double* phi,
                             const double* n3, const double* X, const double* w, int64_t n,
                             const double* poses30, int nb, const double* params6,
                             const int* iparams2, const double* gOut6, double* gTwist6,
                             double* gParams4, double* gX, double* gW)
{
    DenseView d{};
    d.g = gridFromArrays(grid9, shape3);
    d.phi = phi;
    d.nrm = n3;
    ContactParams cp{params6[0], params6[1], params6[2], params6[3], params6[4], iparams2[0],
                     iparams2[1]};
    auto q = [&](V3 y) { return denseQuery(d, y, cp.mode, true); };
    contactBackward(X, w, n, poses30, nb, cp, q, gOut6, gTwist6, gParams4, gX, gW);
}

// Linearisation cache of SDFContactEngine.evaluate (contacts_sdf.py:307-327):
// fill at one reference pose with the dense query of the forward ...
void oracle_contact_linearize(const double* grid9, const int* shape3, const double* phi,
                              const double* n3, const double* X, int64_t n, const double* pose30,
                              const double* params6, const int* iparams2, double* g0,
                              double* grad0, unsigned char* valid0)
{
    DenseView d{};
    d.g = gridFromArrays(grid9, shape3);
    d.phi = phi;
    d.nrm = n3;
    const ContactParams cp{params6[0], params6[1], params6[2], params6[3], params6[4],
                           iparams2[0], iparams2[1]};
    const Pose P = loadPose(pose30);
    for (int64_t i = 0; i < n; ++i)
    {
        const V3 Y = matvec(P.R, load3(X, i)) + v3(P.t[0], P.t[1], P.t[2]);
        QueryOut q{};
        q.valid = false;
        if (!(cp.margin > 0.0 && !(std::abs(Y.z) < cp.margin))) q = denseQuery(d, Y, cp.mode, true);
        valid0[i] = q.valid ? 1 : 0;
        g0[i] = q.valid ? q.gap : 0.0;
        grad0[3 * i] = q.valid ? q.grad.x : 0.0;
        grad0[3 * i + 1] = q.valid ? q.grad.y : 0.0;
        grad0[3 * i + 2] = q.valid ? q.grad.z : 0.0;
    }
}

// ... and the forward on the prediction g = g0 + grad0 . (t - t_ref) with
// the cached gradient (contacts_sdf.py:311-318), gated by the cached validity
// and the narrow filter at the current pose.
void oracle_contact_forward_linearized(const double* g0, const double* grad0,
                                       const unsigned char* valid0, const double* tref3,
                                       const double* X, const double* w, int64_t n,
                                       const double* poses30, int nb, const double* params6,
                                       const int* iparams2, double* out6, int* activeCount)
{
    const ContactParams cp{params6[0], params6[1], params6[2], params6[3], params6[4],
                           iparams2[0], iparams2[1]};
    for (int b = 0; b < nb; ++b)
    {
        const Pose P = loadPose(poses30 + 30 * int64_t(b));
        const V3 dt = v3(P.t[0] - tref3[0], P.t[1] - tref3[1], P.t[2] - tref3[2]);
        double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
        int act = 0;
        for (int64_t i = 0; i < n; ++i)
        {
            auto q = [&](V3) {
                QueryOut r{};
                r.valid = valid0[i] != 0;
                r.grad = load3(grad0, i);
                r.gap = g0[i] + dot(r.grad, dt);
                return r;
            };
            const PointTerms t = pointForward(P, load3(X, i), cp, q);
            if (!t.active) continue;
            ++act;
            const V3 wf = w[i] * t.fq;
            const V3 tq = w[i] * cross(t.off, t.fq);
            F[0] += wf.x; F[1] += wf.y; F[2] += wf.z;
            T[0] += tq.x; T[1] += tq.y; T[2] += tq.z;
        }
        for (int c = 0; c < 3; ++c)
        {
            out6[6 * b + c] = F[c];
            out6[6 * b + 3 + c] = T[c];
        }
        if (activeCount) activeCount[b] = act;
    }
}

// SDFContactEngine.evaluate in the world frame on a dense grid (see
// contactEvaluate).  params6 = kn, c, mu, reg, narrowMargin, unused;
// iparams2 = order, queryMode.
void oracle_contact_evaluate(const double* grid9, const int* shape3, const double* phi,
                             const double* n3, const double* X, const double* w, int64_t n,
                             const double* states36, const double* frames24, int nb,
                             const double* params6, const int* iparams2, const uint8_t* keep,
                             const uint8_t* pairOn, double* out12, int* activeCount)
{
    DenseView d{};
    d.g = gridFromArrays(grid9, shape3);
    d.phi = phi;
    d.nrm = n3;
    ContactParams cp{params6[0], params6[1], params6[2], params6[3], params6[4], iparams2[0],
                     iparams2[1]};
    auto q = [&](V3 y) { return denseQuery(d, y, cp.mode, false); };
    contactEvaluate(X, w, n, states36, frames24, nb, cp, keep, pairOn, q, out12, activeCount);
}

// Sparse Hessian stencil alone (SdfBuilder.cpp:1154-1225) on given bricks.
void oracle_sparse_hessian(const double* grid9, const int* shape3, int blockDim, int nblocks,
                           const int* bijk, const float* nrm, float* hes)
{
    estimateSparseHessian(gridFromArrays(grid9, shape3), blockDim, nblocks, bijk, nrm, hes);
}

// The contact-aware builds' per-voxel work (SdfBuilder.cpp:281-359 behind
// the tile index of :202-258) at selected voxels only, for checking large
// builds by sampling.  sparseBlockDim > 0 selects the sparse build's tile
// influence max(activation, tol + 2h) with activation = band + halo h
// (:575-640) and fills brickMask (one byte per brick of the block grid,
// x fastest) with its activation rule (any overlapping tile list non-empty);
// sparseBlockDim = 0 is the dense build (influence tol + 2h).
// Outputs per voxel: count, then stride entries of phi, normal, branch.
int oracle_contact_aware_probe(const double* verts, int64_t nv, const int* tris, int nt,
                               const double* cornerN, const double* grid9, const int* shape3,
                               int maxCandidates, double featureTol, double mergeDeg, int tileDim,
                               int useTiles, int sparseBlockDim, int halo, int bandVoxels,
                               double bandDistance, const int* ijk, int64_t nvox, uint8_t* ccount,
                               double* cphi, double* cn3, int* cbranch, uint8_t* brickMask)
{
    if (nv <= 0 || nt <= 0 || maxCandidates <= 0) return -1;
    const Grid g = gridFromArrays(grid9, shape3);
    const int S = std::max(0, std::min(maxCandidates, 8));
    const std::vector<CTri> ct = makeContactTris(verts, tris, nt, cornerN);
    Bvh tree;
    tree.build(verts, tris, nt);
    const double h = g.hmax();
    const double tol = featureTol > 0.0 ? featureTol : 1.5 * h;
    const double cosMerge = std::cos(mergeDeg * std::acos(-1.0) / 180.0);
    double infl = tol + 2.0 * h;
    if (sparseBlockDim > 0)
    {
        const double band = bandDistance > 0.0 ? bandDistance : static_cast<double>(bandVoxels) * h;
        infl = smax(band + static_cast<double>(halo) * h, infl);
    }
    std::vector<int> all(nt);
    std::iota(all.begin(), all.end(), 0);
    TileIndex ti;
    if (useTiles) ti = buildTileIndex(g, ct, tileDim, infl);
    if (brickMask && sparseBlockDim > 0 && useTiles && !ti.ids.empty())
    {
        const int bd = sparseBlockDim;
        const int nbx = (g.n[0] + bd - 1) / bd, nby = (g.n[1] + bd - 1) / bd,
                  nbz = (g.n[2] + bd - 1) / bd;
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t bl = 0; bl < int64_t(nbx) * nby * nbz; ++bl)
        {
            const int bi = int(bl % nbx), bj = int((bl / nbx) % nby), bk = int(bl / (int64_t(nbx) * nby));
            const int i0 = bi * bd, j0 = bj * bd, k0 = bk * bd;
            const int i1 = std::min(i0 + bd - 1, g.n[0] - 1);
            const int j1 = std::min(j0 + bd - 1, g.n[1] - 1);
            const int k1 = std::min(k0 + bd - 1, g.n[2] - 1);
            bool act = false;
            for (int tk = k0 / ti.tileDim; tk <= k1 / ti.tileDim && !act; ++tk)
                for (int tj = j0 / ti.tileDim; tj <= j1 / ti.tileDim && !act; ++tj)
                    for (int tii = i0 / ti.tileDim; tii <= i1 / ti.tileDim && !act; ++tii)
                        act = !ti.ids[size_t(ti.lin(tii, tj, tk))].empty();
            brickMask[bl] = act ? 1 : 0;
        }
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < nvox; ++q)
    {
        const int i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
        const std::vector<int>& cand = useTiles ? ti.forVoxel(i, j, k) : all;
        const std::vector<CCand> cs = candidatesForPoint(g.center(i, j, k), ct, cand, tree, tol,
                                                         cosMerge, S);
        const int cnt = std::min(int(cs.size()), S);
        ccount[q] = uint8_t(cnt);
        for (int e = 0; e < S; ++e)
        {
            const bool on = e < cnt;
            cphi[q * S + e] = on ? cs[e].phi : 0.0;
            store3(cn3, q * S + e, on ? cs[e].normal : v3(1, 0, 0));
            cbranch[q * S + e] = on ? cs[e].branch : -1;
        }
    }
    return S;
}

// Analytic-plane SDF contact (sdf_grid.py:4-22): phi = z - zTop, grad = e_z.
void oracle_contact_forward_plane(double zTop, const double* X, const double* w, int64_t n,
                                  const double* poses30, int nb, const double* params6,
                                  const int* iparams2, double* out6, int* activeCount)
{
    ContactParams cp{params6[0], params6[1], params6[2], params6[3], params6[4], iparams2[0],
                     iparams2[1]};
    auto q = [&](V3 y) {
        QueryOut r;
        r.gap = y.z - zTop;
        r.grad = v3(0, 0, 1);
        r.normal = v3(0, 0, 1);
        r.valid = true;
        return r;
    };
    contactForward(X, w, n, poses30, nb, cp, q, out6, activeCount);
}

} // extern "C"

extern "C" void oracle_sort_by_abs(const double* values, int n, int* order)
{
    // std::sort over candidates with |a.phi| < |b.phi| (SdfBuilder.cpp:348-353),
    // applied to indices: the same comparisons and swaps, hence the same order.
    std::vector<int> v(n);
    std::iota(v.begin(), v.end(), 0);
    std::sort(v.begin(), v.end(),
              [&](int a, int b) { return std::abs(values[a]) < std::abs(values[b]); });
    for (int i = 0; i < n; ++i) order[i] = v[i];
}
