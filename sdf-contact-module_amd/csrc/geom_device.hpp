// FP64 device geometry for the SDF build kernels.
//
// These routines restate the reference's nearest-triangle search and
// ray-parity sign (KdTree.h:26-74, 244-441) operation for operation so that,
// compiled without FMA contraction, a GPU build reproduces the CPU FP64
// results bit for bit (distances, witnesses, normals and the nearest
// triangle index, including exact ties, which the reference resolves by its
// near-first traversal order).  Reductions are evaluated ((x + y) + z).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "sdfo_internal.hpp"

namespace sdfo {
namespace dev {

struct D3
{
    double x, y, z;
};

__host__ __device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 add(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 scale(double s, D3 a) { return D3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ D3 divs(D3 a, double s) { return D3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b)
{
    return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double sqnorm(D3 a) { return dot(a, a); }

// KdTree.h:26-74 (Ericson, Real-Time Collision Detection 5.1.5).
__device__ inline D3 closestPoint(D3 p, D3 a, D3 b, D3 c)
{
    const D3 ab = sub(b, a);
    const D3 ac = sub(c, a);
    const D3 ap = sub(p, a);
    const double d1 = dot(ab, ap);
    const double d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    const D3 bp = sub(p, b);
    const double d3v = dot(ab, bp);
    const double d4 = dot(ac, bp);
    if (d3v >= 0.0 && d4 <= d3v) return b;
    const double vc = d1 * d4 - d3v * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3v <= 0.0)
    {
        const double v = d1 / (d1 - d3v);
        return add(a, scale(v, ab));
    }
    const D3 cp = sub(p, c);
    const double d5 = dot(ab, cp);
    const double d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0)
    {
        const double w = d2 / (d2 - d6);
        return add(a, scale(w, ac));
    }
    const double va = d3v * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3v) >= 0.0 && (d5 - d6) >= 0.0)
    {
        const double w = (d4 - d3v) / ((d4 - d3v) + (d5 - d6));
        return add(b, scale(w, sub(c, b)));
    }
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom;
    const double w = vc * denom;
    return add(add(a, D3{ab.x * v, ab.y * v, ab.z * v}), D3{ac.x * w, ac.y * w, ac.z * w});
}

// One axis of the point-box distance of KdTree.h:244-264, its two branches
// taken as selects: inside [lo, hi] the gap is 0 and sq + 0 * 0 is sq, so the
// sum is the branchy one bit for bit, without divergent control flow.
__device__ __forceinline__ double addAxisGap(double sq, double p, double lo, double hi)
{
    const double below = lo - p, above = p - hi;
    const double d = p < lo ? below : (p > hi ? above : 0.0);
    return sq + d * d;
}

// KdTree.h:244-264.
__device__ __forceinline__ double boxSq(D3 p, const BvhNode& n)
{
    double sq = addAxisGap(0.0, p.x, n.bmin[0], n.bmax[0]);
    sq = addAxisGap(sq, p.y, n.bmin[1], n.bmax[1]);
    return addAxisGap(sq, p.z, n.bmin[2], n.bmax[2]);
}

// KdTree.h:335-360 (slab test; std::max / std::min argument order).  The
// swap and the early exit are selects: once an axis empties the interval the
// result is false whatever the later axes give, as with the early return.
__device__ __forceinline__ bool rayBox(D3 o, D3 d, const BvhNode& n)
{
    double tmin = 0.0;
    double tmax = DBL_MAX;
    bool hit = true;
    const double oc[3] = {o.x, o.y, o.z};
    const double dc[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const double invD = 1.0 / dc[a];
        const double t0 = (n.bmin[a] - oc[a]) * invD;
        const double t1 = (n.bmax[a] - oc[a]) * invD;
        const bool swap = t0 > t1;
        const double lo = swap ? t1 : t0, hi = swap ? t0 : t1;
        tmin = (tmin < lo) ? lo : tmin;
        tmax = (hi < tmax) ? hi : tmax;
        hit = hit && !(tmax < tmin);
    }
    return hit;
}

// KdTree.h:362-396 (Moller-Trumbore 1997, eps 1e-12).  The early exits are
// folded into one condition: every quantity is computed from the same
// operands as in the branchy form, and a test that would have returned false
// early makes the conjunction false whatever the later ones give.
__device__ __forceinline__ bool rayTri(D3 o, D3 d, D3 a, D3 b, D3 c)
{
    const double eps = 1.0e-12;
    const D3 e1 = sub(b, a);
    const D3 e2 = sub(c, a);
    const D3 h = cross(d, e2);
    const double det = dot(e1, h);
    const double invDet = 1.0 / det;
    const D3 s = sub(o, a);
    const double u = invDet * dot(s, h);
    const D3 q = cross(s, e1);
    const double v = invDet * dot(d, q);
    const double t = invDet * dot(e2, q);
    return !(fabs(det) < eps) && !(u < -eps || u > 1.0 + eps) && !(v < -eps || u + v > 1.0 + eps) &&
           t > eps;
}

struct Tri
{
    D3 a, b, c;
};

__device__ __forceinline__ Tri loadTri(const double* tri9, int slot)
{
    const double* t = tri9 + 9 * size_t(slot);
    return Tri{d3(t[0], t[1], t[2]), d3(t[3], t[4], t[5]), d3(t[6], t[7], t[8])};
}

struct SignedNearest
{
    double sd;
    D3 cp;
    D3 face;
    int slot;
};

// Per-lane LDS stack: entry e of lane l lives at stack[e * stride + l].
struct LaneStack
{
    int* base;
    int stride;
    __device__ __forceinline__ void set(int e, int v) { base[e * stride] = v; }
    __device__ __forceinline__ int get(int e) const { return base[e * stride]; }
};

// KdTree::searchNearest (KdTree.h:266-328): near-first traversal of one lane,
// strict < in leaves, so the first minimum in traversal order wins.
__device__ inline void refNearest(D3 p, const BvhNode* __restrict__ nodes,
                                  const double* __restrict__ tri9, LaneStack st, double& bestSq,
                                  int& bestSlot, int root = 0)
{
    bestSq = DBL_MAX;
    bestSlot = -1;
    int sp = 0;
    st.set(sp++, root);
    while (sp > 0)
    {
        const int ni = st.get(--sp);
        const BvhNode nd = nodes[ni];
        // Entry test of searchNearest (KdTree.h:275-279); for children it is
        // also the parent's "<= bestSq" test with the same bestSq.
        if (!(boxSq(p, nd) <= bestSq)) continue;
        if (nd.triCount > 0)
        {
            for (int s = nd.triStart; s < nd.triStart + nd.triCount; ++s)
            {
                const Tri t = loadTri(tri9, s);
                const D3 cp = closestPoint(p, t.a, t.b, t.c);
                const double sq = sqnorm(sub(p, cp));
                if (sq < bestSq)
                {
                    bestSq = sq;
                    bestSlot = s;
                }
            }
            continue;
        }
        const double dl = boxSq(p, nodes[nd.left]);
        const double dr = boxSq(p, nodes[nd.right]);
        int first = nd.left, second = nd.right;
        if (dr < dl)
        {
            first = nd.right;
            second = nd.left;
        }
        st.set(sp++, second);
        st.set(sp++, first);
    }
}

// KdTree::countRayIntersections (KdTree.h:398-441) for one lane: every box
// the ray crosses is visited.
__device__ inline int refRayHits(D3 p, D3 dir, const BvhNode* __restrict__ nodes,
                                 const double* __restrict__ tri9, LaneStack st, int root = 0)
{
    int hits = 0;
    int sp = 0;
    st.set(sp++, root);
    while (sp > 0)
    {
        const int ni = st.get(--sp);
        const BvhNode nd = nodes[ni];
        if (!rayBox(p, dir, nd)) continue;
        if (nd.triCount > 0)
        {
            for (int s = nd.triStart; s < nd.triStart + nd.triCount; ++s)
            {
                const Tri tt = loadTri(tri9, s);
                if (rayTri(p, dir, tt.a, tt.b, tt.c)) ++hits;
            }
            continue;
        }
        st.set(sp++, nd.right);
        st.set(sp++, nd.left);
    }
    return hits;
}

// KdTree::findSignedNearest (KdTree.h:143-165) from a nearest slot, its
// squared distance and the ray hit count.
__device__ inline SignedNearest finishSigned(D3 p, const double* __restrict__ tri9, double bestSq,
                                             int bestSlot, int hits)
{
    SignedNearest r;
    r.slot = bestSlot;
    double dist;
    if (bestSlot >= 0)
    {
        const Tri t = loadTri(tri9, bestSlot);
        r.cp = closestPoint(p, t.a, t.b, t.c);
        dist = sqrt(bestSq);
        const D3 n = cross(sub(t.b, t.a), sub(t.c, t.a));
        const double nLen = sqrt(sqnorm(n));
        r.face = nLen > 1.0e-15 ? divs(n, nLen) : d3(0.0, 0.0, 1.0);
    }
    else
    {
        // No triangle compared (non-finite point): ClosestPointResult defaults
        // with distance = DBL_MAX (KdTree.h:18-24, 125-127).
        r.cp = d3(0.0, 0.0, 0.0);
        r.face = d3(0.0, 0.0, 1.0);
        dist = DBL_MAX;
    }
    if (hits > 0)
        r.sd = (hits % 2) == 1 ? -dist : dist;
    else
        r.sd = dot(sub(p, r.cp), r.face) >= 0.0 ? dist : -dist;
    return r;
}

// Per-lane form of findSignedNearest.
__device__ inline SignedNearest signedNearest(D3 p, const BvhNode* __restrict__ nodes,
                                              const double* __restrict__ tri9, D3 dir,
                                              LaneStack st, int root = 0)
{
    double bestSq;
    int bestSlot;
    refNearest(p, nodes, tri9, st, bestSq, bestSlot, root);
    const int hits = refRayHits(p, dir, nodes, tri9, st, root);
    return finishSigned(p, tri9, bestSq, bestSlot, hits);
}

// ---------------------------------------------------------------------------
// Wave-cooperative traversal.  The 64 lanes of a wave (nearby voxels) walk
// ONE shared stack: a node is visited when any lane needs it, node and
// triangle records are wave-uniform (scalar loads), and no lane diverges.
// Results are the per-lane traversal's bit for bit:
//
// * Nearest.  The per-lane search returns, among the triangles of minimal
//   squared distance m, the first in its depth-first order, provided it
//   reaches that triangle t*; at each node that order puts the child with the
//   strictly smaller box distance first (KdTree.h:311-326).  Each lane here
//   knows the path key of every node it is shown (one bit per level: 0 = the
//   child this lane visits first), so its depth-first rank is known wherever
//   the wave goes, and it keeps the lexicographic minimum of (sq, rank, slot).
//   The wave prunes a node only when for every lane its box distance exceeds
//   (sqrt(best) + mu)^2, mu = 2^-30 * (largest coordinate magnitude), far above
//   the rounding of the box and triangle distances, so every triangle with
//   sq <= m is examined.  The per-lane search reaches t* whenever the box
//   distance of t*'s leaf is <= m (its ancestors' boxes contain the leaf box,
//   so their rounded distances are no larger, and its best before t* is > m).
//   A lane for which that check fails is reported (exact = false) and must be
//   re-run with the per-lane search.
//   Path keys need no per-lane storage: a stack entry carries the wave mask
//   of lanes whose bit for it is 1, and the entry's parent key is a prefix of
//   the key of the node popped just before it (that node is the parent or one
//   of its descendants).
// * Ray parity.  The hit count does not depend on order; the wave pushes
//   with each node the mask of lanes whose ray crossed every box on the path,
//   which is the set of nodes the per-lane walk visits.
//
// Must be called by all 64 lanes of the wave in uniform control flow;
// `active` lanes are the ones whose results are used (inactive lanes need a
// finite point).
// ---------------------------------------------------------------------------
constexpr int kWave = 64;

struct WaveStacks
{
    BvhNode* rec;             // kStackDepth node records (nearest walk)
    double* leaf;             // staged leaf: 8 triangle boxes (6 doubles) then 8 triangles (9)
    unsigned long long* mask; // kStackDepth: path-bit mask (nearest) / lane mask (ray)
    int* node;                // kStackDepth: depth (nearest) / node index (ray)
};

constexpr int kLeafMax = 8; // KdTree.h:186
constexpr int kLeafDoubles = kLeafMax * (6 + 9);

// LDS bytes per wave for WaveStacks (a multiple of 16).
constexpr int kWaveStackBytes = kStackDepth * (64 + 8 + 4) + kLeafDoubles * 8;

// Carves one wave's stacks out of 16-byte aligned LDS.
__device__ __forceinline__ WaveStacks carveWaveStacks(unsigned char* base)
{
    WaveStacks ws;
    ws.rec = reinterpret_cast<BvhNode*>(base);
    ws.leaf = reinterpret_cast<double*>(base + kStackDepth * 64);
    ws.mask = reinterpret_cast<unsigned long long*>(base + kStackDepth * 64 + kLeafDoubles * 8);
    ws.node = reinterpret_cast<int*>(base + kStackDepth * 72 + kLeafDoubles * 8);
    return ws;
}

// Brick-pool index (brick-major, x fastest inside a brick of edge bd) of
// lane `lane` of wave unit `u`.  When bd is a multiple of 4 a unit is a
// 4x4x4 sub-brick, whose 64 lanes share more of their nearest-search work
// than the 8x8x1 slab of 64 consecutive voxels; other edges keep 64
// consecutive voxels per unit.
__device__ __forceinline__ int64_t brickUnitVoxel(int bd, int64_t u, int lane)
{
    if ((bd & 3) != 0) return u * 64 + lane;
    const int q = bd >> 2, subs = q * q * q;
    const int64_t b = u / subs;
    const int s = int(u - b * subs);
    const int x = 4 * (s % q) + (lane & 3);
    const int y = 4 * ((s / q) % q) + ((lane >> 2) & 3);
    const int z = 4 * (s / (q * q)) + (lane >> 4);
    return b * int64_t(bd) * bd * bd + x + bd * (y + bd * z);
}

__device__ __forceinline__ int uniformInt(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ unsigned long long uniformMask(unsigned long long v)
{
    const unsigned lo = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(v))));
    const unsigned hi = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(v >> 32))));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

struct WaveNearest
{
    double bestSq;
    int bestSlot;
    bool exact;   // false: the per-lane search must be re-run for this lane
    bool aborted; // wave-uniform: the step budget ran out (bestSq is an upper bound)
};

// Distance^2 from p to a box given as bmin3, bmax3 (same arithmetic as boxSq).
__device__ __forceinline__ double boxSq6(D3 p, const double* b)
{
    double sq = addAxisGap(0.0, p.x, b[0], b[3]);
    sq = addAxisGap(sq, p.y, b[1], b[4]);
    return addAxisGap(sq, p.z, b[2], b[5]);
}

// The stack holds whole node records (copied from the parent's child loads),
// so a level costs one dependent global load (the children), issued as soon
// as the record is popped.  Leaves are staged into LDS by one vector load.
__device__ inline WaveNearest waveNearest(D3 p, bool active, unsigned long long act,
                                          const BvhNode* __restrict__ nodes,
                                          const double* __restrict__ tri9,
                                          const double* __restrict__ box6, double mu,
                                          WaveStacks ws, int lane, int root, int budget = 0)
{
    double bestSq = DBL_MAX, thr = DBL_MAX, bestB = 0.0;
    int steps = 0; // node visits + triangle tests (wave-uniform)
    uint32_t bestRank = 0xffffffffu;
    int bestSlot = -1;
    uint32_t curKey = 0u; // key of the last node popped
    int curDepth = 0;
    const int nAct = __popcll(act);
    ws.rec[0] = nodes[root];
    ws.node[0] = 0;
    ws.mask[0] = 0ull;
    int sp = 1;
    while (sp > 0)
    {
        --sp;
        if (budget > 0 && ++steps > budget) return WaveNearest{bestSq, bestSlot, false, true};
        const BvhNode nd = ws.rec[sp];
        const int depth = uniformInt(ws.node[sp]);
        const unsigned long long bits = uniformMask(ws.mask[sp]);
        const int left = uniformInt(nd.left), right = uniformInt(nd.right);
        const int triStart = uniformInt(nd.triStart), triCount = uniformInt(nd.triCount);
        // children requested before the box test (a leaf reloads itself: harmless)
        const BvhNode L = nodes[triCount == 0 ? left : 0];
        const BvhNode R = nodes[triCount == 0 ? right : 0];
        uint32_t key = 0u;
        if (depth > 0)
            key = ((curKey >> (curDepth - depth + 1)) << 1) | uint32_t((bits >> lane) & 1ull);
        curKey = key;
        curDepth = depth;
        const double bs = boxSq(p, nd);
        if (__ballot(active && bs <= thr) == 0ull) continue;
        if (triCount > 0)
        {
            // stage the leaf's triangle boxes and vertices (<= 120 doubles)
            const int nb = 6 * triCount, nt = 9 * triCount;
            const double* gb = box6 + 6 * size_t(triStart);
            const double* gt = tri9 + 9 * size_t(triStart);
            if (lane < nb) ws.leaf[lane] = gb[lane];
            if (lane < nt) ws.leaf[6 * kLeafMax + lane] = gt[lane];
            if (lane + kWave < nt) ws.leaf[6 * kLeafMax + kWave + lane] = gt[lane + kWave];
            // A triangle whose own box is beyond every lane's threshold can
            // neither win nor tie for any lane (same margin argument).
            unsigned need = 0u;
            for (int q = 0; q < triCount; ++q)
                if (__ballot(active && boxSq6(p, ws.leaf + 6 * q) <= thr) != 0ull) need |= 1u << q;
            const uint32_t rank = key << (31 - depth);
            while (need != 0u)
            {
                const int q = __builtin_ctz(need);
                need &= need - 1u;
                ++steps;
                const double* t = ws.leaf + 6 * kLeafMax + 9 * q;
                const D3 ta = d3(t[0], t[1], t[2]), tb = d3(t[3], t[4], t[5]), tc = d3(t[6], t[7], t[8]);
                const double sq = sqnorm(sub(p, closestPoint(p, ta, tb, tc)));
                const int s = triStart + q;
                const bool better =
                    sq < bestSq ||
                    (sq == bestSq && (rank < bestRank || (rank == bestRank && s < bestSlot)));
                if (better)
                {
                    if (sq < bestSq)
                    {
                        bestSq = sq;
                        const double r = sqrt(sq) + mu;
                        thr = r * r;
                    }
                    bestRank = rank;
                    bestSlot = s;
                    bestB = bs;
                }
            }
            continue;
        }
        const double dl = boxSq(p, L);
        const double dr = boxSq(p, R);
        const bool firstRight = dr < dl; // this lane's per-lane order
        const unsigned long long fr = __ballot(firstRight);
        const bool needL = __ballot(active && dl <= thr) != 0ull;
        const bool needR = __ballot(active && dr <= thr) != 0ull;
        const bool rightFirst = 2 * __popcll(fr & act) > nAct;
        const int childDepth = depth + 1;
        // The child visited second by the wave is pushed first.  Path bit 1 =
        // the child this lane visits second: left has bit 1 where firstRight.
        if (rightFirst)
        {
            if (needL)
            {
                ws.rec[sp] = L;
                ws.node[sp] = childDepth;
                ws.mask[sp] = fr;
                ++sp;
            }
            if (needR)
            {
                ws.rec[sp] = R;
                ws.node[sp] = childDepth;
                ws.mask[sp] = ~fr;
                ++sp;
            }
        }
        else
        {
            if (needR)
            {
                ws.rec[sp] = R;
                ws.node[sp] = childDepth;
                ws.mask[sp] = ~fr;
                ++sp;
            }
            if (needL)
            {
                ws.rec[sp] = L;
                ws.node[sp] = childDepth;
                ws.mask[sp] = fr;
                ++sp;
            }
        }
    }
    return WaveNearest{bestSq, bestSlot, bestSlot >= 0 && bestB <= bestSq, false};
}

__device__ inline int waveRayHits(D3 p, D3 dir, unsigned long long act,
                                  const BvhNode* __restrict__ nodes,
                                  const double* __restrict__ tri9, WaveStacks ws, int lane,
                                  int root)
{
    const unsigned long long me = 1ull << lane;
    int hits = 0;
    ws.node[0] = root;
    ws.mask[0] = act;
    int sp = 1;
    while (sp > 0)
    {
        --sp;
        const int ni = uniformInt(ws.node[sp]);
        const unsigned long long m = uniformMask(ws.mask[sp]);
        const BvhNode nd = nodes[ni];
        const bool in = (m & me) != 0ull && rayBox(p, dir, nd);
        const unsigned long long m2 = __ballot(in);
        if (m2 == 0ull) continue;
        if (nd.triCount > 0)
        {
            for (int s = nd.triStart; s < nd.triStart + nd.triCount; ++s)
            {
                const Tri tt = loadTri(tri9, s);
                if (rayTri(p, dir, tt.a, tt.b, tt.c) && in) ++hits;
            }
            continue;
        }
        ws.node[sp] = nd.right;
        ws.mask[sp] = m2;
        ++sp;
        ws.node[sp] = nd.left;
        ws.mask[sp] = m2;
        ++sp;
    }
    return hits;
}

struct WaveSigned
{
    SignedNearest s;
    bool exact; // false: recompute this lane with signedNearest
};

// findSignedNearest for the 64 lanes of a wave (see above).
__device__ inline WaveSigned waveSignedNearest(D3 p, bool active, const BvhNode* __restrict__ nodes,
                                               const double* __restrict__ tri9,
                                               const double* __restrict__ box6, D3 dir, double mu,
                                               WaveStacks ws, int root = 0)
{
    const int lane = int(threadIdx.x) & (kWave - 1);
    const unsigned long long act = __ballot(active);
    const WaveNearest wn = waveNearest(p, active, act, nodes, tri9, box6, mu, ws, lane, root);
    const int hits = waveRayHits(p, dir, act, nodes, tri9, ws, lane, root);
    return WaveSigned{finishSigned(p, tri9, wn.bestSq, wn.bestSlot, hits), wn.exact};
}

// ---------------------------------------------------------------------------
// Frontier search: one voxel per wave, the lanes expanding the BVH level by
// level in parallel (64 nodes per step), for voxels whose wave search ran
// over its step budget (points equidistant from many triangles).  The
// exactness argument is the wave search's: ranks are the voxel's own
// depth-first path keys, pruning keeps every triangle with sq <= m (margin
// mu), the winner is the lexicographic minimum of (sq, rank, slot) over the
// wave, and the per-lane search reaches it when its leaf box distance is
// <= m.  A full frontier or a failed check reports exact = false.
// ---------------------------------------------------------------------------
struct FrontierLds
{
    int* node[2];     // cap entries each: node << 5 | depth
    uint32_t* key[2]; // cap entries each
    int* counter;     // [0] next-level count, [1] overflow flag
    int cap;
};

// LDS bytes of a frontier of `cap` entries per level.
__host__ __device__ constexpr int frontierBytes(int cap) { return 2 * cap * (4 + 4) + 16; }

__device__ __forceinline__ FrontierLds carveFrontier(unsigned char* base, int cap)
{
    FrontierLds f;
    for (int b = 0; b < 2; ++b)
    {
        f.node[b] = reinterpret_cast<int*>(base + b * cap * 8);
        f.key[b] = reinterpret_cast<uint32_t*>(base + b * cap * 8 + cap * 4);
    }
    f.counter = reinterpret_cast<int*>(base + 2 * cap * 8);
    f.cap = cap;
    return f;
}

// Outcome of a frontier search.
enum FrontierStatus
{
    kFrontierExact = 0,    // result certified
    kFrontierUnsure = 1,   // re-run with the per-lane search
    kFrontierOverflow = 2, // a level outgrew the frontier: retry with a larger one
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

// Lexicographic (sq, rank, slot) minimum over the wave; every lane gets it.
__device__ __forceinline__ void waveLexMin(double& sq, uint32_t& rank, int& slot, double& b)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
        const double s2 = __shfl_xor(sq, o);
        const uint32_t r2 = uint32_t(__shfl_xor(int(rank), o));
        const int t2 = __shfl_xor(slot, o);
        const double b2 = __shfl_xor(b, o);
        const bool take = s2 < sq || (s2 == sq && (r2 < rank || (r2 == rank && unsigned(t2) < unsigned(slot))));
        if (take)
        {
            sq = s2;
            rank = r2;
            slot = t2;
            b = b2;
        }
    }
}

__device__ inline SignedNearest frontierSignedNearest(D3 p, const BvhNode* __restrict__ nodes,
                                                      const double* __restrict__ tri9,
                                                      const double* __restrict__ box6, D3 dir,
                                                      double mu, int root, double& seed,
                                                      FrontierLds f, int& status, int hits)
{
    const int lane = int(threadIdx.x) & (kWave - 1);
    double bestSq = DBL_MAX, bestB = 0.0;
    uint32_t bestRank = 0xffffffffu;
    int bestSlot = -1;
    double m = seed; // wave-uniform bound: a triangle at this squared distance exists
    double thr = DBL_MAX;
    if (m < DBL_MAX)
    {
        const double r = sqrt(m) + mu;
        thr = r * r;
    }
    if (lane == 0)
    {
        f.node[0][0] = root << 5;
        f.key[0][0] = 0u;
        f.counter[1] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int cur = 0, n = 1;
    bool overflow = false;
    while (n > 0 && !overflow)
    {
        if (lane == 0) f.counter[0] = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (int base = 0; base < n; base += kWave)
        {
            const int idx = base + lane;
            if (idx < n)
            {
                const int e = f.node[cur][idx];
                const int ni = e >> 5, depth = e & 31;
                const uint32_t key = f.key[cur][idx];
                const BvhNode nd = nodes[ni];
                const double bs = boxSq(p, nd);
                if (bs <= thr)
                {
                    if (nd.triCount > 0)
                    {
                        const uint32_t rank = key << (31 - depth);
                        for (int s = nd.triStart; s < nd.triStart + nd.triCount; ++s)
                        {
                            if (!(boxSq6(p, box6 + 6 * size_t(s)) <= thr)) continue;
                            const Tri t = loadTri(tri9, s);
                            const double sq = sqnorm(sub(p, closestPoint(p, t.a, t.b, t.c)));
                            if (sq < bestSq || (sq == bestSq && (rank < bestRank || (rank == bestRank && s < bestSlot))))
                            {
                                bestSq = sq;
                                bestRank = rank;
                                bestSlot = s;
                                bestB = bs;
                            }
                        }
                    }
                    else
                    {
                        const double dl = boxSq(p, nodes[nd.left]);
                        const double dr = boxSq(p, nodes[nd.right]);
                        const uint32_t fr = dr < dl ? 1u : 0u; // this voxel's per-lane order
                        if (dl <= thr)
                        {
                            const int q = atomicAdd(&f.counter[0], 1);
                            if (q < f.cap)
                            {
                                f.node[cur ^ 1][q] = (nd.left << 5) | (depth + 1);
                                f.key[cur ^ 1][q] = (key << 1) | fr;
                            }
                            else
                                f.counter[1] = 1;
                        }
                        if (dr <= thr)
                        {
                            const int q = atomicAdd(&f.counter[0], 1);
                            if (q < f.cap)
                            {
                                f.node[cur ^ 1][q] = (nd.right << 5) | (depth + 1);
                                f.key[cur ^ 1][q] = (key << 1) | (fr ^ 1u);
                            }
                            else
                                f.counter[1] = 1;
                        }
                    }
                }
            }
            const double wm = waveMinD(bestSq);
            if (wm < m)
            {
                m = wm;
                const double r = sqrt(m) + mu;
                thr = r * r;
            }
        }
        // the next level's entries were written by other lanes
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        overflow = uniformInt(f.counter[1]) != 0;
        n = uniformInt(f.counter[0]);
        n = n < f.cap ? n : f.cap;
        cur ^= 1;
    }
    waveLexMin(bestSq, bestRank, bestSlot, bestB);
    seed = m; // the bound reached (for a retry)
    status = overflow ? kFrontierOverflow
                      : ((bestSlot >= 0 && bestB <= bestSq) ? kFrontierExact : kFrontierUnsure);
    if (overflow) return SignedNearest{};
    // ray parity: the count the caller has (hits >= 0), else counted by lane 0
    // with a per-lane stack in the frontier buffer
    if (hits < 0)
    {
        hits = 0;
        if (lane == 0) hits = refRayHits(p, dir, nodes, tri9, LaneStack{f.node[0], 1}, root);
        hits = __shfl(hits, 0);
    }
    return finishSigned(p, tri9, bestSq, bestSlot, hits);
}

// SdfBuilder.cpp:1070-1085: sign * (c - w) / |c - w|, or sign * face normal.
__device__ __forceinline__ D3 voxelGradient(D3 c, const SignedNearest& s)
{
    const D3 diff = sub(c, s.cp);
    const double dist = sqrt(sqnorm(diff));
    const double sign = (s.sd >= 0.0) ? 1.0 : -1.0;
    if (dist > 1e-12) return divs(scale(sign, diff), dist);
    return scale(sign, s.face);
}

} // namespace dev
} // namespace sdfo
