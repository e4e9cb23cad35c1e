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

// KdTree.h:244-264.
__device__ __forceinline__ double boxSq(D3 p, const BvhNode& n)
{
    double sq = 0.0;
    const double pc[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        if (pc[a] < n.bmin[a])
        {
            const double d = n.bmin[a] - pc[a];
            sq += d * d;
        }
        else if (pc[a] > n.bmax[a])
        {
            const double d = pc[a] - n.bmax[a];
            sq += d * d;
        }
    }
    return sq;
}

// KdTree.h:335-360 (slab test; std::max / std::min argument order).
__device__ __forceinline__ bool rayBox(D3 o, D3 d, const BvhNode& n)
{
    double tmin = 0.0;
    double tmax = DBL_MAX;
    const double oc[3] = {o.x, o.y, o.z};
    const double dc[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const double invD = 1.0 / dc[a];
        double t0 = (n.bmin[a] - oc[a]) * invD;
        double t1 = (n.bmax[a] - oc[a]) * invD;
        if (t0 > t1)
        {
            const double tt = t0;
            t0 = t1;
            t1 = tt;
        }
        tmin = (tmin < t0) ? t0 : tmin;
        tmax = (t1 < tmax) ? t1 : tmax;
        if (tmax < tmin) return false;
    }
    return true;
}

// KdTree.h:362-396 (Moller-Trumbore 1997, eps 1e-12).
__device__ __forceinline__ bool rayTri(D3 o, D3 d, D3 a, D3 b, D3 c)
{
    const double eps = 1.0e-12;
    const D3 e1 = sub(b, a);
    const D3 e2 = sub(c, a);
    const D3 h = cross(d, e2);
    const double det = dot(e1, h);
    if (fabs(det) < eps) return false;
    const double invDet = 1.0 / det;
    const D3 s = sub(o, a);
    const double u = invDet * dot(s, h);
    if (u < -eps || u > 1.0 + eps) return false;
    const D3 q = cross(s, e1);
    const double v = invDet * dot(d, q);
    if (v < -eps || u + v > 1.0 + eps) return false;
    const double t = invDet * dot(e2, q);
    return t > eps;
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

// KdTree::findSignedNearest (KdTree.h:143-165): near-first nearest search
// (KdTree.h:266-328; strict < in leaves, so the first minimum in traversal
// order wins) and the ray-parity count (KdTree.h:398-441; every box the ray
// crosses is visited).
__device__ inline SignedNearest signedNearest(D3 p, const BvhNode* __restrict__ nodes,
                                              const double* __restrict__ tri9, D3 dir,
                                              LaneStack st)
{
    double bestSq = DBL_MAX;
    int bestSlot = -1;
    int sp = 0;
    st.set(sp++, 0);
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

    int hits = 0;
    sp = 0;
    st.set(sp++, 0);
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
    if (hits > 0)
        r.sd = (hits % 2) == 1 ? -dist : dist;
    else
        r.sd = dot(sub(p, r.cp), r.face) >= 0.0 ? dist : -dist;
    return r;
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
