// Ordering of merged contact-aware candidates by |phi| exactly as the CPU
// path's std::sort orders them (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace sdfo {

// Explicit-stack size of sortPerm (its stack never holds more than
// 2 floor(log2 n) + 1 ranges) and the largest n of the debug sort entry point.
constexpr int kMergeCap = 64;

// Array view with a stride between elements (lists interleaved across lanes).
template <class T>
struct Strided
{
    T* p;
    int64_t s;
    __host__ __device__ __forceinline__ T& operator[](int i) const { return p[int64_t(i) * s]; }
    __host__ __device__ __forceinline__ Strided operator+(int k) const { return Strided{p + int64_t(k) * s, s}; }
};

// --- std::sort(merged, |a.phi| < |b.phi|) as GNU libstdc++ performs it, on a
// permutation of the merged entries (SdfBuilder.cpp:348-353).
template <class Phi>
struct AbsPhiOrderT
{
    Phi phi;
    __host__ __device__ __forceinline__ bool operator()(int a, int b) const
    {
        return std::fabs(phi[a]) < std::fabs(phi[b]);
    }
};
using AbsPhiOrder = AbsPhiOrderT<const double*>;

template <class V, class Lt>
__host__ __device__ inline void pushHeap(V v, int hole, int top, int value, Lt lt)
{
    int parent = (hole - 1) / 2;
    while (hole > top && lt(v[parent], value))
    {
        v[hole] = v[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    v[hole] = value;
}

template <class V, class Lt>
__host__ __device__ inline void adjustHeap(V v, int hole, int len, int value, Lt lt)
{
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2)
    {
        child = 2 * (child + 1);
        if (lt(v[child], v[child - 1])) child--;
        v[hole] = v[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2)
    {
        child = 2 * (child + 1);
        v[hole] = v[child - 1];
        hole = child - 1;
    }
    pushHeap(v, hole, top, value, lt);
}

template <class V, class Lt>
__host__ __device__ inline void heapSortRange(V v, int n, Lt lt)
{
    if (n >= 2)
    {
        for (int parent = (n - 2) / 2;; --parent)
        {
            adjustHeap(v, parent, n, v[parent], lt);
            if (parent == 0) break;
        }
    }
    for (int last = n; last > 1;)
    {
        --last;
        const int value = v[last];
        v[last] = v[0];
        adjustHeap(v, 0, last, value, lt);
    }
}

template <class V>
__host__ __device__ inline void swapi(V v, int a, int b)
{
    const int t = v[a];
    v[a] = v[b];
    v[b] = t;
}

template <class V, class Lt>
__host__ __device__ inline void sortPerm(V v, int n, Lt lt)
{
    if (n <= 1) return;
    // introsort loop; sub-ranges are disjoint, so an explicit stack replaces
    // the recursion on the right part without changing any result.
    int stF[kMergeCap], stL[kMergeCap], stD[kMergeCap];
    int sp = 0;
    int lg = 0;
    while ((2 << lg) <= n) ++lg; // floor(log2(n))
    stF[sp] = 0;
    stL[sp] = n;
    stD[sp] = 2 * lg;
    ++sp;
    while (sp > 0)
    {
        --sp;
        const int first = stF[sp];
        int last = stL[sp];
        int depth = stD[sp];
        while (last - first > 16)
        {
            if (depth == 0)
            {
                heapSortRange(v + first, last - first, lt);
                break;
            }
            --depth;
            // median of (first+1, mid, last-1) moved to first
            const int a = first + 1, b = first + (last - first) / 2, c = last - 1;
            if (lt(v[a], v[b]))
            {
                if (lt(v[b], v[c])) swapi(v, first, b);
                else if (lt(v[a], v[c])) swapi(v, first, c);
                else swapi(v, first, a);
            }
            else if (lt(v[a], v[c])) swapi(v, first, a);
            else if (lt(v[b], v[c])) swapi(v, first, c);
            else swapi(v, first, b);
            // unguarded partition of [first+1, last) around v[first]
            int lo = first + 1, hi = last;
            while (true)
            {
                while (lt(v[lo], v[first])) ++lo;
                --hi;
                while (lt(v[first], v[hi])) --hi;
                if (!(lo < hi)) break;
                swapi(v, lo, hi);
                ++lo;
            }
            stF[sp] = lo;
            stL[sp] = last;
            stD[sp] = depth;
            ++sp;
            last = lo;
        }
    }
    // final insertion sort: guarded on the first 16, unguarded after
    const int head = n > 16 ? 16 : n;
    for (int i = 1; i < head; ++i)
    {
        const int val = v[i];
        if (lt(val, v[0]))
        {
            for (int k = i; k > 0; --k) v[k] = v[k - 1];
            v[0] = val;
        }
        else
        {
            int k = i;
            while (lt(val, v[k - 1]))
            {
                v[k] = v[k - 1];
                --k;
            }
            v[k] = val;
        }
    }
    for (int i = head; i < n; ++i)
    {
        const int val = v[i];
        int k = i;
        while (lt(val, v[k - 1]))
        {
            v[k] = v[k - 1];
            --k;
        }
        v[k] = val;
    }
}

} // namespace sdfo
