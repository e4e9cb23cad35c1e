// Field access and interpolation for the query and contact kernels.
//
// A point is located in FP64 exactly as the reference does it
// (f = (p - bmin)/voxel - 0.5, floor / std::round; SdfOracle.cpp:530-549,
// GridSpec.h:37-45), so the valid mask and the chosen cell/voxel are exact;
// the interpolation itself runs in T (float for the query stream, double for
// the contact reduction) on the FP32 field mirror.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include <cfloat>
#include <cstdint>

namespace sdfo {
namespace dev {

// Everything a query kernel needs, passed by value.
struct QGrid
{
    double bmin[3];
    double voxel[3];
    int n[3];
    int sparse;
    int hasHessian;
    int stride;
    // dense
    const float4* pn;
    // dense, optional: the 8 corner records of each cell packed contiguously
    // (cell (i,j,k) -> cellRec[8 * (i + (nx-1) * (j + (ny-1) * k)) + di + 2 dj + 4 dk]),
    // one 128-byte line per interpolation instead of four to eight.
    const float4* cellRec;
    const float* h6f;
    const uint8_t* ccount;
    const float4* cpn;
    const int32_t* cbranch;
    // sparse
    int bd, nbx, nby, nbz;
    const int32_t* table;
    const float4* spn;
    const float* sh6;
    const uint8_t* sccount;
    const float4* scpn;
    const int32_t* scbranch;
    // sparse, optional: cell records at each brick-pool voxel and their
    // validity bits (maskWords 64-bit words per brick)
    const float4* scellRec;
    const unsigned long long* scellMask;
    int maskWords;
    // dense FP64 fields (the authoritative SdfData arrays), read by the FP64
    // query kernels: phi, n (3 per voxel), H (6 per voxel), candidate phi
    // (S per voxel) and normals (3S per voxel)
    const double *phi64, *n64, *h64, *cphi64, *cn64;
    // dense, optional: FP64 cell records (the 8 corners' (phi, n) per cell)
    const double4* cellRec64;
};

// The 8 packed corner records of cell c (dense or sparse cell records), or
// nullptr when the cell has no record because a corner sample is missing.
// Requires g.cellRec or g.scellRec and a cell inside the grid.
__device__ __forceinline__ const float4* cellRecord(const QGrid& g, int i0, int j0, int k0)
{
    if (!g.sparse)
        return g.cellRec + 8 * (int64_t(i0) + int64_t(g.n[0] - 1) * (j0 + int64_t(g.n[1] - 1) * k0));
    const int bd = g.bd;
    const int bi = i0 / bd, bj = j0 / bd, bk = k0 / bd;
    const int32_t b = __ldg(&g.table[bi + int64_t(g.nbx) * (bj + int64_t(g.nby) * bk)]);
    if (b < 0) return nullptr;
    const int local = (i0 - bi * bd) + bd * ((j0 - bj * bd) + bd * (k0 - bk * bd));
    const unsigned long long m = __ldg(&g.scellMask[int64_t(b) * g.maskWords + local / 64]);
    if (!((m >> (local % 64)) & 1ull)) return nullptr;
    return g.scellRec + 8 * (int64_t(b) * (bd * bd * bd) + local);
}

template <typename T>
struct V3T
{
    T x, y, z;
};

// Voxel record lookup: returns the linear record index (dense voxel or
// brick-pool slot) or -1 when the sample does not exist (SdfData.h:216-247).
__device__ __forceinline__ int64_t recordOf(const QGrid& g, int i, int j, int k)
{
    if (i < 0 || i >= g.n[0] || j < 0 || j >= g.n[1] || k < 0 || k >= g.n[2]) return -1;
    if (!g.sparse) return i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k);
    const int bd = g.bd;
    const int bi = i / bd, bj = j / bd, bk = k / bd;
    const int32_t b = __ldg(&g.table[bi + int64_t(g.nbx) * (bj + int64_t(g.nby) * bk)]);
    if (b < 0) return -1;
    return int64_t(b) * (bd * bd * bd) + (i - bi * bd) + bd * ((j - bj * bd) + bd * (k - bk * bd));
}

// (phi, n) of a voxel record: the FP32 mirror; with F64, a dense grid's FP64
// fields or a sparse grid's FP32 fields widened (the reference's sparse data
// is FP32).
template <bool F64 = false>
__device__ __forceinline__ std::conditional_t<F64, double4, float4> fieldAt(const QGrid& g,
                                                                             int64_t rec)
{
    if constexpr (F64)
    {
        if (g.sparse)
        {
            const float4 f = g.spn[rec];
            return make_double4(f.x, f.y, f.z, f.w);
        }
        return make_double4(g.phi64[rec], g.n64[3 * rec], g.n64[3 * rec + 1], g.n64[3 * rec + 2]);
    }
    else
    {
        return g.sparse ? g.spn[rec] : g.pn[rec];
    }
}

// The 6 Hessian entries of a voxel record (FP64 for a dense grid with F64,
// else FP32 widened); false if the grid stores none.
template <bool F64 = false>
__device__ __forceinline__ bool hessianAt(const QGrid& g, int64_t rec,
                                          std::conditional_t<F64, double, float> (&h)[6])
{
    if (F64 && !g.sparse)
    {
        if (g.h64 == nullptr) return false;
#pragma unroll
        for (int e = 0; e < 6; ++e) h[e] = g.h64[6 * rec + e];
        return true;
    }
    const float* p = g.sparse ? g.sh6 : g.h6f;
    if (p == nullptr) return false;
#pragma unroll
    for (int e = 0; e < 6; ++e) h[e] = p[6 * rec + e];
    return true;
}

// (phi, n) of contact-aware candidate slot q (record * stride + c).
template <bool F64 = false>
__device__ __forceinline__ std::conditional_t<F64, double4, float4> candidateAt(const QGrid& g,
                                                                                int64_t q)
{
    if constexpr (F64)
    {
        if (g.sparse)
        {
            const float4 f = g.scpn[q];
            return make_double4(f.x, f.y, f.z, f.w);
        }
        return make_double4(g.cphi64[q], g.cn64[3 * q], g.cn64[3 * q + 1], g.cn64[3 * q + 2]);
    }
    else
    {
        return g.sparse ? g.scpn[q] : g.cpn[q];
    }
}

// Continuous cell location (SdfOracle.cpp:530-549).
struct Cell
{
    int i0, j0, k0;
    double u, v, w;
    bool valid;
};

__device__ __forceinline__ Cell locateCell(const QGrid& g, double px, double py, double pz)
{
    Cell c;
    const double fx = (px - g.bmin[0]) / g.voxel[0] - 0.5;
    const double fy = (py - g.bmin[1]) / g.voxel[1] - 0.5;
    const double fz = (pz - g.bmin[2]) / g.voxel[2] - 0.5;
    const double flx = floor(fx), fly = floor(fy), flz = floor(fz);
    // int conversion of an out-of-range double is undefined: clamp first.
    c.i0 = int(fmax(fmin(flx, 2.0e9), -2.0e9));
    c.j0 = int(fmax(fmin(fly, 2.0e9), -2.0e9));
    c.k0 = int(fmax(fmin(flz, 2.0e9), -2.0e9));
    c.u = fx - double(c.i0);
    c.v = fy - double(c.j0);
    c.w = fz - double(c.k0);
    c.valid = !(c.i0 < 0 || c.i0 + 1 >= g.n[0] || c.j0 < 0 || c.j0 + 1 >= g.n[1] || c.k0 < 0 ||
                c.k0 + 1 >= g.n[2]) &&
              (fx == fx) && (fy == fy) && (fz == fz);
    return c;
}

// Nearest voxel (GridSpec.h:37-45; std::round = half away from zero).
__device__ __forceinline__ bool nearestVoxel(const QGrid& g, double px, double py, double pz,
                                             int& i, int& j, int& k)
{
    const double fx = round((px - g.bmin[0]) / g.voxel[0] - 0.5);
    const double fy = round((py - g.bmin[1]) / g.voxel[1] - 0.5);
    const double fz = round((pz - g.bmin[2]) / g.voxel[2] - 0.5);
    if (!(fx == fx) || !(fy == fy) || !(fz == fz)) return false;
    i = int(fmax(fmin(fx, 2.0e9), -2.0e9));
    j = int(fmax(fmin(fy, 2.0e9), -2.0e9));
    k = int(fmax(fmin(fz, 2.0e9), -2.0e9));
    return true;
}

// Result of one interpolation.
template <typename T>
struct Interp
{
    T gap;
    V3T<T> grad;
    T hess[6]; // xx, xy, xz, yy, yz, zz (only with WithHess)
};

// a * b + c rounded once, in FP32 or FP64.
template <typename T>
__device__ __forceinline__ T fmadd(T a, T b, T c)
{
    if constexpr (std::is_same_v<T, float>)
        return __builtin_fmaf(a, b, c);
    else
        return __builtin_fma(a, b, c);
}

// Hermite basis (SdfOracle.cpp:976-979, 1009-1012) with first and second
// derivatives, kept in named registers.
template <typename T>
struct Basis
{
    T h0, h1, h2, h3;     // h00, h10, h01, h11
    T d0, d1, d2, d3;     // derivatives
    T e0, e1, e2, e3;     // second derivatives
    __device__ __forceinline__ T v(int corner, int kind) const // kind 0: value weight, 1: slope weight
    {
        return corner ? (kind ? h3 : h2) : (kind ? h1 : h0);
    }
    __device__ __forceinline__ T d(int corner, int kind) const
    {
        return corner ? (kind ? d3 : d2) : (kind ? d1 : d0);
    }
    __device__ __forceinline__ T e(int corner, int kind) const
    {
        return corner ? (kind ? e3 : e2) : (kind ? e1 : e0);
    }
};

template <typename T, bool Second>
__device__ __forceinline__ Basis<T> hermite(T t)
{
    Basis<T> b;
    const T t2 = t * t;
    b.h0 = t2 * (T(2) * t - T(3)) + T(1);
    b.h1 = t * (t * (t - T(2)) + T(1));
    b.h2 = t2 * (T(3) - T(2) * t);
    b.h3 = t2 * (t - T(1));
    b.d0 = T(6) * t * (t - T(1));
    b.d1 = t * (T(3) * t - T(4)) + T(1);
    b.d2 = T(6) * t * (T(1) - t);
    b.d3 = t * (T(3) * t - T(2));
    if (Second)
    {
        b.e0 = T(12) * t - T(6);
        b.e1 = T(6) * t - T(4);
        b.e2 = T(6) - T(12) * t;
        b.e3 = T(6) * t - T(2);
    }
    else
    {
        b.e0 = b.e1 = b.e2 = b.e3 = T(0);
    }
    return b;
}

// The 8 corner records (phi, n) of a cell; false if a sample is missing.
template <bool F64 = false>
__device__ __forceinline__ bool cellCorners(const QGrid& g, const Cell& c,
                                            std::conditional_t<F64, double4, float4> (&f)[8])
{
    if constexpr (F64)
    {
        if (!g.sparse && g.cellRec64 != nullptr)
        {
            const double4* r = g.cellRec64 + 8 * (int64_t(c.i0) + int64_t(g.n[0] - 1) *
                                                                      (c.j0 + int64_t(g.n[1] - 1) * c.k0));
#pragma unroll
            for (int q = 0; q < 8; ++q) f[q] = r[q];
            return true;
        }
    }
    if (!F64 && (g.sparse ? g.scellRec != nullptr : g.cellRec != nullptr))
    {
        if constexpr (!F64)
        {
            const float4* r = cellRecord(g, c.i0, c.j0, c.k0);
            if (!r) return false;
#pragma unroll
            for (int q = 0; q < 8; ++q) f[q] = r[q];
            return true;
        }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
    {
        const int64_t rec = recordOf(g, c.i0 + (q & 1), c.j0 + ((q >> 1) & 1), c.k0 + (q >> 2));
        if (rec < 0) return false;
        f[q] = fieldAt<F64>(g, rec);
    }
    return true;
}

// Tensor-product cubic Hermite on the 2x2x2 cell corners with corner data
// (phi, h * n) and no cross-derivative terms (SdfOracle.cpp:932-1092).
// Returns false if a corner sample is missing (sparse).
template <typename T, bool WithHess, typename F4>
__device__ __forceinline__ void tricubicCorners(const QGrid& g, const Cell& c, const F4 (&f)[8],
                                                Interp<T>& r);

// The value alone of the FP32-corner interpolant, with the operations that
// tricubicCorners uses for its gap (the x, y and z reductions of the value
// terms), for screens that only need the sign of the gap.
template <typename T>
__device__ __forceinline__ T tricubicValueCorners(const QGrid& g, const Cell& c, const float4 (&f)[8])
{
    const Basis<T> A = hermite<T, false>(T(c.u));
    const Basis<T> B = hermite<T, false>(T(c.v));
    const Basis<T> Cz = hermite<T, false>(T(c.w));
    const T vx = T(g.voxel[0]), vy = T(g.voxel[1]), vz = T(g.voxel[2]);
    auto dot4 = [](T a0, T x0, T a1, T x1, T a2, T x2, T a3, T x3) {
        return fmadd(a0, x0, fmadd(a1, x1, fmadd(a2, x2, a3 * x3)));
    };
    auto dot2 = [](T a0, T x0, T a1, T x1) { return fmadd(a0, x0, a1 * x1); };
    const T a0[2] = {A.h0, A.h2}, a1[2] = {A.h1 * vx, A.h3 * vx};
    T X0[4], Xy[4], Xz[4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
    {
        const float4 e = f[2 * p], o = f[2 * p + 1];
        X0[p] = dot4(a0[0], T(e.x), a1[0], T(e.y), a0[1], T(o.x), a1[1], T(o.y));
        Xy[p] = dot2(a0[0], T(e.z), a0[1], T(o.z));
        Xz[p] = dot2(a0[0], T(e.w), a0[1], T(o.w));
    }
    const T b0[2] = {B.h0, B.h2}, b1[2] = {B.h1 * vy, B.h3 * vy};
    T Y0[2], Yz[2];
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
    {
        const int p = 2 * dk, q = 2 * dk + 1;
        Y0[dk] = dot4(b0[0], X0[p], b1[0], Xy[p], b0[1], X0[q], b1[1], Xy[q]);
        Yz[dk] = dot2(b0[0], Xz[p], b0[1], Xz[q]);
    }
    const T c0[2] = {Cz.h0, Cz.h2}, c1[2] = {Cz.h1 * vz, Cz.h3 * vz};
    return dot4(c0[0], Y0[0], c1[0], Yz[0], c0[1], Y0[1], c1[1], Yz[1]);
}

template <typename T, bool WithHess, bool F64 = false>
__device__ __forceinline__ bool tricubic(const QGrid& g, const Cell& c, Interp<T>& r)
{
    std::conditional_t<F64, double4, float4> f[8];
    if (!cellCorners<F64>(g, c, f)) return false;
    tricubicCorners<T, WithHess>(g, c, f, r);
    return true;
}

// The interpolation itself, on corner records already loaded.
template <typename T, bool WithHess, typename F4>
__device__ __forceinline__ void tricubicCorners(const QGrid& g, const Cell& c, const F4 (&f)[8],
                                                Interp<T>& r)
{
    constexpr bool kSeq = std::is_same_v<F4, double4>;
    const Basis<T> A = hermite<T, WithHess>(T(c.u));
    const Basis<T> B = hermite<T, WithHess>(T(c.v));
    const Basis<T> Cz = hermite<T, WithHess>(T(c.w));
    const T vx = T(g.voxel[0]), vy = T(g.voxel[1]), vz = T(g.voxel[2]);
    if constexpr (!kSeq)
    {
        // The same interpolant reduced one axis at a time (x over corner
        // pairs, then y, then z): ~120 multiply-adds instead of ~400 for the
        // corner-by-corner sum.  The voxel size of each axis multiplies that
        // axis's slope weight, so n needs no per-corner scaling.  Explicit
        // fused multiply-adds fix the rounding, so every kernel that inlines
        // this returns the same bits for the same corners.
        auto dot4 = [](T a0, T x0, T a1, T x1, T a2, T x2, T a3, T x3) {
            return fmadd(a0, x0, fmadd(a1, x1, fmadd(a2, x2, a3 * x3)));
        };
        auto dot2 = [](T a0, T x0, T a1, T x1) { return fmadd(a0, x0, a1 * x1); };
        const T a0[2] = {A.h0, A.h2}, a1[2] = {A.h1 * vx, A.h3 * vx};
        const T da0[2] = {A.d0, A.d2}, da1[2] = {A.d1 * vx, A.d3 * vx};
        const T ea0[2] = {A.e0, A.e2}, ea1[2] = {A.e1 * vx, A.e3 * vx};
        T X0[4], Xy[4], Xz[4], dX0[4], dXy[4], dXz[4], eX0[4], eXy[4], eXz[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) // p = dj + 2 dk, corners 2p (di = 0) and 2p + 1
        {
            const F4 e = f[2 * p], o = f[2 * p + 1];
            X0[p] = dot4(a0[0], T(e.x), a1[0], T(e.y), a0[1], T(o.x), a1[1], T(o.y));
            Xy[p] = dot2(a0[0], T(e.z), a0[1], T(o.z));
            Xz[p] = dot2(a0[0], T(e.w), a0[1], T(o.w));
            dX0[p] = dot4(da0[0], T(e.x), da1[0], T(e.y), da0[1], T(o.x), da1[1], T(o.y));
            dXy[p] = dot2(da0[0], T(e.z), da0[1], T(o.z));
            dXz[p] = dot2(da0[0], T(e.w), da0[1], T(o.w));
            if constexpr (WithHess)
            {
                eX0[p] = dot4(ea0[0], T(e.x), ea1[0], T(e.y), ea0[1], T(o.x), ea1[1], T(o.y));
                eXy[p] = dot2(ea0[0], T(e.z), ea0[1], T(o.z));
                eXz[p] = dot2(ea0[0], T(e.w), ea0[1], T(o.w));
            }
        }
        const T b0[2] = {B.h0, B.h2}, b1[2] = {B.h1 * vy, B.h3 * vy};
        const T db0[2] = {B.d0, B.d2}, db1[2] = {B.d1 * vy, B.d3 * vy};
        const T eb0[2] = {B.e0, B.e2}, eb1[2] = {B.e1 * vy, B.e3 * vy};
        T Y0[2], Yz[2], uY0[2], uYz[2], vY0[2], vYz[2];
        T uuY0[2], uuYz[2], vvY0[2], vvYz[2], uvY0[2], uvYz[2];
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
        {
            const int p = 2 * dk, q = 2 * dk + 1; // dj = 0, 1
            Y0[dk] = dot4(b0[0], X0[p], b1[0], Xy[p], b0[1], X0[q], b1[1], Xy[q]);
            Yz[dk] = dot2(b0[0], Xz[p], b0[1], Xz[q]);
            uY0[dk] = dot4(b0[0], dX0[p], b1[0], dXy[p], b0[1], dX0[q], b1[1], dXy[q]);
            uYz[dk] = dot2(b0[0], dXz[p], b0[1], dXz[q]);
            vY0[dk] = dot4(db0[0], X0[p], db1[0], Xy[p], db0[1], X0[q], db1[1], Xy[q]);
            vYz[dk] = dot2(db0[0], Xz[p], db0[1], Xz[q]);
            if constexpr (WithHess)
            {
                uuY0[dk] = dot4(b0[0], eX0[p], b1[0], eXy[p], b0[1], eX0[q], b1[1], eXy[q]);
                uuYz[dk] = dot2(b0[0], eXz[p], b0[1], eXz[q]);
                vvY0[dk] = dot4(eb0[0], X0[p], eb1[0], Xy[p], eb0[1], X0[q], eb1[1], Xy[q]);
                vvYz[dk] = dot2(eb0[0], Xz[p], eb0[1], Xz[q]);
                uvY0[dk] = dot4(db0[0], dX0[p], db1[0], dXy[p], db0[1], dX0[q], db1[1], dXy[q]);
                uvYz[dk] = dot2(db0[0], dXz[p], db0[1], dXz[q]);
            }
        }
        const T c0[2] = {Cz.h0, Cz.h2}, c1[2] = {Cz.h1 * vz, Cz.h3 * vz};
        const T dc0[2] = {Cz.d0, Cz.d2}, dc1[2] = {Cz.d1 * vz, Cz.d3 * vz};
        r.gap = dot4(c0[0], Y0[0], c1[0], Yz[0], c0[1], Y0[1], c1[1], Yz[1]);
        const T fu = dot4(c0[0], uY0[0], c1[0], uYz[0], c0[1], uY0[1], c1[1], uYz[1]);
        const T fv = dot4(c0[0], vY0[0], c1[0], vYz[0], c0[1], vY0[1], c1[1], vYz[1]);
        const T fw = dot4(dc0[0], Y0[0], dc1[0], Yz[0], dc0[1], Y0[1], dc1[1], Yz[1]);
        r.grad = V3T<T>{fu / vx, fv / vy, fw / vz};
        if constexpr (WithHess)
        {
            const T ec0[2] = {Cz.e0, Cz.e2}, ec1[2] = {Cz.e1 * vz, Cz.e3 * vz};
            auto zc = [&](const T (&y0)[2], const T (&yz)[2]) {
                return dot4(c0[0], y0[0], c1[0], yz[0], c0[1], y0[1], c1[1], yz[1]);
            };
            auto zd = [&](const T (&y0)[2], const T (&yz)[2]) {
                return dot4(dc0[0], y0[0], dc1[0], yz[0], dc0[1], y0[1], dc1[1], yz[1]);
            };
            r.hess[0] = zc(uuY0, uuYz) / (vx * vx);
            r.hess[1] = zc(uvY0, uvYz) / (vx * vy);
            r.hess[2] = zd(uY0, uYz) / (vx * vz);
            r.hess[3] = zc(vvY0, vvYz) / (vy * vy);
            r.hess[4] = zd(vY0, vYz) / (vy * vz);
            r.hess[5] = dot4(ec0[0], Y0[0], ec1[0], Yz[0], ec0[1], Y0[1], ec1[1], Yz[1]) / (vz * vz);
        }
        return;
    }
    T gap = 0, fu = 0, fv = 0, fw = 0;
    T huu = 0, huv = 0, huw = 0, hvv = 0, hvw = 0, hww = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
    {
        const int di = q & 1, dj = (q >> 1) & 1, dk = q >> 2;
        const T phi = T(f[q].x);
        const T gx = T(f[q].y) * vx, gy = T(f[q].z) * vy, gz = T(f[q].w) * vz;
        const T P0 = A.v(di, 0), P1 = A.v(di, 1), dP0 = A.d(di, 0), dP1 = A.d(di, 1);
        const T Q0 = B.v(dj, 0), Q1 = B.v(dj, 1), dQ0 = B.d(dj, 0), dQ1 = B.d(dj, 1);
        const T R0 = Cz.v(dk, 0), R1 = Cz.v(dk, 1), dR0 = Cz.d(dk, 0), dR1 = Cz.d(dk, 1);
        // s(A,B,C) = A0 B0 C0 phi + A1 B0 C0 gx + A0 B1 C0 gy + A0 B0 C1 gz.  FP64
        // corners add the four terms one at a time into the running sum, the
        // CPU path's order (SdfOracle.cpp:999-1078), so its results are its bits;
        // FP32 corners sum each corner's terms first (shorter dependency chain).
        auto acc = [&](T& s, T a, T b, T c, T d) {
            if constexpr (kSeq)
            {
                s += a;
                s += b;
                s += c;
                s += d;
            }
            else
            {
                s += a + b + c + d;
            }
        };
        acc(gap, (P0 * Q0 * R0) * phi, (P1 * Q0 * R0) * gx, (P0 * Q1 * R0) * gy, (P0 * Q0 * R1) * gz);
        acc(fu, (dP0 * Q0 * R0) * phi, (dP1 * Q0 * R0) * gx, (dP0 * Q1 * R0) * gy,
            (dP0 * Q0 * R1) * gz);
        acc(fv, (P0 * dQ0 * R0) * phi, (P1 * dQ0 * R0) * gx, (P0 * dQ1 * R0) * gy,
            (P0 * dQ0 * R1) * gz);
        acc(fw, (P0 * Q0 * dR0) * phi, (P1 * Q0 * dR0) * gx, (P0 * Q1 * dR0) * gy,
            (P0 * Q0 * dR1) * gz);
        if (WithHess)
        {
            const T eP0 = A.e(di, 0), eP1 = A.e(di, 1);
            const T eQ0 = B.e(dj, 0), eQ1 = B.e(dj, 1);
            const T eR0 = Cz.e(dk, 0), eR1 = Cz.e(dk, 1);
            acc(huu, (eP0 * Q0 * R0) * phi, (eP1 * Q0 * R0) * gx, (eP0 * Q1 * R0) * gy,
                (eP0 * Q0 * R1) * gz);
            acc(hvv, (P0 * eQ0 * R0) * phi, (P1 * eQ0 * R0) * gx, (P0 * eQ1 * R0) * gy,
                (P0 * eQ0 * R1) * gz);
            acc(hww, (P0 * Q0 * eR0) * phi, (P1 * Q0 * eR0) * gx, (P0 * Q1 * eR0) * gy,
                (P0 * Q0 * eR1) * gz);
            acc(huv, (dP0 * dQ0 * R0) * phi, (dP1 * dQ0 * R0) * gx, (dP0 * dQ1 * R0) * gy,
                (dP0 * dQ0 * R1) * gz);
            acc(huw, (dP0 * Q0 * dR0) * phi, (dP1 * Q0 * dR0) * gx, (dP0 * Q1 * dR0) * gy,
                (dP0 * Q0 * dR1) * gz);
            acc(hvw, (P0 * dQ0 * dR0) * phi, (P1 * dQ0 * dR0) * gx, (P0 * dQ1 * dR0) * gy,
                (P0 * dQ0 * dR1) * gz);
        }
    }
    r.gap = gap;
    r.grad = V3T<T>{fu / vx, fv / vy, fw / vz};
    if (WithHess)
    {
        r.hess[0] = huu / (vx * vx);
        r.hess[1] = huv / (vx * vy);
        r.hess[2] = huw / (vx * vz);
        r.hess[3] = hvv / (vy * vy);
        r.hess[4] = hvw / (vy * vz);
        r.hess[5] = hww / (vz * vz);
    }
}

// Trilinear value and analytic gradient (SdfOracle.cpp:551-583); its only
// non-zero second derivatives are the mixed ones.
template <typename T, bool WithHess, typename F4>
__device__ __forceinline__ void trilinearCorners(const QGrid& g, const Cell& c, const F4 (&f)[8],
                                                 Interp<T>& r);

template <typename T, bool WithHess, bool F64 = false>
__device__ __forceinline__ bool trilinear(const QGrid& g, const Cell& c, Interp<T>& r)
{
    std::conditional_t<F64, double4, float4> f[8];
    if (!cellCorners<F64>(g, c, f)) return false;
    trilinearCorners<T, WithHess>(g, c, f, r);
    return true;
}

template <typename T, bool WithHess, typename F4>
__device__ __forceinline__ void trilinearCorners(const QGrid& g, const Cell& c, const F4 (&f)[8],
                                                 Interp<T>& r)
{
    T p[2][2][2];
#pragma unroll
    for (int q = 0; q < 8; ++q) p[q & 1][(q >> 1) & 1][q >> 2] = T(f[q].x);
    const T u = T(c.u), v = T(c.v), w = T(c.w);
    const T iu = T(1) - u, iv = T(1) - v, iw = T(1) - w;
    const T c00 = p[0][0][0] * iu + p[1][0][0] * u;
    const T c10 = p[0][1][0] * iu + p[1][1][0] * u;
    const T c01 = p[0][0][1] * iu + p[1][0][1] * u;
    const T c11 = p[0][1][1] * iu + p[1][1][1] * u;
    const T c0 = c00 * iv + c10 * v;
    const T c1 = c01 * iv + c11 * v;
    r.gap = c0 * iw + c1 * w;
    const T ex00 = p[1][0][0] - p[0][0][0], ex10 = p[1][1][0] - p[0][1][0];
    const T ex01 = p[1][0][1] - p[0][0][1], ex11 = p[1][1][1] - p[0][1][1];
    const T ey00 = p[0][1][0] - p[0][0][0], ey10 = p[1][1][0] - p[1][0][0];
    const T ey01 = p[0][1][1] - p[0][0][1], ey11 = p[1][1][1] - p[1][0][1];
    const T ez00 = p[0][0][1] - p[0][0][0], ez10 = p[1][0][1] - p[1][0][0];
    const T ez01 = p[0][1][1] - p[0][1][0], ez11 = p[1][1][1] - p[1][1][0];
    const T vx = T(g.voxel[0]), vy = T(g.voxel[1]), vz = T(g.voxel[2]);
    r.grad.x = (ex00 * iv * iw + ex10 * v * iw + ex01 * iv * w + ex11 * v * w) / vx;
    r.grad.y = (ey00 * iu * iw + ey10 * u * iw + ey01 * iu * w + ey11 * u * w) / vy;
    r.grad.z = (ez00 * iu * iv + ez10 * u * iv + ez01 * iu * v + ez11 * u * v) / vz;
    if (WithHess)
    {
        r.hess[0] = T(0);
        r.hess[1] = ((ex10 - ex00) * iw + (ex11 - ex01) * w) / (vx * vy);
        r.hess[2] = ((ex01 - ex00) * iv + (ex11 - ex10) * v) / (vx * vz);
        r.hess[3] = T(0);
        r.hess[4] = ((ey01 - ey00) * iu + (ey11 - ey10) * u) / (vy * vz);
        r.hess[5] = T(0);
    }
}

} // namespace dev
} // namespace sdfo
