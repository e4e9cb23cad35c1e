// Field access and interpolation for the query and contact kernels.
//
// A point is located in FP64 exactly as the reference does it
// (f = (p - bmin)/voxel - 0.5, floor / std::round; SdfOracle.cpp:530-549,
// GridSpec.h:37-45), so the valid mask and the chosen cell/voxel are exact;
// the interpolation itself runs in T (float for the query stream, double for
// the contact reduction) on the FP32 field mirror.
#pragma once

#include <hip/hip_runtime.h>

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
};

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

__device__ __forceinline__ float4 fieldAt(const QGrid& g, int64_t rec)
{
    return g.sparse ? g.spn[rec] : g.pn[rec];
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

// Hermite basis (SdfOracle.cpp:976-979, 1009-1012) and second derivatives.
template <typename T>
__device__ __forceinline__ void hermite(T t, T* h, T* dh, T* ddh)
{
    const T t2 = t * t;
    h[0] = t2 * (T(2) * t - T(3)) + T(1);
    h[1] = t * (t * (t - T(2)) + T(1));
    h[2] = t2 * (T(3) - T(2) * t);
    h[3] = t2 * (t - T(1));
    dh[0] = T(6) * t * (t - T(1));
    dh[1] = t * (T(3) * t - T(4)) + T(1);
    dh[2] = T(6) * t * (T(1) - t);
    dh[3] = t * (T(3) * t - T(2));
    if (ddh)
    {
        ddh[0] = T(12) * t - T(6);
        ddh[1] = T(6) * t - T(4);
        ddh[2] = T(6) - T(12) * t;
        ddh[3] = T(6) * t - T(2);
    }
}

// Tensor-product cubic Hermite on the 2x2x2 cell corners with corner data
// (phi, h * n) and no cross-derivative terms (SdfOracle.cpp:932-1092).
// Returns false if a corner sample is missing (sparse).
template <typename T, bool WithHess>
__device__ __forceinline__ bool tricubic(const QGrid& g, const Cell& c, Interp<T>& r)
{
    T a[4], da[4], ea[4], b[4], db[4], eb[4], cc[4], dc[4], ec[4];
    hermite<T>(T(c.u), a, da, WithHess ? ea : nullptr);
    hermite<T>(T(c.v), b, db, WithHess ? eb : nullptr);
    hermite<T>(T(c.w), cc, dc, WithHess ? ec : nullptr);
    const T vx = T(g.voxel[0]), vy = T(g.voxel[1]), vz = T(g.voxel[2]);
    T gap = 0, fu = 0, fv = 0, fw = 0;
    T huu = 0, huv = 0, huw = 0, hvv = 0, hvw = 0, hww = 0;
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di)
            {
                const int64_t rec = recordOf(g, c.i0 + di, c.j0 + dj, c.k0 + dk);
                if (rec < 0) return false;
                const float4 f = fieldAt(g, rec);
                const T phi = T(f.x);
                const T gx = T(f.y) * vx, gy = T(f.z) * vy, gz = T(f.w) * vz;
                const T P0 = a[2 * di], P1 = a[2 * di + 1];
                const T Q0 = b[2 * dj], Q1 = b[2 * dj + 1];
                const T R0 = cc[2 * dk], R1 = cc[2 * dk + 1];
                const T dP0 = da[2 * di], dP1 = da[2 * di + 1];
                const T dQ0 = db[2 * dj], dQ1 = db[2 * dj + 1];
                const T dR0 = dc[2 * dk], dR1 = dc[2 * dk + 1];
                // s(A,B,C) = A0 B0 C0 phi + A1 B0 C0 gx + A0 B1 C0 gy + A0 B0 C1 gz
                gap += (P0 * Q0 * R0) * phi + (P1 * Q0 * R0) * gx + (P0 * Q1 * R0) * gy +
                       (P0 * Q0 * R1) * gz;
                fu += (dP0 * Q0 * R0) * phi + (dP1 * Q0 * R0) * gx + (dP0 * Q1 * R0) * gy +
                      (dP0 * Q0 * R1) * gz;
                fv += (P0 * dQ0 * R0) * phi + (P1 * dQ0 * R0) * gx + (P0 * dQ1 * R0) * gy +
                      (P0 * dQ0 * R1) * gz;
                fw += (P0 * Q0 * dR0) * phi + (P1 * Q0 * dR0) * gx + (P0 * Q1 * dR0) * gy +
                      (P0 * Q0 * dR1) * gz;
                if (WithHess)
                {
                    const T eP0 = ea[2 * di], eP1 = ea[2 * di + 1];
                    const T eQ0 = eb[2 * dj], eQ1 = eb[2 * dj + 1];
                    const T eR0 = ec[2 * dk], eR1 = ec[2 * dk + 1];
                    huu += (eP0 * Q0 * R0) * phi + (eP1 * Q0 * R0) * gx + (eP0 * Q1 * R0) * gy +
                           (eP0 * Q0 * R1) * gz;
                    hvv += (P0 * eQ0 * R0) * phi + (P1 * eQ0 * R0) * gx + (P0 * eQ1 * R0) * gy +
                           (P0 * eQ0 * R1) * gz;
                    hww += (P0 * Q0 * eR0) * phi + (P1 * Q0 * eR0) * gx + (P0 * Q1 * eR0) * gy +
                           (P0 * Q0 * eR1) * gz;
                    huv += (dP0 * dQ0 * R0) * phi + (dP1 * dQ0 * R0) * gx +
                           (dP0 * dQ1 * R0) * gy + (dP0 * dQ0 * R1) * gz;
                    huw += (dP0 * Q0 * dR0) * phi + (dP1 * Q0 * dR0) * gx +
                           (dP0 * Q1 * dR0) * gy + (dP0 * Q0 * dR1) * gz;
                    hvw += (P0 * dQ0 * dR0) * phi + (P1 * dQ0 * dR0) * gx +
                           (P0 * dQ1 * dR0) * gy + (P0 * dQ0 * dR1) * gz;
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
    return true;
}

// Trilinear value and analytic gradient (SdfOracle.cpp:551-583); its only
// non-zero second derivatives are the mixed ones.
template <typename T, bool WithHess>
__device__ __forceinline__ bool trilinear(const QGrid& g, const Cell& c, Interp<T>& r)
{
    T p[2][2][2];
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di)
            {
                const int64_t rec = recordOf(g, c.i0 + di, c.j0 + dj, c.k0 + dk);
                if (rec < 0) return false;
                p[di][dj][dk] = T(fieldAt(g, rec).x);
            }
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
    return true;
}

} // namespace dev
} // namespace sdfo
