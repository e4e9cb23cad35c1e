// Every ContactAware candidate of a point (SdfOracle.cpp:269-432), FP64 points,
// compiled without FMA contraction.  Dense grids read the FP64 candidate
// arrays, sparse grids their FP32 arrays upcast, as the CPU path does, so the
// candidates (gap, unit normal, witness, branch) are the CPU path's.  Used by
// the per-point C++ API (SdfOracle::queryFast / query); the bulk FP32 query
// keeps returning only the selected candidate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "sdfo_internal.hpp"

namespace sdfo {

namespace {

struct CandGrid
{
    double bmin[3], voxel[3];
    int n[3];
    int stride;
    int sparse;
    // dense (FP64)
    const uint8_t* ccount;
    const double* cphi;
    const double* cn3;
    const int32_t* cbranch;
    // sparse (FP32)
    int bd, nbx, nby, nbz;
    const int32_t* table;
    const uint8_t* sccount;
    const float* scphi;
    const float* scn3;
    const int32_t* scbranch;
};

constexpr int kMaxCand = 8; // kMaxSdfQueryCandidates (SdfCommon.h)

__global__ void __launch_bounds__(256)
candidatesKernel(CandGrid g, const double* __restrict__ pts, int64_t n, uint8_t* __restrict__ count,
                 double* __restrict__ cgap, double* __restrict__ cn3, double* __restrict__ cw3,
                 int32_t* __restrict__ cb, double* __restrict__ gap, double* __restrict__ normal3,
                 double* __restrict__ witness3, uint8_t* __restrict__ valid)
{
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    count[q] = 0;
    const double px = pts[3 * q], py = pts[3 * q + 1], pz = pts[3 * q + 2];
    // GridSpec::indexFromWorld (GridSpec.h:36-44)
    const double fx = round((px - g.bmin[0]) / g.voxel[0] - 0.5);
    const double fy = round((py - g.bmin[1]) / g.voxel[1] - 0.5);
    const double fz = round((pz - g.bmin[2]) / g.voxel[2] - 0.5);
    if (!(fx == fx) || !(fy == fy) || !(fz == fz)) return;
    const int i = int(fmax(fmin(fx, 2.0e9), -2.0e9));
    const int j = int(fmax(fmin(fy, 2.0e9), -2.0e9));
    const int k = int(fmax(fmin(fz, 2.0e9), -2.0e9));
    if (i < 0 || i >= g.n[0] || j < 0 || j >= g.n[1] || k < 0 || k >= g.n[2]) return;
    int64_t rec;
    int cnt;
    if (g.sparse)
    {
        const int32_t b = g.table[(i / g.bd) + int64_t(g.nbx) * ((j / g.bd) + int64_t(g.nby) * (k / g.bd))];
        if (b < 0) return;
        rec = int64_t(b) * g.bd * g.bd * g.bd + (i % g.bd) + g.bd * ((j % g.bd) + g.bd * (k % g.bd));
        cnt = int(g.sccount[rec]);
    }
    else
    {
        rec = i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k);
        cnt = int(g.ccount[rec]);
    }
    cnt = cnt < g.stride ? cnt : g.stride;
    if (cnt <= 0) return; // first-order fallback: the caller keeps the plain query result
    cnt = cnt < kMaxCand ? cnt : kMaxCand;
    const double dx = px - (g.bmin[0] + g.voxel[0] * (i + 0.5));
    const double dy = py - (g.bmin[1] + g.voxel[1] * (j + 0.5));
    const double dz = pz - (g.bmin[2] + g.voxel[2] * (k + 0.5));
    double best = DBL_MAX;
    int sel = 0;
    for (int c = 0; c < cnt; ++c)
    {
        const int64_t e = rec * g.stride + c;
        double ph, nx, ny, nz;
        int32_t br;
        if (g.sparse)
        {
            ph = double(g.scphi[e]);
            nx = double(g.scn3[3 * e]);
            ny = double(g.scn3[3 * e + 1]);
            nz = double(g.scn3[3 * e + 2]);
            br = g.scbranch[e];
        }
        else
        {
            ph = g.cphi[e];
            nx = g.cn3[3 * e];
            ny = g.cn3[3 * e + 1];
            nz = g.cn3[3 * e + 2];
            br = g.cbranch[e];
        }
        const double nn = sqrt(nx * nx + ny * ny + nz * nz);
        if (nn > 1.0e-15)
        {
            nx /= nn;
            ny /= nn;
            nz /= nn;
        }
        else
        {
            nx = 1.0;
            ny = 0.0;
            nz = 0.0;
        }
        const double gp = ph + (nx * dx + ny * dy + nz * dz);
        const int64_t o = q * kMaxCand + c;
        cgap[o] = gp;
        cn3[3 * o] = nx;
        cn3[3 * o + 1] = ny;
        cn3[3 * o + 2] = nz;
        cw3[3 * o] = px - gp * nx;
        cw3[3 * o + 1] = py - gp * ny;
        cw3[3 * o + 2] = pz - gp * nz;
        cb[o] = br;
        const double ag = fabs(gp);
        if (ag < best)
        {
            best = ag;
            sel = c;
        }
    }
    count[q] = uint8_t(cnt);
    const int64_t o = q * kMaxCand + sel;
    gap[q] = cgap[o];
    for (int a = 0; a < 3; ++a)
    {
        normal3[3 * q + a] = cn3[3 * o + a];
        witness3[3 * q + a] = cw3[3 * o + a];
    }
    valid[q] = 1;
}

} // namespace

bool gridHasCandidates(const sdfo_grid& g)
{
    return g.stride > 0 && (g.sparse ? (g.sccount != nullptr) : (g.ccount != nullptr));
}

void launchCandidatesF64(const sdfo_grid& g, const double* pts, int64_t n, uint8_t* count,
                         double* cgap, double* cn3, double* cw3, int32_t* cb, double* gap,
                         double* normal3, double* witness3, uint8_t* valid, hipStream_t s)
{
    if (n <= 0 || !gridHasCandidates(g)) return;
    CandGrid c{};
    for (int a = 0; a < 3; ++a)
    {
        c.bmin[a] = g.spec.bmin[a];
        c.voxel[a] = g.spec.voxel[a];
        c.n[a] = g.spec.n[a];
    }
    c.stride = g.stride;
    c.sparse = g.sparse ? 1 : 0;
    c.ccount = g.ccount;
    c.cphi = g.cphi;
    c.cn3 = g.cn3;
    c.cbranch = g.cbranch;
    c.bd = g.blockDim;
    c.nbx = g.nbx;
    c.nby = g.nby;
    c.nbz = g.nbz;
    c.table = g.table;
    c.sccount = g.sccount;
    c.scphi = g.scphi;
    c.scn3 = g.scn3;
    c.scbranch = g.scbranch;
    hipLaunchKernelGGL(candidatesKernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, c, pts, n,
                       count, cgap, cn3, cw3, cb, gap, normal3, witness3, valid);
    SDFO_HIP(hipGetLastError());
}

} // namespace sdfo
