// SDF build kernels for gfx950.  Compiled with -ffp-contract=off: the FP64
// arithmetic must round exactly like the CPU path it replaces.
#include <hip/hip_runtime.h>

#include "geom_device.hpp"

namespace sdfo {

using dev::D3;

namespace {

constexpr int kFillBlock = 128; // 2 waves; LDS stack = kStackDepth * 128 * 4 B = 16 KB

// Dense fill: one lane per voxel, one 4x4x4 voxel brick per wave so that the
// lanes of a wave walk nearly the same BVH path (SdfBuilder.cpp:1052-1100).
__global__ void __launch_bounds__(kFillBlock)
denseFillKernel(GridSpec g, const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                const int32_t* __restrict__ triIdx, D3 dir, int bx, int by, int bz,
                double* __restrict__ phi, double* __restrict__ n3, double* __restrict__ wit3,
                int32_t* __restrict__ tri)
{
    __shared__ int stack[kStackDepth * kFillBlock];
    const int lane = threadIdx.x & 63;
    const int64_t brick = int64_t(blockIdx.x) * (kFillBlock / 64) + (threadIdx.x >> 6);
    const int64_t nbricks = int64_t(bx) * by * bz;
    if (brick >= nbricks) return;
    const int bi = int(brick % bx);
    const int bj = int((brick / bx) % by);
    const int bk = int(brick / (int64_t(bx) * by));
    const int i = bi * 4 + (lane & 3);
    const int j = bj * 4 + ((lane >> 2) & 3);
    const int k = bk * 4 + (lane >> 4);
    if (i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) return;
    // GridSpec.h:30-34
    const D3 c = dev::d3(g.bmin[0] + g.voxel[0] * (i + 0.5), g.bmin[1] + g.voxel[1] * (j + 0.5),
                         g.bmin[2] + g.voxel[2] * (k + 0.5));
    dev::LaneStack st{stack + threadIdx.x, kFillBlock};
    const dev::SignedNearest s = dev::signedNearest(c, nodes, tri9, dir, st);
    const int64_t idx = i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k);
    const D3 gr = dev::voxelGradient(c, s);
    phi[idx] = s.sd;
    n3[3 * idx + 0] = gr.x;
    n3[3 * idx + 1] = gr.y;
    n3[3 * idx + 2] = gr.z;
    if (wit3)
    {
        wit3[3 * idx + 0] = s.cp.x;
        wit3[3 * idx + 1] = s.cp.y;
        wit3[3 * idx + 2] = s.cp.z;
    }
    if (tri) tri[idx] = s.slot >= 0 ? triIdx[s.slot] : -1;
}

// Central differences of the normal field, clamped at the border, then
// symmetrised (SdfBuilder.cpp:1105-1152).
__global__ void denseHessianKernel(GridSpec g, const double* __restrict__ n3,
                                   double* __restrict__ h6)
{
    const int64_t N = g.total();
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int i = int(idx % nx);
    const int j = int((idx / nx) % ny);
    const int k = int(idx / (int64_t(nx) * ny));
    auto at = [&](int ii, int jj, int kk) {
        ii = ii < 0 ? 0 : (ii > nx - 1 ? nx - 1 : ii);
        jj = jj < 0 ? 0 : (jj > ny - 1 ? ny - 1 : jj);
        kk = kk < 0 ? 0 : (kk > nz - 1 ? nz - 1 : kk);
        const int64_t l = ii + int64_t(nx) * (jj + int64_t(ny) * kk);
        return dev::d3(n3[3 * l], n3[3 * l + 1], n3[3 * l + 2]);
    };
    const D3 dx = dev::divs(dev::sub(at(i + 1, j, k), at(i - 1, j, k)), 2.0 * g.voxel[0]);
    const D3 dy = dev::divs(dev::sub(at(i, j + 1, k), at(i, j - 1, k)), 2.0 * g.voxel[1]);
    const D3 dz = dev::divs(dev::sub(at(i, j, k + 1), at(i, j, k - 1)), 2.0 * g.voxel[2]);
    double* o = h6 + 6 * idx;
    o[0] = 0.5 * (dx.x + dx.x);
    o[1] = 0.5 * (dy.x + dx.y);
    o[2] = 0.5 * (dz.x + dx.z);
    o[3] = 0.5 * (dy.y + dy.y);
    o[4] = 0.5 * (dz.y + dy.z);
    o[5] = 0.5 * (dz.z + dz.z);
}

// FP32 query mirror: float4 (phi, n) per voxel, 6 floats of Hessian.
__global__ void denseMirrorKernel(int64_t N, const double* __restrict__ phi,
                                  const double* __restrict__ n3, const double* __restrict__ h6,
                                  float4* __restrict__ pn, float* __restrict__ h6f)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= N) return;
    pn[i] = make_float4(float(phi[i]), float(n3[3 * i]), float(n3[3 * i + 1]), float(n3[3 * i + 2]));
    if (h6f)
    {
#pragma unroll
        for (int c = 0; c < 6; ++c) h6f[6 * i + c] = h6 ? float(h6[6 * i + c]) : 0.0f;
    }
}

// Packs the 8 corner records of every cell contiguously; one lane per
// (cell, corner) so both the gathers and the stores are coalesced.
__global__ void cellPackKernel(GridSpec g, const float4* __restrict__ pn, float4* __restrict__ rec)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int cx = g.n[0] - 1, cy = g.n[1] - 1, cz = g.n[2] - 1;
    const int64_t cells = int64_t(cx) * cy * cz;
    if (t >= cells * 8) return;
    const int64_t cell = t >> 3;
    const int q = int(t & 7);
    const int i = int(cell % cx) + (q & 1);
    const int j = int((cell / cx) % cy) + ((q >> 1) & 1);
    const int k = int(cell / (int64_t(cx) * cy)) + (q >> 2);
    rec[t] = pn[i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k)];
}

__global__ void candidateMirrorKernel(int64_t count, const double* __restrict__ cphi,
                                      const double* __restrict__ cn3, float4* __restrict__ cpn)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    cpn[i] = make_float4(float(cphi[i]), float(cn3[3 * i]), float(cn3[3 * i + 1]),
                         float(cn3[3 * i + 2]));
}

// Sparse brick activation: |phi(brick centre)| <= activation
// (SdfBuilder.cpp:869-890).  One lane per brick cell.
__global__ void __launch_bounds__(kFillBlock)
sparseActivateKernel(GridSpec g, const BvhNode* __restrict__ nodes,
                     const double* __restrict__ tri9, D3 dir, int bd, int nbx, int nby, int nbz,
                     double activation, uint8_t* __restrict__ mask)
{
    __shared__ int stack[kStackDepth * kFillBlock];
    const int64_t bl = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (bl >= int64_t(nbx) * nby * nbz) return;
    const int bi = int(bl % nbx), bj = int((bl / nbx) % nby), bk = int(bl / (int64_t(nbx) * nby));
    const int i0 = bi * bd, j0 = bj * bd, k0 = bk * bd;
    const int i1 = min(i0 + bd, g.n[0]), j1 = min(j0 + bd, g.n[1]), k1 = min(k0 + bd, g.n[2]);
    const D3 c = dev::d3(g.bmin[0] + g.voxel[0] * (0.5 * double(i0 + i1)),
                         g.bmin[1] + g.voxel[1] * (0.5 * double(j0 + j1)),
                         g.bmin[2] + g.voxel[2] * (0.5 * double(k0 + k1)));
    dev::LaneStack st{stack + threadIdx.x, kFillBlock};
    const dev::SignedNearest s = dev::signedNearest(c, nodes, tri9, dir, st);
    mask[bl] = fabs(s.sd) <= activation ? 1 : 0;
}

// Sparse fill: one lane per brick voxel, brick-major (SdfBuilder.cpp:925-965).
__global__ void __launch_bounds__(kFillBlock)
sparseFillKernel(GridSpec g, const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                 D3 dir, int bd, int64_t total, const int32_t* __restrict__ bijk,
                 float* __restrict__ sphi, float* __restrict__ sn3)
{
    __shared__ int stack[kStackDepth * kFillBlock];
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int bvc = bd * bd * bd;
    const int64_t b = t / bvc;
    const int local = int(t % bvc);
    const int lx = local % bd, ly = (local / bd) % bd, lz = local / (bd * bd);
    const int i = bijk[3 * b] * bd + lx, j = bijk[3 * b + 1] * bd + ly, k = bijk[3 * b + 2] * bd + lz;
    if (i < 0 || i >= g.n[0] || j < 0 || j >= g.n[1] || k < 0 || k >= g.n[2])
    {
        sphi[t] = 0.0f;
        sn3[3 * t + 0] = 1.0f;
        sn3[3 * t + 1] = 0.0f;
        sn3[3 * t + 2] = 0.0f;
        return;
    }
    const D3 c = dev::d3(g.bmin[0] + g.voxel[0] * (i + 0.5), g.bmin[1] + g.voxel[1] * (j + 0.5),
                         g.bmin[2] + g.voxel[2] * (k + 0.5));
    dev::LaneStack st{stack + threadIdx.x, kFillBlock};
    const dev::SignedNearest s = dev::signedNearest(c, nodes, tri9, dir, st);
    const D3 gr = dev::voxelGradient(c, s);
    sphi[t] = float(s.sd);
    sn3[3 * t + 0] = float(gr.x);
    sn3[3 * t + 1] = float(gr.y);
    sn3[3 * t + 2] = float(gr.z);
}

// Sparse Hessian: missing or out-of-grid neighbours fall back to the voxel's
// own normal (SdfBuilder.cpp:1154-1225); lookups through the brick table.
__global__ void sparseHessianKernel(GridSpec g, int bd, int64_t total, int nbx, int nby, int nbz,
                                    const int32_t* __restrict__ bijk,
                                    const int32_t* __restrict__ table,
                                    const float* __restrict__ sn3, float* __restrict__ sh6)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int bvc = bd * bd * bd;
    const int64_t b = t / bvc;
    const int local = int(t % bvc);
    const int lx = local % bd, ly = (local / bd) % bd, lz = local / (bd * bd);
    const int i = bijk[3 * b] * bd + lx, j = bijk[3 * b + 1] * bd + ly, k = bijk[3 * b + 2] * bd + lz;
    if (i < 0 || i >= g.n[0] || j < 0 || j >= g.n[1] || k < 0 || k >= g.n[2]) return;
    const D3 cur = dev::d3(double(sn3[3 * t]), double(sn3[3 * t + 1]), double(sn3[3 * t + 2]));
    auto nf = [&](int ii, int jj, int kk) {
        if (ii < 0 || ii >= g.n[0] || jj < 0 || jj >= g.n[1] || kk < 0 || kk >= g.n[2]) return cur;
        const int32_t nbk = table[(ii / bd) + int64_t(nbx) * ((jj / bd) + int64_t(nby) * (kk / bd))];
        if (nbk < 0) return cur;
        const int64_t q = int64_t(nbk) * bvc + (ii % bd) + bd * ((jj % bd) + bd * (kk % bd));
        return dev::d3(double(sn3[3 * q]), double(sn3[3 * q + 1]), double(sn3[3 * q + 2]));
    };
    const D3 dx = dev::divs(dev::sub(nf(i + 1, j, k), nf(i - 1, j, k)), 2.0 * g.voxel[0]);
    const D3 dy = dev::divs(dev::sub(nf(i, j + 1, k), nf(i, j - 1, k)), 2.0 * g.voxel[1]);
    const D3 dz = dev::divs(dev::sub(nf(i, j, k + 1), nf(i, j, k - 1)), 2.0 * g.voxel[2]);
    float* o = sh6 + 6 * t;
    o[0] = float(0.5 * (dx.x + dx.x));
    o[1] = float(0.5 * (dy.x + dx.y));
    o[2] = float(0.5 * (dz.x + dx.z));
    o[3] = float(0.5 * (dy.y + dy.y));
    o[4] = float(0.5 * (dz.y + dy.z));
    o[5] = float(0.5 * (dz.z + dz.z));
    (void)nbz;
}

__global__ void sparseMirrorKernel(int64_t count, const float* __restrict__ sphi,
                                   const float* __restrict__ sn3, float4* __restrict__ spn)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    spn[i] = make_float4(sphi[i], sn3[3 * i], sn3[3 * i + 1], sn3[3 * i + 2]);
}

inline unsigned blocksFor(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }

} // namespace

void launchDenseFill(const GridSpec& g, const DeviceBvh& bvh, double* phi, double* n3,
                     double* witness3, int32_t* tri, hipStream_t s)
{
    const int bx = (g.n[0] + 3) / 4, by = (g.n[1] + 3) / 4, bz = (g.n[2] + 3) / 4;
    const int64_t nbr = int64_t(bx) * by * bz;
    const D3 dir = dev::d3(bvh.rayDir[0], bvh.rayDir[1], bvh.rayDir[2]);
    hipLaunchKernelGGL(denseFillKernel, dim3(blocksFor(nbr, kFillBlock / 64)), dim3(kFillBlock), 0,
                       s, g, bvh.nodes, bvh.tri9, bvh.triIdx, dir, bx, by, bz, phi, n3, witness3,
                       tri);
    SDFO_HIP(hipGetLastError());
}

void launchDenseHessian(const GridSpec& g, const double* n3, double* h6, hipStream_t s)
{
    hipLaunchKernelGGL(denseHessianKernel, dim3(blocksFor(g.total(), 256)), dim3(256), 0, s, g,
                       n3, h6);
    SDFO_HIP(hipGetLastError());
}

void launchDenseMirror(int64_t N, const double* phi, const double* n3, const double* h6,
                       float4* pn, float* h6f, hipStream_t s)
{
    hipLaunchKernelGGL(denseMirrorKernel, dim3(blocksFor(N, 256)), dim3(256), 0, s, N, phi, n3,
                       h6, pn, h6f);
    SDFO_HIP(hipGetLastError());
}

void launchCellPack(const GridSpec& g, const float4* pn, float4* cellRec, hipStream_t s)
{
    const int64_t cells = int64_t(g.n[0] - 1) * (g.n[1] - 1) * (g.n[2] - 1);
    if (cells <= 0) return;
    hipLaunchKernelGGL(cellPackKernel, dim3(blocksFor(cells * 8, 256)), dim3(256), 0, s, g, pn,
                       cellRec);
    SDFO_HIP(hipGetLastError());
}

void launchCandidateMirror(int64_t count, const double* cphi, const double* cn3, float4* cpn,
                           hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(candidateMirrorKernel, dim3(blocksFor(count, 256)), dim3(256), 0, s, count,
                       cphi, cn3, cpn);
    SDFO_HIP(hipGetLastError());
}

void launchSparseActivate(const GridSpec& g, const DeviceBvh& bvh, int blockDim, int nbx,
                          int nby, int nbz, double activation, uint8_t* mask, hipStream_t s)
{
    const int64_t nb = int64_t(nbx) * nby * nbz;
    const D3 dir = dev::d3(bvh.rayDir[0], bvh.rayDir[1], bvh.rayDir[2]);
    hipLaunchKernelGGL(sparseActivateKernel, dim3(blocksFor(nb, kFillBlock)), dim3(kFillBlock), 0,
                       s, g, bvh.nodes, bvh.tri9, dir, blockDim, nbx, nby, nbz, activation, mask);
    SDFO_HIP(hipGetLastError());
}

void launchSparseFill(const GridSpec& g, const DeviceBvh& bvh, int blockDim, int nb,
                      const int32_t* bijk, float* sphi, float* sn3, hipStream_t s)
{
    const int64_t total = int64_t(nb) * blockDim * blockDim * blockDim;
    if (total <= 0) return;
    const D3 dir = dev::d3(bvh.rayDir[0], bvh.rayDir[1], bvh.rayDir[2]);
    hipLaunchKernelGGL(sparseFillKernel, dim3(blocksFor(total, kFillBlock)), dim3(kFillBlock), 0,
                       s, g, bvh.nodes, bvh.tri9, dir, blockDim, total, bijk, sphi, sn3);
    SDFO_HIP(hipGetLastError());
}

void launchSparseHessian(const GridSpec& g, int blockDim, int nb, int nbx, int nby, int nbz,
                         const int32_t* bijk, const int32_t* table, const float* sn3, float* sh6,
                         hipStream_t s)
{
    const int64_t total = int64_t(nb) * blockDim * blockDim * blockDim;
    if (total <= 0) return;
    hipLaunchKernelGGL(sparseHessianKernel, dim3(blocksFor(total, 256)), dim3(256), 0, s, g,
                       blockDim, total, nbx, nby, nbz, bijk, table, sn3, sh6);
    SDFO_HIP(hipGetLastError());
}

void launchSparseMirror(int64_t count, const float* sphi, const float* sn3, float4* spn,
                        hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(sparseMirrorKernel, dim3(blocksFor(count, 256)), dim3(256), 0, s, count,
                       sphi, sn3, spn);
    SDFO_HIP(hipGetLastError());
}

} // namespace sdfo
