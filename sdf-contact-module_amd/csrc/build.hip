// SDF build kernels for gfx950.  Compiled with -ffp-contract=off: the FP64
// arithmetic must round exactly like the CPU path it replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "geom_device.hpp"

namespace sdfo {

using dev::D3;

namespace {

constexpr int kFillBlock = 256; // 4 waves
constexpr int kFillWaves = kFillBlock / 64;
constexpr int kReplayBlock = 64; // LDS lane stacks: kStackDepth * 64 * 4 B = 8 KB

// One wave's traversal stacks in the block's LDS.
__device__ __forceinline__ dev::WaveStacks waveStacks(unsigned char* lds)
{
    return dev::carveWaveStacks(lds + (threadIdx.x >> 6) * dev::kWaveStackBytes);
}

// Work items whose nearest triangle the wave search could not certify; a
// second pass re-runs them with the per-lane search.
struct Replay
{
    int* count;
    int64_t* ids;
};

__device__ __forceinline__ void pushReplay(const Replay& r, int64_t id)
{
    r.ids[atomicAdd(r.count, 1)] = id;
}

// Work units whose wave search ran over its step budget, with the bound each
// lane had reached; the frontier pass finishes them one voxel per wave.
struct Heavy
{
    int* count;
    int64_t* ids;  // unit ids
    double* seeds; // 64 per unit
    int* hits;     // 64 per unit: each lane's ray parity count, made by the wave
    unsigned long long* queue;
    int cap; // units the list holds
};

// Step budgets of one wave's search (node visits + triangle tests) the
// tuner chooses from; the mean brick of a 32k-triangle ring takes ~300.
constexpr int kBudgetCount = 3;
constexpr int kBudgets[kBudgetCount] = {1024, 512, 2048};
const char* const kBudgetNames[kBudgetCount] = {"1024", "512", "2048"};
constexpr int kBatchedBudget = 8192;

// Returns false when the list is full: the caller replays the unit's voxels.
__device__ __forceinline__ bool pushHeavy(const Heavy& h, int64_t unit, bool active, double bound,
                                          int hits)
{
    int slot = 0;
    if ((threadIdx.x & 63) == 0) slot = atomicAdd(h.count, 1);
    slot = __shfl(slot, 0);
    if (slot >= h.cap) return false;
    if ((threadIdx.x & 63) == 0) h.ids[slot] = unit;
    h.seeds[int64_t(slot) * 64 + (threadIdx.x & 63)] = active ? bound : DBL_MAX;
    h.hits[int64_t(slot) * 64 + (threadIdx.x & 63)] = hits;
    return true;
}

__device__ __forceinline__ int64_t heavyUnits(const Heavy& h)
{
    const int n = *h.count;
    return int64_t(n < h.cap ? n : h.cap) * 64;
}

// Frontier sizes: the first frontier pass runs 2 waves per block with 640
// entries per level each (12 waves per CU fit its LDS); voxels that outgrow
// it go to a pass with one wave of 4000 entries per block.
constexpr int kFrontierSmall = 640, kFrontierLarge = 4000;

// Voxels whose frontier overflowed, with the bound reached.
struct Spill
{
    int* count;
    int64_t* ids; // voxel ids (dense linear index / sparse pool index)
    double* seeds;
    int* hits; // ray parity counts
    unsigned long long* queue;
};

__device__ __forceinline__ D3 voxelCentre(const GridSpec& g, int i, int j, int k)
{
    // GridSpec.h:30-34
    return dev::d3(g.bmin[0] + g.voxel[0] * (i + 0.5), g.bmin[1] + g.voxel[1] * (j + 0.5),
                   g.bmin[2] + g.voxel[2] * (k + 0.5));
}

struct DenseFillOut
{
    const int32_t* triIdx;
    double *phi, *n3, *wit3;
    int32_t* tri;
};

// Voxel bricks of one wave: 2^sx x 2^sy x 2^(6-sx-sy) voxels, lane l at
// (l & (2^sx-1), (l >> sx) & (2^sy-1), l >> (sx+sy)) inside its brick.
struct BrickShape
{
    int sx, sy;     // log2 of the brick's x and y extent
    int bx, by, bz; // bricks per axis
    static BrickShape make(const GridSpec& g, int sx, int sy)
    {
        BrickShape b;
        b.sx = sx;
        b.sy = sy;
        const int ex = 1 << sx, ey = 1 << sy, ez = 64 >> (sx + sy);
        b.bx = (g.n[0] + ex - 1) / ex;
        b.by = (g.n[1] + ey - 1) / ey;
        b.bz = (g.n[2] + ez - 1) / ez;
        return b;
    }
    __host__ __device__ int64_t count() const { return int64_t(bx) * by * bz; }
    __device__ __forceinline__ void voxel(int64_t brick, int lane, int& i, int& j, int& k) const
    {
        const int bi = int(brick % bx);
        const int bj = int((brick / bx) % by);
        const int bk = int(brick / (int64_t(bx) * by));
        i = (bi << sx) + (lane & ((1 << sx) - 1));
        j = (bj << sy) + ((lane >> sx) & ((1 << sy) - 1));
        k = (bk << (6 - sx - sy)) + (lane >> (sx + sy));
    }
};

// One mesh of a batched dense fill.
struct MeshJob
{
    GridSpec g;
    double mu;
    int root;
    BrickShape sh;
    int64_t brickBegin; // first unit of this mesh in the batch's brick sequence
    int64_t voxelBegin; // first voxel id of this mesh in the batch's replay ids
    DenseFillOut o;
};

// SdfBuilder.cpp:1070-1085: phi, normal, witness and nearest triangle of a voxel.
__device__ __forceinline__ void writeDenseVoxel(const DenseFillOut& o, D3 c,
                                                const dev::SignedNearest& s, int64_t idx)
{
    const D3 gr = dev::voxelGradient(c, s);
    o.phi[idx] = s.sd;
    o.n3[3 * idx + 0] = gr.x;
    o.n3[3 * idx + 1] = gr.y;
    o.n3[3 * idx + 2] = gr.z;
    if (o.wit3)
    {
        o.wit3[3 * idx + 0] = s.cp.x;
        o.wit3[3 * idx + 1] = s.cp.y;
        o.wit3[3 * idx + 2] = s.cp.z;
    }
    if (o.tri) o.tri[idx] = s.slot >= 0 ? o.triIdx[s.slot] : -1;
}

// Next work unit of a wave from the launch's queue counter (wave-uniform).
__device__ __forceinline__ int64_t nextUnit(unsigned long long* queue)
{
    unsigned long long u = 0;
    if ((threadIdx.x & 63) == 0) u = atomicAdd(queue, 1ull);
    return int64_t(dev::uniformMask(__shfl(u, 0)));
}

// Dense fill: one lane per voxel, one 4x4x4 voxel brick per wave; the wave
// walks the BVH cooperatively (SdfBuilder.cpp:1052-1100).  Waves take bricks
// from a queue until it is drained, so clusters of expensive bricks (points
// equidistant from many triangles) spread over the whole chip.
__global__ void __launch_bounds__(kFillBlock) __attribute__((amdgpu_waves_per_eu(5)))
denseFillKernel(GridSpec g, const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                const double* __restrict__ box6, D3 dir, double mu, BrickShape sh, DenseFillOut o,
                Replay replay, Heavy heavy, unsigned long long* queue, int budget, int force)
{
    __shared__ __align__(16) unsigned char lds[kFillWaves * dev::kWaveStackBytes];
    const int lane = threadIdx.x & 63;
    const int64_t nbricks = sh.count();
    const dev::WaveStacks ws = waveStacks(lds);
    for (int64_t brick = nextUnit(queue); brick < nbricks; brick = nextUnit(queue))
    {
        int i, j, k;
        sh.voxel(brick, lane, i, j, k);
        const bool active = i < g.n[0] && j < g.n[1] && k < g.n[2];
        const D3 c = voxelCentre(g, min(i, g.n[0] - 1), min(j, g.n[1] - 1), min(k, g.n[2] - 1));
        const unsigned long long act = __ballot(active);
        const dev::WaveNearest wn =
            dev::waveNearest(c, active, act, nodes, tri9, box6, mu, ws, lane, 0, budget);
        const int64_t idx = i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k);
        const int hits = dev::waveRayHits(c, dir, act, nodes, tri9, ws, lane, 0);
        if (wn.aborted)
        {
            if (!pushHeavy(heavy, brick, active, wn.bestSq, hits) && active) pushReplay(replay, idx);
            continue;
        }
        if (!active) continue;
        if (!wn.exact || force)
            pushReplay(replay, idx);
        else
            writeDenseVoxel(o, c, dev::finishSigned(c, tri9, wn.bestSq, wn.bestSlot, hits), idx);
    }
}

// Voxels of the bricks denseFillKernel handed over: one voxel per wave.
template <int Cap, int Waves>
__global__ void __launch_bounds__(64 * Waves)
denseHeavyKernel(GridSpec g, const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                 const double* __restrict__ box6, D3 dir, double mu, BrickShape sh, DenseFillOut o,
                 Replay replay, Heavy heavy, Spill spill, int force)
{
    __shared__ __align__(16) unsigned char lds[Waves * dev::frontierBytes(Cap)];
    const dev::FrontierLds f = dev::carveFrontier(lds + (threadIdx.x >> 6) * dev::frontierBytes(Cap), Cap);
    const int64_t units = heavyUnits(heavy);
    for (int64_t u = nextUnit(heavy.queue); u < units; u = nextUnit(heavy.queue))
    {
        int i, j, k;
        sh.voxel(heavy.ids[u >> 6], int(u & 63), i, j, k);
        if (i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) continue; // wave-uniform
        const D3 c = voxelCentre(g, i, j, k);
        double seed = heavy.seeds[u];
        const int hits = heavy.hits[u];
        int status = 0;
        const dev::SignedNearest s =
            dev::frontierSignedNearest(c, nodes, tri9, box6, dir, mu, 0, seed, f, status, hits);
        if ((threadIdx.x & 63) != 0) continue;
        const int64_t idx = i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k);
        if (status == dev::kFrontierOverflow)
        {
            const int q = atomicAdd(spill.count, 1);
            spill.ids[q] = idx;
            spill.seeds[q] = seed;
            spill.hits[q] = hits;
        }
        else if (status != dev::kFrontierExact || force)
            pushReplay(replay, idx);
        else
            writeDenseVoxel(o, c, s, idx);
    }
}

// Voxels whose first frontier overflowed: one wave with a 4000-entry frontier.
__global__ void __launch_bounds__(64)
denseSpillKernel(GridSpec g, const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                 const double* __restrict__ box6, D3 dir, double mu, DenseFillOut o, Replay replay,
                 Spill spill, int force)
{
    __shared__ __align__(16) unsigned char lds[dev::frontierBytes(kFrontierLarge)];
    const dev::FrontierLds f = dev::carveFrontier(lds, kFrontierLarge);
    const int64_t n = *spill.count;
    for (int64_t q = nextUnit(spill.queue); q < n; q = nextUnit(spill.queue))
    {
        const int64_t idx = spill.ids[q];
        const int i = int(idx % g.n[0]), j = int((idx / g.n[0]) % g.n[1]),
                  k = int(idx / (int64_t(g.n[0]) * g.n[1]));
        const D3 c = voxelCentre(g, i, j, k);
        double seed = spill.seeds[q];
        const int hits = spill.hits[q];
        int status = 0;
        const dev::SignedNearest s =
            dev::frontierSignedNearest(c, nodes, tri9, box6, dir, mu, 0, seed, f, status, hits);
        if ((threadIdx.x & 63) != 0) continue;
        if (status != dev::kFrontierExact || force)
            pushReplay(replay, idx);
        else
            writeDenseVoxel(o, c, s, idx);
    }
}

// Batched dense fill: the bricks of many meshes (each with its own grid and
// its own BVH inside one concatenated node array) in one queue-driven launch.
__device__ __forceinline__ int findJob(const MeshJob* __restrict__ jobs, int nJobs, int64_t unit,
                                       bool byVoxel)
{
    int lo = 0, hi = nJobs - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t b = byVoxel ? jobs[mid].voxelBegin : jobs[mid].brickBegin;
        if (b <= unit)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(kFillBlock) __attribute__((amdgpu_waves_per_eu(5)))
denseFillBatchedKernel(const MeshJob* __restrict__ jobs, int nJobs, int64_t totalBricks,
                       const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                       const double* __restrict__ box6, D3 dir, Replay replay, Heavy heavy,
                       unsigned long long* queue, int budget, int force)
{
    __shared__ __align__(16) unsigned char lds[kFillWaves * dev::kWaveStackBytes];
    const int lane = threadIdx.x & 63;
    const dev::WaveStacks ws = waveStacks(lds);
    for (int64_t unit = nextUnit(queue); unit < totalBricks; unit = nextUnit(queue))
    {
        const int m = dev::uniformInt(findJob(jobs, nJobs, unit, false));
        const MeshJob& J = jobs[m];
        const GridSpec g = J.g;
        const BrickShape sh = J.sh;
        int i, j, k;
        sh.voxel(unit - J.brickBegin, lane, i, j, k);
        const bool active = i < g.n[0] && j < g.n[1] && k < g.n[2];
        const D3 c = voxelCentre(g, min(i, g.n[0] - 1), min(j, g.n[1] - 1), min(k, g.n[2] - 1));
        const unsigned long long act = __ballot(active);
        const dev::WaveNearest wn =
            dev::waveNearest(c, active, act, nodes, tri9, box6, J.mu, ws, lane, J.root, budget);
        const int64_t idx = i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k);
        const int hits = dev::waveRayHits(c, dir, act, nodes, tri9, ws, lane, J.root);
        if (wn.aborted)
        {
            if (!pushHeavy(heavy, unit, active, wn.bestSq, hits) && active)
                pushReplay(replay, J.voxelBegin + idx);
            continue;
        }
        if (!active) continue;
        if (!wn.exact || force)
            pushReplay(replay, J.voxelBegin + idx);
        else
            writeDenseVoxel(J.o, c, dev::finishSigned(c, tri9, wn.bestSq, wn.bestSlot, hits), idx);
    }
}

// Voxels of the bricks the batched fill handed over: one voxel per wave.
template <int Cap, int Waves>
__global__ void __launch_bounds__(64 * Waves)
denseHeavyBatchedKernel(const MeshJob* __restrict__ jobs, int nJobs,
                        const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                        const double* __restrict__ box6, D3 dir, Replay replay, Heavy heavy,
                        Spill spill, int force)
{
    __shared__ __align__(16) unsigned char lds[Waves * dev::frontierBytes(Cap)];
    const dev::FrontierLds f = dev::carveFrontier(lds + (threadIdx.x >> 6) * dev::frontierBytes(Cap), Cap);
    const int64_t units = heavyUnits(heavy);
    for (int64_t u = nextUnit(heavy.queue); u < units; u = nextUnit(heavy.queue))
    {
        const int64_t brick = heavy.ids[u >> 6];
        const MeshJob& J = jobs[dev::uniformInt(findJob(jobs, nJobs, brick, false))];
        const GridSpec& g = J.g;
        int i, j, k;
        J.sh.voxel(brick - J.brickBegin, int(u & 63), i, j, k);
        if (i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) continue; // wave-uniform
        const D3 c = voxelCentre(g, i, j, k);
        double seed = heavy.seeds[u];
        const int hits = heavy.hits[u];
        int status = 0;
        const dev::SignedNearest s =
            dev::frontierSignedNearest(c, nodes, tri9, box6, dir, J.mu, J.root, seed, f, status, hits);
        if ((threadIdx.x & 63) != 0) continue;
        const int64_t idx = i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k);
        if (status == dev::kFrontierOverflow)
        {
            const int q = atomicAdd(spill.count, 1);
            spill.ids[q] = J.voxelBegin + idx;
            spill.seeds[q] = seed;
            spill.hits[q] = hits;
        }
        else if (status != dev::kFrontierExact || force)
            pushReplay(replay, J.voxelBegin + idx);
        else
            writeDenseVoxel(J.o, c, s, idx);
    }
}

// Batched voxels whose first frontier overflowed: one wave, 4000 entries.
__global__ void __launch_bounds__(64)
denseSpillBatchedKernel(const MeshJob* __restrict__ jobs, int nJobs,
                        const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                        const double* __restrict__ box6, D3 dir, Replay replay, Spill spill,
                        int force)
{
    __shared__ __align__(16) unsigned char lds[dev::frontierBytes(kFrontierLarge)];
    const dev::FrontierLds f = dev::carveFrontier(lds, kFrontierLarge);
    const int64_t n = *spill.count;
    for (int64_t q = nextUnit(spill.queue); q < n; q = nextUnit(spill.queue))
    {
        const int64_t id = spill.ids[q];
        const MeshJob& J = jobs[dev::uniformInt(findJob(jobs, nJobs, id, true))];
        const GridSpec& g = J.g;
        const int64_t idx = id - J.voxelBegin;
        const int i = int(idx % g.n[0]), j = int((idx / g.n[0]) % g.n[1]),
                  k = int(idx / (int64_t(g.n[0]) * g.n[1]));
        const D3 c = voxelCentre(g, i, j, k);
        double seed = spill.seeds[q];
        const int hits = spill.hits[q];
        int status = 0;
        const dev::SignedNearest s =
            dev::frontierSignedNearest(c, nodes, tri9, box6, dir, J.mu, J.root, seed, f, status, hits);
        if ((threadIdx.x & 63) != 0) continue;
        if (status != dev::kFrontierExact || force)
            pushReplay(replay, id);
        else
            writeDenseVoxel(J.o, c, s, idx);
    }
}

__global__ void __launch_bounds__(kReplayBlock)
denseReplayBatchedKernel(const MeshJob* __restrict__ jobs, int nJobs,
                         const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                         D3 dir, Replay replay)
{
    __shared__ int stack[kStackDepth * kReplayBlock];
    const int n = *replay.count;
    for (int q = int(blockIdx.x) * kReplayBlock + int(threadIdx.x); q < n;
         q += int(gridDim.x) * kReplayBlock)
    {
        const int64_t id = replay.ids[q];
        const MeshJob& J = jobs[findJob(jobs, nJobs, id, true)];
        const int64_t idx = id - J.voxelBegin;
        const GridSpec& g = J.g;
        const int i = int(idx % g.n[0]), j = int((idx / g.n[0]) % g.n[1]),
                  k = int(idx / (int64_t(g.n[0]) * g.n[1]));
        const D3 c = voxelCentre(g, i, j, k);
        dev::LaneStack st{stack + threadIdx.x, kReplayBlock};
        writeDenseVoxel(J.o, c, dev::signedNearest(c, nodes, tri9, dir, st, J.root), idx);
    }
}

// Per-lane search for the voxels denseFillKernel listed.
__global__ void __launch_bounds__(kReplayBlock)
denseReplayKernel(GridSpec g, const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                  D3 dir, Replay replay, DenseFillOut o)
{
    __shared__ int stack[kStackDepth * kReplayBlock];
    const int n = *replay.count;
    for (int q = int(blockIdx.x) * kReplayBlock + int(threadIdx.x); q < n;
         q += int(gridDim.x) * kReplayBlock)
    {
        const int64_t idx = replay.ids[q];
        const int i = int(idx % g.n[0]), j = int((idx / g.n[0]) % g.n[1]),
                  k = int(idx / (int64_t(g.n[0]) * g.n[1]));
        const D3 c = voxelCentre(g, i, j, k);
        dev::LaneStack st{stack + threadIdx.x, kReplayBlock};
        writeDenseVoxel(o, c, dev::signedNearest(c, nodes, tri9, dir, st), idx);
    }
}

// Central differences of the normal field, clamped at the border, then
// symmetrised (SdfBuilder.cpp:1105-1152).  With pn set it also writes the FP32
// query mirror of the voxel (what denseMirrorKernel writes), saving a pass.
__global__ void denseHessianKernel(GridSpec g, const double* __restrict__ n3,
                                   double* __restrict__ h6, const double* __restrict__ phi,
                                   float4* __restrict__ pn, float* __restrict__ h6f)
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
    const double H[6] = {0.5 * (dx.x + dx.x), 0.5 * (dy.x + dx.y), 0.5 * (dz.x + dx.z),
                         0.5 * (dy.y + dy.y), 0.5 * (dz.y + dy.z), 0.5 * (dz.z + dz.z)};
    double* o = h6 + 6 * idx;
#pragma unroll
    for (int c = 0; c < 6; ++c) o[c] = H[c];
    if (pn)
    {
        pn[idx] = make_float4(float(phi[idx]), float(n3[3 * idx]), float(n3[3 * idx + 1]),
                              float(n3[3 * idx + 2]));
#pragma unroll
        for (int c = 0; c < 6; ++c) h6f[6 * idx + c] = float(H[c]);
    }
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

// FP64 cell records: the 8 corners (phi, n) of each cell from the FP64 fields,
// one thread per (cell, corner), in the cell order of cellPackKernel.
__global__ void cellPack64Kernel(GridSpec g, const double* __restrict__ phi,
                                 const double* __restrict__ n3, double4* __restrict__ rec)
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
    const int64_t v = i + int64_t(g.n[0]) * (j + int64_t(g.n[1]) * k);
    rec[t] = make_double4(phi[v], n3[3 * v], n3[3 * v + 1], n3[3 * v + 2]);
}

// Sparse cell records: for the cell whose lowest corner is brick-pool voxel
// t, the 8 corner records (located through the brick table across brick
// faces) and one validity bit (all 8 corners exist, the rule of
// SparseSdfData::sample, SdfData.h:216-247).  One thread per cell; the
// bits of a wave's 64 cells form one mask word when a brick holds a multiple
// of 64 voxels, else they are or-ed in.
__global__ void __launch_bounds__(256)
sparseCellPackKernel(GridSpec g, int bd, int nbx, int nby, int nbz,
                     const int32_t* __restrict__ bijk, const int32_t* __restrict__ table,
                     const float4* __restrict__ spn, int64_t count, int words,
                     float4* __restrict__ rec, unsigned long long* __restrict__ mask)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if ((t & ~int64_t(63)) >= count) return; // whole waves
    const int bvc = bd * bd * bd;
    bool ok = false;
    int64_t b = 0;
    int local = 0;
    if (t < count)
    {
        b = t / bvc;
        local = int(t % bvc);
        const int i0 = bijk[3 * b] * bd + local % bd;
        const int j0 = bijk[3 * b + 1] * bd + (local / bd) % bd;
        const int k0 = bijk[3 * b + 2] * bd + local / (bd * bd);
        int64_t r[8];
        ok = true;
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            const int i = i0 + (q & 1), j = j0 + ((q >> 1) & 1), k = k0 + (q >> 2);
            r[q] = -1;
            if (i >= 0 && i < g.n[0] && j >= 0 && j < g.n[1] && k >= 0 && k < g.n[2])
            {
                const int bi = i / bd, bj = j / bd, bk = k / bd;
                const int32_t nb = table[bi + int64_t(nbx) * (bj + int64_t(nby) * bk)];
                if (nb >= 0)
                    r[q] = int64_t(nb) * bvc + (i - bi * bd) + bd * ((j - bj * bd) + bd * (k - bk * bd));
            }
            ok = ok && r[q] >= 0;
        }
        if (ok)
        {
#pragma unroll
            for (int q = 0; q < 8; ++q) rec[8 * t + q] = spn[r[q]];
        }
    }
    (void)nbz;
    if (bvc % 64 == 0)
    {
        const unsigned long long m = __ballot(ok);
        if ((threadIdx.x & 63) == 0 && t < count) mask[t >> 6] = m;
    }
    else if (ok)
        atomicOr(&mask[b * words + local / 64], 1ull << (local % 64));
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
// (SdfBuilder.cpp:869-890).  One lane per brick cell, 64 consecutive cells per wave.
__device__ __forceinline__ D3 brickCentre(const GridSpec& g, int bd, int nbx, int nby, int64_t bl)
{
    const int bi = int(bl % nbx), bj = int((bl / nbx) % nby), bk = int(bl / (int64_t(nbx) * nby));
    const int i0 = bi * bd, j0 = bj * bd, k0 = bk * bd;
    const int i1 = min(i0 + bd, g.n[0]), j1 = min(j0 + bd, g.n[1]), k1 = min(k0 + bd, g.n[2]);
    return dev::d3(g.bmin[0] + g.voxel[0] * (0.5 * double(i0 + i1)),
                   g.bmin[1] + g.voxel[1] * (0.5 * double(j0 + j1)),
                   g.bmin[2] + g.voxel[2] * (0.5 * double(k0 + k1)));
}

// Whether some triangle lies within distance a of p, for each lane of a wave
// (every lane calls it).  The brick test |findSignedNearest(c)| <= a
// (SdfBuilder.cpp:887-889) compares the nearest distance sqrt(bestSq), the
// smallest sqrt(sq) over the triangles, so it holds exactly when one
// triangle's sqrt(sq) <= a: neither the sign (ray parity) nor which of several
// nearest triangles wins matters.  One shared stack (entry e in lane e), a
// node visited while some lane without a hit may reach it; a lane prunes boxes
// farther than a + mu (mu is far above the rounding of box distances).
__device__ bool waveWithin(D3 p, bool active, double a, double mu,
                           const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                           const double* __restrict__ box6)
{
    const int lane = threadIdx.x & 63;
    const double reach = a + mu, reach2 = reach * reach;
    bool found = false;
    int stk = 0;
    int sp = 1;
    while (sp > 0 && __ballot(active && !found) != 0ull)
    {
        --sp;
        const int ni = __builtin_amdgcn_readlane(stk, sp);
        const BvhNode nd = nodes[ni];
        const bool need = active && !found && dev::boxSq(p, nd) <= reach2;
        if (__ballot(need) == 0ull) continue;
        if (nd.triCount > 0)
        {
            for (int q = nd.triStart; q < nd.triStart + nd.triCount; ++q)
            {
                const bool mem = need && !found && dev::boxSq6(p, box6 + 6 * size_t(q)) <= reach2;
                if (__ballot(mem) == 0ull) continue;
                const dev::Tri tr = dev::loadTri(tri9, q);
                if (mem && sqrt(dev::sqnorm(dev::sub(p, dev::closestPoint(p, tr.a, tr.b, tr.c)))) <= a)
                    found = true;
            }
            continue;
        }
        // the child most lanes find nearer is popped first
        const bool rf = need && dev::boxSq(p, nodes[nd.right]) < dev::boxSq(p, nodes[nd.left]);
        const bool rightFirst = 2 * __popcll(__ballot(rf)) > __popcll(__ballot(need));
        if (lane == sp) stk = rightFirst ? nd.left : nd.right;
        if (lane == sp + 1) stk = rightFirst ? nd.right : nd.left;
        sp += 2;
    }
    return found;
}

__global__ void __launch_bounds__(kFillBlock)
sparseActivateKernel(GridSpec g, const BvhNode* __restrict__ nodes,
                     const double* __restrict__ tri9, const double* __restrict__ box6, double mu,
                     int bd, int nbx, int nby, int nbz, double activation,
                     uint8_t* __restrict__ mask, Replay replay, int force)
{
    const int64_t ncell = int64_t(nbx) * nby * nbz;
    const int64_t bl0 = int64_t(blockIdx.x) * kFillBlock + threadIdx.x;
    if ((bl0 & ~int64_t(63)) >= ncell) return; // whole waves only
    const bool active = bl0 < ncell;
    const int64_t bl = active ? bl0 : ncell - 1;
    const D3 c = brickCentre(g, bd, nbx, nby, bl);
    if (force)
    {
        if (active) pushReplay(replay, bl);
        return;
    }
    const bool hit = waveWithin(c, active, activation, mu, nodes, tri9, box6);
    if (active) mask[bl] = hit ? 1 : 0;
}

__global__ void __launch_bounds__(kReplayBlock)
sparseActivateReplayKernel(GridSpec g, const BvhNode* __restrict__ nodes,
                           const double* __restrict__ tri9, D3 dir, int bd, int nbx, int nby,
                           double activation, Replay replay, uint8_t* __restrict__ mask)
{
    __shared__ int stack[kStackDepth * kReplayBlock];
    const int n = *replay.count;
    for (int q = int(blockIdx.x) * kReplayBlock + int(threadIdx.x); q < n;
         q += int(gridDim.x) * kReplayBlock)
    {
        const int64_t bl = replay.ids[q];
        const D3 c = brickCentre(g, bd, nbx, nby, bl);
        dev::LaneStack st{stack + threadIdx.x, kReplayBlock};
        mask[bl] = fabs(dev::signedNearest(c, nodes, tri9, dir, st).sd) <= activation ? 1 : 0;
    }
}

__device__ __forceinline__ void brickVoxel(const int32_t* __restrict__ bijk, int bd, int64_t t,
                                           int& i, int& j, int& k)
{
    const int bvc = bd * bd * bd;
    const int64_t b = t / bvc;
    const int local = int(t % bvc);
    i = bijk[3 * b] * bd + local % bd;
    j = bijk[3 * b + 1] * bd + (local / bd) % bd;
    k = bijk[3 * b + 2] * bd + local / (bd * bd);
}

__device__ __forceinline__ void writeSparseVoxel(D3 c, const dev::SignedNearest& s, int64_t t,
                                                 float* __restrict__ sphi, float* __restrict__ sn3)
{
    const D3 gr = dev::voxelGradient(c, s);
    sphi[t] = float(s.sd);
    sn3[3 * t + 0] = float(gr.x);
    sn3[3 * t + 1] = float(gr.y);
    sn3[3 * t + 2] = float(gr.z);
}

// Sparse fill: one lane per brick voxel, brick-major (SdfBuilder.cpp:925-965);
// 64 brick voxels per queue unit, laid out by brickUnitVoxel.
__global__ void __launch_bounds__(kFillBlock) __attribute__((amdgpu_waves_per_eu(5)))
sparseFillKernel(GridSpec g, const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                 const double* __restrict__ box6, D3 dir, double mu, int bd, int64_t total,
                 const int32_t* __restrict__ bijk, float* __restrict__ sphi,
                 float* __restrict__ sn3, Replay replay, Heavy heavy, unsigned long long* queue,
                 int budget, int force)
{
    __shared__ __align__(16) unsigned char lds[kFillWaves * dev::kWaveStackBytes];
    const int64_t nunits = (total + 63) / 64;
    const int lane = threadIdx.x & 63;
    const dev::WaveStacks ws = waveStacks(lds);
    for (int64_t u = nextUnit(queue); u < nunits; u = nextUnit(queue))
    {
        const int64_t t0 = dev::brickUnitVoxel(bd, u, lane);
        const int64_t t = t0 < total ? t0 : total - 1;
        int i, j, k;
        brickVoxel(bijk, bd, t, i, j, k);
        const bool inGrid = i >= 0 && i < g.n[0] && j >= 0 && j < g.n[1] && k >= 0 && k < g.n[2];
        const bool active = t0 < total && inGrid;
        if (t0 < total && !inGrid)
        {
            sphi[t] = 0.0f;
            sn3[3 * t + 0] = 1.0f;
            sn3[3 * t + 1] = 0.0f;
            sn3[3 * t + 2] = 0.0f;
        }
        const D3 c = voxelCentre(g, min(max(i, 0), g.n[0] - 1), min(max(j, 0), g.n[1] - 1),
                                 min(max(k, 0), g.n[2] - 1));
        const unsigned long long act = __ballot(active);
        const dev::WaveNearest wn =
            dev::waveNearest(c, active, act, nodes, tri9, box6, mu, ws, lane, 0, budget);
        const int hits = dev::waveRayHits(c, dir, act, nodes, tri9, ws, lane, 0);
        if (wn.aborted)
        {
            if (!pushHeavy(heavy, u, active, wn.bestSq, hits) && active) pushReplay(replay, t);
            continue;
        }
        if (!active) continue;
        if (!wn.exact || force)
            pushReplay(replay, t);
        else
            writeSparseVoxel(c, dev::finishSigned(c, tri9, wn.bestSq, wn.bestSlot, hits), t, sphi, sn3);
    }
}

template <int Cap, int Waves>
__global__ void __launch_bounds__(64 * Waves)
sparseHeavyKernel(GridSpec g, const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                  const double* __restrict__ box6, D3 dir, double mu, int bd, int64_t total,
                  const int32_t* __restrict__ bijk, float* __restrict__ sphi,
                  float* __restrict__ sn3, Replay replay, Heavy heavy, Spill spill, int force)
{
    __shared__ __align__(16) unsigned char lds[Waves * dev::frontierBytes(Cap)];
    const dev::FrontierLds f = dev::carveFrontier(lds + (threadIdx.x >> 6) * dev::frontierBytes(Cap), Cap);
    const int64_t units = heavyUnits(heavy);
    for (int64_t u = nextUnit(heavy.queue); u < units; u = nextUnit(heavy.queue))
    {
        const int64_t t = dev::brickUnitVoxel(bd, heavy.ids[u >> 6], int(u & 63));
        if (t >= total) continue; // wave-uniform
        int i, j, k;
        brickVoxel(bijk, bd, t, i, j, k);
        if (i < 0 || i >= g.n[0] || j < 0 || j >= g.n[1] || k < 0 || k >= g.n[2]) continue;
        const D3 c = voxelCentre(g, i, j, k);
        double seed = heavy.seeds[u];
        const int hits = heavy.hits[u];
        int status = 0;
        const dev::SignedNearest s =
            dev::frontierSignedNearest(c, nodes, tri9, box6, dir, mu, 0, seed, f, status, hits);
        if ((threadIdx.x & 63) != 0) continue;
        if (status == dev::kFrontierOverflow)
        {
            const int q = atomicAdd(spill.count, 1);
            spill.ids[q] = t;
            spill.seeds[q] = seed;
            spill.hits[q] = hits;
        }
        else if (status != dev::kFrontierExact || force)
            pushReplay(replay, t);
        else
            writeSparseVoxel(c, s, t, sphi, sn3);
    }
}

__global__ void __launch_bounds__(64)
sparseSpillKernel(GridSpec g, const BvhNode* __restrict__ nodes, const double* __restrict__ tri9,
                  const double* __restrict__ box6, D3 dir, double mu, int bd,
                  const int32_t* __restrict__ bijk, float* __restrict__ sphi,
                  float* __restrict__ sn3, Replay replay, Spill spill, int force)
{
    __shared__ __align__(16) unsigned char lds[dev::frontierBytes(kFrontierLarge)];
    const dev::FrontierLds f = dev::carveFrontier(lds, kFrontierLarge);
    const int64_t n = *spill.count;
    for (int64_t q = nextUnit(spill.queue); q < n; q = nextUnit(spill.queue))
    {
        const int64_t t = spill.ids[q];
        int i, j, k;
        brickVoxel(bijk, bd, t, i, j, k);
        const D3 c = voxelCentre(g, i, j, k);
        double seed = spill.seeds[q];
        const int hits = spill.hits[q];
        int status = 0;
        const dev::SignedNearest s =
            dev::frontierSignedNearest(c, nodes, tri9, box6, dir, mu, 0, seed, f, status, hits);
        if ((threadIdx.x & 63) != 0) continue;
        if (status != dev::kFrontierExact || force)
            pushReplay(replay, t);
        else
            writeSparseVoxel(c, s, t, sphi, sn3);
    }
}

__global__ void __launch_bounds__(kReplayBlock)
sparseFillReplayKernel(GridSpec g, const BvhNode* __restrict__ nodes,
                       const double* __restrict__ tri9, D3 dir, int bd,
                       const int32_t* __restrict__ bijk, Replay replay, float* __restrict__ sphi,
                       float* __restrict__ sn3)
{
    __shared__ int stack[kStackDepth * kReplayBlock];
    const int n = *replay.count;
    for (int q = int(blockIdx.x) * kReplayBlock + int(threadIdx.x); q < n;
         q += int(gridDim.x) * kReplayBlock)
    {
        const int64_t t = replay.ids[q];
        int i, j, k;
        brickVoxel(bijk, bd, t, i, j, k);
        const D3 c = voxelCentre(g, i, j, k);
        dev::LaneStack st{stack + threadIdx.x, kReplayBlock};
        writeSparseVoxel(c, dev::signedNearest(c, nodes, tri9, dir, st), t, sphi, sn3);
    }
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

// Pruning margin of the wave-cooperative nearest search: 2^-30 times the
// largest coordinate magnitude of the grid box (which contains the mesh).
double tieMargin(const GridSpec& g)
{
    double L = 0.0;
    for (int a = 0; a < 3; ++a) L = std::max(L, std::max(std::fabs(g.bmin[a]), std::fabs(g.bmax[a])));
    return std::ldexp(std::max(L, DBL_MIN), -30);
}

namespace {

// Stream-ordered replay list sized for every work item of a launch.
struct ReplayBuf
{
    Replay r{nullptr, nullptr};
    Heavy h{nullptr, nullptr, nullptr, nullptr, nullptr, 0};
    Spill sp{nullptr, nullptr, nullptr, nullptr, nullptr};
    unsigned long long* queue = nullptr; // work-queue counter of the fill launch
    hipStream_t s;
    void* owned = nullptr; // the block, when this object allocated it
    // Bytes of the one block a ReplayBuf carves: the 48-byte counter header,
    // then the replay ids (cap items) and, for 64-lane work units, the heavy
    // list, its seeds and the spill list (units * 64 voxels).
    static constexpr size_t kHeaderBytes = 48; // the six queue counters and list lengths
    static size_t bytes(int64_t cap, int64_t units)
    {
        auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
        size_t b = up(48) + up(sizeof(int64_t) * size_t(std::max<int64_t>(cap, 1)));
        if (units > 0)
            b += up(sizeof(int64_t) * size_t(units)) + 3 * up(sizeof(double) * 64 * size_t(units)) +
                 2 * up(sizeof(int) * 64 * size_t(units));
        return b;
    }
    // cap: work items (replay ids); units: 64-lane work units (heavy list).
    // The lists come from `mem` (at least bytes(cap, units), 256-byte aligned,
    // not in use by other work on the stream) or from one allocation.
    // zeroed: the caller cleared the counter header of `mem` on this stream.
    ReplayBuf(int64_t cap, hipStream_t st, int64_t units = 0, void* mem = nullptr,
              bool zeroed = false)
        : s(st)
    {
        if (!mem)
        {
            SDFO_HIP(hipMallocAsync(&owned, bytes(cap, units), s));
            mem = owned;
        }
        auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
        char* p = static_cast<char*>(mem);
        queue = reinterpret_cast<unsigned long long*>(p);
        if (!(zeroed && mem)) SDFO_HIP(hipMemsetAsync(queue, 0, kHeaderBytes, s));
        r.count = reinterpret_cast<int*>(queue + 1);
        h.queue = queue + 2;
        h.count = reinterpret_cast<int*>(queue + 3);
        sp.queue = queue + 4;
        sp.count = reinterpret_cast<int*>(queue + 5);
        p += up(48);
        r.ids = reinterpret_cast<int64_t*>(p);
        p += up(sizeof(int64_t) * size_t(std::max<int64_t>(cap, 1)));
        h.cap = int(std::min<int64_t>(units, INT32_MAX));
        if (units > 0)
        {
            const size_t lane = up(sizeof(double) * 64 * size_t(units));
            h.ids = reinterpret_cast<int64_t*>(p);
            p += up(sizeof(int64_t) * size_t(units));
            h.seeds = reinterpret_cast<double*>(p);
            sp.ids = reinterpret_cast<int64_t*>(p + lane);
            sp.seeds = reinterpret_cast<double*>(p + 2 * lane);
            p += 3 * lane;
            h.hits = reinterpret_cast<int*>(p);
            sp.hits = reinterpret_cast<int*>(p + up(sizeof(int) * 64 * size_t(units)));
        }
    }
    // Copies the list length to the calling thread's pinned counter (read it
    // after synchronising the stream).
    void report()
    {
        int* hc = replayCounterHost();
        if (hc) SDFO_HIP(hipMemcpyAsync(hc, r.count, sizeof(int), hipMemcpyDeviceToHost, s));
    }
    ~ReplayBuf()
    {
        if (owned) (void)hipFreeAsync(owned, s);
    }
};

constexpr int kReplayGrid = 512; // grid-stride blocks of the replay kernels

// Brick shapes the tuner chooses from: (log2 x, log2 y) extents.
constexpr int kFillShapeCount = 4;
constexpr int kFillShapes[kFillShapeCount][2] = {{2, 2}, {3, 3}, {3, 2}, {4, 2}};
const char* const kFillShapeNames[kFillShapeCount] = {"4x4x4", "8x8x1", "8x4x2", "16x4x1"};

// Blocks of a queue-driven fill launch: enough resident waves for the chip,
// never more than there are work units.
int deviceCUs()
{
    static int cus = 0;
    if (cus == 0)
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        if (cus <= 0) cus = 256;
    }
    return cus;
}

unsigned queueGrid(int64_t units)
{
    const int64_t want = int64_t(deviceCUs()) * 8; // 8 blocks of 4 waves per CU
    const int64_t need = (units + kFillWaves - 1) / kFillWaves;
    return unsigned(std::max<int64_t>(1, std::min(want, need)));
}

// Blocks of the frontier (heavy) pass: its unit count is only known on the
// device, so a fixed grid of queue-driven waves.
unsigned heavyGrid() { return queueGrid(int64_t(1) << 40); }

// Blocks of the overflow pass (one wave each): a block per CU, since
// overflowing voxels are few and each block holds a 32 KB frontier.
unsigned spillGrid() { return unsigned(deviceCUs()); }

} // namespace

int* replayCounterHost()
{
    thread_local int* h = nullptr;
    if (!h)
    {
        if (hipHostMalloc(reinterpret_cast<void**>(&h), sizeof(int), hipHostMallocDefault) != hipSuccess)
            h = nullptr;
        else
            *h = 0;
    }
    return h;
}

bool& forceReplay()
{
    thread_local bool f = false;
    return f;
}

int& debugFillBudget()
{
    thread_local int b = 0;
    return b;
}

int& debugHeavyCap()
{
    thread_local int c = 0;
    return c;
}

namespace {

// Fill variants the tuner chooses from: brick shape x step budget.
constexpr int kFillVariants = kFillShapeCount * kBudgetCount;

std::vector<std::string> fillVariantNames()
{
    std::vector<std::string> n;
    for (int v = 0; v < kFillVariants; ++v)
        n.push_back(std::string(kFillShapeNames[v / kBudgetCount]) + "/budget" + kBudgetNames[v % kBudgetCount]);
    return n;
}

void runDenseFill(const GridSpec& g, const DeviceBvh& bvh, const DenseFillOut& o, int variant,
                  bool report, hipStream_t s, void* scratch, bool zeroed = false)
{
    const int shape = variant / kBudgetCount;
    const int budget = debugFillBudget() > 0 ? debugFillBudget() : kBudgets[variant % kBudgetCount];
    const BrickShape sh = BrickShape::make(g, kFillShapes[shape][0], kFillShapes[shape][1]);
    const D3 dir = dev::d3(bvh.rayDir[0], bvh.rayDir[1], bvh.rayDir[2]);
    ReplayBuf rb(g.total(), s, sh.count(), scratch, zeroed);
    const double mu = tieMargin(g);
    hipLaunchKernelGGL(denseFillKernel, dim3(queueGrid(sh.count())), dim3(kFillBlock), 0, s, g,
                       bvh.nodes, bvh.tri9, bvh.box6, dir, mu, sh, o, rb.r, rb.h, rb.queue, budget,
                       forceReplay() ? 1 : 0);
    SDFO_HIP(hipGetLastError());
    hipLaunchKernelGGL((denseHeavyKernel<kFrontierSmall, 2>), dim3(heavyGrid()), dim3(128), 0, s, g,
                       bvh.nodes, bvh.tri9, bvh.box6, dir, mu, sh, o, rb.r, rb.h, rb.sp,
                       forceReplay() ? 1 : 0);
    SDFO_HIP(hipGetLastError());
    hipLaunchKernelGGL(denseSpillKernel, dim3(spillGrid()), dim3(64), 0, s, g, bvh.nodes, bvh.tri9,
                       bvh.box6, dir, mu, o, rb.r, rb.sp, forceReplay() ? 1 : 0);
    SDFO_HIP(hipGetLastError());
    hipLaunchKernelGGL(denseReplayKernel, dim3(kReplayGrid), dim3(kReplayBlock), 0, s, g, bvh.nodes,
                       bvh.tri9, dir, rb.r, o);
    SDFO_HIP(hipGetLastError());
    if (report) rb.report();
}

std::string fillKey(const GridSpec& g, int64_t ntris)
{
    return "dense_fill tris~2^" + std::to_string(log2Bucket(ntris)) + " voxels~2^" +
           std::to_string(log2Bucket(g.total()));
}

} // namespace

// Fill variant (brick shape x step budget) for a size class: the tuner's
// choice, else 4 x 4 x 4 bricks with a budget of 1024 steps.
int fillShapeFor(const GridSpec& g, int64_t ntris)
{
    int v = 0;
    if (tunerLookup(fillKey(g, ntris), v) && v >= 0 && v < kFillVariants) return v;
    return 0;
}

size_t denseFillScratchBytes(const GridSpec& g)
{
    size_t b = 0;
    for (int q = 0; q < kFillShapeCount; ++q)
        b = std::max(b, ReplayBuf::bytes(g.total(),
                                         BrickShape::make(g, kFillShapes[q][0], kFillShapes[q][1]).count()));
    return b;
}

size_t denseFillScratchHeaderBytes() { return ReplayBuf::kHeaderBytes; }

void launchDenseFill(const GridSpec& g, const DeviceBvh& bvh, double* phi, double* n3,
                     double* witness3, int32_t* tri, hipStream_t s, void* scratch,
                     bool scratchZeroed)
{
    const DenseFillOut o{bvh.triIdx, phi, n3, witness3, tri};
    const std::string key = fillKey(g, bvh.ntris);
    int shape = 0;
    if (!tunerLookup(key, shape) && tunerEnabled() && !streamCapturing(s))
    {
        scratchZeroed = false; // the timed runs leave their counters behind
        // time every brick shape on this build (each run writes the same,
        // exact fields) and keep the fastest for the size class
        const std::vector<std::pair<int, double>> t = timeVariants(
            s, kFillVariants, [&](int v) { runDenseFill(g, bvh, o, v, false, s, scratch); });
        const std::vector<std::string> names = fillVariantNames();
        std::vector<const char*> cn;
        for (const std::string& n : names) cn.push_back(n.c_str());
        tunerRecord(key, t, cn.data());
        tunerLookup(key, shape);
    }
    if (shape < 0 || shape >= kFillVariants) shape = 0;
    runDenseFill(g, bvh, o, shape, true, s, scratch, scratchZeroed);
}

void launchDenseFillBatched(const std::vector<BatchMesh>& meshes, const DeviceBvh& bvh,
                            hipStream_t s)
{
    std::vector<MeshJob> jobs(meshes.size());
    int64_t bricks = 0, voxels = 0;
    for (size_t m = 0; m < meshes.size(); ++m)
    {
        const BatchMesh& b = meshes[m];
        MeshJob& J = jobs[m];
        J.g = b.g;
        J.mu = tieMargin(b.g);
        J.root = b.root;
        const int shape = fillShapeFor(b.g, b.ntris) / kBudgetCount;
        J.sh = BrickShape::make(b.g, kFillShapes[shape][0], kFillShapes[shape][1]);
        J.brickBegin = bricks;
        J.voxelBegin = voxels;
        J.o = DenseFillOut{bvh.triIdx, b.phi, b.n3, b.witness3, b.tri};
        bricks += J.sh.count();
        voxels += b.g.total();
    }
    if (jobs.empty() || bricks == 0) return;
    MeshJob* dj = nullptr;
    SDFO_HIP(hipMallocAsync(reinterpret_cast<void**>(&dj), sizeof(MeshJob) * jobs.size(), s));
    SDFO_HIP(hipMemcpyAsync(dj, jobs.data(), sizeof(MeshJob) * jobs.size(), hipMemcpyHostToDevice, s));
    const D3 dir = dev::d3(bvh.rayDir[0], bvh.rayDir[1], bvh.rayDir[2]);
    {
        // bricks over the step budget go to the frontier pass; a list of at
        // most 2^18 bricks, the rest (if any) to the per-lane replay
        const int64_t cap = debugHeavyCap() > 0 ? debugHeavyCap() : int64_t(1) << 18;
        ReplayBuf rb(voxels, s, std::min<int64_t>(bricks, cap));
        const int force = forceReplay() ? 1 : 0;
        // a batch has enough bricks to hide ordinary slow ones: only the
        // extreme ones (points equidistant from thousands of triangles) go to
        // the frontier pass
        const int budget = debugFillBudget() > 0 ? debugFillBudget() : kBatchedBudget;
        hipLaunchKernelGGL(denseFillBatchedKernel, dim3(queueGrid(bricks)), dim3(kFillBlock), 0, s,
                           dj, int(jobs.size()), bricks, bvh.nodes, bvh.tri9, bvh.box6, dir, rb.r,
                           rb.h, rb.queue, budget, force);
        SDFO_HIP(hipGetLastError());
        hipLaunchKernelGGL((denseHeavyBatchedKernel<kFrontierSmall, 2>), dim3(heavyGrid()), dim3(128),
                           0, s, dj, int(jobs.size()), bvh.nodes, bvh.tri9, bvh.box6, dir, rb.r, rb.h,
                           rb.sp, force);
        SDFO_HIP(hipGetLastError());
        hipLaunchKernelGGL(denseSpillBatchedKernel, dim3(spillGrid()), dim3(64), 0, s, dj,
                           int(jobs.size()), bvh.nodes, bvh.tri9, bvh.box6, dir, rb.r, rb.sp, force);
        SDFO_HIP(hipGetLastError());
        hipLaunchKernelGGL(denseReplayBatchedKernel, dim3(kReplayGrid), dim3(kReplayBlock), 0, s, dj,
                           int(jobs.size()), bvh.nodes, bvh.tri9, dir, rb.r);
        SDFO_HIP(hipGetLastError());
        rb.report();
    }
    // the job table is read by both kernels; the host copy must live until the
    // stream has consumed it
    SDFO_HIP(hipStreamSynchronize(s));
    SDFO_HIP(hipFreeAsync(dj, s));
}

void launchDenseHessian(const GridSpec& g, const double* n3, double* h6, hipStream_t s,
                        const double* phi, float4* pn, float* h6f)
{
    hipLaunchKernelGGL(denseHessianKernel, dim3(blocksFor(g.total(), 256)), dim3(256), 0, s, g,
                       n3, h6, phi, pn, h6f);
    SDFO_HIP(hipGetLastError());
}

void launchDenseMirror(int64_t N, const double* phi, const double* n3, const double* h6,
                       float4* pn, float* h6f, hipStream_t s)
{
    hipLaunchKernelGGL(denseMirrorKernel, dim3(blocksFor(N, 256)), dim3(256), 0, s, N, phi, n3,
                       h6, pn, h6f);
    SDFO_HIP(hipGetLastError());
}

void launchCellPack64(const GridSpec& g, const double* phi, const double* n3, double4* rec,
                      hipStream_t s)
{
    const int64_t items = int64_t(g.n[0] - 1) * (g.n[1] - 1) * (g.n[2] - 1) * 8;
    if (items <= 0) return;
    hipLaunchKernelGGL(cellPack64Kernel, dim3(unsigned((items + 255) / 256)), dim3(256), 0, s, g, phi,
                       n3, rec);
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

void launchSparseCellPack(const GridSpec& g, int bd, int nbx, int nby, int nbz,
                          const int32_t* bijk, const int32_t* table, const float4* spn,
                          int64_t count, int words, float4* rec, unsigned long long* mask,
                          hipStream_t s)
{
    if (count <= 0) return;
    if ((bd * bd * bd) % 64 != 0)
        SDFO_HIP(hipMemsetAsync(mask, 0, sizeof(unsigned long long) * size_t(count / (bd * bd * bd)) * words, s));
    hipLaunchKernelGGL(sparseCellPackKernel, dim3(blocksFor(count, 256)), dim3(256), 0, s, g, bd, nbx,
                       nby, nbz, bijk, table, spn, count, words, rec, mask);
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
    ReplayBuf rb(nb, s);
    hipLaunchKernelGGL(sparseActivateKernel, dim3(blocksFor(nb, kFillBlock)), dim3(kFillBlock), 0,
                       s, g, bvh.nodes, bvh.tri9, bvh.box6, tieMargin(g), blockDim, nbx, nby, nbz,
                       activation, mask, rb.r, forceReplay() ? 1 : 0);
    SDFO_HIP(hipGetLastError());
    hipLaunchKernelGGL(sparseActivateReplayKernel, dim3(kReplayGrid), dim3(kReplayBlock), 0, s, g,
                       bvh.nodes, bvh.tri9, dir, blockDim, nbx, nby, activation, rb.r, mask);
    SDFO_HIP(hipGetLastError());
}

void launchSparseFill(const GridSpec& g, const DeviceBvh& bvh, int blockDim, int nb,
                      const int32_t* bijk, float* sphi, float* sn3, hipStream_t s)
{
    const int64_t total = int64_t(nb) * blockDim * blockDim * blockDim;
    if (total <= 0) return;
    const D3 dir = dev::d3(bvh.rayDir[0], bvh.rayDir[1], bvh.rayDir[2]);
    const int64_t units = (total + 63) / 64;
    ReplayBuf rb(total, s, units);
    const double mu = tieMargin(g);
    hipLaunchKernelGGL(sparseFillKernel, dim3(queueGrid(units)), dim3(kFillBlock), 0, s, g, bvh.nodes,
                       bvh.tri9, bvh.box6, dir, mu, blockDim, total, bijk, sphi, sn3, rb.r, rb.h,
                       rb.queue, debugFillBudget() > 0 ? debugFillBudget() : kBudgets[0],
                       forceReplay() ? 1 : 0);
    SDFO_HIP(hipGetLastError());
    hipLaunchKernelGGL((sparseHeavyKernel<kFrontierSmall, 2>), dim3(heavyGrid()), dim3(128), 0, s, g,
                       bvh.nodes, bvh.tri9, bvh.box6, dir, mu, blockDim, total, bijk, sphi, sn3,
                       rb.r, rb.h, rb.sp, forceReplay() ? 1 : 0);
    SDFO_HIP(hipGetLastError());
    hipLaunchKernelGGL(sparseSpillKernel, dim3(spillGrid()), dim3(64), 0, s, g, bvh.nodes, bvh.tri9,
                       bvh.box6, dir, mu, blockDim, bijk, sphi, sn3, rb.r, rb.sp,
                       forceReplay() ? 1 : 0);
    SDFO_HIP(hipGetLastError());
    hipLaunchKernelGGL(sparseFillReplayKernel, dim3(kReplayGrid), dim3(kReplayBlock), 0, s, g,
                       bvh.nodes, bvh.tri9, dir, blockDim, bijk, rb.r, sphi, sn3);
    SDFO_HIP(hipGetLastError());
    rb.report();
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

namespace {

// The box of each slot triangle: the same min / max as the host computes.
__global__ void __launch_bounds__(256)
slotBoxKernel(const double* __restrict__ tri9, int64_t n, double* __restrict__ box6)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double* t = tri9 + 9 * s;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const double a = t[k], b = t[3 + k], c = t[6 + k];
        double lo = (b < a) ? b : a;
        lo = (c < lo) ? c : lo;
        double hi = (a < b) ? b : a;
        hi = (hi < c) ? c : hi;
        box6[6 * s + k] = lo;
        box6[6 * s + 3 + k] = hi;
    }
}

// std::min / std::max as the host tree applies them (the first argument wins
// ties, which fixes the sign of zero bounds).
__device__ __forceinline__ double minOfDev(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double maxOfDev(double a, double b) { return (a < b) ? b : a; }

// Box of a triangle's vertices, folded as the host builder folds it.
__device__ __forceinline__ void triangleBox(const double* __restrict__ verts,
                                            const int32_t* __restrict__ ix, double* lo,
                                            double* hi)
{
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const double a = verts[3 * int64_t(ix[0]) + k], b = verts[3 * int64_t(ix[1]) + k],
                     c = verts[3 * int64_t(ix[2]) + k];
        lo[k] = minOfDev(minOfDev(a, b), c);
        hi[k] = maxOfDev(maxOfDev(a, b), c);
    }
}

// The BVH of KdTree::build (KdTree.h:91-121, 188-242) assembled from the slot
// order: the tree's shape follows from the triangle count (split at
// count / 2, leaves <= 8, pre-order indices), so a node's index, slot range
// and children follow from its path from the root.  Leaf boxes are folded
// over the leaf's slots in order and inner boxes are the union of the
// children's, as the host builder makes them.  The tree is cut at depth
// P.topLevels into 2^topLevels subtrees of at most kSubtreeSlots slots: one
// workgroup per subtree writes its slots (triangle, box) and nodes, then one
// workgroup per mesh writes the nodes above the cut.  Inside a workgroup a
// node is written by the second of its children to finish (counted with a
// workgroup-scope atomic), so no lane waits and no cache traffic between
// XCDs is needed.
constexpr int kAssembleBlock = 256;

struct TreeNodeAt
{
    int me, lo, hi, depth; // mesh-local node index, slot range, depth (root 0)
};

// Node count of the left child of a node at depth d holding c slots: the
// sizes at depth d + 1 are (n >> (d + 1)) or one more, tabulated per mesh.
__device__ __forceinline__ int leftChildNodes(const BvhMeshPart& P, int c, int d)
{
    const int h = c >> 1;
    return h == int(P.nSlots >> (d + 1)) ? P.f0[d + 1] : P.f1[d + 1];
}

// The node reached from t by n steps, step e going right when bit
// (n - 1 - e) of bits is set.
__device__ __forceinline__ TreeNodeAt walkDown(const BvhMeshPart& P, TreeNodeAt t, uint32_t bits, int n)
{
    for (int e = 0; e < n; ++e)
    {
        const int mid = t.lo + ((t.hi - t.lo) >> 1);
        if ((bits >> (n - 1 - e)) & 1u)
        {
            t.me += 1 + leftChildNodes(P, t.hi - t.lo, t.depth);
            t.lo = mid;
        }
        else
        {
            t.me += 1;
            t.hi = mid;
        }
        ++t.depth;
    }
    return t;
}

// Writes inner node t from its two children.
__device__ __forceinline__ void writeInner(const BvhMeshPart& P, BvhNode* nodes, const TreeNodeAt& t)
{
    const int L = t.me + 1, R = t.me + 1 + leftChildNodes(P, t.hi - t.lo, t.depth);
    const BvhNode& nl = nodes[P.nodeBase + L];
    const BvhNode& nr = nodes[P.nodeBase + R];
    BvhNode nd;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        nd.bmin[k] = minOfDev(nl.bmin[k], nr.bmin[k]);
        nd.bmax[k] = maxOfDev(nl.bmax[k], nr.bmax[k]);
    }
    nd.left = int(P.nodeBase + L);
    nd.right = int(P.nodeBase + R);
    nd.triStart = -1;
    nd.triCount = 0;
    nodes[P.nodeBase + t.me] = nd;
}

// The second of a node's children to finish writes it (its flag counts the
// children; workgroup scope: the node and its children belong to one
// workgroup's part of the tree).
__device__ __forceinline__ bool secondChild(int32_t* flag)
{
    return __hip_atomic_fetch_add(flag, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
}

__global__ void __launch_bounds__(kAssembleBlock)
bvhSubtreeKernel(const BvhMeshPart* __restrict__ parts, int nParts,
                 const double* __restrict__ verts, const int32_t* __restrict__ tris,
                 const int32_t* __restrict__ slotTri, double* __restrict__ tri9,
                 double* __restrict__ box6, BvhNode* nodes, int32_t* flags)
{
    int lo = 0, hi = nParts - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (parts[mid].subBase <= int64_t(blockIdx.x))
            lo = mid;
        else
            hi = mid - 1;
    }
    const BvhMeshPart& P = parts[lo];
    const double* V = verts + 3 * P.vertBase;
    const int32_t* T = tris + 3 * P.triBase;
    const TreeNodeAt root = walkDown(P, TreeNodeAt{0, 0, int(P.nSlots), 0},
                                     uint32_t(int64_t(blockIdx.x) - P.subBase), P.topLevels);
    for (int s = root.lo + int(threadIdx.x); s < root.hi; s += kAssembleBlock)
    {
        const int64_t gs = P.slotBase + s;
        const int32_t* ix = T + 3 * int64_t(slotTri[gs]);
        double* t9 = tri9 + 9 * gs;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < 3; ++k) t9[3 * c + k] = V[3 * int64_t(ix[c]) + k];
        double bl[3], bh[3];
        triangleBox(V, ix, bl, bh);
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            box6[6 * gs + k] = bl[k];
            box6[6 * gs + 3 + k] = bh[k];
        }
        // down to the leaf holding s; bit (steps - 1 - e) of path = right at step e
        TreeNodeAt t = root;
        uint32_t path = 0;
        int steps = 0;
        while (t.hi - t.lo > 8)
        {
            const bool right = s >= t.lo + ((t.hi - t.lo) >> 1);
            path = (path << 1) | (right ? 1u : 0u);
            t = walkDown(P, t, right ? 1u : 0u, 1);
            ++steps;
        }
        if (s != t.lo) continue;
        double bmn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, bmx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        for (int q = t.lo; q < t.hi; ++q)
        {
            triangleBox(V, T + 3 * int64_t(slotTri[P.slotBase + q]), bl, bh);
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                bmn[k] = minOfDev(bmn[k], bl[k]);
                bmx[k] = maxOfDev(bmx[k], bh[k]);
            }
        }
        BvhNode nd;
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            nd.bmin[k] = bmn[k];
            nd.bmax[k] = bmx[k];
        }
        nd.left = nd.right = -1;
        nd.triStart = int(P.slotBase + t.lo);
        nd.triCount = t.hi - t.lo;
        nodes[P.nodeBase + t.me] = nd;
        for (int d = steps - 1; d >= 0; --d)
        {
            const TreeNodeAt up = walkDown(P, root, path >> (steps - d), d);
            if (!secondChild(flags + P.nodeBase + up.me)) break;
            writeInner(P, nodes, up);
        }
    }
}

// The nodes above the cut: one workgroup per mesh, one lane per subtree root
// (written by bvhSubtreeKernel), climbing the same way.
__global__ void __launch_bounds__(kAssembleBlock)
bvhTopKernel(const BvhMeshPart* __restrict__ parts, BvhNode* nodes, int32_t* flags)
{
    const BvhMeshPart& P = parts[blockIdx.x];
    const int L = P.topLevels;
    for (int r = int(threadIdx.x); r < (1 << L); r += kAssembleBlock)
    {
        for (int d = L - 1; d >= 0; --d)
        {
            const TreeNodeAt up =
                walkDown(P, TreeNodeAt{0, 0, int(P.nSlots), 0}, uint32_t(r) >> (L - d), d);
            if (!secondChild(flags + P.nodeBase + up.me)) break;
            writeInner(P, nodes, up);
        }
    }
}

} // namespace

void launchSlotBoxes(const double* tri9, int64_t n, double* box6, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(slotBoxKernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, tri9, n, box6);
    SDFO_HIP(hipGetLastError());
}

void launchBvhAssemble(const DeviceBvh& db, hipStream_t s)
{
    if (db.ntris <= 0 || db.nSubtrees <= 0) return;
    hipLaunchKernelGGL(bvhSubtreeKernel, dim3(unsigned(db.nSubtrees)), dim3(kAssembleBlock), 0, s,
                       db.parts, db.nParts, db.verts, db.tris, db.triIdx, db.tri9, db.box6, db.nodes,
                       db.flags);
    SDFO_HIP(hipGetLastError());
    if (db.topLevelsMax > 0)
    {
        hipLaunchKernelGGL(bvhTopKernel, dim3(unsigned(db.nParts)), dim3(kAssembleBlock), 0, s, db.parts,
                           db.nodes, db.flags);
        SDFO_HIP(hipGetLastError());
    }
}

namespace {

// KdTree::findNearest / findSignedNearest (KdTree.h:123-165) for a stream of
// points: one lane per point, the per-lane search of the reference.
constexpr int kMeshBlock = 128;
__global__ void __launch_bounds__(kMeshBlock)
meshQueryKernel(const BvhNode* __restrict__ nodes, int nnodes, const double* __restrict__ tri9,
                const int32_t* __restrict__ slotTri, D3 dir, int signedDist,
                const double* __restrict__ pts, int64_t n, double* __restrict__ dist,
                double* __restrict__ cp3, double* __restrict__ n3, int32_t* __restrict__ tri)
{
    __shared__ int stack[kStackDepth * kMeshBlock];
    const int64_t i = int64_t(blockIdx.x) * kMeshBlock + threadIdx.x;
    if (i >= n) return;
    const D3 p = dev::d3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    const dev::LaneStack st{stack + threadIdx.x, kMeshBlock};
    double bestSq = DBL_MAX;
    int bestSlot = -1, hits = 0;
    if (nnodes > 0) // an empty tree compares nothing (KdTree.h:123-133)
    {
        dev::refNearest(p, nodes, tri9, st, bestSq, bestSlot);
        if (signedDist) hits = dev::refRayHits(p, dir, nodes, tri9, st);
    }
    const dev::SignedNearest r = dev::finishSigned(p, tri9, bestSq, bestSlot, hits);
    if (dist) dist[i] = signedDist ? r.sd : (bestSlot >= 0 ? sqrt(bestSq) : DBL_MAX);
    if (cp3)
    {
        cp3[3 * i] = r.cp.x;
        cp3[3 * i + 1] = r.cp.y;
        cp3[3 * i + 2] = r.cp.z;
    }
    if (n3)
    {
        n3[3 * i] = r.face.x;
        n3[3 * i + 1] = r.face.y;
        n3[3 * i + 2] = r.face.z;
    }
    if (tri) tri[i] = bestSlot >= 0 ? slotTri[bestSlot] : -1;
}

} // namespace

void launchMeshQuery(const DeviceBvh& bvh, bool signedDist, const double* pts, int64_t n,
                     double* dist, double* cp3, double* n3, int32_t* tri, hipStream_t s)
{
    if (n <= 0) return;
    const D3 dir = dev::d3(bvh.rayDir[0], bvh.rayDir[1], bvh.rayDir[2]);
    hipLaunchKernelGGL(meshQueryKernel, dim3(unsigned((n + kMeshBlock - 1) / kMeshBlock)),
                       dim3(kMeshBlock), 0, s, bvh.nodes, bvh.nnodes, bvh.tri9, bvh.triIdx, dir,
                       signedDist ? 1 : 0, pts, n, dist, cp3, n3, tri);
    SDFO_HIP(hipGetLastError());
}

} // namespace sdfo
