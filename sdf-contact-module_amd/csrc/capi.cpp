// C ABI of libsdfo (include/sdfo.h): argument checking, grid handles,
// builders and launch of the query / contact kernels.  Every entry point
// converts exceptions into a status code plus a thread-local message.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sdfo_internal.hpp"

// The opaque batched-query handle of the C ABI.
struct sdfo_query_plan
{
    int device = 0;
    std::unique_ptr<sdfo::QueryPlan> plan;
};

namespace sdfo {

// Scratch buffers of the build kernels come from the device's stream-ordered
// pool; keeping the pool's memory across synchronisations makes the next
// build's allocations free instead of remapping pages every time.
void keepStreamPoolMemory(int device)
{
    static std::mutex lock;
    static std::vector<int> done;
    std::lock_guard<std::mutex> lk(lock);
    if (std::find(done.begin(), done.end(), device) != done.end()) return;
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool)
    {
        uint64_t keep = ~uint64_t(0);
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    done.push_back(device);
}

namespace {
thread_local std::string g_lastError;

// Test switch of the calling thread: grids built or loaded while it is on
// keep no cell records, so every query and contact kernel takes the
// per-corner gather path that grids past cellRecordLimit() use.
thread_local bool g_skipCellRecords = false;
// A thread keeps the host BVH scratch of meshes up to this many triangles
// between builds (about 100 MB at the limit).
constexpr int64_t kKeepScratchTris = int64_t(1) << 20;

bool cellRecordsFit(int64_t bytes) { return !g_skipCellRecords && bytes <= cellRecordLimit(); }
}

int64_t cellRecordLimit()
{
    static std::mutex m;
    static int64_t limits[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return kCellRecordLimit;
    std::lock_guard<std::mutex> lk(m);
    if (limits[dev] == 0)
    {
        size_t freeB = 0, totalB = 0;
        int64_t lim = kCellRecordLimit;
        if (hipMemGetInfo(&freeB, &totalB) == hipSuccess && int64_t(totalB / 8) > lim)
            lim = int64_t(totalB / 8);
        limits[dev] = lim;
    }
    return limits[dev];
}

void setSkipCellRecords(bool on) { g_skipCellRecords = on; }

void setLastError(const std::string& msg) { g_lastError = msg; }

// Builds the FP32 query mirror of a dense grid whose FP64 fields are on the device.
double* lastBuildPhases()
{
    thread_local double phases[kBuildPhases] = {0, 0, 0, 0, 0, 0, 0, 0};
    return phases;
}

void finishDense(sdfo_grid& g, hipStream_t s, bool mirrored)
{
    const int64_t N = g.spec.total();
    if (!g.pn) g.pn = g.alloc<float4>(size_t(N));
    if (!g.h6f) g.h6f = g.alloc<float>(size_t(N) * 6);
    if (!mirrored) launchDenseMirror(N, g.phi, g.n3, g.h6, g.pn, g.h6f, s);
    const int64_t cells = int64_t(g.spec.n[0] - 1) * (g.spec.n[1] - 1) * (g.spec.n[2] - 1);
    if (cells > 0 && cellRecordsFit(cells * 8 * int64_t(sizeof(float4))))
    {
        if (!g.cellRec) g.cellRec = g.alloc<float4>(size_t(cells) * 8);
        launchCellPack(g.spec, g.pn, g.cellRec, s);
    }
    if (g.stride > 0)
    {
        if (!g.cpn) g.cpn = g.alloc<float4>(size_t(N) * g.stride);
        launchCandidateMirror(N * g.stride, g.cphi, g.cn3, g.cpn, s);
    }
}

int64_t denseCells(const GridSpec& sp)
{
    return int64_t(sp.n[0] - 1) * (sp.n[1] - 1) * (sp.n[2] - 1);
}

// Whether a dense grid of this shape keeps cell records (the calling
// thread's switch and the device's record limit).
bool denseRecords(const GridSpec& sp)
{
    const int64_t cells = denseCells(sp);
    return cells > 0 && cellRecordsFit(cells * 8 * int64_t(sizeof(float4)));
}

// Device bytes of a dense grid's arrays carved from one block: the FP64
// fields, the FP32 mirrors and, with `records`, the cell records.
size_t denseGridBytes(const GridSpec& sp, bool records)
{
    const size_t N = size_t(sp.total());
    size_t b = 0;
    for (size_t per : {sizeof(double), 3 * sizeof(double), 6 * sizeof(double), 3 * sizeof(double),
                       sizeof(int32_t), sizeof(float4), 6 * sizeof(float)})
        b += sdfo_grid::arenaBytes(per * N);
    if (records) b += sdfo_grid::arenaBytes(size_t(denseCells(sp)) * 8 * sizeof(float4));
    return b;
}

// Carves the arrays denseGridBytes counts from g's block, in the same order.
void carveDenseGrid(sdfo_grid& g, bool records)
{
    const int64_t N = g.spec.total();
    g.phi = g.alloc<double>(size_t(N));
    g.n3 = g.alloc<double>(size_t(N) * 3);
    g.h6 = g.alloc<double>(size_t(N) * 6);
    g.witness3 = g.alloc<double>(size_t(N) * 3);
    g.tri = g.alloc<int32_t>(size_t(N));
    // the FP32 mirrors and cell records of finishDense
    g.pn = g.alloc<float4>(size_t(N));
    g.h6f = g.alloc<float>(size_t(N) * 6);
    if (records) g.cellRec = g.alloc<float4>(size_t(denseCells(g.spec)) * 8);
}

// Gives g one device block of `bytes` (ordered on s) to carve its arrays from.
void giveArena(sdfo_grid& g, size_t bytes, hipStream_t s)
{
    auto arena = std::make_shared<ArenaBlock>();
    SDFO_HIP(hipMallocAsync(reinterpret_cast<void**>(&arena->base), std::max<size_t>(bytes, 256), s));
    g.arena = arena;
    g.arenaNext = arena->base;
    g.arenaLeft = bytes;
}

// Brick table from the brick list: first brick wins on duplicate keys
// (std::unordered_map::emplace in SparseSdfData::rebuildIndex, SdfData.h:172-182);
// bricks outside the brick grid are unreachable.
std::vector<int32_t> brickTable(const std::vector<int32_t>& bijk, int nbx, int nby, int nbz)
{
    std::vector<int32_t> table(size_t(nbx) * nby * nbz, -1);
    const int nb = int(bijk.size() / 3);
    for (int b = 0; b < nb; ++b)
    {
        const int bi = bijk[3 * b], bj = bijk[3 * b + 1], bk = bijk[3 * b + 2];
        if (bi < 0 || bi >= nbx || bj < 0 || bj >= nby || bk < 0 || bk >= nbz) continue;
        int32_t& slot = table[bi + size_t(nbx) * (bj + size_t(nby) * bk)];
        if (slot < 0) slot = b;
    }
    return table;
}

void finishSparse(sdfo_grid& g, const std::vector<int32_t>& bijk, hipStream_t s)
{
    const std::vector<int32_t> table = brickTable(bijk, g.nbx, g.nby, g.nbz);
    if (!g.table) g.table = g.alloc<int32_t>(table.size());
    SDFO_HIP(hipMemcpyAsync(g.table, table.data(), table.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice, s));
    SDFO_HIP(hipStreamSynchronize(s)); // the host table goes out of scope
    const int64_t bvc = int64_t(g.blockDim) * g.blockDim * g.blockDim;
    const int64_t count = int64_t(g.nb) * bvc;
    if (count > 0)
    {
        if (!g.spn) g.spn = g.alloc<float4>(size_t(count));
        launchSparseMirror(count, g.sphi, g.sn3, g.spn, s);
        if (cellRecordsFit(count * 8 * int64_t(sizeof(float4))) && g.bijk)
        {
            const int words = int((bvc + 63) / 64);
            if (!g.scellRec) g.scellRec = g.alloc<float4>(size_t(count) * 8);
            if (!g.scellMask) g.scellMask = g.alloc<unsigned long long>(size_t(g.nb) * words);
            launchSparseCellPack(g.spec, g.blockDim, g.nbx, g.nby, g.nbz, g.bijk, g.table, g.spn,
                                 count, words, g.scellRec, g.scellMask, s);
        }
        if (g.stride > 0)
        {
            if (!g.scpn) g.scpn = g.alloc<float4>(size_t(count) * g.stride);
            launchSparseMirror(count * g.stride, g.scphi, g.scn3, g.scpn, s);
        }
    }
}

namespace {

void checkMesh(const double* verts, int64_t nv, const int32_t* tris, int64_t nt)
{
    if (nv <= 0 || nt <= 0 || verts == nullptr || tris == nullptr)
        throw Error(SDFO_ERR_EMPTY_MESH, "Empty mesh provided to SdfBuilder");
    if (nt > (int64_t(1) << 30)) throw Error(SDFO_ERR_INVALID_ARGUMENT, "too many triangles");
}

void checkParams(const sdfo_build_params* p)
{
    if (p == nullptr) throw Error(SDFO_ERR_INVALID_ARGUMENT, "params is NULL");
    if (p->voxel_size <= 0.0 && p->target_resolution <= 0)
        throw Error(SDFO_ERR_INVALID_ARGUMENT, "target_resolution must be positive");
    if (p->max_resolution <= 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "max_resolution must be positive");
}

void checkPlanArgs(int32_t ns, const sdfo_grid* const* grids, const int64_t* offsets)
{
    if (ns < 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "n_streams must be >= 0");
    if (ns == 0) return;
    if (!grids || !offsets) throw Error(SDFO_ERR_INVALID_ARGUMENT, "grids and offsets are required");
    if (offsets[0] != 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "offsets[0] must be 0");
    for (int32_t q = 0; q < ns; ++q)
    {
        if (!grids[q]) throw Error(SDFO_ERR_INVALID_ARGUMENT, "grid is NULL");
        if (offsets[q + 1] < offsets[q])
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
    }
}

const int kPlanBlocks[4] = {64, 128, 256, 512};
const char* const kPlanBlockNames[4] = {"64", "128", "256", "512"};

// Runs a plan with the tuner's block size for its size class (timing every
// block size first when the tuner is on and the stream is not capturing).
void runPlanTuned(const QueryPlan& plan, int mode, const float* pts, float* gap, float* grad3,
                  uint8_t* valid, hipStream_t s)
{
    const std::string key = "query_plan mode=" + std::to_string(mode) + " points~2^" +
                            std::to_string(log2Bucket(plan.points));
    int v = 2;
    if (!tunerLookup(key, v) && tunerEnabled() && !streamCapturing(s) && plan.points > 0)
    {
        runQueryPlan(plan, mode, kPlanBlocks[0], pts, gap, grad3, valid, s); // warm
        const std::vector<std::pair<int, double>> t = timeVariants(s, 4, [&](int q) {
            runQueryPlan(plan, mode, kPlanBlocks[q], pts, gap, grad3, valid, s);
        });
        tunerRecord(key, t, kPlanBlockNames);
        tunerLookup(key, v);
    }
    runQueryPlan(plan, mode, kPlanBlocks[std::max(0, std::min(3, v))], pts, gap, grad3, valid, s);
}

int currentDevice()
{
    int d = 0;
    SDFO_HIP(hipGetDevice(&d));
    return d;
}

// Keeps the grid's device current for the duration of a call.
struct DeviceGuard
{
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        int cur = 0;
        SDFO_HIP(hipGetDevice(&cur));
        if (cur != dev)
        {
            prev = cur;
            SDFO_HIP(hipSetDevice(dev));
        }
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

std::unique_ptr<sdfo_grid> buildDense(const double* verts, int64_t nv, const int32_t* tris,
                                      int64_t nt, const sdfo_build_params& p, hipStream_t s)
{
    using clk = std::chrono::steady_clock;
    auto msSince = [](clk::time_point a) {
        return std::chrono::duration<double, std::milli>(clk::now() - a).count();
    };
    double* ph = lastBuildPhases();
    for (int q = 0; q < kBuildPhases; ++q) ph[q] = 0.0;
    const clk::time_point t0 = clk::now();
    checkMesh(verts, nv, tris, nt);
    auto g = std::make_unique<sdfo_grid>();
    g->device = currentDevice();
    g->allocStream = s;
    g->spec = setupGrid(verts, nv, p);
    ph[0] = msSince(t0);
    // A worker thread requests the device memory while this thread builds the
    // host BVH: the grid's arrays in one block, the BVH's block and the fill
    // passes' lists (the allocations cost host time comparable to the tree).
    const bool records = denseRecords(g->spec);
    DeviceBvh db;
    void* fillScratch = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    PoolTask alloc;
    alloc.fn = [&] {
        DeviceGuard dg(g->device);
        keepStreamPoolMemory(g->device);
        const size_t bytes = denseGridBytes(g->spec, records);
        giveArena(*g, bytes, s);
        carveDenseGrid(*g, records);
        // the mesh goes to the device while the host orders the tree's slots
        db.reserveAssembly(bvhNodeCount(nt), nt, nv, 1, s);
        const BvhMeshPart part{0, nt, 0, 0, 0};
        db.stageMeshes(verts, nv, tris, nt, &part, 1, s);
        SDFO_HIP(hipMallocAsync(&fillScratch, denseFillScratchBytes(g->spec), s));
        SDFO_HIP(hipMemsetAsync(fillScratch, 0, denseFillScratchHeaderBytes(), s));
        for (hipEvent_t& e : ev) SDFO_HIP(hipEventCreate(&e));
    };
    bool allocJoined = false;
    auto joinAlloc = [&] {
        if (allocJoined) return;
        allocJoined = true;
        poolWait(&alloc);
    };
    auto cleanUp = [&] {
        (void)hipStreamSynchronize(s);
        if (fillScratch) (void)hipFreeAsync(fillScratch, s);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        db.release();
        // the copies from the scratch arrays finished with the stream
        releaseHostBvhScratch(kKeepScratchTris);
    };
    try
    {
        poolSubmit(&alloc);
        const clk::time_point t1 = clk::now();
        const HostBvhOrderView hb = buildHostBvhOrderScratch(verts, nv, tris, nt);
        ph[1] = msSince(t1);
        const clk::time_point t2 = clk::now();
        joinAlloc();
        db.assemble(hb.slotTri, hb.depth, hb.rayDir, s);
        ph[2] = msSince(t2);
        SDFO_HIP(hipEventRecord(ev[0], s));
        launchDenseFill(g->spec, db, g->phi, g->n3, g->witness3, g->tri, s, fillScratch, true);
        SDFO_HIP(hipEventRecord(ev[1], s));
        launchDenseHessian(g->spec, g->n3, g->h6, s, g->phi, g->pn, g->h6f);
        SDFO_HIP(hipEventRecord(ev[2], s));
        finishDense(*g, s, true);
        SDFO_HIP(hipEventRecord(ev[3], s));
        SDFO_HIP(hipStreamSynchronize(s));
        float f = 0.0f;
        for (int q = 0; q < 3; ++q)
        {
            SDFO_HIP(hipEventElapsedTime(&f, ev[q], ev[q + 1]));
            ph[3 + q] = f;
        }
    }
    catch (...)
    {
        // the worker's allocations must be finished before they are freed
        try
        {
            joinAlloc();
        }
        catch (...)
        {
        }
        cleanUp();
        throw;
    }
    cleanUp();
    ph[6] = msSince(t0);
    const int* rc = replayCounterHost();
    ph[7] = rc ? double(*rc) : -1.0;
    return g;
}

// Page-locked host staging for the batched build's concatenated BVH, kept
// between builds and grown on demand, so its upload is a DMA instead of a
// staged copy of pageable memory.  One batched build holds it at a time.
struct PinnedStaging
{
    std::mutex lock;
    char* p = nullptr;
    size_t bytes = 0;
    char* get(size_t n)
    {
        if (n > bytes)
        {
            if (p) (void)hipHostFree(p);
            p = nullptr;
            bytes = 0;
            SDFO_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), n, hipHostMallocDefault));
            bytes = n;
        }
        return p;
    }
};

PinnedStaging& batchStaging()
{
    static PinnedStaging* s = new PinnedStaging; // kept for the process
    return *s;
}

// Many meshes: host set-up and BVHs in parallel, one concatenated BVH on the
// device, one queue-driven fill launch for all bricks, then the per-mesh
// Hessian and mirror kernels.
std::vector<std::unique_ptr<sdfo_grid>> buildDenseBatched(int32_t nm, const double* verts,
                                                          const int64_t* vo, const int32_t* tris,
                                                          const int64_t* to,
                                                          const sdfo_build_params* params,
                                                          hipStream_t s)
{
    using clk = std::chrono::steady_clock;
    auto msSince = [](clk::time_point a) {
        return std::chrono::duration<double, std::milli>(clk::now() - a).count();
    };
    double* ph = lastBuildPhases();
    for (int q = 0; q < kBuildPhases; ++q) ph[q] = 0.0;
    const clk::time_point t0 = clk::now();
    for (int32_t m = 0; m < nm; ++m)
    {
        if (vo[m + 1] < vo[m] || to[m + 1] < to[m])
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "batched build: offsets must be non-decreasing");
        checkMesh(verts + 3 * vo[m], vo[m + 1] - vo[m], tris + 3 * to[m], to[m + 1] - to[m]);
    }
    // The concatenated BVH: each mesh's node and slot ranges are known from
    // its triangle count, so the meshes build straight into their slices.
    std::vector<GridSpec> specs(static_cast<size_t>(nm));
    std::vector<int64_t> nodeBase(static_cast<size_t>(nm) + 1, 0), slotBase(static_cast<size_t>(nm) + 1, 0);
    for (int32_t m = 0; m < nm; ++m)
    {
        specs[size_t(m)] = setupGrid(verts + 3 * vo[m], vo[m + 1] - vo[m], params[m]);
        nodeBase[size_t(m) + 1] = nodeBase[size_t(m)] + bvhNodeCount(to[m + 1] - to[m]);
        slotBase[size_t(m) + 1] = slotBase[size_t(m)] + (to[m + 1] - to[m]);
    }
    if (nodeBase[size_t(nm)] >= (int64_t(1) << 26) || slotBase[size_t(nm)] > INT32_MAX)
        throw Error(SDFO_ERR_UNSUPPORTED, "batched build: the concatenated BVH is too large");
    // raw arrays: every element is written by the mesh builds (no zero fill)
    const int64_t nAllNodes = nodeBase[size_t(nm)], nAllSlots = slotBase[size_t(nm)];
    // (the stream is synchronised before this function returns, so the
    // staging is free again when the lock is released)
    std::lock_guard<std::mutex> stagingLock(batchStaging().lock);
    const size_t slotBytes = sizeof(int32_t) * size_t(std::max<int64_t>(1, nAllSlots));
    int32_t* const allSlotTri = reinterpret_cast<int32_t*>(batchStaging().get(slotBytes));
    std::vector<int> depths(static_cast<size_t>(nm), 0);
    // meshes are independent: the host worker threads take them largest
    // first, and a mesh's large subtrees go to idle workers
    std::vector<int32_t> order(static_cast<size_t>(nm));
    for (int32_t m = 0; m < nm; ++m) order[size_t(m)] = m;
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return (to[a + 1] - to[a]) > (to[b + 1] - to[b]);
    });
    std::atomic<int32_t> next{0};
    std::vector<std::string> errs(static_cast<size_t>(nm));
    std::vector<int> codes(static_cast<size_t>(nm), SDFO_OK);
    auto work = [&] {
        for (int32_t q = next++; q < nm; q = next++)
        {
            const int32_t m = order[size_t(q)];
            try
            {
                // the slot order only: the tree is assembled on the device
                const size_t sb = size_t(slotBase[size_t(m)]);
                depths[size_t(m)] = buildHostBvhOrder(verts + 3 * vo[m], vo[m + 1] - vo[m],
                                                      tris + 3 * to[m], to[m + 1] - to[m],
                                                      allSlotTri + sb, true);
            }
            catch (const Error& e)
            {
                codes[size_t(m)] = e.code;
                errs[size_t(m)] = e.what();
            }
            catch (const std::exception& e)
            {
                codes[size_t(m)] = SDFO_ERR_INVALID_ARGUMENT;
                errs[size_t(m)] = e.what();
            }
        }
    };
    const int nth = std::max(1, std::min<int>(nm, hostThreads()));
    std::vector<PoolTask> pool(size_t(nth - 1));
    struct PoolJoin
    {
        std::vector<PoolTask>& p;
        ~PoolJoin()
        {
            for (PoolTask& t : p)
            {
                try
                {
                    if (t.fn) poolWait(&t);
                }
                catch (...)
                {
                }
            }
        }
    } poolJoin{pool};
    ph[0] = msSince(t0); // validation and grid set-ups
    const clk::time_point tPool = clk::now();
    for (PoolTask& t : pool)
    {
        t.fn = work;
        poolSubmit(&t);
    }

    // meanwhile this thread allocates every grid's device arrays, all from
    // one device block the grids share
    std::vector<std::unique_ptr<sdfo_grid>> grids(static_cast<size_t>(nm));
    std::vector<BatchMesh> jobs(static_cast<size_t>(nm));
    const int dev = currentDevice();
    keepStreamPoolMemory(dev);
    auto gridBytes = [&](const GridSpec& sp) { return denseGridBytes(sp, denseRecords(sp)); };
    size_t arenaTotal = 0;
    for (int32_t m = 0; m < nm; ++m) arenaTotal += gridBytes(specs[size_t(m)]);
    char* arenaBase = nullptr;
    SDFO_HIP(hipMallocAsync(reinterpret_cast<void**>(&arenaBase), std::max<size_t>(arenaTotal, 256), s));
    const auto arena = std::make_shared<ArenaBlock>();
    arena->base = arenaBase;
    size_t arenaOff = 0;
    // the meshes go to the device beside the BVH work too
    DeviceBvh db;
    struct ReleaseOnExit
    {
        DeviceBvh& d;
        ~ReleaseOnExit() { d.release(); } // stream-ordered, also on an early throw
    } dbRelease{db};
    std::vector<BvhMeshPart> parts(static_cast<size_t>(nm));
    for (int32_t m = 0; m < nm; ++m)
        parts[size_t(m)] = BvhMeshPart{slotBase[size_t(m)], to[m + 1] - to[m], nodeBase[size_t(m)],
                                       vo[m] - vo[0], to[m] - to[0]};
    db.reserveAssembly(nAllNodes, nAllSlots, vo[nm] - vo[0], nm, s);
    db.stageMeshes(verts + 3 * vo[0], vo[nm] - vo[0], tris + 3 * to[0], to[nm] - to[0], parts.data(),
                   nm, s);
    for (int32_t m = 0; m < nm; ++m)
    {
        auto g = std::make_unique<sdfo_grid>();
        g->device = dev;
        g->allocStream = s;
        g->spec = specs[size_t(m)];
        g->arena = arena;
        g->arenaNext = arenaBase + arenaOff;
        g->arenaLeft = gridBytes(g->spec);
        arenaOff += g->arenaLeft;
        carveDenseGrid(*g, denseRecords(g->spec));
        grids[size_t(m)] = std::move(g);
    }
    ph[2] = msSince(tPool); // allocations, overlapping the BVHs
    work();
    for (PoolTask& t : pool) poolWait(&t);
    for (int32_t m = 0; m < nm; ++m)
        if (codes[size_t(m)] != SDFO_OK)
            throw Error(codes[size_t(m)], "mesh " + std::to_string(m) + ": " + errs[size_t(m)]);
    int depth = 0;
    for (int32_t m = 0; m < nm; ++m)
    {
        sdfo_grid& g = *grids[size_t(m)];
        jobs[size_t(m)] = BatchMesh{g.spec, int(nodeBase[size_t(m)]), to[m + 1] - to[m], g.phi, g.n3,
                                    g.witness3, g.tri};
        depth = std::max(depth, depths[size_t(m)]);
    }
    HostBvh dirOnly;
    setRayDir(dirOnly);
    ph[1] = msSince(tPool); // BVHs (with the allocations beside them)
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    try
    {
        for (hipEvent_t& e : ev) SDFO_HIP(hipEventCreate(&e));
        db.assemble(allSlotTri, depth, dirOnly.rayDir, s);
        SDFO_HIP(hipEventRecord(ev[0], s));
        launchDenseFillBatched(jobs, db, s);
        SDFO_HIP(hipEventRecord(ev[1], s));
        for (int32_t m = 0; m < nm; ++m)
        {
            sdfo_grid& g = *grids[size_t(m)];
            launchDenseHessian(g.spec, g.n3, g.h6, s, g.phi, g.pn, g.h6f);
            finishDense(g, s, true);
        }
        SDFO_HIP(hipEventRecord(ev[2], s));
        SDFO_HIP(hipStreamSynchronize(s));
        float f = 0.0f;
        SDFO_HIP(hipEventElapsedTime(&f, ev[0], ev[1]));
        ph[3] = f;
        SDFO_HIP(hipEventElapsedTime(&f, ev[1], ev[2]));
        ph[5] = f; // Hessians, mirrors and their allocations
    }
    catch (...)
    {
        (void)hipStreamSynchronize(s);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        db.release();
        releaseHostBvhScratch(kKeepScratchTris);
        throw;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    db.release();
    // the copies from the scratch arrays finished with the stream
    releaseHostBvhScratch(kKeepScratchTris);
    ph[6] = msSince(t0);
    const int* rc = replayCounterHost();
    ph[7] = rc ? double(*rc) : -1.0;
    return grids;
}

std::unique_ptr<sdfo_grid> buildSparse(const double* verts, int64_t nv, const int32_t* tris,
                                       int64_t nt, const sdfo_build_params& p, hipStream_t s)
{
    checkMesh(verts, nv, tris, nt);
    if (p.block_dim <= 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "Sparse SDF blockDim must be positive");
    auto g = std::make_unique<sdfo_grid>();
    g->device = currentDevice();
    g->allocStream = s;
    g->sparse = true;
    g->spec = setupGrid(verts, nv, p);
    g->blockDim = p.block_dim;
    g->halo = p.halo_voxels;
    g->hasHessian = p.store_hessian != 0;
    // SdfBuilder.cpp:835-842
    const double h = g->spec.hmax();
    g->band = p.band_distance > 0.0 ? p.band_distance : double(p.band_voxels) * h;
    const double blockRadius = 0.5 * std::sqrt(3.0) * double(p.block_dim) * h;
    g->activation = g->band + double(p.halo_voxels) * h + blockRadius;
    const int bd = p.block_dim;
    g->nbx = (g->spec.n[0] + bd - 1) / bd;
    g->nby = (g->spec.n[1] + bd - 1) / bd;
    g->nbz = (g->spec.n[2] + bd - 1) / bd;
    const int64_t cells = int64_t(g->nbx) * g->nby * g->nbz;
    // A worker thread reserves the device BVH and uploads the mesh while this
    // thread orders the tree's slots; the tree is then assembled on the device.
    DeviceBvh db;
    PoolTask stage;
    stage.fn = [&] {
        DeviceGuard dg(g->device);
        keepStreamPoolMemory(g->device);
        db.reserveAssembly(bvhNodeCount(nt), nt, nv, 1, s);
        const BvhMeshPart part{0, nt, 0, 0, 0};
        db.stageMeshes(verts, nv, tris, nt, &part, 1, s);
    };
    bool submitted = false, staged = false;
    try
    {
        poolSubmit(&stage);
        submitted = true;
        const HostBvhOrderView hb = buildHostBvhOrderScratch(verts, nv, tris, nt);
        staged = true;
        poolWait(&stage);
        db.assemble(hb.slotTri, hb.depth, hb.rayDir, s);
        g->activeMask = g->alloc<uint8_t>(size_t(cells));
        launchSparseActivate(g->spec, db, bd, g->nbx, g->nby, g->nbz, g->activation, g->activeMask, s);
        std::vector<uint8_t> mask(static_cast<size_t>(cells));
        SDFO_HIP(hipMemcpyAsync(mask.data(), g->activeMask, mask.size(), hipMemcpyDeviceToHost, s));
        SDFO_HIP(hipStreamSynchronize(s));
        // Blocks in (k, j, i) order (SdfBuilder.cpp:892-920).
        std::vector<int32_t> bijk;
        for (int64_t c = 0; c < cells; ++c)
        {
            if (!mask[size_t(c)]) continue;
            bijk.push_back(int32_t(c % g->nbx));
            bijk.push_back(int32_t((c / g->nbx) % g->nby));
            bijk.push_back(int32_t(c / (int64_t(g->nbx) * g->nby)));
        }
        g->nb = int(bijk.size() / 3);
        const int64_t bvc = int64_t(bd) * bd * bd;
        const int64_t count = int64_t(g->nb) * bvc;
        if (g->nb > 0)
        {
            g->bijk = g->alloc<int32_t>(bijk.size());
            SDFO_HIP(hipMemcpyAsync(g->bijk, bijk.data(), bijk.size() * sizeof(int32_t),
                                    hipMemcpyHostToDevice, s));
            g->sphi = g->alloc<float>(size_t(count));
            g->sn3 = g->alloc<float>(size_t(count) * 3);
            launchSparseFill(g->spec, db, bd, g->nb, g->bijk, g->sphi, g->sn3, s);
        }
        finishSparse(*g, bijk, s);
        if (g->hasHessian && g->nb > 0)
        {
            g->sh6 = g->alloc<float>(size_t(count) * 6);
            SDFO_HIP(hipMemsetAsync(g->sh6, 0, size_t(count) * 6 * sizeof(float), s));
            launchSparseHessian(g->spec, bd, g->nb, g->nbx, g->nby, g->nbz, g->bijk, g->table,
                                g->sn3, g->sh6, s);
        }
        SDFO_HIP(hipStreamSynchronize(s));
    }
    catch (...)
    {
        if (submitted && !staged)
        {
            try
            {
                poolWait(&stage);
            }
            catch (...)
            {
            }
        }
        (void)hipStreamSynchronize(s);
        db.release();
        releaseHostBvhScratch(kKeepScratchTris);
        throw;
    }
    db.release();
    releaseHostBvhScratch(kKeepScratchTris);
    return g;
}

template <typename F>
int guarded(F&& f)
{
    try
    {
        g_lastError.clear();
        f();
        return SDFO_OK;
    }
    catch (const Error& e)
    {
        g_lastError = e.what();
        return e.code;
    }
    catch (const std::bad_alloc&)
    {
        g_lastError = "host allocation failed";
        return SDFO_ERR_OUT_OF_MEMORY;
    }
    catch (const std::exception& e)
    {
        g_lastError = e.what();
        return SDFO_ERR_INVALID_ARGUMENT;
    }
    catch (...)
    {
        g_lastError = "unknown error";
        return SDFO_ERR_INVALID_ARGUMENT;
    }
}

void fillInfo(const sdfo_grid& g, sdfo_grid_info* info)
{
    std::memset(info, 0, sizeof(*info));
    for (int a = 0; a < 3; ++a)
    {
        info->bmin[a] = g.spec.bmin[a];
        info->bmax[a] = g.spec.bmax[a];
        info->voxel[a] = g.spec.voxel[a];
        info->shape[a] = g.spec.n[a];
    }
    info->sparse = g.sparse ? 1 : 0;
    info->has_hessian = g.hasHessian ? 1 : 0;
    info->candidate_stride = g.stride;
    info->block_dim = g.sparse ? g.blockDim : 0;
    info->halo_voxels = g.sparse ? g.halo : 0;
    info->block_count = g.nb;
    info->block_grid[0] = g.nbx;
    info->block_grid[1] = g.nby;
    info->block_grid[2] = g.nbz;
    info->band_distance = g.band;
    info->activation_distance = g.activation;
    int64_t lazy = 0;
    {
        std::lock_guard<std::mutex> lk(g.rec64.lock);
        lazy = g.rec64.rec ? g.rec64.bytes : 0;
    }
    info->device_bytes = g.deviceBytes + lazy;
}

void checkGrid(const sdfo_grid* g)
{
    if (g == nullptr) throw Error(SDFO_ERR_INVALID_ARGUMENT, "grid is NULL");
}

} // namespace

} // namespace sdfo

using namespace sdfo;

extern "C" {

const char* sdfo_version(void) { return "sdfo 0.1.0 (gfx950; SDFO file format 0.1.0)"; }

const char* sdfo_last_error(void) { return sdfo::g_lastError.c_str(); }

void sdfo_default_build_params(sdfo_build_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->padding = 0.1;
    p->voxel_size = 0.0;
    p->target_resolution = 128;
    p->max_resolution = 512;
    p->block_dim = 8;
    p->halo_voxels = 2;
    p->band_voxels = 8;
    p->store_hessian = 1;
    p->band_distance = 0.0;
    p->max_candidates = 4;
    p->tile_dim = 8;
    p->feature_distance_tolerance = 0.0;
    p->normal_merge_angle_degrees = 5.0;
    p->use_tile_acceleration = 1;
}

int sdfo_setup_grid(const double* verts, int64_t n_verts, const sdfo_build_params* params,
                    sdfo_grid_info* out)
{
    return guarded([&] {
        if (!verts || n_verts <= 0) throw Error(SDFO_ERR_EMPTY_MESH, "Empty mesh provided to SdfBuilder");
        checkParams(params);
        if (!out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        sdfo_grid tmp;
        tmp.spec = setupGrid(verts, n_verts, *params);
        fillInfo(tmp, out);
    });
}

int sdfo_build_dense(const double* verts, int64_t n_verts, const int32_t* tris, int64_t n_tris,
                     const sdfo_build_params* params, sdfo_grid** out, sdfo_stream stream)
{
    return guarded([&] {
        if (!out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        *out = nullptr;
        checkParams(params);
        *out = buildDense(verts, n_verts, tris, n_tris, *params, hipStream_t(stream)).release();
    });
}

int sdfo_build_sparse(const double* verts, int64_t n_verts, const int32_t* tris, int64_t n_tris,
                      const sdfo_build_params* params, sdfo_grid** out, sdfo_stream stream)
{
    return guarded([&] {
        if (!out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        *out = nullptr;
        checkParams(params);
        *out = buildSparse(verts, n_verts, tris, n_tris, *params, hipStream_t(stream)).release();
    });
}

void sdfo_grid_free_async(sdfo_grid* grid, sdfo_stream stream)
{
    if (!grid) return;
    grid->releaseOrdered = true;
    grid->releaseStream = hipStream_t(stream);
    sdfo_grid_free(grid);
}

void sdfo_grid_free(sdfo_grid* grid)
{
    if (!grid) return;
    try
    {
        DeviceGuard dg(grid->device);
        delete grid;
    }
    catch (...)
    {
        delete grid;
    }
}

int sdfo_grid_get_info(const sdfo_grid* grid, sdfo_grid_info* info)
{
    return guarded([&] {
        checkGrid(grid);
        if (!info) throw Error(SDFO_ERR_INVALID_ARGUMENT, "info is NULL");
        fillInfo(*grid, info);
    });
}

int sdfo_grid_download_dense(const sdfo_grid* grid, double* phi, double* n3, double* h6,
                             double* witness3, int32_t* tri)
{
    return guarded([&] {
        checkGrid(grid);
        if (grid->sparse) throw Error(SDFO_ERR_INVALID_ARGUMENT, "grid is sparse");
        DeviceGuard dg(grid->device);
        const size_t N = size_t(grid->spec.total());
        if (phi) SDFO_HIP(hipMemcpy(phi, grid->phi, N * sizeof(double), hipMemcpyDeviceToHost));
        if (n3) SDFO_HIP(hipMemcpy(n3, grid->n3, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
        if (h6) SDFO_HIP(hipMemcpy(h6, grid->h6, 6 * N * sizeof(double), hipMemcpyDeviceToHost));
        if (witness3)
            SDFO_HIP(hipMemcpy(witness3, grid->witness3, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
        if (tri)
        {
            if (grid->tri)
                SDFO_HIP(hipMemcpy(tri, grid->tri, N * sizeof(int32_t), hipMemcpyDeviceToHost));
            else
                std::fill(tri, tri + N, -1);
        }
    });
}

int sdfo_grid_download_candidates(const sdfo_grid* grid, uint8_t* count, double* phi, double* n3,
                                  int32_t* branch)
{
    return guarded([&] {
        checkGrid(grid);
        if (grid->sparse || grid->stride <= 0)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "grid has no dense contact-aware candidates");
        DeviceGuard dg(grid->device);
        const size_t N = size_t(grid->spec.total()), NS = N * size_t(grid->stride);
        if (count) SDFO_HIP(hipMemcpy(count, grid->ccount, N, hipMemcpyDeviceToHost));
        if (phi) SDFO_HIP(hipMemcpy(phi, grid->cphi, NS * sizeof(double), hipMemcpyDeviceToHost));
        if (n3) SDFO_HIP(hipMemcpy(n3, grid->cn3, 3 * NS * sizeof(double), hipMemcpyDeviceToHost));
        if (branch)
            SDFO_HIP(hipMemcpy(branch, grid->cbranch, NS * sizeof(int32_t), hipMemcpyDeviceToHost));
    });
}

int sdfo_grid_download_sparse(const sdfo_grid* grid, int32_t* bijk, float* phi, float* n3,
                              float* h6, uint8_t* active_mask, uint8_t* ccount, float* cphi,
                              float* cn3, int32_t* cbranch)
{
    return guarded([&] {
        checkGrid(grid);
        if (!grid->sparse) throw Error(SDFO_ERR_INVALID_ARGUMENT, "grid is dense");
        DeviceGuard dg(grid->device);
        const size_t bvc = size_t(grid->blockDim) * grid->blockDim * grid->blockDim;
        const size_t cnt = size_t(grid->nb) * bvc;
        const size_t cells = size_t(grid->nbx) * grid->nby * grid->nbz;
        const size_t S = size_t(grid->stride);
        auto cp = [](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) SDFO_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        };
        cp(bijk, grid->bijk, size_t(grid->nb) * 3 * sizeof(int32_t));
        cp(phi, grid->sphi, cnt * sizeof(float));
        cp(n3, grid->sn3, cnt * 3 * sizeof(float));
        if (h6)
        {
            if (grid->sh6)
                cp(h6, grid->sh6, cnt * 6 * sizeof(float));
            else
                std::memset(h6, 0, cnt * 6 * sizeof(float));
        }
        if (active_mask)
        {
            if (grid->activeMask)
                cp(active_mask, grid->activeMask, cells);
            else
                std::memset(active_mask, 0, cells);
        }
        if (S > 0)
        {
            cp(ccount, grid->sccount, cnt);
            cp(cphi, grid->scphi, cnt * S * sizeof(float));
            cp(cn3, grid->scn3, cnt * S * 3 * sizeof(float));
            cp(cbranch, grid->scbranch, cnt * S * sizeof(int32_t));
        }
    });
}

int sdfo_grid_from_dense(const double* bmin3, const double* bmax3, const double* voxel3,
                         const int32_t* shape3, const double* phi, const double* n3,
                         const double* h6, const double* witness3, int32_t stride,
                         const uint8_t* ccount, const double* cphi, const double* cn3,
                         const int32_t* cbranch, sdfo_grid** out)
{
    return guarded([&] {
        if (!out || !bmin3 || !bmax3 || !voxel3 || !shape3 || !phi || !n3)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL argument");
        *out = nullptr;
        if (stride < 0 || stride > 8) throw Error(SDFO_ERR_INVALID_ARGUMENT, "Invalid contact-aware candidate stride");
        if (stride > 0 && (!ccount || !cphi || !cn3 || !cbranch))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "candidate arrays are NULL");
        auto g = std::make_unique<sdfo_grid>();
        g->device = currentDevice();
        for (int a = 0; a < 3; ++a)
        {
            g->spec.bmin[a] = bmin3[a];
            g->spec.bmax[a] = bmax3[a];
            g->spec.voxel[a] = voxel3[a];
            g->spec.n[a] = shape3[a];
            if (shape3[a] <= 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "shape must be positive");
        }
        const size_t N = size_t(g->spec.total());
        g->phi = g->alloc<double>(N);
        g->n3 = g->alloc<double>(3 * N);
        g->h6 = g->alloc<double>(6 * N);
        g->witness3 = g->alloc<double>(3 * N);
        SDFO_HIP(hipMemcpy(g->phi, phi, N * sizeof(double), hipMemcpyHostToDevice));
        SDFO_HIP(hipMemcpy(g->n3, n3, 3 * N * sizeof(double), hipMemcpyHostToDevice));
        if (h6)
            SDFO_HIP(hipMemcpy(g->h6, h6, 6 * N * sizeof(double), hipMemcpyHostToDevice));
        else
            SDFO_HIP(hipMemset(g->h6, 0, 6 * N * sizeof(double)));
        if (witness3)
            SDFO_HIP(hipMemcpy(g->witness3, witness3, 3 * N * sizeof(double), hipMemcpyHostToDevice));
        else
            SDFO_HIP(hipMemset(g->witness3, 0, 3 * N * sizeof(double)));
        g->stride = stride;
        if (stride > 0)
        {
            const size_t NS = N * size_t(stride);
            g->ccount = g->alloc<uint8_t>(N);
            g->cphi = g->alloc<double>(NS);
            g->cn3 = g->alloc<double>(3 * NS);
            g->cbranch = g->alloc<int32_t>(NS);
            SDFO_HIP(hipMemcpy(g->ccount, ccount, N, hipMemcpyHostToDevice));
            SDFO_HIP(hipMemcpy(g->cphi, cphi, NS * sizeof(double), hipMemcpyHostToDevice));
            SDFO_HIP(hipMemcpy(g->cn3, cn3, 3 * NS * sizeof(double), hipMemcpyHostToDevice));
            SDFO_HIP(hipMemcpy(g->cbranch, cbranch, NS * sizeof(int32_t), hipMemcpyHostToDevice));
        }
        finishDense(*g, nullptr);
        SDFO_HIP(hipDeviceSynchronize());
        *out = g.release();
    });
}

int sdfo_query(const sdfo_grid* grid, int32_t mode, const float* points, int64_t n, float* gap,
               float* grad3, uint8_t* valid, float* normal3, float* witness3, int32_t* branch,
               sdfo_stream stream)
{
    return guarded([&] {
        checkGrid(grid);
        if (mode < 0 || mode > 4) throw Error(SDFO_ERR_INVALID_ARGUMENT, "invalid query mode");
        if (n < 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "negative point count");
        if (n > 0 && (!points || !gap || !grad3 || !valid))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL point or output pointer");
        DeviceGuard dg(grid->device);
        launchQueryF32(*grid, mode, points, n, gap, grad3, valid, normal3, witness3, branch,
                       hipStream_t(stream));
    });
}

int sdfo_query_f64(const sdfo_grid* grid, int32_t mode, const double* points, int64_t n,
                   double* gap, double* grad3, uint8_t* valid, double* normal3, double* witness3,
                   int32_t* branch, sdfo_stream stream)
{
    return guarded([&] {
        checkGrid(grid);
        if (mode < 0 || mode > 4) throw Error(SDFO_ERR_INVALID_ARGUMENT, "invalid query mode");
        if (n < 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "negative point count");
        if (n > 0 && (!points || !gap || !grad3 || !valid))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL point or output pointer");
        DeviceGuard dg(grid->device);
        launchQueryF64(*grid, mode, points, n, gap, grad3, valid, normal3, witness3, branch,
                       hipStream_t(stream));
    });
}

static void checkContact(const sdfo_grid* grid, const double* X, const double* w, int64_t n,
                         const double* poses, int64_t n_poses, const sdfo_contact_params* p)
{
    checkGrid(grid);
    if (!p) throw Error(SDFO_ERR_INVALID_ARGUMENT, "params is NULL");
    if (p->mode != SDFO_TRILINEAR && p->mode != SDFO_TRICUBIC)
        throw Error(SDFO_ERR_UNSUPPORTED, "contact reduction supports Trilinear and Tricubic fields");
    if (p->order < 0 || p->order > 8) throw Error(SDFO_ERR_INVALID_ARGUMENT, "order must be in [0, 8]");
    if (n < 0 || n_poses < 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "negative size");
    if (n_poses > 65535) throw Error(SDFO_ERR_INVALID_ARGUMENT, "at most 65535 poses per call");
    if ((n > 0 && (!X || !w)) || (n_poses > 0 && !poses))
        throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL input pointer");
}

int sdfo_contact_forward(const sdfo_grid* grid, const double* X, const double* w, int64_t n,
                         const double* poses, int64_t n_poses, const sdfo_contact_params* params,
                         double* out, int32_t* active, sdfo_stream stream)
{
    return guarded([&] {
        checkContact(grid, X, w, n, poses, n_poses, params);
        if (n_poses > 0 && !out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        DeviceGuard dg(grid->device);
        launchContactForward(*grid, X, w, n, poses, n_poses, *params, out, active,
                             hipStream_t(stream));
    });
}

int sdfo_contact_linearize(const sdfo_grid* grid, const double* X, int64_t n, const double* pose,
                           const sdfo_contact_params* params, double* g0, double* grad0,
                           uint8_t* valid0, sdfo_stream stream)
{
    return guarded([&] {
        checkContact(grid, X, X, n, pose, 1, params);
        if (n > 0 && (!g0 || !grad0 || !valid0))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "the cache arrays are required");
        DeviceGuard dg(grid->device);
        launchContactLinearize(*grid, X, n, pose, *params, g0, grad0, valid0, hipStream_t(stream));
    });
}

int sdfo_mesh_build(const double* verts, int64_t n_verts, const int32_t* tris, int64_t n_tris,
                    sdfo_mesh** out)
{
    return guarded([&] {
        if (!out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        *out = nullptr;
        if (n_verts < 0 || n_tris < 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "negative size");
        if ((n_verts > 0 && !verts) || (n_tris > 0 && !tris))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL mesh array");
        if (n_tris > (int64_t(1) << 30)) throw Error(SDFO_ERR_INVALID_ARGUMENT, "too many triangles");
        auto m = std::make_unique<sdfo_mesh>();
        m->device = currentDevice();
        m->ntris = n_tris;
        const HostBvh hb = buildHostBvh(verts, n_verts, tris, n_tris);
        m->bvh.upload(hb, nullptr);
        SDFO_HIP(hipStreamSynchronize(nullptr));
        *out = m.release();
    });
}

void sdfo_mesh_free(sdfo_mesh* mesh) { delete mesh; }

int sdfo_mesh_triangle_count(const sdfo_mesh* mesh, int64_t* count)
{
    return guarded([&] {
        if (!mesh || !count) throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL argument");
        *count = mesh->ntris;
    });
}

int sdfo_mesh_query(const sdfo_mesh* mesh, const double* points, int64_t n,
                    int32_t signed_distance, double* distance, double* closest3, double* normal3,
                    int32_t* triangle, sdfo_stream stream)
{
    return guarded([&] {
        if (!mesh) throw Error(SDFO_ERR_INVALID_ARGUMENT, "mesh is NULL");
        if (n < 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "negative size");
        if (n > 0 && !points) throw Error(SDFO_ERR_INVALID_ARGUMENT, "points is NULL");
        DeviceGuard dg(mesh->device);
        launchMeshQuery(mesh->bvh, signed_distance != 0, points, n, distance, closest3, normal3,
                        triangle, hipStream_t(stream));
    });
}

int sdfo_mesh_query_host(const sdfo_mesh* mesh, const double* points, int64_t n,
                         int32_t signed_distance, double* distance, double* closest3,
                         double* normal3, int32_t* triangle, sdfo_stream stream)
{
    return guarded([&] {
        if (!mesh) throw Error(SDFO_ERR_INVALID_ARGUMENT, "mesh is NULL");
        if (n < 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "negative size");
        if (n == 0) return;
        if (!points) throw Error(SDFO_ERR_INVALID_ARGUMENT, "points is NULL");
        DeviceGuard dg(mesh->device);
        const hipStream_t s = hipStream_t(stream);
        const size_t N = size_t(n);
        // points, distance, closest point, normal, triangle in one device block
        const size_t oP = 0, oD = oP + 24 * N, oC = oD + 8 * N, oN = oC + 24 * N, oT = oN + 24 * N;
        char* base = nullptr;
        SDFO_HIP(hipMallocAsync(reinterpret_cast<void**>(&base), oT + 4 * N, s));
        try
        {
            SDFO_HIP(hipMemcpyAsync(base + oP, points, 24 * N, hipMemcpyHostToDevice, s));
            launchMeshQuery(mesh->bvh, signed_distance != 0, reinterpret_cast<double*>(base + oP),
                            n, reinterpret_cast<double*>(base + oD),
                            reinterpret_cast<double*>(base + oC),
                            reinterpret_cast<double*>(base + oN),
                            reinterpret_cast<int32_t*>(base + oT), s);
            if (distance) SDFO_HIP(hipMemcpyAsync(distance, base + oD, 8 * N, hipMemcpyDeviceToHost, s));
            if (closest3) SDFO_HIP(hipMemcpyAsync(closest3, base + oC, 24 * N, hipMemcpyDeviceToHost, s));
            if (normal3) SDFO_HIP(hipMemcpyAsync(normal3, base + oN, 24 * N, hipMemcpyDeviceToHost, s));
            if (triangle) SDFO_HIP(hipMemcpyAsync(triangle, base + oT, 4 * N, hipMemcpyDeviceToHost, s));
            SDFO_HIP(hipFreeAsync(base, s));
            SDFO_HIP(hipStreamSynchronize(s));
        }
        catch (...)
        {
            (void)hipStreamSynchronize(s);
            (void)hipFreeAsync(base, s);
            throw;
        }
    });
}

static CullArgs cullArgs(const sdfo_point_bvh* bvh, const double* boxes6, int64_t n,
                         double margin)
{
    if (!boxes6) throw Error(SDFO_ERR_INVALID_ARGUMENT, "boxes6 is NULL");
    if (bvh->n != n)
        throw Error(SDFO_ERR_INVALID_ARGUMENT, "the point BVH was built over " +
                                                   std::to_string(bvh->n) + " points, not " +
                                                   std::to_string(n));
    CullArgs c;
    c.leafBox = bvh->leafBox;
    c.leafOf = bvh->leafOf;
    c.nLeaves = bvh->nLeaves;
    c.boxes = boxes6;
    c.margin = margin;
    return c;
}

int sdfo_point_bvh_build(const double* points, int64_t n, int32_t leaf_size, sdfo_point_bvh** out)
{
    return guarded([&] {
        if (!out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        *out = nullptr;
        if (n < 0 || n > INT32_MAX) throw Error(SDFO_ERR_INVALID_ARGUMENT, "n must be in [0, 2^31)");
        if (n > 0 && !points) throw Error(SDFO_ERR_INVALID_ARGUMENT, "points is NULL");
        if (leaf_size < 1) throw Error(SDFO_ERR_INVALID_ARGUMENT, "leaf_size must be positive");
        auto b = std::make_unique<sdfo_point_bvh>();
        buildPointBvh(*b, points, n, leaf_size);
        *out = b.release();
    });
}

void sdfo_point_bvh_free(sdfo_point_bvh* bvh) { delete bvh; }

int sdfo_point_bvh_get_info(const sdfo_point_bvh* bvh, int64_t* n_points, int64_t* n_nodes,
                            int64_t* n_leaves)
{
    return guarded([&] {
        if (!bvh) throw Error(SDFO_ERR_INVALID_ARGUMENT, "bvh is NULL");
        if (n_points) *n_points = bvh->n;
        if (n_nodes) *n_nodes = int64_t(bvh->nodes.size());
        if (n_leaves) *n_leaves = bvh->nLeaves;
    });
}

int sdfo_point_bvh_query(const sdfo_point_bvh* bvh, const double* qmin3, const double* qmax3,
                         double margin, int32_t* indices, int64_t* count)
{
    return guarded([&] {
        if (!bvh || !qmin3 || !qmax3 || !count)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL argument");
        const std::vector<int32_t> r = queryPointBvh(*bvh, qmin3, qmax3, margin);
        *count = int64_t(r.size());
        if (!r.empty())
        {
            if (!indices) throw Error(SDFO_ERR_INVALID_ARGUMENT, "indices is NULL");
            std::memcpy(indices, r.data(), r.size() * sizeof(int32_t));
        }
    });
}

int sdfo_point_bvh_cull(const sdfo_point_bvh* bvh, const double* boxes6, int64_t n_boxes,
                        double margin, uint8_t* mask, sdfo_stream stream)
{
    return guarded([&] {
        if (!bvh) throw Error(SDFO_ERR_INVALID_ARGUMENT, "bvh is NULL");
        if (n_boxes < 0 || n_boxes > 65535)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "n_boxes must be in [0, 65535]");
        if (n_boxes == 0 || bvh->n == 0) return;
        if (!mask) throw Error(SDFO_ERR_INVALID_ARGUMENT, "mask is NULL");
        pointBvhToDevice(*bvh, currentDevice());
        launchCullMask(cullArgs(bvh, boxes6, bvh->n, margin), bvh->n, n_boxes, mask,
                       hipStream_t(stream));
    });
}

int sdfo_contact_forward_culled(const sdfo_grid* grid, const sdfo_point_bvh* bvh,
                                const double* boxes6, const double* X, const double* w, int64_t n,
                                const double* poses, int64_t n_poses,
                                const sdfo_contact_params* params, double* out, int32_t* active,
                                sdfo_stream stream)
{
    return guarded([&] {
        checkContact(grid, X, w, n, poses, n_poses, params);
        if (n_poses > 0 && !out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        if (!bvh) throw Error(SDFO_ERR_INVALID_ARGUMENT, "bvh is NULL");
        DeviceGuard dg(grid->device);
        pointBvhToDevice(*bvh, grid->device);
        const CullArgs c = cullArgs(bvh, boxes6, n, 0.0);
        launchContactForward(*grid, X, w, n, poses, n_poses, *params, out, active,
                             hipStream_t(stream), &c);
    });
}

int sdfo_contact_evaluate(const sdfo_grid* grid, const sdfo_point_bvh* bvh, const double* X,
                          const double* w, int64_t n, const double* states, const double* frames,
                          int64_t n_pairs, const sdfo_contact_params* params,
                          const double* body_boxes9, double* out, int32_t* active,
                          sdfo_stream stream)
{
    return guarded([&] {
        checkContact(grid, X, w, n, states, n_pairs, params);
        if (n_pairs > 0 && !out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        DeviceGuard dg(grid->device);
        CullArgs c;
        if (bvh)
        {
            pointBvhToDevice(*bvh, grid->device);
            // the boxes are computed on the device; this only checks the point count
            c = cullArgs(bvh, X, n, 0.0);
        }
        launchContactEvaluate(*grid, X, w, n, states, frames, n_pairs, *params, bvh ? &c : nullptr,
                              body_boxes9, out, active, hipStream_t(stream));
    });
}

int sdfo_contact_evaluate_noalloc(const sdfo_grid* grid, const sdfo_point_bvh* bvh,
                                  const double* X, const double* w, int64_t n,
                                  const double* states, const double* frames, int64_t n_pairs,
                                  const sdfo_contact_params* params, const double* body_boxes9,
                                  double* out, int32_t* active, sdfo_stream stream)
{
    return guarded([&] {
        checkContact(grid, X, w, n, states, n_pairs, params);
        if (n_pairs > 0 && !out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        DeviceGuard dg(grid->device);
        CullArgs c;
        if (bvh)
        {
            // uploading the tree would allocate: it must already be on this device
            if (bvh->device != grid->device || !bvh->leafBox || !bvh->leafOf)
                throw Error(SDFO_ERR_INVALID_ARGUMENT,
                            "the point BVH is not on the grid's device yet; one "
                            "sdfo_contact_evaluate or sdfo_point_bvh_cull call with it uploads it");
            c = cullArgs(bvh, X, n, 0.0);
        }
        launchContactEvaluateNoAlloc(*grid, X, w, n, states, frames, n_pairs, *params,
                                     bvh ? &c : nullptr, body_boxes9, out, active,
                                     hipStream_t(stream));
    });
}

int sdfo_contact_forward_linearized(const double* g0, const double* grad0, const uint8_t* valid0,
                                    const double* t_ref3, const double* X, const double* w, int64_t n,
                                    const double* poses, int64_t n_poses,
                                    const sdfo_contact_params* params, const sdfo_point_bvh* bvh,
                                    const double* boxes6, double* out, int32_t* active,
                                    sdfo_stream stream)
{
    return guarded([&] {
        if (!params) throw Error(SDFO_ERR_INVALID_ARGUMENT, "params is NULL");
        if (n < 0 || n_poses < 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "negative size");
        if (n > 0 && (!g0 || !grad0 || !valid0 || !X || !w))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "cache and quadrature arrays are required");
        if (n_poses > 0 && (!poses || !out || !t_ref3))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "poses, t_ref and out are required");
        if (params->order < 0 || params->order > 8)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "order must be in [0, 8]");
        if (n_poses > 65535) throw Error(SDFO_ERR_INVALID_ARGUMENT, "at most 65535 poses per call");
        CullArgs c;
        if (bvh)
        {
            pointBvhToDevice(*bvh, currentDevice());
            c = cullArgs(bvh, boxes6, n, 0.0);
        }
        launchContactForwardLinearized(g0, grad0, valid0, t_ref3, X, w, n, poses, n_poses, *params,
                                       out, active, hipStream_t(stream), bvh ? &c : nullptr);
    });
}

int sdfo_contact_backward(const sdfo_grid* grid, const double* X, const double* w, int64_t n,
                          const double* poses, int64_t n_poses, const sdfo_contact_params* params,
                          const double* g_out, double* g_twist, double* g_params, double* g_X,
                          double* g_w, sdfo_stream stream)
{
    return guarded([&] {
        checkContact(grid, X, w, n, poses, n_poses, params);
        if (n_poses > 0 && !g_out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "g_out is NULL");
        DeviceGuard dg(grid->device);
        launchContactBackward(*grid, X, w, n, poses, n_poses, *params, g_out, g_twist, g_params,
                              g_X, g_w, hipStream_t(stream));
    });
}

int64_t sdfo_contact_saved_bytes(int64_t n, int64_t n_poses)
{
    if (n < 0 || n_poses < 0) return -1;
    return contactSavedBytes(n, n_poses);
}

int sdfo_contact_forward_save(const sdfo_grid* grid, const double* X, const double* w, int64_t n,
                              const double* poses, int64_t n_poses,
                              const sdfo_contact_params* params, double* out, int32_t* active,
                              void* saved, int64_t saved_bytes, sdfo_stream stream)
{
    return guarded([&] {
        checkContact(grid, X, w, n, poses, n_poses, params);
        if (n_poses > 0 && !out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        if (saved_bytes != contactSavedBytes(n, n_poses))
            throw Error(SDFO_ERR_INVALID_ARGUMENT,
                        "saved_bytes must be sdfo_contact_saved_bytes(n, n_poses)");
        if (n_poses > 0 && !saved) throw Error(SDFO_ERR_INVALID_ARGUMENT, "saved is NULL");
        DeviceGuard dg(grid->device);
        launchContactForward(*grid, X, w, n, poses, n_poses, *params, out, active,
                             hipStream_t(stream), nullptr, saved);
    });
}

int sdfo_contact_backward_saved(const sdfo_grid* grid, const double* X, const double* w, int64_t n,
                                const double* poses, int64_t n_poses,
                                const sdfo_contact_params* params, const double* g_out,
                                const void* saved, int64_t saved_bytes, double* g_twist,
                                double* g_params, double* g_X, double* g_w, sdfo_stream stream)
{
    return guarded([&] {
        checkContact(grid, X, w, n, poses, n_poses, params);
        if (n_poses > 0 && !g_out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "g_out is NULL");
        if (saved_bytes != contactSavedBytes(n, n_poses))
            throw Error(SDFO_ERR_INVALID_ARGUMENT,
                        "saved_bytes must be sdfo_contact_saved_bytes(n, n_poses)");
        if (n_poses > 0 && !saved) throw Error(SDFO_ERR_INVALID_ARGUMENT, "saved is NULL");
        DeviceGuard dg(grid->device);
        launchContactBackward(*grid, X, w, n, poses, n_poses, *params, g_out, g_twist, g_params,
                              g_X, g_w, hipStream_t(stream), saved);
    });
}

} // extern "C"

// Contact-aware build, batched builds and queries, autotuner.
extern "C" {

int sdfo_build_contact_aware(const double* verts, int64_t n_verts, const int32_t* tris,
                             int64_t n_tris, const double* corner_normals,
                             const sdfo_build_params* params, int32_t sparse, sdfo_grid** out,
                             sdfo_stream stream)
{
    return guarded([&] {
        if (!out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        *out = nullptr;
        checkParams(params);
        *out = buildContactAware(verts, n_verts, tris, n_tris, corner_normals, *params, sparse != 0,
                                 hipStream_t(stream));
    });
}

int sdfo_build_dense_batched(int32_t n_meshes, const double* verts, const int64_t* vert_offsets,
                             const int32_t* tris, const int64_t* tri_offsets,
                             const sdfo_build_params* params, sdfo_grid** out, sdfo_stream stream)
{
    return guarded([&] {
        if (!out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        if (n_meshes < 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "n_meshes must be >= 0");
        for (int32_t m = 0; m < n_meshes; ++m) out[m] = nullptr;
        if (n_meshes == 0) return;
        if (!verts || !vert_offsets || !tris || !tri_offsets || !params)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "batched build: NULL array");
        for (int32_t m = 0; m < n_meshes; ++m) checkParams(&params[m]);
        std::vector<std::unique_ptr<sdfo_grid>> grids =
            buildDenseBatched(n_meshes, verts, vert_offsets, tris, tri_offsets, params,
                              hipStream_t(stream));
        for (int32_t m = 0; m < n_meshes; ++m) out[m] = grids[size_t(m)].release();
    });
}

namespace {
// Pinned, device-mapped, coherent host memory per calling thread for the
// host-pointer FP64 query: calls whose arrays fit kHostStageBytes run the
// kernels directly on it (they read the points and write the results over
// PCIe), so a call costs one launch and one synchronisation plus two host
// copies, instead of a device allocation and one DMA per array.  Each call
// synchronises before it returns, so one buffer per thread is never shared
// by two calls in flight.
constexpr size_t kHostStageBytes = size_t(8) << 20;

struct HostStage
{
    char* host = nullptr;
    char* dev = nullptr;
    size_t cap = 0;
    ~HostStage()
    {
        if (host) (void)hipHostFree(host);
    }
    char* reserve(size_t bytes)
    {
        if (bytes <= cap) return dev;
        if (host) SDFO_HIP(hipHostFree(host));
        host = dev = nullptr;
        cap = 0;
        const size_t want = std::max(bytes, size_t(64) << 10);
        SDFO_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), want,
                               hipHostMallocMapped | hipHostMallocPortable | hipHostMallocCoherent));
        void* d = nullptr;
        SDFO_HIP(hipHostGetDevicePointer(&d, host, 0));
        dev = static_cast<char*>(d);
        cap = want;
        return dev;
    }
};
thread_local HostStage g_hostStage;
} // namespace

int sdfo_query_host_f64(const sdfo_grid* grid, int32_t mode, const double* points, int64_t n,
                        double* gap, double* normal3, double* witness3, uint8_t* valid,
                        int32_t* branch, uint8_t* cand_count, double* cand_gap,
                        double* cand_normal3, double* cand_witness3, int32_t* cand_branch,
                        sdfo_stream stream)
{
    return guarded([&] {
        if (!grid) throw Error(SDFO_ERR_INVALID_ARGUMENT, "grid is NULL");
        if (mode < SDFO_FIRST_ORDER || mode > SDFO_CONTACT_AWARE)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "unknown query mode");
        if (n < 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "n must be >= 0");
        if (n == 0) return;
        if (!points || !gap || !normal3 || !witness3 || !valid)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "points, gap, normal3, witness3, valid are required");
        const bool cands = cand_count != nullptr;
        if (cands && (!cand_gap || !cand_normal3 || !cand_witness3 || !cand_branch))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "candidate outputs must all be given");
        DeviceGuard dg(grid->device);
        const hipStream_t s = hipStream_t(stream);
        const size_t N = size_t(n);
        // one block for every array: points, gap, grad, normal, witness,
        // branch, valid, then the candidate arrays
        size_t off = 0;
        auto carve = [&](size_t bytes) {
            const size_t o = off;
            off += (bytes + 255) & ~size_t(255);
            return o;
        };
        const size_t oP = carve(24 * N), oG = carve(8 * N), oGr = carve(24 * N), oN = carve(24 * N),
                     oW = carve(24 * N), oB = carve(4 * N), oV = carve(N);
        size_t oCC = 0, oCG = 0, oCN = 0, oCW = 0, oCB = 0;
        if (cands)
        {
            oCC = carve(N);
            oCG = carve(64 * N);
            oCN = carve(192 * N);
            oCW = carve(192 * N);
            oCB = carve(32 * N);
        }
        const bool mapped = off <= kHostStageBytes;
        char* base = nullptr;
        if (mapped)
            base = g_hostStage.reserve(off);
        else
            SDFO_HIP(hipMallocAsync(reinterpret_cast<void**>(&base), off, s));
        char* const host = mapped ? g_hostStage.host : nullptr;
        auto D = [&](size_t o) { return base + o; };
        auto run = [&] {
            if (mapped)
            {
                std::memcpy(host + oP, points, 24 * N);
                if (cands) std::memset(host + oCC, 0, N);
            }
            else
            {
                SDFO_HIP(hipMemcpyAsync(D(oP), points, 24 * N, hipMemcpyHostToDevice, s));
                if (cands) SDFO_HIP(hipMemsetAsync(D(oCC), 0, N, s));
            }
            launchQueryF64(*grid, mode, reinterpret_cast<double*>(D(oP)), n,
                           reinterpret_cast<double*>(D(oG)), reinterpret_cast<double*>(D(oGr)),
                           reinterpret_cast<uint8_t*>(D(oV)), reinterpret_cast<double*>(D(oN)),
                           reinterpret_cast<double*>(D(oW)), reinterpret_cast<int32_t*>(D(oB)), s);
            if (cands && mode == SDFO_CONTACT_AWARE)
                launchCandidatesF64(*grid, reinterpret_cast<double*>(D(oP)), n,
                                    reinterpret_cast<uint8_t*>(D(oCC)),
                                    reinterpret_cast<double*>(D(oCG)),
                                    reinterpret_cast<double*>(D(oCN)),
                                    reinterpret_cast<double*>(D(oCW)),
                                    reinterpret_cast<int32_t*>(D(oCB)),
                                    reinterpret_cast<double*>(D(oG)),
                                    reinterpret_cast<double*>(D(oN)),
                                    reinterpret_cast<double*>(D(oW)),
                                    reinterpret_cast<uint8_t*>(D(oV)), s);
            struct Out
            {
                void* dst;
                size_t o, bytes;
            };
            const Out outs[] = {{gap, oG, 8 * N},
                                {normal3, oN, 24 * N},
                                {witness3, oW, 24 * N},
                                {valid, oV, N},
                                {branch, oB, 4 * N},
                                {cand_count, oCC, N},
                                {cand_gap, oCG, 64 * N},
                                {cand_normal3, oCN, 192 * N},
                                {cand_witness3, oCW, 192 * N},
                                {cand_branch, oCB, 32 * N}};
            const size_t nOut = cands ? 10 : 5;
            if (!mapped)
                for (size_t k = 0; k < nOut; ++k)
                    if (outs[k].dst)
                        SDFO_HIP(hipMemcpyAsync(outs[k].dst, D(outs[k].o), outs[k].bytes,
                                                hipMemcpyDeviceToHost, s));
            SDFO_HIP(hipStreamSynchronize(s));
            if (mapped)
                for (size_t k = 0; k < nOut; ++k)
                    if (outs[k].dst) std::memcpy(outs[k].dst, host + outs[k].o, outs[k].bytes);
        };
        try
        {
            run();
        }
        catch (...)
        {
            (void)hipStreamSynchronize(s);
            if (!mapped) (void)hipFreeAsync(base, s);
            throw;
        }
        if (!mapped) SDFO_HIP(hipFreeAsync(base, s));
    });
}

int sdfo_query_plan_create(int32_t n_streams, const sdfo_grid* const* grids,
                           const int64_t* offsets, sdfo_query_plan** out)
{
    return guarded([&] {
        if (!out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "out is NULL");
        *out = nullptr;
        checkPlanArgs(n_streams, grids, offsets);
        auto p = std::make_unique<sdfo_query_plan>();
        p->device = currentDevice();
        p->plan = makeQueryPlan(n_streams, grids, offsets);
        *out = p.release();
    });
}

int sdfo_query_plan_run(const sdfo_query_plan* plan, int32_t mode, const float* points,
                        float* gap, float* grad3, uint8_t* valid, sdfo_stream stream)
{
    return guarded([&] {
        if (!plan) throw Error(SDFO_ERR_INVALID_ARGUMENT, "plan is NULL");
        if (mode < SDFO_FIRST_ORDER || mode > SDFO_CONTACT_AWARE)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "unknown query mode");
        if (plan->plan->points > 0 && (!points || !gap || !grad3 || !valid))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "points, gap, grad3 and valid are required");
        DeviceGuard dg(plan->device);
        runPlanTuned(*plan->plan, mode, points, gap, grad3, valid, hipStream_t(stream));
    });
}

void sdfo_query_plan_free(sdfo_query_plan* plan)
{
    if (!plan) return;
    try
    {
        DeviceGuard dg(plan->device);
        delete plan;
    }
    catch (...)
    {
        delete plan;
    }
}

int sdfo_query_batched(int32_t n_streams, const sdfo_grid* const* grids, int32_t mode,
                       const int64_t* offsets, const float* points, float* gap, float* grad3,
                       uint8_t* valid, sdfo_stream stream)
{
    return guarded([&] {
        if (mode < SDFO_FIRST_ORDER || mode > SDFO_CONTACT_AWARE)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "unknown query mode");
        checkPlanArgs(n_streams, grids, offsets);
        const auto plan = makeQueryPlan(n_streams, grids, offsets);
        if (plan->points > 0 && (!points || !gap || !grad3 || !valid))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "points, gap, grad3 and valid are required");
        runPlanTuned(*plan, mode, points, gap, grad3, valid, hipStream_t(stream));
        SDFO_HIP(hipStreamSynchronize(hipStream_t(stream)));
    });
}

int sdfo_tuner_set_mode(int32_t mode)
{
    return guarded([&] {
        if (mode != SDFO_TUNE_OFF && mode != SDFO_TUNE_ON_FIRST_USE)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "unknown tuner mode");
        tunerSetMode(mode);
    });
}

int sdfo_tuner_reset(void)
{
    return guarded([&] { tunerReset(); });
}

int64_t sdfo_tuner_report(char* buf, int64_t cap)
{
    const std::string r = tunerReport();
    if (buf && cap > 0)
    {
        const size_t n = std::min<size_t>(r.size(), size_t(cap - 1));
        std::memcpy(buf, r.data(), n);
        buf[n] = '\0';
    }
    return int64_t(r.size());
}

} // extern "C"
