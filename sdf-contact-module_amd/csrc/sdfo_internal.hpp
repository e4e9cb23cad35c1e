// Internal types shared by the host code and the HIP kernels of libsdfo.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sdfo.h"

namespace sdfo {

// Error carried to the C ABI: code + the reference's message text.
struct Error : public std::runtime_error
{
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

void setLastError(const std::string& msg);

#define SDFO_HIP(call)                                                                         \
    do                                                                                         \
    {                                                                                          \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::sdfo::Error(SDFO_ERR_DEVICE, std::string(#call) + ": " +                   \
                                                     hipGetErrorString(e_));                   \
    } while (0)

// Cell-centred grid (GridSpec.h:17-60).
struct GridSpec
{
    double bmin[3] = {0, 0, 0};
    double bmax[3] = {1, 1, 1};
    double voxel[3] = {1, 1, 1};
    int n[3] = {1, 1, 1};
    __host__ __device__ int64_t total() const { return int64_t(n[0]) * n[1] * n[2]; }
    __host__ __device__ double hmax() const
    {
        double h = voxel[0];
        if (h < voxel[1]) h = voxel[1];
        if (h < voxel[2]) h = voxel[2];
        return h;
    }
};

GridSpec setupGrid(const double* verts, int64_t nv, const sdfo_build_params& p);

// Device BVH node: own box + links (64 bytes, one half cache line).
struct alignas(16) BvhNode
{
    double bmin[3];
    double bmax[3];
    int32_t left, right, triStart, triCount;
};
static_assert(sizeof(BvhNode) == 64, "BvhNode layout");

// Node count of the tree KdTree::buildRecursive makes over c triangles
// (KdTree.h:188-242): the split is always at c / 2 and leaves hold <= 8, so
// the shape depends on c only.  All nodes at depth d hold floor(c / 2^d) or
// one more, so the count is carried bottom-up for that pair of sizes:
// O(log c) instead of a walk over the whole subtree.
__host__ __device__ inline int64_t medianTreeNodes(int64_t c)
{
    if (c <= 8) return 1;
    int D = 0;
    while ((c >> D) + 1 > 8) ++D;
    int64_t fa = 1, fa1 = 1; // nodes over (c >> D) and (c >> D) + 1 items: single leaves
    for (int d = D - 1; d >= 0; --d)
    {
        const int64_t a = c >> d, b = a >> 1; // children sizes are b or b + 1
        const int64_t fb = fa, fb1 = fa1;
        const int64_t ha = a >> 1, ha1 = (a + 1) >> 1;
        fa = a <= 8 ? 1 : 1 + (ha == b ? fb : fb1) + ((a - ha) == b ? fb : fb1);
        fa1 = a + 1 <= 8 ? 1 : 1 + (ha1 == b ? fb : fb1) + ((a + 1 - ha1) == b ? fb : fb1);
    }
    return fa;
}

// Host BVH, built exactly like KdTree::build (KdTree.h:91-121, 188-242) and
// flattened: nodes in creation (pre-)order, triangles stored per BVH slot.
struct HostBvh
{
    std::vector<BvhNode> nodes;
    std::vector<double> slotTri9; // 9 doubles (a, b, c) per slot
    std::vector<int32_t> slotTri; // original triangle index per slot
    int depth = 0;
    double rayDir[3] = {0, 0, 0};
};

HostBvh buildHostBvh(const double* verts, int64_t nv, const int32_t* tris, int64_t nt);
// The same tree written into caller arrays (bvhNodeCount(nt) nodes, nt slots)
// with node links and slot ranges offset by the bases; returns the depth.
int buildHostBvhInto(const double* verts, int64_t nv, const int32_t* tris, int64_t nt,
                     BvhNode* nodes, double* slotTri9, int32_t* slotTri, int nodeBase, int slotBase,
                     bool parallel);
int64_t bvhNodeCount(int64_t nt);
// Only the slot order of that tree (slotTri[s] = triangle of slot s) and its
// depth: the node boxes and slot triangles are assembled on the device
// (DeviceBvh::assemble), since the tree's shape follows from nt alone.
int buildHostBvhOrder(const double* verts, int64_t nv, const int32_t* tris, int64_t nt,
                      int32_t* slotTri, bool parallel);
void setRayDir(HostBvh& h);
// The same tree in the calling thread's reusable host buffers (valid until
// that thread's next scratch build): no zero fill, no fresh pages per build.
struct HostBvhView
{
    const BvhNode* nodes;
    int64_t nNodes;
    const double* slotTri9;
    const int32_t* slotTri;
    int64_t nSlots;
    int depth;
    double rayDir[3];
};
HostBvhView buildHostBvhScratch(const double* verts, int64_t nv, const int32_t* tris, int64_t nt);
// Only the slot order and depth (buildHostBvhOrder), in the calling thread's
// scratch arrays (valid until its next scratch build).
struct HostBvhOrderView
{
    const int32_t* slotTri;
    int64_t nSlots;
    int depth;
    double rayDir[3];
};
HostBvhOrderView buildHostBvhOrderScratch(const double* verts, int64_t nv, const int32_t* tris,
                                          int64_t nt);
// Frees the calling thread's scratch arrays if they hold more than keepTris
// triangles' worth.
void releaseHostBvhScratch(int64_t keepTris);
// A task for the library's persistent host worker threads.  poolWait runs
// queued tasks until t is done, then rethrows what t threw.
struct PoolTask
{
    std::function<void()> fn;
    std::atomic<bool> done{false};
    std::exception_ptr error;
};
void poolSubmit(PoolTask* t);
void poolWait(PoolTask* t);

// Host worker threads: the CPUs this process may run on, capped by
// OMP_NUM_THREADS when set and by 64.
int hostThreads();

// Runs f(begin, end) over [0, n) in up to 16 chunks of at least `grain`
// items on the worker threads (the calling thread takes the first chunk).
void parallelRange(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& f);

struct BvhMeshPart;

// Device copy of a BVH.
struct DeviceBvh
{
    BvhNode* nodes = nullptr;
    double* tri9 = nullptr;
    double* box6 = nullptr;
    int32_t* triIdx = nullptr;
    int nnodes = 0;
    int64_t ntris = 0;
    int depth = 0;
    double rayDir[3] = {0, 0, 0};
    hipStream_t stream = nullptr; // the block is a stream-ordered allocation on it
    // The four arrays share one device block; reserve() allocates it ahead of
    // the upload (for a tree of nNodes nodes over nSlots triangle slots), so
    // the allocation can overlap the host build of the tree.
    void* block = nullptr;
    int64_t capNodes = 0, capSlots = 0;
    void reserve(int64_t nNodes, int64_t nSlots, hipStream_t s);
    void upload(const HostBvh& h, hipStream_t s);
    // Assembly on the device.  reserveAssembly allocates the block with room
    // for the meshes' vertices and triangles too; stageMeshes uploads them
    // (parts give each mesh's slot, node, vertex and triangle offsets, all
    // meshes' slots concatenated in part order); assemble uploads the slot
    // order (slotTri[s] = the mesh-local triangle of slot s, as
    // buildHostBvhOrder makes it) and builds nodes, slot triangles and slot
    // boxes in one kernel, equal to the host tree bit for bit.
    void reserveAssembly(int64_t nNodes, int64_t nSlots, int64_t nVerts, int nParts, hipStream_t s);
    void stageMeshes(const double* verts, int64_t nVerts, const int32_t* tris, int64_t nTris,
                     const BvhMeshPart* parts, int nParts, hipStream_t s);
    void assemble(const int32_t* slotTri, int treeDepth, const double* dir, hipStream_t s);
    double* verts = nullptr;
    int32_t* tris = nullptr;
    int32_t* flags = nullptr;
    BvhMeshPart* parts = nullptr;
    int nParts = 0;
    int64_t nSubtrees = 0; // assembly workgroups of the staged meshes
    int topLevelsMax = 0;
    int64_t capVerts = 0;
    int capParts = 0;
    // The same from raw host arrays; the slot boxes are derived on the device.
    void upload(const BvhNode* hostNodes, int64_t nNodes, const double* slotTri9,
                const int32_t* slotTri, int64_t nSlots, int treeDepth, const double* dir,
                hipStream_t s);
    void release();
};

// Slot boxes (bmin3, bmax3 of each slot triangle's vertices) on the device (build.hip).
void launchSlotBoxes(const double* tri9, int64_t n, double* box6, hipStream_t s);

// One mesh of a device BVH assembly: its slots [slotBase, slotBase + nSlots)
// and nodes from nodeBase in the concatenated arrays, its vertices and
// triangles from vertBase / triBase in the staged arrays (triangle indices
// are local to the mesh's vertices).
struct BvhMeshPart
{
    int64_t slotBase, nSlots, nodeBase, vertBase, triBase;
    // Filled by DeviceBvh::stageMeshes: the mesh's tree is assembled as
    // 2^topLevels subtrees (workgroups subBase ...) plus the nodes above
    // them; f0[d] / f1[d] are the node counts of the subtrees over
    // nSlots >> d and (nSlots >> d) + 1 slots, the two sizes at depth d.
    int64_t subBase = 0;
    int32_t topLevels = 0, pad = 0;
    int32_t f0[32] = {}, f1[32] = {};
};
// The assembly cuts a mesh's tree into at most 2^kMaxTopLevels subtrees, of
// at least kMinSubtreeSlots slots when the mesh is larger.
constexpr int kMaxTopLevels = 8;
constexpr int64_t kMinSubtreeSlots = 256;
// Nodes, slot triangles and slot boxes of db from its staged meshes and slot
// order; db.flags must be zero (build.hip).
void launchBvhAssemble(const DeviceBvh& db, hipStream_t s);

// Nearest / signed nearest triangle of a point stream (build.hip).
void launchMeshQuery(const DeviceBvh& bvh, bool signedDist, const double* pts, int64_t n,
                     double* dist, double* cp3, double* n3, int32_t* tri, hipStream_t s);

// Maximal BVH traversal stack per lane (LDS).  Host rejects deeper trees.
constexpr int kStackDepth = 32;

} // namespace sdfo

// A triangle mesh's BVH on a device (the KdTree of the C ABI).
struct sdfo_mesh
{
    int device = 0;
    int64_t ntris = 0;
    sdfo::DeviceBvh bvh;
    sdfo_mesh() = default;
    sdfo_mesh(const sdfo_mesh&) = delete;
    sdfo_mesh& operator=(const sdfo_mesh&) = delete;
    ~sdfo_mesh() { bvh.release(); }
};

// The opaque handle of the C ABI.
// A device block shared by the grids of one batched build, returned to the
// pool when the last of them is released.  A grid released stream-ordered
// records an event on its stream; the block is then freed on the last
// release's stream after every recorded event, without a device-wide
// synchronisation.  Any synchronous release makes the block's release
// synchronise the device.
struct ArenaBlock
{
    char* base = nullptr;
    std::mutex m;
    std::vector<hipEvent_t> events;
    hipStream_t last = nullptr;
    bool ordered = true;
    void noteRelease(bool streamOrdered, hipStream_t s)
    {
        std::lock_guard<std::mutex> lk(m);
        if (!streamOrdered)
        {
            ordered = false;
            return;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
        {
            ordered = false;
            return;
        }
        if (hipEventRecord(e, s) != hipSuccess) ordered = false;
        events.push_back(e);
        last = s;
    }
    ~ArenaBlock()
    {
        bool ok = ordered;
        if (ok)
            for (hipEvent_t e : events) ok = ok && hipStreamWaitEvent(last, e, 0) == hipSuccess;
        if (ok)
            (void)hipFreeAsync(base, last);
        else
        {
            (void)hipDeviceSynchronize(); // no work may still use the grids' arrays
            (void)hipFreeAsync(base, nullptr);
        }
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
};

struct sdfo_grid
{
    sdfo::GridSpec spec;
    int device = 0;
    bool sparse = false;
    bool hasHessian = true;
    int stride = 0; // contact-aware candidate stride

    // ---- dense: FP64 authoritative fields (SdfData) --------------------------
    double* phi = nullptr;     // N
    double* n3 = nullptr;      // 3N
    double* h6 = nullptr;      // 6N
    double* witness3 = nullptr; // 3N
    int32_t* tri = nullptr;    // N
    // ---- dense: FP32 query mirror ---------------------------------------------
    float4* pn = nullptr;      // (phi, nx, ny, nz) per voxel
    float* h6f = nullptr;      // 6 per voxel
    float4* cellRec = nullptr; // 8 corner records per cell (optional, see kCellRecordLimit)
    // ---- dense contact-aware candidates (FP64 authoritative + FP32 mirror) ------
    uint8_t* ccount = nullptr;  // N
    double* cphi = nullptr;     // N*S
    double* cn3 = nullptr;      // N*S*3
    int32_t* cbranch = nullptr; // N*S
    float4* cpn = nullptr;      // N*S (phi, nx, ny, nz)

    // ---- sparse bricks (SparseSdfData) -------------------------------------------
    int blockDim = 8, halo = 2;
    double band = 0.0, activation = 0.0;
    int nb = 0;
    int nbx = 0, nby = 0, nbz = 0;
    int32_t* table = nullptr;   // nbx*nby*nbz -> brick or -1
    uint8_t* activeMask = nullptr; // nbx*nby*nbz
    int32_t* bijk = nullptr;    // nb*3
    float* sphi = nullptr;      // nb*B^3
    float* sn3 = nullptr;       // nb*B^3*3
    float* sh6 = nullptr;       // nb*B^3*6
    float4* spn = nullptr;      // nb*B^3 query mirror
    uint8_t* sccount = nullptr; // nb*B^3
    float* scphi = nullptr;     // nb*B^3*S
    float* scn3 = nullptr;      // nb*B^3*S*3
    int32_t* scbranch = nullptr; // nb*B^3*S
    float4* scpn = nullptr;     // nb*B^3*S (phi, nx, ny, nz) query mirror
    // sparse cell records (optional, see kCellRecordLimit): the 8 corner
    // records of the cell at each brick-pool voxel, and a validity bit per cell
    float4* scellRec = nullptr;           // nb*B^3*8
    unsigned long long* scellMask = nullptr; // nb * ceil(B^3 / 64) words

    // ---- dense: FP64 cell records of the exact FP64-point queries, made on
    // the first FP64 Trilinear / Tricubic query of a grid that keeps FP32
    // cell records (cellRecords64 in query.hip), two 128-byte lines per
    // cell instead of about ten scattered lines from phi and n3
    struct Records64
    {
        std::mutex lock;
        double4* rec = nullptr;
        hipEvent_t ready = nullptr;    // the pack kernel's completion
        std::atomic<bool> done{false}; // ready has completed
        bool tried = false;
        int64_t bytes = 0;
    };
    mutable Records64 rec64;

    int64_t deviceBytes = 0;
    std::vector<void*> allocations;
    // Arrays come from the device's stream-ordered pool, ordered on this stream
    // (a builder sets its stream before allocating; loads use the null stream
    // with synchronous copies).
    hipStream_t allocStream = nullptr;
    // Optional slice of a shared device block: alloc() carves 256-byte aligned
    // arrays from it while it has room (the grids of one batched build share
    // one block, returned to the pool with the last of them).
    std::shared_ptr<ArenaBlock> arena;
    char* arenaNext = nullptr;
    size_t arenaLeft = 0;

    static size_t arenaBytes(size_t bytes) { return (bytes + 255) & ~size_t(255); }

    template <typename T>
    T* alloc(size_t count)
    {
        if (count == 0) return nullptr;
        void* p = nullptr;
        const size_t bytes = count * sizeof(T);
        if (arenaBytes(bytes) <= arenaLeft)
        {
            p = arenaNext;
            arenaNext += arenaBytes(bytes);
            arenaLeft -= arenaBytes(bytes);
            deviceBytes += int64_t(bytes);
            return static_cast<T*>(p);
        }
        hipError_t e = hipMallocAsync(&p, bytes, allocStream);
        if (e != hipSuccess)
            throw sdfo::Error(SDFO_ERR_OUT_OF_MEMORY,
                              "hipMallocAsync of " + std::to_string(bytes) + " bytes failed: " +
                                  hipGetErrorString(e));
        allocations.push_back(p);
        deviceBytes += int64_t(bytes);
        return static_cast<T*>(p);
    }
    // Release mode: synchronous (the device is synchronised first, so no work
    // on any stream can still read the arrays) or ordered on releaseStream.
    bool releaseOrdered = false;
    hipStream_t releaseStream = nullptr;
    ~sdfo_grid()
    {
        if (arena) arena->noteRelease(releaseOrdered, releaseStream);
        if (rec64.rec) allocations.push_back(rec64.rec);
        if (rec64.ready) (void)hipEventDestroy(rec64.ready);
        if (allocations.empty()) return;
        if (!releaseOrdered) (void)hipDeviceSynchronize();
        for (void* p : allocations) (void)hipFreeAsync(p, releaseOrdered ? releaseStream : nullptr);
    }
};

// Quadrature-point AABB tree of the contact broad phase (geometry/bvh.py).
struct PointBvhNode
{
    double lo[3], hi[3];
    int32_t left = -1, right = -1; // -1: leaf
    int64_t start = 0, count = 0;  // leaf: its points are order[start, start + count)
};

struct sdfo_point_bvh
{
    int device = -1; // set by the first call that culls with the tree
    int64_t n = 0;
    int leafSize = 16;
    std::vector<PointBvhNode> nodes; // nodes[0] is the root
    std::vector<int32_t> order;      // point indices, leaf by leaf
    int64_t nLeaves = 0;
    std::vector<double> hostLeafBox; // lo[3], hi[3] per leaf (leaves in node order)
    std::vector<int32_t> hostLeafOf; // leaf of each point
    double* leafBox = nullptr;       // device copies of the two tables
    int32_t* leafOf = nullptr;
    std::mutex uploadMutex;
    sdfo_point_bvh() = default;
    sdfo_point_bvh(const sdfo_point_bvh&) = delete;
    sdfo_point_bvh& operator=(const sdfo_point_bvh&) = delete;
    ~sdfo_point_bvh();
};

namespace sdfo {

void buildPointBvh(sdfo_point_bvh& b, const double* points, int64_t n, int leafSize);
void pointBvhToDevice(const sdfo_point_bvh& b, int device);
std::vector<int32_t> queryPointBvh(const sdfo_point_bvh& b, const double* qmin, const double* qmax,
                                   double margin);

// Per-pose broad-phase cull of the contact kernels: point i of pose b takes
// part when its leaf's box overlaps boxes[6b..6b+5] (qmin, qmax) grown by
// margin, or when no leaf does (contacts_sdf.py:270-271).
struct CullArgs
{
    const double* leafBox = nullptr;
    const int32_t* leafOf = nullptr;
    int64_t nLeaves = 0;
    const double* boxes = nullptr;
    double margin = 0.0;
};

// Phase timings of the calling thread's last dense build, in ms:
// grid set-up, host BVH, BVH upload + allocations (host clock), fill kernel
// (wave pass + replay pass), Hessian, FP32 mirrors + cell records (device
// events), total (host clock); then the number of replayed voxels.
constexpr int kBuildPhases = 8;
double* lastBuildPhases();

// Autotuner (tuner.cpp): a process-wide table of size-class -> variant
// decisions, filled by timing every variant on first use when enabled.
bool tunerEnabled();
bool tunerLookup(const std::string& key, int& value);
// Records timings (ms per variant) and keeps the fastest variant for key.
void tunerRecord(const std::string& key, const std::vector<std::pair<int, double>>& timings,
                 const char* const* names);
std::string tunerReport();
void tunerSetMode(int mode);
void tunerReset();
int log2Bucket(int64_t n);
bool streamCapturing(hipStream_t s);
// Runs f (which enqueues work on s) and returns its device time in ms.
double timeOnStream(hipStream_t s, const std::function<void()>& f);
// Times run(v) for every variant v < n in `rounds` interleaved rounds (v = 0,
// 1, ..., n-1, then again) and returns each variant's fastest time in ms, so
// that clock and cache drift during the timing favours no variant.
std::vector<std::pair<int, double>> timeVariants(hipStream_t s, int n,
                                                 const std::function<void(int)>& run,
                                                 int rounds = 3);
// The tuner's variant for key: the recorded one; else, when the tuner is on
// and s is not capturing, run(0) once (warm-up), time all n variants and keep
// the fastest (names label them in the report); else 0, the default.
int tunedVariant(const std::string& key, hipStream_t s, int n, const char* const* names,
                 const std::function<void(int)>& run);
int fillShapeFor(const GridSpec& g, int64_t ntris);

// Build kernels (build.hip): the calling thread's pinned counter that the last
// dense or sparse fill reports its replayed voxels to (valid after the stream
// is synchronised), and a test switch that sends every voxel to the replay pass.
int* replayCounterHost();
bool& forceReplay();
// Test overrides (calling thread): fill step budget and batched heavy-list
// capacity, 0 = the library's choice.
int& debugFillBudget();
int& debugMergeCap();  // register entries of the contact-aware first pass (0: 4)
int& debugSecondCap(); // entries per voxel of its second pass's first launch (0: 64)
int& debugHeavyCap();

// Keeps the device's stream-ordered pool memory across synchronisations.
void keepStreamPoolMemory(int device);

// Grid completion helpers (capi.cpp): FP32 mirrors and the brick table.
// FP32 mirrors and cell records of a dense grid (mirrored: pn and h6f were
// already written by the Hessian pass).
void finishDense(sdfo_grid& g, hipStream_t s, bool mirrored = false);
void finishSparse(sdfo_grid& g, const std::vector<int32_t>& bijk, hipStream_t s);
std::vector<int32_t> brickTable(const std::vector<int32_t>& bijk, int nbx, int nby, int nbz);

// Grids whose packed cell records fit in cellRecordLimit() bytes get them
// (8 x float4 per cell: a random interpolation reads one 128-byte line):
// an eighth of the current device's memory (36 GB on a 288 GB MI355X, so a
// dense 512^3 grid's 17.2 GB of records fit), and never less than 4 GiB.
constexpr int64_t kCellRecordLimit = int64_t(4) << 30;
int64_t cellRecordLimit();

// Contact-aware build (contact_aware.hip); corner normals: 9 doubles per triangle.
sdfo_grid* buildContactAware(const double* verts, int64_t nv, const int32_t* tris, int64_t nt,
                             const double* cornerN, const sdfo_build_params& p, bool sparse,
                             hipStream_t s);

// One mesh of a batched dense build: its grid, its BVH root inside the
// concatenated node array, and its output fields.
struct BatchMesh
{
    GridSpec g;
    int root;
    int64_t ntris;
    double *phi, *n3, *witness3;
    int32_t* tri;
};

// Kernel launchers (build.hip).
void launchDenseFillBatched(const std::vector<BatchMesh>& meshes, const DeviceBvh& bvh,
                            hipStream_t s);
void launchCellPack(const GridSpec& g, const float4* pn, float4* cellRec, hipStream_t s);
// FP64 cell records (8 double4 corner records per cell) from the FP64 fields.
void launchCellPack64(const GridSpec& g, const double* phi, const double* n3, double4* rec,
                      hipStream_t s);
void launchSparseCellPack(const GridSpec& g, int bd, int nbx, int nby, int nbz,
                          const int32_t* bijk, const int32_t* table, const float4* spn,
                          int64_t count, int words, float4* rec, unsigned long long* mask,
                          hipStream_t s);
// `scratch`, when given, holds denseFillScratchBytes(g) bytes (256-byte
// aligned) that no other work on s uses meanwhile; otherwise the lists of the
// fill passes are allocated on s.
// scratch: denseFillScratchBytes(g) of device memory for the fill's lists
// (else allocated per call); scratchZeroed: its first
// denseFillScratchHeaderBytes() were cleared on s beforehand, so the fill
// enqueues no clear of its own (a build clears them while its host BVH runs).
void launchDenseFill(const GridSpec& g, const DeviceBvh& bvh, double* phi, double* n3,
                     double* witness3, int32_t* tri, hipStream_t s, void* scratch = nullptr,
                     bool scratchZeroed = false);
size_t denseFillScratchBytes(const GridSpec& g);
size_t denseFillScratchHeaderBytes();
// With pn (and h6f) set, also writes the FP32 query mirror of every voxel.
void launchDenseHessian(const GridSpec& g, const double* n3, double* h6, hipStream_t s,
                        const double* phi = nullptr, float4* pn = nullptr, float* h6f = nullptr);
void launchDenseMirror(int64_t N, const double* phi, const double* n3, const double* h6,
                       float4* pn, float* h6f, hipStream_t s);
void launchCandidateMirror(int64_t count, const double* cphi, const double* cn3, float4* cpn,
                           hipStream_t s);
void launchSparseActivate(const GridSpec& g, const DeviceBvh& bvh, int blockDim, int nbx,
                          int nby, int nbz, double activation, uint8_t* mask, hipStream_t s);
void launchSparseFill(const GridSpec& g, const DeviceBvh& bvh, int blockDim, int nb,
                      const int32_t* bijk, float* sphi, float* sn3, hipStream_t s);
void launchSparseHessian(const GridSpec& g, int blockDim, int nb, int nbx, int nby, int nbz,
                         const int32_t* bijk, const int32_t* table, const float* sn3, float* sh6,
                         hipStream_t s);
void launchSparseMirror(int64_t count, const float* sphi, const float* sn3, float4* spn,
                        hipStream_t s);

// Query launchers (query.hip).
void launchQueryF32(const sdfo_grid& g, int mode, const float* pts, int64_t n, float* gap,
                    float* grad3, uint8_t* valid, float* normal3, float* witness3,
                    int32_t* branch, hipStream_t s);
// FP64 points on a dense grid's FP64 fields, without FMA contraction: the
// CPU path's results bit for bit (query_f64.hip).
void launchQueryExactF64(const sdfo_grid& g, int mode, const double* pts, int64_t n, double* gap,
                         double* grad3, uint8_t* valid, double* normal3, double* witness3,
                         int32_t* branch, hipStream_t s);
// The FP64 cell records of a dense grid for a query on stream s, made on
// first use (a grid with FP32 cell records, when twice their bytes fit the
// record limit; never inside a stream capture): the pack kernel runs on s
// and a query on another stream waits for it.  nullptr: the query gathers
// its corners from the FP64 fields.
const double4* cellRecords64(const sdfo_grid& g, hipStream_t s);
void launchQueryF64(const sdfo_grid& g, int mode, const double* pts, int64_t n, double* gap,
                    double* grad3, uint8_t* valid, double* normal3, double* witness3,
                    int32_t* branch, hipStream_t s);

// A batched query plan (query.hip): the device table of (grid, point range)
// per stream, uploaded once; running it is one kernel launch with no
// allocation, copy or synchronisation, so it can be captured in a HIP graph.
struct QueryPlan
{
    int streams = 0;
    int64_t points = 0; // total points over all streams
    int64_t units = 0;  // 64-point work units
    void* table = nullptr;
    ~QueryPlan();
};

std::unique_ptr<QueryPlan> makeQueryPlan(int ns, const sdfo_grid* const* grids,
                                         const int64_t* offsets);
// block: threads per block of the batched kernel (64, 128, 256 or 512).
void runQueryPlan(const QueryPlan& plan, int mode, int block, const float* pts, float* gap,
                  float* grad3, uint8_t* valid, hipStream_t s);

// Every ContactAware candidate of FP64 points (query_exact.hip); arrays of
// 8 candidates per point; points without candidates get count 0 and keep
// their gap / normal / witness / valid.
bool gridHasCandidates(const sdfo_grid& g);
void launchCandidatesF64(const sdfo_grid& g, const double* pts, int64_t n, uint8_t* count,
                         double* cgap, double* cn3, double* cw3, int32_t* cb, double* gap,
                         double* normal3, double* witness3, uint8_t* valid, hipStream_t s);

// Best time (us) over reps launches of nReads uniformly random reads of
// recordBytes-byte records (64 or 128) from an arrayBytes array, two per
// lane in 128-lane blocks (query.hip): the memory system's floor for the
// record-path query.  Synchronous, on the null stream.
double randomRecordReadUs(int recordBytes, int64_t nReads, int64_t arrayBytes, int reps);

// Contact launchers (contact.hip).
// With saved (device, contactSavedBytes(n, nb) bytes; plain forward only)
// the forward also records which (pose, point) pairs are in contact, for
// launchContactBackward to read instead of screening every pair again.
void launchContactForward(const sdfo_grid& g, const double* X, const double* w, int64_t n,
                          const double* poses, int64_t nb, const sdfo_contact_params& p,
                          double* out, int32_t* active, hipStream_t s,
                          const CullArgs* cull = nullptr, void* saved = nullptr);
int64_t contactSavedBytes(int64_t n, int64_t nb);
void launchContactEvaluate(const sdfo_grid& g, const double* X, const double* w, int64_t n,
                           const double* states, const double* frames, int64_t nb,
                           const sdfo_contact_params& p, const CullArgs* cull,
                           const double* bodyBoxes9, double* out, int32_t* active, hipStream_t s);
// The same with the device scratch in a static array (no allocation);
// at most 4096 pairs per call.
void launchContactEvaluateNoAlloc(const sdfo_grid& g, const double* X, const double* w, int64_t n,
                                  const double* states, const double* frames, int64_t nb,
                                  const sdfo_contact_params& p, const CullArgs* cull,
                                  const double* bodyBoxes9, double* out, int32_t* active,
                                  hipStream_t s);
void launchCullMask(const CullArgs& cu, int64_t n, int64_t nBoxes, uint8_t* mask, hipStream_t s);
void launchContactLinearize(const sdfo_grid& g, const double* X, int64_t n, const double* pose,
                            const sdfo_contact_params& p, double* g0, double* grad0,
                            uint8_t* valid0, hipStream_t s);
void launchContactForwardLinearized(const double* g0, const double* grad0, const uint8_t* valid0,
                                    const double* tref, const double* X, const double* w, int64_t n,
                                    const double* poses, int64_t nb, const sdfo_contact_params& p,
                                    double* out, int32_t* active, hipStream_t s,
                                    const CullArgs* cull = nullptr);
void launchContactBackward(const sdfo_grid& g, const double* X, const double* w, int64_t n,
                           const double* poses, int64_t nb, const sdfo_contact_params& p,
                           const double* gOut, double* gTwist, double* gParams, double* gX,
                           double* gW, hipStream_t s, const void* saved = nullptr);

} // namespace sdfo
