// Host-side geometry set-up: grid derivation and the BVH that the build
// kernels traverse.  Both run once per build and are O(N_t log N_t); the
// per-voxel work is on the GPU (build.hip).
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>

#include <sched.h>
#include <cstdlib>

#include "sdfo_internal.hpp"

namespace sdfo {

namespace {

// GridSpec.h:65-74: smallest ladder value >= n, capped at maxRes.
int snapResolution(int n, int maxRes)
{
    static const int ladder[] = {8, 16, 32, 64, 128, 256, 512, 1024, 2048};
    for (int level : ladder)
        if (level >= n || level >= maxRes) return level < maxRes ? level : maxRes;
    return maxRes;
}

inline double minOf(double a, double b) { return (b < a) ? b : a; } // std::min
inline double maxOf(double a, double b) { return (a < b) ? b : a; } // std::max

} // namespace

// SdfBuilder::setupGrid (SdfBuilder.cpp:994-1035).
GridSpec setupGrid(const double* verts, int64_t nv, const sdfo_build_params& p)
{
    double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
    for (int64_t v = 0; v < nv; ++v)
        for (int a = 0; a < 3; ++a)
        {
            lo[a] = minOf(lo[a], verts[3 * v + a]);
            hi[a] = maxOf(hi[a], verts[3 * v + a]);
        }
    const double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    const double diag = std::sqrt(ex * ex + ey * ey + ez * ez);
    const double pad = p.padding * diag;
    GridSpec g;
    double size[3];
    for (int a = 0; a < 3; ++a)
    {
        g.bmin[a] = lo[a] - pad;
        size[a] = (hi[a] + pad) - g.bmin[a];
    }
    double vs = p.voxel_size;
    if (vs <= 0.0) vs = maxOf(maxOf(size[0], size[1]), size[2]) / double(p.target_resolution);
    for (int a = 0; a < 3; ++a)
    {
        const int raw = std::max(1, int(std::ceil(size[a] / vs)));
        g.n[a] = snapResolution(raw, p.max_resolution);
        g.voxel[a] = vs;
        g.bmax[a] = g.bmin[a] + double(g.n[a]) * vs;
    }
    return g;
}

namespace {

// Node count of the subtree built over n triangles (the shape depends on n only).
int64_t subtreeNodes(int64_t n) { return medianTreeNodes(n); }

// Persistent worker threads (created on first use, kept for the process):
// tasks wait in one queue; a thread waiting for a task runs queued ones
// meanwhile, so nested submissions never deadlock.
class WorkPool
{
public:
    static WorkPool& get()
    {
        // never destroyed: the workers sleep on the queue until the process ends
        static WorkPool* p = new WorkPool(std::max(0, std::min(16, hostThreads()) - 1));
        return *p;
    }
    void submit(PoolTask* t)
    {
        {
            std::lock_guard<std::mutex> l(m_);
            q_.push_back(t);
        }
        cv_.notify_one();
    }
    void wait(PoolTask* t)
    {
        while (!t->done.load(std::memory_order_acquire))
        {
            PoolTask* o = tryPop();
            if (o)
                run(o);
            else
                std::this_thread::yield();
        }
        if (t->error) std::rethrow_exception(t->error);
    }

private:
    explicit WorkPool(int n)
    {
        for (int i = 0; i < n; ++i) std::thread([this] { loop(); }).detach();
    }
    PoolTask* tryPop()
    {
        std::lock_guard<std::mutex> l(m_);
        if (q_.empty()) return nullptr;
        PoolTask* t = q_.front();
        q_.pop_front();
        return t;
    }
    static void run(PoolTask* t)
    {
        try
        {
            t->fn();
        }
        catch (...)
        {
            t->error = std::current_exception();
        }
        t->done.store(true, std::memory_order_release);
    }
    void loop()
    {
        for (;;)
        {
            PoolTask* t;
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return !q_.empty(); });
                t = q_.front();
                q_.pop_front();
            }
            run(t);
        }
    }
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<PoolTask*> q_;
};

// Runs f(begin, end) over [0, n) in chunks on up to 16 threads (small
// ranges, and sequential builds, on the calling thread).
template <typename F>
void parallelFor(int n, F f, bool parallel = true)
{
    const int threads = std::min(16, hostThreads());
    const int chunks = (n < 16384 || !parallel) ? 1 : threads;
    if (chunks <= 1)
    {
        f(0, n);
        return;
    }
    const int step = (n + chunks - 1) / chunks;
    std::vector<PoolTask> tasks(size_t(chunks - 1));
    for (int c = 1; c < chunks; ++c)
    {
        const int a = c * step, b = std::min(n, a + step);
        tasks[size_t(c - 1)].fn = [=, &f] {
            if (a < b) f(a, b);
        };
        poolSubmit(&tasks[size_t(c - 1)]);
    }
    f(0, std::min(n, step));
    for (PoolTask& t : tasks) poolWait(&t);
}

// A triangle in BVH slot order with its centroid, so that every pass over a
// node's range reads memory sequentially.
struct Item
{
    double c[3];
    int idx;
    int pad;
};

// A growable buffer of trivially constructible elements that keeps its
// memory between uses and leaves new elements uninitialised: per-build host
// arrays then cost neither a fresh page fault per page nor a zero fill.
template <typename T>
struct GrowBuffer
{
    std::unique_ptr<T[]> p;
    size_t cap = 0;
    T* get(size_t n)
    {
        if (n > cap)
        {
            p.reset(new T[n]);
            cap = n;
        }
        return p.get();
    }
    // drops buffers far larger than typical meshes once they are not needed
    void trim(size_t keep)
    {
        if (cap > keep)
        {
            p.reset();
            cap = 0;
        }
    }
};

// Per-triangle boxes and centroid items of the calling thread's builds.  A
// thread waiting for the worker pool runs queued tasks meanwhile, which may
// be another mesh's build: such a nested build finds the thread's scratch in
// use and takes buffers of its own.
struct BvhScratch
{
    GrowBuffer<double> box;
    GrowBuffer<Item> items;
    bool busy = false;
};
thread_local BvhScratch t_bvhScratch;

struct ScratchLease
{
    BvhScratch* s;
    std::unique_ptr<BvhScratch> own;
    ScratchLease()
    {
        if (t_bvhScratch.busy)
        {
            own = std::make_unique<BvhScratch>();
            s = own.get();
        }
        else
            s = &t_bvhScratch;
        s->busy = true;
    }
    ~ScratchLease()
    {
        s->busy = false;
        // keep the scratch of meshes up to 2^19 triangles (46 MB) per thread
        s->box.trim(size_t(6) << 19);
        s->items.trim(size_t(1) << 19);
    }
};

// Nodes of at least this many triangles scan their centroids in parallel.
constexpr int64_t kParallelExtent = 65536;

struct BvhBuilder
{
    BvhNode* nodes;    // this mesh's nodes (indices local to the mesh); null: order only
    Item* items;
    const double* box; // bmin3, bmax3 per triangle (unused when nodes is null)

    // Builds the subtree over items[lo, hi) whose root has pre-order index me;
    // returns the subtree depth.  Subtrees above `par` triangles are built
    // concurrently (their node ranges are disjoint and their index ranges too).
    // A node's box is the union of its children's (min / max are exact, so it
    // equals the box over the node's triangles).
    int build(int lo, int hi, int64_t me, int level, int64_t par)
    {
        const double big = std::numeric_limits<double>::max();
        BvhNode nd{};
        nd.left = nd.right = -1;
        nd.triStart = -1;
        nd.triCount = 0;
        const int count = hi - lo;
        if (count <= 8 && !nodes) return level;
        if (count <= 8)
        {
            double bmn[3] = {big, big, big}, bmx[3] = {-big, -big, -big};
            for (int i = lo; i < hi; ++i)
            {
                const double* r = box + 6 * size_t(items[size_t(i)].idx);
                for (int k = 0; k < 3; ++k)
                {
                    bmn[k] = minOf(bmn[k], r[k]);
                    bmx[k] = maxOf(bmx[k], r[3 + k]);
                }
            }
            for (int k = 0; k < 3; ++k)
            {
                nd.bmin[k] = bmn[k];
                nd.bmax[k] = bmx[k];
            }
            nd.triStart = lo;
            nd.triCount = count;
            nodes[me] = nd;
            return level;
        }
        double cmn[3] = {big, big, big}, cmx[3] = {-big, -big, -big};
        auto extents = [&](int64_t a, int64_t b, double* mn, double* mx) {
            for (int64_t i = a; i < b; ++i)
            {
                const double* c = items[size_t(i)].c;
                for (int k = 0; k < 3; ++k)
                {
                    mn[k] = minOf(mn[k], c[k]);
                    mx[k] = maxOf(mx[k], c[k]);
                }
            }
        };
        if (count >= kParallelExtent && par != INT64_MAX)
        {
            // the top nodes' scans are on every build's critical path: chunks
            // on the workers, merged by min / max (exact, so in any order)
            std::mutex m;
            parallelRange(count, kParallelExtent / 4, [&](int64_t a, int64_t b) {
                double mn[3] = {big, big, big}, mx[3] = {-big, -big, -big};
                extents(lo + a, lo + b, mn, mx);
                std::lock_guard<std::mutex> lk(m);
                for (int k = 0; k < 3; ++k)
                {
                    cmn[k] = minOf(cmn[k], mn[k]);
                    cmx[k] = maxOf(cmx[k], mx[k]);
                }
            });
        }
        else
            extents(lo, hi, cmn, cmx);
        const double ext[3] = {cmx[0] - cmn[0], cmx[1] - cmn[1], cmx[2] - cmn[2]};
        int axis = 0;
        if (ext[1] > ext[0]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        const int mid = lo + count / 2;
        // the same comparisons, hence the same permutation, as nth_element over
        // triangle indices compared by centroid (KdTree.h:226-229)
        std::nth_element(items + lo, items + mid, items + hi,
                         [axis](const Item& x, const Item& y) { return x.c[axis] < y.c[axis]; });
        const int64_t left = me + 1;
        const int64_t right = me + 1 + subtreeNodes(mid - lo);
        int dl, dr;
        if (count > par)
        {
            PoolTask t;
            t.fn = [&] { dl = build(lo, mid, left, level + 1, par); };
            poolSubmit(&t);
            dr = build(mid, hi, right, level + 1, par);
            poolWait(&t);
        }
        else
        {
            dl = build(lo, mid, left, level + 1, par);
            dr = build(mid, hi, right, level + 1, par);
        }
        if (!nodes) return dl > dr ? dl : dr;
        const BvhNode& L = nodes[left];
        const BvhNode& R = nodes[right];
        for (int k = 0; k < 3; ++k)
        {
            nd.bmin[k] = minOf(L.bmin[k], R.bmin[k]);
            nd.bmax[k] = maxOf(L.bmax[k], R.bmax[k]);
        }
        nd.left = int(left);
        nd.right = int(right);
        nodes[me] = nd;
        return dl > dr ? dl : dr;
    }
};

} // namespace

int hostThreads()
{
    static const int n = [] {
        int c = int(std::thread::hardware_concurrency());
        cpu_set_t set;
        CPU_ZERO(&set);
        if (sched_getaffinity(0, sizeof(set), &set) == 0) c = std::min(c, int(CPU_COUNT(&set)));
        if (const char* e = std::getenv("OMP_NUM_THREADS"))
        {
            const int o = std::atoi(e);
            if (o > 0) c = std::min(c, o);
        }
        return std::max(1, std::min(c, 64));
    }();
    return n;
}

int64_t bvhNodeCount(int64_t nt) { return nt > 0 ? subtreeNodes(nt) : 0; }

void poolSubmit(PoolTask* t) { WorkPool::get().submit(t); }
void poolWait(PoolTask* t) { WorkPool::get().wait(t); }

void parallelRange(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& f)
{
    if (n <= 0) return;
    const int64_t chunks = std::max<int64_t>(
        1, std::min<int64_t>(std::min(16, hostThreads()), n / std::max<int64_t>(1, grain)));
    const int64_t step = (n + chunks - 1) / chunks;
    std::vector<PoolTask> tasks(size_t(chunks - 1));
    for (int64_t c = 1; c < chunks; ++c)
    {
        const int64_t a = c * step, b = std::min(n, a + step);
        tasks[size_t(c - 1)].fn = [=, &f] {
            if (a < b) f(a, b);
        };
        poolSubmit(&tasks[size_t(c - 1)]);
    }
    std::exception_ptr first;
    try
    {
        f(0, std::min(n, step));
    }
    catch (...)
    {
        first = std::current_exception();
    }
    for (PoolTask& t : tasks)
    {
        try
        {
            poolWait(&t);
        }
        catch (...)
        {
            if (!first) first = std::current_exception();
        }
    }
    if (first) std::rethrow_exception(first);
}

// KdTree::build / buildRecursive (KdTree.h:91-121, 188-242): object-median
// split on the largest centroid extent via std::nth_element, leaves of <= 8.
// Node indices are the recursion's pre-order creation order, so the tree is
// node-for-node the reference tree; the traversal order (near child first)
// is reproduced by the kernels.  Writes bvhNodeCount(nt) nodes and nt slots
// into the given arrays with node links and slot ranges offset by nodeBase
// and slotBase (one mesh of a concatenated BVH); returns the depth.
int buildHostBvhInto(const double* verts, int64_t nv, const int32_t* tris, int64_t nt,
                     BvhNode* nodes, double* slotTri9, int32_t* slotTri, int nodeBase, int slotBase,
                     bool parallel)
{
    const int n = int(nt);
    // per-triangle boxes (not when only the order is asked for) and
    // centroids (in slot-order items), in parallel chunks that check the
    // vertex indices first
    ScratchLease lease;
    double* box = nodes ? lease.s->box.get(size_t(n) * 6) : nullptr;
    Item* items = lease.s->items.get(size_t(n));
    std::atomic<bool> badIndex{false};
    parallelFor(
        n,
        [&](int t0, int t1) {
            for (int t = t0; t < t1; ++t)
            {
                const int32_t* ix = tris + 3 * size_t(t);
                if (ix[0] < 0 || ix[0] >= nv || ix[1] < 0 || ix[1] >= nv || ix[2] < 0 || ix[2] >= nv)
                {
                    badIndex.store(true, std::memory_order_relaxed);
                    return;
                }
                const double* a = verts + 3 * size_t(ix[0]);
                const double* b = verts + 3 * size_t(ix[1]);
                const double* c = verts + 3 * size_t(ix[2]);
                Item& it = items[size_t(t)];
                for (int k = 0; k < 3; ++k)
                {
                    if (box)
                    {
                        box[6 * size_t(t) + k] = minOf(minOf(a[k], b[k]), c[k]);
                        box[6 * size_t(t) + 3 + k] = maxOf(maxOf(a[k], b[k]), c[k]);
                    }
                    it.c[k] = ((a[k] + b[k]) + c[k]) / 3.0;
                }
                it.idx = t;
                it.pad = 0;
            }
        },
        parallel);
    if (badIndex.load()) throw Error(SDFO_ERR_INVALID_ARGUMENT, "triangle references invalid vertex");
    int depth = 0;
    if (n > 0)
    {
        // Subtrees of more than n/64 triangles (and at least 1024) are handed
        // to the worker threads.
        const int64_t par = parallel ? std::max<int64_t>(1024, n / 64) : INT64_MAX;
        BvhBuilder bb{nodes, items, box};
        depth = bb.build(0, n, 0, 1, par);
        const int64_t nn = bvhNodeCount(n);
        if (nodes && (nodeBase != 0 || slotBase != 0))
            for (int64_t q = 0; q < nn; ++q)
            {
                BvhNode& nd = nodes[q];
                if (nd.triCount > 0)
                    nd.triStart += slotBase;
                else
                {
                    nd.left += nodeBase;
                    nd.right += nodeBase;
                }
            }
    }
    parallelFor(
        n,
        [&](int s0, int s1) {
            for (int s = s0; s < s1; ++s)
            {
                const int t = items[size_t(s)].idx;
                slotTri[s] = t;
                if (!slotTri9) continue;
                const int32_t* ix = tris + 3 * size_t(t);
                for (int c = 0; c < 3; ++c)
                    for (int k = 0; k < 3; ++k)
                        slotTri9[9 * size_t(s) + 3 * c + k] = verts[3 * size_t(ix[c]) + k];
            }
        },
        parallel);
    return depth;
}

int buildHostBvhOrder(const double* verts, int64_t nv, const int32_t* tris, int64_t nt,
                      int32_t* slotTri, bool parallel)
{
    return buildHostBvhInto(verts, nv, tris, nt, nullptr, nullptr, slotTri, 0, 0, parallel);
}

void setRayDir(HostBvh& h)
{
    // KdTree.h:330-333: fixed irrational ray direction, normalised once on the host.
    const double d0[3] = {1.0, 0.3713906763541037, 0.52999894000318};
    const double len = std::sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
    for (int k = 0; k < 3; ++k) h.rayDir[k] = d0[k] / len;
}

namespace {
// The output arrays of the calling thread's scratch builds.
// Only the top-level dense and sparse builds of a user thread use it, and
// those never run as pool tasks, so it is never in use twice on one thread.
struct BvhOut
{
    GrowBuffer<BvhNode> nodes;
    GrowBuffer<double> slotTri9;
    GrowBuffer<int32_t> slotTri;
};
thread_local BvhOut t_bvhOut;
} // namespace

HostBvhView buildHostBvhScratch(const double* verts, int64_t nv, const int32_t* tris, int64_t nt)
{
    HostBvhView v{};
    const size_t nn = size_t(bvhNodeCount(nt));
    BvhNode* nodes = t_bvhOut.nodes.get(std::max<size_t>(1, nn));
    double* tri9 = t_bvhOut.slotTri9.get(std::max<size_t>(1, size_t(nt) * 9));
    int32_t* slot = t_bvhOut.slotTri.get(std::max<size_t>(1, size_t(nt)));
    v.depth = buildHostBvhInto(verts, nv, tris, nt, nodes, tri9, slot, 0, 0, true);
    v.nodes = nodes;
    v.nNodes = int64_t(nn);
    v.slotTri9 = tri9;
    v.slotTri = slot;
    v.nSlots = nt;
    HostBvh dir;
    setRayDir(dir);
    for (int k = 0; k < 3; ++k) v.rayDir[k] = dir.rayDir[k];
    return v;
}

HostBvhOrderView buildHostBvhOrderScratch(const double* verts, int64_t nv, const int32_t* tris,
                                          int64_t nt)
{
    HostBvhOrderView v{};
    int32_t* slot = t_bvhOut.slotTri.get(std::max<size_t>(1, size_t(nt)));
    v.depth = buildHostBvhOrder(verts, nv, tris, nt, slot, true);
    v.slotTri = slot;
    v.nSlots = nt;
    HostBvh dir;
    setRayDir(dir);
    for (int k = 0; k < 3; ++k) v.rayDir[k] = dir.rayDir[k];
    return v;
}

void releaseHostBvhScratch(int64_t keepTris)
{
    t_bvhOut.nodes.trim(size_t(bvhNodeCount(keepTris)));
    t_bvhOut.slotTri9.trim(size_t(keepTris) * 9);
    t_bvhOut.slotTri.trim(size_t(keepTris));
}

HostBvh buildHostBvh(const double* verts, int64_t nv, const int32_t* tris, int64_t nt)
{
    HostBvh h;
    h.nodes.resize(size_t(bvhNodeCount(nt)));
    h.slotTri.resize(size_t(nt));
    h.slotTri9.resize(size_t(nt) * 9);
    h.depth = buildHostBvhInto(verts, nv, tris, nt, h.nodes.data(), h.slotTri9.data(),
                               h.slotTri.data(), 0, 0, true);
    setRayDir(h);
    return h;
}

void DeviceBvh::upload(const HostBvh& h, hipStream_t s)
{
    upload(h.nodes.data(), int64_t(h.nodes.size()), h.slotTri9.data(), h.slotTri.data(),
           int64_t(h.slotTri.size()), h.depth, h.rayDir, s);
}

namespace {
size_t alignUp256(size_t b) { return (b + 255) & ~size_t(255); }
size_t bvhBlockBytes(int64_t nNodes, int64_t nSlots)
{
    const size_t n = size_t(std::max<int64_t>(1, nNodes)), t = size_t(std::max<int64_t>(1, nSlots));
    return alignUp256(sizeof(BvhNode) * n) + alignUp256(sizeof(double) * 9 * t) +
           alignUp256(sizeof(double) * 6 * t) + alignUp256(sizeof(int32_t) * t);
}
} // namespace

void DeviceBvh::reserve(int64_t nNodes, int64_t nSlots, hipStream_t s)
{
    release();
    stream = s;
    SDFO_HIP(hipMallocAsync(&block, bvhBlockBytes(nNodes, nSlots), s));
    capNodes = std::max<int64_t>(1, nNodes);
    capSlots = std::max<int64_t>(1, nSlots);
}

void DeviceBvh::upload(const BvhNode* hostNodes, int64_t nNodes, const double* slotTri9,
                       const int32_t* slotTri, int64_t nSlots, int treeDepth, const double* dir,
                       hipStream_t s)
{
    if (!block || stream != s || nNodes > capNodes || nSlots > capSlots) reserve(nNodes, nSlots, s);
    depth = treeDepth;
    for (int k = 0; k < 3; ++k) rayDir[k] = dir[k];
    if (depth + 2 > kStackDepth)
        throw Error(SDFO_ERR_UNSUPPORTED, "BVH depth " + std::to_string(depth) +
                                              " exceeds the traversal stack");
    if (nNodes >= (1 << 26))
        throw Error(SDFO_ERR_UNSUPPORTED, "BVH with " + std::to_string(nNodes) +
                                              " nodes exceeds the traversal stack encoding");
    nnodes = int(nNodes);
    ntris = nSlots;
    char* b = static_cast<char*>(block);
    nodes = reinterpret_cast<BvhNode*>(b);
    b += alignUp256(sizeof(BvhNode) * size_t(capNodes));
    tri9 = reinterpret_cast<double*>(b);
    b += alignUp256(sizeof(double) * 9 * size_t(capSlots));
    box6 = reinterpret_cast<double*>(b);
    b += alignUp256(sizeof(double) * 6 * size_t(capSlots));
    triIdx = reinterpret_cast<int32_t*>(b);
    SDFO_HIP(hipMemcpyAsync(nodes, hostNodes, sizeof(BvhNode) * size_t(nNodes),
                            hipMemcpyHostToDevice, s));
    SDFO_HIP(hipMemcpyAsync(tri9, slotTri9, sizeof(double) * 9 * size_t(nSlots),
                            hipMemcpyHostToDevice, s));
    SDFO_HIP(hipMemcpyAsync(triIdx, slotTri, sizeof(int32_t) * size_t(nSlots),
                            hipMemcpyHostToDevice, s));
    launchSlotBoxes(tri9, nSlots, box6, s);
}

namespace {
size_t assemblyExtraBytes(int64_t nNodes, int64_t nSlots, int64_t nVerts, int nParts)
{
    return alignUp256(sizeof(double) * 3 * size_t(std::max<int64_t>(1, nVerts))) +
           alignUp256(sizeof(int32_t) * 3 * size_t(std::max<int64_t>(1, nSlots))) +
           alignUp256(sizeof(int32_t) * size_t(std::max<int64_t>(1, nNodes))) +
           alignUp256(sizeof(BvhMeshPart) * size_t(std::max(1, nParts)));
}
} // namespace

void DeviceBvh::reserveAssembly(int64_t nNodes, int64_t nSlots, int64_t nVerts, int nPartsMax,
                                hipStream_t s)
{
    release();
    stream = s;
    SDFO_HIP(hipMallocAsync(&block,
                            bvhBlockBytes(nNodes, nSlots) +
                                assemblyExtraBytes(nNodes, nSlots, nVerts, nPartsMax),
                            s));
    capNodes = std::max<int64_t>(1, nNodes);
    capSlots = std::max<int64_t>(1, nSlots);
    capVerts = std::max<int64_t>(1, nVerts);
    capParts = std::max(1, nPartsMax);
    char* b = static_cast<char*>(block);
    nodes = reinterpret_cast<BvhNode*>(b);
    b += alignUp256(sizeof(BvhNode) * size_t(capNodes));
    tri9 = reinterpret_cast<double*>(b);
    b += alignUp256(sizeof(double) * 9 * size_t(capSlots));
    box6 = reinterpret_cast<double*>(b);
    b += alignUp256(sizeof(double) * 6 * size_t(capSlots));
    triIdx = reinterpret_cast<int32_t*>(b);
    b += alignUp256(sizeof(int32_t) * size_t(capSlots));
    verts = reinterpret_cast<double*>(b);
    b += alignUp256(sizeof(double) * 3 * size_t(capVerts));
    tris = reinterpret_cast<int32_t*>(b);
    b += alignUp256(sizeof(int32_t) * 3 * size_t(capSlots));
    flags = reinterpret_cast<int32_t*>(b);
    b += alignUp256(sizeof(int32_t) * size_t(capNodes));
    parts = reinterpret_cast<BvhMeshPart*>(b);
}

void DeviceBvh::stageMeshes(const double* hostVerts, int64_t nVerts, const int32_t* hostTris,
                            int64_t nTris, const BvhMeshPart* hostParts, int np, hipStream_t s)
{
    if (!verts || stream != s || nVerts > capVerts || nTris > capSlots || np > capParts)
        throw Error(SDFO_ERR_INVALID_ARGUMENT, "device BVH: staging exceeds the reserved block");
    int64_t nodesTotal = 0;
    for (int q = 0; q < np; ++q) nodesTotal = std::max(nodesTotal, hostParts[q].nodeBase + bvhNodeCount(hostParts[q].nSlots));
    if (nodesTotal > capNodes)
        throw Error(SDFO_ERR_INVALID_ARGUMENT, "device BVH: staging exceeds the reserved block");
    nParts = np;
    nnodes = int(nodesTotal);
    ntris = nTris;
    // cut each mesh's tree where its subtrees hold at most kSubtreeSlots slots
    std::vector<BvhMeshPart> pp(hostParts, hostParts + np);
    nSubtrees = 0;
    topLevelsMax = 0;
    for (BvhMeshPart& q : pp)
    {
        if (q.nSlots > INT32_MAX / 2)
            throw Error(SDFO_ERR_UNSUPPORTED, "device BVH: mesh too large");
        for (int d = 0; d < 32; ++d)
        {
            q.f0[d] = int32_t(medianTreeNodes(q.nSlots >> d));
            q.f1[d] = int32_t(medianTreeNodes((q.nSlots >> d) + 1));
        }
        // every node above the cut must be inner: depth L - 1 holds more than 8
        int L = 0;
        while (L < kMaxTopLevels && (q.nSlots >> (L + 1)) >= kMinSubtreeSlots) ++L;
        q.topLevels = L;
        q.subBase = nSubtrees;
        nSubtrees += int64_t(1) << L;
        topLevelsMax = std::max(topLevelsMax, L);
    }
    if (nVerts > 0)
        SDFO_HIP(hipMemcpyAsync(verts, hostVerts, sizeof(double) * 3 * size_t(nVerts),
                                hipMemcpyHostToDevice, s));
    if (nTris > 0)
        SDFO_HIP(hipMemcpyAsync(tris, hostTris, sizeof(int32_t) * 3 * size_t(nTris),
                                hipMemcpyHostToDevice, s));
    SDFO_HIP(hipMemcpyAsync(parts, pp.data(), sizeof(BvhMeshPart) * size_t(np), hipMemcpyHostToDevice, s));
    SDFO_HIP(hipMemsetAsync(flags, 0, sizeof(int32_t) * size_t(std::max<int64_t>(1, nodesTotal)), s));
}

void DeviceBvh::assemble(const int32_t* slotTri, int treeDepth, const double* dir, hipStream_t s)
{
    depth = treeDepth;
    for (int k = 0; k < 3; ++k) rayDir[k] = dir[k];
    if (depth + 2 > kStackDepth)
        throw Error(SDFO_ERR_UNSUPPORTED, "BVH depth " + std::to_string(depth) +
                                              " exceeds the traversal stack");
    if (nnodes >= (1 << 26))
        throw Error(SDFO_ERR_UNSUPPORTED, "BVH with " + std::to_string(nnodes) +
                                              " nodes exceeds the traversal stack encoding");
    if (ntris > 0)
        SDFO_HIP(hipMemcpyAsync(triIdx, slotTri, sizeof(int32_t) * size_t(ntris),
                                hipMemcpyHostToDevice, s));
    launchBvhAssemble(*this, s);
}

void DeviceBvh::release()
{
    if (block) (void)hipFreeAsync(block, stream);
    block = nullptr;
    capNodes = capSlots = 0;
    capVerts = 0;
    capParts = 0;
    nParts = 0;
    verts = nullptr;
    tris = nullptr;
    flags = nullptr;
    parts = nullptr;
    nSubtrees = 0;
    topLevelsMax = 0;
    nodes = nullptr;
    tri9 = nullptr;
    box6 = nullptr;
    triIdx = nullptr;
}

} // namespace sdfo

#include "sdfo_debug.h"

extern "C" int sdfo_debug_bvh_export(const double* verts, int64_t n_verts, const int32_t* tris,
                                     int64_t n_tris, double* nodes6, int32_t* nodeI4,
                                     int32_t* perm, int32_t max_nodes, int32_t* depth)
{
    try
    {
        if (!verts || !tris || n_verts <= 0 || n_tris <= 0) return SDFO_ERR_EMPTY_MESH;
        const sdfo::HostBvh h = sdfo::buildHostBvh(verts, n_verts, tris, n_tris);
        const int nn = int(h.nodes.size());
        if (nn > max_nodes) return -nn;
        for (int i = 0; i < nn; ++i)
        {
            const sdfo::BvhNode& b = h.nodes[i];
            for (int k = 0; k < 3; ++k)
            {
                nodes6[6 * i + k] = b.bmin[k];
                nodes6[6 * i + 3 + k] = b.bmax[k];
            }
            nodeI4[4 * i + 0] = b.left;
            nodeI4[4 * i + 1] = b.right;
            nodeI4[4 * i + 2] = b.triStart;
            nodeI4[4 * i + 3] = b.triCount;
        }
        for (int64_t s = 0; s < n_tris; ++s) perm[s] = h.slotTri[size_t(s)];
        if (depth) *depth = h.depth;
        return nn;
    }
    catch (const sdfo::Error& e)
    {
        sdfo::setLastError(e.what());
        return e.code;
    }
    catch (...)
    {
        return SDFO_ERR_INVALID_ARGUMENT;
    }
}

extern "C" int sdfo_debug_device_bvh_export(int32_t n_meshes, const double* verts,
                                            const int64_t* vo, const int32_t* tris,
                                            const int64_t* to, double* nodes6, int32_t* nodeI4,
                                            int32_t* perm, double* tri9, double* box6,
                                            int32_t max_nodes, int32_t* depth)
{
    sdfo::DeviceBvh db;
    try
    {
        if (n_meshes <= 0 || !verts || !vo || !tris || !to) return SDFO_ERR_EMPTY_MESH;
        std::vector<sdfo::BvhMeshPart> parts(static_cast<size_t>(n_meshes));
        int64_t nodes = 0, slots = 0;
        for (int32_t m = 0; m < n_meshes; ++m)
        {
            const int64_t nt = to[m + 1] - to[m];
            if (nt <= 0 || vo[m + 1] <= vo[m]) return SDFO_ERR_EMPTY_MESH;
            parts[size_t(m)] = sdfo::BvhMeshPart{slots, nt, nodes, vo[m] - vo[0], to[m] - to[0]};
            nodes += sdfo::bvhNodeCount(nt);
            slots += nt;
        }
        if (nodes > max_nodes) return int(-nodes);
        std::vector<int32_t> slot(static_cast<size_t>(slots));
        int d = 0;
        for (int32_t m = 0; m < n_meshes; ++m)
        {
            const sdfo::BvhMeshPart& P = parts[size_t(m)];
            d = std::max(d, sdfo::buildHostBvhOrder(verts + 3 * vo[m], vo[m + 1] - vo[m],
                                                    tris + 3 * to[m], P.nSlots,
                                                    slot.data() + P.slotBase, true));
        }
        sdfo::HostBvh dir;
        sdfo::setRayDir(dir);
        db.reserveAssembly(nodes, slots, vo[n_meshes] - vo[0], n_meshes, nullptr);
        db.stageMeshes(verts + 3 * vo[0], vo[n_meshes] - vo[0], tris + 3 * to[0],
                       to[n_meshes] - to[0], parts.data(), n_meshes, nullptr);
        db.assemble(slot.data(), d, dir.rayDir, nullptr);
        std::vector<sdfo::BvhNode> nd(static_cast<size_t>(nodes));
        SDFO_HIP(hipMemcpy(nd.data(), db.nodes, sizeof(sdfo::BvhNode) * nd.size(), hipMemcpyDeviceToHost));
        SDFO_HIP(hipMemcpy(perm, db.triIdx, sizeof(int32_t) * size_t(slots), hipMemcpyDeviceToHost));
        SDFO_HIP(hipMemcpy(tri9, db.tri9, sizeof(double) * 9 * size_t(slots), hipMemcpyDeviceToHost));
        SDFO_HIP(hipMemcpy(box6, db.box6, sizeof(double) * 6 * size_t(slots), hipMemcpyDeviceToHost));
        SDFO_HIP(hipDeviceSynchronize());
        db.release();
        for (size_t i = 0; i < nd.size(); ++i)
        {
            for (int k = 0; k < 3; ++k)
            {
                nodes6[6 * i + k] = nd[i].bmin[k];
                nodes6[6 * i + 3 + k] = nd[i].bmax[k];
            }
            nodeI4[4 * i + 0] = nd[i].left;
            nodeI4[4 * i + 1] = nd[i].right;
            nodeI4[4 * i + 2] = nd[i].triStart;
            nodeI4[4 * i + 3] = nd[i].triCount;
        }
        if (depth) *depth = d;
        return int(nodes);
    }
    catch (const sdfo::Error& e)
    {
        (void)hipDeviceSynchronize();
        db.release();
        sdfo::setLastError(e.what());
        return e.code;
    }
    catch (...)
    {
        (void)hipDeviceSynchronize();
        db.release();
        return SDFO_ERR_INVALID_ARGUMENT;
    }
}
