// Contact broad phase (geometry/bvh.py:16-75): the AABB tree over a body's
// quadrature points that SDFContactEngine.evaluate culls with
// (contacts_sdf.py:245-272).  The tree is built on the host exactly as the
// reference splits it; the device keeps only what the per-pose cull needs,
// each point's leaf and each leaf's box: a leaf's box lies inside every
// ancestor's box, so the reference's pruned traversal reports exactly the
// leaves whose own box passes the overlap test.
#include <algorithm>
#include <mutex>

#include "sdfo_internal.hpp"

sdfo_point_bvh::~sdfo_point_bvh()
{
    if (leafBox) (void)hipFree(leafBox);
    if (leafOf) (void)hipFree(leafOf);
}

namespace sdfo {

namespace {

struct TreeBuilder
{
    const double* p;
    int leafSize;
    sdfo_point_bvh& b;

    // AABB_BVH._build (bvh.py:27-54); returns the node index.
    int32_t build(std::vector<int32_t>& idx)
    {
        const int32_t id = int32_t(b.nodes.size());
        b.nodes.emplace_back();
        PointBvhNode nd;
        for (int k = 0; k < 3; ++k)
        {
            nd.lo[k] = p[3 * int64_t(idx[0]) + k];
            nd.hi[k] = nd.lo[k];
        }
        for (int32_t i : idx)
            for (int k = 0; k < 3; ++k)
            {
                const double v = p[3 * int64_t(i) + k];
                nd.lo[k] = std::min(nd.lo[k], v); // numpy min/max, NaN aside
                nd.hi[k] = std::max(nd.hi[k], v);
            }
        const int64_t n = int64_t(idx.size());
        if (n <= leafSize)
        {
            nd.start = int64_t(b.order.size());
            nd.count = n;
            b.order.insert(b.order.end(), idx.begin(), idx.end());
            b.nodes[size_t(id)] = nd;
            return id;
        }
        // longest axis, first on ties (np.argmax), midpoint split
        const double ext[3] = {nd.hi[0] - nd.lo[0], nd.hi[1] - nd.lo[1], nd.hi[2] - nd.lo[2]};
        int axis = 0;
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        const double mid = (nd.lo[axis] + nd.hi[axis]) * 0.5;
        std::vector<int32_t> left, right;
        left.reserve(size_t(n));
        right.reserve(size_t(n));
        for (int32_t i : idx) (p[3 * int64_t(i) + axis] <= mid ? left : right).push_back(i);
        if (left.empty() || right.empty())
        {
            // median split on the sorted coordinate (bvh.py:44-49); stable
            // order among equal coordinates
            std::vector<int32_t> s(idx);
            std::stable_sort(s.begin(), s.end(), [&](int32_t a, int32_t c) {
                return p[3 * int64_t(a) + axis] < p[3 * int64_t(c) + axis];
            });
            const size_t half = size_t(n / 2);
            left.assign(s.begin(), s.begin() + half);
            right.assign(s.begin() + half, s.end());
        }
        std::vector<int32_t>().swap(idx);
        nd.left = build(left);
        nd.right = build(right);
        b.nodes[size_t(id)] = nd;
        return id;
    }
};

inline bool overlaps(const double* lo, const double* hi, const double* qmin, const double* qmax,
                     double margin)
{
    for (int k = 0; k < 3; ++k)
        if (hi[k] + margin < qmin[k] || lo[k] - margin > qmax[k]) return false;
    return true;
}

} // namespace

void buildPointBvh(sdfo_point_bvh& b, const double* points, int64_t n, int leafSize)
{
    b.n = n;
    b.leafSize = leafSize;
    b.nodes.clear();
    b.order.clear();
    if (n == 0) return;
    b.order.reserve(size_t(n));
    std::vector<int32_t> idx(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) idx[size_t(i)] = int32_t(i);
    TreeBuilder{points, leafSize, b}.build(idx);

    std::vector<double> box;
    std::vector<int32_t> leafOf(static_cast<size_t>(n), -1);
    int32_t leaf = 0;
    for (const PointBvhNode& nd : b.nodes)
    {
        if (nd.left >= 0) continue;
        box.insert(box.end(), nd.lo, nd.lo + 3);
        box.insert(box.end(), nd.hi, nd.hi + 3);
        for (int64_t j = nd.start; j < nd.start + nd.count; ++j) leafOf[size_t(b.order[size_t(j)])] = leaf;
        ++leaf;
    }
    b.nLeaves = leaf;
    b.hostLeafBox = std::move(box);
    b.hostLeafOf = std::move(leafOf);
}

// The leaf tables go to the device of the first call that culls with the
// tree; later calls must use that device.
void pointBvhToDevice(const sdfo_point_bvh& cb, int device)
{
    auto& b = const_cast<sdfo_point_bvh&>(cb);
    std::lock_guard<std::mutex> lock(b.uploadMutex);
    if (b.device < 0)
    {
        if (b.n > 0)
        {
            SDFO_HIP(hipMalloc(&b.leafBox, b.hostLeafBox.size() * sizeof(double)));
            SDFO_HIP(hipMalloc(&b.leafOf, b.hostLeafOf.size() * sizeof(int32_t)));
            SDFO_HIP(hipMemcpy(b.leafBox, b.hostLeafBox.data(), b.hostLeafBox.size() * sizeof(double),
                               hipMemcpyHostToDevice));
            SDFO_HIP(hipMemcpy(b.leafOf, b.hostLeafOf.data(), b.hostLeafOf.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice));
        }
        b.device = device;
    }
    else if (b.device != device)
        throw Error(SDFO_ERR_INVALID_ARGUMENT, "the point BVH lives on device " +
                                                   std::to_string(b.device) + ", not " +
                                                   std::to_string(device));
}

// AABB_BVH.query (bvh.py:56-75): the root short cut returns every index in
// order; otherwise leaves in the order of a stack that pushes left, then right.
std::vector<int32_t> queryPointBvh(const sdfo_point_bvh& b, const double* qmin, const double* qmax,
                                   double margin)
{
    std::vector<int32_t> out;
    if (b.nodes.empty()) return out;
    const PointBvhNode& r = b.nodes[0];
    bool all = true;
    for (int k = 0; k < 3; ++k)
        all = all && qmin[k] <= r.lo[k] - margin && qmax[k] >= r.hi[k] + margin;
    if (all)
    {
        out.resize(size_t(b.n));
        for (int64_t i = 0; i < b.n; ++i) out[size_t(i)] = int32_t(i);
        return out;
    }
    std::vector<int32_t> stack{0};
    while (!stack.empty())
    {
        const PointBvhNode& nd = b.nodes[size_t(stack.back())];
        stack.pop_back();
        if (!overlaps(nd.lo, nd.hi, qmin, qmax, margin)) continue;
        if (nd.left < 0)
            out.insert(out.end(), b.order.begin() + nd.start, b.order.begin() + nd.start + nd.count);
        else
        {
            stack.push_back(nd.left);
            stack.push_back(nd.right);
        }
    }
    return out;
}

} // namespace sdfo
