// sdf::KdTree -- the reference's exact nearest-triangle tree
// (plugins/SdfOracle/include/SdfOracle/KdTree.h:18-170) over libsdfo: the tree
// is built on the host exactly as the reference builds it and searched on the
// GPU, each point's result bit-identical to the reference's recursive search.
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "sdfo.h"
#include "SdfOracle/SdfCommon.h"
#include "SdfOracle/detail/Bridge.h"

namespace sdf {

struct ClosestPointResult
{
    Vector3d point{0, 0, 0};
    int triangleIndex{-1};
    double distance{0.0};
    Vector3d faceNormal{0, 0, 1};
};

namespace detail {
struct MeshDeleter
{
    void operator()(sdfo_mesh* m) const { sdfo_mesh_free(m); }
};
} // namespace detail

class KdTree
{
public:
    struct SignedResult
    {
        double signedDistance{0.0};
        Vector3d closestPoint{0, 0, 0};
        Vector3d normal{0, 0, 1};
        int triangleIndex{-1};
    };

    KdTree() = default;

    void build(const std::vector<Vector3d>& vertices, const std::vector<std::array<int, 3>>& triangles)
    {
        std::vector<double> V(3 * vertices.size());
        for (size_t i = 0; i < vertices.size(); ++i)
            for (int k = 0; k < 3; ++k) V[3 * i + size_t(k)] = vertices[i][k];
        std::vector<int32_t> T(3 * triangles.size());
        for (size_t i = 0; i < triangles.size(); ++i)
            for (int k = 0; k < 3; ++k) T[3 * i + size_t(k)] = triangles[i][size_t(k)];
        sdfo_mesh* m = nullptr;
        detail::check(sdfo_mesh_build(V.data(), int64_t(vertices.size()), T.data(),
                                      int64_t(triangles.size()), &m));
        m_mesh.reset(m);
        m_count = int(triangles.size());
    }

    ClosestPointResult findNearest(const Vector3d& p) const { return findNearestBatch({p})[0]; }
    SignedResult findSignedNearest(const Vector3d& p) const { return findSignedNearestBatch({p})[0]; }

    // Many points in one launch.
    std::vector<ClosestPointResult> findNearestBatch(const std::vector<Vector3d>& pts) const
    {
        std::vector<ClosestPointResult> out(pts.size());
        Raw r = run(pts, false);
        for (size_t i = 0; i < pts.size(); ++i)
        {
            out[i].point = Vector3d(r.cp[3 * i], r.cp[3 * i + 1], r.cp[3 * i + 2]);
            out[i].triangleIndex = r.tri[i];
            out[i].distance = r.d[i];
            out[i].faceNormal = Vector3d(r.n[3 * i], r.n[3 * i + 1], r.n[3 * i + 2]);
        }
        return out;
    }

    std::vector<SignedResult> findSignedNearestBatch(const std::vector<Vector3d>& pts) const
    {
        std::vector<SignedResult> out(pts.size());
        Raw r = run(pts, true);
        for (size_t i = 0; i < pts.size(); ++i)
        {
            out[i].signedDistance = r.d[i];
            out[i].closestPoint = Vector3d(r.cp[3 * i], r.cp[3 * i + 1], r.cp[3 * i + 2]);
            out[i].normal = Vector3d(r.n[3 * i], r.n[3 * i + 1], r.n[3 * i + 2]);
            out[i].triangleIndex = r.tri[i];
        }
        return out;
    }

    int triangleCount() const { return m_count; }

private:
    struct Raw
    {
        std::vector<double> d, cp, n;
        std::vector<int32_t> tri;
    };

    Raw run(const std::vector<Vector3d>& pts, bool signedDist) const
    {
        if (!m_mesh)
        {
            // an unbuilt tree compares nothing (KdTree.h:123-133)
            sdfo_mesh* m = nullptr;
            detail::check(sdfo_mesh_build(nullptr, 0, nullptr, 0, &m));
            m_mesh.reset(m);
        }
        const size_t n = pts.size();
        std::vector<double> P(3 * n);
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) P[3 * i + size_t(k)] = pts[i][k];
        Raw r{std::vector<double>(n), std::vector<double>(3 * n), std::vector<double>(3 * n),
              std::vector<int32_t>(n)};
        detail::check(sdfo_mesh_query_host(m_mesh.get(), P.data(), int64_t(n), signedDist ? 1 : 0,
                                           r.d.data(), r.cp.data(), r.n.data(), r.tri.data(),
                                           nullptr));
        return r;
    }

    mutable std::unique_ptr<sdfo_mesh, detail::MeshDeleter> m_mesh;
    int m_count = 0;
};

} // namespace sdf
