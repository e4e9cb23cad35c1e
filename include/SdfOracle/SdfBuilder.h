// SDF builders of the C++ API (SdfBuilder.h:24-113 of the reference): same
// parameter structs and static functions.  The build runs on the GPU
// (libsdfo) and the result is downloaded into the host structs, which are
// returned by value as in the reference.  Errors throw std::runtime_error
// with the reference's messages.
#pragma once

#include <array>
#include <chrono>
#include <functional>
#include <string>
#include <vector>

#include "SdfOracle/SdfData.h"
#include "SdfOracle/SdfSurfaceMesh.h"
#include "SdfOracle/detail/Bridge.h"

namespace sdf {

struct BuilderParams
{
    double padding{0.1};       // padding as a fraction of the bounding-box diagonal
    double voxelSize{0.0};     // explicit voxel size (0 = from targetResolution)
    int targetResolution{128}; // voxels along the longest axis
    int maxResolution{512};    // cap on any axis
    std::function<void(const std::string&)> logCallback;
};

struct SparseBuilderParams : public BuilderParams
{
    int blockDim{8};
    int haloVoxels{2};
    int bandVoxels{8};
    double bandDistance{0.0}; // explicit narrow-band half width (0 = bandVoxels * voxel)
    bool storeHessian{true};
};

struct ContactAwareBuilderParams : public BuilderParams
{
    int maxCandidates{4};
    double featureDistanceTolerance{0.0};
    double normalMergeAngleDegrees{5.0};
    int tileDim{8};
    bool useTileAcceleration{true};
    bool storeHessian{true};
    bool sparse{false};
    int blockDim{8};
    int haloVoxels{2};
    int bandVoxels{8};
    double bandDistance{0.0};
};

namespace detail {

inline sdfo_build_params toParams(const BuilderParams& p)
{
    sdfo_build_params q;
    sdfo_default_build_params(&q);
    q.padding = p.padding;
    q.voxel_size = p.voxelSize;
    q.target_resolution = p.targetResolution;
    q.max_resolution = p.maxResolution;
    return q;
}

inline void flatten(const std::vector<Vector3d>& v, const std::vector<std::array<int, 3>>& t,
                    std::vector<double>& V, std::vector<int32_t>& T)
{
    V.resize(3 * v.size());
    T.resize(3 * t.size());
    for (size_t i = 0; i < v.size(); ++i)
        for (int a = 0; a < 3; ++a) V[3 * i + size_t(a)] = v[i][a];
    for (size_t i = 0; i < t.size(); ++i)
        for (int a = 0; a < 3; ++a) T[3 * i + size_t(a)] = t[i][size_t(a)];
}

inline void log(const BuilderParams& p, const std::string& msg)
{
    if (p.logCallback) p.logCallback(msg);
}

inline void logGrid(const BuilderParams& p, const GridSpec& s, double ms)
{
    log(p, "[SdfBuilder] Grid: " + std::to_string(s.nx()) + " x " + std::to_string(s.ny()) + " x " +
               std::to_string(s.nz()) + " = " + std::to_string(s.totalVoxels()) + " voxels");
    log(p, "[SdfBuilder] GPU build: " + std::to_string(ms) + " ms");
    log(p, "[SdfBuilder] Done.");
}

inline double msSince(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

inline GridHandle buildCA(const CornerNormalTriangleMesh& mesh, const ContactAwareBuilderParams& p, bool sparse)
{
    std::vector<double> V;
    std::vector<int32_t> T;
    flatten(mesh.vertices, mesh.triangles, V, T);
    if (!mesh.vertices.empty() && !mesh.triangles.empty() && mesh.cornerNormals.size() != mesh.triangles.size())
        throw std::runtime_error(sparse ? "Sparse contact-aware SDF requires one corner-normal triplet per triangle"
                                        : "Contact-aware SDF requires one corner-normal triplet per triangle");
    std::vector<double> CN(9 * mesh.cornerNormals.size());
    for (size_t t = 0; t < mesh.cornerNormals.size(); ++t)
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) CN[9 * t + size_t(3 * c + a)] = mesh.cornerNormals[t][size_t(c)][a];
    sdfo_build_params q = toParams(p);
    q.max_candidates = p.maxCandidates;
    q.feature_distance_tolerance = p.featureDistanceTolerance;
    q.normal_merge_angle_degrees = p.normalMergeAngleDegrees;
    q.tile_dim = p.tileDim;
    q.use_tile_acceleration = p.useTileAcceleration ? 1 : 0;
    q.store_hessian = p.storeHessian ? 1 : 0;
    q.block_dim = p.blockDim;
    q.halo_voxels = p.haloVoxels;
    q.band_voxels = p.bandVoxels;
    q.band_distance = p.bandDistance;
    sdfo_grid* g = nullptr;
    check(sdfo_build_contact_aware(V.data(), int64_t(mesh.vertices.size()), T.data(), int64_t(mesh.triangles.size()),
                                   CN.empty() ? nullptr : CN.data(), &q, sparse ? 1 : 0, &g, nullptr));
    return GridHandle(g);
}

} // namespace detail

class SdfBuilder
{
public:
    static SdfData build(const std::vector<Vector3d>& vertices, const std::vector<std::array<int, 3>>& triangles,
                         const BuilderParams& params = {})
    {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<double> V;
        std::vector<int32_t> T;
        detail::flatten(vertices, triangles, V, T);
        const sdfo_build_params q = detail::toParams(params);
        sdfo_grid* g = nullptr;
        detail::check(sdfo_build_dense(V.data(), int64_t(vertices.size()), T.data(), int64_t(triangles.size()), &q,
                                       &g, nullptr));
        detail::GridHandle h(g);
        SdfData d = detail::downloadDense(h.get());
        detail::logGrid(params, d.spec, detail::msSince(t0));
        return d;
    }

    static SdfData build(const TriangleMesh& mesh, const BuilderParams& params = {})
    {
        return build(mesh.vertices, mesh.triangles, params);
    }

    static SdfData buildContactAware(const CornerNormalTriangleMesh& mesh,
                                     const ContactAwareBuilderParams& params = {})
    {
        const auto t0 = std::chrono::steady_clock::now();
        detail::GridHandle h = detail::buildCA(mesh, params, false);
        SdfData d = detail::downloadDense(h.get());
        detail::logGrid(params, d.spec, detail::msSince(t0));
        return d;
    }

    static SparseSdfData buildSparseContactAware(const CornerNormalTriangleMesh& mesh,
                                                 const ContactAwareBuilderParams& params = {})
    {
        const auto t0 = std::chrono::steady_clock::now();
        detail::GridHandle h = detail::buildCA(mesh, params, true);
        SparseSdfData d = detail::downloadSparse(h.get());
        detail::logGrid(params, d.spec, detail::msSince(t0));
        return d;
    }

    static SparseSdfData buildSparse(const std::vector<Vector3d>& vertices,
                                     const std::vector<std::array<int, 3>>& triangles,
                                     const SparseBuilderParams& params = {})
    {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<double> V;
        std::vector<int32_t> T;
        detail::flatten(vertices, triangles, V, T);
        sdfo_build_params q = detail::toParams(params);
        q.block_dim = params.blockDim;
        q.halo_voxels = params.haloVoxels;
        q.band_voxels = params.bandVoxels;
        q.band_distance = params.bandDistance;
        q.store_hessian = params.storeHessian ? 1 : 0;
        sdfo_grid* g = nullptr;
        detail::check(sdfo_build_sparse(V.data(), int64_t(vertices.size()), T.data(), int64_t(triangles.size()), &q,
                                        &g, nullptr));
        detail::GridHandle h(g);
        SparseSdfData d = detail::downloadSparse(h.get());
        detail::logGrid(params, d.spec, detail::msSince(t0));
        return d;
    }

    static SparseSdfData buildSparse(const TriangleMesh& mesh, const SparseBuilderParams& params = {})
    {
        return buildSparse(mesh.vertices, mesh.triangles, params);
    }
};

} // namespace sdf
