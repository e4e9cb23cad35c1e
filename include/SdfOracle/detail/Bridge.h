// Internal: conversions between the C++ API's host structs and libsdfo's
// device grids (include/sdfo.h).  Not part of the public API.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sdfo.h"
#include "SdfOracle/SdfData.h"

namespace sdf {
namespace detail {

struct GridDeleter
{
    void operator()(sdfo_grid* g) const { sdfo_grid_free(g); }
};
using GridHandle = std::unique_ptr<sdfo_grid, GridDeleter>;

[[noreturn]] inline void fail() { throw std::runtime_error(sdfo_last_error()); }
inline void check(int rc)
{
    if (rc != SDFO_OK) fail();
}

inline sdfo_grid_info infoOf(const sdfo_grid* g)
{
    sdfo_grid_info info{};
    check(sdfo_grid_get_info(g, &info));
    return info;
}

inline GridSpec specOf(const sdfo_grid_info& info)
{
    GridSpec s;
    for (int a = 0; a < 3; ++a)
    {
        s.bmin[a] = info.bmin[a];
        s.bmax[a] = info.bmax[a];
        s.voxel[a] = info.voxel[a];
        s.shape[size_t(a)] = info.shape[a];
    }
    return s;
}

inline void specArrays(const GridSpec& s, double* bmin, double* bmax, double* voxel, int32_t* shape)
{
    for (int a = 0; a < 3; ++a)
    {
        bmin[a] = s.bmin[a];
        bmax[a] = s.bmax[a];
        voxel[a] = s.voxel[a];
        shape[a] = s.shape[size_t(a)];
    }
}

// Hessians travel as Hxx, Hxy, Hxz, Hyy, Hyz, Hzz.
inline Matrix3d fromH6(const double* h)
{
    Matrix3d m = Matrix3d::Zero();
    const double v[3][3] = {{h[0], h[1], h[2]}, {h[1], h[3], h[4]}, {h[2], h[4], h[5]}};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m(r, c) = v[r][c];
    return m;
}

inline void toH6(const Matrix3d& m, double* h)
{
    h[0] = m(0, 0);
    h[1] = m(0, 1);
    h[2] = m(0, 2);
    h[3] = m(1, 1);
    h[4] = m(1, 2);
    h[5] = m(2, 2);
}

inline SdfData downloadDense(const sdfo_grid* g)
{
    const sdfo_grid_info info = infoOf(g);
    SdfData d;
    d.resize(specOf(info));
    const size_t n = d.phi0.size();
    std::vector<double> n3(3 * n), h6(6 * n), w3(3 * n);
    check(sdfo_grid_download_dense(g, d.phi0.data(), n3.data(), h6.data(), w3.data(), nullptr));
    for (size_t q = 0; q < n; ++q)
    {
        d.n0[q] = Vector3d(n3[3 * q], n3[3 * q + 1], n3[3 * q + 2]);
        d.H0[q] = fromH6(&h6[6 * q]);
        d.witness0[q] = Vector3d(w3[3 * q], w3[3 * q + 1], w3[3 * q + 2]);
    }
    if (info.candidate_stride > 0)
    {
        d.resizeContactCandidates(info.candidate_stride);
        const size_t ns = n * size_t(info.candidate_stride);
        std::vector<double> cn(3 * ns);
        check(sdfo_grid_download_candidates(g, d.contactCandidateCount.data(), d.contactCandidatePhi0.data(),
                                            cn.data(), d.contactCandidateBranchId.data()));
        for (size_t q = 0; q < ns; ++q)
            d.contactCandidateNormal0[q] = Vector3d(cn[3 * q], cn[3 * q + 1], cn[3 * q + 2]);
    }
    return d;
}

inline SparseSdfData downloadSparse(const sdfo_grid* g)
{
    const sdfo_grid_info info = infoOf(g);
    SparseSdfData d;
    d.spec = specOf(info);
    d.blockDim = info.block_dim;
    d.haloVoxels = info.halo_voxels;
    d.bandDistance = info.band_distance;
    d.hasHessian = info.has_hessian != 0;
    d.contactCandidateStride = info.candidate_stride;
    const size_t nb = size_t(info.block_count), B = size_t(info.block_dim) * info.block_dim * info.block_dim;
    const size_t S = size_t(std::max(info.candidate_stride, 0));
    std::vector<int32_t> bijk(3 * nb);
    std::vector<float> phi(nb * B), n3(3 * nb * B), h6(d.hasHessian ? 6 * nb * B : 0);
    std::vector<uint8_t> cc(S ? nb * B : 0);
    std::vector<float> cphi(nb * B * S), cn3(3 * nb * B * S);
    std::vector<int32_t> cbr(nb * B * S);
    check(sdfo_grid_download_sparse(g, bijk.data(), phi.data(), n3.data(), d.hasHessian ? h6.data() : nullptr,
                                    nullptr, S ? cc.data() : nullptr, S ? cphi.data() : nullptr,
                                    S ? cn3.data() : nullptr, S ? cbr.data() : nullptr));
    d.blocks.resize(nb);
    for (size_t b = 0; b < nb; ++b)
    {
        SparseSdfBlock& blk = d.blocks[b];
        blk.index = {bijk[3 * b], bijk[3 * b + 1], bijk[3 * b + 2]};
        blk.phi0.assign(phi.begin() + long(b * B), phi.begin() + long((b + 1) * B));
        blk.normal0.assign(n3.begin() + long(3 * b * B), n3.begin() + long(3 * (b + 1) * B));
        if (d.hasHessian) blk.hessian0.assign(h6.begin() + long(6 * b * B), h6.begin() + long(6 * (b + 1) * B));
        if (S)
        {
            blk.contactCandidateCount.assign(cc.begin() + long(b * B), cc.begin() + long((b + 1) * B));
            blk.contactCandidatePhi0.assign(cphi.begin() + long(b * B * S), cphi.begin() + long((b + 1) * B * S));
            blk.contactCandidateNormal0.assign(cn3.begin() + long(3 * b * B * S),
                                               cn3.begin() + long(3 * (b + 1) * B * S));
            blk.contactCandidateBranchId.assign(cbr.begin() + long(b * B * S), cbr.begin() + long((b + 1) * B * S));
        }
    }
    d.rebuildIndex();
    return d;
}

// Flattened dense arrays in the C ABI layout.
struct DenseArrays
{
    double bmin[3], bmax[3], voxel[3];
    int32_t shape[3];
    std::vector<double> n3, h6, w3, cn3;
    int32_t stride = 0;
    explicit DenseArrays(const SdfData& d)
    {
        specArrays(d.spec, bmin, bmax, voxel, shape);
        const size_t n = d.phi0.size();
        n3.resize(3 * n);
        h6.resize(6 * n);
        w3.resize(3 * n);
        for (size_t q = 0; q < n; ++q)
        {
            for (int a = 0; a < 3; ++a)
            {
                n3[3 * q + size_t(a)] = d.n0[q][a];
                w3[3 * q + size_t(a)] = q < d.witness0.size() ? d.witness0[q][a] : 0.0;
            }
            if (q < d.H0.size()) toH6(d.H0[q], &h6[6 * q]);
        }
        if (d.hasContactCandidates())
        {
            stride = d.contactCandidateStride;
            cn3.resize(3 * d.contactCandidateNormal0.size());
            for (size_t q = 0; q < d.contactCandidateNormal0.size(); ++q)
                for (int a = 0; a < 3; ++a) cn3[3 * q + size_t(a)] = d.contactCandidateNormal0[q][a];
        }
    }
};

// Flattened sparse arrays in the C ABI layout (blocks in stored order).
struct SparseArrays
{
    double bmin[3], bmax[3], voxel[3];
    int32_t shape[3];
    std::vector<int32_t> bijk, cbr;
    std::vector<float> phi, n3, h6, cphi, cn3;
    std::vector<uint8_t> cc;
    int32_t stride = 0;
    explicit SparseArrays(const SparseSdfData& d)
    {
        specArrays(d.spec, bmin, bmax, voxel, shape);
        const size_t B = size_t(d.blockVoxelCount());
        stride = d.hasContactCandidates() ? d.contactCandidateStride : 0;
        const size_t S = size_t(stride);
        for (const SparseSdfBlock& b : d.blocks)
        {
            bijk.insert(bijk.end(), b.index.begin(), b.index.end());
            auto take = [](std::vector<float>& dst, const std::vector<float>& src, size_t n) {
                for (size_t q = 0; q < n; ++q) dst.push_back(q < src.size() ? src[q] : 0.0f);
            };
            take(phi, b.phi0, B);
            take(n3, b.normal0, 3 * B);
            if (d.hasHessian) take(h6, b.hessian0, 6 * B);
            if (S)
            {
                cc.insert(cc.end(), b.contactCandidateCount.begin(), b.contactCandidateCount.begin() + long(B));
                take(cphi, b.contactCandidatePhi0, B * S);
                take(cn3, b.contactCandidateNormal0, 3 * B * S);
                cbr.insert(cbr.end(), b.contactCandidateBranchId.begin(),
                           b.contactCandidateBranchId.begin() + long(B * S));
            }
        }
    }
};

inline GridHandle uploadDense(const SdfData& d)
{
    DenseArrays a(d);
    sdfo_grid* g = nullptr;
    check(sdfo_grid_from_dense(a.bmin, a.bmax, a.voxel, a.shape, d.phi0.data(), a.n3.data(), a.h6.data(),
                               a.w3.data(), a.stride, a.stride ? d.contactCandidateCount.data() : nullptr,
                               a.stride ? d.contactCandidatePhi0.data() : nullptr,
                               a.stride ? a.cn3.data() : nullptr,
                               a.stride ? d.contactCandidateBranchId.data() : nullptr, &g));
    return GridHandle(g);
}

inline GridHandle uploadSparse(const SparseSdfData& d)
{
    SparseArrays a(d);
    sdfo_grid* g = nullptr;
    check(sdfo_grid_from_sparse(a.bmin, a.bmax, a.voxel, a.shape, d.blockDim, d.haloVoxels, d.bandDistance,
                                d.hasHessian ? 1 : 0, a.stride, int32_t(d.blocks.size()), a.bijk.data(),
                                a.phi.data(), a.n3.data(), d.hasHessian ? a.h6.data() : nullptr,
                                a.stride ? a.cc.data() : nullptr, a.stride ? a.cphi.data() : nullptr,
                                a.stride ? a.cn3.data() : nullptr, a.stride ? a.cbr.data() : nullptr, &g));
    return GridHandle(g);
}

} // namespace detail
} // namespace sdf
