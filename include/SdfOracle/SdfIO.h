// "SDFO" 0.1.0 files of the C++ API (SdfIO.h:29-45 of the reference).
// Saving writes the host structs byte for byte as the reference does
// (returns false on failure); loading goes through the GPU grid and is
// downloaded into the host structs; load errors throw with the reference's
// messages.
#pragma once

#include <string>
#include <vector>

#include "SdfOracle/SdfData.h"
#include "SdfOracle/detail/Bridge.h"

namespace sdf {

inline bool saveSdf(const SdfData& data, const std::string& path)
{
    if (data.phi0.empty()) return false;
    detail::DenseArrays a(data);
    const bool c = a.stride > 0;
    return sdfo_save_dense_host(path.c_str(), a.bmin, a.bmax, a.voxel, a.shape, data.phi0.data(), a.n3.data(),
                                a.h6.data(), a.w3.data(), a.stride, c ? data.contactCandidateCount.data() : nullptr,
                                c ? data.contactCandidatePhi0.data() : nullptr, c ? a.cn3.data() : nullptr,
                                c ? data.contactCandidateBranchId.data() : nullptr) == SDFO_OK;
}

inline bool saveSparseSdf(const SparseSdfData& data, const std::string& path)
{
    detail::SparseArrays a(data);
    const bool c = a.stride > 0;
    return sdfo_save_sparse_host(path.c_str(), a.bmin, a.bmax, a.voxel, a.shape, data.blockDim, data.haloVoxels,
                                 data.bandDistance, data.hasHessian ? 1 : 0, a.stride, int32_t(data.blocks.size()),
                                 a.bijk.data(), a.phi.data(), a.n3.data(), data.hasHessian ? a.h6.data() : nullptr,
                                 c ? a.cc.data() : nullptr, c ? a.cphi.data() : nullptr, c ? a.cn3.data() : nullptr,
                                 c ? a.cbr.data() : nullptr) == SDFO_OK;
}

inline bool isValidSdf(const std::string& path) { return sdfo_is_valid_file(path.c_str()) == 1; }
inline bool isSparseSdf(const std::string& path) { return sdfo_is_sparse_file(path.c_str()) == 1; }

inline SdfData loadSdf(const std::string& path)
{
    sdfo_grid* g = nullptr;
    detail::check(sdfo_grid_load(path.c_str(), &g));
    detail::GridHandle h(g);
    if (detail::infoOf(g).sparse) throw std::runtime_error("SdfOracle data file is not dense 0.1.0 layout");
    return detail::downloadDense(g);
}

inline SparseSdfData loadSparseSdf(const std::string& path)
{
    sdfo_grid* g = nullptr;
    detail::check(sdfo_grid_load(path.c_str(), &g));
    detail::GridHandle h(g);
    if (!detail::infoOf(g).sparse) throw std::runtime_error("SdfOracle data file is not sparse block 0.1.0");
    return detail::downloadSparse(g);
}

} // namespace sdf
