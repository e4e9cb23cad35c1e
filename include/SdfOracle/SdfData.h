// Host SDF data of the C++ API (SdfData.h of the reference): the dense
// fields and the sparse brick store, with the same members and accessors.
// Builders fill them from the GPU; SdfOracle uploads them for queries.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "SdfOracle/GridSpec.h"
#include "SdfOracle/SdfCommon.h"

namespace sdf {

struct SdfData
{
    static constexpr int kMaxContactCandidates = kMaxSdfQueryCandidates;

    GridSpec spec;
    std::vector<double> phi0;       // signed distance per voxel
    std::vector<Vector3d> n0;       // unit gradient per voxel
    std::vector<Matrix3d> H0;       // Hessian per voxel
    std::vector<Vector3d> witness0; // closest surface point per voxel

    int contactCandidateStride{0};
    std::vector<uint8_t> contactCandidateCount;
    std::vector<double> contactCandidatePhi0;
    std::vector<Vector3d> contactCandidateNormal0;
    std::vector<int32_t> contactCandidateBranchId;

    void resize(const GridSpec& s)
    {
        spec = s;
        const int n = s.totalVoxels();
        phi0.assign(size_t(n), 0.0);
        n0.assign(size_t(n), Vector3d::UnitX());
        H0.assign(size_t(n), Matrix3d::Zero());
        witness0.assign(size_t(n), Vector3d::Zero());
        clearContactCandidates();
    }

    void resizeContactCandidates(int maxCandidates)
    {
        const int stride = std::max(0, std::min(maxCandidates, kMaxContactCandidates));
        contactCandidateStride = stride;
        const size_t n = size_t(spec.totalVoxels());
        contactCandidateCount.assign(n, 0);
        contactCandidatePhi0.assign(n * size_t(stride), 0.0);
        contactCandidateNormal0.assign(n * size_t(stride), Vector3d::UnitX());
        contactCandidateBranchId.assign(n * size_t(stride), -1);
    }

    void clearContactCandidates()
    {
        contactCandidateStride = 0;
        contactCandidateCount.clear();
        contactCandidatePhi0.clear();
        contactCandidateNormal0.clear();
        contactCandidateBranchId.clear();
    }

    bool hasContactCandidates() const
    {
        const size_t n = size_t(spec.totalVoxels());
        const size_t ns = n * size_t(std::max(contactCandidateStride, 0));
        return contactCandidateStride > 0 && contactCandidateCount.size() == n &&
               contactCandidatePhi0.size() == ns && contactCandidateNormal0.size() == ns &&
               contactCandidateBranchId.size() == ns;
    }

    int contactCandidateIndex(int voxelIndex, int candidateIndex) const
    {
        return voxelIndex * contactCandidateStride + candidateIndex;
    }

    double& phi(int i, int j, int k) { return phi0[size_t(spec.linearIndex(i, j, k))]; }
    double phi(int i, int j, int k) const { return phi0[size_t(spec.linearIndex(i, j, k))]; }
    Vector3d& normal(int i, int j, int k) { return n0[size_t(spec.linearIndex(i, j, k))]; }
    const Vector3d& normal(int i, int j, int k) const { return n0[size_t(spec.linearIndex(i, j, k))]; }
    Matrix3d& hessian(int i, int j, int k) { return H0[size_t(spec.linearIndex(i, j, k))]; }
    const Matrix3d& hessian(int i, int j, int k) const { return H0[size_t(spec.linearIndex(i, j, k))]; }
    Vector3d& witness(int i, int j, int k) { return witness0[size_t(spec.linearIndex(i, j, k))]; }
    const Vector3d& witness(int i, int j, int k) const
    {
        return witness0[size_t(spec.linearIndex(i, j, k))];
    }
};

struct SparseSdfBlock
{
    std::array<int, 3> index{0, 0, 0};
    std::vector<float> phi0;
    std::vector<float> normal0;  // x, y, z per local voxel
    std::vector<float> hessian0; // Hxx, Hxy, Hxz, Hyy, Hyz, Hzz per local voxel
    std::vector<uint8_t> contactCandidateCount;
    std::vector<float> contactCandidatePhi0;
    std::vector<float> contactCandidateNormal0; // x, y, z per local candidate
    std::vector<int32_t> contactCandidateBranchId;

    int localIndex(int x, int y, int z, int blockDim) const { return x + blockDim * (y + blockDim * z); }
    int contactCandidateIndex(int localVoxelIndex, int candidateIndex, int stride) const
    {
        return localVoxelIndex * stride + candidateIndex;
    }
    bool hasContactCandidates(int blockVoxelCount, int stride) const
    {
        const size_t v = size_t(blockVoxelCount), s = size_t(std::max(stride, 0));
        return stride > 0 && contactCandidateCount.size() >= v && contactCandidatePhi0.size() >= v * s &&
               contactCandidateNormal0.size() >= 3 * v * s && contactCandidateBranchId.size() >= v * s;
    }
};

struct SparseSdfData
{
    GridSpec spec;
    int blockDim{8};
    int haloVoxels{2};
    double bandDistance{0.0};
    bool hasHessian{true};
    int contactCandidateStride{0};
    std::vector<SparseSdfBlock> blocks;
    std::unordered_map<uint64_t, size_t> blockLookup;

    // 21 bits per block coordinate
    static uint64_t key(int bi, int bj, int bk)
    {
        return uint64_t(uint32_t(bi)) | (uint64_t(uint32_t(bj)) << 21) | (uint64_t(uint32_t(bk)) << 42);
    }

    // the first block of a key wins
    void rebuildIndex()
    {
        blockLookup.clear();
        blockLookup.reserve(blocks.size());
        for (size_t i = 0; i < blocks.size(); ++i)
            blockLookup.emplace(key(blocks[i].index[0], blocks[i].index[1], blocks[i].index[2]), i);
    }

    bool empty() const { return blocks.empty(); }
    int blockVoxelCount() const { return blockDim * blockDim * blockDim; }
    size_t totalStoredVoxels() const { return blocks.size() * size_t(blockVoxelCount()); }

    bool hasContactCandidates() const
    {
        if (contactCandidateStride <= 0 || blocks.empty()) return false;
        for (const SparseSdfBlock& b : blocks)
            if (!b.hasContactCandidates(blockVoxelCount(), contactCandidateStride)) return false;
        return true;
    }

    const SparseSdfBlock* blockForVoxel(int i, int j, int k) const
    {
        if (!spec.inBounds(i, j, k) || blockDim <= 0) return nullptr;
        const auto it = blockLookup.find(key(i / blockDim, j / blockDim, k / blockDim));
        return it == blockLookup.end() ? nullptr : &blocks[it->second];
    }

    bool sample(int i, int j, int k, double& phi, Vector3d& normal, Matrix3d* hessian = nullptr) const
    {
        const SparseSdfBlock* b = blockForVoxel(i, j, k);
        return b != nullptr && sampleFromBlock(*b, i, j, k, phi, normal, hessian);
    }

    bool sampleFromBlock(const SparseSdfBlock& b, int i, int j, int k, double& phi, Vector3d& normal,
                         Matrix3d* hessian = nullptr) const
    {
        const int lx = i - b.index[0] * blockDim, ly = j - b.index[1] * blockDim,
                  lz = k - b.index[2] * blockDim;
        if (lx < 0 || lx >= blockDim || ly < 0 || ly >= blockDim || lz < 0 || lz >= blockDim) return false;
        const int local = b.localIndex(lx, ly, lz, blockDim);
        const int vc = blockVoxelCount();
        if (local < 0 || local >= vc || b.phi0.size() < size_t(vc) || b.normal0.size() < size_t(3 * vc))
            return false;
        phi = double(b.phi0[size_t(local)]);
        normal = Vector3d(double(b.normal0[3 * size_t(local)]), double(b.normal0[3 * size_t(local) + 1]),
                          double(b.normal0[3 * size_t(local) + 2]));
        if (hessian != nullptr)
        {
            if (!hasHessian || b.hessian0.size() < size_t(6 * vc)) return false;
            const float* h = &b.hessian0[6 * size_t(local)];
            const double v[3][3] = {{h[0], h[1], h[2]}, {h[1], h[3], h[4]}, {h[2], h[4], h[5]}};
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) (*hessian)(r, c) = v[r][c];
        }
        return true;
    }
};

} // namespace sdf
