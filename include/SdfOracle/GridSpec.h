// Cell-centred grid description of the C++ API (GridSpec.h:17-74 of the
// reference): same members and helpers.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>

#include "SdfOracle/Types.h"

namespace sdf {

struct GridSpec
{
    Vector3d bmin{0.0, 0.0, 0.0};
    Vector3d bmax{1.0, 1.0, 1.0};
    Vector3d voxel{1.0, 1.0, 1.0};
    std::array<int, 3> shape{1, 1, 1}; // (nx, ny, nz)

    int nx() const { return shape[0]; }
    int ny() const { return shape[1]; }
    int nz() const { return shape[2]; }
    int totalVoxels() const { return shape[0] * shape[1] * shape[2]; }

    // voxel (i, j, k) is centred at bmin + voxel * (index + 1/2)
    Vector3d centerFromIndex(int i, int j, int k) const
    {
        return Vector3d(bmin[0] + voxel[0] * (i + 0.5), bmin[1] + voxel[1] * (j + 0.5),
                        bmin[2] + voxel[2] * (k + 0.5));
    }

    // nearest voxel centre, rounding half away from zero
    void indexFromWorld(const Vector3d& p, int& i, int& j, int& k) const
    {
        i = static_cast<int>(std::round((p[0] - bmin[0]) / voxel[0] - 0.5));
        j = static_cast<int>(std::round((p[1] - bmin[1]) / voxel[1] - 0.5));
        k = static_cast<int>(std::round((p[2] - bmin[2]) / voxel[2] - 0.5));
    }

    bool inBounds(int i, int j, int k) const
    {
        return i >= 0 && i < shape[0] && j >= 0 && j < shape[1] && k >= 0 && k < shape[2];
    }

    // x fastest
    int linearIndex(int i, int j, int k) const { return i + shape[0] * (j + shape[1] * k); }
};

// Smallest ladder level >= n (8 ... 2048), capped at maxRes; maxRes past the ladder.
inline int snapToStandardResolution(int n, int maxRes)
{
    for (int level = 8; level <= 2048; level *= 2)
        if (level >= n || level >= maxRes) return std::min(level, maxRes);
    return maxRes;
}

} // namespace sdf
