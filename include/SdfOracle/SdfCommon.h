// Query results of the C++ API (SdfCommon.h of the reference).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "SdfOracle/Types.h"

namespace sdf {

inline constexpr int kMaxSdfQueryCandidates = 8;

struct SdfQueryCandidate
{
    double gap{0.0};
    Vector3d gapNormal{1.0, 0.0, 0.0};
    Vector3d witnessPoint{0.0, 0.0, 0.0};
    int branchId{-1};
};

struct SdfQueryResult
{
    double gap{0.0};
    Vector3d gapNormal{1.0, 0.0, 0.0};
    Vector3d witnessPoint{0.0, 0.0, 0.0};
    bool valid{false};
    std::vector<SdfQueryCandidate> candidates;
};

struct SdfFastQueryResult
{
    double gap{0.0};
    Vector3d gapNormal{1.0, 0.0, 0.0};
    Vector3d witnessPoint{0.0, 0.0, 0.0};
    bool valid{false};
    int candidateCount{0};
    std::array<SdfQueryCandidate, kMaxSdfQueryCandidates> candidates{};
};

} // namespace sdf
