// The query engine of the C++ API (SdfOracle.h:11-47 of the reference).  The
// host SdfData / SparseSdfData stays authoritative (data(), sparseData());
// the first query uploads it once into a GPU grid and every query runs
// there (FP64 points, libsdfo's FP64 query kernels; ContactAware candidates
// from the FP64 / FP32 candidate arrays).  Queries never throw: a device
// error yields invalid results, as an invalid point does.  const queries are
// safe to call from several threads.
#pragma once

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "SdfOracle/SdfCommon.h"
#include "SdfOracle/SdfData.h"
#include "SdfOracle/SdfIO.h"
#include "SdfOracle/detail/Bridge.h"

namespace sdf {

enum class QueryMode
{
    FirstOrder,
    SecondOrder,
    Trilinear,
    Tricubic,
    ContactAware
};

class SdfOracle
{
public:
    SdfOracle() = default;
    explicit SdfOracle(SdfData data) : m_data(std::move(data)) {}
    explicit SdfOracle(SparseSdfData data) : m_sparseData(std::move(data)), m_sparse(true) {}

    static SdfOracle loadFromFile(const std::string& path)
    {
        if (isSparseSdf(path)) return SdfOracle(loadSparseSdf(path));
        return SdfOracle(loadSdf(path));
    }

    SdfQueryResult query(const Vector3d& point, QueryMode mode = QueryMode::SecondOrder) const
    {
        return queryBatch(std::vector<Vector3d>{point}, mode)[0];
    }

    SdfFastQueryResult queryFast(const Vector3d& point, QueryMode mode = QueryMode::SecondOrder) const
    {
        std::vector<SdfFastQueryResult> r;
        run(std::vector<Vector3d>{point}, mode, &r, nullptr);
        return r[0];
    }

    std::vector<SdfQueryResult> queryBatch(const std::vector<Vector3d>& points,
                                           QueryMode mode = QueryMode::SecondOrder) const
    {
        std::vector<SdfQueryResult> r;
        run(points, mode, nullptr, &r);
        return r;
    }

    // Many points at once, in the fixed-size result form (no per-point allocation).
    std::vector<SdfFastQueryResult> queryFastBatch(const std::vector<Vector3d>& points,
                                                   QueryMode mode = QueryMode::SecondOrder) const
    {
        std::vector<SdfFastQueryResult> r;
        run(points, mode, &r, nullptr);
        return r;
    }

    const SdfData& data() const { return m_data; }
    const SparseSdfData& sparseData() const { return m_sparseData; }
    bool isSparse() const { return m_sparse; }
    const GridSpec& gridSpec() const { return m_sparse ? m_sparseData.spec : m_data.spec; }
    bool isValid() const { return m_sparse ? !m_sparseData.empty() : !m_data.phi0.empty(); }

private:
    struct Device
    {
        std::once_flag once;
        detail::GridHandle grid;
    };

    const sdfo_grid* device() const
    {
        Device& d = *m_device;
        std::call_once(d.once, [&] {
            try
            {
                d.grid = m_sparse ? detail::uploadSparse(m_sparseData) : detail::uploadDense(m_data);
            }
            catch (...)
            {
                d.grid.reset();
            }
        });
        return d.grid.get();
    }

    bool hasCandidates() const
    {
        return m_sparse ? m_sparseData.hasContactCandidates() : m_data.hasContactCandidates();
    }

    // SdfOracle.cpp:176-444: mode dispatch; every valid result carries at
    // least one candidate (the primary, branch 0) except ContactAware ones,
    // which carry the voxel's candidates.
    void run(const std::vector<Vector3d>& pts, QueryMode mode, std::vector<SdfFastQueryResult>* fast,
             std::vector<SdfQueryResult>* full) const
    {
        const size_t n = pts.size();
        if (fast) fast->assign(n, SdfFastQueryResult{});
        if (full) full->assign(n, SdfQueryResult{});
        if (n == 0 || !isValid()) return;
        const sdfo_grid* g = device();
        if (!g) return;
        std::vector<double> P(3 * n), gap(n), nrm(3 * n), wit(3 * n);
        std::vector<uint8_t> valid(n);
        for (size_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) P[3 * i + size_t(a)] = pts[i][a];
        const bool ca = mode == QueryMode::ContactAware && hasCandidates();
        std::vector<uint8_t> cc(ca ? n : 0);
        std::vector<double> cg(ca ? 8 * n : 0), cn(ca ? 24 * n : 0), cw(ca ? 24 * n : 0);
        std::vector<int32_t> cb(ca ? 8 * n : 0);
        const int rc = sdfo_query_host_f64(g, int32_t(mode), P.data(), int64_t(n), gap.data(), nrm.data(), wit.data(),
                                           valid.data(), nullptr, ca ? cc.data() : nullptr, ca ? cg.data() : nullptr,
                                           ca ? cn.data() : nullptr, ca ? cw.data() : nullptr,
                                           ca ? cb.data() : nullptr, nullptr);
        if (rc != SDFO_OK) return;
        for (size_t i = 0; i < n; ++i)
        {
            SdfFastQueryResult f;
            f.valid = valid[i] != 0;
            if (f.valid)
            {
                f.gap = gap[i];
                f.gapNormal = Vector3d(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
                f.witnessPoint = Vector3d(wit[3 * i], wit[3 * i + 1], wit[3 * i + 2]);
                if (ca && cc[i] > 0)
                {
                    f.candidateCount = int(cc[i]);
                    for (int c = 0; c < f.candidateCount; ++c)
                    {
                        const size_t o = 8 * i + size_t(c);
                        SdfQueryCandidate& q = f.candidates[size_t(c)];
                        q.gap = cg[o];
                        q.gapNormal = Vector3d(cn[3 * o], cn[3 * o + 1], cn[3 * o + 2]);
                        q.witnessPoint = Vector3d(cw[3 * o], cw[3 * o + 1], cw[3 * o + 2]);
                        q.branchId = cb[o];
                    }
                }
                else
                {
                    f.candidateCount = 1;
                    f.candidates[0] = SdfQueryCandidate{f.gap, f.gapNormal, f.witnessPoint, 0};
                }
            }
            if (fast) (*fast)[i] = f;
            if (full)
            {
                SdfQueryResult& r = (*full)[i];
                r.gap = f.gap;
                r.gapNormal = f.gapNormal;
                r.witnessPoint = f.witnessPoint;
                r.valid = f.valid;
                for (int c = 0; c < f.candidateCount; ++c) r.candidates.push_back(f.candidates[size_t(c)]);
            }
        }
    }

    SdfData m_data;
    SparseSdfData m_sparseData;
    bool m_sparse{false};
    std::shared_ptr<Device> m_device{std::make_shared<Device>()};
};

} // namespace sdf
