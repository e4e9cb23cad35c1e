// Test-support entry points (include/sdfo_debug.h).
#include "absphi_sort.hpp"
#include "sdfo_debug.h"
#include "sdfo_internal.hpp"

extern "C" int sdfo_debug_sort_by_abs(const double* values, int32_t n, int32_t* order)
{
    if (n < 0 || n > sdfo::kMergeCap || (n > 0 && (!values || !order))) return SDFO_ERR_INVALID_ARGUMENT;
    int v[sdfo::kMergeCap];
    for (int i = 0; i < n; ++i) v[i] = i;
    sdfo::sortPerm(v, n, sdfo::AbsPhiOrder{values});
    for (int i = 0; i < n; ++i) order[i] = v[i];
    return SDFO_OK;
}

extern "C" int sdfo_debug_build_phases(double* ms, int32_t cap)
{
    const double* ph = sdfo::lastBuildPhases();
    const int n = cap < sdfo::kBuildPhases ? cap : sdfo::kBuildPhases;
    for (int i = 0; i < n && ms; ++i) ms[i] = ph[i];
    return sdfo::kBuildPhases;
}

extern "C" int sdfo_debug_force_replay(int32_t on)
{
    sdfo::forceReplay() = on != 0;
    return SDFO_OK;
}

extern "C" int sdfo_debug_fill_budget(int32_t budget, int32_t heavy_cap)
{
    if (budget < 0 || heavy_cap < 0) return SDFO_ERR_INVALID_ARGUMENT;
    sdfo::debugFillBudget() = budget;
    sdfo::debugHeavyCap() = heavy_cap;
    return SDFO_OK;
}

extern "C" int sdfo_debug_ca_merge_cap(int32_t first, int32_t second)
{
    if ((first != 0 && first != 1 && first != 4 && first != 8) || second < 0)
        return SDFO_ERR_INVALID_ARGUMENT;
    sdfo::debugMergeCap() = first;
    sdfo::debugSecondCap() = second;
    return SDFO_OK;
}

namespace sdfo {
void setSkipCellRecords(bool on);
}

extern "C" int sdfo_debug_skip_cell_records(int32_t on)
{
    sdfo::setSkipCellRecords(on != 0);
    return SDFO_OK;
}

extern "C" int sdfo_debug_random_read_floor(int32_t record_bytes, int64_t n_reads,
                                            int64_t array_bytes, int32_t reps, double* best_us)
{
    try
    {
        if (!best_us) return SDFO_ERR_INVALID_ARGUMENT;
        *best_us = sdfo::randomRecordReadUs(record_bytes, n_reads, array_bytes, reps);
        return SDFO_OK;
    }
    catch (const sdfo::Error& e)
    {
        return e.code;
    }
    catch (...)
    {
        return SDFO_ERR_INVALID_ARGUMENT;
    }
}

extern "C" int sdfo_debug_tuner_pin(const char* key, int32_t variant)
{
    if (!key || variant < 0) return SDFO_ERR_INVALID_ARGUMENT;
    sdfo::tunerRecord(key, {{int(variant), 0.0}}, nullptr);
    return SDFO_OK;
}
