// Test-support entry points (include/sdfo_debug.h).
#include "absphi_sort.hpp"
#include "sdfo_debug.h"

extern "C" int sdfo_debug_sort_by_abs(const double* values, int32_t n, int32_t* order)
{
    if (n < 0 || n > sdfo::kMergeCap || (n > 0 && (!values || !order))) return SDFO_ERR_INVALID_ARGUMENT;
    int v[sdfo::kMergeCap];
    for (int i = 0; i < n; ++i) v[i] = i;
    sdfo::sortPerm(v, n, sdfo::AbsPhiOrder{values});
    for (int i = 0; i < n; ++i) order[i] = v[i];
    return SDFO_OK;
}
