// FP64 point queries (SdfOracle.cpp:176-1092): the query modes of
// query_device.hpp in FP64 arithmetic on a dense grid's FP64 fields (SdfData)
// or a sparse grid's FP32 fields widened (SparseSdfData holds FP32, as in the
// reference), built without FMA contraction, so every result is the CPU
// path's bit for bit (the FP32 kernels of query.hip read the FP32 mirror).
#include <hip/hip_runtime.h>

#include "query_device.hpp"
#include "sdfo_internal.hpp"

namespace sdfo {

namespace dev {
QGrid makeQGrid(const sdfo_grid& g);
} // namespace dev

void launchQueryExactF64(const sdfo_grid& g, int mode, const double* pts, int64_t n, double* gap,
                         double* grad3, uint8_t* valid, double* normal3, double* witness3,
                         int32_t* branch, hipStream_t s)
{
    if (n <= 0) return;
    QGrid q = dev::makeQGrid(g);
    // Trilinear and Tricubic (and ContactAware without candidates, which
    // falls back to Tricubic) read the 8 corners of a cell: from the FP64
    // cell records when the grid has them
    if (!g.sparse && (mode == SDFO_TRILINEAR || mode == SDFO_TRICUBIC ||
                      (mode == SDFO_CONTACT_AWARE && q.stride <= 0)))
        q.cellRec64 = cellRecords64(g, s);
    using V = P3<double>;
    const dim3 grid(unsigned((n + 255) / 256)), block(256);
    const V* p = reinterpret_cast<const V*>(pts);
    V* gr = reinterpret_cast<V*>(grad3);
    V* nm = reinterpret_cast<V*>(normal3);
    V* wt = reinterpret_cast<V*>(witness3);
    switch (mode)
    {
    case SDFO_FIRST_ORDER:
        hipLaunchKernelGGL((queryKernel<double, double, SDFO_FIRST_ORDER, true>), grid, block, 0, s,
                           q, p, n, gap, gr, valid, nm, wt, branch);
        break;
    case SDFO_SECOND_ORDER:
        hipLaunchKernelGGL((queryKernel<double, double, SDFO_SECOND_ORDER, true>), grid, block, 0, s,
                           q, p, n, gap, gr, valid, nm, wt, branch);
        break;
    case SDFO_TRILINEAR:
        hipLaunchKernelGGL((queryKernel<double, double, SDFO_TRILINEAR, true>), grid, block, 0, s, q,
                           p, n, gap, gr, valid, nm, wt, branch);
        break;
    case SDFO_TRICUBIC:
        hipLaunchKernelGGL((queryKernel<double, double, SDFO_TRICUBIC, true>), grid, block, 0, s, q,
                           p, n, gap, gr, valid, nm, wt, branch);
        break;
    default:
        hipLaunchKernelGGL((queryKernel<double, double, SDFO_CONTACT_AWARE, true>), grid, block, 0, s,
                           q, p, n, gap, gr, valid, nm, wt, branch);
        break;
    }
    SDFO_HIP(hipGetLastError());
}

} // namespace sdfo
