/*
 * sdfo_debug.h -- test support entry points of libsdfo (no reference
 * counterpart).  They expose host-side intermediate structures so that CPU
 * tests can compare them with the CPU oracle without a GPU.
 */
#ifndef SDFO_DEBUG_H
#define SDFO_DEBUG_H

#include "sdfo.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host BVH of a mesh, built exactly as KdTree::build (KdTree.h:91-121,
 * 188-242).  nodes6: bmin3,bmax3 per node; nodeI4: left, right, triStart,
 * triCount; perm: BVH slot -> triangle.  Returns the node count (or a
 * negative error; -(needed) when max_nodes is too small).  No device use. */
int sdfo_debug_bvh_export(const double* verts, int64_t n_verts, const int32_t* tris,
                          int64_t n_tris, double* nodes6, int32_t* nodeI4, int32_t* perm,
                          int32_t max_nodes, int32_t* depth);

/* Order of n <= 64 values by absolute value, as the contact-aware builder
 * orders merged candidates (std::sort with |a| < |b|, SdfBuilder.cpp:348-353).
 * order[i] = index of the i-th value.  Host only. */
int sdfo_debug_sort_by_abs(const double* values, int32_t n, int32_t* order);

#ifdef __cplusplus
}
#endif

#endif /* SDFO_DEBUG_H */
