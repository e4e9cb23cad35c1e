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

/* The device BVH as the builds assemble it (host slot order, then nodes,
 * slot triangles and slot boxes built on the current device) for n_meshes
 * meshes concatenated: vertex offsets vo[n_meshes + 1], triangle offsets
 * to[n_meshes + 1], triangle indices local to each mesh's vertices.  Per
 * node: bmin3, bmax3 (nodes6) and left, right, triStart, triCount (nodeI4,
 * links and slot starts into the concatenation); per slot: the mesh-local
 * triangle (perm), its vertices a, b, c (tri9) and box bmin3, bmax3 (box6).
 * Returns the node count, or a negative error (-(needed) when max_nodes is
 * too small).  Synchronous on the current device. */
int sdfo_debug_device_bvh_export(int32_t n_meshes, const double* verts, const int64_t* vo,
                                 const int32_t* tris, const int64_t* to, double* nodes6,
                                 int32_t* nodeI4, int32_t* perm, double* tri9, double* box6,
                                 int32_t max_nodes, int32_t* depth);

/* Order of n <= 64 values by absolute value, as the contact-aware builder
 * orders merged candidates (std::sort with |a| < |b|, SdfBuilder.cpp:348-353).
 * order[i] = index of the i-th value.  Host only. */
int sdfo_debug_sort_by_abs(const double* values, int32_t n, int32_t* order);

/* Phase timings (ms) of the calling thread's last sdfo_build_dense (or
 * sdfo_build_dense_batched: set-ups, BVH pool, allocations beside it, fill,
 * -, Hessians + mirrors, total), in the
 * spirit of the builder's phase log (SdfBuilder.cpp:364-407): grid set-up,
 * host BVH, BVH upload + allocations, fill kernel, Hessian kernel, FP32
 * mirrors + cell records, total; then the number of voxels the fill re-ran
 * with the per-lane search.  Copies min(cap, 8) values; returns 8. */
int sdfo_debug_build_phases(double* ms, int32_t cap);

/* Test switch for the calling thread: when on, the dense / sparse fills send
 * every voxel (and the sparse activation every brick) to the per-lane replay
 * pass, so tests can check that pass against the CPU oracle. */
int sdfo_debug_force_replay(int32_t on);

/* Test switch for the calling thread: the step budget of the fills' wave
 * search (0: the library's choice) and the capacity of the batched fill's
 * list of bricks handed to the frontier pass (0: 2^18).  Small values send
 * many bricks through the frontier pass and, past the capacity, the replay
 * pass, so tests can check both against the CPU oracle. */
int sdfo_debug_fill_budget(int32_t budget, int32_t heavy_cap);

/* Test switch for the calling thread: how many merged candidates the
 * contact-aware builds' first pass keeps per voxel in registers (first: 1, 4
 * or 8; 0: the library's choice, 4), and how many per voxel the second pass,
 * whose lists live in device memory, has room for in its first launch
 * (second: 0 for the library's choice, 64) before a last launch with room for
 * the longest triangle list.  Small values send many voxels through the later
 * passes, so tests can check them. */
int sdfo_debug_ca_merge_cap(int32_t first, int32_t second);

/* Test switch for the calling thread: grids built or loaded while it is on
 * keep no cell records (the packed 8-corner copies that FP32 Trilinear /
 * Tricubic queries and the contact kernels read), as grids whose records
 * would exceed the library's limit (an eighth of device memory, at least
 * 4 GiB) do, so tests can check the
 * per-corner gather path against the CPU oracle on small grids. */
int sdfo_debug_skip_cell_records(int32_t on);

/* Test switch: records variant as the autotuner's decision for key (a key
 * as the tuner report prints it, e.g. "contact_forward points~2^12
 * poses~2^8"), so tests can run every launch variant of a kernel and check
 * that they give the same bits.  sdfo_tuner_reset clears it. */
int sdfo_debug_tuner_pin(const char* key, int32_t variant);

/* Measurement support: the best time in microseconds (best_us) over reps
 * launches (after one warm-up) of n_reads uniformly random reads of
 * record_bytes-byte records (64 or 128) from a device array of array_bytes,
 * two per lane in 128-lane blocks like the record-path query, each launch
 * with its own random records.  It allocates and frees the array; the
 * current device, synchronous. */
int sdfo_debug_random_read_floor(int32_t record_bytes, int64_t n_reads, int64_t array_bytes,
                                 int32_t reps, double* best_us);

#ifdef __cplusplus
}
#endif

#endif /* SDFO_DEBUG_H */
