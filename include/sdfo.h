/*
 * sdfo.h -- C ABI of the MI355X (gfx950) SdfOracle build / query / contact library.
 *
 * This is the drop-in boundary for the OpenMP path of the reference plugin
 * (plugins/SdfOracle) and for the per-point contact reduction of the Python
 * prototype.  The reference has no C ABI of its own; each entry point below
 * names the reference interface it replaces.  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every function returns SDFO_OK (0) or a negative SDFO_ERR_* code; the
 *    message text (the reference's exception text where one exists) is
 *    available from sdfo_last_error() on the calling thread.  No C++
 *    exception crosses this boundary.
 *  - Builds take HOST mesh arrays and produce a grid handle owned by the
 *    library (device resident, freed by sdfo_grid_free).
 *  - Queries and contact calls take DEVICE pointers (e.g. torch tensor
 *    data_ptr()) owned by the caller, and an explicit hipStream_t (NULL =
 *    the null stream).  They enqueue work and return without synchronising.
 *  - Dense fields are x-fastest, linear index i + nx*(j + ny*k)
 *    (GridSpec.h:56-59).  Hessians are 6 values Hxx,Hxy,Hxz,Hyy,Hyz,Hzz.
 */
#ifndef SDFO_H
#define SDFO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDFO_OK 0
#define SDFO_ERR_INVALID_ARGUMENT (-1)
#define SDFO_ERR_EMPTY_MESH (-2)
#define SDFO_ERR_IO (-3)
#define SDFO_ERR_FORMAT (-4)
#define SDFO_ERR_DEVICE (-5)
#define SDFO_ERR_OUT_OF_MEMORY (-6)
#define SDFO_ERR_UNSUPPORTED (-7)

/* QueryMode (SdfOracle.h:11-18), same order. */
#define SDFO_FIRST_ORDER 0
#define SDFO_SECOND_ORDER 1
#define SDFO_TRILINEAR 2
#define SDFO_TRICUBIC 3
#define SDFO_CONTACT_AWARE 4

typedef struct sdfo_grid sdfo_grid;
/* Broad-phase AABB tree over a body's quadrature points (geometry/bvh.py). */
typedef struct sdfo_point_bvh sdfo_point_bvh;
/* A triangle mesh's BVH on the device: the KdTree of KdTree.h. */
typedef struct sdfo_mesh sdfo_mesh;
typedef void* sdfo_stream; /* hipStream_t */

/* BuilderParams / SparseBuilderParams / ContactAwareBuilderParams
 * (SdfBuilder.h:24-55) flattened into one struct.  sdfo_default_build_params
 * fills the reference's defaults. */
typedef struct sdfo_build_params
{
    double padding;            /* fraction of the bbox diagonal (0.1) */
    double voxel_size;         /* 0 = longest side / target_resolution */
    int32_t target_resolution; /* 128 */
    int32_t max_resolution;    /* 512 */
    /* sparse narrow band */
    int32_t block_dim;    /* 8 */
    int32_t halo_voxels;  /* 2 */
    int32_t band_voxels;  /* 8 */
    int32_t store_hessian; /* 1 */
    double band_distance; /* 0 = band_voxels * voxel */
    /* contact aware */
    int32_t max_candidates;             /* 4 */
    int32_t tile_dim;                   /* 8 */
    double feature_distance_tolerance;  /* 0 = 1.5 voxel */
    double normal_merge_angle_degrees;  /* 5 */
    int32_t use_tile_acceleration;      /* 1 */
    int32_t reserved;
} sdfo_build_params;

/* Grid description (GridSpec.h:17-27 plus the sparse / contact-aware
 * header fields of SdfIO.cpp:62-80). */
typedef struct sdfo_grid_info
{
    double bmin[3], bmax[3], voxel[3];
    int32_t shape[3];
    int32_t sparse;          /* 0 dense, 1 sparse bricks */
    int32_t has_hessian;
    int32_t candidate_stride; /* 0 = no contact-aware candidates */
    int32_t block_dim, halo_voxels;
    int32_t block_count;      /* sparse: active bricks */
    int32_t block_grid[3];    /* sparse: ceil(n / block_dim) */
    double band_distance;
    double activation_distance;
    int64_t device_bytes;     /* device memory held by the handle (with the FP64 cell
                                 records its first FP64 Trilinear / Tricubic query made) */
} sdfo_grid_info;

/* Contact law parameters (contacts_sdf.py:300-305; ContactPair fields,
 * ir/contact.py:1-18). */
typedef struct sdfo_contact_params
{
    double kn;            /* normal stiffness */
    double damping;       /* c_n, quadrature_settings['damping'] */
    double mu;            /* friction coefficient */
    double regularizer;   /* quadrature_settings['regularizer'] (1e-6) */
    double narrow_margin; /* |Y.z| < margin filter (contacts_sdf.py:91-98); <= 0 disables */
    int32_t order;        /* k_order (1 or 2) */
    int32_t mode;         /* SDFO_TRILINEAR or SDFO_TRICUBIC */
} sdfo_contact_params;

/* ---- library ----------------------------------------------------------- */
const char* sdfo_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* sdfo_last_error(void);
/* Fills the reference defaults (SdfBuilder.h:24-55). */
void sdfo_default_build_params(sdfo_build_params* params);

/* SdfBuilder::setupGrid (SdfBuilder.cpp:994-1035) + snapToStandardResolution
 * (GridSpec.h:65-74); host only. */
int sdfo_setup_grid(const double* verts, int64_t n_verts, const sdfo_build_params* params,
                    sdfo_grid_info* out);

/* ---- builders ---------------------------------------------------------- */
/* SdfBuilder::build (SdfBuilder.h:58-62, SdfBuilder.cpp:364-407).
 * verts: n_verts x 3 doubles, tris: n_tris x 3 int32 (host). */
int sdfo_build_dense(const double* verts, int64_t n_verts, const int32_t* tris, int64_t n_tris,
                     const sdfo_build_params* params, sdfo_grid** out, sdfo_stream stream);

/* SdfBuilder::buildSparse (SdfBuilder.h:82-86, SdfBuilder.cpp:823-983). */
int sdfo_build_sparse(const double* verts, int64_t n_verts, const int32_t* tris, int64_t n_tris,
                      const sdfo_build_params* params, sdfo_grid** out, sdfo_stream stream);

/* SdfBuilder::buildContactAware / buildSparseContactAware (SdfBuilder.h:70-78,
 * SdfBuilder.cpp:416-821).  corner_normals: n_tris x 3 x 3 doubles. */
int sdfo_build_contact_aware(const double* verts, int64_t n_verts, const int32_t* tris,
                             int64_t n_tris, const double* corner_normals,
                             const sdfo_build_params* params, int32_t sparse, sdfo_grid** out,
                             sdfo_stream stream);

/* Batched dense build of many meshes in one persistent launch (no
 * reference counterpart: the reference builds one mesh per call).
 * Mesh m uses verts[vert_offsets[m]..vert_offsets[m+1]) and
 * tris[tri_offsets[m]..tri_offsets[m+1]) with mesh-local vertex indices;
 * params[m] is its builder configuration.  out[m] receives its grid. */
int sdfo_build_dense_batched(int32_t n_meshes, const double* verts, const int64_t* vert_offsets,
                             const int32_t* tris, const int64_t* tri_offsets,
                             const sdfo_build_params* params, sdfo_grid** out,
                             sdfo_stream stream);

/* ---- grid handle ------------------------------------------------------- */
/* Releases a grid.  The device is synchronised first, so work still in
 * flight on any stream finishes before the arrays return to the pool. */
void sdfo_grid_free(sdfo_grid* grid);
/* Stream-ordered release (no device-wide synchronisation): the arrays return
 * to the pool after the work already enqueued on `stream`.  The caller
 * guarantees that no other stream still uses the grid (work on other streams
 * is ordered before this call's point on `stream`, or complete).  The grids
 * of one batched build share one device block, freed with the last of them
 * after every such release point. */
void sdfo_grid_free_async(sdfo_grid* grid, sdfo_stream stream);
int sdfo_grid_get_info(const sdfo_grid* grid, sdfo_grid_info* info);

/* Copies the FP64 dense fields to host arrays (SdfData.h:27-118 members
 * phi0, n0, H0, witness0; any pointer may be NULL).  tri receives the
 * nearest-triangle index per voxel (-1 when unknown, e.g. after a load). */
int sdfo_grid_download_dense(const sdfo_grid* grid, double* phi, double* n3, double* h6,
                             double* witness3, int32_t* tri);

/* Copies the dense contact-aware candidate arrays (SdfData.h:46-50). */
int sdfo_grid_download_candidates(const sdfo_grid* grid, uint8_t* count, double* phi,
                                  double* n3, int32_t* branch);

/* Copies the sparse bricks (SparseSdfData, SdfData.h:154-303) in (k,j,i)
 * order: bijk[nb*3], phi[nb*B^3], n3[nb*B^3*3], h6[nb*B^3*6],
 * active_mask[bx*by*bz] (brick activation mask, bit exact with the
 * reference rule), and candidates when present.  Any pointer may be NULL. */
int sdfo_grid_download_sparse(const sdfo_grid* grid, int32_t* bijk, float* phi, float* n3,
                              float* h6, uint8_t* active_mask, uint8_t* ccount, float* cphi,
                              float* cn3, int32_t* cbranch);

/* saveSdf / saveSparseSdf (SdfIO.cpp:93-275): byte-identical "SDFO" 0.1.0. */
int sdfo_grid_save(const sdfo_grid* grid, const char* path);
/* SdfOracle::loadFromFile (SdfOracle.cpp:165-172) -> loadSdf / loadSparseSdf
 * (SdfIO.cpp:279-523); the reference's validation errors map to
 * SDFO_ERR_IO / SDFO_ERR_FORMAT with the same message text. */
int sdfo_grid_load(const char* path, sdfo_grid** out);
/* isValidSdf / isSparseSdf (SdfIO.cpp:527-552): 1 true, 0 false. */
int sdfo_is_valid_file(const char* path);
int sdfo_is_sparse_file(const char* path);

/* Creates a dense grid from host FP64 fields (SdfOracle(SdfData),
 * SdfOracle.cpp:154-157).  h6 may be NULL (zero Hessian); candidate arrays
 * may be NULL when stride == 0. */
int sdfo_grid_from_dense(const double* bmin3, const double* bmax3, const double* voxel3,
                         const int32_t* shape3, const double* phi, const double* n3,
                         const double* h6, const double* witness3, int32_t stride,
                         const uint8_t* ccount, const double* cphi, const double* cn3,
                         const int32_t* cbranch, sdfo_grid** out);

/* Uploads a host SparseSdfData (SdfData.h:154-303; blocks in their stored
 * order: bijk[3*nb], per block phi[B], n[3B], H[6B] when has_hessian, and the
 * candidate arrays when stride > 0; B = block_dim^3) as a grid, as
 * SdfOracle(SparseSdfData) does (SdfOracle.cpp:160-163). */
int sdfo_grid_from_sparse(const double* bmin3, const double* bmax3, const double* voxel3,
                          const int32_t* shape3, int32_t block_dim, int32_t halo, double band,
                          int32_t has_hessian, int32_t stride, int32_t n_blocks,
                          const int32_t* bijk, const float* phi, const float* n3, const float* h6,
                          const uint8_t* ccount, const float* cphi, const float* cn3,
                          const int32_t* cbranch, sdfo_grid** out);

/* saveSdf / saveSparseSdf (SdfIO.cpp:93-275) of host arrays (the layouts of
 * sdfo_grid_from_dense / sdfo_grid_from_sparse; h6 NULL writes zeros). */
int sdfo_save_dense_host(const char* path, const double* bmin3, const double* bmax3,
                         const double* voxel3, const int32_t* shape3, const double* phi,
                         const double* n3, const double* h6, const double* witness3, int32_t stride,
                         const uint8_t* ccount, const double* cphi, const double* cn3,
                         const int32_t* cbranch);
int sdfo_save_sparse_host(const char* path, const double* bmin3, const double* bmax3,
                          const double* voxel3, const int32_t* shape3, int32_t block_dim,
                          int32_t halo, double band, int32_t has_hessian, int32_t stride,
                          int32_t n_blocks, const int32_t* bijk, const float* phi, const float* n3,
                          const float* h6, const uint8_t* ccount, const float* cphi,
                          const float* cn3, const int32_t* cbranch);

/* ---- queries ----------------------------------------------------------- */
/* SdfOracle::queryBatch / queryFast (SdfOracle.cpp:176-444) for FP32 point
 * streams, FP32 arithmetic with FP64 cell location.
 * points: n x 3 floats (device).  gap[n], grad3[n*3] (raw gradient, the
 * Python contract of sdf_grid.py:138-204), valid[n] (device, required);
 * normal3 / witness3 / branch (device, optional, NULL to skip). */
int sdfo_query(const sdfo_grid* grid, int32_t mode, const float* points, int64_t n, float* gap,
               float* grad3, uint8_t* valid, float* normal3, float* witness3, int32_t* branch,
               sdfo_stream stream);

/* Same for FP64 point streams with FP64 outputs, in FP64 arithmetic without
 * FMA contraction, on a dense grid's FP64 fields (phi, n, H, candidates) or a
 * sparse grid's FP32 fields widened (sparse data is FP32, as in the
 * reference): every output equals the CPU path's (SdfOracle.cpp:176-1092) bit
 * for bit.  Used by the source-compatible C++ SdfOracle mirror. */
int sdfo_query_f64(const sdfo_grid* grid, int32_t mode, const double* points, int64_t n,
                   double* gap, double* grad3, uint8_t* valid, double* normal3,
                   double* witness3, int32_t* branch, sdfo_stream stream);

/* SdfOracle::queryFast / queryBatch (SdfOracle.cpp:176-444) on host arrays
 * of FP64 points, for per-point callers: copies in, runs the FP64 query
 * kernels, copies out, synchronises the stream.  gap[n], normal3[3n],
 * witness3[3n], valid[n] are required; branch[n] optional.  With
 * cand_count != NULL and mode SDFO_CONTACT_AWARE on a grid with candidates,
 * every candidate is returned too (cand_gap[8n], cand_normal3[24n],
 * cand_witness3[24n], cand_branch[8n]; cand_count[i] = 0 where the point took
 * the tricubic / first-order fallback), computed from the FP64 (dense) or
 * FP32 (sparse) candidate arrays without FMA contraction. */
int sdfo_query_host_f64(const sdfo_grid* grid, int32_t mode, const double* points, int64_t n,
                        double* gap, double* normal3, double* witness3, uint8_t* valid,
                        int32_t* branch, uint8_t* cand_count, double* cand_gap,
                        double* cand_normal3, double* cand_witness3, int32_t* cand_branch,
                        sdfo_stream stream);

/* Many query streams against many grids in one launch (no reference
 * counterpart).  Stream s queries grids[s] with points
 * [offsets[s], offsets[s+1]) of the packed arrays; all pointers device
 * except grids/offsets (host).  Convenience form: builds a plan, runs it and
 * synchronises the stream. */
int sdfo_query_batched(int32_t n_streams, const sdfo_grid* const* grids, int32_t mode,
                       const int64_t* offsets, const float* points, float* gap, float* grad3,
                       uint8_t* valid, sdfo_stream stream);

/* A reusable batched query (no reference counterpart): the per-stream grid
 * and point-range table is uploaded once by sdfo_query_plan_create (host
 * arrays grids[n_streams], offsets[n_streams + 1]); sdfo_query_plan_run is a
 * single kernel launch that allocates, copies and synchronises nothing, so it
 * can be captured in a HIP graph (tune it, if the tuner is on, by one run
 * outside the capture).  The grids must outlive the plan. */
typedef struct sdfo_query_plan sdfo_query_plan;
int sdfo_query_plan_create(int32_t n_streams, const sdfo_grid* const* grids,
                           const int64_t* offsets, sdfo_query_plan** out);
int sdfo_query_plan_run(const sdfo_query_plan* plan, int32_t mode, const float* points,
                        float* gap, float* grad3, uint8_t* valid, sdfo_stream stream);
void sdfo_query_plan_free(sdfo_query_plan* plan);

/* ---- autotuner (no reference counterpart) ------------------------------ */
/* SDFO_TUNE_ON_FIRST_USE: the first dense build of a size class (triangle
 * and voxel counts, powers of two) times every voxel-brick shape x step
 * budget of the fill and keeps the fastest; the first run of a batched query
 * plan of a size class, and the first FP32 Trilinear / Tricubic query of a
 * size class on a grid with cell records, time every block size.  A tuning
 * call synchronises its stream.  Results never depend on the choice. */
#define SDFO_TUNE_OFF 0
#define SDFO_TUNE_ON_FIRST_USE 1
int sdfo_tuner_set_mode(int32_t mode);
/* Clears every decision. */
int sdfo_tuner_reset(void);
/* One line per decision: "key: variant=ms, ..." with the choice starred.
 * Copies at most cap-1 bytes plus a NUL; returns the full length. */
int64_t sdfo_tuner_report(char* buf, int64_t cap);

/* ---- contact ----------------------------------------------------------- */
/* Per-pose penalty + damping + regularised Coulomb friction reduction of
 * SDFContactEngine.evaluate (contacts_sdf.py:329-385), evaluated in the SDF
 * (body B) frame, FP64 arithmetic, deterministic FP64 sums.
 * X: n x 3 quadrature points in A's frame, w: n weights (device, FP64).
 * poses: B x 30 doubles per pose: R(9 row major), t(3) [Y = R X + t];
 *        M(9), m(3) [offset from A's centre of mass, off = M X + m];
 *        v0(3), Om(3) [relative velocity u = v0 + Om x off].
 * out: B x 6 (F_A, tau_A) (device).  active: B int32 (device, optional). */
int sdfo_contact_forward(const sdfo_grid* grid, const double* X, const double* w, int64_t n,
                         const double* poses, int64_t n_poses, const sdfo_contact_params* params,
                         double* out, int32_t* active, sdfo_stream stream);

/* SDFContactEngine.evaluate (contacts_sdf.py:225-385) for many contact pairs,
 * in the world frame, FP64 arithmetic, deterministic FP64 sums.
 * X: n x 3 quadrature points in body A's (action marker) frame, w: n weights
 * (device, FP64; the engine normalises them to sum 1 at registration).
 * states: n_pairs x 36 doubles (device), the body states of each pair
 *   (contacts_sdf.py:231-240): rA(3) RA(9, row major) vA(3) wA(3)
 *   rB(3) RB(9) vB(3) wB(3); A carries the points, B the SDF.
 * frames: n_pairs x 24 doubles (device) or NULL: the action marker relative
 *   to the base marker, R_ab(9) t_ab(3) (Y = R_ab X + t_ab), and the action
 *   marker's world pose RAm(9) rAm(3) (x_w = rAm + RAm X)
 *   (contacts_sdf.py:243-249, 275-279); NULL uses the body frames
 *   (R_sdf = RB^T RA, t_sdf = RB^T (rA - rB), :280-283).
 * The narrow filter |Y_z| < params->narrow_margin always applies (the
 * engine's default margin is 0.01; a margin <= 0 rejects every point, as in
 * the reference).  bvh (built over the same X, or NULL for no cull): the
 * points of the leaves overlapping the SDF box (grid bmin/bmax mapped into A's
 * frame, grown by max(narrow_margin, 0.05)) take part, or every point when
 * no leaf overlaps (:243-271).  body_boxes9 (host, or NULL to skip): A's local
 * AABB centre(3) and half extents(3), then B's half extents(3): pairs failing
 * _check_broad_phase (:68-89) return zeros.
 * out: n_pairs x 12 (device) = F_A, tau_A, F_B, tau_B in the world frame,
 * tau about each body's reference point; active: n_pairs int32 (optional),
 * the number of points in the sums. */
int sdfo_contact_evaluate(const sdfo_grid* grid, const sdfo_point_bvh* bvh, const double* X,
                          const double* w, int64_t n, const double* states, const double* frames,
                          int64_t n_pairs, const sdfo_contact_params* params,
                          const double* body_boxes9, double* out, int32_t* active,
                          sdfo_stream stream);

/* sdfo_contact_evaluate without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The pair
 * records, cull boxes and partial sums live in a static device array of
 * about 2.3 MB per GPU, held until the process ends; at most 4096 pairs per
 * call (more is rejected with SDFO_ERR_INVALID_ARGUMENT).  Every call on a GPU
 * uses that GPU's one array, so calls that overlap on one GPU (other streams,
 * other threads) race: serialise them.  A point BVH must already be resident
 * on the grid's device (one sdfo_contact_evaluate or sdfo_point_bvh_cull call
 * with it uploads it).  Arguments and results are sdfo_contact_evaluate's. */
int sdfo_contact_evaluate_noalloc(const sdfo_grid* grid, const sdfo_point_bvh* bvh,
                                  const double* X, const double* w, int64_t n,
                                  const double* states, const double* frames, int64_t n_pairs,
                                  const sdfo_contact_params* params, const double* body_boxes9,
                                  double* out, int32_t* active, sdfo_stream stream);

/* Linearisation cache of SDFContactEngine.evaluate (contacts_sdf.py:307-327),
 * used by the Newton iterations of implicit_be.py:474-500.
 * sdfo_contact_linearize fills it at one reference pose (a device record of
 * 30 doubles, only R and t read): g0[n], grad0[3n] (raw gradient), valid0[n]
 * (0 where the narrow filter or the field rejects the point), with the same
 * transform and interpolation as sdfo_contact_forward. */
int sdfo_contact_linearize(const sdfo_grid* grid, const double* X, int64_t n, const double* pose,
                           const sdfo_contact_params* params, double* g0, double* grad0,
                           uint8_t* valid0, sdfo_stream stream);

/* Linearised contact forward: the SDF is not queried; point i's gap under a
 * pose is g0[i] + grad0[i] . (t - t_ref) with gradient grad0[i], and it takes
 * part only where valid0[i] and the narrow filter at the current pose pass;
 * t_ref3 (host, 3 doubles) is the reference pose's t.  With bvh != NULL the
 * broad-phase cull of sdfo_contact_forward_culled applies too.  The rest of
 * the law and the output are sdfo_contact_forward's. */
int sdfo_contact_forward_linearized(const double* g0, const double* grad0, const uint8_t* valid0,
                                    const double* t_ref3, const double* X, const double* w, int64_t n,
                                    const double* poses, int64_t n_poses,
                                    const sdfo_contact_params* params, const sdfo_point_bvh* bvh,
                                    const double* boxes6, double* out, int32_t* active,
                                    sdfo_stream stream);

/* ---- triangle-mesh distance queries ----------------------------------------- */
/* KdTree::build (KdTree.h:91-121, 188-242) over a host mesh; the tree is
 * built on the host exactly as the builders build it and kept on the current
 * device.  An empty mesh is allowed (every query then reports distance
 * DBL_MAX and triangle -1, KdTree.h:123-133). */
int sdfo_mesh_build(const double* verts, int64_t n_verts, const int32_t* tris, int64_t n_tris,
                    sdfo_mesh** out);
void sdfo_mesh_free(sdfo_mesh* mesh);
/* KdTree::triangleCount (KdTree.h:167-170). */
int sdfo_mesh_triangle_count(const sdfo_mesh* mesh, int64_t* count);

/* KdTree::findNearest (signed_distance = 0, KdTree.h:123-133) or
 * findSignedNearest (signed_distance != 0, ray-parity sign, KdTree.h:143-165)
 * for n FP64 device points (n x 3): distance[n] (signed when asked),
 * closest3[3n], normal3[3n] (the nearest triangle's face normal, (0, 0, 1)
 * when degenerate) and triangle[n] (device; any may be NULL).  Bit-identical
 * to the reference's search for each point. */
int sdfo_mesh_query(const sdfo_mesh* mesh, const double* points, int64_t n,
                    int32_t signed_distance, double* distance, double* closest3, double* normal3,
                    int32_t* triangle, sdfo_stream stream);

/* The same on host arrays (copies in and out, synchronises the stream); for
 * per-point callers such as the C++ KdTree mirror. */
int sdfo_mesh_query_host(const sdfo_mesh* mesh, const double* points, int64_t n,
                         int32_t signed_distance, double* distance, double* closest3,
                         double* normal3, int32_t* triangle, sdfo_stream stream);

/* ---- contact broad phase --------------------------------------------------- */
/* AABB_BVH(points, leaf_size) (geometry/bvh.py:19-54) over a body's host FP64
 * quadrature points (n x 3): longest-axis midpoint splits, a median split
 * where one side would be empty, leaves of at most leaf_size points (the
 * engine uses 16, contacts_sdf.py:63).  Host only: the leaf tables go to the
 * device of the first call that culls with the tree, and later culling calls
 * must run on that device. */
int sdfo_point_bvh_build(const double* points, int64_t n, int32_t leaf_size, sdfo_point_bvh** out);
void sdfo_point_bvh_free(sdfo_point_bvh* bvh);
int sdfo_point_bvh_get_info(const sdfo_point_bvh* bvh, int64_t* n_points, int64_t* n_nodes,
                            int64_t* n_leaves);

/* AABB_BVH.query (bvh.py:56-75) on the host: indices (capacity n_points) of
 * the points in every leaf whose box, grown by margin, overlaps
 * [qmin3, qmax3], in the reference's order; *count receives their number. */
int sdfo_point_bvh_query(const sdfo_point_bvh* bvh, const double* qmin3, const double* qmax3,
                         double margin, int32_t* indices, int64_t* count);

/* The same test for many boxes on the device: boxes6 = n_boxes x (qmin3,
 * qmax3), mask[b * n_points + i] = 1 where point i is in a reported leaf. */
int sdfo_point_bvh_cull(const sdfo_point_bvh* bvh, const double* boxes6, int64_t n_boxes,
                        double margin, uint8_t* mask, sdfo_stream stream);

/* sdfo_contact_forward behind the broad phase of SDFContactEngine.evaluate
 * (contacts_sdf.py:245-272): for pose b only the points of the leaves of
 * bvh (built over the same X) that overlap boxes6[6b..6b+5] (device; the
 * engine's bmin_local, bmax_local, margin included) take part, or every
 * point when no leaf overlaps (contacts_sdf.py:270-271). */
int sdfo_contact_forward_culled(const sdfo_grid* grid, const sdfo_point_bvh* bvh,
                                const double* boxes6, const double* X, const double* w, int64_t n,
                                const double* poses, int64_t n_poses,
                                const sdfo_contact_params* params, double* out, int32_t* active,
                                sdfo_stream stream);

/* Hand-written backward of sdfo_contact_forward (no reference counterpart).
 * g_out: B x 6 upstream gradient.  Outputs (device, each optional):
 * g_twist: B x 6 gradient w.r.t. a left twist (omega, dt) of A's pose,
 * g_params: 3 (kn, damping, mu) summed over poses, g_X: n x 3, g_w: n. */
int sdfo_contact_backward(const sdfo_grid* grid, const double* X, const double* w, int64_t n,
                          const double* poses, int64_t n_poses,
                          const sdfo_contact_params* params, const double* g_out,
                          double* g_twist, double* g_params, double* g_X, double* g_w,
                          sdfo_stream stream);

/* A forward that keeps, for its backward, which (pose, point) pairs are in
 * contact (no reference counterpart; what a training step's forward saves).
 * sdfo_contact_saved_bytes: the size of that record for n points and
 * n_poses poses (-1 for negative sizes).  sdfo_contact_forward_save is
 * sdfo_contact_forward (same outputs, bit for bit) that also writes the
 * record into saved (device, saved_bytes = sdfo_contact_saved_bytes(n,
 * n_poses)).  sdfo_contact_backward_saved is sdfo_contact_backward reading
 * the record instead of screening every pair again.  The record also holds
 * every buffer of the two calls (about 37 bytes per (pose, point), the point
 * gradients' partials for up to 64 MiB of poses at a time), so neither call
 * allocates device memory; it must get the same
 * grid, X, w, poses and params as the forward that wrote the record (the
 * backward re-evaluates the recorded pairs exactly, but pairs in contact
 * under other inputs and absent from the record would be missed).  Its
 * results equal sdfo_contact_backward's up to the order of the sums. */
int64_t sdfo_contact_saved_bytes(int64_t n, int64_t n_poses);
int sdfo_contact_forward_save(const sdfo_grid* grid, const double* X, const double* w, int64_t n,
                              const double* poses, int64_t n_poses,
                              const sdfo_contact_params* params, double* out, int32_t* active,
                              void* saved, int64_t saved_bytes, sdfo_stream stream);
int sdfo_contact_backward_saved(const sdfo_grid* grid, const double* X, const double* w, int64_t n,
                                const double* poses, int64_t n_poses,
                                const sdfo_contact_params* params, const double* g_out,
                                const void* saved, int64_t saved_bytes, double* g_twist,
                                double* g_params, double* g_X, double* g_w, sdfo_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* SDFO_H */
