/*
 * casl_reinit.h -- C ABI of the MI355X level-set reinitialization library.
 *
 * Plain pointers and sizes only.  Every call returns an int status (CASL_OK
 * or a negative CASL_ERR_* code); no C++ exception crosses this boundary.
 * Device pointers are HIP device allocations on the current device; `stream`
 * is a hipStream_t passed as void* (NULL = the null stream).
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   - my_p4est_level_set_t::reinitialize_{1st_order, 2nd_order,
 *     1st_order_time_2nd_order_space, 2nd_order_time_1st_order_space}
 *     and their *_above_threshold forms        src/my_p4est_level_set.h:434-613
 *     all of which forward to reinitialize_within_range_of_phi_0
 *                                              src/my_p4est_level_set.cpp:290-557
 *   - my_p4est_level_set_t::reinitialize_2nd_order_with_mask
 *                                              src/my_p4est_level_set.h:487-530
 *   - my_p4est_node_neighbors_t::second_derivatives_central(Vec, Vec[DIM])
 *                                              src/my_p4est_node_neighbors.h:561-565,
 *                                              src/my_p4est_node_neighbors.cpp:1422-1641
 *   - my_p4est_node_neighbors_t::init_neighbors (stencil storage)
 *                                              src/my_p4est_node_neighbors.cpp:33-61
 *   - host-side grid construction that the reference takes from p4est:
 *     my_p4est_new + my_p4est_refine(refine_levelset_cf)
 *                                              src/my_p4est_refine_coarsen.cpp:17-89,
 *                                              examples/reinitialization/main_2d.cpp:165-173
 *     my_p4est_nodes_new                       src/my_p4est_nodes.c:383-446, 614-671
 *     construct_neighbors                      src/my_p4est_node_neighbors.cpp:114-1201
 *     sample_cf_on_nodes                       src/my_p4est_utils.cpp:2129-2170
 */
#ifndef CASL_REINIT_H
#define CASL_REINIT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define CASL_OK                     0
#define CASL_ERR_INVALID_ARGUMENT  -1 /* reference: std::invalid_argument   */
#define CASL_ERR_NO_DEVICE         -2 /* no HIP device / library not usable */
#define CASL_ERR_HIP               -3 /* a HIP runtime call failed          */
#define CASL_ERR_OUT_OF_MEMORY     -4
#define CASL_ERR_INEXACT_DISTANCES -5 /* distances not exact in FP32 where
                                         FP32 is required (batches, grids
                                         built on the device)            */
#define CASL_ERR_INVALID_STENCIL   -6 /* index out of range, bad wall arm   */
#define CASL_ERR_NOT_BUILT         -7 /* forest step called out of order    */

/* Last error message of the calling thread (static storage, never NULL). */
const char *casl_last_error(void);

/* ------------------------------------------------------- precision modes */
#define CASL_PRECISION_FP64  0 /* every operation in FP64                 */
#define CASL_PRECISION_MIXED 1 /* FP32 bulk Godunov/ENO updates and second
                                  derivatives; FP64 subcell root finding
                                  and FP64 convergence norm               */

/* ---------------------------------------------------- flattened stencils
 * One record per node, structure-of-arrays, field for field the per-node
 * struct quad_neighbor_nodes_of_node_t (my_p4est_quad_neighbor_nodes_of_node.h:1191-1225).
 * Faces are ordered m00, p00, 0m0, 0p0[, 00m, 00p]; the corners of a face
 * are ordered mm, pm[, mp, pp] over its two transverse axes (lower axis
 * first); the transverse distances of a face are ordered m0, p0[, 0m, 0p].
 * With F = 2*dim, C = 2^(dim-1), T = 2*(dim-1), N = n_nodes:
 *   nbr [(f*C + c)*N + n]  = node_{f}_{c}      (int32 node index)
 *   d   [f*N + n]          = d_{f}             (negative on a wall arm,
 *                                               construct_neighbors :609-719)
 *   d_tr[(f*T + t)*N + n]  = d_{f}_{t}
 *   wall[n]                bit f set <=> is_node_{x,y,z}{m,p}Wall(node)
 *                                       (my_p4est_utils.cpp:1815-1905)
 */
typedef struct casl_flat_stencil {
  int32_t        dim;     /* 2 or 3 */
  int64_t        n_nodes;
  const int32_t *nbr;
  const double  *d;
  const double  *d_tr;
  const uint8_t *wall;
  double         zero_distance_threshold; /* EPS*min(tree dimensions),
                                             my_p4est_level_set.h:388 */
} casl_flat_stencil;

/* ------------------------------------------------- device stencil handle */
typedef struct casl_stencil casl_stencil_t;

/* Validates the host tables and uploads them to the current device
 * (replaces init_neighbors' cached qnnn vector).  Distances are kept in FP32
 * (exact for trees of dyadic size); when FP32 does not hold them exactly (a
 * box such as [0, 2 pi]) an FP64 copy is kept too, and FP64 calls run the
 * grid-wide kernels on it (mixed calls use the rounded FP32 copy). */
int casl_stencil_create(const casl_flat_stencil *tables, casl_stencil_t **out);
int casl_stencil_destroy(casl_stencil_t *st);
int casl_stencil_info(const casl_stencil_t *st, int32_t *dim, int64_t *n_nodes);

/* Stage kernel selection.  CASL_PATH_TILED: one fused kernel per
 * pseudo-time stage over Morton tiles with LDS-staged halos (the default when
 * the stencil could be tiled into at least one tile per compute unit);
 * CASL_PATH_TWO_PASS: grid-wide pass A + pass B kernels (the default for
 * smaller grids).  A stencil without a tiled encoding (cells that are not
 * cubes, or tile limits exceeded; casl_stencil_tile_info gives the reason)
 * runs CASL_PATH_TWO_PASS whatever was set; get_path returns the path that
 * runs. */
#define CASL_PATH_TILED    0
#define CASL_PATH_TWO_PASS 1
int casl_stencil_set_path(casl_stencil_t *st, int path);
int casl_stencil_get_path(const casl_stencil_t *st, int32_t *path);
/* tiled: 1 if the tiled encoding exists; counts of tiles, halo entries
 * (H1 + H2 over all tiles), records (T + H1 over all tiles) and records of
 * nodes with a hanging arm. */
int casl_stencil_tile_info(const casl_stencil_t *st, int32_t *tiled,
                           int64_t *n_tiles, int64_t *n_halo,
                           int64_t *n_records, int64_t *n_irregular);
/* No counterpart in the reference: the tiled encoding of host tables built
 * and checked on the host alone (no device): counts of tiles, records (T + H1
 * over all tiles), records with hanging faces, tile nodes with hanging faces
 * and records with two hanging faces (3D).  Tables that cannot be tiled
 * return CASL_ERR_INVALID_STENCIL with the reason in casl_last_error; a
 * stencil made from them runs the two-pass path. */
int casl_tile_plan(const casl_flat_stencil *tables, int64_t *n_tiles,
                   int64_t *n_records, int64_t *n_irregular,
                   int64_t *n_irregular_nodes, int64_t *n_two_hanging);

/* No counterpart in the reference: a check that the stencil's tiled encoding
 * (built on the device for casl_device_grid_stencil, on the host for
 * casl_stencil_create) equals, byte for byte, the host builder's encoding of
 * the given host tables of the same grid.  *equal = 1 when every array is
 * the same (or neither could be tiled), else 0 with the first difference in
 * casl_last_error.  Copies the encoding to the host: a test, not a step of a
 * solve. */
int casl_stencil_tile_check(const casl_stencil_t *st, const casl_flat_stencil *tables,
                            int32_t *equal);

/* second_derivatives_central: fdd[a*N + n] = d^2 f / dx_a^2 at node n.
 * f and fdd are FP64 device arrays. */
int casl_second_derivatives(casl_stencil_t *st, const double *f, double *fdd,
                            void *stream);

/* first_derivatives_central: fd[a*N + n] = d f / dx_a at node n -- the central
 * derivative on linearly interpolated face values, corrected by the second
 * derivatives where a hanging arm needs it (src/my_p4est_node_neighbors.h:568-600,
 * src/my_p4est_quad_neighbor_nodes_of_node.cpp:638-760).  FP64 device arrays. */
int casl_first_derivatives(casl_stencil_t *st, const double *f, double *fd,
                           void *stream);
/* compute_normals (src/my_p4est_utils.cpp:1720-1740): normals[a*N + n] =
 * first derivatives / their norm, 0 where the norm is below EPS. */
int casl_normals(casl_stencil_t *st, const double *phi, double *normals,
                 void *stream);
/* compute_mean_curvature(neighbors, normals, kappa) (src/my_p4est_utils.cpp:
 * 1648-1690; qnnn::get_curvature(normals), src/my_p4est_quad_neighbor_nodes_of_node.h:
 * 2167-2171): kappa[n] = sum over a of the central derivative of normals[a]
 * along axis a. */
int casl_mean_curvature(casl_stencil_t *st, const double *normals, double *kappa,
                        void *stream);

/* advect_in_normal_direction(phi, dt, vn[, vn_np1]) (src/my_p4est_level_set.cpp:
 * 559-726, 785-815): one TVD-RK2 step of phi_t + vn |grad phi| = 0 with the
 * Godunov / ENO-minmod stencil of the reinitialization, the second stage with
 * vn_np1 when given (may be NULL).  phi, vn, vn_np1: FP64 device arrays. */
int casl_advect_in_normal_direction(casl_stencil_t *st, double *phi,
                                    const double *vn, const double *vn_np1,
                                    double dt, void *stream);
/* The time step of advect_in_normal_direction(const CF&, phi)
 * (src/my_p4est_level_set.cpp:744-783): min over nodes of 0.8 |s_min / vn|,
 * s_min the node's shortest arm.  *dt is host. */
int casl_advect_cfl_dt(casl_stencil_t *st, const double *vn, double *dt,
                       void *stream);

/* reinitialize_within_range_of_phi_0(phi, order_space, order_pseudotime,
 * phi_0_limit_high, phi_0_limit_low, number_of_iterations): phi is an FP64
 * device array of n_nodes values, updated in place.  Orders outside {1,2}
 * -> CASL_ERR_INVALID_ARGUMENT (reference throws std::invalid_argument).
 * Pass +/-DBL_MAX limits for the plain reinitialize_* entry points. */
int casl_reinit(casl_stencil_t *st, double *phi, int order_space,
                int order_pseudotime, int n_iter, double limit_high,
                double limit_low, int precision, void *stream);

/* casl_reinit (mask == NULL) or casl_reinit_masked (orders 2/2 only) with
 * phi and mask in host memory -- the arrays of the caller's PETSc Vecs
 * (src/my_p4est_level_set.cpp:318-333) -- phi updated in place; returns when
 * phi holds the result.  phi travels through a pinned staging buffer kept
 * with the stencil (allocated on the first call, like the device copy of
 * phi), in chunks whose host copies overlap the DMA transfers.  Same results
 * as casl_reinit / casl_reinit_masked on device copies. */
int casl_reinit_host(casl_stencil_t *st, double *phi, const uint8_t *mask,
                     int order_space, int order_pseudotime, int n_iter,
                     double limit_high, double limit_low, int precision,
                     void *stream);

/* casl_reinit recorded into a hipGraph on its first call and replayed on
 * later calls with the same arguments (phi address included) and the same
 * stage path: one graph launch instead of 2 + 2 * n_iter kernel launches,
 * which matters for small grids.  Results equal casl_reinit's.  A call with
 * other arguments records again.  No counterpart in the reference (the
 * reinitialize_* methods are host loops over PETSc calls). */
int casl_reinit_graph(casl_stencil_t *st, double *phi, int order_space,
                      int order_pseudotime, int n_iter, double limit_high,
                      double limit_low, int precision, void *stream);

/* reinitialize_2nd_order_with_mask: nodes with mask[n] == 0 keep their
 * current value at every stage (orders 2/2).  mask is a device uint8 array. */
int casl_reinit_masked(casl_stencil_t *st, double *phi, const uint8_t *mask,
                       int n_iter, double limit_high, double limit_low,
                       int precision, void *stream);

/* No counterpart in the reference: casl_reinit for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments (no mask) and the same
 * results bit for bit; the call's buffers (the phi_0 copy, flags, subcell
 * distances, stage buffers) come from one fixed device array of
 * CASL_NOALLOC_SCRATCH_BYTES per GPU, which the library holds until the
 * process ends, instead of the stencil's own scratch.  Calls that need more
 * (66 B per node in 3D FP64 on the tiled path, 114 B on the two-pass
 * path) return CASL_ERR_INVALID_ARGUMENT before anything runs.  The stencil
 * itself is made beforehand (casl_stencil_create allocates).  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams or
 * host threads) race.  The first call on a device looks the array up there. */
#define CASL_NOALLOC_SCRATCH_BYTES (64u << 20)
int casl_reinit_noalloc(casl_stencil_t *st, double *phi, int order_space,
                        int order_pseudotime, int n_iter, double limit_high,
                        double limit_low, int precision, void *stream);

/* L-infinity norm of (a - b) over nodes with |ref| < band (FP64 reduction on
 * the device; used for the mixed-precision error report).  Result on host. */
int casl_band_linf(casl_stencil_t *st, const double *a, const double *b,
                   const double *ref, double band, double *result, void *stream);

/* --------------------------------------------------------- kernel timing
 * When enabled, every kernel launch of the library is bracketed by HIP
 * events recorded on its launch stream.  casl_profile_query returns the
 * summed event time (ms) and the launch count of one kernel ("k_dd",
 * "k_prep", "k_tile_prep", "k_stage", "k_tile_stage", "k_batch", "k_gemm") since the last reset; it synchronizes the
 * recorded events.  Off by default (no events, no overhead). */
int casl_profile_enable(int enable);
int casl_profile_reset(void);
int casl_profile_query(const char *kernel, double *total_ms, int64_t *launches);

/* ------------------------------------------------- device block cache
 * No counterpart in the reference (its grids live in host memory).  Device
 * memory the library releases (a destroyed grid's or stencil's tables, the
 * temporaries of a device grid build or tiling) goes back to the driver at
 * once by default.  With a limit above 0, released blocks up to that many
 * bytes per device are kept and reused by later allocations of the library
 * of the same size (at most 25% larger): a regrid loop that builds a grid on
 * the device, drops the previous one and builds again then stops waiting on
 * the driver's hipMalloc of recently freed memory (seconds at hundreds of
 * GB).  A block enters the cache after a device-wide synchronization, as
 * hipFree would wait.  Cached memory is unavailable to other allocators
 * (torch's included) until casl_device_cache_trim frees it; an allocation of
 * the library that fails trims its device's cache and tries once more.
 * Lowering the limit below what a device holds trims that device. */
int casl_device_cache_set_limit(size_t bytes_per_device);
int casl_device_cache_trim(void);
int casl_device_cache_held(int device, size_t *bytes);

/* ------------------------------------------------- batched small trees
 * Many independent trees reinitialized by one persistent kernel.  The trees'
 * stencils are given as one concatenated casl_flat_stencil per tree; phi is
 * the concatenation of the trees' node values in the same order. */
typedef struct casl_batch casl_batch_t;
int casl_batch_create(const casl_flat_stencil *trees, int32_t n_trees,
                      casl_batch_t **out);
int casl_batch_destroy(casl_batch_t *b);
int casl_batch_info(const casl_batch_t *b, int64_t *total_nodes,
                    int32_t *n_trees, int32_t *max_tree_nodes);
/* Launch schedule.  nodes_per_thread 1 or 2: the kernel that gathers through
 * the grid-wide tables, block_threads in {64, 128, 256, 512};
 * nodes_per_thread 0: the kernel that runs each tree's stages tile by tile in
 * LDS (compact records), block_threads 512 (or 256 or 1024 in 2D).
 * block_threads 0 picks the default for the kind. */
int casl_batch_set_schedule(casl_batch_t *b, int32_t block_threads,
                            int32_t nodes_per_thread);
/* tiled = 1 when every tree has a tiled encoding (nodes_per_thread 0 usable);
 * n_tiles = tiles over all trees. */
int casl_batch_tile_info(const casl_batch_t *b, int32_t *tiled,
                         int64_t *n_tiles);
int casl_batch_get_schedule(const casl_batch_t *b, int32_t *block_threads,
                            int32_t *nodes_per_thread);
/* Times every supported launch shape (both kernels) on a copy of phi (phi is
 * not modified), keeps the fastest as the batch's schedule and reports it. */
int casl_batch_autotune(casl_batch_t *b, const double *phi, int n_iter,
                        int precision, void *stream, int32_t *best_block,
                        int32_t *best_nodes_per_thread, float *best_ms);
/* reinitialize_2nd_order (orders 2/2, limits +/-DBL_MAX) of every tree. */
int casl_batch_reinit(casl_batch_t *b, double *phi, int n_iter,
                      int precision, void *stream);
/* Same, recorded once into a hipGraph and replayed (no per-call launches
 * or allocations after the first call with a given n_iter/precision). */
int casl_batch_reinit_graph(casl_batch_t *b, double *phi, int n_iter,
                            int precision, void *stream);

/* ------------------------------------------- learned curvature (h*kappa)
 * The reference infers h*kappa with a CPU MLP (OpenBLAS sgemm) whose
 * training happens outside the repository (src/my_p4est_curvature_ml.cpp:
 * 214-286, 483-556).  Here the network is trained and evaluated on the GPU
 * by BF16-MFMA kernels with FP32 accumulation.  All arrays below are device
 * arrays unless marked host. */

/* Numerical curvature samples on n_patch uniform (M+1)x(M+1) patches of
 * spacing h with origin (x0, y0).  grid[i*(M+1)+j] is the patch-local node
 * index of (x0 + i h, y0 + j h); phi holds the patches' node values one after
 * another (n_nodes each).  For every node with an opposite-sign face
 * neighbour and 4 <= i, j <= M-4 one sample is appended (order unspecified):
 * feat[s][0..8] = 3x3 phi stencil / h (x slowest), feat[s][9] = h*kappa
 * bilinearly interpolated at the projection x - phi n (kappa = div n, central
 * differences: compute_normals + qnnn::get_curvature), proj[s] = that
 * projection, node[s] = p*n_nodes + node index.  *n_samples (host) = number
 * found; only the first max_samples are written. */
int casl_curvature_samples(const double *phi, int n_patch, int64_t n_nodes,
                           int M, const int32_t *grid, double h, double x0,
                           double y0, float *feat, double *proj,
                           int64_t *node, int64_t max_samples,
                           int64_t *n_samples, void *stream);

/* Network inputs of n samples (feat [n][10] as above): out[2s] = the sample
 * negated when feat[s][9] > 0 and rotated so its centre gradient lies in the
 * first quadrant, out[2s+1] = that reflected about y = x.  Replaces the input
 * preparation of kml::Curvature::_computeHybridHK
 * (src/my_p4est_curvature_ml.cpp:483-520, :327-398). */
int casl_curvature_normalize(const float *feat, int64_t n, float *out,
                             void *stream);
/* Hybrid h*kappa from the network outputs pred[2n] of the normalised inputs
 * (src/my_p4est_curvature_ml.cpp:520-556): feat[s][9] where |feat[s][9]| < lo,
 * otherwise the mean prediction (+ -|ihk| when residual != 0) blended towards
 * -|ihk| up to hi, with the sign of feat[s][9]. */
int casl_curvature_hybrid(const float *feat, const float *pred, int64_t n,
                          float lo, float hi, int residual, float *hk,
                          void *stream);

/* Learned-curvature samples on any grid (2D or 3D), through the stencil
 * tables: every node with |phi| < band whose 3^dim block of nodes at spacing
 * h is reachable by face arms of length exactly h that end on a node (x,
 * then y, then z).  feat[s] = phi/h over the block (x slowest, z fastest),
 * then h*kappa interpolated inside the block at the projection x - phi*n;
 * offset[s] =
 * -phi*n (dim doubles); node[s] = node index.  phi, normals ([dim][N], from
 * casl_normals) and kappa (from casl_mean_curvature) are FP64 device arrays;
 * feat is [cap][3^dim + 1] FP32.  *n_samples may exceed max_samples (then
 * only max_samples rows were written, in no particular order); *n_skipped
 * counts band nodes without a uniform block.  The 2D features equal those of
 * kml::Curvature's stencils (my_p4est_curvature_ml.cpp:414-480); the
 * reference has no 3D counterpart. */
int casl_curvature_grid_samples(casl_stencil_t *st, const double *phi,
                                const double *normals, const double *kappa,
                                double h, double band, float *feat,
                                double *offset, int64_t *node,
                                int64_t max_samples, int64_t *n_samples,
                                int64_t *n_skipped, void *stream);
/* 3D network inputs: [n][28] samples -> [n][28], negated when ihk > 0, axes
 * reflected so the centred differences of phi are >= 0 and ordered by
 * decreasing difference (one row per sample; every symmetry of the cube maps
 * a stencil to the same row). */
int casl_curvature_normalize_3d(const float *feat, int64_t n, float *out,
                                void *stream);
/* 3D hybrid h*kappa: as casl_curvature_hybrid with one prediction per
 * sample and ihk = feat[s][27]; with residual != 0, predicted corrections
 * with |pred| <= deadband are dropped (the numerical value is kept where the
 * network's correction is within its noise floor). */
int casl_curvature_hybrid_3d(const float *feat, const float *pred, int64_t n,
                             float lo, float hi, int residual, float deadband,
                             float *hk, void *stream);

typedef struct casl_mlp casl_mlp_t;
/* The learned h*kappa of n samples in one launch, for a network of widths
 * (10|28)-128-128-128-128-1 (2D | 3D inputs): casl_curvature_normalize
 * (_3d), casl_mlp_forward and casl_curvature_hybrid(_3d) fused, the weights
 * and activations held in LDS, every layer on BF16 MFMA with FP32
 * accumulation.  The same values, bit for bit, as those calls in a row
 * (deadband applies to 3D only), without their intermediate arrays.  Replaces
 * kml::Curvature::_computeHybridHK's network evaluation and blend
 * (src/my_p4est_curvature_ml.cpp:483-556, the network itself :214-286). */
int casl_curvature_learned(casl_mlp_t *m, const float *feat, int64_t n,
                           float lo, float hi, int residual, float deadband,
                           float *hk, void *stream);
/* Fully connected ReLU network with widths dims[0..n_layers] (no ReLU on the
 * output), He-normal initialisation from seed, zero biases. */
int casl_mlp_create(const int32_t *dims, int n_layers, uint64_t seed,
                    casl_mlp_t **out);
int casl_mlp_destroy(casl_mlp_t *m);
int casl_mlp_num_params(const casl_mlp_t *m, int64_t *n_weights,
                        int64_t *n_biases);
/* input standardisation x -> (x - mean)/stdev (host arrays of dims[0]) */
int casl_mlp_set_scaler(casl_mlp_t *m, const float *mean, const float *stdev);
/* host arrays: weights per layer [in][out] concatenated, biases per layer */
int casl_mlp_get_params(const casl_mlp_t *m, float *weights, float *biases);
int casl_mlp_set_params(casl_mlp_t *m, const float *weights,
                        const float *biases);
int casl_mlp_get_grads(const casl_mlp_t *m, float *grad_weights,
                       float *grad_biases);
/* y[n][dims[L]] = net(x[n][dims[0]]) */
int casl_mlp_forward(casl_mlp_t *m, const float *x, int n, float *y,
                     void *stream);
/* gradients of mean((y - t)^2) w.r.t. all parameters; *loss is host */
int casl_mlp_backward(casl_mlp_t *m, const float *x, const float *t, int n,
                      float *loss, void *stream);
int casl_mlp_adam_step(casl_mlp_t *m, float lr, float beta1, float beta2,
                       float eps, void *stream);
/* C = op(A) op(B) (+ bias[N]) (ReLU), BF16 operands (raw uint16), FP32
 * accumulation; C32 and/or C16 (BF16) written with leading dimension ldc.
 * op(A)(i,k) = trans_a ? A[k*lda+i] : A[i*lda+k]; op(B) likewise. */
int casl_bf16_gemm(int M, int N, int K, const uint16_t *A, int lda,
                   int trans_a, const uint16_t *B, int ldb, int trans_b,
                   const float *bias, int relu, float *C32, uint16_t *C16,
                   int ldc, void *stream);

/* ------------------------------------------------ host forest (p4est role)
 * A brick of n_trees[0] x n_trees[1] (x n_trees[2]) trees on the box
 * [box_min, box_max], periodic or not per axis (my_p4est_brick_new,
 * src/my_p4est_tools.c:37-170).  casl_forest_new is the brick of one
 * non-periodic tree. */
typedef struct casl_forest casl_forest_t;

/* Level-set descriptions for refinement and sampling. */
#define CASL_LS_CALLBACK  0 /* fn(xyz, ctx)                                  */
#define CASL_LS_STAR      1 /* 2D: (|x-c| - (r0 + A cos(p*theta)))*(1 + k|x-c|^2)
                               params: cx cy r0 A p k                       */
#define CASL_LS_SPHERES   2 /* 3D: min_i(|x-c_i| - r_i)*(1 + m sin(2 pi x) cos(2 pi y))
                               params: m, then n_items x (cx cy cz r)        */
#define CASL_LS_GYROID    3 /* 3D: (sin(w x)cos(w y) + sin(w y)cos(w z)
                                    + sin(w z)cos(w x) - c)*(1 + e x y z)
                               params: w c e                                 */
#define CASL_LS_SPHERE    4 /* 2D/3D signed distance |x-c| - r; params: c.., r */
#define CASL_LS_PLANE     5 /* 2D/3D a.x + b; params: a.., b                 */
#define CASL_LS_MODULATED 6 /* 2D: (|x-c| - r)*(1 + q sin(5x) cos(4y)) (the
                               reference reinitialization example's phi0);
                               params: cx cy r q                             */

typedef struct casl_levelset {
  int32_t       kind;
  int32_t       n_items;  /* spheres: number of spheres */
  const double *params;
  double (*fn)(const double *xyz, void *ctx);
  void         *ctx;
} casl_levelset;

int casl_levelset_eval(const casl_levelset *ls, int dim, int64_t n,
                       const double *xyz /* [n][dim] */, double *out);

int casl_forest_new(int dim, const double *box_min, const double *box_max,
                    casl_forest_t **out);
/* n_trees: trees per axis (1..1024, NULL: 1), periodic: 0/1 per axis (NULL:
 * none).  Trees are numbered along a Morton curve of their brick positions. */
int casl_forest_new_brick(int dim, const double *box_min, const double *box_max,
                          const int32_t *n_trees, const int32_t *periodic,
                          casl_forest_t **out);
/* Number of trees, brick position of each tree [T][3], tree of each leaf [L]
 * and of each node [N] (the node's canonical tree: the lowest-numbered tree
 * that holds it); any output may be NULL. */
int casl_forest_trees(const casl_forest_t *f, int32_t *n_trees,
                      int32_t *tree_pos, int32_t *leaf_tree,
                      int32_t *node_tree);
int casl_forest_destroy(casl_forest_t *f);
/* refine_levelset_cf with splitting_criteria_cf_t(min_lvl, max_lvl, phi, lip,
 * band), applied max_lvl times (non-recursive refinement, no 2:1 balance). */
int casl_forest_refine_levelset(casl_forest_t *f, int min_lvl, int max_lvl,
                                double lip, double band,
                                const casl_levelset *phi);
/* my_p4est_nodes_new: canonical clamped corner nodes, by tree then in Morton
 * order. */
int casl_forest_build_nodes(casl_forest_t *f);
int casl_forest_sizes(const casl_forest_t *f, int64_t *n_leaves,
                      int64_t *n_nodes, int32_t *max_level);
/* leaves: xyz [L][dim] logical coords in their tree, level [L],
 * local_nodes [L][2^dim] */
int casl_forest_get_leaves(const casl_forest_t *f, int32_t *xyz,
                           int8_t *level, int32_t *local_nodes);
/* nodes: logical (unclamped) coords in their canonical tree [N][dim] and
 * physical coords [N][dim] */
int casl_forest_get_nodes(const casl_forest_t *f, int32_t *ixyz,
                          double *xyz);
/* sample_cf_on_nodes */
int casl_forest_sample(const casl_forest_t *f, const casl_levelset *phi,
                       double *out);
/* init_neighbors / construct_neighbors for every node.  The table pointers
 * stay valid until the forest is destroyed. */
int casl_forest_init_neighbors(casl_forest_t *f, casl_flat_stencil *out);

/* ------------------------------------------ grid construction on the device
 * The same grid as casl_forest_new + casl_forest_refine_levelset +
 * casl_forest_build_nodes + casl_forest_init_neighbors, built on the current
 * device: refine_levelset_cf level by level (my_p4est_refine_coarsen.cpp:17-89),
 * p4est's node numbering (my_p4est_nodes.c:383-446, 614-671) and
 * construct_neighbors (my_p4est_node_neighbors.cpp:114-1201).  The tables,
 * node order and node coordinates equal the host forest's whenever the level
 * set evaluates to the same bits on host and device (always for SPHERE and
 * PLANE; sin, cos and atan2 may differ in the last place, which can move a
 * refinement decision taken exactly at its threshold).  CALLBACK level sets
 * are host-only (CASL_ERR_INVALID).  Synchronous on `stream`. */
typedef struct casl_device_grid casl_device_grid_t;
int casl_device_grid_build(int dim, const double *box_min,
                           const double *box_max, int min_lvl, int max_lvl,
                           double lip, double band, const casl_levelset *phi,
                           void *stream, casl_device_grid_t **out);
int casl_device_grid_destroy(casl_device_grid_t *g);
int casl_device_grid_sizes(const casl_device_grid_t *g, int64_t *n_leaves,
                           int64_t *n_nodes, int32_t *max_level);
/* Device pointers to the tables, laid out as casl_flat_stencil describes;
 * valid until the grid is destroyed. */
int casl_device_grid_tables(const casl_device_grid_t *g,
                            casl_flat_stencil *out);
/* Copies the tables (and the unclamped logical node coordinates, [N][dim])
 * to host arrays; any pointer may be NULL to skip that array. */
int casl_device_grid_download(const casl_device_grid_t *g, int32_t *nbr,
                              double *d, double *d_tr, uint8_t *wall,
                              int32_t *ixyz);
/* sample_cf_on_nodes on the device: phi[n] (device FP64) at every node. */
int casl_device_grid_sample(const casl_device_grid_t *g,
                            const casl_levelset *phi, double *out,
                            void *stream);
/* A device stencil handle (as casl_stencil_create) for the grid's tables,
 * with its tiled encoding built on the device: the tables never leave it. */
int casl_device_grid_stencil(const casl_device_grid_t *g,
                             casl_stencil_t **out);

#ifdef __cplusplus
}
#endif

#endif /* CASL_REINIT_H */
