/* dmpm -- MI355X-native differentiable MPM stepper, C ABI.
 *
 * This is the drop-in boundary for the reference's `diffmpm_bindings` module
 * (bind/bind.cpp:54-559).  Every entry point below replaces one binding (or
 * the C++ member it forwards to); the comment names it.  The Python module
 * `diffmpm_bindings` in this repository re-exposes the same surface on top of
 * these functions.
 *
 * Conventions
 *  - Every function returns DMPM_OK (0) or a negative error code; nothing
 *    throws across the ABI.  dmpm_last_error() describes the last failure on
 *    the calling thread.
 *  - Particle arrays are in the caller's particle order, row-major:
 *      x, v: (N,3)   F, C, dFc, gradients: (N,3,3), element (r,c) at 9p+3r+c.
 *    `dtype` says whether they are float32 (DMPM_DTYPE_F32) or float64
 *    (DMPM_DTYPE_F64); `on_device` says whether the pointer is device memory
 *    (1) or host memory (0).  Pointers are borrowed for the call only.
 *  - Grid arrays are dense, node (i,j,k) at i*ny*nz + j*nz + k (Grid.h:89-94).
 *  - A dmpm_sim owns all its device memory.  Calls on one dmpm_sim must be
 *    serialised by the caller; distinct sims are independent.
 *  - Work is queued on the sim's stream; functions that return host values
 *    (losses, norms, exported host arrays) synchronise that stream.
 */
#ifndef DMPM_H
#define DMPM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMPM_OK 0
#define DMPM_ERR_ARG -1      /* bad argument (shape, index, null pointer) */
#define DMPM_ERR_INDEX -2    /* layer index out of range (IndexError)     */
#define DMPM_ERR_HIP -3      /* HIP runtime error                          */
#define DMPM_ERR_STATE -4    /* call not valid in the sim's current state  */
#define DMPM_ERR_NOMEM -5    /* device allocation failed                   */

#define DMPM_DTYPE_F32 0
#define DMPM_DTYPE_F64 1

#define DMPM_PREC_F32 0  /* FP32 state and arithmetic; FP64 det/log J/F^-T */
#define DMPM_PREC_F64 1  /* FP64 everywhere */

#define DMPM_ORDER_QUADRATIC 2
#define DMPM_ORDER_CUBIC 3     /* the reference's B-spline (Interpolation.h:5-30) */

#define DMPM_MODEL_FCR 0       /* fixed-corotated (Elasticity.cpp:45-55) */
#define DMPM_MODEL_NEOHOOKEAN 1

#define DMPM_ADJOINT_EXACT 0
#define DMPM_ADJOINT_REFERENCE 1  /* Back_Timestep's formulas verbatim (BackPropagation.cpp:9-144) */

/* Layer fields for dmpm_get_layer / dmpm_get_layer_grad. */
#define DMPM_FIELD_X 0
#define DMPM_FIELD_V 1
#define DMPM_FIELD_F 2
#define DMPM_FIELD_C 3
#define DMPM_FIELD_DFC 4      /* control dFc of the layer            */
#define DMPM_FIELD_FTOTAL 5   /* F + dFc                             */
#define DMPM_FIELD_M 6
#define DMPM_FIELD_VOL 7
#define DMPM_FIELD_LAM 8
#define DMPM_FIELD_MU 9
#define DMPM_FIELD_P 10         /* PK1 stress P(F + dFc) (MaterialPoint::P after P_op_1) */
#define DMPM_FIELD_ADAM_M 11    /* Adam state of the layer's control (MaterialPoint::momentum) */
#define DMPM_FIELD_ADAM_V 12    /* (MaterialPoint::vector)     */
#define DMPM_FIELD_ADAM_VMAX 13 /* (MaterialPoint::vector_max) */

typedef struct dmpm_sim dmpm_sim;

/* Simulation configuration.  Replaces the grid/material part of OptInput
 * (bind.cpp:30-52) plus the constants the reference hard-codes
 * (ForwardSimulation.cpp:13: zeta = 0.25, J_min = 0.6).  dt, drag,
 * smoothing and f_ext start at CompGraph's defaults (CompGraph.h:80-83) when
 * created through dmpm_config_init. */
typedef struct {
    int dims[3];            /* grid nodes per axis (Grid(nx,ny,nz,...), bind.cpp:159-160) */
    double dx;
    double grid_min[3];
    double dt, drag, smoothing;
    double f_ext[3];
    double zeta;            /* bulk viscosity coefficient */
    double jmin;            /* trust-region J_min */
    int order;              /* DMPM_ORDER_* */
    int model;              /* DMPM_MODEL_* */
    int trust_region;       /* 1: J trust region of P_op_2 on */
    int wall_cells;         /* sticky walls: v = 0 on nodes within this many cells of a face (0: off) */
    int adjoint;            /* DMPM_ADJOINT_* */
    int precision;          /* DMPM_PREC_* */
    int checkpoint_interval;/* K: store every K-th layer and recompute the rest (1: store all) */
    int lds_bytes;          /* LDS budget per block for the grid window (0: default 30 KiB with FP32
                               state, 48 KiB with FP64; max 64 KiB) */
    int block;              /* threads per block: 64, 128, 192 or 256 (0: default 256) */
    int unfused;            /* 1: separate G2P and P2G kernels instead of the fused forward step */
    int deterministic;      /* 1: two-pass scatters (sort by cell, node-centric sums in a fixed
                               order): bitwise reproducible, slower */
    int recompute_grids;    /* with checkpoint_interval 1: 0 keeps every layer's P2G grid (m, p)
                               for the backward pass, as CompGraph keeps each layer's Grid
                               (CompGraph.h:13-30), at 4 values per node per layer; 1 recomputes
                               each layer's P2G in the backward pass instead */
    int adjoint_recompute;  /* 0: the adjoint of step t reads G2P's v' and C' from layer t+1 when
                               that layer is the forward step of layer t as it stands (the
                               reference reads layer n+1's C, BackPropagation.cpp:27-31) and the
                               trust region is off or the adjoint is the reference's; 1: always
                               recompute them by a G2P from layer t's grid */
    int split_nodes;        /* thread blocks of the particle kernels take up to `block` consecutive
                               particles (Hilbert order of the uploaded positions), cut where the
                               box of their stencils would exceed this many grid nodes, so that
                               each block's LDS window stays small (0: 704; negative: no cuts) */
} dmpm_config;

/* Fill `cfg` with the reference's physics (cubic, fixed-corotated, zeta
 * 0.25, J_min 0.6, trust region on, reference adjoint) and CompGraph's
 * default dt = 1/120, drag = 0.5, smoothing = 0.1 (CompGraph.h:80-83). */
void dmpm_config_init(dmpm_config* cfg, int nx, int ny, int nz, double dx, const double grid_min[3]);

const char* dmpm_last_error(void);
/* The configuration a sim was created with. */
int dmpm_get_config(const dmpm_sim* sim, dmpm_config* out);
const char* dmpm_version(void);

/* --- lifetime -------------------------------------------------------------- */
/* CompGraph(pc, grid, target_grid) (bind.cpp:164, CompGraph.cpp:21-28) is
 * dmpm_create + dmpm_set_particles + dmpm_set_target_grid_mass. */
int dmpm_create(const dmpm_config* cfg, dmpm_sim** out);
int dmpm_destroy(dmpm_sim* sim);
/* Use `stream` (a hipStream_t) for all work; 0 = the library's own stream. */
int dmpm_set_stream(dmpm_sim* sim, void* stream);
int dmpm_synchronize(dmpm_sim* sim);

/* Layer 0 particles (the PointCloud handed to CompGraph).  dFc may be NULL.
 * Particles are re-ordered internally (Hilbert order of their stencil base cell); every
 * export returns them in the caller's order. */
int dmpm_set_particles(dmpm_sim* sim, int n, const void* x, const void* v, const void* F, const void* C,
                       const void* dFc, const void* m, const void* vol, const void* lam, const void* mu,
                       int dtype, int on_device);
int dmpm_num_particles(const dmpm_sim* sim, int* n);
/* Target node masses (the target Grid after calculate_point_cloud_volumes). */
int dmpm_set_target_grid_mass(dmpm_sim* sim, const void* m, int dtype, int on_device);

/* CompGraph's private simulation parameters, set by run_optimization
 * (CompGraph.cpp:220-224). */
int dmpm_set_sim_params(dmpm_sim* sim, double dt, double drag, double smoothing, const double f_ext[3]);

/* --- CompGraph ------------------------------------------------------------- */
/* set_up_comp_graph(num_layers) (bind.cpp:345-347, CompGraph.cpp:30-40):
 * every layer becomes a copy of layer 0. */
int dmpm_set_up_comp_graph(dmpm_sim* sim, int num_layers);
/* get_num_layers() (bind.cpp:173-175). */
int dmpm_get_num_layers(const dmpm_sim* sim, int* num_layers);
/* compute_forward_pass(start_layer, current_episode) (bind.cpp:349-353,
 * CompGraph.cpp:131-141). */
int dmpm_compute_forward_pass(dmpm_sim* sim, int start_layer);
/* Forward pass that keeps every layer's dFc: a rollout driven by per-step
 * controls (the configs' control forces).  No counterpart in the reference,
 * whose forward pass zeroes dFc of every layer it writes
 * (ForwardSimulation.cpp:156). */
int dmpm_rollout_forward(dmpm_sim* sim, int start_layer);
/* end_layer_mass_loss() (bind.cpp:183-186, CompGraph.cpp:42-129). */
int dmpm_end_layer_mass_loss(dmpm_sim* sim, double* loss);
/* compute_backward_pass(control_layer) (bind.cpp:355-358,
 * CompGraph.cpp:143-211), including the stored render-gradient injection. */
int dmpm_compute_backward_pass(dmpm_sim* sim, int control_layer);
/* Read a field of layer `layer` (get_point_cloud(i).get_positions() etc.,
 * bind.cpp:84-157).  `out` has n*ncomp elements of `dtype`. */
int dmpm_get_layer(dmpm_sim* sim, int layer, int field, void* out, int dtype, int on_device);
/* Read dL/dx, dL/dv, dL/dF or dL/dC (field X, V, F, C) of the layer the
 * last backward pass stopped at (the control layer), or of the last layer.
 * DMPM_FIELD_DFC reads dL/d dFc of the control layer. */
int dmpm_get_layer_grad(dmpm_sim* sim, int layer, int field, void* out, int dtype, int on_device);
/* accumulate_render_grads(dLdF, dLdx) (bind.cpp:189-279): add to the last
 * layer's adjoint now.  Either pointer may be NULL. */
int dmpm_accumulate_last_layer_grad(dmpm_sim* sim, const void* dLdF, const void* dLdx, const void* dLdv,
                                    const void* dLdC, int dtype, int on_device);
/* Replace the last layer's adjoint (seed for an autograd backward). */
int dmpm_set_last_layer_grad(dmpm_sim* sim, const void* dLdx, const void* dLdv, const void* dLdF,
                             const void* dLdC, int dtype, int on_device);
/* set_render_gradients / clear_render_gradients / has_render_gradients /
 * get_render_gradient_info (bind.cpp:359-528). */
int dmpm_set_render_gradients(dmpm_sim* sim, int n, const void* dLdF, const void* dLdx, int dtype, int on_device);
int dmpm_clear_render_gradients(dmpm_sim* sim);
int dmpm_render_gradient_info(const dmpm_sim* sim, int* has, int* num_points, double* norm_F, double* norm_x);
/* Reorder the particle storage by the Hilbert order of layer `layer`'s
 * positions (thread blocks and cell runs rebuilt as at upload).  Particles
 * drift away from the order they were uploaded in, which widens the blocks'
 * LDS windows and breaks the runs the scatters merge; sorting by a layer in
 * the middle of an episode halves the drift.  Every per-particle array (all
 * layers, controls, Adam state, gradients) is permuted; results do not
 * change beyond summation order, and getters keep the caller's order.
 * Refused (DMPM_ERR_STATE) while a captured graph or batched episode of the
 * sim is alive: those hold its device buffers.  No counterpart in the
 * reference (its particles have no device order). */
int dmpm_sort_particles(dmpm_sim* sim, int layer);
/* Reorder the particle storage per episode segment: the layers are split
 * into `segments` runs of equal length (multiples of the checkpoint
 * interval), and every per-particle array of a run's layers is stored in the
 * Hilbert order of the run's middle layer, its thread blocks cut so that
 * their windows also fit at the stored layers bounding the run.  A body
 * that is advected or deformed drifts away from any single order within a
 * few dozen steps (runs of equal cells break up, windows grow); segment
 * orders keep every layer near an order of its own.  The forward step into
 * a new segment runs in the previous order and writes its output into the
 * new one's rows; the adjoint's carry goes back the same way (and the
 * boundary layer's v and C are permuted back, where the adjoint reads
 * them).  Besides that the sim holds per order a permutation,
 * its block table and the four static planes, and three index maps.
 * Results do not change beyond summation order, and getters,
 * setters and gradients keep the caller's order.  segments = 1 is one order
 * for the whole episode.  dmpm_sort_particles, dmpm_set_particles,
 * dmpm_set_up_comp_graph, dmpm_promote_last_as_initial and
 * dmpm_batch_capture return the sim to one storage order.  Refused
 * (DMPM_ERR_STATE) while a captured graph or batched episode of the sim is
 * alive, and in deterministic mode.  No counterpart in the reference (its
 * particles have no device order). */
int dmpm_sort_segments(dmpm_sim* sim, int segments);
/* promote_last_as_initial(carry_grid) (bind.cpp:282-291). */
int dmpm_promote_last_as_initial(dmpm_sim* sim, int carry_grid);

/* --- controls and optimisation ------------------------------------------- */
/* Set/read dFc of a layer (the per-step control).  NULL clears it. */
int dmpm_set_control(dmpm_sim* sim, int layer, const void* dFc, int dtype, int on_device);
/* Keep dL/d dFc of `layer` from the next backward passes (1) or not (0);
 * read it with dmpm_get_control_grad. */
int dmpm_request_control_grad(dmpm_sim* sim, int layer, int on);
/* The Adam state of `layer`'s control (MaterialPoint::momentum, vector,
 * vector_max, PointCloud.cpp:83-147), e.g. when resuming an optimisation;
 * read it back with dmpm_get_layer(..., DMPM_FIELD_ADAM_*, ...). */
int dmpm_set_adam_state(dmpm_sim* sim, int layer, const void* m, const void* v, const void* vmax, int dtype,
                        int on_device);
int dmpm_get_control_grad(dmpm_sim* sim, int layer, void* out, int dtype, int on_device);
/* PointCloud::Compute_dLdF_Norm of the control layer's gradient (PointCloud.cpp:24-35). */
int dmpm_control_grad_norm(dmpm_sim* sim, double* norm);
/* PointCloud::Descend_Adam on layer `layer` (PointCloud.cpp:83-147), using
 * the control gradient of the last backward pass. */
int dmpm_descend_adam(dmpm_sim* sim, int layer, double alpha, double grad_norm, double beta1, double beta2,
                      double eps, int timestep);
/* optimize_single_timestep(...) (bind.cpp:294-317, CompGraph.cpp:369-465). */
int dmpm_optimize_single_timestep(dmpm_sim* sim, int timestep_idx, int max_gd_iters, int current_episode,
                                  double initial_alpha, int max_line_search_iters);
/* The Adam step count OptimizeSingleTimestep shares across calls and
 * graphs: the reference keeps it in a function-local static
 * (CompGraph.cpp:396), so it lives for the process.  Writes it to *count
 * (may be NULL), then sets it to 0 when reset is 1 (a fresh process). */
int dmpm_single_timestep_adam_count(int* count, int reset);
/* run_optimization(opt) (bind.cpp:165-172, CompGraph.cpp:213-368). */
int dmpm_run_optimization(dmpm_sim* sim, int num_steps, double dt, double drag, const double f_ext[3],
                          int control_stride, int max_gd_iters, int max_line_search_iters, double initial_alpha,
                          double gd_tol, double smoothing, int current_episodes, double* final_loss);

/* --- free functions (bind.cpp:531-559) ------------------------------------- */
/* p2g(pc, grid) (bind.cpp:546-549): the node masses of a dt = 0 P2G of the
 * n particles x (n,3), m (n) are ADDED to grid_m (G = nx*ny*nz values), as
 * the reference's P2G adds onto the Grid it is given.  Computed in the
 * precision of dtype; with on_device = 1 every array is device memory and
 * nothing is copied to the host. */
int dmpm_p2g_mass(const dmpm_config* cfg, int n, const void* x, const void* m, void* grid_m, int dtype,
                  int on_device);
/* calculate_point_cloud_volumes(pc, grid) (bind.cpp:551-554,
 * ForwardSimulation.cpp:234-271): the P2G of dmpm_p2g_mass adds the masses
 * to grid_m (zeros for a fresh Grid), then vol_p = m_p dx^3 / sum_i w_ip m_i
 * over the summed node masses (0 where that sum is not above FLT_EPSILON
 * dx^3). */
int dmpm_calculate_volumes(const dmpm_config* cfg, int n, const void* x, const void* m, void* vol, void* grid_m,
                           int dtype, int on_device);
/* dmpm_calculate_volumes for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference.  Same results as
 * dmpm_calculate_volumes with on_device = 1: x (n,3), m (n), vol (n) and
 * grid_m (nx*ny*nz, the P2G adds to it) are device arrays of dtype; the work
 * is queued on `stream` (a hipStream_t, 0 = the null stream) and the call
 * returns without waiting.  The particle planes the kernels use live in one
 * device array per GPU of DMPM_NOALLOC_SCRATCH_BYTES (4 planes: at most
 * 1,048,576 particles in FP32, 524,288 in FP64; more is refused with
 * DMPM_ERR_ARG); nothing is allocated.  Calls that overlap on one GPU share
 * that array and race: serialise them (one stream, or events). */
#define DMPM_NOALLOC_SCRATCH_BYTES (16u << 20)
int dmpm_calculate_volumes_noalloc(const dmpm_config* cfg, int n, const void* x, const void* m, void* vol,
                                   void* grid_m, int dtype, void* stream);
/* calculate_lame_parameters(E, nu) (bind.cpp:540-544, Elasticity.cpp:149-153). */
int dmpm_lame_parameters(double E, double nu, double* lam, double* mu);

/* --- hipGraph capture of episodes ------------------------------------------ */
/* No counterpart in the reference.  Captures, for each of `count` sims, the
 * forward pass from `start_layer` (keeping every layer's dFc, as
 * dmpm_rollout_forward) and the backward pass to `stop_layer`, each on its
 * sim's stream, into one graph whose branches run concurrently (the batched
 * workload: many independent instances in one launch).  Sims must have
 * distinct streams and must not be in deterministic mode.  The graph reads
 * the sims' buffers: set controls / seeds / particles through the usual calls
 * (outside a launch) and relaunch; changing the number of layers, the
 * particle count or the checkpoint interval needs a new capture. */
/* --- on-disk particle state ------------------------------------------------
 * DMPM_FILE_CEREAL: the reference's PointCloud binary file
 *   (PointCloud::WriteEntirePointCloudToBinaryFile / ReadEntirePointCloudFromBinaryFile,
 *   PointCloud.cpp:382-410, cereal BinaryOutputArchive of MaterialPoint.h:55-71):
 *   uint64 n, then per point x, v, F, C, P, m, vol, lam, mu, dFc, dLdx, dLdv,
 *   dLdF, momentum, vector, vector_max; a 3-vector is int32 3 + 3 float32, a
 *   3x3 matrix int32 3, int32 3 + 9 float32 row by row, scalars float32
 *   (432 bytes per point, little endian).
 * DMPM_FILE_NATIVE: this library's checkpoint format, little endian:
 *   header (64 bytes): char magic[8] = "DMPMSTAT"; uint32 version = 1;
 *     uint32 dtype (DMPM_DTYPE_*); uint64 n; uint32 nfields; int32 layer the
 *     file was written from (-1: from arrays); 32 reserved bytes (0);
 *   field table: nfields x { char name[16] (NUL padded); uint32 ncomp;
 *     uint32 reserved };
 *   data: per field in table order, n x ncomp values of dtype, particle-major,
 *     3x3 matrices row by row.
 *   Field names as in the cereal record (x v F C P m vol lam mu dFc dLdx dLdv
 *   dLdF momentum vector vector_max); readers require x v F C m vol lam mu and
 *   take absent optional fields as zeros.
 * Arrays are host memory in the caller's particle order: vectors (n,3),
 * matrices (n,3,3), scalars (n,). */
#define DMPM_FILE_NATIVE 0
#define DMPM_FILE_CEREAL 1
typedef struct {
    void *x, *v, *F, *C, *P, *m, *vol, *lam, *mu, *dFc, *dLdx, *dLdv, *dLdF, *adam_m, *adam_v, *adam_vmax;
} dmpm_particle_arrays;
/* Particle count (and value type) of a file. */
int dmpm_file_info(const char* path, int format, int* n, int* dtype);
/* Write arrays (x v F C m vol lam mu required; NULL optional fields are
 * written as zeros).  A cereal file always holds float32. */
int dmpm_write_file(const char* path, int format, int n, const dmpm_particle_arrays* arrays, int dtype);
/* Read a file of exactly n particles into the non-NULL arrays. */
int dmpm_read_file(const char* path, int format, int n, dmpm_particle_arrays* arrays, int dtype);
/* Write one layer of a sim: its state, control, statics, stress P(F + dFc),
 * gradients where the sim keeps them (last layer, control layer of the last
 * backward pass; zeros elsewhere) and Adam state (PointCloud.cpp:382-394). */
int dmpm_write_layer(dmpm_sim* sim, int layer, const char* path, int format);
/* dmpm_set_particles from a file, plus layer 0's Adam state when the file
 * holds a non-zero one (PointCloud.cpp:396-410). */
int dmpm_read_particles(dmpm_sim* sim, const char* path, int format);

/* --- particle initialisation -------------------------------------------------
 * Generalised winding numbers w[np] of points P (np x 3) with respect to the
 * triangle mesh V (nv x 3), F (nf x 3, 0-based vertex indices), computed on
 * the device (host arrays in and out).  A lattice point is inside where
 * w >= 1/2: the test of GeometryLoading::GeneratePointCloudFromWatertightTriangleMesh
 * (GeometryLoading.cpp:10-83, libigl signed distance with sign type
 * SIGNED_DISTANCE_TYPE_WINDING_NUMBER), which load_point_cloud_from_obj
 * (bind.cpp:531-538) runs on the lattice of spacing grid_dx /
 * points_per_cell_cuberoot over the mesh's bounding box. */
int dmpm_winding_numbers(const double* V, int nv, const int* F, int nf, const double* P, int np, double* w);

/* --- batched episodes ----------------------------------------------------------
 * Many independent sims (different particle counts, grids and materials; the
 * same precision, B-spline order, model, adjoint mode, launch shape and
 * episode structure) stepped together: every kernel of the episode runs as
 * one launch over all instances (each thread block belongs to one instance),
 * and the whole episode -- forward from start_layer keeping per-layer
 * controls, backward to stop_layer -- is one hipGraph.  No counterpart in the
 * reference, which runs one CompGraph at a time (CompGraph.cpp:131-211).
 * Capture records each sim's episode without running it (host bookkeeping
 * such as which layers are valid is updated as for a run); launch replays it
 * on the batch's own stream.  Synchronise the sims before a launch and the
 * batch before reading the sims.  Inputs may change between launches through
 * the usual calls; particle counts and layer counts may not. */
typedef struct dmpm_batch dmpm_batch;
int dmpm_batch_capture(dmpm_sim** sims, int count, int start_layer, int stop_layer, dmpm_batch** out);
int dmpm_batch_launch(dmpm_batch* batch);
int dmpm_batch_synchronize(dmpm_batch* batch);
int dmpm_batch_destroy(dmpm_batch* batch);

typedef struct dmpm_graph dmpm_graph;
int dmpm_graph_capture(dmpm_sim** sims, int count, int start_layer, int stop_layer, dmpm_graph** out);
int dmpm_graph_launch(dmpm_graph* graph);
int dmpm_graph_synchronize(dmpm_graph* graph);
int dmpm_graph_destroy(dmpm_graph* graph);

/* --- measurement ----------------------------------------------------------- */
/* Time of the last forward/backward pass in ms (HIP events), and the number
 * of particle steps it covered. */
int dmpm_last_timing(const dmpm_sim* sim, double* fwd_ms, double* bwd_ms);
/* Per-kernel timing with HIP events on the sim's stream.  capacity > 0
 * starts recording (up to `capacity` launches, events pre-created here);
 * 0 stops.  Kinds: DMPM_KERNEL_*.  dmpm_kernel_stats synchronises and
 * returns the summed duration and the launch count of one kind. */
#define DMPM_KERNEL_P2G 0
#define DMPM_KERNEL_G2P 1
#define DMPM_KERNEL_ADJ_SCATTER 2
#define DMPM_KERNEL_ADJ_GATHER 3
#define DMPM_KERNEL_GRID_CLEAR 4
#define DMPM_KERNEL_STEP 5
#define DMPM_KERNEL_ADJ_FUSED 6  /* adjoint gather of step t + adjoint scatter of step t-1 in one launch */
int dmpm_kernel_timer(dmpm_sim* sim, int capacity);
int dmpm_kernel_stats(dmpm_sim* sim, int kind, double* total_ms, int* count);
/* Scatter windows counted since the last reset: blocks whose stencil box
 * was held in LDS, and blocks that fell back to global atomics. */
int dmpm_window_stats(dmpm_sim* sim, long long* lds_blocks, long long* fallback_blocks, int reset);
/* Window counting on (1) or off (0, the default).  Counting costs one
 * device-scope atomic per block of every scatter launch (about 2 % of an
 * episode at 1M particles); launches recorded into a graph or a batched
 * episode keep the setting they were recorded with.  No counterpart in the
 * reference (a diagnostic of this library's LDS windows). */
int dmpm_window_counters(dmpm_sim* sim, int on);
/* Device bytes the sim holds (buffers in use plus freed buffers it keeps
 * for reuse: setting up again with the same particle count takes its
 * buffers from them), and the peak.  Buffers of an old particle count are given back when
 * dmpm_set_particles changes it; dmpm_destroy gives back everything. */
int dmpm_memory(const dmpm_sim* sim, size_t* current, size_t* peak);

#ifdef __cplusplus
}
#endif
#endif
