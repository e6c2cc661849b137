/*
 * warpvbd.h -- C ABI of the MI355X-native Vertex Block Descent solver.
 *
 * Plain pointers and sizes only.  Host arrays are caller-owned; every device
 * buffer is owned by the handle.  Vertex and element arrays use the caller's
 * (original) numbering; the colour-major renumbering is internal.  A handle is
 * bound to the HIP device that was current at vbd_create and to its own stream
 * (default priority); one handle is not re-entrant, separate handles are
 * independent.  Every call returns 0 on success or a negative VBD_E* code;
 * vbd_last_error() gives a message for the calling thread.  Non-convergence
 * is not an error: *iters_out == max_iter then.
 *
 * Reference interfaces replaced (WarpVBD, paths relative to its repository):
 *   vbd_create            VBDSolver.__init__            src/vbd_solver.py:367-499
 *   vbd_reset             VBDSolver.reset               src/vbd_solver.py:501-506
 *   vbd_step              VBDSolver.step                src/vbd_solver.py:509-612
 *   vbd_get_state /
 *   vbd_set_state         VBDSolver.old_positions / .old_velocities (attributes read
 *                         and assigned by callers, src/vbd_solver.py:403-404,535,609-610)
 *   vbd_get_colors        coloring returned by graph_coloring, src/coloring.py:40-72,
 *                         as stored in VBDSolver.color_groups (src/vbd_solver.py:470-483)
 *   vbd_get_incidence     VBDSolver.adj_v2e             src/vbd_solver.py:485-498
 *   vbd_color             graph_coloring_wp (the "parallel" colouring slot)
 *                         src/coloring.py:85-133, returning graph_coloring's
 *                         result (src/coloring.py:24-72) bit-exact
 *   vbd_color_graph       graph_coloring(adjacency)     src/coloring.py:40-72
 *   vbd_use_device        the device= argument          src/vbd_solver.py:383
 */
#ifndef WARPVBD_H
#define WARPVBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vbd_solver vbd_solver_t;

enum {
    VBD_OK = 0,
    VBD_EINVAL = -1,      /* bad argument (null pointer, size, dt <= 0, max_iter < 1) */
    VBD_EHIP = -2,        /* a HIP runtime call failed */
    VBD_ENOMEM = -3,      /* device allocation failed */
    VBD_EMESH = -4,       /* element index outside [0, n_vertices), or a mesh above 2^26 - 1 vertices or
                             vertex-tet incidences (16.7 M tets: the sweep's 32-bit offsets) */
    VBD_ECOLORS = -5,     /* a vertex needs more than 256 colours (src/coloring.py:50-68) */
    VBD_EVALENCE = -6,    /* a vertex has more than 32 incident elements (src/vbd_solver.py:12,529) */
    VBD_EDEGENERATE = -7  /* an element has a singular rest shape (src/vbd_solver.py:336-343) */
};

enum { VBD_MATERIAL_SNH = 0, VBD_MATERIAL_STVK = 1 };

/* Build a solver.  rest_xyz: f64[n_vertices*3] rest (= initial) positions;
 * tets: i32[n_tets*4]; density/youngs/poisson: f64[n_tets] per element;
 * active: u8[n_vertices] (NULL = all active); gravity: f64[3];
 * damping: the reference's damping_coefficient; material: VBD_MATERIAL_*.
 * Rest shape, Lame parameters, colouring and the colour-major slot layout are
 * computed on the device.  Replaces src/vbd_solver.py:367-499. */
int vbd_create(const double *rest_xyz, int32_t n_vertices, const int32_t *tets, int32_t n_tets,
               const double *density, const double *youngs, const double *poisson, const uint8_t *active,
               const double *gravity, double damping, int32_t material, vbd_solver_t **out);

void vbd_destroy(vbd_solver_t *h);

/* Restore x_prev = rest positions, v_prev = 0.  src/vbd_solver.py:501-506. */
int vbd_reset(vbd_solver_t *h);

/* One implicit-Euler time step.  positions: f64[n*3] x_n to step from, or NULL
 * to step from the solver's own old_positions (the usual caller pattern,
 * examples/run_cantilever.py:79-80).  Iterates colour sweeps until
 * max_v |dx_x+dx_y+dx_z| < tol or max_iter iterations.  out_xyz (nullable)
 * receives the new positions; iters_out the number of iterations executed;
 * metric_out the last iteration's convergence metric.  The host waits once per
 * chunk of iterations, not per iteration.  Replaces src/vbd_solver.py:509-612. */
int vbd_step(vbd_solver_t *h, const double *positions, double dt, double tol, int32_t max_iter,
             double *out_xyz, int32_t *iters_out, double *metric_out);

/* vbd_step on device memory, for callers that hold positions in GPU tensors
 * (replaces the wp.array in/out of VBDSolver.step, src/vbd_solver.py:509-612).
 * positions_dev (nullable: start from x_prev) and out_dev (nullable) are f64
 * [N*3] device pointers in original vertex order, on the handle's device.
 * stream: the caller's hipStream_t (NULL = the null stream); the step is
 * ordered after work already queued there, and work queued there afterwards
 * sees out_dev.  The call still returns only when the iterations are done
 * (the iteration count is read back). */
int vbd_step_dev(vbd_solver_t *h, const double *positions_dev, double dt, double tol, int32_t max_iter,
                 double *out_dev, void *stream, int32_t *iters_out, double *metric_out);

/* After a step that ran its last allowed iteration (iters == max_iter): the
 * largest |dx| component of that iteration over the swept vertices, the value
 * the reference prints in its non-convergence warning ("Final dx max:
 * abs(dxs).max()", src/vbd_solver.py:581-582); NaN after a step that stopped
 * earlier.  Single-body solvers (vbd_step, vbd_step_dev, vbd_step_warm). */
int vbd_last_dx_max(vbd_solver_t *h, double *dx_max);

/* Change the material of a solver whose elements share one (E, nu): the
 * slot records are rescaled in place (no rebuild).  Replaces constructing a
 * new VBDSolver with other youngs_modulus / poissons_ratio arrays. */
int vbd_set_uniform_material(vbd_solver_t *h, double youngs, double poisson);
/* Change the gravity vector of a single-body solver (the gravity argument,
 * src/vbd_solver.py:378). */
int vbd_set_gravity(vbd_solver_t *h, const double *gravity);

/* Read / overwrite x_prev and v_prev (either pointer may be NULL). */
int vbd_get_state(vbd_solver_t *h, double *x_prev, double *v_prev);
int vbd_set_state(vbd_solver_t *h, const double *x_prev, const double *v_prev);

/* Device snapshots of the current positions, so a returned position array stays
 * valid after later steps without a host copy.  slot in [0, VBD_MAX_SNAPSHOTS). */
#define VBD_MAX_SNAPSHOTS 16
int vbd_snapshot(vbd_solver_t *h, int32_t slot);
int vbd_read_snapshot(vbd_solver_t *h, int32_t slot, double *out_xyz);
int vbd_load_snapshot(vbd_solver_t *h, int32_t slot); /* x_prev <- snapshot */

/* ------------------------------------------------------------------ batches
 * Many bodies in one handle (the reference's parallel environments,
 * examples/run_mass_spring_parallel.py:35-119, concatenate bodies into one
 * mesh with one global convergence test; here each body converges on its own
 * and its result equals its single-body solve).  Body b owns vertices
 * [body_vertex_offsets[b], body_vertex_offsets[b+1]); every element must lie in
 * one body.  gravity: f64[n_bodies*3].  Per-body materials come through the
 * per-element arrays.  vbd_step/get/set_state also work on a batched handle
 * (vbd_step then reports the largest per-body iteration count). */
int vbd_batch_create(const double *rest_xyz, int32_t n_vertices, const int32_t *tets, int32_t n_tets,
                     const int32_t *body_vertex_offsets, int32_t n_bodies, const double *density,
                     const double *youngs, const double *poisson, const uint8_t *active, const double *gravity,
                     double damping, int32_t material, vbd_solver_t **out);
/* One step of every body; iters_out / metric_out: per body (nullable). */
int vbd_batch_step(vbd_solver_t *h, double dt, double tol, int32_t max_iter, int32_t *iters_out, double *metric_out);
/* Per-body iterations and metrics of the last step (after vbd_step or vbd_step_warm). */
int vbd_batch_results(vbd_solver_t *h, int32_t *iters_out, double *metric_out);
/* Execution form: VBD_BATCH_LAUNCHES (one launch per colour across all bodies,
 * iterations replayed from a captured hipGraph when use_graph; once most
 * bodies have converged the launches cover only the live bodies' vertex
 * ranges), VBD_BATCH_RESIDENT (one workgroup of block_size threads per body
 * runs its whole step), or VBD_BATCH_LAUNCHES_FULL (colour launches over every
 * body in every iteration, without the live-range lists).  All forms give the
 * same result for each body. */
#define VBD_BATCH_LAUNCHES 1
#define VBD_BATCH_RESIDENT 2
#define VBD_BATCH_LAUNCHES_FULL 3
int vbd_batch_set_mode(vbd_solver_t *h, int32_t mode, int32_t lanes_per_vertex, int32_t block_size,
                       int32_t use_graph);
int vbd_batch_get_mode(vbd_solver_t *h, int32_t *mode, int32_t *lanes_per_vertex, int32_t *block_size,
                       int32_t *use_graph);
/* Autotuner: times every (form, lanes per vertex 2/4/8, block size) candidate
 * for trial_steps steps from the current state (each run twice, the second
 * timed, so that graph captures stay out of the timing), restores the state,
 * keeps the fastest and returns its ms per step.  Short trials favour the
 * one-workgroup-per-body form early in a trajectory; trials as long as the
 * run to be tuned pick best. */
int vbd_batch_autotune(vbd_solver_t *h, double dt, double tol, int32_t max_iter, int32_t trial_steps,
                       double *best_ms_per_step);

/* ------------------------------------------------------------------ warm start
 * A 4-layer message-passing network (width 64, BF16 weights on MFMA, FP32
 * accumulation) on the vertex graph of the solver's mesh predicts each step's
 * converged correction x_{n+1} - x_inertia from the state at the start of the
 * step.  No counterpart in WarpVBD.  The network guess is
 * x_inertia + h + dt^2 |g| out, where the history part h is c_prev, the
 * previous step's converged correction, or 2 c_prev - c_prev2, its linear
 * extrapolation, when the step before that one was consecutive (each scaled
 * by (dt / dt of its step)^2; h = 0 right after create, vbd_reset or
 * vbd_load_snapshot), so the network predicts how the correction departs from
 * that extrapolation.  Training samples come from the solver's
 * own plain steps (vbd_mpnn_collect) and live in device memory
 * (sample_capacity of them, 76 bytes per vertex each). */
typedef struct vbd_mpnn vbd_mpnn_t;
int vbd_mpnn_create(vbd_solver_t *h, int32_t sample_capacity, vbd_mpnn_t **out);
void vbd_mpnn_destroy(vbd_mpnn_t *m);
int vbd_mpnn_num_params(void);
/* flat FP32 parameters: W_in[64x32] b_in[64] (W_l[64x128] b_l[64]) x 4 W_out[3x64] b_out[3] */
int vbd_mpnn_set_params(vbd_mpnn_t *m, const float *params);
int vbd_mpnn_get_params(vbd_mpnn_t *m, float *params);
/* material features: one (E, nu) for every body, or one per body (arrays of n_bodies) */
int vbd_mpnn_set_material(vbd_mpnn_t *m, double youngs, double poisson);
int vbd_mpnn_set_body_materials(vbd_mpnn_t *m, const double *youngs, const double *poisson);
/* one plain solver step that stores (features before, correction after minus h) in `slot` */
int vbd_mpnn_collect(vbd_mpnn_t *m, int32_t slot, double dt, double tol, int32_t max_iter, int32_t *iters_out);
/* loss of a stored sample; grads (nullable, vbd_mpnn_num_params floats) receives its gradient */
int vbd_mpnn_loss_grad(vbd_mpnn_t *m, int32_t slot, float *loss_out, float *grads);
/* Adam over stored samples [0, n_samples) */
int vbd_mpnn_train(vbd_mpnn_t *m, int32_t n_samples, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                   float *loss_per_epoch);
/* network output [n,3] (original vertex order, units of dt^2 |g|; the guess adds h) for sample `slot`, or
 * the current state if slot < 0 */
int vbd_mpnn_predict(vbd_mpnn_t *m, int32_t slot, double dt, float *out);
/* a stored sample and the graph, in original vertex order (each pointer nullable) */
int vbd_mpnn_read_sample(vbd_mpnn_t *m, int32_t slot, float *features, float *targets, int32_t *nb_off, int32_t *nb,
                         int32_t *n_edges);
/* Which initial guess vbd_step_warm takes.  CHOOSE (default): per body the
 * one with the smallest first-sweep residual among the plain inertia
 * prediction, the network guess and the history guess x_inertia + h
 * (when there is a c_prev).  NETWORK: always the network's.  PLAIN: always the
 * plain one (the network is not evaluated).  ADAPTIVE: CHOOSE, but after a
 * step where fewer than 2% of the bodies took the network guess the network
 * is skipped for 1, 2, 4, ... up to 16 steps, which then choose between the
 * plain and the history guess only.  HISTORY: that choice on every step,
 * without the network.  In CHOOSE, ADAPTIVE and HISTORY, on steps with a
 * history guess the plain guess is compared only on check steps: after a
 * check where no body took it, the next 1, 2, 4, ... up to 16 such steps
 * leave it out (a HISTORY step then takes the history guess without a
 * comparison); any change from outside a step (state set or passed to a
 * step, material, gravity, reset, snapshot load) makes the next step a check
 * step. */
#define VBD_WARM_CHOOSE 0
#define VBD_WARM_NETWORK 1
#define VBD_WARM_PLAIN 2
#define VBD_WARM_ADAPTIVE 3
#define VBD_WARM_HISTORY 4
int vbd_mpnn_set_policy(vbd_mpnn_t *m, int32_t policy);
/* bodies that started the last vbd_step_warm from the plain, network and
 * history guess (counts[3]) */
int vbd_mpnn_guess_counts(vbd_mpnn_t *m, int32_t *counts);
/* vbd_step (or vbd_batch_step on a batched handle) from x_prev with, per body,
 * the initial guess the policy picks; first-sweep residual = Jacobi
 * max|dx_x+dx_y+dx_z| over the body; *used_network_out: the number of bodies
 * that started from the network guess. */
int vbd_step_warm(vbd_solver_t *h, vbd_mpnn_t *m, double dt, double tol, int32_t max_iter, int32_t *iters_out,
                  double *metric_out, int32_t *used_network_out);

/* Slot record layout.  Each slot is a 16-byte compact record (the tet's other
 * three vertices and a dictionary entry); the entry holds its 44 bytes of
 * constants.  Meshes whose slots share constants (voxel meshes: 5 rest
 * shapes x 4 local indices per material) share entries, so a sweep reads 16
 * bytes per slot from HBM; otherwise each slot has its own entry (64 bytes).
 * The values, and so the results, are the same either way.  A shared
 * dictionary of at most 64 entries (or a batch's) also holds each entry's
 * rest edges, from which SNH sweeps form cof(F) b and J - 1 with fewer
 * operations (a solver's sweeps all use the same form).
 * vbd_use_record_dictionary(0) gives solvers created afterwards one entry per
 * slot (process-wide; default on).  vbd_records reports the number of entries
 * and the HBM bytes per slot a sweep reads.  No counterpart in WarpVBD (its
 * per-tet arrays, src/vbd_solver.py:295-343, are what the records hold). */
void vbd_use_record_dictionary(int32_t enable);
int vbd_records(vbd_solver_t *h, int32_t *n_dictionary, int32_t *bytes_per_slot);
/* Neighbour staging: colours of at least 40,960 vertices swept at the default
 * launch (4 lanes per vertex, 256 threads) read each block's neighbours once,
 * form each (vertex, neighbour) difference once into LDS with the dictionary,
 * and read a 4-byte record per slot (the results are the same bit for bit).
 * It needs a dictionary of at most 64 entries, or, in a batch, at most 64 per
 * 64-vertex block (a batch of many materials lists each block's own).  In a
 * single-body solver those colours also give each vertex 4, 2 or 1 lanes by
 * its slot count (more than 10, 6..10, at most 5), so that a lane walks at
 * most 5 slots of a voxel mesh; the blocks of 1-lane vertices read their
 * neighbours from global memory.  *n_colors: the colours that use staging.
 * No counterpart in WarpVBD. */
int vbd_staged_colors(vbd_solver_t *h, int32_t *n_colors);
/* Neighbour staging in solvers created afterwards (process-wide; default on;
 * off, the big colours keep their lanes per vertex and read every neighbour
 * from global memory). */
void vbd_use_neighbour_staging(int32_t enable);
/* The tail sweep: a single body's trailing colours below small_color (the
 * last 2 to 4 of them; the 256x32x32 beam's colours of 407, 407 and 1
 * vertices) swept in one launch instead of one per colour, when their
 * vertices split into clusters joined by no edge (at most 32 vertices of a
 * colour and 1,024 distinct neighbours each) and the record dictionary is
 * shared: one 1024-thread block per cluster reads its neighbours once into
 * LDS and sweeps the colours one after another from there (the results are
 * the same bit for bit).  *n_colors: the colours it sweeps (0: none),
 * *n_clusters: its blocks.  No counterpart in WarpVBD. */
int vbd_tail_sweep(vbd_solver_t *h, int32_t *n_colors, int32_t *n_clusters);
/* The tail sweep in solvers created, and launch settings changed, afterwards
 * (process-wide; default on). */
void vbd_use_tail_sweep(int32_t enable);
/* The rest-edge form of the SNH slot arithmetic: with a shared record
 * dictionary of at most 64 entries, or a batch's, each entry also holds its
 * rest edges E0 = X_o0 - X_o2, E1 = X_o1 - X_o2, R = X_self - X_o2 (scaled by
 * k = det[b_o0 b_o1 b_self]), and every sweep of the solver forms cof(F) b and
 * J - 1 from them and the displacement differences (15 VALU instructions per
 * slot fewer than from F's rows; another rounding of the same quantities).
 * Larger dictionaries are read from global memory by every slot, where three
 * more loads per slot cost more than they save, so those solvers keep F's
 * rows.  *on: whether the solver uses the form.  No counterpart in WarpVBD
 * (its src/vbd_solver.py:101-124 forms the SNH terms from F). */
int vbd_edge_form(vbd_solver_t *h, int32_t *on);
/* The rest-edge form in solvers created afterwards (process-wide; default on;
 * off, every sweep forms cof(F) b and J - 1 from F's rows). */
void vbd_use_edge_form(int32_t enable);
/* bytes[n_colors]: the HBM bytes one sweep of each colour reads and writes
 * with this layout, sum_v (68 + bytes_per_slot k_v) + 16 |gathered neighbours|:
 * per swept vertex its 16-byte packed position, 32-byte FP64 inertia target,
 * 4-byte slot offset and 16-byte write-back; per gathered neighbour its packed
 * position; with neighbour staging (vbd_staged_colors) 4 more per vertex,
 * 4 bytes per slot and 4 per staged (vertex, neighbour) entry
 * (vbd_color_bytes is the algorithmic model: 40
 * bytes per vertex, the 68 bytes per slot of the reference's per-tet arrays,
 * 12 per neighbour). */
int vbd_color_layout_bytes(vbd_solver_t *h, double *bytes);

/* Mesh facts computed at create. */
int vbd_info(vbd_solver_t *h, int32_t *n_vertices, int32_t *n_tets, int32_t *n_colors, int32_t *max_incident,
             int32_t *max_degree, int64_t *n_slots);
/* colors: i32[n_vertices]; -1 for a vertex above the largest referenced index
 * (src/coloring.py:28). */
int vbd_get_colors(vbd_solver_t *h, int32_t *colors);
/* color_sizes: i32[n_colors] active vertices per colour (the sweep sizes). */
int vbd_get_color_sizes(vbd_solver_t *h, int32_t *color_sizes);
/* adj: i32[n_vertices*max_incident], ascending element ids padded with -1. */
int vbd_get_incidence(vbd_solver_t *h, int32_t *adj);

/* Standalone GPU colouring: sequential first-fit in vertex order, bit-exact
 * (src/coloring.py:24-72).  colors: i32[n_vertices]. */
int vbd_color(const int32_t *tets, int32_t n_tets, int32_t n_vertices, int32_t *colors, int32_t *n_colors_out,
              int32_t *max_degree_out);

/* GPU first-fit colouring of an explicit adjacency list (CSR: adj_off[n+1],
 * adj[adj_off[n]]): colour[v] = lowest colour not used by a neighbour u < v in
 * v's list -- graph_coloring(adjacency), src/coloring.py:40-72, bit-exact. */
int vbd_color_graph(const int32_t *adj_off, const int32_t *adj, int32_t n_vertices, int32_t *colors,
                    int32_t *n_colors_out);

/* vbd_color without any allocation.  No counterpart in WarpVBD: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  tets_dev: i32[n_tets*4] and
 * colors_dev: i32[n_vertices], device memory of the current device; stream:
 * a hipStream_t (NULL = the null stream).  Same colouring as vbd_color.
 * Every intermediate array lives in one 64 MiB device array per GPU that the
 * library holds for the life of the process: 20 bytes per vertex, 16 per tet
 * and 4 per distinct lower-indexed neighbour of each vertex (at most 24 per
 * tet) must fit -- the 256x32x32 beam (1.31 M tets) needs about 37 MB; a
 * larger mesh is rejected with VBD_EINVAL before any of its data is read.
 * Calls on one GPU share that array, so calls that overlap on one GPU race.
 * The call waits on the stream (it reads counts back) and returns when the
 * colours are in colors_dev. */
int vbd_color_noalloc(const int32_t *tets_dev, int32_t n_tets, int32_t n_vertices, int32_t *colors_dev,
                      int32_t *n_colors_out, int32_t *max_degree_out, void *stream);

/* Make HIP device `device` current for the calling thread (the reference's
 * device= argument, src/vbd_solver.py:383). */
int vbd_use_device(int32_t device);

/* Timing of the colour-sweep kernels on the handle's stream: when enabled,
 * each sweep launch of a colour with at least min_vertices active vertices is
 * bracketed by HIP events.  vbd_profile_read returns the summed kernel time
 * (ms), the number of timed launches and their algorithmic bytes (the byte
 * model sum over the colour of 40 + 68 k_v per active vertex plus 12 per
 * distinct gathered neighbour), then resets the counters. */
int vbd_profile_enable(vbd_solver_t *h, int32_t enable, int32_t min_vertices);
int vbd_profile_read(vbd_solver_t *h, double *ms, int64_t *launches, double *bytes);
/* Per-colour algorithmic bytes of one sweep (f64[n_colors]). */
int vbd_color_bytes(vbd_solver_t *h, double *bytes);

/* Execution knobs: lanes per vertex (1,2,4,8,16,32), threads per block (64..1024),
 * vertices below which consecutive colours run fused in one workgroup, and the
 * iteration chunk between host waits (0 = adaptive). */
int vbd_set_launch(vbd_solver_t *h, int32_t lanes_per_vertex, int32_t block_size, int32_t small_color,
                   int32_t chunk);

/* Replay the iterations of a step from captured hipGraphs (default on; launches
 * are enqueued one by one while profiling is enabled). */
int vbd_set_graph(vbd_solver_t *h, int32_t use_graph);

const char *vbd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
