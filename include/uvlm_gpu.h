/*
 * uvlm_gpu.h -- C-ABI of the MI355X (gfx950) UVLM hot path.
 *
 * Plain C types only: pointers to host arrays of double, sizes and flags.
 * Every entry point returns UVLM_OK (0) or a negative error code; the text
 * of the last error is available from uvlm_last_error().  Host arrays are
 * borrowed for the duration of the call only.  Unless stated otherwise a
 * call is host-synchronous: its outputs are valid when it returns.
 *
 * Array layouts follow the reference's MultipleSheet (sheet-major,
 * row-major, K = j + cols*(i + rows*n)) with 3 doubles per point:
 *   nodes  [num][rows+1][cols+1][3]   ring corners of a lattice
 *   gamma  [num][rows][cols]          ring circulations
 * Ring (n,i,j) has corners p0=(n,i,j) p1=(n,i,j+1) p2=(n,i+1,j+1) p3=(n,i+1,j).
 *
 * Each entry point names the reference (xeno14/UVLM) function it replaces.
 * One context per host thread; a context is not reentrant.
 */
#ifndef UVLM_GPU_H
#define UVLM_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct uvlm_ctx uvlm_ctx;

/* return codes */
#define UVLM_OK 0
#define UVLM_ERR_ARG (-1)      /* invalid argument or shape */
#define UVLM_ERR_HIP (-2)      /* HIP runtime error */
#define UVLM_ERR_OOM (-3)      /* device allocation failed */
#define UVLM_ERR_SINGULAR (-4) /* zero pivot in the dense solve */
#define UVLM_ERR_STATE (-5)    /* call out of order (e.g. sim not initialised) */
#define UVLM_ERR_CAPACITY (-6) /* wake row capacity exhausted */
#define UVLM_ERR_NOCONV (-7)   /* iterative refinement did not converge */

/* arithmetic of the Biot-Savart sums (flags & 0x0f) */
#define UVLM_PREC_F64_EXACT 0x0 /* reference operation order, no FMA, IEEE div/sqrt */
#define UVLM_PREC_F64 0x1       /* FP64 with FMA and refined hardware rsqrt/rcp */
#define UVLM_PREC_F32C 0x2      /* FP32 tile-relative, FP64-carried accumulation,
                                   FP64 re-evaluation of near-core segments */
/* ring summation variants (uvlm_ring_induced_velocity) */
#define UVLM_SCALE_PER_RING 0x10 /* sum 4 segments, then x Gamma (vortex.cpp:41-50) */
#define UVLM_CHORDWISE_ONLY 0x20 /* segments 0 and 2 only (vortex.cpp:52-65) */
/* one sequential sum per target (no source splitting): bitwise identical to
 * the reference order in F64_EXACT mode */
#define UVLM_SEQUENTIAL 0x100

/* dense solve modes */
#define UVLM_SOLVE_F64 0      /* FP64 LU, partial pivoting, FP64 MFMA updates */
#define UVLM_SOLVE_F32_IR 1   /* FP32 LU + FP64 iterative refinement */
#define UVLM_SOLVE_BF16_IR 2  /* BF16-input MFMA LU updates + FP64 refinement */
#define UVLM_SOLVE_F64_FULLPIV 3 /* FP64, full pivoting in Eigen::FullPivLU's elimination
                                   order: bitwise equal to a sequential FullPivLU */
#define UVLM_SOLVE_F64_INV 4 /* FP64 iterative refinement x += X (b - A x) on a kept
                               approximate inverse X (Newton-Schulz refreshed); the
                               simulation warm-starts from the previous step's
                               circulations.  Stops at max|dx| <= tol max|x|. */

/* wake advection schemes (AdvectFactory names, advect_factory.cpp:22-27) */
#define UVLM_ADVECT_EULER 0
#define UVLM_ADVECT_RK2 1
#define UVLM_ADVECT_RK4 2 /* reference update k1+2k2+2k3+k3 (advect.cpp:172-173) */
#define UVLM_ADVECT_AB2 3
#define UVLM_ADVECT_RK4_CLASSIC 4 /* textbook k1+2k2+2k3+k4 (no reference counterpart) */

/* --- context ------------------------------------------------------------- */
int uvlm_create(int device, uvlm_ctx** out);
int uvlm_destroy(uvlm_ctx* ctx);
const char* uvlm_last_error(const uvlm_ctx* ctx);
/* Core cut-off of the context's Biot-Savart sums (vortex.cpp:31 hard-codes
 * eps = 1e-10; a configured core such as 1e-6 keeps the rule's shape: a
 * segment contributes zero when |r1 x r2|^2 < eps, |r1| < eps or
 * |r2| < eps).  Applies to later calls with UVLM_PREC_F64_EXACT or
 * UVLM_PREC_F64 (ring sums, segment chains, influence matrices); the F32C
 * sums, the far-field split and batches keep the reference's 1e-10 and
 * refuse other values (UVLM_ERR_ARG).  eps in (0, 1e-2). */
int uvlm_set_cutoff(uvlm_ctx* ctx, double eps);

/* The stateless sums (uvlm_induced_velocity, uvlm_ring_induced_velocity)
 * keep the last source set packed on the device, with a host copy of its
 * input: a call whose sources equal it byte for byte (same shape and record
 * layout) skips the upload and the packing, so querying one lattice target
 * by target costs the targets' transfers only.  Counts of source uploads
 * and of such reuses since the context was created (no reference
 * counterpart). */
int uvlm_source_cache_stats(uvlm_ctx* ctx, long long* uploads, long long* hits);

/* Wait for all work queued on the context's stream. */
int uvlm_sync(uvlm_ctx* ctx);

/* --- stateless kernels (host arrays) ---------------------------------------- */

/* Replaces UVLM::BiotSavartLaw(result, start, end, pos) (vortex.cpp:21-39),
 * evaluated for n independent (start, end, pos) triples [n][3] each. */
int uvlm_biot_savart(uvlm_ctx* ctx, const double* start, const double* end,
                     const double* pos, size_t n, unsigned flags, double* out);

/* Replaces UVLM::InducedVelocity(x, pos, gamma) (velocity.h:14-24) for
 * n_targets targets [n][3]: sum over rings in (n,i,j) order of VORING
 * (vortex.cpp:145-160, Gamma applied per segment). */
int uvlm_induced_velocity(uvlm_ctx* ctx, const double* targets, size_t n_targets,
                          const double* nodes, const double* gamma, size_t num,
                          size_t rows, size_t cols, unsigned flags, double* out);

/* Replaces the legacy ring sums UVLM::InducedVelocity(result, pos, first, last)
 * and ChordwiseInducedVelocity (shed.inl:22-33,66-77) over n_rings 4-node
 * rings [n_rings][4][3] with segments node[(k+1)%4] -> node[k]
 * (vortex.h:132-140).  Use UVLM_SCALE_PER_RING for VortexRing::BiotSavartLaw
 * semantics and UVLM_CHORDWISE_ONLY for the chordwise variant. */
int uvlm_ring_induced_velocity(uvlm_ctx* ctx, const double* targets,
                               size_t n_targets, const double* rings,
                               const double* gamma, size_t n_rings,
                               unsigned flags, double* out);

/* Replaces SimpleSimulator::CalcMatrix (simple_simulator.cpp:76-101):
 * A[K][L] = VORING_L(cpos_K; Gamma=1) . normals_K, written row-major [N][N].
 * flags: UVLM_PREC_F64_EXACT evaluates the reference's operation order;
 * UVLM_PREC_F64 and UVLM_PREC_F32C the FP64 rsqrt form of the simulation's
 * fast modes (one rsqrt per corner, one reciprocal per segment, ~1e-14 of
 * max |A| from the reference order). */
int uvlm_influence_matrix(uvlm_ctx* ctx, const double* nodes, size_t num,
                          size_t rows, size_t cols, const double* cpos,
                          const double* normals, unsigned flags, double* A);

/* Allocation-free variant of uvlm_induced_velocity, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  targets [nt][3], nodes [num][rows+1][cols+1][3], gamma
 * [num][rows][cols] and out [nt][3] are DEVICE pointers; the launches go on
 * `stream` (a hipStream_t; NULL = the context's stream) and the call returns
 * without waiting for them.  Source records and partial sums live in a
 * fixed 64 MiB device array per GPU (16 MiB of records: at most 262,144
 * chain elements, or 131,072 rings with UVLM_PREC_F64_EXACT; 48 MiB of
 * partial sums: 3 x chunks x nt doubles); larger calls return
 * UVLM_ERR_CAPACITY before launching anything.  Results are bitwise those
 * of uvlm_induced_velocity with the same flags.  Every call on one GPU
 * shares that array: calls that overlap in time on one GPU (two streams,
 * two threads) race. */
int uvlm_induced_velocity_noalloc(uvlm_ctx* ctx, const double* targets, size_t nt,
                                  const double* nodes, const double* gamma, size_t num,
                                  size_t rows, size_t cols, unsigned flags, double* out,
                                  void* stream);

/* Replaces the legacy internal::CalcMatrix (linear.inl:14-32):
 * A[i][j] = ring_j.BiotSavartLaw(targets_i, Gamma=1) . normals_i for 4-node
 * rings [n][4][3] (segments node[(k+1)%4] -> node[k]), row-major [n][n]. */
int uvlm_ring_influence_matrix(uvlm_ctx* ctx, const double* rings, size_t n,
                               const double* targets, const double* normals,
                               double* A);

/* Replaces Eigen::FullPivLU<MatrixXd>(A).solve(rhs) (simple_simulator.cpp
 * :263-268, linear.inl:125-129): A row-major [n][n].  iters_out (nullable)
 * receives the refinement sweeps of the mixed-precision modes, or the
 * numerical rank (FullPivLU's default threshold) in UVLM_SOLVE_F64_FULLPIV. */
int uvlm_solve(uvlm_ctx* ctx, size_t n, const double* A, const double* rhs,
               double* x, int mode, double tol, int max_iter, int* iters_out);

/* --- device-resident SimpleSimulator state ------------------------------
 * The wake (nodes and circulations) lives in device memory as a row ring
 * buffer, so MultipleSheet::prepend_row is O(cols).  The host keeps the
 * kinematics (Morphing) and hands over nodes and surface velocities. */

/* Allocate the state of num wings of rows x cols rings for at most
 * max_steps shedding steps.  flags: UVLM_PREC_* of the Biot-Savart sums. */
int uvlm_sim_init(uvlm_ctx* ctx, size_t num, size_t rows, size_t cols,
                  size_t max_steps, unsigned flags);

/* Morphed wing nodes [num][rows+1][cols+1][3] of the current step
 * (result of Morphing::Perfome, simple_simulator.cpp:247-253).  Also saves
 * the previous circulations (wing_gamma_prev_, :243).  Asynchronous. */
int uvlm_sim_set_wing(uvlm_ctx* ctx, const double* nodes);

/* Replaces SimpleSimulator::Shed(step) (simple_simulator.cpp:171-188):
 * prepend the trailing-edge node row and, when step > 0, the trailing-edge
 * circulation row of the previous solve.  Asynchronous. */
int uvlm_sim_shed(uvlm_ctx* ctx, size_t step);

/* Replaces CalcMatrix + CalcRhs + FullPivLU (simple_simulator.cpp:76-124,
 * 263-268): vmorph_cp [N][3] is Morphing::Velocity at the reference
 * collocation points (cols major as wing_gamma), ff the forward-flight
 * velocity.  gamma_out [N] (nullable) receives the circulations (sync). */
int uvlm_sim_solve(uvlm_ctx* ctx, const double* vmorph_cp, const double ff[3],
                   int solve_mode, double* gamma_out);

/* Replaces SimpleSimulator::CalcLoad (simple_simulator.cpp:195-233,
 * joukowski.cpp:12-92).  vmorph_edges [num][rows*(2*cols+1)][3] is
 * Morphing::Velocity at the reference midpoints of the wing's unique
 * non-trailing-edge edges: first the spanwise edges (i,j)-(i,j+1) for
 * i<rows, j<cols, then the chordwise edges (i,j)-(i+1,j) for i<rows,
 * j<=cols.  ref_area [num] = chord*span.  The coefficients (CD, CL) per wing
 * are appended to the device history (read with uvlm_sim_get_loads).
 * Asynchronous. */
int uvlm_sim_loads(uvlm_ctx* ctx, const double* vmorph_edges, const double ff[3],
                   double dt, const double* ref_area);

/* Replaces advect::Advection::Advect(&wake_pos_, ...) with the scheme of
 * AdvectFactory (advect.cpp:11-216): Euler, RK2, RK4 (reference formula),
 * AB2, or classic RK4; node rows i <= ignore_steps use Euler in RK2/RK4.
 * Asynchronous. */
int uvlm_sim_advect(uvlm_ctx* ctx, int scheme, size_t ignore_steps,
                    const double ff[3], double dt);

/* Replaces SimpleSimulator::EraseOldestWake (simple_simulator.cpp:235-238). */
int uvlm_sim_erase_oldest_wake(uvlm_ctx* ctx);

/* Wake shape: node rows (s+1 at step s) and circulation rows (s). */
int uvlm_sim_wake_shape(uvlm_ctx* ctx, size_t* node_rows, size_t* gamma_rows);
/* Copy the wake out: nodes [num][node_rows][cols+1][3], gamma
 * [num][gamma_rows][cols] (either nullable).  Synchronous. */
int uvlm_sim_get_wake(uvlm_ctx* ctx, double* nodes, double* gamma);
/* Restore a wake and the current wing circulations (e.g. a snapshot). */
int uvlm_sim_set_wake(uvlm_ctx* ctx, size_t node_rows, size_t gamma_rows,
                      const double* nodes, const double* gamma,
                      const double* wing_gamma);
/* Current wing circulations [N].  Synchronous. */
int uvlm_sim_get_gamma(uvlm_ctx* ctx, double* gamma);
/* Load history: count_inout in = capacity in steps, out = steps recorded;
 * out [count][num][2] (CD, CL).  Synchronous. */
int uvlm_sim_get_loads(uvlm_ctx* ctx, double* out, size_t* count_inout);
/* Biot-Savart pairs evaluated by the simulation so far (one pair = one
 * segment evaluation) and the pairs the reference's own loop structure
 * would have evaluated for the same steps. */
int uvlm_sim_pair_counts(uvlm_ctx* ctx, double* evaluated, double* reference);

/* --- far-field split of the wake sum (no reference counterpart) -----------
 * The reference sums every wake ring at every target (advect.h:16-46).  The
 * split cuts the wake lattice into blocks of 4 x 16 rings merged 2 x 2 into
 * a tree, tiles the targets 64 at a time in Morton order, keeps the sum
 * exact for the blocks near a tile and replaces it, for blocks farther than
 * theta times their radius (beyond the tile's radius), by the multipole
 * expansion of the block's vector potential in its vorticity moments
 * (orders 0..3). */

/* Wake-induced velocity at n_targets targets of the lattice nodes
 * [num][rows+1][cols+1][3], gamma [num][rows][cols] (theta > 1). */
int uvlm_farfield_velocity(uvlm_ctx* ctx, const double* targets, size_t n_targets,
                           const double* nodes, const double* gamma, size_t num,
                           size_t rows, size_t cols, double theta, double* out);
/* Use the split for the wake's velocity in the simulation (theta <= 0
 * switches it off): in the advection stages at the wake's nodes with theta,
 * at the collocation points (RHS) and the edge midpoints (loads) with
 * 2 theta; the tree is built once per wake state.  Every audit_every steps
 * (0: never) the first advection stage is also summed exactly (wakes of
 * 32768 nodes or more: every 8th node, a different eighth each audit; the
 * stage keeps the split's velocities, where a full audit continues with the
 * exact ones); if the largest difference exceeds tol times the largest
 * velocity, the split is switched off for the rest of the run. */
int uvlm_sim_set_farfield(uvlm_ctx* ctx, double theta, int audit_every, double tol);
/* Graded criterion for the advection stages (reach <= 0 switches it off):
 * a tile of wake nodes whose gap g to the bounding sphere of the wings
 * (radius rw) exceeds reach * rw uses theta * reach * rw / g, not below
 * theta_min (> 1).  Velocity errors far downstream move wake whose
 * influence on the wings falls like 1/g^3; the audit weighs the difference
 * at each node by min(1, reach rw / g)^3 accordingly. */
int uvlm_sim_set_farfield_focus(uvlm_ctx* ctx, double reach, double theta_min);
/* Relative difference found by the last audit (-1 before the first), audits
 * run, and whether the split is active. */
int uvlm_sim_farfield_stats(uvlm_ctx* ctx, double* last_audit_err, int* audits, int* active);
/* Counters of UVLM_SOLVE_F64_INV in the simulation: refinement sweeps and
 * Newton-Schulz refreshes since the context was made, rebuilds of the
 * inverse after a diverging refinement, sweeps of the last solve.  No
 * reference counterpart (the reference factors the matrix every step with
 * Eigen::FullPivLU, src/simulator/simple_simulator.cpp:265-266). */
int uvlm_sim_inverse_stats(uvlm_ctx* ctx, long long* sweeps, long long* refreshes,
                           long long* rebuilds, int* last_sweeps);
/* Counters of UVLM_SOLVE_F32_IR / UVLM_SOLVE_BF16_IR in the simulation:
 * solves, FP64 refinement sweeps in total and in the last solve.  No
 * reference counterpart. */
int uvlm_sim_refine_stats(uvlm_ctx* ctx, long long* solves, long long* sweeps,
                          int* last_sweeps);

/* --- far-field correction network (no reference counterpart) -----------
 * A small MLP (32 -> 128 -> 128 -> 128 -> 30, SiLU), evaluated once per
 * (target tile, far wake cluster), that predicts the remainder of the
 * cluster's moment expansion over the tile: its value, gradient and Hessian
 * at the tile centre c_T in delta = (x - c_T) / (0.15 D), D = |c_T - c|, in units
 * of the leading omitted term s rho^4 / (4 pi D^2), from features of the
 * pair (offset direction, r / D, normalised moments Q0..Q3).  Output order:
 * value [3], gradient [i][j] (9), Hessian [i][jk] for jk = 00, 01, 02, 11,
 * 12, 22 (18).  Parameters: FP32 vector of UVLM_MLP_PARAMS floats,
 * W1 [128][32], b1, W2 [128][128], b2, W3, b3, W4 [30][128], b4 [30], two
 * pad words.  The reference sums every wake ring exactly
 * (src/uvlm/advect.h:16-46); the network corrects the far part of the split
 * sum (uvlm_farfield_velocity, uvlm_sim_set_farfield). */
#define UVLM_MLP_IN 32
#define UVLM_MLP_OUT 30
#define UVLM_MLP_PARAMS 41120
#define UVLM_MLP_FP32 0 /* training arithmetic */
#define UVLM_MLP_BF16 1 /* inference: BF16 weights and activations, FP32 accumulate (MFMA) */

/* y [n][30] of features [n][32]. */
int uvlm_mlp_forward(uvlm_ctx* ctx, const float* weights, const float* feats, size_t n,
                     float* y, int mode);
/* loss = mean over the samples of 1/2 |y - targets|^2 (targets [n][30])
 * and its gradient [UVLM_MLP_PARAMS] (FP32, deterministic order). */
int uvlm_mlp_grad(uvlm_ctx* ctx, const float* weights, const float* feats,
                  const float* targets, size_t n, float* grad, float* loss);
/* Adam (beta 0.9 / 0.999, eps 1e-8) for `iters` mini-batches of `batch`
 * samples, drawn without replacement per epoch from a seeded permutation;
 * weights updated in place; losses (nullable) [iters]. */
int uvlm_mlp_train(uvlm_ctx* ctx, float* weights, const float* feats, const float* targets,
                   size_t n, size_t iters, size_t batch, float lr, unsigned seed,
                   float* losses);

/* Network correction of the far part of the split (uvlm_farfield_velocity
 * and the simulation's split): weights NULL switches it off.  Pairs (tile,
 * cluster) of clusters up to 16 level-0 blocks (3 merge levels), with c_T
 * outside the cluster's sphere and a tile radius at most 0.3 D, among the
 * first 256 far clusters of a tile, get the network's polynomial added to
 * their expansion at every target of the tile; a pair with an output
 * |y| > ymax (outside the trained range) is summed exactly instead. */
int uvlm_farfield_set_mlp(uvlm_ctx* ctx, const float* weights, double ymax);
/* Order in which the split tiles n_targets targets (no reference
 * counterpart; exposed for tests): the targets' 30-bit Morton keys in
 * their bounding box, 10 bits per axis interleaved x, y, z from bit 0
 * (keys[n_targets], unsorted), and the target indices sorted by key, equal
 * keys in index order, padded with -1 to a multiple of 64
 * (order[64 * ceil(n_targets / 64)]). */
int uvlm_farfield_target_order(uvlm_ctx* ctx, const double* targets, size_t n_targets,
                               unsigned* keys, int* order);
/* (target, cluster) pairs corrected and pairs tripped to the exact sum
 * since the network was set. */
int uvlm_farfield_mlp_stats(uvlm_ctx* ctx, unsigned long long* pairs,
                            unsigned long long* trips);
/* Training samples of the split at theta: up to *count (tile, far cluster)
 * pairs that the network corrects, drawn with a seeded generator; features
 * [k][32] at the tile centre and the remainder's Taylor coefficients
 * [k][30] in the output order above, by central differences of step 0.02 D
 * of (exact - expansion) / (s rho^4 / (4 pi D^2)); *count <- k.  The _sim_
 * form draws them from the simulation's wake, its nodes being the targets,
 * with the lists of its advection call (graded criterion, per-target test). */
int uvlm_farfield_samples(uvlm_ctx* ctx, const double* targets, size_t nt,
                          const double* nodes, const double* gamma, size_t num, size_t rows,
                          size_t cols, double theta, unsigned seed, float* feats, float* resid,
                          size_t* count);
int uvlm_sim_farfield_samples(uvlm_ctx* ctx, double theta, unsigned seed, float* feats,
                              float* resid, size_t* count);
/* In-run learning: draw up to max_samples samples of the split at theta
 * from the simulation's current wake, train the network on them from
 * `weights` (in/out, warm start; Adam as uvlm_mlp_train) and switch the
 * correction on with ymax = twice the largest sample target; last_loss
 * (nullable) <- the final mini-batch loss. */
int uvlm_sim_farfield_learn(uvlm_ctx* ctx, double theta, size_t max_samples, size_t iters,
                            size_t batch, float lr, unsigned seed, float* weights,
                            float* last_loss);

/* --- batched cases (no reference counterpart) ---------------------------
 * The reference runs one SimpleSimulator per process (simple_simulator.cpp
 * :288-313).  A batch advances many independent single-wing cases of
 * different lattice sizes in lockstep: the whole timestep of every case is
 * one fixed launch sequence (kinematics, shedding, matrix, batched LU,
 * loads, wake advection), replayed from a hipGraph.  Kinematics follow
 * Morphing::Perfome (morphing.cpp:23-43) with
 *   plug(t) = plug[0] cos(plug[1] t + plug[2]),
 *   flap(t) = flap[0] cos(flap[1] t + flap[2]),
 * attack angle alpha, no twist or bend, origin = Morphing::origin.  Sums
 * use UVLM_PREC_F64 segment chains; the solve is FP64 LU with partial
 * pivoting. */
typedef struct uvlm_batch_case {
  size_t rows, cols; /* cols even (symmetric wing) */
  double chord, span; /* reference area chord*span */
  double dt, alpha;
  double origin[3], ff[3], plug[3], flap[3];
} uvlm_batch_case;

/* Allocate a batch of n cases.  wing_nodes: the cases' reference-frame
 * node lattices [(rows+1)][(cols+1)][3], concatenated in case order
 * (e.g. from uvlm_host_build_wing).  At most max_steps+1 steps may run.
 * scheme: UVLM_ADVECT_EULER, _RK2 or _RK4.  Replaces the previous batch of
 * the context. */
int uvlm_batch_init(uvlm_ctx* ctx, const uvlm_batch_case* cases, size_t n,
                    const double* wing_nodes, size_t max_steps, int scheme,
                    size_t ignore_steps);
/* Host-only check of a batch plan (no device work, no context): lays out
 * the n cases as uvlm_batch_init does and replays the store and load
 * extents of every kernel of steps 0..max_steps against the allocations.
 * 0 when every access is in bounds, else UVLM_ERR_ARG (invalid case) or
 * UVLM_ERR_CAPACITY with the first offending access written to msg
 * (nullable, msg_len bytes).  uvlm_batch_init runs the same check. */
int uvlm_batch_check_plan(const uvlm_batch_case* cases, size_t n, size_t max_steps, char* msg,
                          size_t msg_len);
/* Enqueue `steps` timesteps of every case (MainLoop, simple_simulator.cpp
 * :240-286, per case).  use_graph: replay one captured hipGraph per step.
 * The host waits for the device after every step, so that at most one
 * step's launches (a few hundred on the eager path) are queued at a time.
 * Returns after the last step completed. */
int uvlm_batch_run(uvlm_ctx* ctx, size_t steps, int use_graph);
/* Time every compiled launch shape of the batch's N-body on the current
 * state and keep the fastest.  chosen (nullable) receives its index,
 * times_ms (nullable, max_variants entries) the measured times. */
int uvlm_batch_autotune(uvlm_ctx* ctx, int* chosen, double* times_ms, int max_variants);
/* Steps run so far. */
int uvlm_batch_steps(uvlm_ctx* ctx, size_t* steps);
/* Per-case results (synchronous): circulations [rows*cols]; wing nodes
 * [(rows+1)(cols+1)][3]; wake nodes [steps][cols+1][3] and circulations
 * [steps-1][cols] (MultipleSheet order, newest row first); load history
 * [count][2] (CD, CL) with count_inout as for uvlm_sim_get_loads. */
int uvlm_batch_get_gamma(uvlm_ctx* ctx, size_t c, double* gamma);
int uvlm_batch_get_wing(uvlm_ctx* ctx, size_t c, double* nodes);
int uvlm_batch_get_wake(uvlm_ctx* ctx, size_t c, double* nodes, double* gamma);
int uvlm_batch_get_loads(uvlm_ctx* ctx, size_t c, double* out, size_t* count_inout);
/* Segment evaluations of all steps run so far. */
int uvlm_batch_pairs(uvlm_ctx* ctx, double* pairs);
/* Options.  UVLM_BATCH_FORCE_PIVOTED (value 0/1): solve every case through
 * the partially pivoted fallback instead of the checked no-pivot path (the
 * results agree; used to test the fallback). */
#define UVLM_BATCH_FORCE_PIVOTED 1
int uvlm_batch_set_option(uvlm_ctx* ctx, int option, int value);
int uvlm_batch_free(uvlm_ctx* ctx);

/* --- measurement hook (no reference counterpart) ------------------------
 * When enabled, every Biot-Savart N-body launch of the context is bracketed
 * by HIP events on the context's stream.  uvlm_profile_read synchronises
 * and returns the launches recorded since the last read: their count, the
 * summed kernel time in milliseconds and the summed segment evaluations. */
int uvlm_profile_enable(uvlm_ctx* ctx, int enable);
int uvlm_profile_read(uvlm_ctx* ctx, size_t* launches, double* kernel_ms,
                      double* pairs);
/* The same for the near-block sums of the far-field split (k_ff_near): its
 * launches since the last read, their summed time and their (target,
 * segment) pairs, counted on the device. */
int uvlm_profile_read_near(uvlm_ctx* ctx, size_t* launches, double* kernel_ms,
                           double* pairs);

#ifdef __cplusplus
}
#endif
#endif
