/*
 * msbgx.h -- C ABI of the MI355X-native MSBG FLIP pressure projection.
 *
 * Drop-in boundary for the reference solver entry points used by flip_sim and
 * msbg_demo.  Every function returns MSBGX_OK (0) or a negative error code;
 * msbgx_last_error() returns a message for the calling thread's last error.
 * No torch or HIP types cross this boundary: plain pointers and sizes only.
 * Device memory is owned by the handle and never exposed.
 *
 * Each entry point names the reference interface it replaces (src/ paths are
 * relative to the reference repository).
 */
#ifndef MSBGX_H
#define MSBGX_H

#include <stdint.h>
#include "msbgx_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MSBGX_OK 0
#define MSBGX_EINVAL (-1)   /* invalid argument                         */
#define MSBGX_ENOMEM (-2)   /* device or host allocation failed         */
#define MSBGX_EHIP (-3)     /* HIP runtime error                        */
#define MSBGX_ESTATE (-4)   /* call out of order (e.g. solve w/o prepare) */

/* Domain boundary kinds (src/msbg.h:191-192). */
#define MSBGX_DBC_OPEN 0
#define MSBGX_DBC_SOLID 1

/* Channel ids for inspection and the per-kernel entry points. */
#define MSBGX_CH_PRESSURE 0     /* CH_PRESSURE                          */
#define MSBGX_CH_DIVERGENCE 1   /* CH_DIVERGENCE (right-hand side)      */
#define MSBGX_CH_R 2            /* CH_FLOAT_2   PCG residual / MG rhs   */
#define MSBGX_CH_Z 3            /* CH_FLOAT_3   preconditioned residual */
#define MSBGX_CH_D 4            /* CH_FLOAT_4   search direction        */
#define MSBGX_CH_Q 5            /* CH_FLOAT_6   A*d / MG residual       */
#define MSBGX_CH_FACE_U 7       /* CH_FACE_AREA x                       */
#define MSBGX_CH_FACE_V 8
#define MSBGX_CH_FACE_W 9
#define MSBGX_CH_DIAGONAL 10    /* CH_DIAGONAL (inverse diagonal)       */
#define MSBGX_CH_FLAGS 11       /* CH_CELL_FLAGS (uint16 per cell)      */

typedef struct msbgx_grid msbgx_grid;

/* Reference: MSBG::FlipPressureCallbacks (src/msbg.h:829-835). */
typedef struct msbgx_callbacks {
  uint16_t (*sample_cell_type)(void *user, int x, int y, int z);
  float (*sample_face_coeff)(void *user, int dir, int x, int y, int z);
  float (*sample_rhs)(void *user, int x, int y, int z);
  void *user;
} msbgx_callbacks;

/* Reference: MSBG::FlipPressureSolveParams (src/msbg.h:837-844). */
typedef struct msbgx_solve_params {
  int n_pre;        /* 3    */
  int n_post;       /* 3    */
  int n_coarse;     /* 24   */
  int max_iter;     /* 500  */
  float tol;        /* 5e-4 */
} msbgx_solve_params;

/* Storage precision of the smoother's operands (precision form, a build
 * addition without reference counterpart): the block Gauss-Seidel smoother
 * reads 16-bit copies of the face coefficients and the inverse diagonal and
 * computes in FP32; the Krylov operator q = A d keeps the FP32 coefficients. */
#define MSBGX_PREC_FP32 0
#define MSBGX_PREC_FP16 1
#define MSBGX_PREC_BF16 2

/* Solver option bits for msbgx_set_options(). */
#define MSBGX_OPT_LITERAL_VCYCLE 0x1u  /* run residual/restriction/prolongation even when every
                                          coarser level has no relaxation block             */
#define MSBGX_OPT_NO_GRAPH 0x2u        /* launch PCG iterations eagerly instead of replaying a
                                          captured hipGraph                                  */
#define MSBGX_OPT_NO_COMPACT 0x4u      /* batched solves: keep converged problems' blocks in the
                                          launches instead of dropping them                  */
#define MSBGX_OPT_UNFUSED_COARSE 0x8u  /* learned coarse form with two directions on 32^3
                                          blocks: their residuals one at a time and the
                                          Galerkin sums in a pass of their own (the form of
                                          other block sizes and scale counts)                */
#define MSBGX_OPT_FP32_DIRECTIONS 0x10u /* learned coarse form, two-direction pass: keep the
                                          directions and their residuals in FP32 (default:
                                          BF16 storage, FP32 arithmetic, A zc on the PCG
                                          blocks only)                                       */

const char *msbgx_last_error(void);
void msbgx_default_solve_params(msbgx_solve_params *p);

/* Reference: MultiresSparseGrid::create (src/msbg.h:859-875, src/msbg.cpp:117-1106),
 * with blockSize0 given and nResLevels in [1,5].  Grid dims must be multiples
 * of bsx0; bsx0 in {4,8,16,32} and bsx0 >> (n_levels-1) >= 2. */
int msbgx_create(msbgx_grid **out, int sx0, int sy0, int sz0, int bsx0, int n_levels);
/* Reference: MultiresSparseGrid::destroy (src/msbg.h:877). */
void msbgx_destroy(msbgx_grid *g);
int msbgx_num_blocks(const msbgx_grid *g);
int msbgx_set_options(msbgx_grid *g, unsigned options);

/* Reference: setDomainBoundarySpec_ (src/msbg.h:935-940, src/msbg.cpp:2462). */
int msbgx_set_domain_bc(msbgx_grid *g, int left, int right, int bottom, int top, int front, int back);
/* Reference: regularizeRefinementMap (src/msbg.h:922-926, src/msbg.cpp:1202-1420).
 * `out` may equal `in`. */
int msbgx_regularize_refinement_map(const msbgx_grid *g, const int *in, int *out);
/* Reference: setRefinementMap (src/msbg.h:928-934, src/msbg.cpp:1428-1736),
 * doResetChannels = true. */
int msbgx_set_refinement_map(msbgx_grid *g, const int *levels);

/* Reference: preparePressureSolveFLIP (src/msbg.h:849-852, src/msbg3.cpp:3041-3095).
 * Callbacks are invoked serially on the calling thread with level-local
 * coordinates, as in the reference (src/msbg3.cpp:2494-2576). */
int msbgx_prepare_pressure_solve(msbgx_grid *g, const msbgx_callbacks *cb);
/* Bulk variant of the above: callbacks that read fine-grid arrays of
 * sx0*sy0*sz0 values (x fastest).  rhs may be NULL (all zero). */
int msbgx_prepare_pressure_solve_dense(msbgx_grid *g, const uint16_t *flags, const float *face_u,
                                       const float *face_v, const float *face_w, const float *rhs);
/* Bulk variant over an analytic scene: cell types and face coefficients are
 * sampled on the host, the divergence right-hand side is computed on the GPU
 * from the scene's MAC velocity (flipRhsCB, src/flip_sim.cpp:329-338). */
int msbgx_prepare_pressure_solve_scene(msbgx_grid *g, const msbgx_scene *scene);

/* Reference: solvePressureFLIP (src/msbg.h:854-857, src/msbg3.cpp:3098-3271).
 * Returns the iteration count in *iters (max_iter when not converged) and
 * |r|/|b| in *rel_residual. */
int msbgx_solve_pressure(msbgx_grid *g, const msbgx_solve_params *p, int *iters, float *rel_residual);
/* Precision form (no reference counterpart): selects the smoother operand
 * storage (MSBGX_PREC_*); takes effect immediately, also after prepare.
 * With FP16/BF16 operands the V-cycle's coarsest "solve" (n_coarse relax
 * iterations, src/msbg3.cpp:2423-2430) runs in FP64 (x and arithmetic in
 * double over the FP32 coefficients) for coarsest block edges up to 16. */
int msbgx_set_precision(msbgx_grid *g, int prec);

/* Mixed-precision iterative refinement around the MG-PCG (no reference
 * counterpart).  x64 = 0; repeat: r = b - A x64 in FP64 over the level-0
 * unknowns (coarse-block unknowns held at 0, as in the reference's PCG);
 * stop when |r| <= tol |b|; solve A e = r with the MG-PCG to inner->tol;
 * x64 += e.  The pressure channel receives (float) x64; the FP64 solution
 * of the level-0 blocks is copied to x64_out (in PCG-block order, see
 * msbgx_level0_blocks) when non-NULL.  Returns outer iterations, the summed
 * inner iterations and the final FP64 relative residual. */
int msbgx_solve_pressure_refined(msbgx_grid *g, const msbgx_solve_params *inner, double tol, int max_outer,
                                 int *outer_iters, int *inner_iters, double *rel_residual, double *x64_out);
/* The level-0 blocks the PCG vector operations cover, in their pool order
 * (the order of msbgx_solve_pressure_refined's x64_out); returns the count. */
int msbgx_level0_blocks(const msbgx_grid *g, int *out, int cap);

/* Batched form (no reference counterpart; the reference solves one grid per
 * call, src/msbg3.cpp:3262-3271).  Merges n prepared problems that share the
 * block size, level count and domain boundary spec into one block set; each
 * smoother, apply and transfer launch then covers the blocks of every
 * problem, and each problem keeps its own PCG scalars, smoother weight and
 * convergence test.  The problems are copied (their handles stay usable and
 * may be destroyed).  Block `b` of problem `q` is block
 * msbgx_batch_block_base(batch, q) + b of the batch handle, so the channel
 * inspection functions and the per-kernel entry points work on the batch.
 * Each problem's iterates equal those of msbgx_solve_pressure on its own
 * handle.  Domain spec, refinement map, prepare, msbgx_solve_pressure and the
 * refined solve are rejected on a batch handle. */
int msbgx_batch_create(msbgx_grid **out, msbgx_grid *const *problems, int n);
int msbgx_batch_size(const msbgx_grid *batch);
int msbgx_batch_block_base(const msbgx_grid *batch, int problem);
/* Per-problem iteration counts and |r|/|b| into iters[n], rel_residual[n]. */
int msbgx_batch_solve(msbgx_grid *batch, const msbgx_solve_params *p, int *iters, float *rel_residual);
/* Per-problem relative residual history of the last batched solve. */
int msbgx_batch_residual_history(const msbgx_grid *batch, int problem, double *out, int cap);

/* Autotuner (no reference counterpart): measures, on the handle's own
 * prepared problem (single or batch), the smoother kernel per block size
 * (bit k of *smoother_mask set = the LDS-tile kernel for blocks of
 * bsx0 >> k, FP32 operands, blocks up to 16^3; for 32^3 blocks with 16-bit
 * operands, bit 0 set = the record-streaming kernel instead of the
 * register-resident one) and the PCG iterations per graph replay
 * (*graph_chunk), and keeps the fastest.  The choices change
 * launch shapes only: results stay bit-identical.  Runs trial solves with
 * `p`; the pressure channel is zero on return. */
int msbgx_autotune(msbgx_grid *g, const msbgx_solve_params *p, int *graph_chunk, int *smoother_mask);
/* Explicit settings for the same two knobs (graph_chunk in [1,64]). */
int msbgx_set_launch_config(msbgx_grid *g, int graph_chunk, int smoother_mask);

/* Learned correction of the preconditioner (no reference counterpart).
 * A 4-layer 3x3x3 convolutional network without biases (1 -> 16 -> 16 -> 16
 * -> 1 channels, ReLU after the first three; BF16 weights and activations,
 * FP32 accumulation) maps the residual the V-cycle leaves on a level-0 16^3
 * block plus its one-voxel halo to an error estimate:
 *   z = V(r) + N(r - A V(r)).
 * Parameters, FP32, in this order: W1[27][16] (tap, out), W2[27][16][16],
 * W3[27][16][16] (tap, in, out), W4[27][16] (tap, in); tap = 9 (dz+1) +
 * 3 (dy+1) + (dx+1).  The fine form needs 16^3 base blocks; the coarse form
 * works with any block size. */
#define MSBGX_NET_PARAMS 14688
typedef struct msbgx_net msbgx_net;
/* Where the network acts (msbgx_net_set_mode; the training set is cleared
 * when the form changes):
 *   MSBGX_NET_FINE    z = V(r) + N(r - A V(r)) on every level-0 16^3 block;
 *   MSBGX_NET_COARSE  a coarse correction: the network runs on the grid of
 *                     block means of r (one cell per block, tiled in 16^3
 *                     tiles with a one-cell halo), its output is extended
 *                     constant over each level-0 PCG block's unknowns and
 *                     scaled by the Galerkin step <zc, r> / <zc, A zc> (zc),
 *                     and z = zc + V(r - A zc).  Trained on block means of
 *                     r_k (input) and of x* - x_k over the unknowns (target).
 * Setting the form also sets its default stall test (fine: 5 iterations,
 * 10 %; coarse: 32 iterations, 2 %). */
#define MSBGX_NET_FINE 0
#define MSBGX_NET_COARSE 1
/* He-normal initial weights from `seed`. */
int msbgx_net_create(msbgx_net **out, uint64_t seed);
void msbgx_net_destroy(msbgx_net *n);
int msbgx_net_get_weights(msbgx_net *n, float *w);        /* MSBGX_NET_PARAMS floats */
int msbgx_net_set_weights(msbgx_net *n, const float *w);  /* also resets the optimiser */
/* Replace the training set by `count` host samples: x [count][18^3] (block
 * plus halo, x fastest), target and mask [count][16^3]. */
int msbgx_net_set_samples(msbgx_net *n, int count, const float *x, const float *target, const uint8_t *mask);
/* Masked mean-square loss over the training set, its gradient (NULL to skip,
 * MSBGX_NET_PARAMS floats) and the network outputs ([count][16^3], NULL to
 * skip), computed by the GPU kernels used in training. */
int msbgx_net_loss_grad(msbgx_net *n, double *loss, float *grad, float *out);
/* Append training samples from the plain MG-PCG on a prepared grid: after
 * every `every`-th iteration of a solve (params p), `blocks_per_pair` random
 * level-0 PCG blocks give (input r - A V(r), target x* - x_k - V(r)) divided
 * by the target's RMS on the block; x* is a prior solve to p->tol.  Returns the number of
 * iterates sampled in *pairs.  The grid's pressure is zero afterwards. */
int msbgx_net_collect(msbgx_net *n, msbgx_grid *g, const msbgx_solve_params *p, int every, int max_pairs,
                      int blocks_per_pair, uint64_t seed, int *pairs);
int msbgx_net_set_mode(msbgx_net *n, int mode);
int msbgx_net_get_mode(const msbgx_net *n);
/* Coarse form: how the network's block values extend over a PCG block's
 * unknowns -- constant (the block's own value, the default), or trilinear
 * between block centres (cell-centred), the weights renormalised over the
 * PCG blocks among the eight nearest centres.  Training is the same for
 * both (targets are block means). */
#define MSBGX_NET_PROLONG_CONSTANT 0
#define MSBGX_NET_PROLONG_LINEAR 1
int msbgx_net_set_prolongation(msbgx_net *n, int kind);
int msbgx_net_get_prolongation(const msbgx_net *n);
/* Coarse form: the number of scales (1 to MSBGX_NET_MAX_SCALES, default 1)
 * the network is applied at -- the dense block grid coarsened 1x, 2x, 4x
 * (means over the PCG blocks of each cell), one direction per scale.  With
 * one scale the correction is scaled by its Galerkin step; with several it
 * is the combination of the directions that minimises the A-norm error
 * over their span (coefficients from their Gram matrix, FP64).  Training
 * samples (msbgx_net_collect) are drawn at every scale; changing the number
 * drops the collected samples. */
#define MSBGX_NET_MAX_SCALES 3
int msbgx_net_set_coarse_scales(msbgx_net *n, int scales);
int msbgx_net_get_coarse_scales(const msbgx_net *n);
/* Stall test of msbgx_solve_pressure_learned: a stall is |r| dropping by
 * less than `drop` (relative) over `window` iterations (0 <= window <= 32;
 * 0 disables the test; defaults per form, see above).  The fine form falls
 * back at its first stall; the coarse form restarts its directions at the
 * first and falls back at the second. */
int msbgx_net_set_stall(msbgx_net *n, int window, float drop);
int msbgx_net_num_samples(const msbgx_net *n);
/* Adam on random minibatches of `batch` samples; first-step and smoothed
 * last-step losses; msbgx_net_last_train_ms gives the GPU time. */
int msbgx_net_train(msbgx_net *n, int steps, int batch, float lr, uint64_t seed, double *loss_first,
                    double *loss_last);
double msbgx_net_last_train_ms(const msbgx_net *n);
/* Per-kernel entry point on the level-0 PCG blocks (MG level 0 channels).
 * Fine form: ch_out += N(ch_in), the halo read from level-0 neighbours
 * (16^3 blocks only).  Coarse form: ch_out += P N(R ch_in), R the mean
 * over each PCG block's cells on the dense block grid (0 elsewhere), N the
 * network over 16^3 tiles of that grid, P the net's prolongation onto the
 * block's unknowns. */
int msbgx_net_apply(msbgx_net *n, msbgx_grid *g, int ch_in, int ch_out);
/* Measurement: mean duration of one forward 16 -> 16 MFMA layer launch over
 * `blocks` patches (HIP events, `reps` launches) and its algorithmic FLOP
 * (blocks * 16^3 cells * 27 * 16 * 16 * 2). */
int msbgx_net_time_conv(msbgx_net *n, int blocks, int reps, double *ms_per_launch, double *flop_per_launch);
/* MG-PCG with the learned correction (flexible beta).  A non-finite network
 * output, a corrected direction with r.z <= 0, or a stalled residual (see
 * msbgx_net_set_stall) turns the correction off for the rest of the solve
 * and restarts the plain MG-PCG; *fallback_iter is the iteration where that
 * happened, -1 if never. */
int msbgx_solve_pressure_learned(msbgx_grid *g, msbgx_net *n, const msbgx_solve_params *p, int *iters,
                                 float *rel_residual, int *fallback_iter);

/* Relative residual after each PCG iteration of the last solve (up to cap);
 * returns the history length (out may be NULL to query it). */
int msbgx_residual_history(const msbgx_grid *g, double *out, int cap);
/* Device time of the last solve in milliseconds (HIP events around the PCG). */
double msbgx_last_solve_ms(const msbgx_grid *g);

/* Inspection.  Block `bid` of channel `ch` at MG level `mg`, at the block's
 * level there: edge^3 floats (uint16 for MSBGX_CH_FLAGS).  Reference access
 * path: getFloatChannel(ch, level, mg)->getBlockDataPtr(bid) (src/msbg.h:2999-3140).
 * Returns the block edge (>0), 0 when the channel has no data there. */
int msbgx_get_channel_block(const msbgx_grid *g, int ch, int mg, int bid, void *dst);
int msbgx_set_channel_block(msbgx_grid *g, int ch, int mg, int bid, const float *src);
/* Per-block level and BLK_* flags at MG level mg (src/msbg.h:558-576, 3560-3590). */
int msbgx_block_info(const msbgx_grid *g, int mg, int *level, uint32_t *flags);
/* Relaxation block list at MG level mg (src/msbg3.cpp:2868-2879); returns its length. */
int msbgx_relax_list(const msbgx_grid *g, int mg, int *out, int cap);
float msbgx_omega(const msbgx_grid *g);

/* Per-kernel entry points (each is one reference operation at MG level mg):
 *   relax        MultiresSparseGrid::relax           src/msbg3.cpp:2245-2369
 *   multiply     multiplyLaplacianMatrixOpt          src/msbg3.cpp:1757-1912
 *   downsample   downsampleChannel<float,float>      src/msbg4.cpp:1513-1915
 *   prolongate   prolongateChannel<float>            src/msbg4.cpp:1921-2029
 *   zero         zeroChannelInPlace                  src/msbg3.cpp:2375-2399
 *   vcycle       z=0; vCycle(0, z, r, q)             src/msbg3.cpp:2412-2465
 * Channels follow the MSBGX_CH_* ids; at MG levels >= 1 only R, Z and Q exist. */
int msbgx_relax(msbgx_grid *g, int mg, int reverse, int n_iter, int ch_x, int ch_b);
int msbgx_multiply(msbgx_grid *g, int mg, int residual, int ch_x, int ch_b, int ch_dst);
int msbgx_downsample(msbgx_grid *g, int mg_coarse, int ch_src, int ch_dst);
int msbgx_prolongate(msbgx_grid *g, int mg_fine, int ch);
int msbgx_zero_channel(msbgx_grid *g, int ch, int mg);
int msbgx_copy_channel(msbgx_grid *g, int ch_src, int ch_dst);
int msbgx_vcycle(msbgx_grid *g, int n_pre, int n_post, int n_coarse);

/* Measurement: runs `reps` launches of the kernel `which` on the handle's
 * stream between HIP events and returns the mean duration per launch in ms
 * and the algorithmic HBM bytes one launch moves.
 *   which 0: red-black colour passes of the smoother over the block size
 *            with the most relaxation cells at MG0, the two colour lists in
 *            turn (reps rounded up to whole pairs); cells per launch are the
 *            mean of the two lists; 30 B per cell with FP32 operands (x read
 *            and written, b, D, W_x, W_y, W_z, flags), 22 B with FP16/BF16
 *            operands (D and W 2 B each)
 *   which 1: residual r = b - A x over the MG0 fluid blocks of the finest
 *            level (30 B per cell)
 *   which 2: PCG update x += a d, r -= a q, |r|^2 (24 B per cell)
 *   which 3: FLIP pressure-gradient velocity update over the level-0 blocks
 *            (per cell flags 2 + velocity write 12, pressure 4 on fluid
 *            cells, velocity and beta 8 per face with a fluid side)
 *   which 4: FLIP divergence right-hand side from the velocity (per cell
 *            flags 2 + b 4, velocity 4 per face of a fluid cell) 
 * The channel contents after the call are unspecified (z, q, x, r, velocity
 * and divergence used). */
int msbgx_time_kernel(msbgx_grid *g, int which, int reps, double *ms_per_launch, double *bytes_per_launch,
                      int *cells_per_launch);

/* FLIP caller steps on the device, either side of the solve in
 * FlipSimulation::pressureProjection (src/flip_sim.cpp:1281-1350).  They need
 * a prepared single grid.  The MAC velocity lives in three device channels
 * over the level-0 blocks -- the reference's CH_VEC3_4 at resolution level 0
 * (src/flip_sim.cpp:363-375): component c of cell (x,y,z) is the velocity on
 * the cell's -c face; blocks of other levels hold no velocity, read as 0
 * (getVC, src/flip_sim.cpp:118-125).  A cell is fluid (isFluid(mass)) iff the
 * cell type sampled at prepare holds neither SOLID nor VOID (flipCellTypeCB,
 * src/flip_sim.cpp:313-321); beta is the sampled face coefficient
 * (flipFaceCoeffCB, src/flip_sim.cpp:323-327).  Host arrays are dense
 * sx0*sy0*sz0 fine grids, x fastest. */
#define MSBGX_FLIP_SIGN_REFERENCE 0  /* v -= dt*beta*(pR - pL) with the solved p, as written in
                                        src/flip_sim.cpp:1317-1349                          */
#define MSBGX_FLIP_SIGN_CORRECTED 1  /* the same update with -p: the MSBG path solves
                                        A p = +div/dt (flipRhsCB), the update assumes the
                                        compact path's A p = -div/dt (src/flip_sim.cpp:936-939) */
/* Velocity channels from host arrays (cells of non-level-0 blocks ignored). */
int msbgx_flip_set_velocity(msbgx_grid *g, const float *u, const float *v, const float *w);
/* Velocity channels from a scene's MAC field (msbgx_scene_mac), on the GPU. */
int msbgx_flip_set_velocity_scene(msbgx_grid *g, const msbgx_scene *scene);
/* Velocity channels to host arrays (0 on cells of non-level-0 blocks). */
int msbgx_flip_get_velocity(msbgx_grid *g, float *u, float *v, float *w);
/* Reference: flipRhsCB (src/flip_sim.cpp:329-338) as sampled by the prepare
 * (src/msbg3.cpp:2551-2579): the divergence channel of the level-0 blocks
 * becomes isFluid ? div(v)/dt : 0, from the device velocity. */
int msbgx_flip_rhs_from_velocity(msbgx_grid *g, float dt);
/* Reference: the pressure-gradient update of FlipSimulation::pressureProjection
 * (src/flip_sim.cpp:1317-1349) with the pressure channel: per face, 0 at
 * index 0 along its axis or when neither side is fluid, else
 * v - dt*beta*(pR - pL) (a non-fluid side's pressure 0, non-finite results 0). */
int msbgx_flip_update_velocity(msbgx_grid *g, float dt, int sign_mode);

/* Particle-to-grid transfer (reference: FlipSimulation::particleToGrid,
 * src/flip_sim.cpp:579-828) on the GPU.  A transfer handle snapshots a grid's
 * raster and level map (call after msbgx_set_refinement_map, before the
 * prepare that will use its outputs).  Outputs live on the level-0 blocks,
 * as the reference's resolution-level-0 grids: cell mass (particle count),
 * face velocities (weighted mean, 0 where the weight sum <= mass_eps), face
 * beta from the face density phi with the liquid/air interface clamp, and
 * the active blocks (a cell with mass > mass_eps).  The reference's sums run
 * in its thread partition's order; here each face gathers its 18 source
 * cells in (z, y, x) order and each cell's particles in index order, so the
 * result is deterministic. */
typedef struct msbgx_particle {   /* layout of FlipParticle (src/flip_sim.h:48) */
  float pos[3];                   /* cell units: cell (i,j,k) spans [i, i+1) x ... */
  float vel[3];
  int32_t phase;                  /* 0 liquid, 1 air */
} msbgx_particle;
typedef struct msbgx_p2g_params { /* FlipConfig fields (src/flip_sim.h:34-42)   */
  float rho_l;                    /* 1000 */
  float rho_g;                    /* 1    */
  float mass_eps;                 /* 1e-5 */
  float beta_min;                 /* 1e-6 */
  float alpha_phi;                /* 1    */
  int32_t particles_per_voxel;    /* 8    */
} msbgx_p2g_params;
typedef struct msbgx_p2g msbgx_p2g;
void msbgx_default_p2g_params(msbgx_p2g_params *p);
int msbgx_p2g_create(msbgx_p2g **out, const msbgx_grid *g);
void msbgx_p2g_destroy(msbgx_p2g *t);
/* Transfer of n particles from a host array (copied to the device first). */
int msbgx_p2g_run(msbgx_p2g *t, int64_t n, const msbgx_particle *particles, const msbgx_p2g_params *p);
/* Transfer of a scene's particles (msbgx_scene_particle for every liquid
 * cell, generated on the GPU); *n_particles receives their count. */
int msbgx_p2g_run_scene(msbgx_p2g *t, const msbgx_scene *scene, const msbgx_p2g_params *p, int64_t *n_particles);
/* Outputs as dense sx0*sy0*sz0 host arrays, x fastest, 0 outside the level-0
 * blocks; any pointer may be NULL. */
int msbgx_p2g_get(msbgx_p2g *t, float *mass, float *u, float *v, float *w, float *beta_u, float *beta_v,
                  float *beta_w);
/* Active blocks (ascending), up to cap; returns their count. */
int msbgx_p2g_active_blocks(msbgx_p2g *t, int *out, int cap);
/* Device time of the last transfer (HIP events): whole transfer and the
 * gather kernel alone, in ms, and the particle count. */
int msbgx_p2g_last_times(const msbgx_p2g *t, double *ms_total, double *ms_gather, int64_t *n_particles);
/* The last transfer's velocity into a prepared grid's FLIP velocity channels
 * (same raster and level map). */
int msbgx_p2g_to_velocity(msbgx_p2g *t, msbgx_grid *g);

/* Compact MIC(0)-PCG (the FLIP caller's "paper baseline" solver kind,
 * src/flip_sim.cpp:858-1151) on the GPU, as a comparator: the system over the
 * fluid cells of the level-0 blocks in (z, y, x) order with diag = sum of the
 * six face coefficients (0 at the domain boundary and towards cells without
 * level-0 data), off-diagonals -beta between fluid cells, rhs = -div/dt (the
 * divergence channel), MIC(0) with sigma = 0.25, PCG from x = 0 to
 * |r| <= tol |b| (FP64 dots, float alpha/beta).  The solution (compact sign,
 * A p = -div/dt) goes to the pressure channel, 0 elsewhere.  Returns the
 * converged iteration, or the iterations run, and |r|/|b|;
 * msbgx_last_solve_ms gives the device time. */
int msbgx_solve_pressure_compact(msbgx_grid *g, int max_iter, float tol, int *iters, float *rel_residual);
/* Reference: buildMIC0Compact + applyMIC0Compact (src/flip_sim.cpp:1027-1097)
 * on the compact system: the factor and z = M^-1 rhs per row; returns the row
 * count (cap = 0 queries it). */
int msbgx_compact_mic0(msbgx_grid *g, float *precon, float *z, int cap);

/* Reference: MultiresSparseGrid::applyChannelPdeFast<RenderDensity>
 * (src/msbg4.cpp:60-960) as the msbg_demo solver entry calls it
 * (src/msbg_demo.cpp:716-727): mean-curvature smoothing of the uint16 render
 * density (RSURF_16_BIT) with the 8-colour in-place block scheme over the
 * listed blocks of level 0, one-voxel halo blocks with Neumann values where
 * no level-0 block holds the voxel (src/halo.cpp:267-480), values clamped to
 * [0, 1], num_iter rounded up to even, time step time_step.  density is the
 * dense sx0*sy0*sz0 channel (x fastest), smoothed in place on the GPU; only
 * voxels of level-0 blocks are read or written.  Needs the level map, not a
 * prepare; block edge <= 32.  *ms (may be NULL) receives the device time. */
int msbgx_pde_mean_curvature(msbgx_grid *g, uint16_t *density, const int *active_blocks, int n_active, int num_iter,
                             float time_step, double *ms);

/* Reference: MultiresSparseGrid::applyChannelPdeFast<Data_T>
 * (src/msbg4.cpp:60-960, declared src/msbg.h:2176-2198) over a density
 * channel, with the in-place multicolour block schemes.  The parameters keep
 * the reference's meaning:
 *   lapl_typ      the reference's encoding: -(type + MSBGX_PDE_OPT_8_COLOR)
 *                 for 8 colour passes, -type for red-black passes (colour
 *                 (bx+by+bz)&1 = 1 first); type 4 mean curvature
 *                 (PDE_MEAN_CURVATURE), 2 bi-Laplacian (two-voxel halo),
 *                 1 Laplacian times |grad|, 3 simple diffusion, 14
 *                 Perona-Malik, 15 its Jacobi form, 17 Jacobi diffusion
 *                 (src/msbg4.cpp:555-800);
 *   do_constr_zero_one, dens_thr, dens_thr_tol1, dens_thr_tol2
 *                 the [0, 1] clamp, or hard constraints around the input
 *                 when dens_thr or dens_thr_tol1 > 1e-12
 *                 (src/msbg4.cpp:806-820);
 *   num_iter      rounded up to even; grid_spacing must be 1;
 *   time_step, taubin (Taubin alternation for types 4 and 1),
 *   anis_diff_alpha (types 14/15; < 0 selects the exponential conductance).
 * Rejected with MSBGX_EINVAL, as undefined or not built: laplTyp >= 0 (the
 * unordered double-buffered scheme), types 1/3/17 at block edges other than
 * 16 (their fixed 18-voxel row stride, src/msbg4.cpp:688-693), red-black
 * passes with types 2 and 4 (their stencils read same-colour diagonal
 * blocks while those are written), type 2 at block edge 32, a uint16
 * channel without a clamp, anis_diff_alpha 0 for types 14/15, negative
 * num_iter or time step.  The reference's lambda, lambdaSDF, redist,
 * velocity-dependent and stochastic-rounding (uint8) options are not built
 * (lambda is forced to 0 there, src/msbg4.cpp:164).
 * dtype MSBGX_PDE_U16: uint16 render density read as v/65535, stored
 * rounded to nearest even; MSBGX_PDE_F32: float.  channel is the dense
 * sx0*sy0*sz0 array (x fastest), smoothed in place on the GPU over the
 * listed level-0 blocks; only voxels of level-0 blocks are read or
 * written.  Needs the level map, not a prepare; block edge <= 32.  *ms
 * (may be NULL) receives the device time. */
#define MSBGX_PDE_OPT_8_COLOR (1 << 29)
#define MSBGX_PDE_U16 0
#define MSBGX_PDE_F32 1
typedef struct msbgx_pde_params {
  int lapl_typ;
  int do_constr_zero_one;
  float dens_thr, dens_thr_tol1, dens_thr_tol2;
  int num_iter;
  float grid_spacing;
  float time_step;
  float taubin;
  float anis_diff_alpha;
} msbgx_pde_params;
int msbgx_pde_apply(msbgx_grid *g, int dtype, void *channel, const int *active_blocks, int n_active,
                    const msbgx_pde_params *params, double *ms);

/* Allocation-free variant of msbgx_pde_mean_curvature.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments (no timing output), same results bit for bit;
 * it calls no hipMalloc/hipFree and creates no stream or event: its device
 * work space is one static array of MSBGX_PDE_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends.  A call needs
 *   2 sx0*sy0*sz0 + 2 (level-0 blocks) bsx0^3 + 12 (blocks) + 8 n_active
 * bytes (each of the seven parts rounded up to 256) of it; larger inputs are
 * rejected with MSBGX_EINVAL.  It runs on the calling thread's per-thread
 * default stream and returns when `density` holds the result.  Calls that
 * overlap on one GPU share the array and race. */
#define MSBGX_PDE_NOALLOC_SCRATCH_BYTES (80u << 20)
int msbgx_pde_mean_curvature_noalloc(msbgx_grid *g, uint16_t *density, const int *active_blocks, int n_active,
                                     int num_iter, float time_step);

/* Caller-side block level policy (src/flip_sim.cpp:480-548) over a scene:
 * 26-neighbourhood block distance d from seed blocks (mode 0: blocks holding
 * liquid; mode 1: blocks holding a liquid cell with an air face neighbour);
 * level 0 for d <= d0, 1 for d == d0+1, else n_levels-1.  Not regularised. */
int msbgx_level_map_from_scene(const msbgx_scene *s, int bsx0, int n_levels, int mode, int d0, int *out);

#ifdef __cplusplus
}
#endif

#endif /* MSBGX_H */
