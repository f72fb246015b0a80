/*
 * pawan_gpu.h -- C-ABI of the MI355X (gfx950) implementation of PAWAN's
 * particle-interaction and time-integration hot path.
 *
 * The reference keeps this path in-process C++ (no FFI): the integrators
 * call the abstract solver __system (src/system/system.h:20-90) through
 * setStates / solve / getRates, and __interaction implements it over a list
 * of __wake objects (src/interaction/interaction.h:22-261).  The C++ class
 * API shipped with this library (pawan_amd/csrc/host/pawan/ headers) keeps those
 * class names and virtual shapes and calls the functions below; every entry
 * point names the reference member it replaces.
 *
 * Conventions
 *   - Every function returns an int status: PW_OK (0) or a PW_ERR_* code.
 *     pw_last_error() returns a thread-local message for the last failure.
 *     No C++ exception crosses this boundary.
 *   - Host pointers are borrowed for the duration of the call only.
 *   - Per-particle 3-vectors are N x 3 row-major doubles (the layout of the
 *     reference's gsl_matrix rows, src/wake/wake.h:49-52).
 *   - A context owns its device memory and one HIP stream; work is
 *     asynchronous on that stream and every function that returns host
 *     data synchronises it.  A context is driven by one host thread.
 */
#ifndef PAWAN_GPU_H_
#define PAWAN_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pw_ctx pw_ctx;

/* status codes */
enum {
  PW_OK = 0,
  PW_ERR_ARG = 1,         /* bad argument (null pointer, size, index)      */
  PW_ERR_HIP = 2,         /* a HIP runtime call failed                      */
  PW_ERR_STATE = 3,       /* call not valid in the context's current state  */
  PW_ERR_NONFINITE = 4,   /* a rate or state came out NaN / Inf             */
  PW_ERR_UNSUPPORTED = 5, /* option not available for this configuration    */
  PW_ERR_NOMEM = 6        /* device allocation failed                       */
};

/* pair arithmetic */
enum {
  PW_FP64 = 0,  /* FP64 pair math and accumulation (the plain form)         */
  PW_FP32C = 1, /* FP32 pair math, compensated per-target accumulation,
                   FP64 state, stage combination and diagnostics           */
  PW_LEARNED = 2, /* tree path: direct FP32C near field over Morton-sorted
                     256-particle leaves; far field by the analytic
                     expansion to quadrupole order, or by monopole +
                     dipole plus the learned residual network (pw_ff_*);
                     coupled systems only                                  */
  PW_FP64_REF = 3 /* FP64 with the reference's kernel formula evaluated for
                     every pair, operation by operation (PW_FP64 replaces
                     it by its closed far-field limit where the Gaussian
                     terms are below FP64 rounding, rho > 9.49 sigma)      */
};

/* time integration schemes (the value is the reference's order) */
enum { PW_EULER = 1, PW_RK3 = 3, PW_RK4 = 4 };

/* Smoothing kernel (the reference selects one header when it is compiled:
 * interaction_utils.h:22-26, -DGAUSSIAN or -DHIGHORDER). */
enum {
  PW_KERNEL_GAUSSIAN = 0,  /* gaussian.h:20-92 (default)                     */
  PW_KERNEL_HIGHORDER = 1  /* high-order algebraic, highorder.h:20-106       */
};

/* Library version string. */
const char *pw_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char *pw_last_error(void);

/* Number of visible HIP devices. */
int pw_device_count(int *count);

/* Create a solver context on HIP device `device`.
 * Replaces __interaction::__interaction() (src/interaction/interaction.cpp:
 * 11-22): the reference hard-codes nu = 1.56e-5 there; the context starts
 * with the same value and pw_set_nu changes it. */
int pw_create(pw_ctx **out, int device);

/* Release the context and all of its device memory (~__interaction,
 * interaction.cpp:37-42). */
int pw_destroy(pw_ctx *ctx);

/* Append a wake of n particles (__interaction::addWake,
 * interaction.cpp:44-52).  maxn >= n is the wake's _maxnumParticles, which
 * fixes its block of the state vector: 6*maxn doubles laid out
 * [pos(maxn x 3) ; alpha(maxn x 3)] (src/wake/wake.cpp:856-893).
 * pos, alpha: n x 3; radius, volume: n.  Returns the wake index in *wake
 * (may be NULL).  Wakes can only be added before the first solve/step. */
int pw_add_wake(pw_ctx *ctx, int64_t n, int64_t maxn, const double *pos,
                const double *alpha, const double *radius,
                const double *volume, int *wake);

/* Kinematic viscosity of the PSE term (__interaction::_nu,
 * interaction.h:25; hard-coded in the reference, settable here). */
int pw_set_nu(pw_ctx *ctx, double nu);
int pw_get_nu(pw_ctx *ctx, double *nu);

/* coupled = 1 (default): all wakes interact, as __interaction::interact()
 * does for its wake list (interaction.cpp:196-210).  coupled = 0: every
 * wake is an independent system (the reference equivalent is one
 * __interaction object per wake); they are advanced together, each wake's
 * targets seeing only its own wake's sources. */
int pw_set_coupled(pw_ctx *ctx, int coupled);

/* Smoothing kernel of every later solve, step, diagnosis, point velocity
 * and ring resolve: PW_KERNEL_GAUSSIAN (default) or PW_KERNEL_HIGHORDER
 * (KERNEL, QSIG, ENST, ENSTF of highorder.h:20-106 in place of gaussian.h's;
 * the reference makes this choice at compile time,
 * interaction_utils.h:22-26).  PW_FP64 and PW_FP64_REF both run its FP64
 * form, PW_FP32C its compensated FP32 form; PW_LEARNED runs it in the near
 * field and, in the far field, keeps each node's deviation of the kernel
 * from the singular law, (15/8)(sigma/r)^4 of it at most, within the
 * tolerance (pw_ff_set_params).  May be changed between steps. */
int pw_set_kernel(pw_ctx *ctx, int kernel);
int pw_get_kernel(pw_ctx *ctx, int *kernel);

/* enable = 0: PW_FP32C sums every pair of a target in one plain FP32
 * accumulator, with no chunking and no two-sum (one 256-target block per
 * work item, grid launch, Gaussian kernel only).  A measuring aid that
 * shows what the compensated accumulation buys; not for production.
 * No counterpart in the reference (FP64 only).  Default 1. */
int pw_set_compensation(pw_ctx *ctx, int enable);

/* ---- vortex ring initial condition -------------------------------------
 * pw_vring_particles: the particles of __vring(R, core, nlayers, nazi, sig)
 * (src/wake/vring.cpp:50-100), nazi (2 nlayers + 1)^2 of them, station by
 * station: pos (3 per particle), radius (= sig) and volume.  Their
 * vorticity is zero until resolved (pass zeros to pw_add_wake).
 * nlayers in [0, 15]. */
int pw_vring_particles(double R, double core, int nlayers, int nazi, double sig, double *pos,
                       double *radius, double *volume);
/* pw_vring_resolve: __resolve::rebuild (src/resolve/resolve.cpp:12-63) for
 * a ring held by wake `wake` of the context: the (2L+1)^2 cell strengths
 * whose smoothed vorticity field at the first station's particles matches
 * the Gaussian core profile of __vring::getIdealRates (vring.cpp:223-242),
 * found with the influence matrix built on the device and an LU solve with
 * partial pivoting (gsl_linalg_LU_decomp/solve).  The strengths are written
 * into the wake's vorticity on the device (setStates, vring.cpp:201-205) and
 * returned in strength[(2L+1)^2] if non-null; influence, if non-null,
 * receives the matrix, row-major.  Uses the context's smoothing kernel.
 * PW_ERR_NONFINITE if the matrix is singular (the wake is then unchanged).
 * The .wakeinfluence file the reference writes along the way is not
 * produced. */
int pw_vring_resolve(pw_ctx *ctx, int wake, int nlayers, int nazi, double R, double core,
                     double *strength, double *influence);

/* Total particles and wakes in the context. */
int pw_get_size(pw_ctx *ctx, int64_t *num_particles, int *num_wakes);

/* Per-wake state and rates, n x 3 each (the wake's _position/_vorticity and
 * _velocity/_retvorcity rows, src/wake/wake.h:49-52). */
int pw_set_wake_state(pw_ctx *ctx, int wake, const double *pos,
                      const double *alpha);
/* The volumes of wake `wake` (n doubles), as __wake::_volume after a merge
 * zeroes the absorbed particles' (wake.cpp:147). */
int pw_set_wake_volume(pw_ctx *ctx, int wake, const double *volume);
int pw_get_wake_state(pw_ctx *ctx, int wake, double *pos, double *alpha);
int pw_get_wake_rates(pw_ctx *ctx, int wake, double *vel, double *retvor);

/* Whole-system state / rate vectors in the reference layout: wakes
 * concatenated, wake w occupying 6*maxn_w doubles [pos ; alpha] or
 * [vel ; retvor]; padding rows are read as zero and written as zero.
 * len must equal sum_w 6*maxn_w (__system::_totalmaxsize).
 * Replace __interaction::setStates / getStates / getRates
 * (interaction.cpp:337-365 with wake.cpp:856-893). */
int pw_set_states(pw_ctx *ctx, const double *state, int64_t len);
int pw_get_states(pw_ctx *ctx, double *state, int64_t len);
int pw_get_rates(pw_ctx *ctx, double *rate, int64_t len);

/* Velocity at m points (points: m x 3 AoS, host) induced by every particle
 * of the context (all wakes), FP64, with each particle's own radius as
 * sigma; written to u (m x 3, host).  Replaces __interaction::getVi
 * (src/interaction/interaction.cpp:453-476), which adds the same sum to its
 * output vector for one point (its KERNEL call at :465 takes the source's
 * radius).  A point coinciding with a particle gets no contribution from it.
 * Blocking; allocates its point buffers per call. */
int pw_velocity_at(pw_ctx *ctx, const double *points, int64_t m, double *u);

/* pw_velocity_at for callers that cannot allocate (code under HIP graph
 * capture, tight serving or coupling loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments and results
 * as pw_velocity_at (bit-identical: the same kernel), but the points and
 * velocities are staged through a static device array of the library, so
 * the call makes no hipMalloc / hipFree.  Limits: m <= 65,536 points
 * (PW_ERR_ARG beyond); the context must already hold its particles on the
 * device (its first pw_solve, pw_step or pw_velocity_at puts them there),
 * else PW_ERR_STATE.  Each GPU has one staging array, shared by every
 * context on it: calls that overlap on one GPU (from several host threads)
 * race.  Blocking. */
int pw_velocity_at_noalloc(pw_ctx *ctx, const double *points, int64_t m, double *u);

/* ---- particle management of one wake ----------------------------------
 * __wake::split and __wake::merge (src/wake/wake.cpp:61-110, 112-159; relax
 * is split then merge, :161-165), on the device, for the arrays of one wake
 * as __wake holds them (host pointers, updated in place; rows of 3 for pos
 * and alpha).  Both visit particles from the first while they are older than
 * 45 steps (stepnum - active > 45, the first younger one ends the loop).
 * Results are bit-for-bit the reference's sequential ones.
 *
 * pw_wake_split: a particle with |a| >= 2 birth and birth > 1e-4 becomes two
 * with a/2 and birth |a|/2 each, 0.25 sigma apart along a; the new ones are
 * appended after the last particle, in the order of their parents, and
 * *num_particles and *npidx advance.  Arrays have maxn rows.  The reference
 * writes new particles at _npidx and wraps at MAXNUMPARTICLES, overwriting
 * its first particles: this call requires *npidx == *num_particles and
 * returns PW_ERR_UNSUPPORTED, changing nothing, if *npidx differs or the new
 * particles would not fit in maxn rows.  *nsplit (may be NULL) gets the
 * count.
 *
 * pw_wake_merge: particle i absorbs every later j within (sigma_i +
 * sigma_j)/8 whose strength is not aligned with a_i (1 - cos >= 1e-4):
 * x_i moves to the |a|-weighted mean, a_i += a_j, and j keeps its row with
 * zero strength and volume.  *nmerged (may be NULL) gets the count. */
int pw_wake_split(int device, double stepnum, int64_t maxn, int64_t *num_particles, int64_t *npidx,
                  double *pos, double *alpha, double *radius, double *volume, double *birthstrength,
                  double *active, int64_t *nsplit);
int pw_wake_merge(int device, double stepnum, int64_t num_particles, double *pos, double *alpha,
                  const double *radius, double *volume, const double *active, int64_t *nmerged);

/* One right-hand-side evaluation: zero the rates, then all self and (when
 * coupled) cross-wake pair interactions, leaving velocity and vorticity
 * rate of every particle on the device (__interaction::solve ->
 * interact(), interaction.cpp:54-57, 196-210; __parallel::interact,
 * parallel.cpp:14-86).  precision is PW_FP64 or PW_FP32C. */
int pw_solve(pw_ctx *ctx, int precision);

/* Advance the device-resident state by nsteps steps of dt with scheme
 * PW_EULER (__integration::step, src/integration/integration.cpp:126-136),
 * PW_RK3 (__rk3::step, src/integration/rk3.cpp:12-65) or PW_RK4
 * (__rk4::step, src/integration/rk4.cpp:12-77).  The stage combination is
 * done in FP64 in the reference's operation order.  After the call the
 * rates on the device are those of the last stage evaluated, as in the
 * reference (the wake rates hold k of the final solve()). */
int pw_step(pw_ctx *ctx, int scheme, double dt, int precision,
            int64_t nsteps);

/* ---- per-wake blocks (the class API's overridable hooks) ---------------
 * The reference's __interaction::interact() and influence() visit their
 * wakes block by block through private virtuals that subclasses override
 * (interact(W) / interact(W1, W2), influence(W) / influence(W1, W2),
 * interaction.h:133-157; __parallel does, parallel.h:48-70), and diagnose()
 * calls the protected calculate* virtuals per wake and per wake pair
 * (interaction.h:41-111, interaction.cpp:64-132).  These entry points give
 * the same blocks on the device, so that a subclass can replace some
 * blocks and leave the others to the GPU.  A block is a pair (a, b) of
 * wake indices, a <= b: a == b the pairs within wake a, a < b the pairs
 * between wakes a and b (coupled contexts only).
 *
 * pw_interact_wakes: the rates of the listed blocks only (each block adds
 * what interact(W) or interact(W1, W2) adds to its wakes,
 * interaction.cpp:262-302); particles of wakes in no block get zero.
 * precision PW_FP64, PW_FP64_REF or PW_FP32C.  pairs holds 2 * npairs
 * ints.  The rates are read back with pw_get_wake_rates / pw_get_rates. */
int pw_interact_wakes(pw_ctx *ctx, int precision, const int *pairs, int npairs);

/* pw_vorticity_field: the smoothed vorticity field of the listed blocks
 * (pairs NULL: every block of the system) at every particle, field[3 N]
 * (__interaction::influence(), interaction.cpp:212-260, with
 * INFLUENCE / SELFINFLUENCE, interaction_utils.h:626-703: sum over
 * sources of ZETASIG(rho, sigma_ts) a_s, and ZETASIG(0, s_t) a_t for the
 * particle itself). */
int pw_vorticity_field(pw_ctx *ctx, const int *pairs, int npairs, double *field, int64_t len);

/* pw_diagnose_wakes: the diagnostics of one block, 16 doubles.
 * wa == wb: calculateTotalVorticity (3), calculateLinearImpulse (3),
 * calculateAngularImpulse (3), calculateHelicity, calculateEnstrophy,
 * calculateEnstrophyF, calculateKineticEnergy, calculateKineticEnergyF of
 * wake wa (interaction.cpp:489-718), then the centroid sums
 * sum |a| z (x^2 + y^2) and sum |a| (x^2 + y^2) of the wake
 * (interaction.cpp:135-157).  wa < wb: zeros, then the five
 * calculate*(W1, W2) cross terms, then zeros. */
int pw_diagnose_wakes(pw_ctx *ctx, int wa, int wb, double *out, int64_t out_len);

/* FP64 diagnostics of the current state, 15 doubles in the order
 * __interaction::writediagnosis writes them (interaction.cpp:325-335):
 * total vorticity (3), linear impulse (3), angular impulse (3), helicity,
 * enstrophy, div-free enstrophy, kinetic energy, div-free kinetic energy,
 * centroid Zc (__interaction::diagnose, interaction.cpp:64-194).
 * For an uncoupled context out holds 15 doubles per wake, each wake
 * diagnosed as its own system. */
int pw_diagnose(pw_ctx *ctx, double *out, int64_t out_len);

/* Wait for all queued work of the context. */
int pw_synchronize(pw_ctx *ctx);

/* Kernel timing with HIP events on the context's stream.  While enabled,
 * every launch of the pair-interaction kernel is bracketed by events; the
 * totals accumulate until reset.  pw_get_pair_timing synchronises and
 * returns the summed kernel time in milliseconds and the number of
 * launches. */
int pw_set_profiling(pw_ctx *ctx, int enable);
int pw_get_pair_timing(pw_ctx *ctx, double *total_ms, int64_t *launches);

/* How the pair kernel of `precision` is launched.  parts: source ranges
 * each 256-target block is split into (partial sums reduced per target in
 * FP64, in a fixed order); persistent: resident blocks pull work items from
 * an atomic queue (largest first) instead of one block per item; variant,
 * a combination of
 *   PW_VARIANT_WAVES8       (FP32C) the kernel built for 8 waves per SIMD
 *                           instead of 7 (64 vector registers);
 *   PW_VARIANT_SCALAR_FEED  (FP32C and PW_FP64 with the Gaussian kernel)
 *                           each wave of 64 targets streams the sources
 *                           through scalar loads, chunk by chunk, instead
 *                           of the block's four waves sharing LDS tiles;
 *   PW_VARIANT_TWO_PER_LANE (FP32C with SCALAR_FEED) two targets per lane:
 *                           a wave owns 128 targets and every source record
 *                           it reads feeds two pairs (WAVES8 is ignored);
 * bits that do not apply to the precision are ignored.
 * No counterpart in the reference (its OpenMP loops have no launch shape);
 * defaults are parts 1, grid launch, variant 0.  A variant outside those
 * bits is PW_ERR_ARG. */
enum {
  PW_VARIANT_WAVES8 = 1,
  PW_VARIANT_SCALAR_FEED = 2,
  PW_VARIANT_TWO_PER_LANE = 4
};
int pw_set_pair_config(pw_ctx *ctx, int precision, int parts, int persistent,
                       int variant);
int pw_get_pair_config(pw_ctx *ctx, int precision, int *parts,
                       int *persistent, int *variant);

/* Time the candidate launch configurations of `precision` on the
 * context's current state (one right-hand side each) and keep the fastest;
 * *best_ms (may be NULL) gets its time.  Results are cached per problem
 * shape and smoothing kernel for the process, and in the text file named
 * by $PAWAN_TUNE_CACHE when set.  Leaves the rates of the current state on
 * the device, as pw_solve does. */
int pw_autotune(pw_ctx *ctx, int precision, double *best_ms);

/* enable = 1: pw_step captures one step as a HIP graph (per scheme,
 * precision and dt) and replays it for every step, instead of launching
 * each kernel.  Ignored while kernel profiling is on. */
int pw_set_graph(pw_ctx *ctx, int enable);

/* ---- learned far field (PW_LEARNED) -------------------------------------
 * No counterpart in the reference, which has only the direct sum
 * (interaction.cpp:196-210).  Particles are sorted along a Morton curve
 * into leaves of 256, grouped 8 to a node; for each target leaf a node is
 * far when r_node + r_leaf < theta * d, it lies beyond the Gaussian core,
 * and the self-estimated error of its far evaluation is at most tol.  With
 * use_mlp = 0 far nodes are evaluated by the analytic expansion to
 * quadrupole order, whose error is estimated as coef_md * s3 / (d - r_leaf)^5
 * (s3 bounds sum |alpha| |x - centre|^3); with use_mlp = 1 by monopole +
 * dipole plus the network's residual (the network stands in for the
 * quadrupole and beyond), estimated as coef_mlp * s2 / (d - r_leaf)^4
 * (s2 = sum |alpha| |x - centre|^2).  Every other pair is summed directly
 * (FP32C).  tol <= 0 disables the error test.  With the
 * high-order kernel the error a node's far evaluation is held to also
 * counts (15/8) sigma_max^4 s0 / (4 pi gap^6), the largest deviation of the
 * kernel from the singular law over the node (gap = d - r_node - r_leaf,
 * s0 = sum |alpha|); with tol <= 0 a far node needs
 * gap^2 > 5600 sigma_max^2 instead, where that deviation is below 2^-24. */
int pw_ff_set_params(pw_ctx *ctx, double theta, double tol, int use_mlp);
int pw_ff_get_params(pw_ctx *ctx, double *theta, double *tol, int *use_mlp);
/* enable = 1 (default): remove the mean vorticity rate after every
 * PW_LEARNED right-hand side.  The direct sum conserves total vorticity to
 * rounding because its pair terms are antisymmetric
 * (interaction_utils.h:493-495); one-sided far-field evaluations are not. */
int pw_ff_set_conserve(pw_ctx *ctx, int enable);
int pw_ff_get_conserve(pw_ctx *ctx, int *enable);
/* error coefficients of the analytic expansion (quadrupole order, against
 * s3 / d^5) and of monopole + dipole + network (against s2 / d^4), set by
 * pw_ff_train from held-out samples; may be overridden */
int pw_ff_set_err_coef(pw_ctx *ctx, double coef_md, double coef_mlp);
int pw_ff_get_err_coef(pw_ctx *ctx, double *coef_md, double *coef_mlp);
/* build the tree and the lists of the current state (done by every
 * PW_LEARNED right-hand side; for inspection) */
int pw_ff_build(pw_ctx *ctx);
/* out[0..13): n, leaves, nodes, levels, near capacity, far capacity,
 * training samples held, near pairs, far (target, node) evaluations,
 * list overflow flag, max near leaves, max far nodes, near pairs of leaf
 * pairs within the Gaussian core range, far-list entries (network
 * evaluations, one per target leaf and far node) (of the lists last built;
 * a shorter len takes the leading values) */
int pw_ff_info(pw_ctx *ctx, int64_t *out, int64_t len);
/* Stage times of the PW_LEARNED right-hand sides evaluated while kernel
 * timing is on (pw_set_profiling, which also resets them), from HIP events
 * on the context's stream: out[0] = right-hand sides, out[1..5] = total
 * milliseconds of tree and lists, near field, analytic far field, network
 * residual, and the rest (vorticity conservation, scatter). */
int pw_ff_get_timing(pw_ctx *ctx, double *out, int64_t len);
/* copy an internal array to host memory (bytes must match exactly):
 * PW_FF_PERM int32[n] (sorted position -> particle), PW_FF_NODES
 * double[nodes][24] (centre 3, r, s0, s2, sigma, M 3, D 9 (row-major
 * sum a_i delta_j), first, last, level, child), PW_FF_NEAR_CNT /
 * PW_FF_FAR_CNT int32[leaves], PW_FF_NEAR / PW_FF_FAR int32[leaves][cap],
 * PW_FF_SAMPLES_X float[ns][16], PW_FF_SAMPLES_Y float[ns][12],
 * PW_FF_SAMPLES_AUX double[ns][12] (u scale, gradient scale, |x - c|, s2,
 * |monopole + dipole error of 4 pi u|, |4 pi u|, node radius, node, target
 * (sorted index), target leaf, |error of 4 pi u to quadrupole order|, s3),
 * PW_FF_QUAD double[nodes][20] (quadrupole
 * sum a_p delta_j delta_l for p = x, y, z, each as jl = xx yy zz xy xz yz;
 * then the bounds s1 of sum |a| |delta| and s3 of sum |a| |delta|^3) */
enum {
  PW_FF_PERM = 0, PW_FF_NODES = 1, PW_FF_NEAR_CNT = 2, PW_FF_NEAR = 3,
  PW_FF_FAR_CNT = 4, PW_FF_FAR = 5, PW_FF_SAMPLES_X = 6, PW_FF_SAMPLES_Y = 7,
  PW_FF_SAMPLES_AUX = 8, PW_FF_QUAD = 9
};
int pw_ff_export(pw_ctx *ctx, int what, void *out, int64_t bytes);

/* Training samples from the context's current state: nsamples (target,
 * far node) pairs drawn with the seed (leaf uniform, far node uniform in its
 * list, target uniform in the leaf), the node's exact sum at the target
 * (FP64, the reference's Gaussian kernel) minus the analytic expansion, in
 * the network's scaled units; appended to the context's sample set. */
int pw_ff_sample(pw_ctx *ctx, int64_t nsamples, uint64_t seed);
int pw_ff_clear_samples(pw_ctx *ctx);
/* network parameters (FP32 master copy, 36,748 floats: W1[16][128] b1
 * W2[128][128] b2 W3[128][128] b3 W4[128][12] b4) */
int pw_ff_init_weights(pw_ctx *ctx, uint64_t seed);
int pw_ff_set_weights(pw_ctx *ctx, const float *params, int64_t n);
int pw_ff_get_weights(pw_ctx *ctx, float *params, int64_t n);
/* Adam training on the held samples (the last tenth held out): `steps`
 * minibatches of `batch`, learning rate lr with cosine decay to lr/20.
 * Then the BF16 inference network is packed, evaluated on the held-out
 * samples, and the error coefficients are calibrated (99th percentile of
 * |error| / (s3/d^5) for the analytic expansion, |error| / (s2/d^4) for
 * monopole + dipole + network).  report (may be NULL) gets up to 10 doubles: final
 * train loss, held-out loss, held-out loss of the zero residual (expansion
 * only), rms velocity error ratio network/expansion, the same for the
 * gradient, coef_md, coef_mlp, steps, held-out samples, milliseconds. */
int pw_ff_train(pw_ctx *ctx, int64_t steps, int batch, double lr, uint64_t seed,
                double *report, int64_t report_len);
/* The inference network (BF16 matrix cores) on n scaled inputs: X
 * float[n][16] -> Y float[n][12]. */
int pw_ff_mlp_forward(pw_ctx *ctx, const float *X, int64_t n, float *Y);
/* Loss (1/n) sum (f(x) - y)^2 and its gradient w.r.t. the FP32 master
 * parameters, by the training kernels (FP32). */
int pw_ff_mlp_grad(pw_ctx *ctx, const float *X, const float *Y, int64_t n,
                   float *grad, double *loss);

/* Measurement only (no counterpart in the reference): the vector-issue rate
 * the current device reaches for one instruction form, in wave instructions
 * per CU per nanosecond, from long chains of that form on every CU with
 * waves_per_cu (4..32, a multiple of 4) waves resident.  form 0: FP32 FMA,
 * vector-register operands; 1: FP32 FMA with a scalar-register operand;
 * 2: FP64 FMA, vector-register operands; 3: FP32 multiply.  The ceilings
 * the pair kernels' issue rates are quoted against.  cus (may be NULL) gets
 * the device's CU count. */
int pw_valu_ceiling(int form, int waves_per_cu, double *rate, int *cus);

#ifdef __cplusplus
}
#endif
#endif
