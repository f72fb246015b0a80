/*
 * pawan_oracle.h -- CPU restatement of PAWAN's particle-interaction and
 * time-integration hot path (Gaussian kernel, symmetric pair interaction,
 * self/cross wake loops, Euler/RK3/RK4 steppers, FP64 diagnostics).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.
 * The product (pawan_amd/) never links, loads or calls anything here.
 *
 * Particle sets are passed as one concatenation of all wakes:
 *   off[0..nw]      wake offsets (off[0] = 0, off[nw] = N), int64
 *   pos, alpha      N x 3 row-major (AoS, like the reference's gsl_matrix
 *                   rows, src/wake/wake.h:49-52)
 *   radius, volume  N (src/wake/wake.h:53-54)
 *
 * Pinning: every function is checked in tests/test_oracle_*.py against
 * closed forms, quadrature, invariants or the __float128 twin; see the
 * comment of each function for what pins it.
 */
#ifndef PAWAN_ORACLE_H_
#define PAWAN_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar pieces ---------------------------------------------------- */
double or_dnrm2(const double *x);
void or_kernel(double rho, double sigma, double *q, double *F, double *Z,
               double *n);
double or_qsig(double rho, double sigma);

/* Smoothing kernel of every or_* function below (except the __float128
 * twins, which are Gaussian only): gaussian.h or highorder.h, the choice the
 * reference makes at compile time (interaction_utils.h:22-26).  Global to
 * the process; OR_GAUSSIAN at load. */
enum { OR_GAUSSIAN = 0, OR_HIGHORDER = 1 };
void or_set_smoothing(int kind);
int or_get_smoothing(void);
/* Enstrophy factors: ENST(rho, sigma, F1, F2), ENSTF(rho, sigma, F1),
 * ENST(sigma), ENSTF(sigma) (highorder.h:77-106; zero for the Gaussian). */
void or_enst_pair(double rho, double sigma, double *F1, double *F2);
double or_enstf_pair(double rho, double sigma);
double or_enst_self(double sigma);
double or_enstf_self(double sigma);
void or_kernel_q128(double rho, double sigma, double *q, double *F,
                    double *Z);

/* One symmetric pair, INTERACT (src/interaction/interaction_utils.h:456-501).
 * Accumulates into dr_src, dr_trg, da_src, da_trg (each a 3-vector). */
/* getVi (interaction.cpp:453-476): velocity at m points, added to u (m x 3). */
void or_velocity_at(int64_t n, const double *pos, const double *alpha, const double *radius,
                    int64_t m, const double *pts, double *u);
void or_interact(double nu, double s_src, double s_trg, const double *r_src,
                 const double *r_trg, const double *a_src,
                 const double *a_trg, double v_src, double v_trg,
                 double *dr_src, double *dr_trg, double *da_src,
                 double *da_trg);

/* ---- right-hand side ---------------------------------------------------- */
/* Serial reference loops: zero, self pairs per wake, cross pairs i<j.
 * (src/interaction/interaction.cpp:196-210, :262-302) */
void or_solve(int nw, const int64_t *off, const double *pos,
              const double *alpha, const double *radius,
              const double *volume, double nu, double *vel, double *retvor);

/* OpenMP restatement of src/interaction/parallel.cpp:14-86: serial source
 * loop, inner parallel-for over targets with a reduction of the source
 * terms.  Results depend on the thread count (as in the reference). */
void or_solve_omp(int nw, const int64_t *off, const double *pos,
                  const double *alpha, const double *radius,
                  const double *volume, double nu, double *vel,
                  double *retvor);

/* Same as or_solve_omp but only for the listed source rows of a single
 * wake (nw must be 1): the unit of work the CPU baseline samples. */
void or_solve_omp_rows(int64_t n, const double *pos, const double *alpha,
                       const double *radius, const double *volume, double nu,
                       const int64_t *rows, int64_t nrows, double *vel,
                       double *retvor);

/* Target-side restatement: rates of the listed targets, each summed over
 * sources in the order the serial loops visit them (own wake first, then
 * the other wakes in ascending order).  Bitwise equal to or_solve for
 * those targets; parallel over targets when use_omp != 0. */
void or_rates_targets(int nw, const int64_t *off, const double *pos,
                      const double *alpha, const double *radius,
                      const double *volume, double nu, const int64_t *targets,
                      int64_t ntargets, double *vel, double *retvor,
                      int use_omp);

/* All-pairs right-hand side in __float128 (libquadmath); small N only. */
void or_rates_targets_q128(int nw, const int64_t *off, const double *pos,
                           const double *alpha, const double *radius,
                           const double *volume, double nu,
                           const int64_t *targets, int64_t ntargets,
                           double *vel, double *retvor);
void or_solve_q128(int nw, const int64_t *off, const double *pos,
                   const double *alpha, const double *radius,
                   const double *volume, double nu, double *vel,
                   double *retvor);

/* ---- time integration --------------------------------------------------- */
/* scheme: 1 = Euler (src/integration/integration.cpp:126-136), 3 = RK3
 * (src/integration/rk3.cpp:12-65), 4 = RK4 (src/integration/rk4.cpp:12-77).
 * Advances pos/alpha in place by nsteps steps of dt.  The right-hand side
 * is or_rates_targets over all particles (bitwise equal to or_solve).
 * Returns 0 on success, -1 on a bad scheme. */
int or_step(int scheme, double dt, int64_t nsteps, int nw, const int64_t *off,
            double *pos, double *alpha, const double *radius,
            const double *volume, double nu, int use_omp);

/* The same steppers over any flat state vector y[len] with a caller
 * supplied autonomous right-hand side f(len, y, k, ctx) (the role of
 * __system::setStates/solve/getRates in rk4.cpp:24-56).  Used by the
 * fake linear-system test of the stage arithmetic. */
typedef void (*or_rhs_fn)(int64_t len, const double *y, double *k, void *ctx);
int or_step_generic(int scheme, double dt, int64_t nsteps, int64_t len,
                    double *y, or_rhs_fn f, void *ctx);

/* out = k*c + y, the stage update of rk4.cpp:31-34. */
void or_axpy_stage(int64_t len, double c, const double *k, const double *y,
                   double *out);

/* ---- diagnostics -------------------------------------------------------- */
/* out[15] = Omega(3), I(3), A(3), H, E, EF, KE, KEF, Zc in the order
 * __interaction::writediagnosis writes them
 * (src/interaction/interaction.cpp:325-335).  zc_prev is the value Zc keeps
 * when every wake has a zero denominator (interaction.cpp:153-155). */
void or_diagnose(int nw, const int64_t *off, const double *pos,
                 const double *alpha, const double *radius, double zc_prev,
                 double *out, int use_omp);

/* Vortex ring (vring.cpp:50-100, 223-242) and __resolve::rebuild
 * (resolve.cpp:12-63): particles of nAzi (2L+1)^2, the ideal core profile,
 * the (2L+1)^2 x (2L+1)^2 influence matrix and the pivoted LU solve. */
void or_vring(double R, double core, int nLayers, int nAzi, double sig,
              double *pos, double *alpha, double *radius, double *volume);
void or_vring_ideal(double R, double core, int nLayers, double *rate);
void or_vring_influence(int nLayers, int nAzi, const double *pos, const double *radius,
                        double *M);
int or_lu_solve(int n, double *A, const double *b, double *x);

/* Particle management of one wake (__wake::split, merge, relax,
 * src/wake/wake.cpp:61-165), on arrays of maxn rows (pos, alpha: x 3).
 * or_wake_split appends each new particle at *npidx, which must equal
 * *num_particles (the coupling's bookkeeping; the capacity must not wrap):
 * returns the number of particles split, or -1 (nothing changed) if
 * *npidx != *num_particles or the new particles would not fit in maxn rows.
 * or_wake_merge returns the number of merges. */
int64_t or_wake_split(double stepnum, int64_t maxn, int64_t *num_particles, int64_t *npidx,
                      double *pos, double *alpha, double *radius, double *volume,
                      double *birth, double *active);
int64_t or_wake_merge(double stepnum, int64_t n, double *pos, double *alpha,
                      const double *radius, double *volume, const double *active);

#ifdef __cplusplus
}
#endif
#endif
