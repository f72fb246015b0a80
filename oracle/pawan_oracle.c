/*
 * pawan_oracle.c -- FP64 CPU restatement of PAWAN's interaction and
 * integration hot path.  TEST INFRASTRUCTURE ONLY (see pawan_oracle.h).
 *
 * Reference paths are relative to the reference repository root.  The
 * arithmetic follows the reference operation by operation (same operand
 * order, no fused multiply-add: built with -ffp-contract=off) so that the
 * per-pair results match what the reference computes; only the GSL
 * containers are replaced by plain arrays.
 *
 * Third-party arithmetic restated here: GSL's cblas dnrm2/ddot/dscal and
 * gsl_vector_add/sub (GSL version unpinned by the reference,
 * CMakeLists.txt:21).  dnrm2 is the scaled sum-of-squares algorithm of the
 * netlib reference BLAS DNRM2, which GSL's cblas implements; ddot is the
 * plain left-to-right sum.  glibc libm supplies erf/exp/pow/sqrt, as in the
 * reference build.
 */
#include "pawan_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_1_PI
#define M_1_PI 0.31830988618379067154
#endif
#ifndef M_SQRT1_2
#define M_SQRT1_2 0.70710678118654752440
#endif

/* ---------------------------------------------------------------------- */
/* GSL BLAS-1 pieces used on the hot path                                  */
/* ---------------------------------------------------------------------- */

/* gsl_blas_dnrm2 on a 3-vector: netlib DNRM2 scaled sum of squares.
 * Used at src/interaction/interaction_utils.h:412,474,545 and
 * src/interaction/interaction.cpp:144.
 * Pinned: tests compare against math.hypot / sqrt(sum x^2) to 1 ulp and
 * against exact Pythagorean triples. */
double or_dnrm2(const double *x) {
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < 3; ++i) {
    if (x[i] != 0.0) {
      const double ax = fabs(x[i]);
      if (scale < ax) {
        ssq = 1.0 + ssq * (scale / ax) * (scale / ax);
        scale = ax;
      } else {
        ssq += (ax / scale) * (ax / scale);
      }
    }
  }
  return scale * sqrt(ssq);
}

/* gsl_blas_ddot: r = 0; r += x[i]*y[i], i = 0,1,2. */
static inline double ddot3(const double *x, const double *y) {
  double r = 0.0;
  r += x[0] * y[0];
  r += x[1] * y[1];
  r += x[2] * y[2];
  return r;
}

/* gsl_cross, src/utils/gsl_utils.h:142-153. */
static inline void cross3(const double *a, const double *b, double *c) {
  const double ax = a[0], ay = a[1], az = a[2];
  const double bx = b[0], by = b[1], bz = b[2];
  c[0] = ay * bz - az * by;
  c[1] = az * bx - ax * bz;
  c[2] = ax * by - ay * bx;
}

/* ---------------------------------------------------------------------- */
/* Gaussian smoothing kernel, src/interaction/gaussian.h                   */
/* ---------------------------------------------------------------------- */

/* ZETASIG, gaussian.h:20-23. */
static inline double zetasig(double rho, double sigma) {
  double rho_bar = rho / sigma;
  return exp(-0.5 * rho_bar * rho_bar) / pow(sigma, 3.0) /
         pow(2.0 * M_PI, 1.5);
}

/* KERNEL, gaussian.h:47-60.
 * Pinned: q against the quadrature (1/rho^3) * int_0^rho Z(r) r^2 dr,
 * F against the finite difference q'(rho)/rho, Z against its normalisation
 * int Z 4 pi r^2 dr = 1, all three against the __float128 twin where no
 * cancellation occurs, and the small/large rho_bar series limits. */
static void gaussian_kernel(double rho, double sigma, double *q, double *F,
                            double *Z, double *n) {
  double rho_bar = rho / sigma;
  double sig3 = sigma * sigma * sigma;
  *Z = zetasig(rho, sigma);
  double phi = 0.25 * M_1_PI * erf(M_SQRT1_2 * rho_bar) / sig3;
  *q = (phi / rho_bar - *Z) / (rho_bar * rho_bar);
  *n = zetasig(rho, sigma);
  *F = (*Z - 3.0 * *q) / (rho * rho);
}

/* QSIG, gaussian.h:30-37 (used by HELICITY). */
static double gaussian_qsig(double rho, double sigma) {
  double rho_bar = rho / sigma;
  double sig3 = sigma * sigma * sigma;
  double Z = zetasig(rho, sigma);
  double phi = 0.25 * M_1_PI * erf(M_SQRT1_2 * rho_bar) / sig3;
  return (phi / rho_bar - Z) / (rho_bar * rho_bar);
}

/* ---------------------------------------------------------------------- */
/* High-order algebraic smoothing kernel, src/interaction/highorder.h      */
/* ---------------------------------------------------------------------- */
/* The reference chooses highorder.h or gaussian.h when it is compiled
 * (HIGHORDER / GAUSSIAN, interaction_utils.h:22-26); the oracle takes the
 * same choice at run time through or_set_smoothing.
 * Pinned (tests/test_oracle_highorder.py): Z by its normalisation
 * int Z 4 pi r^2 dr = 1; q against (1/rho^3) int_0^rho Z(r) r^2 dr by
 * quadrature; F against q'(rho)/rho by finite differences; the PSE kernel
 * eta by the second-moment condition int eta rho_bar^2 dV = 3 that makes
 * (2/sigma^2) int (a_s - a_t) eta the Laplacian; ENSTF's pair factor
 * equals ZETASIG; the enstrophy pair factors tend to the self terms as
 * rho -> 0. */
static int smoothing = OR_GAUSSIAN;

void or_set_smoothing(int kind) { smoothing = kind; }
int or_get_smoothing(void) { return smoothing; }

static inline double cube(double x) { return x * x * x; } /* gsl_pow_3 */
static inline double pow7(double x) {                      /* gsl_pow_7 */
  double x2 = x * x, x4 = x2 * x2;
  return x * x2 * x4;
}

/* ZETASIG, highorder.h:20-24. */
static inline double ho_zetasig(double rho, double sigma) {
  double rho_bar = rho / sigma;
  return 1.875 * M_1_PI / pow(rho_bar * rho_bar + 1.0, 3.5) / pow(sigma, 3);
}

/* QSIG, highorder.h:31-39. */
static double ho_qsig(double rho, double sigma) {
  double rho_bar = rho / sigma;
  double rho_bar2 = rho_bar * rho_bar;
  double rho3 = rho * rho * rho;
  return 0.25 * M_1_PI * rho_bar2 * rho_bar * (rho_bar2 + 2.5) /
         pow(rho_bar2 + 1.0, 2.5) / rho3;
}

/* ETASIG, highorder.h:41-44. */
static inline double ho_etasig(double rho, double sigma) {
  double rho_bar = rho / sigma;
  return 13.125 * M_1_PI / pow(rho_bar * rho_bar + 1.0, 4.5) / pow(sigma, 3);
}

/* KERNEL, highorder.h:56-67. */
static void ho_kernel(double rho, double sigma, double *q, double *F,
                      double *Z, double *n) {
  *Z = ho_zetasig(rho, sigma);
  *q = ho_qsig(rho, sigma);
  *n = ho_etasig(rho, sigma);
  *F = (*Z - 3.0 * *q) / (rho * rho);
}

void or_kernel(double rho, double sigma, double *q, double *F, double *Z,
               double *n) {
  if (smoothing == OR_HIGHORDER)
    ho_kernel(rho, sigma, q, F, Z, n);
  else
    gaussian_kernel(rho, sigma, q, F, Z, n);
}

double or_qsig(double rho, double sigma) {
  return smoothing == OR_HIGHORDER ? ho_qsig(rho, sigma)
                                   : gaussian_qsig(rho, sigma);
}

/* ENST(rho, sigma, F1, F2), highorder.h:77-87; zero for the Gaussian
 * (gaussian.h:69-75).  Pinned: the pair term F1 a1.a2 + F2 (a1.x)(a2.x) is
 * a1.(Z I - grad grad G) a2 with G the Newton potential of the core, so
 * F1 = Z + M/rho^3 and F2 = (Z - 3 M/rho^3)/rho^2 with M the enclosed core,
 * checked by quadrature of Z alone; F1 at rho -> 0 equals ENST(sigma). */
void or_enst_pair(double rho, double sigma, double *F1, double *F2) {
  if (smoothing != OR_HIGHORDER) {
    *F1 = 0.0;
    *F2 = 0.0;
    return;
  }
  double rho2 = rho * rho;
  double sig2 = sigma * sigma;
  double factor = (M_1_PI / 8.0) / sqrt(pow7(rho2 + sig2));
  *F1 = factor * (2.0 * rho2 * rho2 + 7.0 * sig2 * rho2 + 20.0 * sig2 * sig2);
  *F2 = factor * (-6.0 * rho2 * rho2 - 27.0 * sig2 * rho2 - 21.0 * sig2 * sig2) /
        (rho2 + sig2);
}

/* ENSTF(rho, sigma, F1), highorder.h:89-96; zero for the Gaussian
 * (gaussian.h:77-81). */
double or_enstf_pair(double rho, double sigma) {
  if (smoothing != OR_HIGHORDER) return 0.0;
  double rho2 = rho * rho;
  double sig2 = sigma * sigma;
  return (15.0 * M_1_PI * sig2 * sig2 / 8.0) / sqrt(pow7(rho2 + sig2));
}

/* ENST(sigma) and ENSTF(sigma), highorder.h:98-106 and gaussian.h:87-92. */
double or_enst_self(double sigma) {
  return smoothing == OR_HIGHORDER ? 5.0 * M_1_PI / 2.0 / cube(sigma) : 0.0;
}
double or_enstf_self(double sigma) {
  return smoothing == OR_HIGHORDER ? 15.0 * M_1_PI / 8.0 / cube(sigma) : 0.0;
}

/* ---------------------------------------------------------------------- */
/* Pair interaction                                                        */
/* ---------------------------------------------------------------------- */

/* INTERACT, src/interaction/interaction_utils.h:456-501, with VELOCITY
 * (:55-62), VORSTRETCH (:73-105) and DIFFUSION (:344-378) inlined in the
 * same operation order.
 * Pinned: antisymmetry of the vorticity rate, velocity = regularised
 * Biot-Savart (far-field limit), stretching = transposed contraction of the
 * finite-difference velocity gradient, PSE against nu * Laplacian on a
 * lattice, rotation/translation equivariance (tests/test_oracle_pair.py). */
void or_interact(double nu, double s_src, double s_trg, const double *r_src,
                 const double *r_trg, const double *a_src,
                 const double *a_trg, double v_src, double v_trg,
                 double *dr_src, double *dr_trg, double *da_src,
                 double *da_trg) {
  double d[3] = {r_trg[0] - r_src[0], r_trg[1] - r_src[1],
                 r_trg[2] - r_src[2]};
  double rho = or_dnrm2(d);
  double q, F, Z, n;
  double sigma = sqrt(0.5 * (s_src * s_src + s_trg * s_trg));
  or_kernel(rho, sigma, &q, &F, &Z, &n);

  double c[3];
  /* VELOCITY(q, a_src, d): target */
  cross3(a_src, d, c);
  for (int i = 0; i < 3; ++i) dr_trg[i] += q * c[i];
  /* VELOCITY(-q, a_trg, d): source */
  cross3(a_trg, d, c);
  for (int i = 0; i < 3; ++i) dr_src[i] += (-q) * c[i];

  /* VORSTRETCH */
  double txs[3];
  cross3(a_trg, a_src, txs);
  double roaxa = ddot3(d, txs);
  double da[3];
  for (int i = 0; i < 3; ++i) da[i] = q * txs[i] + (F * roaxa) * d[i];
  /* DIFFUSION */
  double sig12 = 0.5 * sigma * sigma;
  double coef = n * nu / sig12;
  for (int i = 0; i < 3; ++i) {
    double dva = v_trg * a_src[i] - v_src * a_trg[i];
    da[i] += coef * dva;
  }
  for (int i = 0; i < 3; ++i) {
    da_trg[i] += da[i];
    da_src[i] -= da[i];
  }
}

/* getVi, src/interaction/interaction.cpp:453-476: the velocity induced at a
 * point r by every particle of every wake, added to vi.  The kernel is
 * evaluated with the source's own radius as sigma (the call at
 * interaction.cpp:465 passes it, with 5 arguments that match no KERNEL
 * overload; the 6-argument KERNEL of gaussian.h:47-60 is used with the same
 * arguments), and VELOCITY (interaction_utils.h:55-62) with
 * d = r - x_particle.  A point that coincides with a particle gets no
 * contribution from it (alpha x d = 0 there; the reference's kernel would
 * give 0/0).
 * Pinned: the singular Biot-Savart law far from the particles
 * (tests/test_oracle_point.py), and, for wakes of one radius, the velocity
 * or_solve gives the particles themselves (there the pair sigma
 * sqrt((s^2 + s^2)/2) is the particle's own). */
void or_velocity_at(int64_t n, const double *pos, const double *alpha, const double *radius,
                    int64_t m, const double *pts, double *u) {
  for (int64_t k = 0; k < m; ++k) {
    double vi[3] = {u[3 * k], u[3 * k + 1], u[3 * k + 2]};
    for (int64_t i = 0; i < n; ++i) {
      double d[3] = {pts[3 * k] - pos[3 * i], pts[3 * k + 1] - pos[3 * i + 1], pts[3 * k + 2] - pos[3 * i + 2]};
      double rho = or_dnrm2(d);
      if (rho == 0.0) continue;
      double q, F, Z, nn;
      or_kernel(rho, radius[i], &q, &F, &Z, &nn);
      double c[3];
      cross3(&alpha[3 * i], d, c);
      for (int j = 0; j < 3; ++j) vi[j] += q * c[j];
    }
    for (int j = 0; j < 3; ++j) u[3 * k + j] = vi[j];
  }
}

/* Target-side half of INTERACT: the contribution of source j to target t.
 * Written with t in the reference's "target" role.  Because every
 * operation in INTERACT is sign-symmetric under swapping the roles
 * (d -> -d, a_trg x a_src -> -(a_src x a_trg)), this equals, bit for bit,
 * the contribution the reference gives particle t when t is in the
 * "source" role of the pair. */
static inline void pair_target(double nu, const double *rt, const double *at,
                               double st, double vt, const double *rs,
                               const double *as, double ss, double vs,
                               double *u, double *w) {
  double d[3] = {rt[0] - rs[0], rt[1] - rs[1], rt[2] - rs[2]};
  double rho = or_dnrm2(d);
  double q, F, Z, n;
  double sigma = sqrt(0.5 * (ss * ss + st * st));
  or_kernel(rho, sigma, &q, &F, &Z, &n);
  double c[3];
  cross3(as, d, c);
  for (int i = 0; i < 3; ++i) u[i] += q * c[i];
  double txs[3];
  cross3(at, as, txs);
  double roaxa = ddot3(d, txs);
  double sig12 = 0.5 * sigma * sigma;
  double coef = n * nu / sig12;
  for (int i = 0; i < 3; ++i) {
    double da = q * txs[i] + (F * roaxa) * d[i];
    da += coef * (vt * as[i] - vs * at[i]);
    w[i] += da;
  }
}

/* ---------------------------------------------------------------------- */
/* Right-hand side                                                         */
/* ---------------------------------------------------------------------- */

/* __interaction::interact(), interaction.cpp:196-210, with the serial
 * interact(W) (:262-281) and interact(W1,W2) (:283-302). */
void or_solve(int nw, const int64_t *off, const double *pos,
              const double *alpha, const double *radius,
              const double *volume, double nu, double *vel, double *retvor) {
  const int64_t N = off[nw];
  memset(vel, 0, sizeof(double) * 3 * N);
  memset(retvor, 0, sizeof(double) * 3 * N);
  for (int w = 0; w < nw; ++w) {
    for (int64_t s = off[w]; s < off[w + 1]; ++s)
      for (int64_t t = s + 1; t < off[w + 1]; ++t)
        or_interact(nu, radius[s], radius[t], pos + 3 * s, pos + 3 * t,
                    alpha + 3 * s, alpha + 3 * t, volume[s], volume[t],
                    vel + 3 * s, vel + 3 * t, retvor + 3 * s, retvor + 3 * t);
  }
  for (int wi = 0; wi < nw; ++wi)
    for (int wj = wi + 1; wj < nw; ++wj)
      for (int64_t s = off[wi]; s < off[wi + 1]; ++s)
        for (int64_t t = off[wj]; t < off[wj + 1]; ++t)
          or_interact(nu, radius[s], radius[t], pos + 3 * s, pos + 3 * t,
                      alpha + 3 * s, alpha + 3 * t, volume[s], volume[t],
                      vel + 3 * s, vel + 3 * t, retvor + 3 * s,
                      retvor + 3 * t);
}

/* The variant of INTERACT that returns the source terms in scalars,
 * interaction_utils.h:523-576 (identical arithmetic to or_interact). */
static inline void interact3(double nu, double s_src, double s_trg,
                             const double *r_src, const double *r_trg,
                             const double *a_src, const double *a_trg,
                             double v_src, double v_trg, double *dr_trg,
                             double *da_trg, double *vs, double *qs) {
  double z_u[3] = {0, 0, 0}, z_w[3] = {0, 0, 0};
  or_interact(nu, s_src, s_trg, r_src, r_trg, a_src, a_trg, v_src, v_trg, z_u,
              dr_trg, z_w, da_trg);
  for (int i = 0; i < 3; ++i) {
    vs[i] = z_u[i];
    qs[i] = z_w[i];
  }
}

/* One source row of __parallel::interact(W), parallel.cpp:15-48. */
static void omp_row(int64_t s, int64_t t0, int64_t t1, const double *pos,
                    const double *alpha, const double *radius,
                    const double *volume, double nu, double *vel,
                    double *retvor) {
  double vx = 0.0, vy = 0.0, vz = 0.0, qx = 0.0, qy = 0.0, qz = 0.0;
#pragma omp parallel for reduction(+ : vx, vy, vz, qx, qy, qz)
  for (int64_t t = t0; t < t1; ++t) {
    double vsrc[3], qsrc[3];
    interact3(nu, radius[s], radius[t], pos + 3 * s, pos + 3 * t,
              alpha + 3 * s, alpha + 3 * t, volume[s], volume[t],
              vel + 3 * t, retvor + 3 * t, vsrc, qsrc);
    vx += vsrc[0];
    vy += vsrc[1];
    vz += vsrc[2];
    qx += qsrc[0];
    qy += qsrc[1];
    qz += qsrc[2];
  }
  vel[3 * s + 0] = vx + vel[3 * s + 0];
  vel[3 * s + 1] = vy + vel[3 * s + 1];
  vel[3 * s + 2] = vz + vel[3 * s + 2];
  retvor[3 * s + 0] = qx + retvor[3 * s + 0];
  retvor[3 * s + 1] = qy + retvor[3 * s + 1];
  retvor[3 * s + 2] = qz + retvor[3 * s + 2];
}

/* __parallel::interact(W) / interact(W1,W2), parallel.cpp:14-86, driven by
 * __interaction::interact(), interaction.cpp:196-210. */
void or_solve_omp(int nw, const int64_t *off, const double *pos,
                  const double *alpha, const double *radius,
                  const double *volume, double nu, double *vel,
                  double *retvor) {
  const int64_t N = off[nw];
  memset(vel, 0, sizeof(double) * 3 * N);
  memset(retvor, 0, sizeof(double) * 3 * N);
  for (int w = 0; w < nw; ++w)
    for (int64_t s = off[w]; s < off[w + 1]; ++s)
      omp_row(s, s + 1, off[w + 1], pos, alpha, radius, volume, nu, vel,
              retvor);
  for (int wi = 0; wi < nw; ++wi)
    for (int wj = wi + 1; wj < nw; ++wj)
      for (int64_t s = off[wi]; s < off[wi + 1]; ++s)
        omp_row(s, off[wj], off[wj + 1], pos, alpha, radius, volume, nu, vel,
                retvor);
}

void or_solve_omp_rows(int64_t n, const double *pos, const double *alpha,
                       const double *radius, const double *volume, double nu,
                       const int64_t *rows, int64_t nrows, double *vel,
                       double *retvor) {
  for (int64_t r = 0; r < nrows; ++r)
    omp_row(rows[r], rows[r] + 1, n, pos, alpha, radius, volume, nu, vel,
            retvor);
}

static void target_rates(int nw, const int64_t *off, const double *pos,
                         const double *alpha, const double *radius,
                         const double *volume, double nu, int64_t t,
                         double *u, double *w) {
  int wt = 0;
  while (!(t >= off[wt] && t < off[wt + 1])) ++wt;
  u[0] = u[1] = u[2] = 0.0;
  w[0] = w[1] = w[2] = 0.0;
  const double *rt = pos + 3 * t, *at = alpha + 3 * t;
  const double st = radius[t], vt = volume[t];
  /* own wake first (interact(W)), then the cross-wake pairs in the order
   * the i<j loop of interaction.cpp:205-209 reaches them */
  for (int64_t s = off[wt]; s < off[wt + 1]; ++s) {
    if (s == t) continue;
    pair_target(nu, rt, at, st, vt, pos + 3 * s, alpha + 3 * s, radius[s],
                volume[s], u, w);
  }
  for (int ws = 0; ws < nw; ++ws) {
    if (ws == wt) continue;
    for (int64_t s = off[ws]; s < off[ws + 1]; ++s)
      pair_target(nu, rt, at, st, vt, pos + 3 * s, alpha + 3 * s, radius[s],
                  volume[s], u, w);
  }
}

void or_rates_targets(int nw, const int64_t *off, const double *pos,
                      const double *alpha, const double *radius,
                      const double *volume, double nu, const int64_t *targets,
                      int64_t ntargets, double *vel, double *retvor,
                      int use_omp) {
  if (use_omp) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t k = 0; k < ntargets; ++k)
      target_rates(nw, off, pos, alpha, radius, volume, nu, targets[k],
                   vel + 3 * k, retvor + 3 * k);
  } else {
    for (int64_t k = 0; k < ntargets; ++k)
      target_rates(nw, off, pos, alpha, radius, volume, nu, targets[k],
                   vel + 3 * k, retvor + 3 * k);
  }
}

/* ---------------------------------------------------------------------- */
/* Time integration                                                        */
/* ---------------------------------------------------------------------- */

/* Rates of the whole particle system: k = [vel ; retvor] for the state
 * y = [pos ; alpha] (the per-wake state layout of wake.cpp:856-893 with the
 * wakes concatenated and no padding rows). */
struct particle_ctx {
  int nw;
  const int64_t *off;
  const double *radius, *volume;
  double nu;
  int use_omp;
  int64_t *idx;
  double *u, *w;
};

static void particle_rhs(int64_t len, const double *y, double *k, void *vctx) {
  struct particle_ctx *c = (struct particle_ctx *)vctx;
  const int64_t N = c->off[c->nw];
  or_rates_targets(c->nw, c->off, y, y + 3 * N, c->radius, c->volume, c->nu,
                   c->idx, N, c->u, c->w, c->use_omp);
  memcpy(k, c->u, sizeof(double) * 3 * N);
  memcpy(k + 3 * N, c->w, sizeof(double) * 3 * N);
}

/* out = k*c + y  (gsl_vector_memcpy(out,k); gsl_vector_scale(out,c);
 * gsl_vector_add(out,y)) */
void or_axpy_stage(int64_t len, double c, const double *k, const double *y,
                   double *out) {
  for (int64_t i = 0; i < len; ++i) out[i] = k[i] * c + y[i];
}

int or_step_generic(int scheme, double dt, int64_t nsteps, int64_t L,
                    double *y, or_rhs_fn f, void *ctx) {
  if (scheme != 1 && scheme != 3 && scheme != 4) return -1;
  size_t bytes = sizeof(double) * (L > 0 ? L : 1);
  double *x1 = (double *)malloc(bytes);
  double *k1 = (double *)malloc(bytes);
  double *k2 = (double *)malloc(bytes);
  double *k3 = (double *)malloc(bytes);
  double *k4 = (double *)malloc(bytes);
  for (int64_t it = 0; it < nsteps; ++it) {
    if (scheme == 1) {
      /* __integration::step, integration.cpp:126-136 */
      f(L, y, k1, ctx);
      for (int64_t i = 0; i < L; ++i) k1[i] = k1[i] * dt;
      for (int64_t i = 0; i < L; ++i) y[i] = y[i] + k1[i];
    } else if (scheme == 4) {
      /* __rk4::step, rk4.cpp:12-77 */
      f(L, y, k1, ctx);
      or_axpy_stage(L, 0.5 * dt, k1, y, x1);
      f(L, x1, k2, ctx);
      or_axpy_stage(L, 0.5 * dt, k2, y, x1);
      f(L, x1, k3, ctx);
      or_axpy_stage(L, dt, k3, y, x1);
      f(L, x1, k4, ctx);
      const double c6 = dt / 6., c3 = dt / 3.;
      for (int64_t i = 0; i < L; ++i) {
        double a = (k1[i] + k4[i]) * c6;
        double b = (k2[i] + k3[i]) * c3;
        y[i] = y[i] + (a + b);
      }
    } else {
      /* __rk3::step, rk3.cpp:12-65 */
      f(L, y, k1, ctx);
      or_axpy_stage(L, 0.5 * dt, k1, y, x1);
      f(L, x1, k2, ctx);
      for (int64_t i = 0; i < L; ++i) {
        double x2 = k2[i] * (2. * dt) + y[i];
        k1[i] = k1[i] * dt;
        x1[i] = x2 - k1[i];
      }
      f(L, x1, k3, ctx);
      const double c6 = dt / 6.0, s6 = 1.0 / 6.0, c23 = dt * 2.0 / 3.0;
      for (int64_t i = 0; i < L; ++i) {
        double a = k1[i] * s6 + k3[i] * c6;
        double b = k2[i] * c23;
        y[i] = y[i] + (a + b);
      }
    }
  }
  free(x1);
  free(k1);
  free(k2);
  free(k3);
  free(k4);
  return 0;
}

int or_step(int scheme, double dt, int64_t nsteps, int nw, const int64_t *off,
            double *pos, double *alpha, const double *radius,
            const double *volume, double nu, int use_omp) {
  const int64_t N = off[nw];
  const int64_t L = 6 * N;
  struct particle_ctx c;
  c.nw = nw;
  c.off = off;
  c.radius = radius;
  c.volume = volume;
  c.nu = nu;
  c.use_omp = use_omp;
  c.idx = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
  c.u = (double *)malloc(sizeof(double) * 3 * (N > 0 ? N : 1));
  c.w = (double *)malloc(sizeof(double) * 3 * (N > 0 ? N : 1));
  for (int64_t i = 0; i < N; ++i) c.idx[i] = i;
  double *y = (double *)malloc(sizeof(double) * (L > 0 ? L : 1));
  memcpy(y, pos, sizeof(double) * 3 * N);
  memcpy(y + 3 * N, alpha, sizeof(double) * 3 * N);
  int rc = or_step_generic(scheme, dt, nsteps, L, y, particle_rhs, &c);
  memcpy(pos, y, sizeof(double) * 3 * N);
  memcpy(alpha, y + 3 * N, sizeof(double) * 3 * N);
  free(y);
  free(c.idx);
  free(c.u);
  free(c.w);
  return rc;
}

/* ---------------------------------------------------------------------- */
/* Diagnostics, interaction.cpp:64-194 and interaction_utils.h:107-331     */
/* ---------------------------------------------------------------------- */

/* HELICITY, interaction_utils.h:214-242 (x1,a1 = particle I; x2,a2 = J). */
static inline double helicity_pair(double s1, double s2, const double *x1,
                                   const double *x2, const double *a1,
                                   const double *a2) {
  double x12[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
  double rho = or_dnrm2(x12);
  double sigma = sqrt(0.5 * (s1 * s1 + s2 * s2));
  double q = or_qsig(rho, sigma);
  double c[3];
  cross3(a1, a2, c);
  return q * ddot3(x12, c);
}

/* KINETICENERGY pair, interaction_utils.h:278-306. */
static inline double ke_pair(double s1, double s2, const double *x1,
                             const double *x2, const double *a1,
                             const double *a2) {
  double x12[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
  double rho2 = ddot3(x12, x12);
  double sigma2 = 0.5 * (s1 * s1 + s2 * s2);
  double a1a2 = ddot3(a1, a2);
  double x12a1 = ddot3(x12, a1);
  double x12a2 = ddot3(x12, a2);
  return ((rho2 + 2.0 * sigma2) * a1a2 + (x12a1 * x12a2)) /
         pow(rho2 + sigma2, 1.5) / 16.0 / M_PI;
}

/* KINETICENERGYF pair, interaction_utils.h:308-331. */
static inline double kef_pair(double s1, double s2, const double *x1,
                              const double *x2, const double *a1,
                              const double *a2) {
  double x12[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
  double rho2 = ddot3(x12, x12);
  double sigma2 = 0.5 * (s1 * s1 + s2 * s2);
  double a1a2 = ddot3(a1, a2);
  return (rho2 + 1.5 * sigma2) * a1a2 / pow(rho2 + sigma2, 1.5) / 8.0 / M_PI;
}

/* KINETICENERGY / KINETICENERGYF self terms, interaction_utils.h:249-267. */
static inline double ke_self(double s, const double *a) {
  return M_1_PI * ddot3(a, a) / s / 8.0;
}
static inline double kef_self(double s, const double *a) {
  return 3 * M_1_PI * ddot3(a, a) / (s * 16.0);
}

/* ENSTROPHY / ENSTROPHYF pair, interaction_utils.h:141-202. */
static inline double enst_pair(double s1, double s2, const double *x1,
                               const double *x2, const double *a1,
                               const double *a2) {
  double x12[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
  double rho = or_dnrm2(x12);
  double sigma = sqrt(0.5 * (s1 * s1 + s2 * s2));
  double F1, F2;
  or_enst_pair(rho, sigma, &F1, &F2);
  double a1a2 = ddot3(a1, a2);
  double a1x12 = ddot3(a1, x12);
  double a2x12 = ddot3(a2, x12);
  return F1 * a1a2 + F2 * a1x12 * a2x12;
}
static inline double enstf_pair(double s1, double s2, const double *x1,
                                const double *x2, const double *a1,
                                const double *a2) {
  double x12[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
  double rho = or_dnrm2(x12);
  double sigma = sqrt(0.5 * (s1 * s1 + s2 * s2));
  return or_enstf_pair(rho, sigma) * ddot3(a1, a2);
}

enum { Q_H = 0, Q_KE = 1, Q_KEF = 2, Q_E = 3, Q_EF = 4 };

/* One pair term of quadratic diagnostic `which` (unscaled). */
static inline double quad_pair(int which, double s1, double s2,
                               const double *x1, const double *x2,
                               const double *a1, const double *a2) {
  switch (which) {
    case Q_H: return helicity_pair(s1, s2, x1, x2, a1, a2);
    case Q_KE: return ke_pair(s1, s2, x1, x2, a1, a2);
    case Q_KEF: return kef_pair(s1, s2, x1, x2, a1, a2);
    case Q_E: return enst_pair(s1, s2, x1, x2, a1, a2);
    default: return enstf_pair(s1, s2, x1, x2, a1, a2);
  }
}

/* Self term of quadratic diagnostic `which` (ENSTROPHY(s, a) etc.,
 * interaction_utils.h:112-130, 249-267); helicity has none. */
static inline double quad_self(int which, double s, const double *a) {
  switch (which) {
    case Q_KE: return ke_self(s, a);
    case Q_KEF: return kef_self(s, a);
    case Q_E: return or_enst_self(s) * ddot3(a, a);
    case Q_EF: return or_enstf_self(s) * ddot3(a, a);
    default: return 0.0;
  }
}

/* Sum over J of one row I of a quadratic diagnostic (the J loop of
 * interaction.cpp:608-718 / parallel.cpp:196-317). */
static double quad_row(int which, int64_t I, int64_t j0, int64_t j1,
                       const double *pos, const double *alpha,
                       const double *radius) {
  double acc = 0.0;
  const double *rI = pos + 3 * I, *aI = alpha + 3 * I;
  for (int64_t J = j0; J < j1; ++J) {
    const double *rJ = pos + 3 * J, *aJ = alpha + 3 * J;
    acc += 2.0 * quad_pair(which, radius[I], radius[J], rI, rJ, aI, aJ);
  }
  return acc;
}

/* calculateHelicity / calculateKineticEnergy(F) for one wake (j1 == -1) or
 * a wake pair.  Serial form (use_omp == 0) accumulates every term straight
 * into the total, as interaction.cpp:608-718; the OpenMP form sums each row
 * first and then adds the row, as parallel.cpp:196-317. */
static double quad_block(int which, int64_t i0, int64_t i1, int64_t j0,
                         int64_t j1, int self, const double *pos,
                         const double *alpha, const double *radius,
                         int use_omp) {
  double tot = 0.0;
  if (!use_omp) {
    for (int64_t I = i0; I < i1; ++I) {
      const double *rI = pos + 3 * I, *aI = alpha + 3 * I;
      if (self && which != Q_H) tot += quad_self(which, radius[I], aI);
      int64_t js = self ? I + 1 : j0, je = self ? i1 : j1;
      for (int64_t J = js; J < je; ++J) {
        const double *rJ = pos + 3 * J, *aJ = alpha + 3 * J;
        tot += 2.0 * quad_pair(which, radius[I], radius[J], rI, rJ, aI, aJ);
      }
    }
    return tot;
  }
  int64_t n = i1 - i0;
  double *row = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t k = 0; k < n; ++k) {
    int64_t I = i0 + k;
    int64_t js = self ? I + 1 : j0, je = self ? i1 : j1;
    row[k] = quad_row(which, I, js, je, pos, alpha, radius);
  }
  for (int64_t k = 0; k < n; ++k) {
    int64_t I = i0 + k;
    if (self && which != Q_H) tot += quad_self(which, radius[I], alpha + 3 * I);
    tot += row[k];
  }
  free(row);
  return tot;
}

static double quad_total(int which, int nw, const int64_t *off,
                         const double *pos, const double *alpha,
                         const double *radius, int use_omp) {
  /* interaction.cpp:93-131: per wake, then each pair i<j */
  double tot = 0.0;
  for (int i = 0; i < nw; ++i) {
    tot += quad_block(which, off[i], off[i + 1], 0, 0, 1, pos, alpha, radius,
                      use_omp);
    for (int j = i + 1; j < nw; ++j)
      tot += quad_block(which, off[i], off[i + 1], off[j], off[j + 1], 0, pos,
                        alpha, radius, use_omp);
  }
  return tot;
}

void or_diagnose(int nw, const int64_t *off, const double *pos,
                 const double *alpha, const double *radius, double zc_prev,
                 double *out, int use_omp) {
  double O[3] = {0, 0, 0}, I[3] = {0, 0, 0}, A[3] = {0, 0, 0};
  for (int w = 0; w < nw; ++w) {
    double o[3] = {0, 0, 0}, li[3] = {0, 0, 0}, ai[3] = {0, 0, 0};
    for (int64_t i = off[w]; i < off[w + 1]; ++i) {
      const double *r = pos + 3 * i, *a = alpha + 3 * i;
      double rxa[3], rxrxa[3];
      /* calculateTotalVorticity, interaction.cpp:489-496 */
      for (int k = 0; k < 3; ++k) o[k] += a[k];
      /* calculateLinearImpulse, :498-510 */
      cross3(r, a, rxa);
      for (int k = 0; k < 3; ++k) li[k] += rxa[k];
      /* calculateAngularImpulse, :512-527 */
      cross3(r, rxa, rxrxa);
      for (int k = 0; k < 3; ++k) ai[k] += rxrxa[k];
    }
    for (int k = 0; k < 3; ++k) {
      li[k] = li[k] * 0.5;
      ai[k] = ai[k] * (1.0 / 3.0);
      O[k] += o[k];
      I[k] += li[k];
      A[k] += ai[k];
    }
  }
  double H = quad_total(Q_H, nw, off, pos, alpha, radius, use_omp);
  /* calculateEnstrophy(F), interaction.cpp:528-606: identically zero for
   * the Gaussian kernel (gaussian.h:69-92) */
  double E = 0.0, EF = 0.0;
  if (smoothing == OR_HIGHORDER) {
    E = quad_total(Q_E, nw, off, pos, alpha, radius, use_omp);
    EF = quad_total(Q_EF, nw, off, pos, alpha, radius, use_omp);
  }
  double KE = quad_total(Q_KE, nw, off, pos, alpha, radius, use_omp);
  double KEF = quad_total(Q_KEF, nw, off, pos, alpha, radius, use_omp);
  /* Centroid, interaction.cpp:135-157: the last wake with a non-zero
   * denominator wins. */
  double Zc = zc_prev;
  for (int w = 0; w < nw; ++w) {
    double num = 0.0, den = 0.0;
    for (int64_t i = off[w]; i < off[w + 1]; ++i) {
      double oy = or_dnrm2(alpha + 3 * i);
      double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
      num += oy * z * (x * x + y * y);
      den += oy * (x * x + y * y);
    }
    if (den != 0.0) Zc = num / den;
  }
  for (int k = 0; k < 3; ++k) {
    out[k] = O[k];
    out[3 + k] = I[k];
    out[6 + k] = A[k];
  }
  out[9] = H;
  out[10] = E;
  out[11] = EF;
  out[12] = KE;
  out[13] = KEF;
  out[14] = Zc;
}

/* ---------------------------------------------------------------------- */
/* Vortex ring initial condition and its strength resolution              */
/* ---------------------------------------------------------------------- */

/* __vring(radius, core, nLayers, nAzi, sig), src/wake/vring.cpp:50-100:
 * nAzi azimuthal stations psi_a = a 2pi/nAzi, each carrying the
 * (2L+1)^2 cells of one cross-section: a centre particle and L layers, the
 * j-th with 8j particles at r_j = r0 (12 j^2 + 1)/(6 j) around the core
 * centre, r0 = 0.35/(2L+1).  Vorticity = strength x (-sin psi, cos psi, 0)
 * with every strength 0 at construction (vring.cpp:25-48, _strength is
 * calloc'd).  Writes pos, alpha, radius, volume of nAzi (2L+1)^2 particles.
 * Pinned (tests/test_oracle_vring.py): positions on circles of radius r_j
 * about the core centre at angle k 2pi/(8j); per-station volume
 * pi 0.35^2 R dpsi (the cell volumes telescope). */
void or_vring(double R, double core, int nLayers, int nAzi, double sig,
              double *pos, double *alpha, double *radius, double *volume) {
  (void)core;
  const double r0 = 0.35 / (2 * nLayers + 1);
  const double dPsi = 2.0 * M_PI / nAzi;
  int64_t index = 0;
  for (int i = 0; i < nAzi; ++i) {
    double psi = i * dPsi;
    double sPsi = sin(psi), cPsi = cos(psi);
    pos[3 * index] = R * cPsi;
    pos[3 * index + 1] = R * sPsi;
    pos[3 * index + 2] = 0.0;
    radius[index] = sig;
    volume[index] = M_PI * (r0 * r0) * R * dPsi;
    index++;
    for (int j = 1; j <= nLayers; ++j) {
      double r = r0 * (12. * (double)(j * j) + 1.0) / (6. * j);
      int nCore = 8 * j;
      double dTheta = 2. * M_PI / nCore;
      for (int k = 0; k < nCore; ++k) {
        double theta = k * dTheta;
        pos[3 * index] = (R + r * cos(theta)) * cPsi;
        pos[3 * index + 1] = (R + r * cos(theta)) * sPsi;
        pos[3 * index + 2] = r * sin(theta);
        radius[index] = sig;
        volume[index] = 2.0 * r0 * dPsi *
                        (2.0 * dTheta * j * R * r0 +
                         (12.0 * (double)(j * j) + 1.0) * (r0 * r0) *
                             (sin(theta + dTheta) - sin(theta)) / 3.0);
        index++;
      }
    }
  }
  for (int64_t p = 0; p < index; ++p) alpha[3 * p] = alpha[3 * p + 1] = alpha[3 * p + 2] = 0.0;
}

/* __vring::getIdealRates, vring.cpp:223-242: the Gaussian core profile at
 * the (2L+1)^2 cells of one cross-section. */
void or_vring_ideal(double R, double core, int nLayers, double *rate) {
  int64_t index = 0;
  double mag0 = 0.5 / (M_PI * core * core);
  double r0 = 0.35 / (2 * nLayers + 1);
  rate[index++] = mag0;
  for (int j = 1; j <= nLayers; ++j) {
    double r = r0 * (12. * (double)(j * j) + 1.0) / (6. * j);
    int nCore = 8 * j;
    double dTheta = 2. * M_PI / nCore;
    double decay = exp(-0.5 * r * r / (core * core));
    for (int k = 0; k < nCore; ++k) {
      double theta = k * dTheta;
      rate[index++] = mag0 * (1.0 + r * cos(theta) / R) * decay;
    }
  }
}

static inline double zeta_any(double rho, double sigma) {
  return smoothing == OR_HIGHORDER ? ho_zetasig(rho, sigma) : zetasig(rho, sigma);
}

/* The influence matrix of __resolve::rebuild (src/resolve/resolve.cpp:12-40)
 * for one ring of m = (2L+1)^2 cells over nAzi stations: column i is
 * getRates (vring.cpp:208-215: |(k_x, k_y)| of the vorticity field at the
 * first station's m particles) after setStates(e_i) and resolve(), i.e.
 * influence() (interaction.cpp:212-243) with INFLUENCE / SELFINFLUENCE
 * (interaction_utils.h:626-703).  The field at particle p collects the
 * terms Z(rho_pq, sigma_pq) a_q in ascending q, the self term
 * ZETASIG(0, s_p) a_p at q = p, as the reference's triangular loop adds
 * them; particles outside cell i carry a = 0 and add exact zeros, so only
 * cell i's particles are visited.  M is m x m, row-major M[p * m + i].
 * Pinned: against an independent numpy evaluation of the smoothed field
 * omega(x_p) = sum_q zeta_sigma(x_p - x_q) a_q (tests/test_oracle_vring.py). */
void or_vring_influence(int nLayers, int nAzi, const double *pos, const double *radius,
                        double *M) {
  const int m = (2 * nLayers + 1) * (2 * nLayers + 1);
  const double dPsi = 2.0 * M_PI / nAzi;
  for (int p = 0; p < m; ++p) {
    for (int i = 0; i < m; ++i) {
      double k[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < nAzi; ++a) {
        int64_t q = (int64_t)a * m + i;
        double psi = a * dPsi;
        double t[3] = {-sin(psi) * 1.0, cos(psi) * 1.0, 0.0};
        double Z;
        if (q == p) {
          Z = zeta_any(0.0, radius[p]);
        } else {
          double d[3] = {pos[3 * q] - pos[3 * p], pos[3 * q + 1] - pos[3 * p + 1],
                         pos[3 * q + 2] - pos[3 * p + 2]};
          double sigma = sqrt(0.5 * (radius[p] * radius[p] + radius[q] * radius[q]));
          Z = zeta_any(or_dnrm2(d), sigma);
        }
        for (int c = 0; c < 3; ++c) k[c] += Z * t[c];
      }
      M[(int64_t)p * m + i] = sqrt(k[0] * k[0] + k[1] * k[1]);
    }
  }
}

/* gsl_linalg_LU_decomp + gsl_linalg_LU_solve (resolve.cpp:45-49): LU with
 * partial pivoting (the row of largest |a_ij| at or below the diagonal,
 * first one on ties), unit lower triangle, then forward and back
 * substitution on the permuted right-hand side.  A (n x n, row-major) is
 * overwritten by its factors.  Returns 1 if a pivot is zero or not finite
 * (GSL flags only zero pivots and carries NaN/Inf into the solution; the
 * device solve reports both, and so does this restatement). 
 * Pinned: numpy.linalg.solve (LAPACK dgesv) on random and influence
 * matrices (tests/test_oracle_vring.py). */
int or_lu_solve(int n, double *A, const double *b, double *x) {
  int *perm = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) perm[i] = i;
  int singular = 0;
  for (int j = 0; j < n; ++j) {
    int ip = j;
    double mx = fabs(A[(int64_t)j * n + j]);
    for (int i = j + 1; i < n; ++i) {
      double v = fabs(A[(int64_t)i * n + j]);
      if (v > mx) { mx = v; ip = i; }
    }
    if (ip != j) {
      for (int c = 0; c < n; ++c) {
        double tmp = A[(int64_t)j * n + c];
        A[(int64_t)j * n + c] = A[(int64_t)ip * n + c];
        A[(int64_t)ip * n + c] = tmp;
      }
      int t = perm[j]; perm[j] = perm[ip]; perm[ip] = t;
    }
    double piv = A[(int64_t)j * n + j];
    /* a zero pivot (GSL's singular case) or a non-finite one (a NaN or Inf
       entry, which GSL would carry into the solution) */
    if (!(fabs(piv) > 0.0) || !isfinite(piv)) { singular = 1; continue; }
    for (int i = j + 1; i < n; ++i) {
      double l = A[(int64_t)i * n + j] / piv;
      A[(int64_t)i * n + j] = l;
      for (int c = j + 1; c < n; ++c) A[(int64_t)i * n + c] -= l * A[(int64_t)j * n + c];
    }
  }
  for (int i = 0; i < n; ++i) x[i] = b[perm[i]];
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < i; ++c) x[i] -= A[(int64_t)i * n + c] * x[c];
  for (int i = n - 1; i >= 0; --i) {
    for (int c = i + 1; c < n; ++c) x[i] -= A[(int64_t)i * n + c] * x[c];
    x[i] /= A[(int64_t)i * n + i];
  }
  free(perm);
  return singular;
}

/* ---------------------------------------------------------------------- */
/* Particle management: __wake::split and __wake::merge                    */
/* ---------------------------------------------------------------------- */

/* src/wake/wake.cpp:61-110.  The loop visits particles from the first while
 * they are older than 45 steps (stepnum - active > 45; the first younger one
 * ends it) and re-reads the particle count, which grows as particles are
 * appended.  A particle splits when |a| >= 2 birth and birth > 1e-4: both
 * halves carry a/2 and birth |a|/2, the new one at x + dx and the old one at
 * x - dx, dx = a (0.25 sigma / |a|), with the old one's radius, volume and
 * age.  The new particle goes to row *npidx; the reference then advances
 * *npidx, wraps it to 0 at MAXNUMPARTICLES (wake.cpp:96-99) and sets the
 * count to it.  Restated here for *npidx == *num_particles and no wrap. */
int64_t or_wake_split(double stepnum, int64_t maxn, int64_t *num_particles, int64_t *npidx,
                      double *pos, double *alpha, double *radius, double *volume,
                      double *birth, double *active) {
  if (*npidx != *num_particles) return -1;
  /* the new particles must fit: count the splits first (a split does not
     change any other particle's test) */
  int64_t want = 0;
  for (int64_t i = 0; i < *num_particles && stepnum - active[i] > 45; ++i) {
    const double m = or_dnrm2(&alpha[3 * i]);
    if (m >= 2 * birth[i] && birth[i] > 0.0001) ++want;
  }
  if (*npidx + want > maxn) return -1;
  int64_t nsplit = 0;
  for (int64_t i = 0; i < *num_particles && stepnum - active[i] > 45; ++i) {
    const double ib = birth[i];
    double *iv = &alpha[3 * i];
    const double im = or_dnrm2(iv);
    if (im >= 2 * ib && ib > 0.0001) {
      double *ip = &pos[3 * i];
      const double ir = radius[i], vol = volume[i];
      double dp[3] = {iv[0], iv[1], iv[2]};
      const double sc = 0.25 * ir / im; /* gsl_blas_dscal(0.25*iradius/ivormag, deltapos) */
      for (int c = 0; c < 3; ++c) dp[c] *= sc;
      const int64_t k = *npidx;
      for (int c = 0; c < 3; ++c) {
        pos[3 * k + c] = ip[c];
        pos[3 * k + c] += dp[c];
        alpha[3 * k + c] = iv[c];
        alpha[3 * k + c] *= 0.5;
      }
      radius[k] = ir;
      volume[k] = vol;
      birth[k] = 0.5 * im;
      active[k] = active[i];
      for (int c = 0; c < 3; ++c) {
        ip[c] -= dp[c];
        iv[c] *= 0.5;
      }
      birth[i] = 0.5 * im;
      ++*npidx;
      *num_particles = *npidx;
      ++nsplit;
    }
  }
  return nsplit;
}

/* src/wake/wake.cpp:112-159.  Outer particles i from the first while older
 * than 45 steps; inner j over all later particles.  |a_i| is taken once per
 * i (the merges below grow a_i but not that norm); a pair merges when
 * |x_i - x_j| < (sigma_i + sigma_j)/8, both strengths are non-zero and
 * 1 - a_i.a_j / |a_i| / |a_j| >= 1e-4: x_i moves to (|a_i| x_i + |a_j| x_j)
 * scaled by 1/(|a_i| + |a_j|), a_i += a_j, and particle j keeps its place
 * with zero strength and volume. */
int64_t or_wake_merge(double stepnum, int64_t n, double *pos, double *alpha,
                      const double *radius, double *volume, const double *active) {
  int64_t nmerge = 0;
  for (int64_t i = 0; i < n && stepnum - active[i] > 45; ++i) {
    double *ip = &pos[3 * i], *iv = &alpha[3 * i];
    const double im = or_dnrm2(iv);
    const double ir = radius[i];
    for (int64_t j = i + 1; j < n; ++j) {
      double *jp = &pos[3 * j], *jv = &alpha[3 * j];
      const double jm = or_dnrm2(jv);
      const double jr = radius[j];
      double d[3] = {ip[0], ip[1], ip[2]};
      for (int c = 0; c < 3; ++c) d[c] -= jp[c];
      const double rho = or_dnrm2(d);
      if (rho < 0.125 * (ir + jr) && jm != 0.0 && im != 0.0) {
        double al = ddot3(iv, jv);
        al = 1 - al / im / jm;
        if (al >= 0.0001) {
          double m[3] = {ip[0], ip[1], ip[2]}, t[3] = {jp[0], jp[1], jp[2]};
          for (int c = 0; c < 3; ++c) {
            m[c] *= im;
            t[c] *= jm;
          }
          for (int c = 0; c < 3; ++c) m[c] += t[c];
          const double sc = 1 / (im + jm);
          for (int c = 0; c < 3; ++c) m[c] *= sc;
          for (int c = 0; c < 3; ++c) {
            ip[c] = m[c];
            iv[c] += jv[c];
            jv[c] = 0.0;
          }
          volume[j] = 0.0;
          ++nmerge;
        }
      }
    }
  }
  return nmerge;
}
