/*
 * pawan_oracle_q128.c -- __float128 twin of the oracle's kernel and
 * all-pairs right-hand side (libquadmath).  TEST INFRASTRUCTURE ONLY.
 *
 * Same formulas as gaussian.h:20-60 and interaction_utils.h:456-501, but
 * with 113-bit significands, so the rho_bar -> 0 cancellation in q and F
 * (which costs FP64 about log10(15/rho_bar^4) digits) leaves ~25 correct
 * digits.  Used as the "truth" against which both the FP64 oracle and the
 * GPU paths are measured on small systems.
 */
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

#include "pawan_oracle.h"

typedef __float128 qd;

static void kernel_q(qd rho, qd sigma, qd *q, qd *F, qd *Z) {
  const qd pi = M_PIq;
  qd rho_bar = rho / sigma;
  qd sig3 = sigma * sigma * sigma;
  *Z = expq(-0.5Q * rho_bar * rho_bar) / sig3 / powq(2.0Q * pi, 1.5Q);
  qd phi = 0.25Q / pi * erfq(M_SQRT1_2q * rho_bar) / sig3;
  *q = (phi / rho_bar - *Z) / (rho_bar * rho_bar);
  *F = (*Z - 3.0Q * *q) / (rho * rho);
}

void or_kernel_q128(double rho, double sigma, double *q, double *F,
                    double *Z) {
  qd qq, FF, ZZ;
  kernel_q((qd)rho, (qd)sigma, &qq, &FF, &ZZ);
  *q = (double)qq;
  *F = (double)FF;
  *Z = (double)ZZ;
}

/* Rates of the listed targets, all sources, in __float128. */
void or_rates_targets_q128(int nw, const int64_t *off, const double *pos,
                           const double *alpha, const double *radius,
                           const double *volume, double nu,
                           const int64_t *targets, int64_t ntargets,
                           double *vel, double *retvor) {
  const int64_t N = off[nw];
  const qd qnu = (qd)nu;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t k = 0; k < ntargets; ++k) {
    const int64_t t = targets[k];
    qd u[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    const double *rt = pos + 3 * t, *at = alpha + 3 * t;
    for (int64_t s = 0; s < N; ++s) {
      if (s == t) continue;
      const double *rs = pos + 3 * s, *as = alpha + 3 * s;
      qd d[3] = {(qd)rt[0] - rs[0], (qd)rt[1] - rs[1], (qd)rt[2] - rs[2]};
      qd rho = sqrtq(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      qd ss = radius[s], st = radius[t];
      qd sigma = sqrtq(0.5Q * (ss * ss + st * st));
      qd q, F, Z;
      kernel_q(rho, sigma, &q, &F, &Z);
      qd a_s[3] = {as[0], as[1], as[2]}, a_t[3] = {at[0], at[1], at[2]};
      /* u_t += q (a_s x d) */
      u[0] += q * (a_s[1] * d[2] - a_s[2] * d[1]);
      u[1] += q * (a_s[2] * d[0] - a_s[0] * d[2]);
      u[2] += q * (a_s[0] * d[1] - a_s[1] * d[0]);
      /* a_t x a_s */
      qd c[3] = {a_t[1] * a_s[2] - a_t[2] * a_s[1],
                 a_t[2] * a_s[0] - a_t[0] * a_s[2],
                 a_t[0] * a_s[1] - a_t[1] * a_s[0]};
      qd dc = d[0] * c[0] + d[1] * c[1] + d[2] * c[2];
      qd coef = 2.0Q * qnu * Z / (sigma * sigma);
      for (int j = 0; j < 3; ++j)
        w[j] += q * c[j] + F * dc * d[j] +
                coef * ((qd)volume[t] * a_s[j] - (qd)volume[s] * a_t[j]);
    }
    for (int j = 0; j < 3; ++j) {
      vel[3 * k + j] = (double)u[j];
      retvor[3 * k + j] = (double)w[j];
    }
  }
}

void or_solve_q128(int nw, const int64_t *off, const double *pos,
                   const double *alpha, const double *radius,
                   const double *volume, double nu, double *vel,
                   double *retvor) {
  const int64_t N = off[nw];
  int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
  for (int64_t i = 0; i < N; ++i) idx[i] = i;
  or_rates_targets_q128(nw, off, pos, alpha, radius, volume, nu, idx, N, vel,
                        retvor);
  free(idx);
}
