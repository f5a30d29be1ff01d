/*
 * fel_oracle.c -- CPU restatement (plain C, FP64, serial) of the per-timestep
 * kernels of toruneko1129/fel-multi-phase-flow.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * Each orc_<name> keeps the argument list of the Fortran subroutine <name>
 * and evaluates every floating-point expression in the reference's order
 * (left-to-right sums, x**2 as x*x, no contraction), so that for routines
 * whose reference source compiles here the results are bit-identical to the
 * compiled reference (oracle/_ref).  Loops run in the reference's k,j,i order
 * on one thread, which fixes the order of the global sums.
 *
 * Pinning status per routine is recorded in the header comment of each
 * function: "pinned: _ref" means bit-compared against the compiled reference
 * source in tests/test_oracle_ref.py; other pins are closed forms or library
 * routines named in the test.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fel_oracle.h"

/* ------------------------------------------------------------------------- */
/* VOF: vertex gradients, MTHINC fluxes, update, mass fix                     */
/* ------------------------------------------------------------------------- */

/* srcs/cal_grad_p2a.f:30-71.  Vertex-centred gradient on [-2,n+2]^3.
 * pinned: _ref, plus the affine-field closed form. */
void orc_cal_grad_p2a(const int *ID, const int *svall, const int *ni_, const int *nj_, const int *nk_,
                      const double *dxinv_, const double *dyinv_array, const double *dzinv_,
                      const double *q, double *qx, double *qy, double *qz) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  (void)svall;
  for (int k = -2; k <= nk + 2; k++)
    for (int j = -2; j <= nj + 2; j++)
      for (int i = -2; i <= ni + 2; i++) {
        const int dis = j + nj * (*ID);
        const double dyinv = dyinv_array[dis + 2];
        const double q000 = q[OIX(i, j, k)], qp00 = q[OIX(i + 1, j, k)];
        const double q0p0 = q[OIX(i, j + 1, k)], qpp0 = q[OIX(i + 1, j + 1, k)];
        const double q00p = q[OIX(i, j, k + 1)], qp0p = q[OIX(i + 1, j, k + 1)];
        const double q0pp = q[OIX(i, j + 1, k + 1)], qppp = q[OIX(i + 1, j + 1, k + 1)];
        double gx = -q000 + qp00 - q0p0 + qpp0 - q00p + qp0p - q0pp + qppp;
        double gy = -q000 - qp00 + q0p0 + qpp0 - q00p - qp0p + q0pp + qppp;
        double gz = -q000 - qp00 - q0p0 - qpp0 + q00p + qp0p + q0pp + qppp;
        qx[OIX(i, j, k)] = gx * 0.25 * dxinv;
        qy[OIX(i, j, k)] = gy * 0.25 * dyinv;
        qz[OIX(i, j, k)] = gz * 0.25 * dzinv;
      }
}

/* srcs/solphi_mthinc1.f:1-292 (+ mthinc_x.h, mthinc_tanh.h).  Face fluxes of
 * the MTHINC scheme, already multiplied by dt/dx.
 * pinned: _ref; symmetric d=0 case and the uniform-field upwind closed form. */
void orc_solphi_mthinc1(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                        const double *dyinv_, const double *dzinv_, const double *dt_,
                        const double *bet_mthinc, const double *phi, const double *phix,
                        const double *phiy, const double *phiz, const double *u, const double *v,
                        const double *w, double *flphix, double *flphiy, double *flphiz) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_, dt = *dt_;
  ODIMS(ni, nj, nk);
  const double dx = 1.0 / dxinv, dy = 1.0 / dyinv, dz = 1.0 / dzinv;
  for (int k = -1; k <= nk + 1; k++)
    for (int j = -1; j <= nj + 1; j++)
      for (int i = -1; i <= ni + 1; i++) {
        flphix[OIX(i, j, k)] = 0.0;
        flphiy[OIX(i, j, k)] = 0.0;
        flphiz[OIX(i, j, k)] = 0.0;
      }
  const double bet = *bet_mthinc;
  const double gp1 = -(0.5 / sqrt(3.0));
  const double gp2 = 0.5 / sqrt(3.0);
  const double dcellinv = 1.0 / sqrt((dx * dx + dy * dy + dz * dz) / 3.0);
  const double ch = cosh(-(bet * 1.5));
  const double eps_mthinc = bet / (2.0 * (ch * ch));
  const double eps_trunc = 0.5 * (tanh(-(bet * 2.5)) + 1.0);
  double xs[8];

  for (int k = 0; k <= nk + 1; k++)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        double rnx, rny, rnz;
        cell_normal(sx, sy, phix, phiy, phiz, i, j, k, &rnx, &rny, &rnz);
        const double q00 = sqrt(rnx * rnx + rny * rny + rnz * rnz);
        if (q00 <= eps_mthinc * dcellinv) {
          double phi00 = phi[OIX(i, j, k)];
          if (phi00 <= eps_trunc) phi00 = 0.0;
          if (1.0 - phi00 <= eps_trunc) phi00 = 1.0;
          const double um = u[OIX(i - 1, j, k)], up = u[OIX(i, j, k)];
          const double vm = v[OIX(i, j - 1, k)], vp = v[OIX(i, j, k)];
          const double wm = w[OIX(i, j, k - 1)], wp = w[OIX(i, j, k)];
          if (um <= 0.0) flphix[OIX(i - 1, j, k)] = um * phi00 * dt * dxinv;
          if (up >= 0.0) flphix[OIX(i, j, k)] = up * phi00 * dt * dxinv;
          if (vm <= 0.0) flphiy[OIX(i, j - 1, k)] = vm * phi00 * dt * dyinv;
          if (vp >= 0.0) flphiy[OIX(i, j, k)] = vp * phi00 * dt * dyinv;
          if (wm <= 0.0) flphiz[OIX(i, j, k - 1)] = wm * phi00 * dt * dzinv;
          if (wp >= 0.0) flphiz[OIX(i, j, k)] = wp * phi00 * dt * dzinv;
        } else {
          const double coef = 1.0 / q00;
          rnx = rnx * coef;
          rny = rny * coef;
          rnz = rnz * coef;
          const double phic = phi[OIX(i, j, k)];
          double ddd = 0.0;
          for (int inr = 1; inr <= 10; inr++) {
            const double wt = mthinc_tanh_sum(rnx, rny, rnz, gp1, gp2, gp1, gp2, gp1, gp2, 0.125,
                                              bet, ddd, xs);
            const double f = wt - phic;
            double s = 0.0;
            for (int m = 0; m < 8; m++) {
              const double c = cosh(bet * (xs[m] + ddd));
              const double t = 0.125 / (c * c);
              s = (m == 0) ? t : s + t;
            }
            const double df = 0.5 * bet * s;
            ddd = ddd - f * df / (df * df + 1.0e-99);
          }
          const double um = u[OIX(i - 1, j, k)], up = u[OIX(i, j, k)];
          const double vm = v[OIX(i, j - 1, k)], vp = v[OIX(i, j, k)];
          const double wm = w[OIX(i, j, k - 1)], wp = w[OIX(i, j, k)];
          if (um < 0.0) {
            const double vd = um * dt * dxinv, st = -0.5, en = -0.5 - vd;
            const double g1 = (st + en) * 0.5 + (-st + en) * gp1;
            const double g2 = (st + en) * 0.5 + (-st + en) * gp2;
            flphix[OIX(i - 1, j, k)] =
                mthinc_tanh_sum(rnx, rny, rnz, g1, g2, gp1, gp2, gp1, gp2, 0.125 * vd, bet, ddd, xs);
          }
          if (up > 0.0) {
            const double vd = up * dt * dxinv, st = 0.5 - vd, en = 0.5;
            const double g1 = (st + en) * 0.5 + (-st + en) * gp1;
            const double g2 = (st + en) * 0.5 + (-st + en) * gp2;
            flphix[OIX(i, j, k)] =
                mthinc_tanh_sum(rnx, rny, rnz, g1, g2, gp1, gp2, gp1, gp2, 0.125 * vd, bet, ddd, xs);
          }
          if (vm < 0.0) {
            const double vd = vm * dt * dyinv, st = -0.5, en = -0.5 - vd;
            const double g1 = (st + en) * 0.5 + (-st + en) * gp1;
            const double g2 = (st + en) * 0.5 + (-st + en) * gp2;
            flphiy[OIX(i, j - 1, k)] =
                mthinc_tanh_sum(rnx, rny, rnz, gp1, gp2, g1, g2, gp1, gp2, 0.125 * vd, bet, ddd, xs);
          }
          if (vp > 0.0) {
            const double vd = vp * dt * dyinv, st = 0.5 - vd, en = 0.5;
            const double g1 = (st + en) * 0.5 + (-st + en) * gp1;
            const double g2 = (st + en) * 0.5 + (-st + en) * gp2;
            flphiy[OIX(i, j, k)] =
                mthinc_tanh_sum(rnx, rny, rnz, gp1, gp2, g1, g2, gp1, gp2, 0.125 * vd, bet, ddd, xs);
          }
          if (wm < 0.0) {
            const double vd = wm * dt * dzinv, st = -0.5, en = -0.5 - vd;
            const double g1 = (st + en) * 0.5 + (-st + en) * gp1;
            const double g2 = (st + en) * 0.5 + (-st + en) * gp2;
            flphiz[OIX(i, j, k - 1)] =
                mthinc_tanh_sum(rnx, rny, rnz, gp1, gp2, gp1, gp2, g1, g2, 0.125 * vd, bet, ddd, xs);
          }
          if (wp > 0.0) {
            const double vd = wp * dt * dzinv, st = 0.5 - vd, en = 0.5;
            const double g1 = (st + en) * 0.5 + (-st + en) * gp1;
            const double g2 = (st + en) * 0.5 + (-st + en) * gp2;
            flphiz[OIX(i, j, k)] =
                mthinc_tanh_sum(rnx, rny, rnz, gp1, gp2, gp1, gp2, g1, g2, 0.125 * vd, bet, ddd, xs);
          }
        }
      }
}

/* srcs/solphi_mthinc2.f:1-48.  Conservative update plus clamp to [0,1].
 * pinned: _ref, and the telescoping-sum invariant. */
void orc_solphi_mthinc2(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                        const double *dyinv_, const double *dzinv_, const double *dt_,
                        const double *u, const double *v, const double *w, const double *flphix,
                        const double *flphiy, const double *flphiz, const double *phi,
                        double *phin) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_, dt = *dt_;
  ODIMS(ni, nj, nk);
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const double div00 = +(-u[OIX(i - 1, j, k)] + u[OIX(i, j, k)]) * dxinv +
                             (-v[OIX(i, j - 1, k)] + v[OIX(i, j, k)]) * dyinv +
                             (-w[OIX(i, j, k - 1)] + w[OIX(i, j, k)]) * dzinv;
        double p = phi[OIX(i, j, k)] * (1.0 + dt * div00) + flphix[OIX(i - 1, j, k)] -
                   flphix[OIX(i, j, k)] + flphiy[OIX(i, j - 1, k)] - flphiy[OIX(i, j, k)] +
                   flphiz[OIX(i, j, k - 1)] - flphiz[OIX(i, j, k)];
        p = fmin(p, 1.0);
        p = fmax(p, 0.0);
        phin[OIX(i, j, k)] = p;
      }
}

/* srcs/solphi_mthinc3.f:1-163.  Truncation of nearly flat cells, clamp, and
 * the global mass correction phin += offset*phin^2(1-phin).
 * Unbuildable reference (needs the MPI header); pinned by the volume
 * invariant |sum(phin)-sum(phi)| <= 1e-12*sum(phi) and a brute-force
 * truncation check in tests/test_oracle_math.py. */
void orc_solphi_mthinc3(const int *ipara, const int *ni_, const int *nj_, const int *nk_,
                        const double *dxinv_, const double *dyinv_, const double *dzinv_,
                        const double *bet_mthinc, const double *phix, const double *phiy,
                        const double *phiz, const double *phi, double *phin) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)ipara;
  const double dx = 1.0 / *dxinv_, dy = 1.0 / *dyinv_, dz = 1.0 / *dzinv_;
  const double bet = *bet_mthinc;
  const double dcellinv = 1.0 / sqrt((dx * dx + dy * dy + dz * dz) / 3.0);
  const double ch = cosh(-(bet * 1.5));
  const double eps_mthinc = bet / (2.0 * (ch * ch));
  const double eps_trunc = 0.5 * (tanh(-(bet * 2.5)) + 1.0);
  for (int k = 0; k <= nk + 1; k++)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        double rnx, rny, rnz;
        cell_normal(sx, sy, phix, phiy, phiz, i, j, k, &rnx, &rny, &rnz);
        const double q00 = sqrt(rnx * rnx + rny * rny + rnz * rnz);
        if (q00 <= eps_mthinc * dcellinv) {
          if (phin[OIX(i, j, k)] <= eps_trunc) phin[OIX(i, j, k)] = 0.0;
          if (1.0 - phin[OIX(i, j, k)] <= eps_trunc) phin[OIX(i, j, k)] = 1.0;
        }
      }
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        double p = fmin(phin[OIX(i, j, k)], 1.0);
        phin[OIX(i, j, k)] = fmax(p, 0.0);
      }
  double sum1 = 0.0, sum2 = 0.0, den2 = 0.0;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const double p = phin[OIX(i, j, k)];
        sum1 = sum1 + phi[OIX(i, j, k)];
        sum2 = sum2 + p;
        den2 = den2 + p * p * (1.0 - p);
      }
  const double offset = (sum1 - sum2) / (den2 + 1.0e-99);
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        double p = phin[OIX(i, j, k)];
        p = p + offset * (p * p) * (1.0 - p);
        p = fmin(p, 1.0);
        phin[OIX(i, j, k)] = fmax(p, 0.0);
      }
}

/* srcs/summation.f:1-54 (nbub slices, slice 0 = sum clamped in the interior).
 * pinned: _ref. */
void orc_summation(const int *ni_, const int *nj_, const int *nk_, double *qk, const int *nbub_) {
  const int ni = *ni_, nj = *nj_, nk = *nk_, nbub = *nbub_;
  ODIMS(ni, nj, nk);
  for (ptrdiff_t m = 0; m < n3; m++) qk[m] = qk[n3 + m];
  for (int l = 2; l <= nbub; l++)
    for (ptrdiff_t m = 0; m < n3; m++) qk[m] = qk[m] + qk[(ptrdiff_t)l * n3 + m];
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) qk[OIX(i, j, k)] = fmin(fmax(qk[OIX(i, j, k)], 0.0), 1.0);
}

/* ------------------------------------------------------------------------- */
/* Momentum advection: WENO5 in derivative form (srcs/weno5.h:1-71)          */
/* ------------------------------------------------------------------------- */

static inline double weno5(double dfm5, double dfm3, double dfm1, double dfp1, double dfp3,
                           double dfp5, double vel00) {
  const double verysmall = 1.0e-80, d1_12 = 1.0 / 12.0;
  const double dfm_1 = dfm5 * 2.0 - dfm3 * 7.0 + dfm1 * 11.0;
  const double dfm_2 = -dfm3 + dfm1 * 5.0 + dfp1 * 2.0;
  const double dfm_3 = dfm1 * 2.0 + dfp1 * 5.0 - dfp3;
  const double dfp_1 = dfp5 * 2.0 - dfp3 * 7.0 + dfp1 * 11.0;
  const double dfp_2 = -dfp3 + dfp1 * 5.0 + dfm1 * 2.0;
  const double dfp_3 = dfp1 * 2.0 + dfm1 * 5.0 - dfm3;
  double a, b;
  a = dfm5 - dfm3 * 2.0 + dfm1;
  b = dfm5 - dfm3 * 4.0 + dfm1 * 3.0;
  const double sm_1 = a * a * 13.0 + b * b * 3.0;
  a = dfm3 - dfm1 * 2.0 + dfp1;
  b = dfm3 - dfp1;
  const double sm_2 = a * a * 13.0 + b * b * 3.0;
  a = dfm1 - dfp1 * 2.0 + dfp3;
  b = dfm1 * 3.0 - dfp1 * 4.0 + dfp3;
  const double sm_3 = a * a * 13.0 + b * b * 3.0;
  a = dfp5 - dfp3 * 2.0 + dfp1;
  b = dfp5 - dfp3 * 4.0 + dfp1 * 3.0;
  const double sp_1 = a * a * 13.0 + b * b * 3.0;
  a = dfp3 - dfp1 * 2.0 + dfm1;
  b = dfp3 - dfm1;
  const double sp_2 = a * a * 13.0 + b * b * 3.0;
  a = dfp1 - dfm1 * 2.0 + dfm3;
  b = dfp1 * 3.0 - dfm1 * 4.0 + dfm3;
  const double sp_3 = a * a * 13.0 + b * b * 3.0;
  double am_1 = sm_2 * sm_3, am_2 = sm_1 * sm_3, am_3 = sm_1 * sm_2;
  double ap_1 = sp_2 * sp_3, ap_2 = sp_1 * sp_3, ap_3 = sp_1 * sp_2;
  am_1 = am_1 * am_1;
  am_2 = am_2 * am_2;
  am_3 = am_3 * am_3;
  ap_1 = ap_1 * ap_1;
  ap_2 = ap_2 * ap_2;
  ap_3 = ap_3 * ap_3;
  am_2 = am_2 * 6.0;
  am_3 = am_3 * 3.0;
  ap_2 = ap_2 * 6.0;
  ap_3 = ap_3 * 3.0;
  const double deninv_am = 1.0 / (am_1 + am_2 + am_3 + verysmall);
  const double deninv_ap = 1.0 / (ap_1 + ap_2 + ap_3 + verysmall);
  const double wgm_1 = am_1 * deninv_am, wgm_3 = am_3 * deninv_am;
  const double wgp_1 = ap_1 * deninv_ap, wgp_3 = ap_3 * deninv_ap;
  const double wgm_2 = 1.0 - wgm_1 - wgm_3;
  const double wgp_2 = 1.0 - wgp_1 - wgp_3;
  const double velab = fabs(vel00);
  const double velm1 = vel00 + velab, velp1 = vel00 - velab;
  return (velm1 * (wgm_1 * dfm_1 + wgm_2 * dfm_2 + wgm_3 * dfm_3) +
          velp1 * (wgp_1 * dfp_1 + wgp_2 * dfp_2 + wgp_3 * dfp_3)) *
         d1_12;
}

/* One WENO5 derivative of field f along unit stride st (srcs/cal_advu_weno5.f:86-92). */
static inline double weno_dir(const double *f, ptrdiff_t c, ptrdiff_t st, double vel00) {
  const double dfm5 = -f[c - 3 * st] + f[c - 2 * st];
  const double dfm3 = -f[c - 2 * st] + f[c - st];
  const double dfm1 = -f[c - st] + f[c];
  const double dfp1 = -f[c] + f[c + st];
  const double dfp3 = -f[c + st] + f[c + 2 * st];
  const double dfp5 = -f[c + 2 * st] + f[c + 3 * st];
  return weno5(dfm5, dfm3, dfm1, dfp1, dfp3, dfp5, vel00);
}

/* srcs/cal_advu_weno5.f:1-257.  pinned: _ref; linear-field exactness. */
void orc_cal_advu_weno5(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                        const double *dyinv_, const double *dzinv_, const double *u,
                        const double *v, const double *w, double *adv_u, double *adv_v,
                        double *adv_w) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  memset(adv_u, 0, sizeof(double) * n3);
  memset(adv_v, 0, sizeof(double) * n3);
  memset(adv_w, 0, sizeof(double) * n3);
  const ptrdiff_t SX = 1, SY = sx, SZ = sx * sy;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        double a, cm, cp;
        cm = (u[c - SX] + u[c]) * 0.5;
        cp = (u[c] + u[c + SX]) * 0.5;
        a = weno_dir(u, c, SX, (cm + cp) * 0.5) * dxinv;
        cm = (v[c - SY] + v[c + SX - SY]) * 0.5;
        cp = (v[c] + v[c + SX]) * 0.5;
        a = a + weno_dir(u, c, SY, (cm + cp) * 0.5) * dyinv;
        cm = (w[c - SZ] + w[c + SX - SZ]) * 0.5;
        cp = (w[c] + w[c + SX]) * 0.5;
        a = a + weno_dir(u, c, SZ, (cm + cp) * 0.5) * dzinv;
        adv_u[c] = a;
      }
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        double a, cm, cp;
        cm = (u[c - SX] + u[c - SX + SY]) * 0.5;
        cp = (u[c] + u[c + SY]) * 0.5;
        a = weno_dir(v, c, SX, (cm + cp) * 0.5) * dxinv;
        cm = (v[c - SY] + v[c]) * 0.5;
        cp = (v[c] + v[c + SY]) * 0.5;
        a = a + weno_dir(v, c, SY, (cm + cp) * 0.5) * dyinv;
        cm = (w[c - SZ] + w[c + SY - SZ]) * 0.5;
        cp = (w[c] + w[c + SY]) * 0.5;
        a = a + weno_dir(v, c, SZ, (cm + cp) * 0.5) * dzinv;
        adv_v[c] = a;
      }
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        double a, cm, cp;
        cm = (u[c - SX] + u[c - SX + SZ]) * 0.5;
        cp = (u[c] + u[c + SZ]) * 0.5;
        a = weno_dir(w, c, SX, (cm + cp) * 0.5) * dxinv;
        cm = (v[c - SY] + v[c - SY + SZ]) * 0.5;
        cp = (v[c] + v[c + SZ]) * 0.5;
        a = a + weno_dir(w, c, SY, (cm + cp) * 0.5) * dyinv;
        cm = (w[c - SZ] + w[c]) * 0.5;
        cp = (w[c] + w[c + SZ]) * 0.5;
        a = a + weno_dir(w, c, SZ, (cm + cp) * 0.5) * dzinv;
        adv_w[c] = a;
      }
}

/* ------------------------------------------------------------------------- */
/* Properties, pressure predictor, explicit viscous terms                    */
/* ------------------------------------------------------------------------- */

/* srcs/cal_arith_mean.f:1-29 (phi=1 takes the "g" value).  pinned: _ref. */
void orc_cal_arith_mean(const int *ni_, const int *nj_, const int *nk_, const double *phik,
                        const double *ql_, const double *qg_, double *qk) {
  ODIMS(*ni_, *nj_, *nk_);
  const double ql = *ql_, qg = *qg_;
  for (ptrdiff_t m = 0; m < n3; m++) {
    const double pl = 1.0 - phik[m], pg = phik[m];
    qk[m] = ql * pl + qg * pg;
  }
}

/* srcs/calphat.f:1-28.  pinned: _ref. */
void orc_calphat(const int *ni_, const int *nj_, const int *nk_, const double *p, const double *po,
                 double *phat) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  for (int k = -2; k <= nk + 3; k++)
    for (int j = -2; j <= nj + 3; j++)
      for (int i = -1; i <= ni + 3; i++) phat[OIX(i, j, k)] = p[OIX(i, j, k)] * 2.0 - po[OIX(i, j, k)];
}

/* srcs/cal_grad_prs.f:1-42.  pinned: _ref. */
void orc_cal_grad_prs(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                      const double *dyinv_, const double *dzinv_, const double *q, double *qx_u,
                      double *qy_v, double *qz_w) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  for (int k = -2; k <= nk + 2; k++)
    for (int j = -2; j <= nj + 2; j++)
      for (int i = -2; i <= ni + 2; i++) {
        const double q0 = q[OIX(i, j, k)];
        qx_u[OIX(i, j, k)] = (-q0 + q[OIX(i + 1, j, k)]) * dxinv;
        qy_v[OIX(i, j, k)] = (-q0 + q[OIX(i, j + 1, k)]) * dyinv;
        qz_w[OIX(i, j, k)] = (-q0 + q[OIX(i, j, k + 1)]) * dzinv;
      }
}

/* srcs/calsij.f:1-143, strain rate s(…,1..6).  pinned: _ref. */
void orc_calsij(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                const double *dyinv_, const double *dzinv_, const double *uk, const double *vk,
                const double *wk, double *sk) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  double *s1 = sk, *s2 = sk + n3, *s3 = sk + 2 * n3, *s4 = sk + 3 * n3, *s5 = sk + 4 * n3,
         *s6 = sk + 5 * n3;
  for (int k = -2; k <= nk + 3; k++)
    for (int j = -2; j <= nj + 3; j++)
      for (int i = -1; i <= ni + 3; i++)
        s1[OIX(i, j, k)] = (-uk[OIX(i - 1, j, k)] + uk[OIX(i, j, k)]) * dxinv;
  for (int k = -2; k <= nk + 3; k++)
    for (int j = -1; j <= nj + 3; j++)
      for (int i = -2; i <= ni + 3; i++)
        s2[OIX(i, j, k)] = (-vk[OIX(i, j - 1, k)] + vk[OIX(i, j, k)]) * dyinv;
  for (int k = -1; k <= nk + 3; k++)
    for (int j = -1; j <= nj + 3; j++)
      for (int i = -1; i <= ni + 3; i++)
        s3[OIX(i, j, k)] = (-wk[OIX(i, j, k - 1)] + wk[OIX(i, j, k)]) * dzinv;
  for (int k = -2; k <= nk + 3; k++)
    for (int j = -2; j <= nj + 2; j++)
      for (int i = -2; i <= ni + 2; i++)
        s4[OIX(i, j, k)] = (+(-uk[OIX(i, j, k)] + uk[OIX(i, j + 1, k)]) * dyinv +
                            (-vk[OIX(i, j, k)] + vk[OIX(i + 1, j, k)]) * dxinv) *
                           0.5;
  for (int k = -2; k <= nk + 2; k++)
    for (int j = -2; j <= nj + 3; j++)
      for (int i = -2; i <= ni + 2; i++)
        s5[OIX(i, j, k)] = (+(-uk[OIX(i, j, k)] + uk[OIX(i, j, k + 1)]) * dzinv +
                            (-wk[OIX(i, j, k)] + wk[OIX(i + 1, j, k)]) * dxinv) *
                           0.5;
  for (int k = -2; k <= nk + 2; k++)
    for (int j = -2; j <= nj + 2; j++)
      for (int i = -2; i <= ni + 3; i++)
        s6[OIX(i, j, k)] = ((-vk[OIX(i, j, k)] + vk[OIX(i, j, k + 1)]) * dzinv +
                            (-wk[OIX(i, j, k)] + wk[OIX(i, j + 1, k)]) * dyinv) *
                           0.5;
}

/* srcs/cal_arith_tau.f:1-121, tau = 2 mu S with 4-point mu on shear faces.
 * pinned: _ref. */
void orc_cal_arith_tau(const int *ni_, const int *nj_, const int *nk_, const double *s,
                       const double *rmu, double *tau) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  for (int l = 0; l < 3; l++)
    for (ptrdiff_t m = 0; m < n3; m++) tau[l * n3 + m] = s[l * n3 + m] * rmu[m] * 2.0;
  for (int k = -2; k <= nk + 3; k++)
    for (int j = -2; j <= nj + 2; j++)
      for (int i = -2; i <= ni + 2; i++) {
        const double m = +0.25 * rmu[OIX(i, j, k)] + 0.25 * rmu[OIX(i + 1, j, k)] +
                         0.25 * rmu[OIX(i, j + 1, k)] + 0.25 * rmu[OIX(i + 1, j + 1, k)];
        tau[3 * n3 + OIX(i, j, k)] = s[3 * n3 + OIX(i, j, k)] * m * 2.0;
      }
  for (int k = -2; k <= nk + 2; k++)
    for (int j = -2; j <= nj + 3; j++)
      for (int i = -2; i <= ni + 2; i++) {
        const double m = +0.25 * rmu[OIX(i, j, k)] + 0.25 * rmu[OIX(i + 1, j, k)] +
                         0.25 * rmu[OIX(i, j, k + 1)] + 0.25 * rmu[OIX(i + 1, j, k + 1)];
        tau[4 * n3 + OIX(i, j, k)] = s[4 * n3 + OIX(i, j, k)] * m * 2.0;
      }
  for (int k = -2; k <= nk + 2; k++)
    for (int j = -2; j <= nj + 2; j++)
      for (int i = -2; i <= ni + 3; i++) {
        const double m = +0.25 * rmu[OIX(i, j, k)] + 0.25 * rmu[OIX(i, j + 1, k)] +
                         0.25 * rmu[OIX(i, j, k + 1)] + 0.25 * rmu[OIX(i, j + 1, k + 1)];
        tau[5 * n3 + OIX(i, j, k)] = s[5 * n3 + OIX(i, j, k)] * m * 2.0;
      }
}

/* srcs/cal_div_tensor.f:1-90, vis = div(tau) at u, v, w points.
 * pinned: _ref. */
void orc_cal_div_tensor(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                        const double *dyinv_, const double *dzinv_, const double *tauk,
                        double *uk, double *vk, double *wk) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  const double *t1 = tauk, *t2 = tauk + n3, *t3 = tauk + 2 * n3, *t4 = tauk + 3 * n3,
               *t5 = tauk + 4 * n3, *t6 = tauk + 5 * n3;
  memset(uk, 0, sizeof(double) * n3);
  memset(vk, 0, sizeof(double) * n3);
  memset(wk, 0, sizeof(double) * n3);
  for (int k = -1; k <= nk + 3; k++)
    for (int j = -1; j <= nj + 3; j++)
      for (int i = -2; i <= ni + 2; i++)
        uk[OIX(i, j, k)] = +(-t1[OIX(i, j, k)] + t1[OIX(i + 1, j, k)]) * dxinv +
                           (-t4[OIX(i, j - 1, k)] + t4[OIX(i, j, k)]) * dyinv +
                           (-t5[OIX(i, j, k - 1)] + t5[OIX(i, j, k)]) * dzinv;
  for (int k = -1; k <= nk + 3; k++)
    for (int j = -2; j <= nj + 2; j++)
      for (int i = -1; i <= ni + 3; i++)
        vk[OIX(i, j, k)] = +(-t4[OIX(i - 1, j, k)] + t4[OIX(i, j, k)]) * dxinv +
                           (-t2[OIX(i, j, k)] + t2[OIX(i, j + 1, k)]) * dyinv +
                           (-t6[OIX(i, j, k - 1)] + t6[OIX(i, j, k)]) * dzinv;
  for (int k = -2; k <= nk + 2; k++)
    for (int j = -1; j <= nj + 3; j++)
      for (int i = -1; i <= ni + 3; i++)
        wk[OIX(i, j, k)] = +(-t5[OIX(i - 1, j, k)] + t5[OIX(i, j, k)]) * dxinv +
                           (-t6[OIX(i, j - 1, k)] + t6[OIX(i, j, k)]) * dyinv +
                           (-t3[OIX(i, j, k)] + t3[OIX(i, j, k + 1)]) * dzinv;
}

/* ------------------------------------------------------------------------- */
/* CSF surface tension                                                       */
/* ------------------------------------------------------------------------- */

/* srcs/init_q.f:1-29.  pinned: trivially (zero fill). */
void orc_init_q(const int *ni_, const int *nj_, const int *nk_, double *q1, double *q2, double *q3) {
  ODIMS(*ni_, *nj_, *nk_);
  memset(q1, 0, sizeof(double) * n3);
  memset(q2, 0, sizeof(double) * n3);
  memset(q3, 0, sizeof(double) * n3);
}

/* srcs/cal_grad_p2uvw.f:1-47, face gradients at u, v, w points.
 * pinned: _ref. */
void orc_cal_grad_p2uvw(const int *ID, const int *svall, const int *ni_, const int *nj_,
                        const int *nk_, const double *dxinv_, const double *dyinv_array,
                        const double *dzinv_, const double *q, double *qx_u, double *qy_v,
                        double *qz_w) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  (void)svall;
  for (int k = -2; k <= nk + 2; k++)
    for (int j = -2; j <= nj + 2; j++)
      for (int i = -2; i <= ni + 2; i++) {
        const double dyinv = dyinv_array[j + nj * (*ID) + 2];
        const double q0 = q[OIX(i, j, k)];
        qx_u[OIX(i, j, k)] = (-q0 + q[OIX(i + 1, j, k)]) * dxinv;
        qy_v[OIX(i, j, k)] = (-q0 + q[OIX(i, j + 1, k)]) * dyinv;
        qz_w[OIX(i, j, k)] = (-q0 + q[OIX(i, j, k + 1)]) * dzinv;
      }
}

/* srcs/calrn.f:1-95, unit normal averaged from vertex gradients onto the
 * staggering (ip,jp,kp).  pinned: _ref. */
void orc_calrn(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
               const double *dyinv_, const double *dzinv_, const int *ip_, const int *jp_,
               const int *kp_, const double *phix, const double *phiy, const double *phiz,
               double *rn) {
  const int ni = *ni_, nj = *nj_, nk = *nk_, ip = *ip_, jp = *jp_, kp = *kp_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  const double verysmall = sqrt(dxinv * dxinv + dyinv * dyinv + dzinv * dzinv) * 1.0e-30;
  memset(rn, 0, sizeof(double) * 3 * n3);
  const double *a[3] = {phix, phiy, phiz};
  for (int k = -2; k <= nk + 3 - kp; k++)
    for (int j = -2; j <= nj + 3 - jp; j++)
      for (int i = -2; i <= ni + 3 - ip; i++) {
        const int ii = i + ip, jj = j + jp, kk = k + kp;
        double r[3];
        for (int c = 0; c < 3; c++) {
          const double *p = a[c];
          r[c] = (+p[OIX(i, j, k)] + p[OIX(ii, j, k)] + p[OIX(i, jj, k)] + p[OIX(ii, jj, k)] +
                  p[OIX(i, j, kk)] + p[OIX(ii, j, kk)] + p[OIX(i, jj, kk)] + p[OIX(ii, jj, kk)]) *
                 0.125;
        }
        const double arn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + verysmall;
        const double inv = 1.0 / arn;
        rn[OIX(ii, jj, kk)] = r[0] * inv;
        rn[n3 + OIX(ii, jj, kk)] = r[1] * inv;
        rn[2 * n3 + OIX(ii, jj, kk)] = r[2] * inv;
      }
}

/* srcs/calrkap.f:1-102, curvature as divergence of the normals at the
 * u, v and w points.  pinned: _ref. */
void orc_calrkap(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                 const double *dyinv_, const double *dzinv_, const double *rn_pp,
                 const double *rn_12, const double *rn_13, const double *rn_23, double *rkap) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  memset(rkap, 0, sizeof(double) * 3 * n3);
  const ptrdiff_t C2 = n3, C3 = 2 * n3;
  for (int k = -1; k <= nk + 3; k++)
    for (int j = -1; j <= nj + 3; j++)
      for (int i = -2; i <= ni + 2; i++)
        rkap[OIX(i, j, k)] = +(-rn_pp[OIX(i, j, k)] + rn_pp[OIX(i + 1, j, k)]) * dxinv +
                             (-rn_12[C2 + OIX(i, j - 1, k)] + rn_12[C2 + OIX(i, j, k)]) * dyinv +
                             (-rn_13[C3 + OIX(i, j, k - 1)] + rn_13[C3 + OIX(i, j, k)]) * dzinv;
  for (int k = -1; k <= nk + 3; k++)
    for (int j = -2; j <= nj + 2; j++)
      for (int i = -1; i <= ni + 3; i++)
        rkap[C2 + OIX(i, j, k)] = +(-rn_12[OIX(i - 1, j, k)] + rn_12[OIX(i, j, k)]) * dxinv +
                                  (-rn_pp[C2 + OIX(i, j, k)] + rn_pp[C2 + OIX(i, j + 1, k)]) * dyinv +
                                  (-rn_23[C3 + OIX(i, j, k - 1)] + rn_23[C3 + OIX(i, j, k)]) * dzinv;
  for (int k = -2; k <= nk + 2; k++)
    for (int j = -1; j <= nj + 3; j++)
      for (int i = -1; i <= ni + 3; i++)
        rkap[C3 + OIX(i, j, k)] = +(-rn_13[OIX(i - 1, j, k)] + rn_13[OIX(i, j, k)]) * dxinv +
                                  (-rn_23[C2 + OIX(i, j - 1, k)] + rn_23[C2 + OIX(i, j, k)]) * dyinv +
                                  (-rn_pp[C3 + OIX(i, j, k)] + rn_pp[C3 + OIX(i, j, k + 1)]) * dzinv;
}

/* srcs/calfst.f:1-50, CSF force -2 sigma/(rho_l+rho_g) kappa grad(phi).
 * pinned: _ref. */
void orc_calfst(const int *ni_, const int *nj_, const int *nk_, const double *rhol,
                const double *rhog, const double *surface_tension, const double *phix_u,
                const double *phiy_v, const double *phiz_w, const double *rkap, double *fst_u,
                double *fst_v, double *fst_w) {
  ODIMS(*ni_, *nj_, *nk_);
  const double st = 2.0 * (*surface_tension) / (*rhol + *rhog);
  for (ptrdiff_t m = 0; m < n3; m++) {
    fst_u[m] = -(st * rkap[m] * phix_u[m]);
    fst_v[m] = -(st * rkap[n3 + m] * phiy_v[m]);
    fst_w[m] = -(st * rkap[2 * n3 + m] * phiz_w[m]);
  }
}

/* srcs/sum_fst.f:1-37.  pinned: _ref. */
void orc_sum_fst(const int *ni_, const int *nj_, const int *nk_, const double *fu,
                 const double *fv, const double *fw, double *su, double *sv, double *sw) {
  ODIMS(*ni_, *nj_, *nk_);
  for (ptrdiff_t m = 0; m < n3; m++) {
    su[m] = su[m] + fu[m];
    sv[m] = sv[m] + fv[m];
    sw[m] = sw[m] + fw[m];
  }
}

/* srcs/cal_av.f:1-39, interior mean.  Unbuildable reference (MPI header);
 * pinned against numpy.mean in tests/test_oracle_math.py. */
void orc_cal_av(const int *ipara, const int *ndiv, const int *ni_, const int *nj_, const int *nk_,
                const double *q, double *av) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)ipara;
  double sum = 0.0;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) sum = sum + q[OIX(i, j, k)];
  *av = sum / (double)(ni * nj * nk * (*ndiv));
}

/* srcs/bnd_avzero.f:1-27.  pinned via orc_cal_av. */
void orc_bnd_avzero(const int *ipara, const int *ndiv, const int *ni_, const int *nj_,
                    const int *nk_, double *q) {
  double av;
  orc_cal_av(ipara, ndiv, ni_, nj_, nk_, q, &av);
  ODIMS(*ni_, *nj_, *nk_);
  for (ptrdiff_t m = 0; m < n3; m++) q[m] = q[m] - av;
}

/* srcs/cal_srcu.f:1-141, explicit right-hand side of the momentum step
 * (AB2 advection, pressure predictor, CN half of viscosity, CSF, gravity).
 * pinned: _ref. */
void orc_cal_srcu(const int *nID, const int *ni_, const int *nj_, const int *nk_,
                  const double *rho_av, const double *dt_, const double *grvb_,
                  const double *grvp_, const double *u, const double *v, const double *w,
                  const double *rho, const double *rhon, const double *adv_u, const double *adv_v,
                  const double *adv_w, const double *adv_uo, const double *adv_vo,
                  const double *adv_wo, const double *prs_u, const double *prs_v,
                  const double *prs_w, const double *vis_u, const double *vis_v,
                  const double *vis_w, const double *fst_u, const double *fst_v,
                  const double *fst_w, const double *fst_un, const double *fst_vn,
                  const double *fst_wn, double *src_u, double *src_v, double *src_w) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)rho_av;
  const int jen = (nID[ORC_Y_PLUS - 1] < 0) ? nj - 1 : nj;
  const double dtinv = 1.0 / (*dt_), grvb = *grvb_, grvp = *grvp_;
  const ptrdiff_t SX = 1, SY = sx, SZ = sx * sy;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        const double r = (rho[c] + rho[c + SX] + rhon[c] + rhon[c + SX]) * 0.25;
        src_u[c] = +u[c] * dtinv - adv_u[c] * 1.5 + adv_uo[c] * 0.5 - prs_u[c] / r +
                   vis_u[c] / r * 0.5 + fst_u[c] * 0.5 + fst_un[c] * 0.5 + grvb;
      }
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= jen; j++)
      for (int i = 1; i <= ni; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        const double r = (rho[c] + rho[c + SY] + rhon[c] + rhon[c + SY]) * 0.25;
        src_v[c] = +v[c] * dtinv - adv_v[c] * 1.5 + adv_vo[c] * 0.5 - prs_v[c] / r +
                   vis_v[c] / r * 0.5 + fst_v[c] * 0.5 + fst_vn[c] * 0.5 + grvp;
      }
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        const double r = (rho[c] + rho[c + SZ] + rhon[c] + rhon[c + SZ]) * 0.25;
        src_w[c] = +w[c] * dtinv - adv_w[c] * 1.5 + adv_wo[c] * 0.5 - prs_w[c] / r +
                   vis_w[c] / r * 0.5 + fst_w[c] * 0.5 + fst_wn[c] * 0.5;
      }
}

/* ------------------------------------------------------------------------- */
/* Semi-implicit viscous step                                                */
/* ------------------------------------------------------------------------- */

/* srcs/cal_arith_coef_vis.f:1-411 with the 45 arrays gathered in a[45].
 * pinned: _ref (through the argument-list wrapper below). */
static void coef_vis(int ni, int nj, int nk, double dxinv, double dyinv, double dzinv,
                     const double *rmuk, double *const a[NCOEF]) {
  ODIMS(ni, nj, nk);
  for (int c = 0; c < NCOEF; c++) memset(a[c], 0, sizeof(double) * n3);
  const double dx2 = dxinv * dxinv, dy2 = dyinv * dyinv, dz2 = dzinv * dzinv;
  const double dxy = dxinv * dyinv, dxz = dxinv * dzinv, dyz = dyinv * dzinv;
#define R(i, j, k) rmuk[OIX(i, j, k)]
  for (int k = -1; k <= nk + 2; k++)
    for (int j = -1; j <= nj + 2; j++)
      for (int i = -1; i <= ni + 2; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        const double xm = R(i, j, k), xp = R(i + 1, j, k);
        const double ym = 0.25 * (+R(i, j - 1, k) + R(i + 1, j - 1, k) + R(i, j, k) + R(i + 1, j, k));
        const double yp = 0.25 * (+R(i, j, k) + R(i + 1, j, k) + R(i, j + 1, k) + R(i + 1, j + 1, k));
        const double zm = 0.25 * (+R(i, j, k - 1) + R(i + 1, j, k - 1) + R(i, j, k) + R(i + 1, j, k));
        const double zp = 0.25 * (+R(i, j, k) + R(i + 1, j, k) + R(i, j, k + 1) + R(i + 1, j, k + 1));
        const double b = dz2 * zm, s = dy2 * ym, w = dx2 * xm * 2.0, e = dx2 * xp * 2.0;
        const double n = dy2 * yp, t = dz2 * zp;
        a[CU_B][c] = b; a[CU_S][c] = s; a[CU_W][c] = w; a[CU_E][c] = e; a[CU_N][c] = n; a[CU_T][c] = t;
        a[CU_VWS][c] = dxy * ym; a[CU_VES][c] = -(dxy * ym);
        a[CU_VWN][c] = -(dxy * yp); a[CU_VEN][c] = dxy * yp;
        a[CU_WWB][c] = dxz * zm; a[CU_WEB][c] = -(dxz * zm);
        a[CU_WWT][c] = -(dxz * zp); a[CU_WET][c] = dxz * zp;
        a[CU_P][c] = +b + s + w + e + n + t;
      }
  for (int k = -1; k <= nk + 2; k++)
    for (int j = -1; j <= nj + 2; j++)
      for (int i = -1; i <= ni + 2; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        const double xm = 0.25 * (+R(i - 1, j, k) + R(i, j, k) + R(i - 1, j + 1, k) + R(i, j + 1, k));
        const double xp = 0.25 * (+R(i, j, k) + R(i + 1, j, k) + R(i, j + 1, k) + R(i + 1, j + 1, k));
        const double ym = R(i, j, k), yp = R(i, j + 1, k);
        const double zm = 0.25 * (+R(i, j, k - 1) + R(i, j + 1, k - 1) + R(i, j, k) + R(i, j + 1, k));
        const double zp = 0.25 * (+R(i, j, k) + R(i, j + 1, k) + R(i, j, k + 1) + R(i, j + 1, k + 1));
        const double b = dz2 * zm, s = dy2 * ym * 2.0, w = dx2 * xm, e = dx2 * xp;
        const double n = dy2 * yp * 2.0, t = dz2 * zp;
        a[CV_B][c] = b; a[CV_S][c] = s; a[CV_W][c] = w; a[CV_E][c] = e; a[CV_N][c] = n; a[CV_T][c] = t;
        a[CV_USW][c] = dxy * xm; a[CV_UNW][c] = -(dxy * xm);
        a[CV_USE][c] = -(dxy * xp); a[CV_UNE][c] = dxy * xp;
        a[CV_WSB][c] = dyz * zm; a[CV_WNB][c] = -(dyz * zm);
        a[CV_WST][c] = -(dyz * zp); a[CV_WNT][c] = dyz * zp;
        a[CV_P][c] = +b + s + w + e + n + t;
      }
  for (int k = -1; k <= nk + 2; k++)
    for (int j = -1; j <= nj + 2; j++)
      for (int i = -1; i <= ni + 2; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        const double xm = 0.25 * (+R(i - 1, j, k) + R(i, j, k) + R(i - 1, j, k + 1) + R(i, j, k + 1));
        const double xp = 0.25 * (+R(i, j, k) + R(i + 1, j, k) + R(i, j, k + 1) + R(i + 1, j, k + 1));
        const double ym = 0.25 * (+R(i, j - 1, k) + R(i, j, k) + R(i, j - 1, k + 1) + R(i, j, k + 1));
        const double yp = 0.25 * (+R(i, j, k) + R(i, j + 1, k) + R(i, j, k + 1) + R(i, j + 1, k + 1));
        const double zm = R(i, j, k), zp = R(i, j, k + 1);
        const double b = dz2 * zm * 2.0, s = dy2 * ym, w = dx2 * xm, e = dx2 * xp;
        const double n = dy2 * yp, t = dz2 * zp * 2.0;
        a[CW_B][c] = b; a[CW_S][c] = s; a[CW_W][c] = w; a[CW_E][c] = e; a[CW_N][c] = n; a[CW_T][c] = t;
        a[CW_UBW][c] = dxz * xm; a[CW_UTW][c] = -(dxz * xm);
        a[CW_UBE][c] = -(dxz * xp); a[CW_UTE][c] = dxz * xp;
        a[CW_VBS][c] = dyz * ym; a[CW_VTS][c] = -(dyz * ym);
        a[CW_VBN][c] = -(dyz * yp); a[CW_VTN][c] = dyz * yp;
        a[CW_P][c] = +b + s + w + e + n + t;
      }
#undef R
}

void orc_cal_arith_coef_vis(const int *ni, const int *nj, const int *nk, const double *dxinv,
                            const double *dyinv, const double *dzinv, const double *rmuk,
                            double *a0, double *a1, double *a2, double *a3, double *a4, double *a5,
                            double *a6, double *a7, double *a8, double *a9, double *a10,
                            double *a11, double *a12, double *a13, double *a14, double *a15,
                            double *a16, double *a17, double *a18, double *a19, double *a20,
                            double *a21, double *a22, double *a23, double *a24, double *a25,
                            double *a26, double *a27, double *a28, double *a29, double *a30,
                            double *a31, double *a32, double *a33, double *a34, double *a35,
                            double *a36, double *a37, double *a38, double *a39, double *a40,
                            double *a41, double *a42, double *a43, double *a44) {
  double *const a[NCOEF] = {a0,  a1,  a2,  a3,  a4,  a5,  a6,  a7,  a8,  a9,  a10, a11,
                            a12, a13, a14, a15, a16, a17, a18, a19, a20, a21, a22, a23,
                            a24, a25, a26, a27, a28, a29, a30, a31, a32, a33, a34, a35,
                            a36, a37, a38, a39, a40, a41, a42, a43, a44};
  coef_vis(*ni, *nj, *nk, *dxinv, *dyinv, *dzinv, rmuk, a);
}

/* srcs/cpy.f:1-24.  pinned: trivially. */
void orc_cpy(const int *ni_, const int *nj_, const int *nk_, const double *q1, double *q2) {
  ODIMS(*ni_, *nj_, *nk_);
  memcpy(q2, q1, sizeof(double) * n3);
}

/* srcs/bnd_periodic.f:1-46: x-periodic then z-periodic ghost fill.
 * pinned: _ref. */
void orc_bnd_periodic(const int *ni_, const int *nj_, const int *nk_, double *qk) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  for (int k = -2; k <= nk + 3; k++)
    for (int j = -2; j <= nj + 3; j++) {
      qk[OIX(-2, j, k)] = qk[OIX(ni - 2, j, k)];
      qk[OIX(-1, j, k)] = qk[OIX(ni - 1, j, k)];
      qk[OIX(0, j, k)] = qk[OIX(ni, j, k)];
      qk[OIX(ni + 1, j, k)] = qk[OIX(1, j, k)];
      qk[OIX(ni + 2, j, k)] = qk[OIX(2, j, k)];
      qk[OIX(ni + 3, j, k)] = qk[OIX(3, j, k)];
    }
  for (int j = -2; j <= nj + 3; j++)
    for (int i = -2; i <= ni + 3; i++) {
      qk[OIX(i, j, -2)] = qk[OIX(i, j, nk - 2)];
      qk[OIX(i, j, -1)] = qk[OIX(i, j, nk - 1)];
      qk[OIX(i, j, 0)] = qk[OIX(i, j, nk)];
      qk[OIX(i, j, nk + 1)] = qk[OIX(i, j, 1)];
      qk[OIX(i, j, nk + 2)] = qk[OIX(i, j, 2)];
      qk[OIX(i, j, nk + 3)] = qk[OIX(i, j, 3)];
    }
}

/* srcs/bnd_neumann.f:4-46: mirror ghosts at the j walls.
 * pinned: _ref. */
void orc_bnd_neumann(const int *nID, const int *ni_, const int *nj_, const int *nk_, double *qk) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  if (nID[ORC_Y_MINUS - 1] < 0)
    for (int k = -2; k <= nk + 3; k++)
      for (int i = -2; i <= ni + 3; i++) {
        qk[OIX(i, -2, k)] = qk[OIX(i, 3, k)];
        qk[OIX(i, -1, k)] = qk[OIX(i, 2, k)];
        qk[OIX(i, 0, k)] = qk[OIX(i, 1, k)];
      }
  if (nID[ORC_Y_PLUS - 1] < 0)
    for (int k = -2; k <= nk + 3; k++)
      for (int i = -2; i <= ni + 3; i++) {
        qk[OIX(i, nj + 1, k)] = qk[OIX(i, nj, k)];
        qk[OIX(i, nj + 2, k)] = qk[OIX(i, nj - 1, k)];
        qk[OIX(i, nj + 3, k)] = qk[OIX(i, nj - 2, k)];
      }
}

/* srcs/bnd_comm.f:1-120: the halo exchange is a no-op on one rank
 * (bnd_comm.f:18).  pinned: by definition. */
void orc_bnd_comm(const int *ipara, const int *nID, const int *ni, const int *nj, const int *nk,
                  const int *key, double *sendjb, double *recvjb, double *qk) {
  (void)ipara; (void)nID; (void)ni; (void)nj; (void)nk; (void)key; (void)sendjb; (void)recvjb; (void)qk;
}

/* srcs/bndu.f:1-161.  Navier slip at the j walls with the phase-weighted slip
 * length, then the x- and z-periodic fills with the off-by-one for u in x and
 * for w in z that the reference has (bndu.f:107-109,144-146).
 * pinned: _ref. */
void orc_bndu(const int *nID, const int *ni_, const int *nj_, const int *nk_, double *uk,
              double *vk, double *wk, const double *uwall_, const double *dy_, const double *l1,
              const double *l2, const double *phi) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double uwall = *uwall_, dy = *dy_;
  ODIMS(ni, nj, nk);
  if (nID[ORC_Y_MINUS - 1] < 0)
    for (int k = -2; k <= nk + 3; k++)
      for (int i = -2; i <= ni + 3; i++) {
        const double phi_av = (phi[OIX(i, 1, k)] + phi[OIX((i + 1) % (ni + 3), 1, k)]) / 2.0;
        const double inv1 = phi_av / l1[OIX(i, 1, k)];
        const double inv2 = (1.0 - phi_av) / l2[OIX(i, 1, k)];
        const double ls = 1.0 / (inv1 + inv2);
        const double coef1 = (2.0 * dy) / (2.0 * ls + dy);
        const double coef2 = (2.0 * ls - dy) / (2.0 * ls + dy);
        uk[OIX(i, 0, k)] = coef1 * (-uwall) + coef2 * uk[OIX(i, 1, k)];
        uk[OIX(i, -1, k)] = 2.0 * uk[OIX(i, 0, k)] - uk[OIX(i, 1, k)];
        uk[OIX(i, -2, k)] = 2.0 * uk[OIX(i, -1, k)] - uk[OIX(i, 0, k)];
        vk[OIX(i, -2, k)] = -vk[OIX(i, 2, k)];
        vk[OIX(i, -1, k)] = -vk[OIX(i, 1, k)];
        vk[OIX(i, 0, k)] = 0.0;
        wk[OIX(i, 0, k)] = coef2 * wk[OIX(i, 1, k)];
        wk[OIX(i, -1, k)] = 2.0 * wk[OIX(i, 0, k)] - wk[OIX(i, 1, k)];
        wk[OIX(i, -2, k)] = 2.0 * wk[OIX(i, -1, k)] - wk[OIX(i, 0, k)];
      }
  if (nID[ORC_Y_PLUS - 1] < 0)
    for (int k = -2; k <= nk + 3; k++)
      for (int i = -2; i <= ni + 3; i++) {
        const double phi_av = (phi[OIX(i, nj, k)] + phi[OIX((i + 1) % (ni + 3), nj, k)]) / 2.0;
        const double inv1 = phi_av / l1[OIX(i, nj, k)];
        const double inv2 = (1.0 - phi_av) / l2[OIX(i, nj, k)];
        const double ls = 1.0 / (inv1 + inv2);
        const double coef1 = (2.0 * dy) / (2.0 * ls + dy);
        const double coef2 = (2.0 * ls - dy) / (2.0 * ls + dy);
        uk[OIX(i, nj + 1, k)] = coef1 * uwall + coef2 * uk[OIX(i, nj, k)];
        uk[OIX(i, nj + 2, k)] = 2.0 * uk[OIX(i, nj + 1, k)] - uk[OIX(i, nj, k)];
        uk[OIX(i, nj + 3, k)] = 2.0 * uk[OIX(i, nj + 2, k)] - uk[OIX(i, nj + 1, k)];
        vk[OIX(i, nj, k)] = 0.0;
        vk[OIX(i, nj + 1, k)] = -vk[OIX(i, nj - 1, k)];
        vk[OIX(i, nj + 2, k)] = -vk[OIX(i, nj - 2, k)];
        vk[OIX(i, nj + 3, k)] = -vk[OIX(i, nj - 3, k)];
        wk[OIX(i, nj + 1, k)] = coef2 * wk[OIX(i, nj, k)];
        wk[OIX(i, nj + 2, k)] = 2.0 * wk[OIX(i, nj + 1, k)] - wk[OIX(i, nj, k)];
        wk[OIX(i, nj + 3, k)] = 2.0 * wk[OIX(i, nj + 2, k)] - wk[OIX(i, nj + 1, k)];
      }
  for (int j = -2; j <= nj + 3; j++)
    for (int k = -2; k <= nk + 3; k++) {
      uk[OIX(0, j, k)] = uk[OIX(ni + 1, j, k)];
      uk[OIX(-1, j, k)] = uk[OIX(ni, j, k)];
      uk[OIX(-2, j, k)] = uk[OIX(ni - 1, j, k)];
      vk[OIX(0, j, k)] = vk[OIX(ni, j, k)];
      vk[OIX(-1, j, k)] = vk[OIX(ni - 1, j, k)];
      vk[OIX(-2, j, k)] = vk[OIX(ni - 2, j, k)];
      wk[OIX(0, j, k)] = wk[OIX(ni, j, k)];
      wk[OIX(-1, j, k)] = wk[OIX(ni - 1, j, k)];
      wk[OIX(-2, j, k)] = wk[OIX(ni - 2, j, k)];
      uk[OIX(ni + 2, j, k)] = uk[OIX(2, j, k)];
      uk[OIX(ni + 3, j, k)] = uk[OIX(3, j, k)];
      vk[OIX(ni + 1, j, k)] = vk[OIX(1, j, k)];
      vk[OIX(ni + 2, j, k)] = vk[OIX(2, j, k)];
      vk[OIX(ni + 3, j, k)] = vk[OIX(3, j, k)];
      wk[OIX(ni + 1, j, k)] = wk[OIX(1, j, k)];
      wk[OIX(ni + 2, j, k)] = wk[OIX(2, j, k)];
      wk[OIX(ni + 3, j, k)] = wk[OIX(3, j, k)];
    }
  for (int i = -2; i <= ni + 3; i++)
    for (int j = -2; j <= nj + 3; j++) {
      uk[OIX(i, j, 0)] = uk[OIX(i, j, nk)];
      uk[OIX(i, j, -1)] = uk[OIX(i, j, nk - 1)];
      uk[OIX(i, j, -2)] = uk[OIX(i, j, nk - 2)];
      vk[OIX(i, j, 0)] = vk[OIX(i, j, nk)];
      vk[OIX(i, j, -1)] = vk[OIX(i, j, nk - 1)];
      vk[OIX(i, j, -2)] = vk[OIX(i, j, nk - 2)];
      wk[OIX(i, j, 0)] = wk[OIX(i, j, nk + 1)];
      wk[OIX(i, j, -1)] = wk[OIX(i, j, nk)];
      wk[OIX(i, j, -2)] = wk[OIX(i, j, nk - 1)];
      uk[OIX(i, j, nk + 1)] = uk[OIX(i, j, 1)];
      uk[OIX(i, j, nk + 2)] = uk[OIX(i, j, 2)];
      uk[OIX(i, j, nk + 3)] = uk[OIX(i, j, 3)];
      vk[OIX(i, j, nk + 1)] = vk[OIX(i, j, 1)];
      vk[OIX(i, j, nk + 2)] = vk[OIX(i, j, 2)];
      vk[OIX(i, j, nk + 3)] = vk[OIX(i, j, 3)];
      wk[OIX(i, j, nk + 2)] = wk[OIX(i, j, 2)];
      wk[OIX(i, j, nk + 3)] = wk[OIX(i, j, 3)];
    }
}

/* Iteration count and residual ratio of the last orc_solu_sor4 call, for the
 * monitor and the tests (the reference prints them, srcs/solu_sor4.f:298-305). */
int orc_sor_last_lp = -1;
double orc_sor_last_ratio = 0.0;

/* srcs/solu_sor4.f:1-308.  CN viscous solve: block Gauss-Seidel u -> v -> w,
 * each over 8 parity colours in the order (111,222,112,221,121,212,211,122),
 * omega = 1, BC refresh (bndu + bnd_periodic of the swept component) after
 * every second colour; stop when err/err0 <= 1e-3 or after 21 sweeps.
 * Unbuildable reference (MPI header).  pinned: bit-identical to an
 * independent numpy colour sweep that calls the compiled reference
 * bndu/bnd_periodic between colour pairs, and its converged state satisfies
 * the Crank-Nicolson equation evaluated through the explicit stress chain
 * (tests/test_oracle_math.py). */
static void sor4_impl(const int *ipara, const int *ID, const int *nID, const int *ndiv,
                      const int *ni_, const int *nj_, const int *nk_, const int *key,
                      double *sendjb, double *recvjb, const int *nstep, const int *imon_t,
                      const double *dt_, const double *rho, const double *rhon,
                      double *const *coef, const double *src_u, const double *src_v,
                      const double *src_w, double *un, double *vn, double *wn, const double *uwall,
                      const double *dy, const double *l1, const double *l2, const double *phi) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  static const int isor[8] = {1, 2, 1, 2, 1, 2, 2, 1};
  static const int jsor[8] = {1, 2, 1, 2, 2, 1, 1, 2};
  static const int ksor[8] = {1, 2, 2, 1, 1, 2, 1, 2};
  const double dtinv = 1.0 / (*dt_);
  const int jen = (nID[ORC_Y_PLUS - 1] < 0) ? nj - 1 : nj;
  const ptrdiff_t SX = 1, SY = sx, SZ = sx * sy;
  double err = 0.0, err0 = 1.0;
  int lp;
  for (lp = 0; lp <= 20; lp++) {
    err = 0.0;
    for (int c = 0; c < 8; c++) {
      for (int k = ksor[c]; k <= nk; k += 2)
        for (int j = jsor[c]; j <= nj; j += 2)
          for (int i = isor[c]; i <= ni; i += 2) {
            const ptrdiff_t m = OIX(i, j, k);
            const double r00 = (rho[m] + rho[m + SX] + rhon[m] + rhon[m + SX]) * 0.25;
            const double visn =
                (+coef[CU_B][m] * un[m - SZ] + coef[CU_S][m] * un[m - SY] + coef[CU_W][m] * un[m - SX] +
                 coef[CU_E][m] * un[m + SX] + coef[CU_N][m] * un[m + SY] + coef[CU_T][m] * un[m + SZ] +
                 coef[CU_VWS][m] * vn[m - SY] + coef[CU_VES][m] * vn[m + SX - SY] +
                 coef[CU_VWN][m] * vn[m] + coef[CU_VEN][m] * vn[m + SX] +
                 coef[CU_WWB][m] * wn[m - SZ] + coef[CU_WEB][m] * wn[m + SX - SZ] +
                 coef[CU_WWT][m] * wn[m] + coef[CU_WET][m] * wn[m + SX]) /
                r00;
            const double du =
                (visn * 0.5 + src_u[m]) / (dtinv + coef[CU_P][m] / r00 * 0.5) - un[m];
            un[m] = un[m] + du;
            err = err + du * du;
          }
      if (c % 2 == 1) {
        orc_bndu(nID, ni_, nj_, nk_, un, vn, wn, uwall, dy, l1, l2, phi);
        orc_bnd_periodic(ni_, nj_, nk_, un);
        orc_bnd_comm(ipara, nID, ni_, nj_, nk_, key, sendjb, recvjb, un);
      }
    }
    for (int c = 0; c < 8; c++) {
      for (int k = ksor[c]; k <= nk; k += 2)
        for (int j = jsor[c]; j <= jen; j += 2)
          for (int i = isor[c]; i <= ni; i += 2) {
            const ptrdiff_t m = OIX(i, j, k);
            const double r00 = (rho[m] + rho[m + SY] + rhon[m] + rhon[m + SY]) * 0.25;
            const double visn =
                (+coef[CV_B][m] * vn[m - SZ] + coef[CV_S][m] * vn[m - SY] + coef[CV_W][m] * vn[m - SX] +
                 coef[CV_E][m] * vn[m + SX] + coef[CV_N][m] * vn[m + SY] + coef[CV_T][m] * vn[m + SZ] +
                 coef[CV_USW][m] * un[m - SX] + coef[CV_USE][m] * un[m] +
                 coef[CV_UNW][m] * un[m - SX + SY] + coef[CV_UNE][m] * un[m + SY] +
                 coef[CV_WSB][m] * wn[m - SZ] + coef[CV_WNB][m] * wn[m + SY - SZ] +
                 coef[CV_WST][m] * wn[m] + coef[CV_WNT][m] * wn[m + SY]) /
                r00;
            const double dv =
                (visn * 0.5 + src_v[m]) / (dtinv + coef[CV_P][m] / r00 * 0.5) - vn[m];
            vn[m] = vn[m] + dv;
            err = err + dv * dv;
          }
      if (c % 2 == 1) {
        orc_bndu(nID, ni_, nj_, nk_, un, vn, wn, uwall, dy, l1, l2, phi);
        orc_bnd_periodic(ni_, nj_, nk_, vn);
        orc_bnd_comm(ipara, nID, ni_, nj_, nk_, key, sendjb, recvjb, vn);
      }
    }
    for (int c = 0; c < 8; c++) {
      for (int k = ksor[c]; k <= nk; k += 2)
        for (int j = jsor[c]; j <= nj; j += 2)
          for (int i = isor[c]; i <= ni; i += 2) {
            const ptrdiff_t m = OIX(i, j, k);
            const double r00 = (rho[m] + rho[m + SZ] + rhon[m] + rhon[m + SZ]) * 0.25;
            const double visn =
                (+coef[CW_B][m] * wn[m - SZ] + coef[CW_S][m] * wn[m - SY] + coef[CW_W][m] * wn[m - SX] +
                 coef[CW_E][m] * wn[m + SX] + coef[CW_N][m] * wn[m + SY] + coef[CW_T][m] * wn[m + SZ] +
                 coef[CW_UBW][m] * un[m - SX] + coef[CW_UBE][m] * un[m] +
                 coef[CW_UTW][m] * un[m - SX + SZ] + coef[CW_UTE][m] * un[m + SZ] +
                 coef[CW_VBS][m] * vn[m - SY] + coef[CW_VBN][m] * vn[m] +
                 coef[CW_VTS][m] * vn[m - SY + SZ] + coef[CW_VTN][m] * vn[m + SZ]) /
                r00;
            const double dw =
                (visn * 0.5 + src_w[m]) / (dtinv + coef[CW_P][m] / r00 * 0.5) - wn[m];
            wn[m] = wn[m] + dw;
            err = err + dw * dw;
          }
      if (c % 2 == 1) {
        orc_bndu(nID, ni_, nj_, nk_, un, vn, wn, uwall, dy, l1, l2, phi);
        orc_bnd_periodic(ni_, nj_, nk_, wn);
        orc_bnd_comm(ipara, nID, ni_, nj_, nk_, key, sendjb, recvjb, wn);
      }
    }
    err = sqrt(err / (double)(ni * nj * nk * (*ndiv))) + 1.0e-99;
    if (lp == 0) err0 = err;
    if (err / err0 <= 1.0e-3) break;
  }
  /* the loop variable on exit: 21 after 21 sweeps without convergence */
  orc_sor_last_lp = lp;
  orc_sor_last_ratio = err / err0;
  if (*imon_t > 0 && (*nstep % *imon_t) == 0 && *ID == 0)
    printf("Err_solu %10d%10d%20.10E%20.10E%20.10E\n", *nstep, lp, err / err0, err, err0);
}

/* Argument-list form of srcs/solu_sor4.f:1-17 (45 coefficient arrays). */
void orc_solu_sor4(const int *ipara, const int *ID, const int *nID, const int *ndiv,
                   const int *ni, const int *nj, const int *nk, const int *key, double *sendjb,
                   double *recvjb, const int *nstep, const int *imon_t, const double *dt,
                   const double *rho, const double *rhon, double *c0, double *c1, double *c2, double *c3, double *c4, double *c5, double *c6, double *c7, double *c8, double *c9, double *c10, double *c11, double *c12, double *c13, double *c14, double *c15, double *c16, double *c17, double *c18, double *c19, double *c20, double *c21, double *c22, double *c23, double *c24, double *c25, double *c26, double *c27, double *c28, double *c29, double *c30, double *c31, double *c32, double *c33, double *c34, double *c35, double *c36, double *c37, double *c38, double *c39, double *c40, double *c41, double *c42, double *c43, double *c44,
                   const double *src_u, const double *src_v, const double *src_w, double *un,
                   double *vn, double *wn, const double *uwall, const double *dy,
                   const double *l1, const double *l2, const double *phi) {
  double *const coef[NCOEF] = {c0, c1, c2, c3, c4, c5, c6, c7, c8, c9, c10, c11, c12, c13, c14, c15, c16, c17, c18, c19, c20, c21, c22, c23, c24, c25, c26, c27, c28, c29, c30, c31, c32, c33, c34, c35, c36, c37, c38, c39, c40, c41, c42, c43, c44};
  sor4_impl(ipara, ID, nID, ndiv, ni, nj, nk, key, sendjb, recvjb, nstep, imon_t, dt, rho, rhon,
            coef, src_u, src_v, src_w, un, vn, wn, uwall, dy, l1, l2, phi);
}

/* ------------------------------------------------------------------------- */
/* Pressure projection                                                       */
/* ------------------------------------------------------------------------- */

/* srcs/caldiv.f:1-70.  Unbuildable reference (MPI header); pinned by the
 * closed-form divergence of polynomial fields. */
void orc_caldiv(const int *ipara, const int *ndiv, const int *ni_, const int *nj_, const int *nk_,
                const double *dxinv_, const double *dyinv_, const double *dzinv_,
                const double *dt_, const double *uk, const double *vk, const double *wk,
                double *div, double *err_div) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  (void)ipara;
  const double dtinv = 1.0 / (*dt_);
  for (int k = -1; k <= nk + 3; k++)
    for (int j = -1; j <= nj + 3; j++)
      for (int i = -1; i <= ni + 3; i++)
        div[OIX(i, j, k)] = ((-uk[OIX(i - 1, j, k)] + uk[OIX(i, j, k)]) * dxinv +
                             (-vk[OIX(i, j - 1, k)] + vk[OIX(i, j, k)]) * dyinv +
                             (-wk[OIX(i, j, k - 1)] + wk[OIX(i, j, k)]) * dzinv) *
                            dtinv;
  double e = 0.0;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) e = e + div[OIX(i, j, k)] * div[OIX(i, j, k)];
  *err_div = sqrt(e / (double)(ni * nj * nk * (*ndiv))) + 1.0e-99;
}

/* In-place radix-2 complex FFT of length n (power of two) with stride 1,
 * exp(sign*2*pi*i*m*k/n), twiddles from cos/sin of the exact angle.  This is
 * the DFT that srcs/fourn.f computes; the oracle does not restate fourn's
 * trigonometric recurrence.  pinned: numpy.fft. */
static void fft1(double *re, double *im, int n, int sign) {
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int h = len >> 1;
    for (int m = 0; m < h; m++) {
      const double ang = (double)sign * 2.0 * M_PI * (double)m / (double)len;
      const double wr = cos(ang), wi = sin(ang);
      for (int s = 0; s < n; s += len) {
        const int a = s + m, b = s + m + h;
        const double tr = re[b] * wr - im[b] * wi, ti = re[b] * wi + im[b] * wr;
        re[b] = re[a] - tr; im[b] = im[a] - ti;
        re[a] = re[a] + tr; im[a] = im[a] + ti;
      }
    }
  }
}

/* 2-D DFT of a (nk rows, ni columns) plane, x fastest. */
static void fft2(double *re, double *im, int ni, int nk, int sign) {
  double *cr = malloc(sizeof(double) * nk), *ci = malloc(sizeof(double) * nk);
  for (int k = 0; k < nk; k++) fft1(re + (size_t)k * ni, im + (size_t)k * ni, ni, sign);
  for (int i = 0; i < ni; i++) {
    for (int k = 0; k < nk; k++) { cr[k] = re[(size_t)k * ni + i]; ci[k] = im[(size_t)k * ni + i]; }
    fft1(cr, ci, nk, sign);
    for (int k = 0; k < nk; k++) { re[(size_t)k * ni + i] = cr[k]; im[(size_t)k * ni + i] = ci[k]; }
  }
  free(cr);
  free(ci);
}

/* srcs/solp_fft_tdma1.f:1-50: per j-plane forward 2-D DFT of div, divided by
 * ni*nk, stored at centred mode indices phir(-ni/2:ni/2-1, nj, -nk/2:nk/2-1).
 * pinned: numpy.fft.fft2 and the compiled reference (fourn) to 1e-13. */
void orc_solp_fft_tdma1(const int *ni_, const int *nj_, const int *nk_, double *fftdata,
                        const double *div, double *phir_r, double *phir_i) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)fftdata;
  double *re = malloc(sizeof(double) * ni * nk), *im = malloc(sizeof(double) * ni * nk);
  const double nrm = (double)(ni * nk);
  for (int j = 1; j <= nj; j++) {
    for (int k = 0; k < nk; k++)
      for (int i = 0; i < ni; i++) {
        re[(size_t)k * ni + i] = div[OIX(i + 1, j, k + 1)];
        im[(size_t)k * ni + i] = 0.0;
      }
    fft2(re, im, ni, nk, -1);
    for (int k = -nk / 2; k <= nk / 2 - 1; k++) {
      const int kk = (k + nk) % nk;
      for (int i = -ni / 2; i <= ni / 2 - 1; i++) {
        const int ii = (i + ni) % ni;
        const size_t o = (size_t)(i + ni / 2) + (size_t)ni * ((size_t)(j - 1) + (size_t)nj * (size_t)(k + nk / 2));
        phir_r[o] = re[(size_t)kk * ni + ii] / nrm;
        phir_i[o] = im[(size_t)kk * ni + ii] / nrm;
      }
    }
  }
  free(re);
  free(im);
}

/* srcs/trans_r2w.f:23-38 and srcs/trans_w2r.f:23-38 on one rank: relabel of
 * the k mode index, which leaves the memory image unchanged.  pinned: by the
 * index identity (both arrays start at the same flat offset). */
void orc_trans_r2w(const int *ipara, const int *ID, const int *ndiv, const int *svall,
                   const int *key, double *sendbuf, double *recvbuf, const double *r, double *w) {
  (void)ipara; (void)ID; (void)ndiv; (void)key; (void)sendbuf; (void)recvbuf;
  const size_t n = (size_t)svall[0] * svall[1] * svall[2];
  if (w != r) memcpy(w, r, sizeof(double) * n);
}

void orc_trans_w2r(const int *ipara, const int *ID, const int *ndiv, const int *svall,
                   const int *key, double *sendbuf, double *recvbuf, const double *w, double *r) {
  (void)ipara; (void)ID; (void)ndiv; (void)key; (void)sendbuf; (void)recvbuf;
  const size_t n = (size_t)svall[0] * svall[1] * svall[2];
  if (w != r) memcpy(r, w, sizeof(double) * n);
}

/* srcs/solp_fft_tdma2.f:4-91: one Thomas solve along j per (kx,kz) mode
 * with the modal diagonal shift; the singular (0,0) mode is pinned by
 * apinv = -d/(d^2+1e-99).  pinned: _ref (bitwise) and dense solves. */
void orc_solp_fft_tdma2(const int *ID, const int *ndiv, const int *svall, const double *rho_,
                        const double *dxinv_, const double *dzinv_, const double *as_p_fft,
                        const double *an_p_fft, const double *ap_p_fft, double *atdma,
                        double *btdma_r, double *btdma_i, double *phiw_r, double *phiw_i) {
  const int ni = svall[0], nj = svall[1], nk = svall[2] / (*ndiv);
  const double rho = *rho_, dxinv = *dxinv_, dzinv = *dzinv_;
  const double pi = atan(1.0) * 4.0;
  const double dx2r = dxinv * dxinv / rho, dz2r = dzinv * dzinv / rho;
  const double pi2_ni = pi * 2.0 / (double)svall[0], pi2_nk = pi * 2.0 / (double)svall[2];
  /* 1-D arrays are declared (-2:nj+3); index j lives at [j+2]. */
  double *a = atdma + 2, *br = btdma_r + 2, *bi = btdma_i + 2;
  const double *as = as_p_fft + 2, *an = an_p_fft + 2, *ap0 = ap_p_fft + 2;
  for (int k = 1; k <= nk; k++)
    for (int i = -ni / 2; i <= ni / 2 - 1; i++) {
      const double c1x = cos((double)i * pi2_ni) * 2.0;
      const double c1z = cos((double)(k + (*ID) * nk - svall[2] / 2 - 1) * pi2_nk) * 2.0;
      a[0] = 0.0; br[0] = 0.0; bi[0] = 0.0;
      a[nj + 1] = 0.0; br[nj + 1] = 0.0; bi[nj + 1] = 0.0;
      for (int j = 1; j <= nj; j++) {
        const size_t o = (size_t)(i + ni / 2) + (size_t)ni * ((size_t)(j - 1) + (size_t)nj * (size_t)(k - 1));
        const double ap = ap0[j] + dx2r * (2.0 - c1x) + dz2r * (2.0 - c1z);
        const double d00 = -ap + as[j] * a[j - 1];
        const double apinv = -d00 / (d00 * d00 + 1.0e-99);
        a[j] = apinv * an[j];
        br[j] = apinv * (-phiw_r[o] + as[j] * br[j - 1]);
        bi[j] = apinv * (-phiw_i[o] + as[j] * bi[j - 1]);
      }
      {
        const size_t o = (size_t)(i + ni / 2) + (size_t)ni * ((size_t)(nj - 1) + (size_t)nj * (size_t)(k - 1));
        phiw_r[o] = br[nj];
        phiw_i[o] = bi[nj];
      }
      for (int j = nj - 1; j >= 1; j--) {
        const size_t o = (size_t)(i + ni / 2) + (size_t)ni * ((size_t)(j - 1) + (size_t)nj * (size_t)(k - 1));
        phiw_r[o] = br[j] + a[j] * phiw_r[o + ni];
        phiw_i[o] = bi[j] + a[j] * phiw_i[o + ni];
      }
    }
}

/* srcs/solp_fft_tdma3.f:1-51: inverse (unnormalised) 2-D DFT per j-plane,
 * dp = real part.  pinned: numpy.fft.ifft2 and the compiled reference. */
void orc_solp_fft_tdma3(const int *ni_, const int *nj_, const int *nk_, double *fftdata,
                        const double *phir_r, const double *phir_i, double *dp) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)fftdata;
  double *re = malloc(sizeof(double) * ni * nk), *im = malloc(sizeof(double) * ni * nk);
  for (int j = 1; j <= nj; j++) {
    for (int k = -nk / 2; k <= nk / 2 - 1; k++) {
      const int kk = (k + nk) % nk;
      for (int i = -ni / 2; i <= ni / 2 - 1; i++) {
        const int ii = (i + ni) % ni;
        const size_t o = (size_t)(i + ni / 2) + (size_t)ni * ((size_t)(j - 1) + (size_t)nj * (size_t)(k + nk / 2));
        re[(size_t)kk * ni + ii] = phir_r[o];
        im[(size_t)kk * ni + ii] = phir_i[o];
      }
    }
    fft2(re, im, ni, nk, +1);
    for (int k = 0; k < nk; k++)
      for (int i = 0; i < ni; i++) dp[OIX(i + 1, j, k + 1)] = re[(size_t)k * ni + i];
  }
  free(re);
  free(im);
}

/* Relative residual of the discrete Poisson problem, srcs/solp_fft_tdma4.f:20-93
 * (the reference prints it every imon_t steps).  Returned for tests. */
double orc_poisson_residual(int ni, int nj, int nk, double rho, double dxinv, double dyinv,
                            double dzinv, const double *div, const double *dp, double *err_abs) {
  ODIMS(ni, nj, nk);
  const double aw = dxinv * dxinv / rho, ae = dxinv * dxinv / rho;
  const double as = dyinv * dyinv / rho, an = dyinv * dyinv / rho;
  const double ab = dzinv * dzinv / rho, at = dzinv * dzinv / rho;
  const double ap = aw + ae + as + an + ab + at;
  double e0 = 0.0, e = 0.0;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) e0 = e0 + div[OIX(i, j, k)] * div[OIX(i, j, k)];
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const double d = +dp[OIX(i, j, k - 1)] * ab + dp[OIX(i, j - 1, k)] * as +
                         dp[OIX(i - 1, j, k)] * aw - dp[OIX(i, j, k)] * ap +
                         dp[OIX(i + 1, j, k)] * ae + dp[OIX(i, j + 1, k)] * an +
                         dp[OIX(i, j, k + 1)] * at - div[OIX(i, j, k)];
        e = e + d * d;
      }
  e0 = sqrt(e0 / (double)(ni * nj * nk)) + 1.0e-99;
  e = sqrt(e / (double)(ni * nj * nk)) + 1.0e-99;
  if (err_abs) *err_abs = e;
  return e / e0;
}

/* srcs/solp_fft_tdma4.f:1-98: residual monitor, printed every imon_t steps.
 * Unbuildable reference (MPI header); pinned by the discrete-eigenfunction
 * test (residual of an exact solution is 0 to rounding). */
void orc_solp_fft_tdma4(const int *ipara, const int *ID, const int *ndiv, const int *ni,
                        const int *nj, const int *nk, const int *nstep, const int *imon_t,
                        const double *rho, const double *dxinv, const double *dyinv,
                        const double *dzinv, const double *div, const double *dp) {
  (void)ipara; (void)ndiv;
  if (*imon_t > 0 && (*nstep % *imon_t) == 0) {
    double e;
    const double r = orc_poisson_residual(*ni, *nj, *nk, *rho, *dxinv, *dyinv, *dzinv, div, dp, &e);
    if (*ID == 0) printf("Err_solp %10d%20.10E%20.10E%20.10E\n", *nstep, r, e, e / r);
  }
}

/* srcs/corunp_explicit.f:1-78: projection of u, v, w and pn = phat + dp.
 * pinned: _ref. */
void orc_corunp_explicit(const int *nID, const int *ni_, const int *nj_, const int *nk_,
                         const double *rho_, const double *dxinv_, const double *dyinv_,
                         const double *dzinv_, const double *dt_, const double *dp,
                         const double *phat, double *un, double *vn, double *wn, double *pn) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double rho = *rho_, dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_, dt = *dt_;
  ODIMS(ni, nj, nk);
  const int jen = (nID[ORC_Y_PLUS - 1] < 0) ? nj - 1 : nj;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++)
        un[OIX(i, j, k)] = un[OIX(i, j, k)] + (dp[OIX(i, j, k)] - dp[OIX(i + 1, j, k)]) / rho * dxinv * dt;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= jen; j++)
      for (int i = 1; i <= ni; i++)
        vn[OIX(i, j, k)] = vn[OIX(i, j, k)] + (dp[OIX(i, j, k)] - dp[OIX(i, j + 1, k)]) / rho * dyinv * dt;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++)
        wn[OIX(i, j, k)] = wn[OIX(i, j, k)] + (dp[OIX(i, j, k)] - dp[OIX(i, j, k + 1)]) / rho * dzinv * dt;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) pn[OIX(i, j, k)] = phat[OIX(i, j, k)] + dp[OIX(i, j, k)];
}

/* Pressure matrix coefficients, srcs/cal_coef_prs.f:1-57 (setup).
 * pinned: _ref. */
void orc_cal_coef_prs(const int *nj_, const double *rho_, const double *dxinv_,
                      const double *dyinv_, const double *dzinv_, double *aw, double *ae,
                      double *as, double *an, double *ab, double *at, double *ap) {
  const int nj = *nj_;
  const double rho = *rho_, dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  aw += 2; ae += 2; as += 2; an += 2; ab += 2; at += 2; ap += 2;
  for (int j = 1; j <= nj; j++) {
    aw[j] = dxinv * dxinv / rho;
    ae[j] = dxinv * dxinv / rho;
    as[j] = dyinv * dyinv / rho;
    an[j] = dyinv * dyinv / rho;
    ab[j] = dzinv * dzinv / rho;
    at[j] = dzinv * dzinv / rho;
    ap[j] = +aw[j] + ae[j] + as[j] + an[j] + ab[j] + at[j];
  }
  ap[1] = ap[1] - as[1];
  as[1] = 0.0;
  ap[nj] = ap[nj] - an[nj];
  an[nj] = 0.0;
}

/* srcs/cal_coef_solp_fft.f:1-42 (setup).  pinned: _ref. */
void orc_cal_coef_solp_fft(const int *nj_, const double *as_p, const double *an_p,
                           double *as_fft, double *an_fft, double *ap_fft) {
  const int nj = *nj_;
  as_p += 2; an_p += 2; as_fft += 2; an_fft += 2; ap_fft += 2;
  for (int j = 1; j <= nj; j++) {
    as_fft[j] = as_p[j];
    an_fft[j] = an_p[j];
    ap_fft[j] = as_fft[j] + an_fft[j];
  }
  ap_fft[1] = ap_fft[1] - as_fft[1];
  as_fft[1] = 0.0;
  ap_fft[nj] = ap_fft[nj] - an_fft[nj];
  an_fft[nj] = 0.0;
}

/* ------------------------------------------------------------------------- */
/* Contact line                                                              */
/* ------------------------------------------------------------------------- */

/* srcs/gnbc.f90:1-99: dynamic contact angle from the wall-adjacent contact
 * line speed, theta = acos(clamp(cos(theta0) - zeta*u_cl/sigma)), written at
 * theta(i,1,k) and theta(i,nj,k).  pinned: _ref. */
void orc_gnbc(const int *nID, const int *ni_, const int *nj_, const int *nk_, const double *uk,
              const double *wk, const double *uwall_, const double *theta_0_array,
              const double *surface_tension, const double *zeta_array, double *theta_array) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double uwall = *uwall_, st = *surface_tension;
  ODIMS(ni, nj, nk);
  const double pi = acos(-1.0), eps = 1.0e-12;
  for (int side = 0; side < 2; side++) {
    if (side == 0 && !(nID[ORC_Y_MINUS - 1] < 0)) continue;
    if (side == 1 && !(nID[ORC_Y_PLUS - 1] < 0)) continue;
    const int j0 = side == 0 ? 0 : nj, j1 = side == 0 ? 1 : nj + 1, jw = side == 0 ? 1 : nj;
    const double uw = side == 0 ? uwall : -uwall;
    for (int k = -2; k <= nk + 3; k++)
      for (int i = -2; i <= ni + 3; i++) {
        const int mi = (i + ni + 2) % (ni + 3);
        const double ux = (uk[OIX(mi, j0, k)] + uk[OIX(mi, j1, k)] + uk[OIX(i, j0, k)] + uk[OIX(i, j1, k)]) / 4.0 + uw;
        const double uz = (wk[OIX(mi, j0, k)] + wk[OIX(mi, j1, k)] + wk[OIX(i, j0, k)] + wk[OIX(i, j1, k)]) / 4.0;
        double ucl = copysign(sqrt(ux * ux + uz * uz + eps), ux);
        if (i <= ni / 2) ucl = -ucl;
        const double th0 = theta_0_array[OIX(i, jw, k)] * (pi / 180.0);
        double ct = cos(th0) - zeta_array[OIX(i, jw, k)] * ucl / st;
        ct = fmin(1.0, fmax(-1.0, ct));
        theta_array[OIX(i, jw, k)] = acos(ct) * (180.0 / pi);
      }
  }
}

/* Wall-row read that returns 0 for i or k outside -2..n+3.  The reference
 * reads past the row or plane at the outermost ghost column/plane
 * (srcs/bnd_contact_angle.f90:45-46,77-78); those ghosts are overwritten by
 * bnd_periodic or never reach an interior value, and this restatement
 * defines the read as 0. */
static inline double rd_guard(const double *q, ptrdiff_t sx, ptrdiff_t sy, int ni, int nk, int i,
                              int j, int k) {
  return (i < -2 || i > ni + 3 || k < -2 || k > nk + 3) ? 0.0 : q[OIX(i, j, k)];
}

/* srcs/bnd_contact_angle.f90:1-88: contact-angle ghosts of phi at the j
 * walls, normal n = (s sin(th) dy/dx, cos(th), s sin(th) dy/dz)/|.| with
 * s = sign(i - ni/2), then two linear extrapolations.  Unbuildable
 * reference (mixed-kind SIGN); pinned by the per-cell closed form in
 * tests/test_oracle_math.py. */
void orc_bnd_contact_angle(const int *nID, const int *ni_, const int *nj_, const int *nk_,
                           double *qk, const double *theta_array, const double *dx_,
                           const double *dy_, const double *dz_) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  const double pi = acos(-1.0), eps = 1.0e-20;
  const double ax = *dy_ / *dx_, az = *dy_ / *dz_;
  for (int side = 0; side < 2; side++) {
    if (side == 0 && !(nID[ORC_Y_MINUS - 1] < 0)) continue;
    if (side == 1 && !(nID[ORC_Y_PLUS - 1] < 0)) continue;
    const int jw = side == 0 ? 1 : nj, jg = side == 0 ? 0 : nj + 1, js = side == 0 ? -1 : 1;
    for (int k = -2; k <= nk + 3; k++)
      for (int i = -2; i <= ni + 3; i++) {
        const double th = theta_array[OIX(i, jw, k)] * pi / 180.0;
        const double ct = cos(th), st = sin(th);
        const double sg = copysign(1.0, (double)(i - ni / 2));
        double nx = sg * st * ax, nz = sg * st * az, ny = ct;
        const double nrm = sqrt(nx * nx + ny * ny + nz * nz + eps);
        nx = nx / nrm;
        ny = ny / nrm;
        nz = nz / nrm;
        const double q1 = qk[OIX(i, jw, k)];
        const double g = q1 - (ny / (nx + eps)) * (rd_guard(qk, sx, sy, ni, nk, i + 1, jw, k) - rd_guard(qk, sx, sy, ni, nk, i - 1, jw, k)) * 0.5 * ax -
                         (ny / (nz + eps)) * (rd_guard(qk, sx, sy, ni, nk, i, jw, k + 1) - rd_guard(qk, sx, sy, ni, nk, i, jw, k - 1)) * 0.5 * az;
        qk[OIX(i, jg, k)] = g;
        qk[OIX(i, jg + js, k)] = 2.0 * g - q1;
        qk[OIX(i, jg + 2 * js, k)] = 2.0 * qk[OIX(i, jg + js, k)] - g;
      }
  }
}

/* ------------------------------------------------------------------------- */
/* Per-step diagnostics                                                      */
/* ------------------------------------------------------------------------- */

/* srcs/cal_center.f:1-120: phase-weighted centroid with lap tracking.
 * Unbuildable reference (MPI header); pinned by the closed-form centroid of
 * a uniform slab. */
void orc_cal_center(const int *ipara, const int *ID, const int *svall, const int *ni_,
                    const int *nj_, const int *nk_, const double *xl_, const double *yl_,
                    const double *zl_, const double *dt_, const double *rhog, const double *phi,
                    double *center, double *center_pre1, double *center_pre2, int *area_pre1,
                    int *area_pre2, int *lap, double *velocity) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)ipara; (void)zl_; (void)rhog;
  const double xl = *xl_, yl = *yl_;
  double weight = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
  const double delta = yl / (double)(*svall);
  const double s1 = xl * 0.25, s2 = xl * 0.75;
  if (*area_pre1 == 1 && *area_pre2 == 3) *lap = *lap + 1;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const double p = phi[OIX(i, j, k)];
        weight = weight + p;
        double lx = delta * ((double)i - 0.5);
        const double ly = delta * ((double)(j + (*ID) * nj) - 0.5);
        const double lz = delta * ((double)k - 0.5);
        if (*area_pre1 == 3 && lx < s1) lx = lx + xl;
        if (*area_pre1 == 1 && lx > s2) lx = lx - xl;
        lx = lx + xl * (double)(*lap);
        wx = wx + p * lx;
        wy = wy + p * ly;
        wz = wz + p * lz;
      }
  center[0] = wx / weight;
  center[1] = wy / weight;
  center[2] = wz / weight;
  *velocity = (center[0] - *center_pre1) / (*dt_);
  *center_pre2 = *center_pre1;
  *center_pre1 = center[0];
  double cc = center[0];
  *area_pre2 = *area_pre1;
  if (cc > xl) {
    do { cc = cc - xl; } while (!(cc < xl));
  }
  if (cc < s1) *area_pre1 = 1;
  if (cc > s2) *area_pre1 = 3;
  if (s1 <= cc && cc <= s2) *area_pre1 = 2;
}

/* srcs/cal_vel.f:1-83: phase-weighted mean velocity and Reynolds numbers.
 * Unbuildable reference (MPI header); pinned by a uniform-field closed form. */
void orc_cal_vel(const int *ipara, const int *ni_, const int *nj_, const int *nk_,
                 const double *radius, const double *rhol, const double *rmul, const double *u,
                 const double *v, const double *w, const double *phi, double *vel, double *re) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)ipara;
  double sp = 0.0, pu = 0.0, pv = 0.0, pw = 0.0;
  const double diameter = *radius * 2.0;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const ptrdiff_t m = OIX(i, j, k);
        sp = sp + phi[m];
        pu = pu + phi[m] * u[m];
        pv = pv + phi[m] * v[m];
        pw = pw + phi[m] * w[m];
      }
  vel[0] = pu / sp;
  vel[1] = pv / sp;
  vel[2] = pw / sp;
  for (int c = 0; c < 3; c++) re[c] = *rhol * vel[c] * diameter / *rmul;
}

/* srcs/caldt.f:1-116: CFL time step from the largest advective rate
 * |u|/dx, |v|/dy (v on j <= jen: nj-1 below a wall at Y_PLUS), |w|/dz over
 * the interior (floor 1e-20), the viscous limit cfl*0.5*dcell^2/nu*100 and
 * the capillary limit cfl*sqrt(dcell^3/(2 sigma/(rho_l+rho_g))), with
 * dcell = (dxinv*dyinv*dzinv)^(-1/3) and nu the larger kinematic
 * viscosity; dt = min of the three, time += dt.  The reference prints a
 * monitor line every imon_t steps on ID 0.  Unbuildable reference (MPI
 * header); pinned by closed forms (each limit selected in turn) and a
 * numpy brute force in tests/test_oracle_math.py.  Returns dt1..dt3 in
 * lim[0..2] when lim is not NULL. */
void orc_caldt(const int *ipara, const int *nID, const int *ID, const int *ndiv, const int *ni_,
               const int *nj_, const int *nk_, const int *nstep, const int *imon_t,
               const double *dxinv, const double *dyinv, const double *dzinv, const double *cfl,
               const double *rhol, const double *rhog, const double *rmul, const double *rmug,
               const double *surface_tension, const double *u, const double *v, const double *w,
               double *dt, double *time, double *lim) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)ipara; (void)ndiv;
  const double dcell = pow(*dxinv * *dyinv * *dzinv, -1.0 / 3.0);
  const double rnu = fmax(*rmul / *rhol, *rmug / *rhog);
  const double st_rhominv = 2.0 * *surface_tension / (*rhol + *rhog);
  const int jen = nID[4] < 0 ? nj - 1 : nj;
  double advmax = 1.0e-20;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) advmax = fmax(fabs(u[OIX(i, j, k)] * *dxinv), advmax);
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= jen; j++)
      for (int i = 1; i <= ni; i++) advmax = fmax(fabs(v[OIX(i, j, k)] * *dyinv), advmax);
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) advmax = fmax(fabs(w[OIX(i, j, k)] * *dzinv), advmax);
  const double dt1 = *cfl / advmax;
  const double dt2 = *cfl * 0.5 * (dcell * dcell) / rnu * 1.0e2;
  const double dt3 = *cfl * sqrt((dcell * dcell * dcell) / st_rhominv);
  double d = fmin(dt1, dt2);
  d = fmin(d, dt3);
  *dt = d;
  *time = *time + d;
  if (lim) {
    lim[0] = dt1;
    lim[1] = dt2;
    lim[2] = dt3;
  }
  if (*imon_t > 0 && *nstep % *imon_t == 0 && *ID == 0)
    printf("time=%17.10E dt=%17.10E dt_adv=%17.10E dt_vis=%17.10E dt_srf=%17.10E\n", *time, d,
           dt1, dt2, dt3);
}
