/*
 * fel_oracle_adj.c -- CPU restatement (plain C, FP64, serial) of the adjoints
 * (reverse-mode derivatives) of the linear routines of the velocity-pressure
 * update of toruneko1129/fel-multi-phase-flow.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library's *_b_ entry
 * points, never part of the product.
 *
 * The reference has no adjoint.  orc_<name>_b restates the transpose of the
 * linear map that the forward routine orc_<name> (fel_oracle.c, itself pinned
 * to the compiled reference) applies to its active arrays: every active array
 * argument of the forward list is replaced by its adjoint, of the same shape;
 * arrays that only parameterise the map (densities, coefficients, slip
 * lengths, the volume fraction) are passed as in the forward.  Adjoints of
 * the forward's inputs are incremented; adjoints of what the forward
 * overwrites are consumed (set to zero); in/out arrays are updated in place.
 *
 * Each routine is written as the forward's assignments taken in reverse
 * order, one scatter per assignment ("x = a*y + b*z" becomes y_b += a*x_b,
 * z_b += b*x_b, x_b = 0), so it is the exact transpose of the forward's
 * sequence.  pinned: the dot-product identity <J v, w> = <v, J^T w> against
 * the forward oracle on seeded random v, w (tests/test_oracle_adjoint.py);
 * any dropped term, wrong index or transposed coefficient breaks it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fel_oracle.h"
#include "fel_oracle_api.h"

/* srcs/caldiv.f:27-38: div(c) = ((u(c)-u(c-x))dxinv + (v(c)-v(c-y))dyinv +
 * (w(c)-w(c-z))dzinv)/dt on [-1,n+3]^3.  The RMS output (caldiv.f:43-67) is
 * a diagnostic and has no adjoint. */
void orc_caldiv_b(const int *ipara, const int *ndiv, const int *ni_, const int *nj_, const int *nk_,
                  const double *dxinv_, const double *dyinv_, const double *dzinv_,
                  const double *dt_, double *ukb, double *vkb, double *wkb, double *divb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)ipara; (void)ndiv;
  const double dtinv = 1.0 / (*dt_);
  const double fx = *dxinv_ * dtinv, fy = *dyinv_ * dtinv, fz = *dzinv_ * dtinv;
  for (int k = nk + 3; k >= -1; k--)
    for (int j = nj + 3; j >= -1; j--)
      for (int i = ni + 3; i >= -1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        const double g = divb[c];
        ukb[c] += g * fx;
        ukb[OIX(i - 1, j, k)] -= g * fx;
        vkb[c] += g * fy;
        vkb[OIX(i, j - 1, k)] -= g * fy;
        wkb[c] += g * fz;
        wkb[OIX(i, j, k - 1)] -= g * fz;
        divb[c] = 0.0;
      }
}

/* srcs/solp_fft_tdma1.f:25-47: phir = W div(interior) / (ni nk), W the
 * forward DFT (exp(-2 pi i ...)) of each j-plane at centred mode indices.
 * Transpose: div_b(interior) += Re(W^H (phir_rb + i phir_ib)) / (ni nk),
 * i.e. the unnormalised inverse transform solp_fft_tdma3 applies, divided by
 * ni nk (a power of two: exact). */
void orc_solp_fft_tdma1_b(const int *ni_, const int *nj_, const int *nk_, double *fftdata,
                          double *divb, double *phir_rb, double *phir_ib) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  const size_t nm = (size_t)ni * nj * nk;
  double *t = calloc((size_t)n3, sizeof(double));
  orc_solp_fft_tdma3(ni_, nj_, nk_, fftdata, phir_rb, phir_ib, t);
  const double nrm = (double)(ni * nk);
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) divb[OIX(i, j, k)] += t[OIX(i, j, k)] / nrm;
  memset(phir_rb, 0, sizeof(double) * nm);
  memset(phir_ib, 0, sizeof(double) * nm);
  free(t);
}

/* srcs/trans_r2w.f / trans_w2r.f on one rank: a copy w = r (none when the
 * arrays coincide).  Transpose: r_b += w_b, w_b = 0. */
void orc_trans_r2w_b(const int *ipara, const int *ID, const int *ndiv, const int *svall,
                     const int *key, double *sendbuf, double *recvbuf, double *rb, double *wb) {
  (void)ipara; (void)ID; (void)ndiv; (void)key; (void)sendbuf; (void)recvbuf;
  if (rb == wb) return;
  const size_t n = (size_t)svall[0] * svall[1] * svall[2];
  for (size_t m = 0; m < n; m++) {
    rb[m] += wb[m];
    wb[m] = 0.0;
  }
}

void orc_trans_w2r_b(const int *ipara, const int *ID, const int *ndiv, const int *svall,
                     const int *key, double *sendbuf, double *recvbuf, double *wb, double *rb) {
  orc_trans_r2w_b(ipara, ID, ndiv, svall, key, sendbuf, recvbuf, wb, rb);
}

/* srcs/solp_fft_tdma2.f:46-87: x = T^-1 b per (kx,kz) mode, T tridiagonal
 * along j with sub-diagonal as(j), super-diagonal an(j) and diagonal
 * -(ap0(j) + dx2r (2 - c1x) + dz2r (2 - c1z)).  Transpose: the Thomas solve
 * of T^T (sub-diagonal an(j-1), super-diagonal as(j+1)), in place on the
 * adjoints of the real and imaginary parts.  The singular (0,0) mode (c1x =
 * c1z = 2), whose last pivot is zero in exact arithmetic, is solved pinned:
 * x(nj) = 0 and the leading nj-1 rows, which is the transpose of the pinned
 * solve and agrees with any other solution of a compatible system up to the
 * constant the mean removal after the solve takes out. */
void orc_solp_fft_tdma2_b(const int *ID, const int *ndiv, const int *svall, const double *rho_,
                          const double *dxinv_, const double *dzinv_, const double *as_p_fft,
                          const double *an_p_fft, const double *ap_p_fft, double *atdma,
                          double *btdma_r, double *btdma_i, double *phiw_rb, double *phiw_ib) {
  const int ni = svall[0], nj = svall[1], nk = svall[2] / (*ndiv);
  const double rho = *rho_, dxinv = *dxinv_, dzinv = *dzinv_;
  const double pi = atan(1.0) * 4.0;
  const double dx2r = dxinv * dxinv / rho, dz2r = dzinv * dzinv / rho;
  const double pi2_ni = pi * 2.0 / (double)svall[0], pi2_nk = pi * 2.0 / (double)svall[2];
  double *a = atdma + 2, *br = btdma_r + 2, *bi = btdma_i + 2;
  const double *as = as_p_fft + 2, *an = an_p_fft + 2, *ap0 = ap_p_fft + 2;
  for (int k = 1; k <= nk; k++)
    for (int i = -ni / 2; i <= ni / 2 - 1; i++) {
      const double c1x = cos((double)i * pi2_ni) * 2.0;
      const double c1z = cos((double)(k + (*ID) * nk - svall[2] / 2 - 1) * pi2_nk) * 2.0;
      const int singular = c1x == 2.0 && c1z == 2.0;
      a[0] = 0.0; br[0] = 0.0; bi[0] = 0.0;
      for (int j = 1; j <= nj; j++) {
        const size_t o = (size_t)(i + ni / 2) + (size_t)ni * ((size_t)(j - 1) + (size_t)nj * (size_t)(k - 1));
        const double sub = j > 1 ? an[j - 1] : 0.0;
        const double sup = j < nj ? as[j + 1] : 0.0;
        const double ap = ap0[j] + dx2r * (2.0 - c1x) + dz2r * (2.0 - c1z);
        const double d00 = -ap + sub * a[j - 1];
        double apinv = -d00 / (d00 * d00 + 1.0e-99);
        if (singular && j == nj) apinv = 0.0;
        a[j] = apinv * sup;
        br[j] = apinv * (-phiw_rb[o] + sub * br[j - 1]);
        bi[j] = apinv * (-phiw_ib[o] + sub * bi[j - 1]);
      }
      {
        const size_t o = (size_t)(i + ni / 2) + (size_t)ni * ((size_t)(nj - 1) + (size_t)nj * (size_t)(k - 1));
        phiw_rb[o] = br[nj];
        phiw_ib[o] = bi[nj];
      }
      for (int j = nj - 1; j >= 1; j--) {
        const size_t o = (size_t)(i + ni / 2) + (size_t)ni * ((size_t)(j - 1) + (size_t)nj * (size_t)(k - 1));
        phiw_rb[o] = br[j] + a[j] * phiw_rb[o + ni];
        phiw_ib[o] = bi[j] + a[j] * phiw_ib[o + ni];
      }
    }
}

/* srcs/solp_fft_tdma3.f:25-44: dp(interior) = Re(W^H phir), the
 * unnormalised inverse DFT.  Transpose: phir_rb + i phir_ib += W dp_b
 * (the unnormalised forward DFT: solp_fft_tdma1's transform times ni nk,
 * exact), dp_b(interior) = 0. */
void orc_solp_fft_tdma3_b(const int *ni_, const int *nj_, const int *nk_, double *fftdata,
                          double *phir_rb, double *phir_ib, double *dpb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  const size_t nm = (size_t)ni * nj * nk;
  double *tr = malloc(sizeof(double) * nm), *ti = malloc(sizeof(double) * nm);
  orc_solp_fft_tdma1(ni_, nj_, nk_, fftdata, dpb, tr, ti);
  const double nrm = (double)(ni * nk);
  for (size_t m = 0; m < nm; m++) {
    phir_rb[m] += tr[m] * nrm;
    phir_ib[m] += ti[m] * nrm;
  }
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) dpb[OIX(i, j, k)] = 0.0;
  free(tr);
  free(ti);
}

/* srcs/bnd_avzero.f:10-23: q -= mean(q over the interior) on the whole
 * padded array.  Transpose: q_b(interior) -= (sum of q_b over the padded
 * array) / (ni nj nk ndiv). */
void orc_bnd_avzero_b(const int *ipara, const int *ndiv, const int *ni_, const int *nj_,
                      const int *nk_, double *qb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)ipara;
  double s = 0.0;
  for (ptrdiff_t m = 0; m < n3; m++) s = s + qb[m];
  const double av = s / (double)(ni * nj * nk * (*ndiv));
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) qb[OIX(i, j, k)] -= av;
}

/* srcs/bnd_neumann.f:14-41: mirror ghosts q(i,-2..0,k) = q(i,3..1,k) and
 * q(i,nj+1..nj+3,k) = q(i,nj..nj-2,k).  Transpose: each ghost's adjoint is
 * added to its source and cleared (top wall first: reverse order). */
void orc_bnd_neumann_b(const int *nID, const int *ni_, const int *nj_, const int *nk_, double *qb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  if (nID[ORC_Y_PLUS - 1] < 0)
    for (int k = nk + 3; k >= -2; k--)
      for (int i = ni + 3; i >= -2; i--)
        for (int g = 3; g >= 1; g--) {
          qb[OIX(i, nj + 1 - g, k)] += qb[OIX(i, nj + g, k)];
          qb[OIX(i, nj + g, k)] = 0.0;
        }
  if (nID[ORC_Y_MINUS - 1] < 0)
    for (int k = nk + 3; k >= -2; k--)
      for (int i = ni + 3; i >= -2; i--)
        for (int g = 0; g >= -2; g--) {
          /* forward order: -2 <- 3, -1 <- 2, 0 <- 1; reversed */
          qb[OIX(i, 1 - g, k)] += qb[OIX(i, g, k)];
          qb[OIX(i, g, k)] = 0.0;
        }
}

/* srcs/bnd_periodic.f:14-43: x ghosts of every (j,k) line, then z ghosts of
 * every (i,j) line (the z pass copies the x ghosts too).  Transpose: the z
 * pass reversed, then the x pass reversed. */
void orc_bnd_periodic_b(const int *ni_, const int *nj_, const int *nk_, double *qb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  for (int j = nj + 3; j >= -2; j--)
    for (int i = ni + 3; i >= -2; i--) {
      const int gk[6] = {-2, -1, 0, nk + 1, nk + 2, nk + 3};
      const int sk[6] = {nk - 2, nk - 1, nk, 1, 2, 3};
      for (int m = 5; m >= 0; m--) {
        qb[OIX(i, j, sk[m])] += qb[OIX(i, j, gk[m])];
        qb[OIX(i, j, gk[m])] = 0.0;
      }
    }
  for (int k = nk + 3; k >= -2; k--)
    for (int j = nj + 3; j >= -2; j--) {
      const int gi[6] = {-2, -1, 0, ni + 1, ni + 2, ni + 3};
      const int si[6] = {ni - 2, ni - 1, ni, 1, 2, 3};
      for (int m = 5; m >= 0; m--) {
        qb[OIX(si[m], j, k)] += qb[OIX(gi[m], j, k)];
        qb[OIX(gi[m], j, k)] = 0.0;
      }
    }
}

/* srcs/corunp_explicit.f:26-74: u += (dp(c) - dp(c+x))/rho dxinv dt on the
 * interior (v to jen), pn = phat + dp on the interior.  Transpose: the
 * velocity adjoints pass through; dp_b and phat_b gather them; pn_b is
 * consumed on the interior. */
void orc_corunp_explicit_b(const int *nID, const int *ni_, const int *nj_, const int *nk_,
                           const double *rho_, const double *dxinv_, const double *dyinv_,
                           const double *dzinv_, const double *dt_, double *dpb, double *phatb,
                           double *unb, double *vnb, double *wnb, double *pnb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double rho = *rho_, dt = *dt_;
  ODIMS(ni, nj, nk);
  const int jen = (nID[ORC_Y_PLUS - 1] < 0) ? nj - 1 : nj;
  const double fx = 1.0 / rho * *dxinv_ * dt, fy = 1.0 / rho * *dyinv_ * dt,
               fz = 1.0 / rho * *dzinv_ * dt;
  for (int k = nk; k >= 1; k--)
    for (int j = nj; j >= 1; j--)
      for (int i = ni; i >= 1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        phatb[c] += pnb[c];
        dpb[c] += pnb[c];
        pnb[c] = 0.0;
      }
  for (int k = nk; k >= 1; k--)
    for (int j = nj; j >= 1; j--)
      for (int i = ni; i >= 1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        dpb[c] += wnb[c] * fz;
        dpb[OIX(i, j, k + 1)] -= wnb[c] * fz;
      }
  for (int k = nk; k >= 1; k--)
    for (int j = jen; j >= 1; j--)
      for (int i = ni; i >= 1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        dpb[c] += vnb[c] * fy;
        dpb[OIX(i, j + 1, k)] -= vnb[c] * fy;
      }
  for (int k = nk; k >= 1; k--)
    for (int j = nj; j >= 1; j--)
      for (int i = ni; i >= 1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        dpb[c] += unb[c] * fx;
        dpb[OIX(i + 1, j, k)] -= unb[c] * fx;
      }
}

/* Navier-slip coefficients of srcs/bndu.f:39-47 (73-81) at wall row jw. */
static void slip_coef(ptrdiff_t sx, ptrdiff_t sy, int ni, int i, int jw, int k, double dy,
                      const double *l1, const double *l2, const double *phi, double *c1,
                      double *c2) {
  const double phi_av = (phi[OIX(i, jw, k)] + phi[OIX((i + 1) % (ni + 3), jw, k)]) / 2.0;
  const double inv1 = phi_av / l1[OIX(i, jw, k)];
  const double inv2 = (1.0 - phi_av) / l2[OIX(i, jw, k)];
  const double ls = 1.0 / (inv1 + inv2);
  *c1 = (2.0 * dy) / (2.0 * ls + dy);
  *c2 = (2.0 * ls - dy) / (2.0 * ls + dy);
}

/* srcs/bndu.f:29-158 (orc_bndu): the j-wall rows, then the x fills, then
 * the z fills, each written as ghost = function of earlier values.
 * Transpose in reverse: z fills, x fills, wall rows; uwall_b gathers the
 * wall rows' -c1 (bottom) and +c1 (top) weights.  The slip lengths and phi
 * parameterise the map; their adjoints are not formed. */
void orc_bndu_b(const int *nID, const int *ni_, const int *nj_, const int *nk_, double *ukb,
                double *vkb, double *wkb, double *uwallb, const double *dy_, const double *l1,
                const double *l2, const double *phi) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dy = *dy_;
  ODIMS(ni, nj, nk);
  /* z fills (bndu.f:130-155), reversed */
  for (int i = ni + 3; i >= -2; i--)
    for (int j = nj + 3; j >= -2; j--) {
#define ZB(q, g, s)                                        \
  do {                                                     \
    q[OIX(i, j, s)] += q[OIX(i, j, g)];                    \
    q[OIX(i, j, g)] = 0.0;                                 \
  } while (0)
      ZB(wkb, nk + 3, 3); ZB(wkb, nk + 2, 2);
      ZB(vkb, nk + 3, 3); ZB(vkb, nk + 2, 2); ZB(vkb, nk + 1, 1);
      ZB(ukb, nk + 3, 3); ZB(ukb, nk + 2, 2); ZB(ukb, nk + 1, 1);
      ZB(wkb, -2, nk - 1); ZB(wkb, -1, nk); ZB(wkb, 0, nk + 1);
      ZB(vkb, -2, nk - 2); ZB(vkb, -1, nk - 1); ZB(vkb, 0, nk);
      ZB(ukb, -2, nk - 2); ZB(ukb, -1, nk - 1); ZB(ukb, 0, nk);
#undef ZB
    }
  /* x fills (bndu.f:104-127), reversed */
  for (int j = nj + 3; j >= -2; j--)
    for (int k = nk + 3; k >= -2; k--) {
#define XB(q, g, s)                                        \
  do {                                                     \
    q[OIX(s, j, k)] += q[OIX(g, j, k)];                    \
    q[OIX(g, j, k)] = 0.0;                                 \
  } while (0)
      XB(wkb, ni + 3, 3); XB(wkb, ni + 2, 2); XB(wkb, ni + 1, 1);
      XB(vkb, ni + 3, 3); XB(vkb, ni + 2, 2); XB(vkb, ni + 1, 1);
      XB(ukb, ni + 3, 3); XB(ukb, ni + 2, 2);
      XB(wkb, -2, ni - 2); XB(wkb, -1, ni - 1); XB(wkb, 0, ni);
      XB(vkb, -2, ni - 2); XB(vkb, -1, ni - 1); XB(vkb, 0, ni);
      XB(ukb, -2, ni - 1); XB(ukb, -1, ni); XB(ukb, 0, ni + 1);
#undef XB
    }
  /* wall rows (bndu.f:29-98), top wall first (reverse order) */
  double ub = 0.0;
  if (nID[ORC_Y_PLUS - 1] < 0)
    for (int k = nk + 3; k >= -2; k--)
      for (int i = ni + 3; i >= -2; i--) {
        double c1, c2;
        slip_coef(sx, sy, ni, i, nj, k, dy, l1, l2, phi, &c1, &c2);
        double *q = wkb;
        q[OIX(i, nj + 2, k)] += 2.0 * q[OIX(i, nj + 3, k)];
        q[OIX(i, nj + 1, k)] -= q[OIX(i, nj + 3, k)];
        q[OIX(i, nj + 3, k)] = 0.0;
        q[OIX(i, nj + 1, k)] += 2.0 * q[OIX(i, nj + 2, k)];
        q[OIX(i, nj, k)] -= q[OIX(i, nj + 2, k)];
        q[OIX(i, nj + 2, k)] = 0.0;
        q[OIX(i, nj, k)] += c2 * q[OIX(i, nj + 1, k)];
        q[OIX(i, nj + 1, k)] = 0.0;
        q = vkb;
        q[OIX(i, nj - 3, k)] -= q[OIX(i, nj + 3, k)];
        q[OIX(i, nj + 3, k)] = 0.0;
        q[OIX(i, nj - 2, k)] -= q[OIX(i, nj + 2, k)];
        q[OIX(i, nj + 2, k)] = 0.0;
        q[OIX(i, nj - 1, k)] -= q[OIX(i, nj + 1, k)];
        q[OIX(i, nj + 1, k)] = 0.0;
        q[OIX(i, nj, k)] = 0.0;
        q = ukb;
        q[OIX(i, nj + 2, k)] += 2.0 * q[OIX(i, nj + 3, k)];
        q[OIX(i, nj + 1, k)] -= q[OIX(i, nj + 3, k)];
        q[OIX(i, nj + 3, k)] = 0.0;
        q[OIX(i, nj + 1, k)] += 2.0 * q[OIX(i, nj + 2, k)];
        q[OIX(i, nj, k)] -= q[OIX(i, nj + 2, k)];
        q[OIX(i, nj + 2, k)] = 0.0;
        q[OIX(i, nj, k)] += c2 * q[OIX(i, nj + 1, k)];
        ub += c1 * q[OIX(i, nj + 1, k)];
        q[OIX(i, nj + 1, k)] = 0.0;
      }
  if (nID[ORC_Y_MINUS - 1] < 0)
    for (int k = nk + 3; k >= -2; k--)
      for (int i = ni + 3; i >= -2; i--) {
        double c1, c2;
        slip_coef(sx, sy, ni, i, 1, k, dy, l1, l2, phi, &c1, &c2);
        double *q = wkb;
        q[OIX(i, -1, k)] += 2.0 * q[OIX(i, -2, k)];
        q[OIX(i, 0, k)] -= q[OIX(i, -2, k)];
        q[OIX(i, -2, k)] = 0.0;
        q[OIX(i, 0, k)] += 2.0 * q[OIX(i, -1, k)];
        q[OIX(i, 1, k)] -= q[OIX(i, -1, k)];
        q[OIX(i, -1, k)] = 0.0;
        q[OIX(i, 1, k)] += c2 * q[OIX(i, 0, k)];
        q[OIX(i, 0, k)] = 0.0;
        q = vkb;
        q[OIX(i, 0, k)] = 0.0;
        q[OIX(i, 1, k)] -= q[OIX(i, -1, k)];
        q[OIX(i, -1, k)] = 0.0;
        q[OIX(i, 2, k)] -= q[OIX(i, -2, k)];
        q[OIX(i, -2, k)] = 0.0;
        q = ukb;
        q[OIX(i, -1, k)] += 2.0 * q[OIX(i, -2, k)];
        q[OIX(i, 0, k)] -= q[OIX(i, -2, k)];
        q[OIX(i, -2, k)] = 0.0;
        q[OIX(i, 0, k)] += 2.0 * q[OIX(i, -1, k)];
        q[OIX(i, 1, k)] -= q[OIX(i, -1, k)];
        q[OIX(i, -1, k)] = 0.0;
        q[OIX(i, 1, k)] += c2 * q[OIX(i, 0, k)];
        ub -= c1 * q[OIX(i, 0, k)];
        q[OIX(i, 0, k)] = 0.0;
      }
  *uwallb += ub;
}

/* One colour's update of component q, reversed (srcs/solu_sor4.f:131-175):
 * x_new = (0.5 S/r00 + src)/den with S = sum of a_m * neighbour_m and den =
 * dtinv + a_p/r00*0.5, for every cell x of the colour in the component's
 * range.  Transpose: src_b(x) += x_b/den; neighbour_b += a_m * 0.5/(r00
 * den) * x_b; x_b = 0 (the old value of x does not enter its update). */
static const int kIsor[8] = {1, 2, 1, 2, 1, 2, 2, 1};
static const int kJsor[8] = {1, 2, 1, 2, 2, 1, 1, 2};
static const int kKsor[8] = {1, 2, 2, 1, 1, 2, 1, 2};

static void colour_b(int q, int col, int ni, int nj, int nk, int jen, double dtinv,
                     const double *rho, const double *rhon, double *const *cf, double *srcb,
                     double *ub, double *vb, double *wb) {
  ODIMS(ni, nj, nk);
  const ptrdiff_t X = 1, Y = sx, Z = sx * sy;
  const int jend = q == 1 ? jen : nj;
  const ptrdiff_t F = q == 0 ? X : (q == 1 ? Y : Z);
  double *const *c = cf + 15 * q;
  double *self = q == 0 ? ub : (q == 1 ? vb : wb);
  for (int k = kKsor[col]; k <= nk; k += 2)
    for (int j = kJsor[col]; j <= jend; j += 2)
      for (int i = kIsor[col]; i <= ni; i += 2) {
        const ptrdiff_t m = OIX(i, j, k);
        const double r00 = (rho[m] + rho[m + F] + rhon[m] + rhon[m + F]) * 0.25;
        const double den = dtinv + c[6][m] / r00 * 0.5;
        const double g = self[m] / den;
        srcb[m] += g;
        const double h = g * 0.5 / r00;
        self[m] = 0.0;
        /* own component: w, e, s, n, b, t */
        self[m - X] += c[0][m] * h;
        self[m + X] += c[1][m] * h;
        self[m - Y] += c[2][m] * h;
        self[m + Y] += c[3][m] * h;
        self[m - Z] += c[4][m] * h;
        self[m + Z] += c[5][m] * h;
        if (q == 0) {
          vb[m - Y] += c[7][m] * h;
          vb[m + X - Y] += c[8][m] * h;
          vb[m] += c[9][m] * h;
          vb[m + X] += c[10][m] * h;
          wb[m - Z] += c[11][m] * h;
          wb[m + X - Z] += c[12][m] * h;
          wb[m] += c[13][m] * h;
          wb[m + X] += c[14][m] * h;
        } else if (q == 1) {
          ub[m - X] += c[7][m] * h;
          ub[m - X + Y] += c[8][m] * h;
          ub[m] += c[9][m] * h;
          ub[m + Y] += c[10][m] * h;
          wb[m - Z] += c[11][m] * h;
          wb[m + Y - Z] += c[12][m] * h;
          wb[m] += c[13][m] * h;
          wb[m + Y] += c[14][m] * h;
        } else {
          ub[m - X] += c[7][m] * h;
          ub[m - X + Z] += c[8][m] * h;
          ub[m] += c[9][m] * h;
          ub[m + Z] += c[10][m] * h;
          vb[m - Y] += c[11][m] * h;
          vb[m - Y + Z] += c[12][m] * h;
          vb[m] += c[13][m] * h;
          vb[m + Z] += c[14][m] * h;
        }
      }
}

/* srcs/solu_sor4.f:131-293 with the sweep count of the forward call frozen:
 * `lp` is the forward's loop variable on exit (felgpu_sor_last_ /
 * orc_sor_last_lp), so min(lp+1, 21) sweeps are reversed.  For fixed
 * densities and coefficients the sweeps are affine in (un, vn, wn, src_*,
 * uwall); this is their transpose: for each sweep from the last, the
 * components w, v, u, each colour from the eighth, the refresh after every
 * second colour (bndu, then bnd_periodic of the component) reversed before
 * the colour's update.  Adjoint arguments: src_*b incremented, un/vn/wnb in
 * place (the adjoint of the final state in, of the initial state out),
 * *uwallb incremented. */
static void sor4_b_impl(const int *nID, const int *ni_, const int *nj_, const int *nk_,
                        const double *dt_, const double *rho, const double *rhon,
                        double *const *coef, double *src_ub, double *src_vb, double *src_wb,
                        double *unb, double *vnb, double *wnb, double *uwallb, const double *dy,
                        const double *l1, const double *l2, const double *phi, const int *lp) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dtinv = 1.0 / (*dt_);
  const int jen = (nID[ORC_Y_PLUS - 1] < 0) ? nj - 1 : nj;
  const int nsw = *lp < 21 ? *lp + 1 : 21;
  double *srcb[3] = {src_ub, src_vb, src_wb};
  double *fld[3] = {unb, vnb, wnb};
  for (int s = nsw - 1; s >= 0; s--)
    for (int q = 2; q >= 0; q--)
      for (int col = 7; col >= 0; col--) {
        if (col % 2 == 1) {
          orc_bnd_periodic_b(ni_, nj_, nk_, fld[q]);
          orc_bndu_b(nID, ni_, nj_, nk_, unb, vnb, wnb, uwallb, dy, l1, l2, phi);
        }
        colour_b(q, col, ni, nj, nk, jen, dtinv, rho, rhon, coef, srcb[q], unb, vnb, wnb);
      }
}

/* Argument-list form: srcs/solu_sor4.f:1-17 with the active arrays replaced
 * by their adjoints, *uwallb after the velocities, and lp appended. */
void orc_solu_sor4_b(const int *ipara, const int *ID, const int *nID, const int *ndiv,
                     const int *ni, const int *nj, const int *nk, const int *key, double *sendjb,
                     double *recvjb, const int *nstep, const int *imon_t, const double *dt,
                     const double *rho, const double *rhon, double *c0, double *c1, double *c2,
                     double *c3, double *c4, double *c5, double *c6, double *c7, double *c8,
                     double *c9, double *c10, double *c11, double *c12, double *c13, double *c14,
                     double *c15, double *c16, double *c17, double *c18, double *c19, double *c20,
                     double *c21, double *c22, double *c23, double *c24, double *c25, double *c26,
                     double *c27, double *c28, double *c29, double *c30, double *c31, double *c32,
                     double *c33, double *c34, double *c35, double *c36, double *c37, double *c38,
                     double *c39, double *c40, double *c41, double *c42, double *c43, double *c44,
                     double *src_ub, double *src_vb, double *src_wb, double *unb, double *vnb,
                     double *wnb, double *uwallb, const double *dy, const double *l1,
                     const double *l2, const double *phi, const int *lp) {
  (void)ipara; (void)ID; (void)ndiv; (void)key; (void)sendjb; (void)recvjb; (void)nstep; (void)imon_t;
  double *const coef[NCOEF] = {c0, c1, c2, c3, c4, c5, c6, c7, c8, c9, c10, c11, c12, c13, c14,
                               c15, c16, c17, c18, c19, c20, c21, c22, c23, c24, c25, c26, c27,
                               c28, c29, c30, c31, c32, c33, c34, c35, c36, c37, c38, c39, c40,
                               c41, c42, c43, c44};
  sor4_b_impl(nID, ni, nj, nk, dt, rho, rhon, coef, src_ub, src_vb, src_wb, unb, vnb, wnb, uwallb,
              dy, l1, l2, phi, lp);
}

/* ------------------------------------------------------------------------- */
/* Explicit momentum terms                                                   */
/* ------------------------------------------------------------------------- */

/* srcs/calphat.f:16-22: phat = 2p - po on i in [-1,ni+3], j, k in
 * [-2,n+3].  Transpose: p_b += 2 phat_b, po_b -= phat_b, phat_b = 0 there. */
void orc_calphat_b(const int *ni_, const int *nj_, const int *nk_, double *pb, double *pob,
                   double *phatb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  for (int k = nk + 3; k >= -2; k--)
    for (int j = nj + 3; j >= -2; j--)
      for (int i = ni + 3; i >= -1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        pb[c] += phatb[c] * 2.0;
        pob[c] -= phatb[c];
        phatb[c] = 0.0;
      }
}

/* srcs/cal_grad_prs.f:24-37: q_x(c) = (q(c+x) - q(c)) dxinv (likewise y, z)
 * on [-2,n+2]^3. */
void orc_cal_grad_prs_b(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                        const double *dyinv_, const double *dzinv_, double *qb, double *qx_ub,
                        double *qy_vb, double *qz_wb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  for (int k = nk + 2; k >= -2; k--)
    for (int j = nj + 2; j >= -2; j--)
      for (int i = ni + 2; i >= -2; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        qb[c] -= qx_ub[c] * dxinv + qy_vb[c] * dyinv + qz_wb[c] * dzinv;
        qb[OIX(i + 1, j, k)] += qx_ub[c] * dxinv;
        qb[OIX(i, j + 1, k)] += qy_vb[c] * dyinv;
        qb[OIX(i, j, k + 1)] += qz_wb[c] * dzinv;
        qx_ub[c] = 0.0;
        qy_vb[c] = 0.0;
        qz_wb[c] = 0.0;
      }
}

/* srcs/calsij.f:28-139: the six strain-rate components on their boxes
 * (orc_calsij), transposed component by component in reverse order. */
void orc_calsij_b(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                  const double *dyinv_, const double *dzinv_, double *ukb, double *vkb,
                  double *wkb, double *skb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  double *s1 = skb, *s2 = skb + n3, *s3 = skb + 2 * n3, *s4 = skb + 3 * n3, *s5 = skb + 4 * n3,
         *s6 = skb + 5 * n3;
  for (int k = nk + 2; k >= -2; k--)
    for (int j = nj + 2; j >= -2; j--)
      for (int i = ni + 3; i >= -2; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        const double h = s6[c] * 0.5;
        vkb[c] -= h * dzinv;
        vkb[OIX(i, j, k + 1)] += h * dzinv;
        wkb[c] -= h * dyinv;
        wkb[OIX(i, j + 1, k)] += h * dyinv;
        s6[c] = 0.0;
      }
  for (int k = nk + 2; k >= -2; k--)
    for (int j = nj + 3; j >= -2; j--)
      for (int i = ni + 2; i >= -2; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        const double h = s5[c] * 0.5;
        ukb[c] -= h * dzinv;
        ukb[OIX(i, j, k + 1)] += h * dzinv;
        wkb[c] -= h * dxinv;
        wkb[OIX(i + 1, j, k)] += h * dxinv;
        s5[c] = 0.0;
      }
  for (int k = nk + 3; k >= -2; k--)
    for (int j = nj + 2; j >= -2; j--)
      for (int i = ni + 2; i >= -2; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        const double h = s4[c] * 0.5;
        ukb[c] -= h * dyinv;
        ukb[OIX(i, j + 1, k)] += h * dyinv;
        vkb[c] -= h * dxinv;
        vkb[OIX(i + 1, j, k)] += h * dxinv;
        s4[c] = 0.0;
      }
  for (int k = nk + 3; k >= -1; k--)
    for (int j = nj + 3; j >= -1; j--)
      for (int i = ni + 3; i >= -1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        wkb[OIX(i, j, k - 1)] -= s3[c] * dzinv;
        wkb[c] += s3[c] * dzinv;
        s3[c] = 0.0;
      }
  for (int k = nk + 3; k >= -2; k--)
    for (int j = nj + 3; j >= -1; j--)
      for (int i = ni + 3; i >= -2; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        vkb[OIX(i, j - 1, k)] -= s2[c] * dyinv;
        vkb[c] += s2[c] * dyinv;
        s2[c] = 0.0;
      }
  for (int k = nk + 3; k >= -2; k--)
    for (int j = nj + 3; j >= -2; j--)
      for (int i = ni + 3; i >= -1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        ukb[OIX(i - 1, j, k)] -= s1[c] * dxinv;
        ukb[c] += s1[c] * dxinv;
        s1[c] = 0.0;
      }
}

/* srcs/cal_arith_tau.f:21-118: tau_l = 2 mu s_l (l = 1..3, every cell),
 * tau_l = 2 mu_edge s_l on the shear boxes (l = 4..6), mu fixed. */
void orc_cal_arith_tau_b(const int *ni_, const int *nj_, const int *nk_, double *sb,
                         const double *rmu, double *taub) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  for (int k = nk + 2; k >= -2; k--)
    for (int j = nj + 2; j >= -2; j--)
      for (int i = ni + 3; i >= -2; i--) {
        const double m = +0.25 * rmu[OIX(i, j, k)] + 0.25 * rmu[OIX(i, j + 1, k)] +
                         0.25 * rmu[OIX(i, j, k + 1)] + 0.25 * rmu[OIX(i, j + 1, k + 1)];
        sb[5 * n3 + OIX(i, j, k)] += taub[5 * n3 + OIX(i, j, k)] * m * 2.0;
        taub[5 * n3 + OIX(i, j, k)] = 0.0;
      }
  for (int k = nk + 2; k >= -2; k--)
    for (int j = nj + 3; j >= -2; j--)
      for (int i = ni + 2; i >= -2; i--) {
        const double m = +0.25 * rmu[OIX(i, j, k)] + 0.25 * rmu[OIX(i + 1, j, k)] +
                         0.25 * rmu[OIX(i, j, k + 1)] + 0.25 * rmu[OIX(i + 1, j, k + 1)];
        sb[4 * n3 + OIX(i, j, k)] += taub[4 * n3 + OIX(i, j, k)] * m * 2.0;
        taub[4 * n3 + OIX(i, j, k)] = 0.0;
      }
  for (int k = nk + 3; k >= -2; k--)
    for (int j = nj + 2; j >= -2; j--)
      for (int i = ni + 2; i >= -2; i--) {
        const double m = +0.25 * rmu[OIX(i, j, k)] + 0.25 * rmu[OIX(i + 1, j, k)] +
                         0.25 * rmu[OIX(i, j + 1, k)] + 0.25 * rmu[OIX(i + 1, j + 1, k)];
        sb[3 * n3 + OIX(i, j, k)] += taub[3 * n3 + OIX(i, j, k)] * m * 2.0;
        taub[3 * n3 + OIX(i, j, k)] = 0.0;
      }
  for (int l = 2; l >= 0; l--)
    for (ptrdiff_t m = n3 - 1; m >= 0; m--) {
      sb[l * n3 + m] += taub[l * n3 + m] * rmu[m] * 2.0;
      taub[l * n3 + m] = 0.0;
    }
}

/* srcs/cal_div_tensor.f:20-86: vis = div(tau) at the u, v, w points on
 * their boxes, zero elsewhere (the arrays are overwritten whole). */
void orc_cal_div_tensor_b(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                          const double *dyinv_, const double *dzinv_, double *taukb, double *ukb,
                          double *vkb, double *wkb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  double *t1 = taukb, *t2 = taukb + n3, *t3 = taukb + 2 * n3, *t4 = taukb + 3 * n3,
         *t5 = taukb + 4 * n3, *t6 = taukb + 5 * n3;
  for (int k = nk + 2; k >= -2; k--)
    for (int j = nj + 3; j >= -1; j--)
      for (int i = ni + 3; i >= -1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        const double g = wkb[c];
        t5[OIX(i - 1, j, k)] -= g * dxinv;
        t5[c] += g * dxinv;
        t6[OIX(i, j - 1, k)] -= g * dyinv;
        t6[c] += g * dyinv;
        t3[c] -= g * dzinv;
        t3[OIX(i, j, k + 1)] += g * dzinv;
      }
  for (int k = nk + 3; k >= -1; k--)
    for (int j = nj + 2; j >= -2; j--)
      for (int i = ni + 3; i >= -1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        const double g = vkb[c];
        t4[OIX(i - 1, j, k)] -= g * dxinv;
        t4[c] += g * dxinv;
        t2[c] -= g * dyinv;
        t2[OIX(i, j + 1, k)] += g * dyinv;
        t6[OIX(i, j, k - 1)] -= g * dzinv;
        t6[c] += g * dzinv;
      }
  for (int k = nk + 3; k >= -1; k--)
    for (int j = nj + 3; j >= -1; j--)
      for (int i = ni + 2; i >= -2; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        const double g = ukb[c];
        t1[c] -= g * dxinv;
        t1[OIX(i + 1, j, k)] += g * dxinv;
        t4[OIX(i, j - 1, k)] -= g * dyinv;
        t4[c] += g * dyinv;
        t5[OIX(i, j, k - 1)] -= g * dzinv;
        t5[c] += g * dzinv;
      }
  memset(ukb, 0, sizeof(double) * n3);
  memset(vkb, 0, sizeof(double) * n3);
  memset(wkb, 0, sizeof(double) * n3);
}

/* srcs/cal_srcu.f:51-122: src = u/dt - 1.5 adv + 0.5 adv_o - prs/r00 +
 * 0.5 vis/r00 + 0.5 fst + 0.5 fst_n (+ gravity) on the interior (v to jen),
 * densities fixed.  Gravity is a constant of the map and has no adjoint. */
void orc_cal_srcu_b(const int *nID, const int *ni_, const int *nj_, const int *nk_,
                    const double *rho_av, const double *dt_, const double *grvb_,
                    const double *grvp_, double *ub, double *vb, double *wb, const double *rho,
                    const double *rhon, double *adv_ub, double *adv_vb, double *adv_wb,
                    double *adv_uob, double *adv_vob, double *adv_wob, double *prs_ub,
                    double *prs_vb, double *prs_wb, double *vis_ub, double *vis_vb,
                    double *vis_wb, double *fst_ub, double *fst_vb, double *fst_wb,
                    double *fst_unb, double *fst_vnb, double *fst_wnb, double *src_ub,
                    double *src_vb, double *src_wb) {
  (void)rho_av; (void)grvb_; (void)grvp_;
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  const int jen = (nID[ORC_Y_PLUS - 1] < 0) ? nj - 1 : nj;
  const double dtinv = 1.0 / (*dt_);
  const ptrdiff_t F[3] = {1, sx, sx * sy};
  double *vel[3] = {ub, vb, wb}, *adv[3] = {adv_ub, adv_vb, adv_wb},
         *advo[3] = {adv_uob, adv_vob, adv_wob}, *prs[3] = {prs_ub, prs_vb, prs_wb},
         *vis[3] = {vis_ub, vis_vb, vis_wb}, *fst[3] = {fst_ub, fst_vb, fst_wb},
         *fstn[3] = {fst_unb, fst_vnb, fst_wnb}, *src[3] = {src_ub, src_vb, src_wb};
  for (int q = 2; q >= 0; q--) {
    const int jend = q == 1 ? jen : nj;
    for (int k = nk; k >= 1; k--)
      for (int j = jend; j >= 1; j--)
        for (int i = ni; i >= 1; i--) {
          const ptrdiff_t c = OIX(i, j, k);
          const double r = (rho[c] + rho[c + F[q]] + rhon[c] + rhon[c + F[q]]) * 0.25;
          const double s = src[q][c];
          vel[q][c] += s * dtinv;
          adv[q][c] -= s * 1.5;
          advo[q][c] += s * 0.5;
          prs[q][c] -= s / r;
          vis[q][c] += s / r * 0.5;
          fst[q][c] += s * 0.5;
          fstn[q][c] += s * 0.5;
          src[q][c] = 0.0;
        }
  }
}

/* Reverse of weno5 (srcs/weno5.h:24-60, orc weno5 in fel_oracle.c): given the
 * six differences, the transport velocity and the adjoint rb of the result,
 * adds the adjoints of the differences to db[6] (dfm5, dfm3, dfm1, dfp1,
 * dfp3, dfp5) and returns the adjoint of the velocity.  The derivative of
 * |vel| at vel = 0 is taken as 0. */
static double weno5_b(const double *df, double vel00, double rb, double *db) {
  const double verysmall = 1.0e-80, d1_12 = 1.0 / 12.0;
  const double dfm5 = df[0], dfm3 = df[1], dfm1 = df[2], dfp1 = df[3], dfp3 = df[4], dfp5 = df[5];
  /* forward */
  const double m1 = dfm5 * 2.0 - dfm3 * 7.0 + dfm1 * 11.0, m2 = -dfm3 + dfm1 * 5.0 + dfp1 * 2.0,
               m3 = dfm1 * 2.0 + dfp1 * 5.0 - dfp3;
  const double p1 = dfp5 * 2.0 - dfp3 * 7.0 + dfp1 * 11.0, p2 = -dfp3 + dfp1 * 5.0 + dfm1 * 2.0,
               p3 = dfp1 * 2.0 + dfm1 * 5.0 - dfm3;
  /* smoothness indicators: 13 a^2 + 3 b^2 with (a, b) per stencil */
  double a[6], b[6], sm[6];
  a[0] = dfm5 - dfm3 * 2.0 + dfm1; b[0] = dfm5 - dfm3 * 4.0 + dfm1 * 3.0;
  a[1] = dfm3 - dfm1 * 2.0 + dfp1; b[1] = dfm3 - dfp1;
  a[2] = dfm1 - dfp1 * 2.0 + dfp3; b[2] = dfm1 * 3.0 - dfp1 * 4.0 + dfp3;
  a[3] = dfp5 - dfp3 * 2.0 + dfp1; b[3] = dfp5 - dfp3 * 4.0 + dfp1 * 3.0;
  a[4] = dfp3 - dfp1 * 2.0 + dfm1; b[4] = dfp3 - dfm1;
  a[5] = dfp1 - dfm1 * 2.0 + dfm3; b[5] = dfp1 * 3.0 - dfm1 * 4.0 + dfm3;
  for (int t = 0; t < 6; t++) sm[t] = a[t] * a[t] * 13.0 + b[t] * b[t] * 3.0;
  double smb[6] = {0, 0, 0, 0, 0, 0};
  const double velab = fabs(vel00);
  const double velm1 = vel00 + velab, velp1 = vel00 - velab;
  const double sg = vel00 > 0.0 ? 1.0 : (vel00 < 0.0 ? -1.0 : 0.0);
  double velb = 0.0;
  double dm[3], dp[3];
  for (int side = 0; side < 2; side++) {
    const double *s = sm + 3 * side;
    const double f1 = side ? p1 : m1, f2 = side ? p2 : m2, f3 = side ? p3 : m3;
    const double P1 = s[1] * s[2], P2 = s[0] * s[2], P3 = s[0] * s[1];
    const double am1 = P1 * P1, am2 = P2 * P2 * 6.0, am3 = P3 * P3 * 3.0;
    const double deninv = 1.0 / (am1 + am2 + am3 + verysmall);
    const double w1 = am1 * deninv, w3 = am3 * deninv, w2 = 1.0 - w1 - w3;
    const double R = w1 * f1 + w2 * f2 + w3 * f3;
    const double vfac = side ? velp1 : velm1;
    velb += R * d1_12 * rb * (side ? (1.0 - sg) : (1.0 + sg));
    const double Rb = vfac * d1_12 * rb;
    double *d = side ? dp : dm;
    d[0] = Rb * w1; d[1] = Rb * w2; d[2] = Rb * w3;
    const double w1b = Rb * (f1 - f2), w3b = Rb * (f3 - f2);
    const double am1b0 = w1b * deninv, am3b0 = w3b * deninv;
    const double deninvb = w1b * am1 + w3b * am3;
    const double denb = -deninvb * deninv * deninv;
    const double am1b = am1b0 + denb, am2b = denb, am3b = am3b0 + denb;
    const double P1b = 2.0 * P1 * am1b, P2b = 12.0 * P2 * am2b, P3b = 6.0 * P3 * am3b;
    double *sb = smb + 3 * side;
    sb[0] += P2b * s[2] + P3b * s[1];
    sb[1] += P1b * s[2] + P3b * s[0];
    sb[2] += P1b * s[1] + P2b * s[0];
  }
  /* differences from the candidate stencils (dfm_k, dfp_k) */
  db[0] += dm[0] * 2.0;
  db[1] += -dm[0] * 7.0 - dm[1] - dp[2];
  db[2] += dm[0] * 11.0 + dm[1] * 5.0 + dm[2] * 2.0 + dp[1] * 2.0 + dp[2] * 5.0;
  db[3] += dm[1] * 2.0 + dm[2] * 5.0 + dp[0] * 11.0 + dp[1] * 5.0 + dp[2] * 2.0;
  db[4] += -dm[2] - dp[0] * 7.0 - dp[1];
  db[5] += dp[0] * 2.0;
  /* differences from the smoothness indicators */
  static const double ca[6][6] = {{1, -2, 1, 0, 0, 0}, {0, 1, -2, 1, 0, 0}, {0, 0, 1, -2, 1, 0},
                                  {0, 0, 0, 1, -2, 1}, {0, 0, 1, -2, 1, 0}, {0, 1, -2, 1, 0, 0}};
  static const double cb[6][6] = {{1, -4, 3, 0, 0, 0}, {0, 1, 0, -1, 0, 0}, {0, 0, 3, -4, 1, 0},
                                  {0, 0, 0, 3, -4, 1}, {0, 0, -1, 0, 1, 0}, {0, 1, -4, 3, 0, 0}};
  for (int t = 0; t < 6; t++) {
    const double ab = 26.0 * a[t] * smb[t], bb = 6.0 * b[t] * smb[t];
    for (int m = 0; m < 6; m++) db[m] += ab * ca[t][m] + bb * cb[t][m];
  }
  return velb;
}

/* srcs/cal_advu_weno5.f:1-257 reversed (orc_cal_advu_weno5): for each
 * component, cell and direction, the adjoint of the WENO5 derivative
 * goes to the seven cells of its differences and to the four velocities of
 * its transport velocity.  Not linear: u, v, w are the forward's inputs;
 * ub, vb, wb are incremented; the adv adjoints are consumed (the forward
 * overwrites the whole arrays). */
void orc_cal_advu_weno5_b(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                          const double *dyinv_, const double *dzinv_, const double *u,
                          double *ub, const double *v, double *vb, const double *w, double *wb,
                          double *adv_ub, double *adv_vb, double *adv_wb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dinv[3] = {*dxinv_, *dyinv_, *dzinv_};
  ODIMS(ni, nj, nk);
  const ptrdiff_t E[3] = {1, sx, sx * sy};
  const double *vel[3] = {u, v, w};
  double *velb[3] = {ub, vb, wb};
  double *advb[3] = {adv_ub, adv_vb, adv_wb};
  for (int q = 2; q >= 0; q--)
    for (int k = nk; k >= 1; k--)
      for (int j = nj; j >= 1; j--)
        for (int i = ni; i >= 1; i--) {
          const ptrdiff_t c = OIX(i, j, k);
          const double ab = advb[q][c];
          advb[q][c] = 0.0;
          for (int dd = 2; dd >= 0; dd--) {
            const ptrdiff_t st = E[dd];
            /* transport velocity: component dd at c - e_d, c - e_d + e_q, c, c + e_q */
            const double *t = vel[dd];
            const ptrdiff_t o[4] = {c - st, c - st + E[q], c, c + E[q]};
            const double cm = (t[o[0]] + t[o[1]]) * 0.5, cp = (t[o[2]] + t[o[3]]) * 0.5;
            const double vel00 = (cm + cp) * 0.5;
            const double *f = vel[q];
            double df[6];
            for (int m = 0; m < 6; m++) df[m] = -f[c + (m - 3) * st] + f[c + (m - 2) * st];
            double db[6] = {0, 0, 0, 0, 0, 0};
            const double vb00 = weno5_b(df, vel00, ab * dinv[dd], db);
            double *fb = velb[q];
            for (int m = 0; m < 6; m++) {
              fb[c + (m - 3) * st] -= db[m];
              fb[c + (m - 2) * st] += db[m];
            }
            for (int m = 0; m < 4; m++) velb[dd][o[m]] += vb00 * 0.25;
          }
        }
  /* the forward also zeroes the ghost cells of adv: consumed everywhere */
  for (int q = 0; q < 3; q++) memset(advb[q], 0, sizeof(double) * n3);
}

/* srcs/cpy.f:14-20: q2 = q1 on the padded array.  Transpose: q1_b += q2_b,
 * q2_b = 0. */
void orc_cpy_b(const int *ni_, const int *nj_, const int *nk_, double *q1b, double *q2b) {
  ODIMS(*ni_, *nj_, *nk_);
  if (q1b == q2b) return;
  for (ptrdiff_t m = n3 - 1; m >= 0; m--) {
    q1b[m] += q2b[m];
    q2b[m] = 0.0;
  }
}

/* ------------------------------------------------------------------------- */
/* Volume-fraction transport (MTHINC)                                        */
/* ------------------------------------------------------------------------- */

/* srcs/cal_grad_p2a.f:30-71 transposed: the 8-point vertex gradients on
 * [-2,n+2]^3 (linear). */
void orc_cal_grad_p2a_b(const int *ID, const int *svall, const int *ni_, const int *nj_,
                        const int *nk_, const double *dxinv_, const double *dyinv_array,
                        const double *dzinv_, double *qb, double *qxb, double *qyb, double *qzb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  (void)svall;
  for (int k = nk + 2; k >= -2; k--)
    for (int j = nj + 2; j >= -2; j--)
      for (int i = ni + 2; i >= -2; i--) {
        const double dyinv = dyinv_array[j + nj * (*ID) + 2];
        const ptrdiff_t c = OIX(i, j, k);
        const double gx = qxb[c] * 0.25 * dxinv, gy = qyb[c] * 0.25 * dyinv,
                     gz = qzb[c] * 0.25 * dzinv;
        for (int m = 0; m < 8; m++) {
          const int a = m & 1, b = (m >> 1) & 1, e = (m >> 2) & 1;
          const double sxg = a ? 1.0 : -1.0, syg = b ? 1.0 : -1.0, szg = e ? 1.0 : -1.0;
          qb[OIX(i + a, j + b, k + e)] += sxg * gx + syg * gy + szg * gz;
        }
        qxb[c] = 0.0;
        qyb[c] = 0.0;
        qzb[c] = 0.0;
      }
}

static inline double sig2(double bet, double z) { return 1.0 / (1.0 + exp(-(2.0 * bet * z))); }

/* srcs/solphi_mthinc1.f:1-292 reversed at its inputs (orc_solphi_mthinc1).
 * Flat cells (upwind branch): flux = u phi00 dt/dx, phi00 the truncated phi
 * (derivative 0 where truncated).  Interface cells: flux = sum over the
 * face sub-cell's 8 Gauss points of (vd/8) s(n.x + d), s(z) = 1/(1 +
 * exp(-2 bet z)), n = r/|r| (r the mean of the 8 vertex gradients), the
 * sub-cell's Gauss points moving with vd = u dt/dx; d solves
 * sum_g s(n.x_g + d)/8 = phi over the cell's Gauss points, and its
 * derivative is taken by the implicit-function theorem at the Newton
 * iterate: dd = (dphi - dn.G_n)/G_d.  A face both neighbours write (flat
 * cells, u = 0) takes the later writer's derivative, as the forward's loop
 * order leaves its value.  Branches and upwind directions are those of the
 * forward at its inputs.  The flux adjoints are consumed on [-1,n+1]^3. */
void orc_solphi_mthinc1_b(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                          const double *dyinv_, const double *dzinv_, const double *dt_,
                          const double *bet_mthinc, const double *phi, double *phib,
                          const double *phix, double *phixb, const double *phiy, double *phiyb,
                          const double *phiz, double *phizb, const double *u, double *ub,
                          const double *v, double *vb, const double *w, double *wb,
                          double *flphixb, double *flphiyb, double *flphizb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dinv[3] = {*dxinv_, *dyinv_, *dzinv_}, dt = *dt_;
  ODIMS(ni, nj, nk);
  const double dx = 1.0 / dinv[0], dy = 1.0 / dinv[1], dz = 1.0 / dinv[2];
  const double bet = *bet_mthinc;
  const double gp1 = -(0.5 / sqrt(3.0)), gp2 = 0.5 / sqrt(3.0);
  const double gp[2] = {gp1, gp2};
  const double dcellinv = 1.0 / sqrt((dx * dx + dy * dy + dz * dz) / 3.0);
  const double ch = cosh(-(bet * 1.5));
  const double eps_mthinc = bet / (2.0 * (ch * ch));
  const double eps_trunc = 0.5 * (tanh(-(bet * 2.5)) + 1.0);
  const ptrdiff_t E[3] = {1, sx, sx * sy};
  const double *vel[3] = {u, v, w};
  double *velb[3] = {ub, vb, wb};
  double *flb[3] = {flphixb, flphiyb, flphizb};
  double *gb[3] = {phixb, phiyb, phizb};
  double xs[8];
  for (int k = nk + 1; k >= 0; k--)
    for (int j = nj + 1; j >= 0; j--)
      for (int i = ni + 1; i >= 0; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        double r[3];
        cell_normal(sx, sy, phix, phiy, phiz, i, j, k, &r[0], &r[1], &r[2]);
        const double q00 = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        /* faces in the forward's order: minus then plus face of x, y, z;
         * reversed here */
        if (q00 <= eps_mthinc * dcellinv) {
          double phi00 = phi[c];
          double dphi = 1.0;
          if (phi00 <= eps_trunc) { phi00 = 0.0; dphi = 0.0; }
          if (1.0 - phi00 <= eps_trunc) { phi00 = 1.0; dphi = 0.0; }
          for (int a = 2; a >= 0; a--) {
            const double vm = vel[a][c - E[a]], vp = vel[a][c];
            if (vp >= 0.0) {
              const double fbv = flb[a][c];
              velb[a][c] += fbv * phi00 * dt * dinv[a];
              phib[c] += fbv * vp * dt * dinv[a] * dphi;
              flb[a][c] = 0.0;
            }
            if (vm <= 0.0) {
              const double fbv = flb[a][c - E[a]];
              velb[a][c - E[a]] += fbv * phi00 * dt * dinv[a];
              phib[c] += fbv * vm * dt * dinv[a] * dphi;
              flb[a][c - E[a]] = 0.0;
            }
          }
          continue;
        }
        const double coef = 1.0 / q00;
        const double n[3] = {r[0] * coef, r[1] * coef, r[2] * coef};
        const double phic = phi[c];
        double ddd = 0.0;
        for (int inr = 1; inr <= 10; inr++) {
          const double wt = mthinc_tanh_sum(n[0], n[1], n[2], gp1, gp2, gp1, gp2, gp1, gp2,
                                                0.125, bet, ddd, xs);
          const double f = wt - phic;
          double s = 0.0;
          for (int m = 0; m < 8; m++) {
            const double cc = cosh(bet * (xs[m] + ddd));
            const double t = 0.125 / (cc * cc);
            s = (m == 0) ? t : s + t;
          }
          const double df = 0.5 * bet * s;
          ddd = ddd - f * df / (df * df + 1.0e-99);
        }
        /* the cell's own Gauss points: G_d, G_n */
        double Gd = 0.0, Gn[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < 8; m++) {
          const double X[3] = {gp[m & 1], gp[(m >> 1) & 1], gp[(m >> 2) & 1]};
          const double z = n[0] * X[0] + n[1] * X[1] + n[2] * X[2] + ddd;
          const double sg = sig2(bet, z);
          const double sp = 0.125 * 2.0 * bet * sg * (1.0 - sg);
          Gd += sp;
          for (int q = 0; q < 3; q++) Gn[q] += sp * X[q];
        }
        double nb[3] = {0.0, 0.0, 0.0}, db = 0.0;
        for (int a = 2; a >= 0; a--) {
          for (int side = 1; side >= 0; side--) {
            /* side 1: plus face (vp > 0), side 0: minus face (vm < 0) */
            const double vface = side ? vel[a][c] : vel[a][c - E[a]];
            if (side ? !(vface > 0.0) : !(vface < 0.0)) continue;
            const ptrdiff_t fc = side ? c : c - E[a];
            const double fbv = flb[a][fc];
            flb[a][fc] = 0.0;
            const double vd = vface * dt * dinv[a];
            double st, en, dst, den_;
            if (side) { st = 0.5 - vd; en = 0.5; dst = -1.0; den_ = 0.0; }
            else { st = -0.5; en = -0.5 - vd; dst = 0.0; den_ = -1.0; }
            const double g[2] = {(st + en) * 0.5 + (-st + en) * gp1, (st + en) * 0.5 + (-st + en) * gp2};
            const double dg[2] = {(dst + den_) * 0.5 + (-dst + den_) * gp1,
                                  (dst + den_) * 0.5 + (-dst + den_) * gp2};
            const double W = 0.125 * vd;
            double S = 0.0, Sp = 0.0, Spx[3] = {0.0, 0.0, 0.0}, Svd = 0.0;
            for (int m = 0; m < 8; m++) {
              const int ch3[3] = {m & 1, (m >> 1) & 1, (m >> 2) & 1};
              double X[3];
              for (int q = 0; q < 3; q++) X[q] = q == a ? g[ch3[q]] : gp[ch3[q]];
              const double z = n[0] * X[0] + n[1] * X[1] + n[2] * X[2] + ddd;
              const double sg = sig2(bet, z);
              const double sp = 2.0 * bet * sg * (1.0 - sg);
              S += sg;
              Sp += sp;
              for (int q = 0; q < 3; q++) Spx[q] += sp * X[q];
              Svd += sp * n[a] * dg[ch3[a]];
            }
            const double vdb = fbv * (0.125 * S + W * Svd);
            velb[a][fc] += vdb * dt * dinv[a];
            for (int q = 0; q < 3; q++) nb[q] += fbv * W * Spx[q];
            db += fbv * W * Sp;
          }
        }
        /* through d (implicit function of n and phi) */
        phib[c] += db / Gd;
        for (int q = 0; q < 3; q++) nb[q] -= db * Gn[q] / Gd;
        /* through n = r/|r| */
        const double ndn = n[0] * nb[0] + n[1] * nb[1] + n[2] * nb[2];
        for (int q = 0; q < 3; q++) {
          const double rb = (nb[q] - ndn * n[q]) * coef;
          for (int m = 0; m < 8; m++) {
            const int a = m & 1, b = (m >> 1) & 1, e = (m >> 2) & 1;
            gb[q][OIX(i - 1 + a, j - 1 + b, k - 1 + e)] += rb * 0.125;
          }
        }
      }
  for (int k = -1; k <= nk + 1; k++)
    for (int j = -1; j <= nj + 1; j++)
      for (int i = -1; i <= ni + 1; i++) {
        flphixb[OIX(i, j, k)] = 0.0;
        flphiyb[OIX(i, j, k)] = 0.0;
        flphizb[OIX(i, j, k)] = 0.0;
      }
}

/* srcs/solphi_mthinc2.f:30-41 reversed: phin = clamp(phi (1 + dt div) +
 * flux differences) on the interior; the clamp passes the derivative where
 * the value is in [0,1] (fmin/fmax select their first operand on ties). */
void orc_solphi_mthinc2_b(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                          const double *dyinv_, const double *dzinv_, const double *dt_,
                          const double *u, double *ub, const double *v, double *vb,
                          const double *w, double *wb, const double *flphix, double *flphixb,
                          const double *flphiy, double *flphiyb, const double *flphiz,
                          double *flphizb, const double *phi, double *phib, double *phinb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_, dt = *dt_;
  ODIMS(ni, nj, nk);
  for (int k = nk; k >= 1; k--)
    for (int j = nj; j >= 1; j--)
      for (int i = ni; i >= 1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        const double div00 = +(-u[OIX(i - 1, j, k)] + u[c]) * dxinv +
                             (-v[OIX(i, j - 1, k)] + v[c]) * dyinv +
                             (-w[OIX(i, j, k - 1)] + w[c]) * dzinv;
        const double p = phi[c] * (1.0 + dt * div00) + flphix[OIX(i - 1, j, k)] - flphix[c] +
                         flphiy[OIX(i, j - 1, k)] - flphiy[c] + flphiz[OIX(i, j, k - 1)] - flphiz[c];
        const double g = (p <= 1.0 && p >= 0.0) ? phinb[c] : 0.0;
        phinb[c] = 0.0;
        phib[c] += g * (1.0 + dt * div00);
        const double h = g * phi[c] * dt;
        ub[c] += h * dxinv; ub[OIX(i - 1, j, k)] -= h * dxinv;
        vb[c] += h * dyinv; vb[OIX(i, j - 1, k)] -= h * dyinv;
        wb[c] += h * dzinv; wb[OIX(i, j, k - 1)] -= h * dzinv;
        flphixb[OIX(i - 1, j, k)] += g; flphixb[c] -= g;
        flphiyb[OIX(i, j - 1, k)] += g; flphiyb[c] -= g;
        flphizb[OIX(i, j, k - 1)] += g; flphizb[c] -= g;
      }
}

/* srcs/solphi_mthinc3.f:43-160 reversed (orc_solphi_mthinc3): truncation of
 * flat cells on [0,n+1]^3 (derivative 0 where it sets 0 or 1), clamp, the
 * mass correction p + offset p^2 (1-p) with offset = (sum phi - sum p) /
 * (sum p^2 (1-p) + 1e-99) over the interior, clamp.  phin is the input
 * phin of the forward (in), phinb the adjoint of its output in and of its
 * input out; phib incremented (through sum phi). */
void orc_solphi_mthinc3_b(const int *ipara, const int *ni_, const int *nj_, const int *nk_,
                          const double *dxinv_, const double *dyinv_, const double *dzinv_,
                          const double *bet_mthinc, const double *phix, const double *phiy,
                          const double *phiz, const double *phi, double *phib, const double *phin,
                          double *phinb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  (void)ipara;
  const double dx = 1.0 / *dxinv_, dy = 1.0 / *dyinv_, dz = 1.0 / *dzinv_;
  const double bet = *bet_mthinc;
  const double dcellinv = 1.0 / sqrt((dx * dx + dy * dy + dz * dz) / 3.0);
  const double ch = cosh(-(bet * 1.5));
  const double eps_mthinc = bet / (2.0 * (ch * ch));
  const double eps_trunc = 0.5 * (tanh(-(bet * 2.5)) + 1.0);
  /* forward replay: p after truncation (mask t) and clamp (mask b) */
  double *p = malloc(sizeof(double) * n3);
  unsigned char *mt = calloc((size_t)n3, 1), *mb = calloc((size_t)n3, 1), *md = calloc((size_t)n3, 1);
  memcpy(p, phin, sizeof(double) * n3);
  for (ptrdiff_t m = 0; m < n3; m++) mt[m] = 1;
  for (int k = 0; k <= nk + 1; k++)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        double nx, ny, nz;
        cell_normal(sx, sy, phix, phiy, phiz, i, j, k, &nx, &ny, &nz);
        const double q00 = sqrt(nx * nx + ny * ny + nz * nz);
        if (q00 <= eps_mthinc * dcellinv) {
          const ptrdiff_t c = OIX(i, j, k);
          if (p[c] <= eps_trunc) { p[c] = 0.0; mt[c] = 0; }
          if (1.0 - p[c] <= eps_trunc) { p[c] = 1.0; mt[c] = 0; }
        }
      }
  double sum1 = 0.0, sum2 = 0.0, den2 = 0.0;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        mb[c] = p[c] <= 1.0 && p[c] >= 0.0;
        p[c] = fmax(fmin(p[c], 1.0), 0.0);
        sum1 = sum1 + phi[c];
        sum2 = sum2 + p[c];
        den2 = den2 + p[c] * p[c] * (1.0 - p[c]);
      }
  const double den = den2 + 1.0e-99;
  const double offset = (sum1 - sum2) / den;
  /* reverse of the correction and its clamp */
  double offb = 0.0;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const ptrdiff_t c = OIX(i, j, k);
        const double q = p[c] + offset * (p[c] * p[c]) * (1.0 - p[c]);
        md[c] = q <= 1.0 && q >= 0.0;
        if (md[c]) offb += phinb[c] * (p[c] * p[c]) * (1.0 - p[c]);
      }
  const double s1b = offb / den, s2b = -offb / den, denb = -offb * offset / den;
  for (int k = nk; k >= 1; k--)
    for (int j = nj; j >= 1; j--)
      for (int i = ni; i >= 1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        const double pc = p[c], gp = 2.0 * pc - 3.0 * pc * pc;
        double l = md[c] ? phinb[c] * (1.0 + offset * gp) : 0.0;
        l += s2b + denb * gp;
        phib[c] += s1b;
        phinb[c] = mb[c] ? l : 0.0;
      }
  for (ptrdiff_t m = 0; m < n3; m++)
    if (!mt[m]) phinb[m] = 0.0;
  free(p); free(mt); free(mb); free(md);
}

/* srcs/summation.f:27-50 reversed: qk(.,0) = sum of the nbub slices on
 * the padded array, clamped to [0,1] on the interior (qk holds the forward's
 * input slices; qkb(.,0) is consumed into qkb(.,1..nbub)). */
void orc_summation_b(const int *ni_, const int *nj_, const int *nk_, const double *qk, double *qkb,
                     const int *nbub_) {
  const int ni = *ni_, nj = *nj_, nk = *nk_, nbub = *nbub_;
  ODIMS(ni, nj, nk);
  for (ptrdiff_t m = 0; m < n3; m++) {
    double s = qk[n3 + m];
    for (int l = 2; l <= nbub; l++) s = s + qk[(ptrdiff_t)l * n3 + m];
    const ptrdiff_t pl = m % (sx * sy), i = pl % sx - 2, j = pl / sx - 2, k = m / (sx * sy) - 2;
    const int inner = i >= 1 && i <= ni && j >= 1 && j <= nj && k >= 1 && k <= nk;
    const double g = (!inner || (s <= 1.0 && s >= 0.0)) ? qkb[m] : 0.0;
    qkb[m] = 0.0;
    for (int l = 1; l <= nbub; l++) qkb[(ptrdiff_t)l * n3 + m] += g;
  }
}

/* ------------------------------------------------------------------------- */
/* Contact line                                                              */
/* ------------------------------------------------------------------------- */

/* srcs/gnbc.f90:22-95 reversed: theta = acos(clamp(cos(theta_0) -
 * zeta u_cl/sigma)) in degrees at the wall rows, u_cl = +-sign(ux) sqrt(ux^2 +
 * uz^2 + 1e-12) from 4-point averages of u (+-uwall) and w.  The clamp and
 * acos pass the derivative only strictly inside (-1,1).  theta_arrayb is
 * consumed on the two wall rows. */
void orc_gnbc_b(const int *nID, const int *ni_, const int *nj_, const int *nk_, const double *uk,
                double *ukb, const double *wk, double *wkb, const double *uwall_, double *uwallb,
                const double *theta_0_array, double *theta_0b, const double *surface_tension,
                const double *zeta_array, double *zetab, double *theta_arrayb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  const double pi = acos(-1.0), eps = 1.0e-12, sigma = *surface_tension, uwall = *uwall_;
  for (int side = 1; side >= 0; side--) {
    if (side == 0 && !(nID[ORC_Y_MINUS - 1] < 0)) continue;
    if (side == 1 && !(nID[ORC_Y_PLUS - 1] < 0)) continue;
    const int jw = side ? nj : 1, ja = side ? nj : 0, jb = side ? nj + 1 : 1;
    const double uws = side ? -uwall : uwall;
    for (int k = nk + 3; k >= -2; k--)
      for (int i = ni + 3; i >= -2; i--) {
        const int mi = (i + ni + 2) % (ni + 3);
        const double ux = (uk[OIX(mi, ja, k)] + uk[OIX(mi, jb, k)] + uk[OIX(i, ja, k)] +
                           uk[OIX(i, jb, k)]) / 4.0 + uws;
        const double uz = (wk[OIX(mi, ja, k)] + wk[OIX(mi, jb, k)] + wk[OIX(i, ja, k)] +
                           wk[OIX(i, jb, k)]) / 4.0;
        const double R = sqrt(ux * ux + uz * uz + eps);
        double sgn = ux >= 0.0 ? 1.0 : -1.0;
        if (i <= ni / 2) sgn = -sgn;
        const double ucl = sgn * R;
        const double th0 = theta_0_array[OIX(i, jw, k)] * (pi / 180.0);
        const double zeta = zeta_array[OIX(i, jw, k)];
        const double ct = cos(th0) - zeta * ucl / sigma;
        const ptrdiff_t cw = OIX(i, jw, k);
        const double tb = theta_arrayb[cw];
        theta_arrayb[cw] = 0.0;
        if (!(ct > -1.0 && ct < 1.0)) continue;
        const double ctb = tb * (180.0 / pi) * (-1.0 / sqrt(1.0 - ct * ct));
        theta_0b[cw] += ctb * (-sin(th0)) * (pi / 180.0);
        zetab[cw] += ctb * (-ucl / sigma);
        const double uclb = ctb * (-zeta / sigma);
        const double uxb = uclb * sgn * ux / R, uzb = uclb * sgn * uz / R;
        ukb[OIX(mi, ja, k)] += uxb / 4.0; ukb[OIX(mi, jb, k)] += uxb / 4.0;
        ukb[OIX(i, ja, k)] += uxb / 4.0; ukb[OIX(i, jb, k)] += uxb / 4.0;
        wkb[OIX(mi, ja, k)] += uzb / 4.0; wkb[OIX(mi, jb, k)] += uzb / 4.0;
        wkb[OIX(i, ja, k)] += uzb / 4.0; wkb[OIX(i, jb, k)] += uzb / 4.0;
        *uwallb += side ? -uxb : uxb;
      }
  }
}

/* Derivative in theta (degrees) of R = n_y/(n_t + 1e-20) of the contact-angle
 * normal n = (s sin t ax, cos t, s sin t az)/|.| (t in radians, the norm with
 * +1e-20 inside the root), t = the tangential component x (comp 0) or z (2). */
static double ca_ratio(double th_deg, double sg, double ax, double az, int comp, double *dr) {
  const double pi = acos(-1.0), eps = 1.0e-20;
  const double th = th_deg * pi / 180.0, ct = cos(th), st = sin(th);
  const double rx = sg * st * ax, ry = ct, rz = sg * st * az;
  const double N = sqrt(rx * rx + ry * ry + rz * rz + eps);
  const double drx = sg * ct * ax, dry = -st, drz = sg * ct * az;
  const double dN = (rx * drx + ry * dry + rz * drz) / N;
  const double ny = ry / N, dny = (dry * N - ry * dN) / (N * N);
  const double rt = comp == 0 ? rx : rz, drt = comp == 0 ? drx : drz;
  const double nt = rt / N, dnt = (drt * N - rt * dN) / (N * N);
  const double b = nt + eps;
  *dr = (dny * b - ny * dnt) / (b * b) * (pi / 180.0);
  return ny / b;
}

/* srcs/bnd_contact_angle.f90:17-72 reversed (orc_bnd_contact_angle): the
 * three ghost rows of each wall column, from the wall row's value, its x and
 * z neighbours (outside the padded range read as 0) and theta.  qk holds the
 * forward's input; qkb the adjoint of its output on entry, of its input on
 * return; theta_arrayb incremented. */
void orc_bnd_contact_angle_b(const int *nID, const int *ni_, const int *nj_, const int *nk_,
                             const double *qk, double *qkb, const double *theta_array,
                             double *theta_arrayb, const double *dx_, const double *dy_,
                             const double *dz_) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  const double ax = *dy_ / *dx_, az = *dy_ / *dz_;
  for (int side = 1; side >= 0; side--) {
    if (side == 0 && !(nID[ORC_Y_MINUS - 1] < 0)) continue;
    if (side == 1 && !(nID[ORC_Y_PLUS - 1] < 0)) continue;
    const int jw = side == 0 ? 1 : nj, jg = side == 0 ? 0 : nj + 1, js = side == 0 ? -1 : 1;
    for (int k = nk + 3; k >= -2; k--)
      for (int i = ni + 3; i >= -2; i--) {
        const double sg = copysign(1.0, (double)(i - ni / 2));
        double dRx, dRz;
        const double Rx = ca_ratio(theta_array[OIX(i, jw, k)], sg, ax, az, 0, &dRx);
        const double Rz = ca_ratio(theta_array[OIX(i, jw, k)], sg, ax, az, 2, &dRz);
        const double qp = (i + 1 > ni + 3) ? 0.0 : qk[OIX(i + 1, jw, k)];
        const double qm = (i - 1 < -2) ? 0.0 : qk[OIX(i - 1, jw, k)];
        const double qt = (k + 1 > nk + 3) ? 0.0 : qk[OIX(i, jw, k + 1)];
        const double qb = (k - 1 < -2) ? 0.0 : qk[OIX(i, jw, k - 1)];
        const double Dx = (qp - qm) * 0.5 * ax, Dz = (qt - qb) * 0.5 * az;
        /* g2 = 2 g1 - g0, g1 = 2 g0 - q1, g0 = q1 - Rx Dx - Rz Dz */
        double l2 = qkb[OIX(i, jg + 2 * js, k)], l1 = qkb[OIX(i, jg + js, k)], l0 = qkb[OIX(i, jg, k)];
        qkb[OIX(i, jg + 2 * js, k)] = 0.0;
        qkb[OIX(i, jg + js, k)] = 0.0;
        qkb[OIX(i, jg, k)] = 0.0;
        l1 += 2.0 * l2;
        l0 -= l2;
        l0 += 2.0 * l1;
        qkb[OIX(i, jw, k)] += l0 - l1;
        if (i + 1 <= ni + 3) qkb[OIX(i + 1, jw, k)] -= l0 * Rx * 0.5 * ax;
        if (i - 1 >= -2) qkb[OIX(i - 1, jw, k)] += l0 * Rx * 0.5 * ax;
        if (k + 1 <= nk + 3) qkb[OIX(i, jw, k + 1)] -= l0 * Rz * 0.5 * az;
        if (k - 1 >= -2) qkb[OIX(i, jw, k - 1)] += l0 * Rz * 0.5 * az;
        theta_arrayb[OIX(i, jw, k)] += l0 * (-Dx * dRx - Dz * dRz);
      }
  }
}
