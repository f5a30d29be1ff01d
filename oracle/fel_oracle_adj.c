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
