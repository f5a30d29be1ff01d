/*
 * fel_oracle_adj.c -- CPU restatement (plain C, FP64, serial) of the adjoints
 * (reverse-mode derivatives) of the routines of the timestep of
 * toruneko1129/fel-multi-phase-flow.
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
 * any dropped term, wrong index or transposed coefficient breaks it.  The
 * nonlinear routines (WENO5, MTHINC, gnbc, the contact angle, the normals,
 * calfst) and the all-argument adjoints (_bp: the parameterising arrays'
 * adjoints too) are pinned the same way against Richardson central
 * differences of the forward oracle, and the reverse of whole windows of
 * steps likewise.
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

/* Adjoint of the slip coefficients of one wall cell (slip_coef) given
 * c1_b, c2_b: ls = 1/(phi_av/l1 + (1-phi_av)/l2), c1 = 2dy/(2ls+dy),
 * c2 = (2ls-dy)/(2ls+dy); increments l1_b, l2_b at the cell and phi_b at
 * the two cells of phi_av. */
static void slip_coef_b(ptrdiff_t sx, ptrdiff_t sy, int ni, int i, int jw, int k, double dy,
                        const double *l1, const double *l2, const double *phi, double c1b,
                        double c2b, double *l1b, double *l2b, double *phib) {
  const ptrdiff_t m0 = OIX(i, jw, k), m1 = OIX((i + 1) % (ni + 3), jw, k);
  const double phi_av = (phi[m0] + phi[m1]) / 2.0;
  const double inv1 = phi_av / l1[m0], inv2 = (1.0 - phi_av) / l2[m0];
  const double ls = 1.0 / (inv1 + inv2);
  const double dd = 2.0 * ls + dy;
  const double lsb = c1b * (-4.0 * dy / (dd * dd)) + c2b * (4.0 * dy / (dd * dd));
  const double sb = -lsb * ls * ls; /* adjoint of inv1 + inv2 */
  l1b[m0] -= sb * phi_av / (l1[m0] * l1[m0]);
  l2b[m0] -= sb * (1.0 - phi_av) / (l2[m0] * l2[m0]);
  const double pab = sb / l1[m0] - sb / l2[m0];
  phib[m0] += pab / 2.0;
  phib[m1] += pab / 2.0;
}

/* srcs/bndu.f:29-158 (orc_bndu): the j-wall rows, then the x fills, then
 * the z fills, each written as ghost = function of earlier values.
 * Transpose in reverse: z fills, x fills, wall rows; uwall_b gathers the
 * wall rows' -c1 (bottom) and +c1 (top) weights.  With uk != NULL the
 * adjoints of the slip lengths and phi are formed too (bndu_bp): at the
 * point where a wall row's ghost u0 = c1 (-+uwall) + c2 u1 (and w0 = c2 w1)
 * is reversed, its adjoint A (the rows beyond folded in) gives c1_b =
 * A (-+uwall), c2_b = A_u u1 + A_w w1 with u1, w1 the call's input values. */
static void bndu_b_core(const int *nID, int ni, int nj, int nk, double *ukb, double *vkb,
                        double *wkb, double *uwallb, double dy, const double *l1,
                        const double *l2, const double *phi, const double *uk,
                        const double *wk, double uwall, double *l1b, double *l2b,
                        double *phib);

void orc_bndu_b(const int *nID, const int *ni_, const int *nj_, const int *nk_, double *ukb,
                double *vkb, double *wkb, double *uwallb, const double *dy_, const double *l1,
                const double *l2, const double *phi) {
  bndu_b_core(nID, *ni_, *nj_, *nk_, ukb, vkb, wkb, uwallb, *dy_, l1, l2, phi, NULL, NULL, 0.0,
              NULL, NULL, NULL);
}

/* All-argument adjoint of bndu: orc_bndu_b plus the adjoints of l1, l2 and
 * phi (incremented); uk, wk are the call's input velocities. */
void orc_bndu_bp(const int *nID, const int *ni_, const int *nj_, const int *nk_,
                 const double *uk, double *ukb, double *vkb, const double *wk, double *wkb,
                 const double *uwall, double *uwallb, const double *dy_, const double *l1,
                 double *l1b, const double *l2, double *l2b, const double *phi, double *phib) {
  bndu_b_core(nID, *ni_, *nj_, *nk_, ukb, vkb, wkb, uwallb, *dy_, l1, l2, phi, uk, wk, *uwall,
              l1b, l2b, phib);
}

static void bndu_b_core(const int *nID, int ni, int nj, int nk, double *ukb, double *vkb,
                        double *wkb, double *uwallb, double dy, const double *l1,
                        const double *l2, const double *phi, const double *uk,
                        const double *wk, double uwall, double *l1b, double *l2b,
                        double *phib) {
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
        double c1b = 0.0, c2b = 0.0;
        double *q = wkb;
        q[OIX(i, nj + 2, k)] += 2.0 * q[OIX(i, nj + 3, k)];
        q[OIX(i, nj + 1, k)] -= q[OIX(i, nj + 3, k)];
        q[OIX(i, nj + 3, k)] = 0.0;
        q[OIX(i, nj + 1, k)] += 2.0 * q[OIX(i, nj + 2, k)];
        q[OIX(i, nj, k)] -= q[OIX(i, nj + 2, k)];
        q[OIX(i, nj + 2, k)] = 0.0;
        q[OIX(i, nj, k)] += c2 * q[OIX(i, nj + 1, k)];
        if (uk) c2b += q[OIX(i, nj + 1, k)] * wk[OIX(i, nj, k)];
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
        if (uk) {
          c1b += q[OIX(i, nj + 1, k)] * uwall;
          c2b += q[OIX(i, nj + 1, k)] * uk[OIX(i, nj, k)];
          slip_coef_b(sx, sy, ni, i, nj, k, dy, l1, l2, phi, c1b, c2b, l1b, l2b, phib);
        }
        q[OIX(i, nj + 1, k)] = 0.0;
      }
  if (nID[ORC_Y_MINUS - 1] < 0)
    for (int k = nk + 3; k >= -2; k--)
      for (int i = ni + 3; i >= -2; i--) {
        double c1, c2;
        slip_coef(sx, sy, ni, i, 1, k, dy, l1, l2, phi, &c1, &c2);
        double c1b = 0.0, c2b = 0.0;
        double *q = wkb;
        q[OIX(i, -1, k)] += 2.0 * q[OIX(i, -2, k)];
        q[OIX(i, 0, k)] -= q[OIX(i, -2, k)];
        q[OIX(i, -2, k)] = 0.0;
        q[OIX(i, 0, k)] += 2.0 * q[OIX(i, -1, k)];
        q[OIX(i, 1, k)] -= q[OIX(i, -1, k)];
        q[OIX(i, -1, k)] = 0.0;
        q[OIX(i, 1, k)] += c2 * q[OIX(i, 0, k)];
        if (uk) c2b += q[OIX(i, 0, k)] * wk[OIX(i, 1, k)];
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
        if (uk) {
          c1b -= q[OIX(i, 0, k)] * uwall;
          c2b += q[OIX(i, 0, k)] * uk[OIX(i, 1, k)];
          slip_coef_b(sx, sy, ni, i, 1, k, dy, l1, l2, phi, c1b, c2b, l1b, l2b, phib);
        }
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
        /* through d (implicit function of n and phi).  Gd = 0: every
         * sigmoid of the cell saturated (phi at 0 or 1 to rounding), the
         * Newton step vanished and the fluxes do not vary with d: no term */
        if (Gd > 0.0) {
          phib[c] += db / Gd;
          for (int q = 0; q < 3; q++) nb[q] -= db * Gn[q] / Gd;
        }
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

/* ------------------------------------------------------------------------- */
/* Parameter adjoints (<name>_bp: the adjoint with respect to every real      */
/* argument -- what <name>_b does, plus the adjoints of the arrays it takes   */
/* as fixed, incremented; primal arrays the derivative needs come before      */
/* their adjoints)                                                            */
/* ------------------------------------------------------------------------- */

/* The 14 neighbour terms of the Gauss-Seidel update of component q in the
 * order srcs/solu_sor4.f:137-152 (186-201, 236-251) sums them: coefficient
 * index within the component's 15, component of the neighbour, offset. */
static void sor_terms(int q, ptrdiff_t X, ptrdiff_t Y, ptrdiff_t Z, int idx[14], int comp[14],
                      ptrdiff_t off[14]) {
  static const int own_idx[6] = {4, 2, 0, 1, 3, 5};
  const ptrdiff_t own_off[6] = {-Z, -Y, -X, X, Y, Z};
  for (int t = 0; t < 6; t++) {
    idx[t] = own_idx[t];
    comp[t] = q;
    off[t] = own_off[t];
  }
  int ci[8];
  int cc[8];
  ptrdiff_t co[8];
  if (q == 0) {
    const int i0[8] = {7, 8, 9, 10, 11, 12, 13, 14};
    const int c0_[8] = {1, 1, 1, 1, 2, 2, 2, 2};
    const ptrdiff_t o0[8] = {-Y, X - Y, 0, X, -Z, X - Z, 0, X};
    memcpy(ci, i0, sizeof ci); memcpy(cc, c0_, sizeof cc); memcpy(co, o0, sizeof co);
  } else if (q == 1) {
    const int i1[8] = {7, 9, 8, 10, 11, 12, 13, 14};
    const int c1_[8] = {0, 0, 0, 0, 2, 2, 2, 2};
    const ptrdiff_t o1[8] = {-X, 0, -X + Y, Y, -Z, Y - Z, 0, Y};
    memcpy(ci, i1, sizeof ci); memcpy(cc, c1_, sizeof cc); memcpy(co, o1, sizeof co);
  } else {
    const int i2[8] = {7, 9, 8, 10, 11, 13, 12, 14};
    const int c2_[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    const ptrdiff_t o2[8] = {-X, 0, -X + Z, Z, -Y, 0, -Y + Z, Z};
    memcpy(ci, i2, sizeof ci); memcpy(cc, c2_, sizeof cc); memcpy(co, o2, sizeof co);
  }
  for (int t = 0; t < 8; t++) {
    idx[6 + t] = ci[t];
    comp[6 + t] = cc[t];
    off[6 + t] = co[t];
  }
}

/* One colour's forward update of component q (srcs/solu_sor4.f:131-175),
 * the same arithmetic as orc_solu_sor4. */
static void colour_fwd(int q, int col, int ni, int nj, int nk, int jen, double dtinv,
                       const double *rho, const double *rhon, double *const *cf,
                       const double *src, double *const *x) {
  ODIMS(ni, nj, nk);
  const ptrdiff_t X = 1, Y = sx, Z = sx * sy;
  const int jend = q == 1 ? jen : nj;
  const ptrdiff_t F = q == 0 ? X : (q == 1 ? Y : Z);
  double *const *c = cf + 15 * q;
  int idx[14], comp[14];
  ptrdiff_t off[14];
  sor_terms(q, X, Y, Z, idx, comp, off);
  for (int k = kKsor[col]; k <= nk; k += 2)
    for (int j = kJsor[col]; j <= jend; j += 2)
      for (int i = kIsor[col]; i <= ni; i += 2) {
        const ptrdiff_t m = OIX(i, j, k);
        const double r00 = (rho[m] + rho[m + F] + rhon[m] + rhon[m + F]) * 0.25;
        double sum = 0.0;
        for (int t = 0; t < 14; t++) sum = sum + c[idx[t]][m] * x[comp[t]][m + off[t]];
        const double visn = sum / r00;
        const double d = (visn * 0.5 + src[m]) / (dtinv + c[6][m] / r00 * 0.5) - x[q][m];
        x[q][m] = x[q][m] + d;
      }
}

/* colour_b with the parameter adjoints: for x = (0.5 S/r00 + src)/den,
 * den = dtinv + 0.5 a_p/r00 and S = sum a_t y_t over the state y before the
 * colour (xs): a_t_b += h y_t, a_p_b -= h x, r00_b = g (0.5 x a_p - 0.5 S)/r00^2
 * (g = x_b/den, h = 0.5 g/r00), r00 = (rho + rho(+F) + rhon + rhon(+F))/4. */
static void colour_bp(int q, int col, int ni, int nj, int nk, int jen, double dtinv,
                      const double *rho, double *rhob, const double *rhon, double *rhonb,
                      double *const *cf, double *const *cfb, const double *src, double *srcb,
                      double *const *xs, double *const *xb) {
  ODIMS(ni, nj, nk);
  const ptrdiff_t X = 1, Y = sx, Z = sx * sy;
  const int jend = q == 1 ? jen : nj;
  const ptrdiff_t F = q == 0 ? X : (q == 1 ? Y : Z);
  double *const *c = cf + 15 * q;
  double *const *cb = cfb + 15 * q;
  int idx[14], comp[14];
  ptrdiff_t off[14];
  sor_terms(q, X, Y, Z, idx, comp, off);
  for (int k = kKsor[col]; k <= nk; k += 2)
    for (int j = kJsor[col]; j <= jend; j += 2)
      for (int i = kIsor[col]; i <= ni; i += 2) {
        const ptrdiff_t m = OIX(i, j, k);
        const double r00 = (rho[m] + rho[m + F] + rhon[m] + rhon[m + F]) * 0.25;
        const double den = dtinv + c[6][m] / r00 * 0.5;
        double sum = 0.0;
        for (int t = 0; t < 14; t++) sum = sum + c[idx[t]][m] * xs[comp[t]][m + off[t]];
        const double x = (sum / r00 * 0.5 + src[m]) / den;
        const double g = xb[q][m] / den;
        const double h = g * 0.5 / r00;
        srcb[m] += g;
        for (int t = 0; t < 14; t++) cb[idx[t]][m] += h * xs[comp[t]][m + off[t]];
        cb[6][m] -= h * x;
        const double r00b = g * (0.5 * x * c[6][m] - 0.5 * sum) / (r00 * r00);
        rhob[m] += 0.25 * r00b;
        rhob[m + F] += 0.25 * r00b;
        rhonb[m] += 0.25 * r00b;
        rhonb[m + F] += 0.25 * r00b;
        xb[q][m] = 0.0;
        for (int t = 0; t < 14; t++) xb[comp[t]][m + off[t]] += c[idx[t]][m] * h;
      }
}

/* All-argument adjoint of solu_sor4 with the forward's sweep count lp
 * frozen: the forward sweeps are recomputed from the call's input state
 * (un, vn, wn, src_*) keeping the state before every colour and before every
 * refresh; the reverse is sor4_b_impl's with each colour through colour_bp
 * and each refresh's bndu through its all-argument adjoint at the kept
 * state.  Adds the adjoints of rho, rhon, the 45 coefficients, l1, l2 and
 * phi to what orc_solu_sor4_b gives. */
void orc_solu_sor4_bp(const int *ipara, const int *ID, const int *nID, const int *ndiv,
                      const int *ni_, const int *nj_, const int *nk_, const int *key,
                      double *sendjb, double *recvjb, const int *nstep, const int *imon_t,
                      const double *dt_, const double *rho, double *rhob, const double *rhon,
                      double *rhonb, double *c0, double *c1, double *c2, double *c3, double *c4, double *c5, double *c6, double *c7, double *c8, double *c9, double *c10, double *c11, double *c12, double *c13, double *c14, double *c15, double *c16, double *c17, double *c18, double *c19, double *c20, double *c21, double *c22, double *c23, double *c24, double *c25, double *c26, double *c27, double *c28, double *c29, double *c30, double *c31, double *c32, double *c33, double *c34, double *c35, double *c36, double *c37, double *c38, double *c39, double *c40, double *c41, double *c42, double *c43, double *c44, double *c0b, double *c1b, double *c2b, double *c3b, double *c4b, double *c5b, double *c6b, double *c7b, double *c8b, double *c9b, double *c10b, double *c11b, double *c12b, double *c13b, double *c14b, double *c15b, double *c16b, double *c17b, double *c18b, double *c19b, double *c20b, double *c21b, double *c22b, double *c23b, double *c24b, double *c25b, double *c26b, double *c27b, double *c28b, double *c29b, double *c30b, double *c31b, double *c32b, double *c33b, double *c34b, double *c35b, double *c36b, double *c37b, double *c38b, double *c39b, double *c40b, double *c41b, double *c42b, double *c43b, double *c44b,
                      const double *src_u, double *src_ub, const double *src_v, double *src_vb,
                      const double *src_w, double *src_wb, const double *un, double *unb,
                      const double *vn, double *vnb, const double *wn, double *wnb,
                      const double *uwall, double *uwallb, const double *dy, const double *l1,
                      double *l1b, const double *l2, double *l2b, const double *phi,
                      double *phib, const int *lp) {
  (void)ipara; (void)ID; (void)ndiv; (void)key; (void)sendjb; (void)recvjb; (void)nstep; (void)imon_t;
  double *const coef[NCOEF] = {c0, c1, c2, c3, c4, c5, c6, c7, c8, c9, c10, c11, c12, c13, c14, c15, c16, c17, c18, c19, c20, c21, c22, c23, c24, c25, c26, c27, c28, c29, c30, c31, c32, c33, c34, c35, c36, c37, c38, c39, c40, c41, c42, c43, c44};
  double *const coefb[NCOEF] = {c0b, c1b, c2b, c3b, c4b, c5b, c6b, c7b, c8b, c9b, c10b, c11b, c12b, c13b, c14b, c15b, c16b, c17b, c18b, c19b, c20b, c21b, c22b, c23b, c24b, c25b, c26b, c27b, c28b, c29b, c30b, c31b, c32b, c33b, c34b, c35b, c36b, c37b, c38b, c39b, c40b, c41b, c42b, c43b, c44b};
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  const double dtinv = 1.0 / (*dt_);
  const int jen = (nID[ORC_Y_PLUS - 1] < 0) ? nj - 1 : nj;
  const int nsw = *lp < 21 ? *lp + 1 : 21;
  const size_t blk = 3 * (size_t)n3;
  double *cur = malloc(sizeof(double) * blk);
  double *snap_c = malloc(sizeof(double) * blk * (size_t)nsw * 24);
  double *snap_r = malloc(sizeof(double) * blk * (size_t)nsw * 12);
  memcpy(cur, un, sizeof(double) * n3);
  memcpy(cur + n3, vn, sizeof(double) * n3);
  memcpy(cur + 2 * n3, wn, sizeof(double) * n3);
  double *const x[3] = {cur, cur + n3, cur + 2 * n3};
  const double *src[3] = {src_u, src_v, src_w};
  double *srcb[3] = {src_ub, src_vb, src_wb};
  double *xb[3] = {unb, vnb, wnb};
  for (int s = 0; s < nsw; s++)
    for (int q = 0; q < 3; q++)
      for (int col = 0; col < 8; col++) {
        memcpy(snap_c + blk * ((size_t)(s * 3 + q) * 8 + col), cur, sizeof(double) * blk);
        colour_fwd(q, col, ni, nj, nk, jen, dtinv, rho, rhon, coef, src[q], x);
        if (col % 2 == 1) {
          memcpy(snap_r + blk * ((size_t)(s * 3 + q) * 4 + col / 2), cur, sizeof(double) * blk);
          orc_bndu(nID, ni_, nj_, nk_, x[0], x[1], x[2], uwall, dy, l1, l2, phi);
          orc_bnd_periodic(ni_, nj_, nk_, x[q]);
        }
      }
  for (int s = nsw - 1; s >= 0; s--)
    for (int q = 2; q >= 0; q--)
      for (int col = 7; col >= 0; col--) {
        if (col % 2 == 1) {
          const double *r = snap_r + blk * ((size_t)(s * 3 + q) * 4 + col / 2);
          orc_bnd_periodic_b(ni_, nj_, nk_, xb[q]);
          bndu_b_core(nID, ni, nj, nk, unb, vnb, wnb, uwallb, *dy, l1, l2, phi, r, r + 2 * n3,
                      *uwall, l1b, l2b, phib);
        }
        double *c0p = snap_c + blk * ((size_t)(s * 3 + q) * 8 + col);
        double *const xs[3] = {c0p, c0p + n3, c0p + 2 * n3};
        colour_bp(q, col, ni, nj, nk, jen, dtinv, rho, rhob, rhon, rhonb, coef, coefb, src[q],
                  srcb[q], xs, xb);
      }
  free(cur);
  free(snap_c);
  free(snap_r);
}

/* All-argument adjoint of cal_arith_tau (orc_cal_arith_tau_b plus the
 * viscosity's adjoint): tau = 2 s mu_face, mu_face the cell value (normal
 * stresses) or the mean of four cells (shear stresses). */
void orc_cal_arith_tau_bp(const int *ni_, const int *nj_, const int *nk_, const double *s,
                          double *sb, const double *rmu, double *rmub, double *taub) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  const ptrdiff_t X = 1, Y = sx, Z = sx * sy;
  /* blocks 3..5: the shear stresses, their four-cell means and boxes */
  const ptrdiff_t da[3] = {X, X, Y}, db[3] = {Y, Z, Z};
  const int hi[3][3] = {{ni + 2, nj + 2, nk + 3}, {ni + 2, nj + 3, nk + 2}, {ni + 3, nj + 2, nk + 2}};
  for (int l = 5; l >= 3; l--) {
    const int b = l - 3;
    for (int k = hi[b][2]; k >= -2; k--)
      for (int j = hi[b][1]; j >= -2; j--)
        for (int i = hi[b][0]; i >= -2; i--) {
          const ptrdiff_t c = OIX(i, j, k);
          const double m = +0.25 * rmu[c] + 0.25 * rmu[c + da[b]] + 0.25 * rmu[c + db[b]] +
                           0.25 * rmu[c + da[b] + db[b]];
          const double t = taub[l * n3 + c];
          sb[l * n3 + c] += t * m * 2.0;
          const double mb = t * s[l * n3 + c] * 2.0;
          rmub[c] += 0.25 * mb;
          rmub[c + da[b]] += 0.25 * mb;
          rmub[c + db[b]] += 0.25 * mb;
          rmub[c + da[b] + db[b]] += 0.25 * mb;
          taub[l * n3 + c] = 0.0;
        }
  }
  for (int l = 2; l >= 0; l--)
    for (ptrdiff_t m = n3 - 1; m >= 0; m--) {
      sb[l * n3 + m] += taub[l * n3 + m] * rmu[m] * 2.0;
      rmub[m] += taub[l * n3 + m] * s[l * n3 + m] * 2.0;
      taub[l * n3 + m] = 0.0;
    }
}

/* All-argument adjoint of cal_srcu (orc_cal_srcu_b plus the densities'
 * adjoints): the face density r enters as -prs/r + 0.5 vis/r, so
 * r_b = src_b (prs - 0.5 vis)/r^2, spread over its four cells. */
void orc_cal_srcu_bp(const int *nID, const int *ni_, const int *nj_, const int *nk_,
                     const double *rho_av, const double *dt_, const double *grvb_,
                     const double *grvp_, double *ub, double *vb, double *wb, const double *rho,
                     double *rhob, const double *rhon, double *rhonb, double *adv_ub,
                     double *adv_vb, double *adv_wb, double *adv_uob, double *adv_vob,
                     double *adv_wob, const double *prs_u, double *prs_ub, const double *prs_v,
                     double *prs_vb, const double *prs_w, double *prs_wb, const double *vis_u,
                     double *vis_ub, const double *vis_v, double *vis_vb, const double *vis_w,
                     double *vis_wb, double *fst_ub, double *fst_vb, double *fst_wb,
                     double *fst_unb, double *fst_vnb, double *fst_wnb, double *src_ub,
                     double *src_vb, double *src_wb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  const int jen = (nID[ORC_Y_PLUS - 1] < 0) ? nj - 1 : nj;
  const ptrdiff_t F[3] = {1, sx, sx * sy};
  const double *prs[3] = {prs_u, prs_v, prs_w}, *vis[3] = {vis_u, vis_v, vis_w};
  double *src[3] = {src_ub, src_vb, src_wb};
  for (int q = 2; q >= 0; q--) {
    const int jend = q == 1 ? jen : nj;
    for (int k = nk; k >= 1; k--)
      for (int j = jend; j >= 1; j--)
        for (int i = ni; i >= 1; i--) {
          const ptrdiff_t c = OIX(i, j, k);
          const double r = (rho[c] + rho[c + F[q]] + rhon[c] + rhon[c + F[q]]) * 0.25;
          const double rb = src[q][c] * (prs[q][c] - 0.5 * vis[q][c]) / (r * r);
          rhob[c] += 0.25 * rb;
          rhob[c + F[q]] += 0.25 * rb;
          rhonb[c] += 0.25 * rb;
          rhonb[c + F[q]] += 0.25 * rb;
        }
  }
  orc_cal_srcu_b(nID, ni_, nj_, nk_, rho_av, dt_, grvb_, grvp_, ub, vb, wb, rho, rhon, adv_ub,
                 adv_vb, adv_wb, adv_uob, adv_vob, adv_wob, prs_ub, prs_vb, prs_wb, vis_ub,
                 vis_vb, vis_wb, fst_ub, fst_vb, fst_wb, fst_unb, fst_vnb, fst_wnb, src_ub,
                 src_vb, src_wb);
}

/* srcs/cal_arith_coef_vis.f:1-411 (coef_vis in fel_oracle.c), linear in
 * the viscosity: each coefficient of cell c is a multiple of one of six
 * face viscosities (xm, xp, ym, yp, zm, zp: a cell value or a mean of four)
 * and a_p their weighted sum.  Transpose: face adjoints from the 15
 * coefficient adjoints of each component, spread to the cells of each
 * face, all coefficient adjoints consumed (the forward overwrites every
 * coefficient array whole). */
void orc_cal_arith_coef_vis_b(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                              const double *dyinv_, const double *dzinv_, double *rmukb,
                              double *c0b, double *c1b, double *c2b, double *c3b, double *c4b, double *c5b, double *c6b, double *c7b, double *c8b, double *c9b, double *c10b, double *c11b, double *c12b, double *c13b, double *c14b, double *c15b, double *c16b, double *c17b, double *c18b, double *c19b, double *c20b, double *c21b, double *c22b, double *c23b, double *c24b, double *c25b, double *c26b, double *c27b, double *c28b, double *c29b, double *c30b, double *c31b, double *c32b, double *c33b, double *c34b, double *c35b, double *c36b, double *c37b, double *c38b, double *c39b, double *c40b, double *c41b, double *c42b, double *c43b, double *c44b) {
  double *const a[NCOEF] = {c0b, c1b, c2b, c3b, c4b, c5b, c6b, c7b, c8b, c9b, c10b, c11b, c12b, c13b, c14b, c15b, c16b, c17b, c18b, c19b, c20b, c21b, c22b, c23b, c24b, c25b, c26b, c27b, c28b, c29b, c30b, c31b, c32b, c33b, c34b, c35b, c36b, c37b, c38b, c39b, c40b, c41b, c42b, c43b, c44b};
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  ODIMS(ni, nj, nk);
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  const double dx2 = dxinv * dxinv, dy2 = dyinv * dyinv, dz2 = dzinv * dzinv;
  const double dxy = dxinv * dyinv, dxz = dxinv * dzinv, dyz = dyinv * dzinv;
#define RB(i, j, k) rmukb[OIX(i, j, k)]
  for (int k = nk + 2; k >= -1; k--)
    for (int j = nj + 2; j >= -1; j--)
      for (int i = ni + 2; i >= -1; i--) {
        const ptrdiff_t c = OIX(i, j, k);
        /* w component (cal_arith_coef_vis.f, third block) */
        {
          const double P = a[CW_P][c];
          const double bb = a[CW_B][c] + P, sb = a[CW_S][c] + P, wb = a[CW_W][c] + P;
          const double eb = a[CW_E][c] + P, nb = a[CW_N][c] + P, tb = a[CW_T][c] + P;
          const double xmb = dx2 * wb + dxz * (a[CW_UBW][c] - a[CW_UTW][c]);
          const double xpb = dx2 * eb + dxz * (-a[CW_UBE][c] + a[CW_UTE][c]);
          const double ymb = dy2 * sb + dyz * (a[CW_VBS][c] - a[CW_VTS][c]);
          const double ypb = dy2 * nb + dyz * (-a[CW_VBN][c] + a[CW_VTN][c]);
          const double zmb = dz2 * bb * 2.0, zpb = dz2 * tb * 2.0;
          RB(i, j, k) += zmb;
          RB(i, j, k + 1) += zpb;
          RB(i - 1, j, k) += 0.25 * xmb; RB(i, j, k) += 0.25 * xmb;
          RB(i - 1, j, k + 1) += 0.25 * xmb; RB(i, j, k + 1) += 0.25 * xmb;
          RB(i, j, k) += 0.25 * xpb; RB(i + 1, j, k) += 0.25 * xpb;
          RB(i, j, k + 1) += 0.25 * xpb; RB(i + 1, j, k + 1) += 0.25 * xpb;
          RB(i, j - 1, k) += 0.25 * ymb; RB(i, j, k) += 0.25 * ymb;
          RB(i, j - 1, k + 1) += 0.25 * ymb; RB(i, j, k + 1) += 0.25 * ymb;
          RB(i, j, k) += 0.25 * ypb; RB(i, j + 1, k) += 0.25 * ypb;
          RB(i, j, k + 1) += 0.25 * ypb; RB(i, j + 1, k + 1) += 0.25 * ypb;
        }
        /* v component */
        {
          const double P = a[CV_P][c];
          const double bb = a[CV_B][c] + P, sb = a[CV_S][c] + P, wb = a[CV_W][c] + P;
          const double eb = a[CV_E][c] + P, nb = a[CV_N][c] + P, tb = a[CV_T][c] + P;
          const double xmb = dx2 * wb + dxy * (a[CV_USW][c] - a[CV_UNW][c]);
          const double xpb = dx2 * eb + dxy * (-a[CV_USE][c] + a[CV_UNE][c]);
          const double ymb = dy2 * sb * 2.0, ypb = dy2 * nb * 2.0;
          const double zmb = dz2 * bb + dyz * (a[CV_WSB][c] - a[CV_WNB][c]);
          const double zpb = dz2 * tb + dyz * (-a[CV_WST][c] + a[CV_WNT][c]);
          RB(i, j, k) += ymb;
          RB(i, j + 1, k) += ypb;
          RB(i - 1, j, k) += 0.25 * xmb; RB(i, j, k) += 0.25 * xmb;
          RB(i - 1, j + 1, k) += 0.25 * xmb; RB(i, j + 1, k) += 0.25 * xmb;
          RB(i, j, k) += 0.25 * xpb; RB(i + 1, j, k) += 0.25 * xpb;
          RB(i, j + 1, k) += 0.25 * xpb; RB(i + 1, j + 1, k) += 0.25 * xpb;
          RB(i, j, k - 1) += 0.25 * zmb; RB(i, j + 1, k - 1) += 0.25 * zmb;
          RB(i, j, k) += 0.25 * zmb; RB(i, j + 1, k) += 0.25 * zmb;
          RB(i, j, k) += 0.25 * zpb; RB(i, j + 1, k) += 0.25 * zpb;
          RB(i, j, k + 1) += 0.25 * zpb; RB(i, j + 1, k + 1) += 0.25 * zpb;
        }
        /* u component */
        {
          const double P = a[CU_P][c];
          const double bb = a[CU_B][c] + P, sb = a[CU_S][c] + P, wb = a[CU_W][c] + P;
          const double eb = a[CU_E][c] + P, nb = a[CU_N][c] + P, tb = a[CU_T][c] + P;
          const double xmb = dx2 * wb * 2.0, xpb = dx2 * eb * 2.0;
          const double ymb = dy2 * sb + dxy * (a[CU_VWS][c] - a[CU_VES][c]);
          const double ypb = dy2 * nb + dxy * (-a[CU_VWN][c] + a[CU_VEN][c]);
          const double zmb = dz2 * bb + dxz * (a[CU_WWB][c] - a[CU_WEB][c]);
          const double zpb = dz2 * tb + dxz * (-a[CU_WWT][c] + a[CU_WET][c]);
          RB(i, j, k) += xmb;
          RB(i + 1, j, k) += xpb;
          RB(i, j - 1, k) += 0.25 * ymb; RB(i + 1, j - 1, k) += 0.25 * ymb;
          RB(i, j, k) += 0.25 * ymb; RB(i + 1, j, k) += 0.25 * ymb;
          RB(i, j, k) += 0.25 * ypb; RB(i + 1, j, k) += 0.25 * ypb;
          RB(i, j + 1, k) += 0.25 * ypb; RB(i + 1, j + 1, k) += 0.25 * ypb;
          RB(i, j, k - 1) += 0.25 * zmb; RB(i + 1, j, k - 1) += 0.25 * zmb;
          RB(i, j, k) += 0.25 * zmb; RB(i + 1, j, k) += 0.25 * zmb;
          RB(i, j, k) += 0.25 * zpb; RB(i + 1, j, k) += 0.25 * zpb;
          RB(i, j, k + 1) += 0.25 * zpb; RB(i + 1, j, k + 1) += 0.25 * zpb;
        }
      }
#undef RB
  for (int cc = 0; cc < NCOEF; cc++) memset(a[cc], 0, sizeof(double) * n3);
}

/* srcs/cal_arith_mean.f:18-26: q = ql (1-phi) + qg phi on every element.
 * Transpose: phi_b += (qg - ql) q_b, q_b = 0. */
void orc_cal_arith_mean_b(const int *ni_, const int *nj_, const int *nk_, double *phikb,
                          const double *ql_, const double *qg_, double *qkb) {
  ODIMS(*ni_, *nj_, *nk_);
  const double ql = *ql_, qg = *qg_;
  for (ptrdiff_t m = n3 - 1; m >= 0; m--) {
    const double plb = ql * qkb[m], pgb = qg * qkb[m];
    phikb[m] += -plb + pgb;
    qkb[m] = 0.0;
  }
}

/* srcs/calrn.f:1-95 (orc_calrn), nonlinear: rn = r/(|r| + verysmall) with r
 * the mean of eight vertex gradients, on the box of the staggering
 * (ip, jp, kp); zero elsewhere.  Reverse: r_b = rn_b/a - (rn_b . r) r/(a^2
 * |r|) with a = |r| + verysmall (the |r| term taken as 0 at r = 0), spread
 * over the eight vertices; rn_b consumed whole. */
void orc_calrn_b(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                 const double *dyinv_, const double *dzinv_, const int *ip_, const int *jp_,
                 const int *kp_, const double *phix, double *phixb, const double *phiy,
                 double *phiyb, const double *phiz, double *phizb, double *rnb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_, ip = *ip_, jp = *jp_, kp = *kp_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  const double verysmall = sqrt(dxinv * dxinv + dyinv * dyinv + dzinv * dzinv) * 1.0e-30;
  const double *a[3] = {phix, phiy, phiz};
  double *ab[3] = {phixb, phiyb, phizb};
  for (int k = nk + 3 - kp; k >= -2; k--)
    for (int j = nj + 3 - jp; j >= -2; j--)
      for (int i = ni + 3 - ip; i >= -2; i--) {
        const int ii = i + ip, jj = j + jp, kk = k + kp;
        const ptrdiff_t v[8] = {OIX(i, j, k), OIX(ii, j, k), OIX(i, jj, k), OIX(ii, jj, k),
                                OIX(i, j, kk), OIX(ii, j, kk), OIX(i, jj, kk), OIX(ii, jj, kk)};
        double r[3];
        for (int c = 0; c < 3; c++) {
          const double *p = a[c];
          r[c] = (+p[v[0]] + p[v[1]] + p[v[2]] + p[v[3]] + p[v[4]] + p[v[5]] + p[v[6]] + p[v[7]]) *
                 0.125;
        }
        const double nr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double inv = 1.0 / (nr + verysmall);
        const ptrdiff_t o = OIX(ii, jj, kk);
        const double gb[3] = {rnb[o], rnb[n3 + o], rnb[2 * n3 + o]};
        const double invb = gb[0] * r[0] + gb[1] * r[1] + gb[2] * r[2];
        const double nrb = -invb * inv * inv;
        for (int c = 0; c < 3; c++) {
          const double rb = gb[c] * inv + (nr > 0.0 ? nrb * r[c] / nr : 0.0);
          for (int t = 0; t < 8; t++) ab[c][v[t]] += rb * 0.125;
        }
        rnb[o] = 0.0;
        rnb[n3 + o] = 0.0;
        rnb[2 * n3 + o] = 0.0;
      }
  memset(rnb, 0, sizeof(double) * 3 * n3);
}

/* srcs/calrkap.f:1-102 (orc_calrkap), linear: the three curvature
 * components are differences of normal components on their boxes, zero
 * elsewhere.  Transpose per component, rkap_b consumed whole. */
void orc_calrkap_b(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_,
                   const double *dyinv_, const double *dzinv_, double *rn_ppb, double *rn_12b,
                   double *rn_13b, double *rn_23b, double *rkapb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  const ptrdiff_t C2 = n3, C3 = 2 * n3;
  for (int k = nk + 2; k >= -2; k--)
    for (int j = nj + 3; j >= -1; j--)
      for (int i = ni + 3; i >= -1; i--) {
        const double t = rkapb[C3 + OIX(i, j, k)];
        rn_13b[OIX(i - 1, j, k)] -= t * dxinv;
        rn_13b[OIX(i, j, k)] += t * dxinv;
        rn_23b[C2 + OIX(i, j - 1, k)] -= t * dyinv;
        rn_23b[C2 + OIX(i, j, k)] += t * dyinv;
        rn_ppb[C3 + OIX(i, j, k)] -= t * dzinv;
        rn_ppb[C3 + OIX(i, j, k + 1)] += t * dzinv;
      }
  for (int k = nk + 3; k >= -1; k--)
    for (int j = nj + 2; j >= -2; j--)
      for (int i = ni + 3; i >= -1; i--) {
        const double t = rkapb[C2 + OIX(i, j, k)];
        rn_12b[OIX(i - 1, j, k)] -= t * dxinv;
        rn_12b[OIX(i, j, k)] += t * dxinv;
        rn_ppb[C2 + OIX(i, j, k)] -= t * dyinv;
        rn_ppb[C2 + OIX(i, j + 1, k)] += t * dyinv;
        rn_23b[C3 + OIX(i, j, k - 1)] -= t * dzinv;
        rn_23b[C3 + OIX(i, j, k)] += t * dzinv;
      }
  for (int k = nk + 3; k >= -1; k--)
    for (int j = nj + 3; j >= -1; j--)
      for (int i = ni + 2; i >= -2; i--) {
        const double t = rkapb[OIX(i, j, k)];
        rn_ppb[OIX(i, j, k)] -= t * dxinv;
        rn_ppb[OIX(i + 1, j, k)] += t * dxinv;
        rn_12b[C2 + OIX(i, j - 1, k)] -= t * dyinv;
        rn_12b[C2 + OIX(i, j, k)] += t * dyinv;
        rn_13b[C3 + OIX(i, j, k - 1)] -= t * dzinv;
        rn_13b[C3 + OIX(i, j, k)] += t * dzinv;
      }
  memset(rkapb, 0, sizeof(double) * 3 * n3);
}

/* srcs/cal_grad_p2uvw.f:1-47 (orc_cal_grad_p2uvw), linear: forward
 * differences on [-2,n+2]^3; the rest of the outputs is left as it was.
 * Transpose on the box, the outputs' adjoints consumed there. */
void orc_cal_grad_p2uvw_b(const int *ID, const int *svall, const int *ni_, const int *nj_,
                          const int *nk_, const double *dxinv_, const double *dyinv_array,
                          const double *dzinv_, double *qb, double *qx_ub, double *qy_vb,
                          double *qz_wb) {
  const int ni = *ni_, nj = *nj_, nk = *nk_;
  const double dxinv = *dxinv_, dzinv = *dzinv_;
  ODIMS(ni, nj, nk);
  (void)svall;
  for (int k = nk + 2; k >= -2; k--)
    for (int j = nj + 2; j >= -2; j--)
      for (int i = ni + 2; i >= -2; i--) {
        const double dyinv = dyinv_array[j + nj * (*ID) + 2];
        const ptrdiff_t c = OIX(i, j, k);
        const double gz = qz_wb[c] * dzinv, gy = qy_vb[c] * dyinv, gx = qx_ub[c] * dxinv;
        qb[OIX(i, j, k + 1)] += gz;
        qb[OIX(i, j + 1, k)] += gy;
        qb[OIX(i + 1, j, k)] += gx;
        qb[c] += -gz - gy - gx;
        qx_ub[c] = 0.0;
        qy_vb[c] = 0.0;
        qz_wb[c] = 0.0;
      }
}

/* srcs/calfst.f:1-50 (orc_calfst), bilinear: f = -(st kappa g) with st =
 * 2 sigma/(rho_l + rho_g) on every element.  Reverse: g_b += -st kappa f_b,
 * kappa_b += -st g f_b, f_b = 0. */
void orc_calfst_b(const int *ni_, const int *nj_, const int *nk_, const double *rhol,
                  const double *rhog, const double *surface_tension, const double *phix_u,
                  double *phix_ub, const double *phiy_v, double *phiy_vb, const double *phiz_w,
                  double *phiz_wb, const double *rkap, double *rkapb, double *fst_ub,
                  double *fst_vb, double *fst_wb) {
  ODIMS(*ni_, *nj_, *nk_);
  const double st = 2.0 * (*surface_tension) / (*rhol + *rhog);
  const double *g[3] = {phix_u, phiy_v, phiz_w};
  double *gb[3] = {phix_ub, phiy_vb, phiz_wb}, *fb[3] = {fst_ub, fst_vb, fst_wb};
  for (ptrdiff_t m = n3 - 1; m >= 0; m--)
    for (int c = 2; c >= 0; c--) {
      const double t = fb[c][m];
      gb[c][m] += -(st * rkap[c * n3 + m] * t);
      rkapb[c * n3 + m] += -(st * g[c][m] * t);
      fb[c][m] = 0.0;
    }
}

/* srcs/sum_fst.f:1-37: s = s + f.  Transpose: f_b += s_b (s_b unchanged). */
void orc_sum_fst_b(const int *ni_, const int *nj_, const int *nk_, double *fub, double *fvb,
                   double *fwb, const double *sub, const double *svb, const double *swb) {
  ODIMS(*ni_, *nj_, *nk_);
  for (ptrdiff_t m = 0; m < n3; m++) {
    fub[m] += sub[m];
    fvb[m] += svb[m];
    fwb[m] += swb[m];
  }
}

/* All-argument adjoint of cal_arith_coef_vis followed by solu_sor4 (the
 * coefficients formed from the viscosity rmu): the composition of
 * orc_solu_sor4_bp and orc_cal_arith_coef_vis_b, with the forward's 45
 * coefficients from orc_cal_arith_coef_vis.  rmub incremented. */
void orc_solu_sor4_bpv(const int *ipara, const int *ID, const int *nID, const int *ndiv,
                       const int *ni_, const int *nj_, const int *nk_, const int *key,
                       double *sendjb, double *recvjb, const int *nstep, const int *imon_t,
                       const double *dt_, const double *rho, double *rhob, const double *rhon,
                       double *rhonb, const double *dxinv, const double *dyinv,
                       const double *dzinv, const double *rmu, double *rmub,
                       const double *src_u, double *src_ub, const double *src_v, double *src_vb,
                       const double *src_w, double *src_wb, const double *un, double *unb,
                       const double *vn, double *vnb, const double *wn, double *wnb,
                       const double *uwall, double *uwallb, const double *dy, const double *l1,
                       double *l1b, const double *l2, double *l2b, const double *phi,
                       double *phib, const int *lp) {
  ODIMS(*ni_, *nj_, *nk_);
  double *cf[NCOEF], *cb[NCOEF];
  for (int m = 0; m < NCOEF; m++) {
    cf[m] = calloc((size_t)n3, sizeof(double));
    cb[m] = calloc((size_t)n3, sizeof(double));
  }
  orc_cal_arith_coef_vis(ni_, nj_, nk_, dxinv, dyinv, dzinv, rmu, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5], cf[6], cf[7], cf[8], cf[9], cf[10], cf[11], cf[12], cf[13], cf[14], cf[15], cf[16], cf[17], cf[18], cf[19], cf[20], cf[21], cf[22], cf[23], cf[24], cf[25], cf[26], cf[27], cf[28], cf[29], cf[30], cf[31], cf[32], cf[33], cf[34], cf[35], cf[36], cf[37], cf[38], cf[39], cf[40], cf[41], cf[42], cf[43], cf[44]);
  orc_solu_sor4_bp(ipara, ID, nID, ndiv, ni_, nj_, nk_, key, sendjb, recvjb, nstep, imon_t, dt_,
                   rho, rhob, rhon, rhonb, cf[0], cf[1], cf[2], cf[3], cf[4], cf[5], cf[6], cf[7], cf[8], cf[9], cf[10], cf[11], cf[12], cf[13], cf[14], cf[15], cf[16], cf[17], cf[18], cf[19], cf[20], cf[21], cf[22], cf[23], cf[24], cf[25], cf[26], cf[27], cf[28], cf[29], cf[30], cf[31], cf[32], cf[33], cf[34], cf[35], cf[36], cf[37], cf[38], cf[39], cf[40], cf[41], cf[42], cf[43], cf[44],
                   cb[0], cb[1], cb[2], cb[3], cb[4], cb[5], cb[6], cb[7], cb[8], cb[9], cb[10], cb[11], cb[12], cb[13], cb[14], cb[15], cb[16], cb[17], cb[18], cb[19], cb[20], cb[21], cb[22], cb[23], cb[24], cb[25], cb[26], cb[27], cb[28], cb[29], cb[30], cb[31], cb[32], cb[33], cb[34], cb[35], cb[36], cb[37], cb[38], cb[39], cb[40], cb[41], cb[42], cb[43], cb[44],
                   src_u, src_ub, src_v, src_vb, src_w, src_wb, un, unb, vn, vnb, wn, wnb, uwall,
                   uwallb, dy, l1, l1b, l2, l2b, phi, phib, lp);
  orc_cal_arith_coef_vis_b(ni_, nj_, nk_, dxinv, dyinv, dzinv, rmub, cb[0], cb[1], cb[2], cb[3], cb[4], cb[5], cb[6], cb[7], cb[8], cb[9], cb[10], cb[11], cb[12], cb[13], cb[14], cb[15], cb[16], cb[17], cb[18], cb[19], cb[20], cb[21], cb[22], cb[23], cb[24], cb[25], cb[26], cb[27], cb[28], cb[29], cb[30], cb[31], cb[32], cb[33], cb[34], cb[35], cb[36], cb[37], cb[38], cb[39], cb[40], cb[41], cb[42], cb[43], cb[44]);
  for (int m = 0; m < NCOEF; m++) {
    free(cf[m]);
    free(cb[m]);
  }
}
