/*
 * fel_oracle_step.c -- CPU restatement of the reference's start-up boundary
 * fills (srcs/main.f90:303-366,467-468) and of one timestep
 * (srcs/main.f90:473-748), composed of the orc_* routines of fel_oracle.c in
 * the reference's call order.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the GPU path; see fel_oracle.h).
 *
 * The state owns every array the reference allocates for the step
 * (srcs/allocate.h:7-165) that the step touches.  Fields are exchanged with
 * the caller through orc_state_field(), which returns the padded array.
 *
 * Deviation from the reference, by necessity: srcs/main.f90:337-339,623,625,650
 * pass phi(-2,-2,-2,l) after the loop "do l=1,nbub", i.e. l = nbub+1, one
 * slice past phi(...,0:nbub).  This restatement passes slice 1 (the only
 * defined choice; with l1 == l2 the slip length does not depend on it).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fel_oracle.h"

#include "fel_oracle_api.h"

extern int orc_sor_last_lp;
extern double orc_sor_last_ratio;

/* Physical and numerical parameters of one run (srcs/main.f90:91-249). */
typedef struct {
  int ni, nj, nk;
  int imon_t;
  double dx, dy, dz, dt;
  double bet_mthinc, rhol, rhog, rmul, rmug, surface_tension, uwall, grvb, grvp;
  double xl, yl, zl, particle_radius;
  double cfl; /* > 0: adaptive dt by caldt each step (srcs/main.f90:479); 0: fixed dt */
} orc_cfg;

enum {
  F_U, F_V, F_W, F_UO, F_VO, F_WO, F_UN, F_VN, F_WN, F_P, F_PO, F_PN, F_PHAT, F_DP, F_DIV,
  F_RHO, F_RMU, F_RHON, F_RMUN, F_ADVU, F_ADVV, F_ADVW, F_ADVUO, F_ADVVO, F_ADVWO,
  F_PRSU, F_PRSV, F_PRSW, F_VISU, F_VISV, F_VISW, F_FSTU, F_FSTV, F_FSTW, F_FSTUN, F_FSTVN,
  F_FSTWN, F_SFU, F_SFV, F_SFW, F_SFUN, F_SFVN, F_SFWN, F_SRCU, F_SRCV, F_SRCW,
  F_PHIX, F_PHIY, F_PHIZ, F_FLX, F_FLY, F_FLZ, F_TH0, F_L1, F_L2, F_ZETA, F_THETA,
  NF3
};
static const char *f3names[NF3] = {
  "u", "v", "w", "uo", "vo", "wo", "un", "vn", "wn", "p", "po", "pn", "phat", "dp", "div",
  "rho", "rmu", "rhon", "rmun", "adv_u", "adv_v", "adv_w", "adv_uo", "adv_vo", "adv_wo",
  "prs_u", "prs_v", "prs_w", "vis_u", "vis_v", "vis_w", "fst_u", "fst_v", "fst_w", "fst_un",
  "fst_vn", "fst_wn", "sum_fst_u", "sum_fst_v", "sum_fst_w", "sum_fst_un", "sum_fst_vn",
  "sum_fst_wn", "src_u", "src_v", "src_w", "phix", "phiy", "phiz", "flphix", "flphiy",
  "flphiz", "theta_0_array", "l1_array", "l2_array", "zeta_array", "theta_array"};

typedef struct {
  orc_cfg c;
  int svall[3], nID[6], ID, ipara, ndiv, nbub, nstep;
  ptrdiff_t n3;
  double time, dxinv, dyinv, dzinv, err_div0;
  double *f3[NF3];
  double *phi, *phin;         /* (…,0:1) */
  double *s, *tau;            /* (…,6) */
  double *rn[4], *rnn[4], *rkap, *rkapn; /* (…,3) */
  double *coef[45];
  double *dyinv_array;        /* (-2:nj+3) */
  double *pcoef[7];           /* aw,ae,as,an,ab,at,ap (-2:nj+3) */
  double *as_fft, *an_fft, *ap_fft, *atdma, *btr, *bti;
  double *phir_r, *phir_i, *phiw_r, *phiw_i, *fftdata;
  /* diagnostics state (srcs/main.f90:84-86) */
  double center[3], center_pre1, center_pre2, velocity, vel[3], re[3];
  int area_pre1, area_pre2, lap;
  int last_sor_lp;
  double last_sor_ratio;
  double dt_cur; /* this step's dt */
} orc_state;

static double *zalloc(size_t n) {
  double *p = calloc(n, sizeof(double));
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

orc_state *orc_state_create(const orc_cfg *cfg) {
  orc_state *S = calloc(1, sizeof(orc_state));
  S->c = *cfg;
  const int ni = cfg->ni, nj = cfg->nj, nk = cfg->nk;
  S->svall[0] = ni; S->svall[1] = nj; S->svall[2] = nk;
  for (int m = 0; m < 6; m++) S->nID[m] = -1;
  S->ID = 0; S->ipara = 0; S->ndiv = 1; S->nbub = 1; S->nstep = 0;
  S->n3 = (ptrdiff_t)(ni + 6) * (nj + 6) * (nk + 6);
  S->dxinv = 1.0 / cfg->dx; S->dyinv = 1.0 / cfg->dy; S->dzinv = 1.0 / cfg->dz;
  for (int m = 0; m < NF3; m++) S->f3[m] = zalloc(S->n3);
  S->phi = zalloc(2 * S->n3);
  S->phin = zalloc(2 * S->n3);
  S->s = zalloc(6 * S->n3);
  S->tau = zalloc(6 * S->n3);
  for (int m = 0; m < 4; m++) { S->rn[m] = zalloc(3 * S->n3); S->rnn[m] = zalloc(3 * S->n3); }
  S->rkap = zalloc(3 * S->n3);
  S->rkapn = zalloc(3 * S->n3);
  for (int m = 0; m < 45; m++) S->coef[m] = zalloc(S->n3);
  const int nj1 = nj + 6, nmax1 = (nj > nk ? nj : nk) + 6;
  S->dyinv_array = zalloc(nj1);
  for (int j = 0; j < nj1; j++) S->dyinv_array[j] = 1.0 / cfg->dy;
  for (int m = 0; m < 7; m++) S->pcoef[m] = zalloc(nj1);
  S->as_fft = zalloc(nj1); S->an_fft = zalloc(nj1); S->ap_fft = zalloc(nj1);
  S->atdma = zalloc(nmax1); S->btr = zalloc(nmax1); S->bti = zalloc(nmax1);
  const size_t nm = (size_t)ni * nj * nk;
  S->phir_r = zalloc(nm); S->phir_i = zalloc(nm); S->phiw_r = zalloc(nm); S->phiw_i = zalloc(nm);
  S->fftdata = zalloc(2 * nm);
  return S;
}

void orc_state_destroy(orc_state *S) {
  if (!S) return;
  for (int m = 0; m < NF3; m++) free(S->f3[m]);
  free(S->phi); free(S->phin); free(S->s); free(S->tau);
  for (int m = 0; m < 4; m++) { free(S->rn[m]); free(S->rnn[m]); }
  free(S->rkap); free(S->rkapn);
  for (int m = 0; m < 45; m++) free(S->coef[m]);
  free(S->dyinv_array);
  for (int m = 0; m < 7; m++) free(S->pcoef[m]);
  free(S->as_fft); free(S->an_fft); free(S->ap_fft); free(S->atdma); free(S->btr); free(S->bti);
  free(S->phir_r); free(S->phir_i); free(S->phiw_r); free(S->phiw_i); free(S->fftdata);
  free(S);
}

/* Padded (-2:n+3)^3 array by name; "phi0"/"phi1"/"phin0"/"phin1" select
 * slices; "coef<m>" the m-th viscous coefficient array. */
double *orc_state_field(orc_state *S, const char *name) {
  for (int m = 0; m < NF3; m++)
    if (!strcmp(name, f3names[m])) return S->f3[m];
  if (!strcmp(name, "phi0")) return S->phi;
  if (!strcmp(name, "phi1")) return S->phi + S->n3;
  if (!strcmp(name, "phin0")) return S->phin;
  if (!strcmp(name, "phin1")) return S->phin + S->n3;
  if (!strncmp(name, "coef", 4)) return S->coef[atoi(name + 4)];
  if (!strcmp(name, "rkap")) return S->rkap;
  if (!strcmp(name, "rkapn")) return S->rkapn;
  return NULL;
}

int orc_state_nstep(const orc_state *S) { return S->nstep; }
int orc_state_last_sor_lp(const orc_state *S) { return S->last_sor_lp; }
double orc_state_last_sor_ratio(const orc_state *S) { return S->last_sor_ratio; }
double orc_state_err_div0(const orc_state *S) { return S->err_div0; }
void orc_state_diag(const orc_state *S, double out[8]) {
  out[0] = S->center[0]; out[1] = S->center[1]; out[2] = S->center[2]; out[3] = S->velocity;
  out[4] = S->vel[0]; out[5] = S->vel[1]; out[6] = S->vel[2]; out[7] = S->time;
}

double orc_state_dt(const orc_state *S) { return S->dt_cur; }

#define FLD(x) (S->f3[x])

/* Start-up boundary fills, srcs/main.f90:322-366, then the pressure
 * coefficients of srcs/main.f90:467-468.  The caller has stored u, v, w, p,
 * phi1 and the wall maps (theta_0, l1, l2, zeta) beforehand (the reference's
 * init + init_phi + init_array_x_stripe). */
void orc_state_startup(orc_state *S) {
  const int *ni = &S->c.ni, *nj = &S->c.nj, *nk = &S->c.nk;
  double *phi1 = S->phi + S->n3, *phi0 = S->phi;
  orc_cpy(ni, nj, nk, FLD(F_U), FLD(F_UN));
  orc_cpy(ni, nj, nk, FLD(F_V), FLD(F_VN));
  orc_cpy(ni, nj, nk, FLD(F_W), FLD(F_WN));
  orc_gnbc(S->nID, ni, nj, nk, FLD(F_U), FLD(F_W), &S->c.uwall, FLD(F_TH0), &S->c.surface_tension,
           FLD(F_ZETA), FLD(F_THETA));
  orc_bnd_contact_angle(S->nID, ni, nj, nk, phi1, FLD(F_THETA), &S->c.dx, &S->c.dy, &S->c.dz);
  orc_bnd_periodic(ni, nj, nk, phi1);
  orc_summation(ni, nj, nk, S->phi, &S->nbub);
  orc_bndu(S->nID, ni, nj, nk, FLD(F_U), FLD(F_V), FLD(F_W), &S->c.uwall, &S->c.dy, FLD(F_L1),
           FLD(F_L2), phi1);
  orc_bndu(S->nID, ni, nj, nk, FLD(F_UO), FLD(F_VO), FLD(F_WO), &S->c.uwall, &S->c.dy, FLD(F_L1),
           FLD(F_L2), phi1);
  orc_bndu(S->nID, ni, nj, nk, FLD(F_UN), FLD(F_VN), FLD(F_WN), &S->c.uwall, &S->c.dy, FLD(F_L1),
           FLD(F_L2), phi1);
  const int per[9] = {F_U, F_V, F_W, F_UO, F_VO, F_WO, F_UN, F_VN, F_WN};
  for (int m = 0; m < 9; m++) orc_bnd_periodic(ni, nj, nk, FLD(per[m]));
  orc_gnbc(S->nID, ni, nj, nk, FLD(F_U), FLD(F_W), &S->c.uwall, FLD(F_TH0), &S->c.surface_tension,
           FLD(F_ZETA), FLD(F_THETA));
  orc_bnd_contact_angle(S->nID, ni, nj, nk, phi0, FLD(F_THETA), &S->c.dx, &S->c.dy, &S->c.dz);
  orc_bnd_periodic(ni, nj, nk, phi0);
  orc_cal_coef_prs(&S->svall[1], &S->c.rhog, &S->dxinv, &S->dyinv, &S->dzinv, S->pcoef[0],
                   S->pcoef[1], S->pcoef[2], S->pcoef[3], S->pcoef[4], S->pcoef[5], S->pcoef[6]);
  orc_cal_coef_solp_fft(&S->svall[1], S->pcoef[2], S->pcoef[3], S->as_fft, S->an_fft, S->ap_fft);
  S->time = 0.0;
  S->nstep = 0;
  S->area_pre1 = S->area_pre2 = S->lap = 0;
  S->center_pre1 = S->center_pre2 = 0.0;
}

/* One timestep, srcs/main.f90:473-748 (output cadences excluded). */
static void orc_state_step1(orc_state *S) {
  const orc_cfg *c = &S->c;
  const int *ni = &c->ni, *nj = &c->nj, *nk = &c->nk;
  const int *nID = S->nID;
  double *phi1 = S->phi + S->n3, *phi0 = S->phi;
  double *phin1 = S->phin + S->n3, *phin0 = S->phin;
  const double *dxinv = &S->dxinv, *dyinv = &S->dyinv, *dzinv = &S->dzinv, *dt = &S->dt_cur;
  int key[8] = {0};
  double sendjb[1], recvjb[1];
  S->nstep++;
  if (c->cfl > 0.0) {
    /* srcs/main.f90:479 (commented out in the reference, which runs a fixed dt) */
    orc_caldt(&S->ipara, nID, &S->ID, &S->ndiv, ni, nj, nk, &S->nstep, &c->imon_t, dxinv, dyinv,
              dzinv, &c->cfl, &c->rhol, &c->rhog, &c->rmul, &c->rmug, &c->surface_tension,
              FLD(F_U), FLD(F_V), FLD(F_W), &S->dt_cur, &S->time, NULL);
  } else {
    S->dt_cur = c->dt;
    S->time = S->time + c->dt;
  }
  /* VOF (main.f90:489-504) */
  orc_cal_grad_p2a(&S->ID, &S->svall[1], ni, nj, nk, dxinv, S->dyinv_array, dzinv, phi1,
                   FLD(F_PHIX), FLD(F_PHIY), FLD(F_PHIZ));
  orc_solphi_mthinc1(ni, nj, nk, dxinv, dyinv, dzinv, dt, &c->bet_mthinc, phi1, FLD(F_PHIX),
                     FLD(F_PHIY), FLD(F_PHIZ), FLD(F_U), FLD(F_V), FLD(F_W), FLD(F_FLX),
                     FLD(F_FLY), FLD(F_FLZ));
  orc_solphi_mthinc2(ni, nj, nk, dxinv, dyinv, dzinv, dt, FLD(F_U), FLD(F_V), FLD(F_W),
                     FLD(F_FLX), FLD(F_FLY), FLD(F_FLZ), phi1, phin1);
  orc_cal_grad_p2a(&S->ID, &S->svall[1], ni, nj, nk, dxinv, S->dyinv_array, dzinv, phin1,
                   FLD(F_PHIX), FLD(F_PHIY), FLD(F_PHIZ));
  orc_solphi_mthinc3(&S->ipara, ni, nj, nk, dxinv, dyinv, dzinv, &c->bet_mthinc, FLD(F_PHIX),
                     FLD(F_PHIY), FLD(F_PHIZ), phi1, phin1);
  orc_gnbc(nID, ni, nj, nk, FLD(F_U), FLD(F_W), &c->uwall, FLD(F_TH0), &c->surface_tension,
           FLD(F_ZETA), FLD(F_THETA));
  orc_bnd_contact_angle(nID, ni, nj, nk, phin1, FLD(F_THETA), &c->dx, &c->dy, &c->dz);
  orc_bnd_periodic(ni, nj, nk, phin1);
  /* main.f90:509-517 */
  orc_summation(ni, nj, nk, S->phin, &S->nbub);
  orc_bnd_neumann(nID, ni, nj, nk, phin0);
  orc_gnbc(nID, ni, nj, nk, FLD(F_U), FLD(F_W), &c->uwall, FLD(F_TH0), &c->surface_tension,
           FLD(F_ZETA), FLD(F_THETA));
  orc_bnd_contact_angle(nID, ni, nj, nk, phi0, FLD(F_THETA), &c->dx, &c->dy, &c->dz);
  orc_bnd_periodic(ni, nj, nk, phin0);
  if (c->imon_t > 0 && S->nstep % c->imon_t == 0) {
    double av;
    orc_cal_av(&S->ipara, &S->ndiv, ni, nj, nk, phin0, &av);
    printf("AV_PHI   %10d%20.10E\n", S->nstep, av);
  }
  /* properties (main.f90:526-529) */
  orc_cal_arith_mean(ni, nj, nk, phi0, &c->rhol, &c->rhog, FLD(F_RHO));
  orc_cal_arith_mean(ni, nj, nk, phin0, &c->rhol, &c->rhog, FLD(F_RHON));
  orc_cal_arith_mean(ni, nj, nk, phi0, &c->rmul, &c->rmug, FLD(F_RMU));
  orc_cal_arith_mean(ni, nj, nk, phin0, &c->rmul, &c->rmug, FLD(F_RMUN));
  /* explicit terms (main.f90:532-541) */
  orc_cal_advu_weno5(ni, nj, nk, dxinv, dyinv, dzinv, FLD(F_U), FLD(F_V), FLD(F_W), FLD(F_ADVU),
                     FLD(F_ADVV), FLD(F_ADVW));
  orc_cal_advu_weno5(ni, nj, nk, dxinv, dyinv, dzinv, FLD(F_UO), FLD(F_VO), FLD(F_WO),
                     FLD(F_ADVUO), FLD(F_ADVVO), FLD(F_ADVWO));
  orc_calphat(ni, nj, nk, FLD(F_P), FLD(F_PO), FLD(F_PHAT));
  orc_cal_grad_prs(ni, nj, nk, dxinv, dyinv, dzinv, FLD(F_PHAT), FLD(F_PRSU), FLD(F_PRSV),
                   FLD(F_PRSW));
  orc_calsij(ni, nj, nk, dxinv, dyinv, dzinv, FLD(F_U), FLD(F_V), FLD(F_W), S->s);
  orc_cal_arith_tau(ni, nj, nk, S->s, FLD(F_RMU), S->tau);
  orc_cal_div_tensor(ni, nj, nk, dxinv, dyinv, dzinv, S->tau, FLD(F_VISU), FLD(F_VISV),
                     FLD(F_VISW));
  /* CSF (main.f90:545-573) */
  orc_init_q(ni, nj, nk, FLD(F_SFU), FLD(F_SFV), FLD(F_SFW));
  orc_init_q(ni, nj, nk, FLD(F_SFUN), FLD(F_SFVN), FLD(F_SFWN));
  static const int ipj[4][3] = {{1, 1, 1}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}};
  orc_cal_grad_p2a(&S->ID, &S->svall[1], ni, nj, nk, dxinv, S->dyinv_array, dzinv, phi1,
                   FLD(F_PHIX), FLD(F_PHIY), FLD(F_PHIZ));
  for (int m = 0; m < 4; m++)
    orc_calrn(ni, nj, nk, dxinv, dyinv, dzinv, &ipj[m][0], &ipj[m][1], &ipj[m][2], FLD(F_PHIX),
              FLD(F_PHIY), FLD(F_PHIZ), S->rn[m]);
  orc_cal_grad_p2a(&S->ID, &S->svall[1], ni, nj, nk, dxinv, S->dyinv_array, dzinv, phin1,
                   FLD(F_PHIX), FLD(F_PHIY), FLD(F_PHIZ));
  for (int m = 0; m < 4; m++)
    orc_calrn(ni, nj, nk, dxinv, dyinv, dzinv, &ipj[m][0], &ipj[m][1], &ipj[m][2], FLD(F_PHIX),
              FLD(F_PHIY), FLD(F_PHIZ), S->rnn[m]);
  orc_calrkap(ni, nj, nk, dxinv, dyinv, dzinv, S->rn[0], S->rn[1], S->rn[2], S->rn[3], S->rkap);
  orc_calrkap(ni, nj, nk, dxinv, dyinv, dzinv, S->rnn[0], S->rnn[1], S->rnn[2], S->rnn[3],
              S->rkapn);
  orc_cal_grad_p2uvw(&S->ID, &S->svall[1], ni, nj, nk, dxinv, S->dyinv_array, dzinv, phi1,
                     FLD(F_PHIX), FLD(F_PHIY), FLD(F_PHIZ));
  orc_calfst(ni, nj, nk, &c->rhol, &c->rhog, &c->surface_tension, FLD(F_PHIX), FLD(F_PHIY),
             FLD(F_PHIZ), S->rkap, FLD(F_FSTU), FLD(F_FSTV), FLD(F_FSTW));
  orc_cal_grad_p2uvw(&S->ID, &S->svall[1], ni, nj, nk, dxinv, S->dyinv_array, dzinv, phin1,
                     FLD(F_PHIX), FLD(F_PHIY), FLD(F_PHIZ));
  orc_calfst(ni, nj, nk, &c->rhol, &c->rhog, &c->surface_tension, FLD(F_PHIX), FLD(F_PHIY),
             FLD(F_PHIZ), S->rkapn, FLD(F_FSTUN), FLD(F_FSTVN), FLD(F_FSTWN));
  orc_sum_fst(ni, nj, nk, FLD(F_FSTU), FLD(F_FSTV), FLD(F_FSTW), FLD(F_SFU), FLD(F_SFV),
              FLD(F_SFW));
  orc_sum_fst(ni, nj, nk, FLD(F_FSTUN), FLD(F_FSTVN), FLD(F_FSTWN), FLD(F_SFUN), FLD(F_SFVN),
              FLD(F_SFWN));
  double rho_av, rhon_av;
  orc_cal_av(&S->ipara, &S->ndiv, ni, nj, nk, FLD(F_RHO), &rho_av);
  orc_cal_av(&S->ipara, &S->ndiv, ni, nj, nk, FLD(F_RHON), &rhon_av);
  rho_av = (rho_av + rhon_av) * 0.5;
  /* source (main.f90:579-588) */
  orc_cal_srcu(nID, ni, nj, nk, &rho_av, dt, &c->grvb, &c->grvp, FLD(F_U), FLD(F_V), FLD(F_W),
               FLD(F_RHO), FLD(F_RHON), FLD(F_ADVU), FLD(F_ADVV), FLD(F_ADVW), FLD(F_ADVUO),
               FLD(F_ADVVO), FLD(F_ADVWO), FLD(F_PRSU), FLD(F_PRSV), FLD(F_PRSW), FLD(F_VISU),
               FLD(F_VISV), FLD(F_VISW), FLD(F_SFU), FLD(F_SFV), FLD(F_SFW), FLD(F_SFUN),
               FLD(F_SFVN), FLD(F_SFWN), FLD(F_SRCU), FLD(F_SRCV), FLD(F_SRCW));
  /* viscous step (main.f90:590-631) */
  {
    double *a[45];
    for (int m = 0; m < 45; m++) a[m] = S->coef[m];
    orc_cal_arith_coef_vis(ni, nj, nk, dxinv, dyinv, dzinv, FLD(F_RMUN), a[0], a[1], a[2], a[3],
                           a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11], a[12], a[13], a[14],
                           a[15], a[16], a[17], a[18], a[19], a[20], a[21], a[22], a[23], a[24],
                           a[25], a[26], a[27], a[28], a[29], a[30], a[31], a[32], a[33], a[34],
                           a[35], a[36], a[37], a[38], a[39], a[40], a[41], a[42], a[43], a[44]);
    orc_cpy(ni, nj, nk, FLD(F_U), FLD(F_UN));
    orc_cpy(ni, nj, nk, FLD(F_V), FLD(F_VN));
    orc_cpy(ni, nj, nk, FLD(F_W), FLD(F_WN));
    orc_solu_sor4(&S->ipara, &S->ID, nID, &S->ndiv, ni, nj, nk, key, sendjb, recvjb, &S->nstep,
                  &c->imon_t, dt, FLD(F_RHO), FLD(F_RHON), a[0], a[1], a[2], a[3], a[4], a[5],
                  a[6], a[7], a[8], a[9], a[10], a[11], a[12], a[13], a[14], a[15], a[16], a[17],
                  a[18], a[19], a[20], a[21], a[22], a[23], a[24], a[25], a[26], a[27], a[28],
                  a[29], a[30], a[31], a[32], a[33], a[34], a[35], a[36], a[37], a[38], a[39],
                  a[40], a[41], a[42], a[43], a[44], FLD(F_SRCU), FLD(F_SRCV), FLD(F_SRCW),
                  FLD(F_UN), FLD(F_VN), FLD(F_WN), &c->uwall, &c->dy, FLD(F_L1), FLD(F_L2), phi1);
    S->last_sor_lp = orc_sor_last_lp;
    S->last_sor_ratio = orc_sor_last_ratio;
  }
  orc_bndu(nID, ni, nj, nk, FLD(F_UN), FLD(F_VN), FLD(F_WN), &c->uwall, &c->dy, FLD(F_L1),
           FLD(F_L2), phi1);
  orc_bnd_periodic(ni, nj, nk, FLD(F_UN));
  orc_bnd_periodic(ni, nj, nk, FLD(F_VN));
  orc_bnd_periodic(ni, nj, nk, FLD(F_WN));
  /* projection (main.f90:633-661) */
  orc_caldiv(&S->ipara, &S->ndiv, ni, nj, nk, dxinv, dyinv, dzinv, dt, FLD(F_UN), FLD(F_VN),
             FLD(F_WN), FLD(F_DIV), &S->err_div0);
  orc_solp_fft_tdma1(ni, nj, nk, S->fftdata, FLD(F_DIV), S->phir_r, S->phir_i);
  /* trans_r2w / trans_w2r on one rank leave the memory image unchanged */
  orc_solp_fft_tdma2(&S->ID, &S->ndiv, S->svall, &c->rhog, dxinv, dzinv, S->as_fft, S->an_fft,
                     S->ap_fft, S->atdma, S->btr, S->bti, S->phir_r, S->phir_i);
  orc_solp_fft_tdma3(ni, nj, nk, S->fftdata, S->phir_r, S->phir_i, FLD(F_DP));
  orc_bnd_avzero(&S->ipara, &S->ndiv, ni, nj, nk, FLD(F_DP));
  orc_bnd_neumann(nID, ni, nj, nk, FLD(F_DP));
  orc_bnd_periodic(ni, nj, nk, FLD(F_DP));
  orc_solp_fft_tdma4(&S->ipara, &S->ID, &S->ndiv, ni, nj, nk, &S->nstep, &c->imon_t, &c->rhog,
                     dxinv, dyinv, dzinv, FLD(F_DIV), FLD(F_DP));
  orc_corunp_explicit(nID, ni, nj, nk, &c->rhog, dxinv, dyinv, dzinv, dt, FLD(F_DP),
                      FLD(F_PHAT), FLD(F_UN), FLD(F_VN), FLD(F_WN), FLD(F_PN));
  orc_bndu(nID, ni, nj, nk, FLD(F_UN), FLD(F_VN), FLD(F_WN), &c->uwall, &c->dy, FLD(F_L1),
           FLD(F_L2), phi1);
  orc_bnd_periodic(ni, nj, nk, FLD(F_UN));
  orc_bnd_periodic(ni, nj, nk, FLD(F_VN));
  orc_bnd_periodic(ni, nj, nk, FLD(F_WN));
  orc_bnd_avzero(&S->ipara, &S->ndiv, ni, nj, nk, FLD(F_PN));
  orc_bnd_neumann(nID, ni, nj, nk, FLD(F_PN));
  orc_bnd_periodic(ni, nj, nk, FLD(F_PN));
  /* diagnostics (main.f90:663-684) */
  orc_cal_center(&S->ipara, &S->ID, &S->svall[1], ni, nj, nk, &c->xl, &c->yl, &c->zl, dt,
                 &c->rhog, phi1, S->center, &S->center_pre1, &S->center_pre2, &S->area_pre1,
                 &S->area_pre2, &S->lap, &S->velocity);
  orc_cal_vel(&S->ipara, ni, nj, nk, &c->particle_radius, &c->rhol, &c->rmul, FLD(F_U), FLD(F_V),
              FLD(F_W), phi1, S->vel, S->re);
  if (c->imon_t > 0 && S->nstep % c->imon_t == 0) {
    double err_div;
    orc_caldiv(&S->ipara, &S->ndiv, ni, nj, nk, dxinv, dyinv, dzinv, dt, FLD(F_UN), FLD(F_VN),
               FLD(F_WN), FLD(F_DIV), &err_div);
    printf("Err_div  %10d%20.10E%20.10E%20.10E\n", S->nstep, err_div / S->err_div0, err_div,
           S->err_div0);
  }
  /* rotation (main.f90:686-697) */
  orc_cpy(ni, nj, nk, FLD(F_U), FLD(F_UO));
  orc_cpy(ni, nj, nk, FLD(F_V), FLD(F_VO));
  orc_cpy(ni, nj, nk, FLD(F_W), FLD(F_WO));
  orc_cpy(ni, nj, nk, FLD(F_P), FLD(F_PO));
  orc_cpy(ni, nj, nk, FLD(F_UN), FLD(F_U));
  orc_cpy(ni, nj, nk, FLD(F_VN), FLD(F_V));
  orc_cpy(ni, nj, nk, FLD(F_WN), FLD(F_W));
  orc_cpy(ni, nj, nk, FLD(F_PN), FLD(F_P));
  orc_cpy(ni, nj, nk, phin0, phi0);
  orc_cpy(ni, nj, nk, phin1, phi1);
}

/* Restart from a checkpoint in the reference's format, srcs/main.f90:375-404:
 * the record of srcs/datain.f:24-38 (one unformatted sequential record with
 * 4-byte length markers: u, v, w, p, uo, vo, wo, po, phi(...,0:nbub), time,
 * nstep) replaces those arrays, time and nstep of a started state; then the
 * restart fills: bndu of (u,v,w) and (uo,vo,wo), bnd_periodic of the eight
 * fields, gnbc and the contact-angle ghosts of phi(...,0) with its periodic
 * fill.  Returns 0, or -1 when the file does not hold this grid's record. */
int orc_state_datain(orc_state *S, const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  const size_t n3 = (size_t)S->n3;
  const size_t body = (8 + 2) * n3 * sizeof(double) + sizeof(double) + sizeof(int);
  int head = 0, tail = 0, step = 0;
  double t = 0.0;
  int ok = fread(&head, sizeof head, 1, f) == 1 && (size_t)head == body;
  const int order[8] = {F_U, F_V, F_W, F_P, F_UO, F_VO, F_WO, F_PO};
  for (int m = 0; m < 8 && ok; m++) ok = fread(S->f3[order[m]], sizeof(double), n3, f) == n3;
  if (ok) ok = fread(S->phi, sizeof(double), 2 * n3, f) == 2 * n3;
  if (ok) ok = fread(&t, sizeof t, 1, f) == 1 && fread(&step, sizeof step, 1, f) == 1;
  if (ok) ok = fread(&tail, sizeof tail, 1, f) == 1 && tail == head;
  fclose(f);
  if (!ok) return -1;
  S->time = t;
  S->nstep = step;
  const int *ni = &S->c.ni, *nj = &S->c.nj, *nk = &S->c.nk;
  double *phi0 = S->phi, *phi1 = S->phi + S->n3;
  orc_bndu(S->nID, ni, nj, nk, FLD(F_U), FLD(F_V), FLD(F_W), &S->c.uwall, &S->c.dy, FLD(F_L1),
           FLD(F_L2), phi1);
  orc_bndu(S->nID, ni, nj, nk, FLD(F_UO), FLD(F_VO), FLD(F_WO), &S->c.uwall, &S->c.dy, FLD(F_L1),
           FLD(F_L2), phi1);
  for (int m = 0; m < 8; m++) orc_bnd_periodic(ni, nj, nk, FLD(order[m]));
  orc_gnbc(S->nID, ni, nj, nk, FLD(F_U), FLD(F_W), &S->c.uwall, FLD(F_TH0), &S->c.surface_tension,
           FLD(F_ZETA), FLD(F_THETA));
  orc_bnd_contact_angle(S->nID, ni, nj, nk, phi0, FLD(F_THETA), &S->c.dx, &S->c.dy, &S->c.dz);
  orc_bnd_periodic(ni, nj, nk, phi0);
  return 0;
}

void orc_state_step(orc_state *S, int nsteps) {
  for (int n = 0; n < nsteps; n++) orc_state_step1(S);
  fflush(stdout);
}
