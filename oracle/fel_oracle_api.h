/* Prototypes of the oracle routines in fel_oracle.c (test infrastructure). */
#ifndef FEL_ORACLE_API_H
#define FEL_ORACLE_API_H
void orc_cal_grad_p2a(const int *ID, const int *svall, const int *ni_, const int *nj_, const int
    *nk_, const double *dxinv_, const double *dyinv_array, const double *dzinv_, const double *q,
    double *qx, double *qy, double *qz);
void orc_solphi_mthinc1(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_, const
    double *dyinv_, const double *dzinv_, const double *dt_, const double *bet_mthinc, const double
    *phi, const double *phix, const double *phiy, const double *phiz, const double *u, const double
    *v, const double *w, double *flphix, double *flphiy, double *flphiz);
void orc_solphi_mthinc2(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_, const
    double *dyinv_, const double *dzinv_, const double *dt_, const double *u, const double *v, const
    double *w, const double *flphix, const double *flphiy, const double *flphiz, const double *phi,
    double *phin);
void orc_solphi_mthinc3(const int *ipara, const int *ni_, const int *nj_, const int *nk_, const
    double *dxinv_, const double *dyinv_, const double *dzinv_, const double *bet_mthinc, const
    double *phix, const double *phiy, const double *phiz, const double *phi, double *phin);
void orc_summation(const int *ni_, const int *nj_, const int *nk_, double *qk, const int *nbub_);
void orc_cal_advu_weno5(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_, const
    double *dyinv_, const double *dzinv_, const double *u, const double *v, const double *w, double
    *adv_u, double *adv_v, double *adv_w);
void orc_cal_arith_mean(const int *ni_, const int *nj_, const int *nk_, const double *phik, const
    double *ql_, const double *qg_, double *qk);
void orc_calphat(const int *ni_, const int *nj_, const int *nk_, const double *p, const double *po,
    double *phat);
void orc_cal_grad_prs(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_, const
    double *dyinv_, const double *dzinv_, const double *q, double *qx_u, double *qy_v, double
    *qz_w);
void orc_calsij(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_, const double
    *dyinv_, const double *dzinv_, const double *uk, const double *vk, const double *wk, double
    *sk);
void orc_cal_arith_tau(const int *ni_, const int *nj_, const int *nk_, const double *s, const double
    *rmu, double *tau);
void orc_cal_div_tensor(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_, const
    double *dyinv_, const double *dzinv_, const double *tauk, double *uk, double *vk, double *wk);
void orc_init_q(const int *ni_, const int *nj_, const int *nk_, double *q1, double *q2, double *q3);
void orc_cal_grad_p2uvw(const int *ID, const int *svall, const int *ni_, const int *nj_, const int
    *nk_, const double *dxinv_, const double *dyinv_array, const double *dzinv_, const double *q,
    double *qx_u, double *qy_v, double *qz_w);
void orc_calrn(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_, const double
    *dyinv_, const double *dzinv_, const int *ip_, const int *jp_, const int *kp_, const double
    *phix, const double *phiy, const double *phiz, double *rn);
void orc_calrkap(const int *ni_, const int *nj_, const int *nk_, const double *dxinv_, const double
    *dyinv_, const double *dzinv_, const double *rn_pp, const double *rn_12, const double *rn_13,
    const double *rn_23, double *rkap);
void orc_calfst(const int *ni_, const int *nj_, const int *nk_, const double *rhol, const double
    *rhog, const double *surface_tension, const double *phix_u, const double *phiy_v, const double
    *phiz_w, const double *rkap, double *fst_u, double *fst_v, double *fst_w);
void orc_sum_fst(const int *ni_, const int *nj_, const int *nk_, const double *fu, const double *fv,
    const double *fw, double *su, double *sv, double *sw);
void orc_cal_av(const int *ipara, const int *ndiv, const int *ni_, const int *nj_, const int *nk_,
    const double *q, double *av);
void orc_bnd_avzero(const int *ipara, const int *ndiv, const int *ni_, const int *nj_, const int
    *nk_, double *q);
void orc_cal_srcu(const int *nID, const int *ni_, const int *nj_, const int *nk_, const double
    *rho_av, const double *dt_, const double *grvb_, const double *grvp_, const double *u, const
    double *v, const double *w, const double *rho, const double *rhon, const double *adv_u, const
    double *adv_v, const double *adv_w, const double *adv_uo, const double *adv_vo, const double
    *adv_wo, const double *prs_u, const double *prs_v, const double *prs_w, const double *vis_u,
    const double *vis_v, const double *vis_w, const double *fst_u, const double *fst_v, const double
    *fst_w, const double *fst_un, const double *fst_vn, const double *fst_wn, double *src_u, double
    *src_v, double *src_w);
void orc_cal_arith_coef_vis(const int *ni, const int *nj, const int *nk, const double *dxinv, const
    double *dyinv, const double *dzinv, const double *rmuk, double *a0, double *a1, double *a2,
    double *a3, double *a4, double *a5, double *a6, double *a7, double *a8, double *a9, double *a10,
    double *a11, double *a12, double *a13, double *a14, double *a15, double *a16, double *a17,
    double *a18, double *a19, double *a20, double *a21, double *a22, double *a23, double *a24,
    double *a25, double *a26, double *a27, double *a28, double *a29, double *a30, double *a31,
    double *a32, double *a33, double *a34, double *a35, double *a36, double *a37, double *a38,
    double *a39, double *a40, double *a41, double *a42, double *a43, double *a44);
void orc_cpy(const int *ni_, const int *nj_, const int *nk_, const double *q1, double *q2);
void orc_bnd_periodic(const int *ni_, const int *nj_, const int *nk_, double *qk);
void orc_bnd_neumann(const int *nID, const int *ni_, const int *nj_, const int *nk_, double *qk);
void orc_bnd_comm(const int *ipara, const int *nID, const int *ni, const int *nj, const int *nk,
    const int *key, double *sendjb, double *recvjb, double *qk);
void orc_bndu(const int *nID, const int *ni_, const int *nj_, const int *nk_, double *uk, double
    *vk, double *wk, const double *uwall_, const double *dy_, const double *l1, const double *l2,
    const double *phi);
void orc_solu_sor4(const int *ipara, const int *ID, const int *nID, const int *ndiv, const int *ni,
    const int *nj, const int *nk, const int *key, double *sendjb, double *recvjb, const int *nstep,
    const int *imon_t, const double *dt, const double *rho, const double *rhon, double *c0, double
    *c1, double *c2, double *c3, double *c4, double *c5, double *c6, double *c7, double *c8, double
    *c9, double *c10, double *c11, double *c12, double *c13, double *c14, double *c15, double *c16,
    double *c17, double *c18, double *c19, double *c20, double *c21, double *c22, double *c23,
    double *c24, double *c25, double *c26, double *c27, double *c28, double *c29, double *c30,
    double *c31, double *c32, double *c33, double *c34, double *c35, double *c36, double *c37,
    double *c38, double *c39, double *c40, double *c41, double *c42, double *c43, double *c44, const
    double *src_u, const double *src_v, const double *src_w, double *un, double *vn, double *wn,
    const double *uwall, const double *dy, const double *l1, const double *l2, const double *phi);
void orc_caldiv(const int *ipara, const int *ndiv, const int *ni_, const int *nj_, const int *nk_,
    const double *dxinv_, const double *dyinv_, const double *dzinv_, const double *dt_, const
    double *uk, const double *vk, const double *wk, double *div, double *err_div);
void orc_solp_fft_tdma1(const int *ni_, const int *nj_, const int *nk_, double *fftdata, const
    double *div, double *phir_r, double *phir_i);
void orc_trans_r2w(const int *ipara, const int *ID, const int *ndiv, const int *svall, const int
    *key, double *sendbuf, double *recvbuf, const double *r, double *w);
void orc_trans_w2r(const int *ipara, const int *ID, const int *ndiv, const int *svall, const int
    *key, double *sendbuf, double *recvbuf, const double *w, double *r);
void orc_solp_fft_tdma2(const int *ID, const int *ndiv, const int *svall, const double *rho_, const
    double *dxinv_, const double *dzinv_, const double *as_p_fft, const double *an_p_fft, const
    double *ap_p_fft, double *atdma, double *btdma_r, double *btdma_i, double *phiw_r, double
    *phiw_i);
void orc_solp_fft_tdma3(const int *ni_, const int *nj_, const int *nk_, double *fftdata, const
    double *phir_r, const double *phir_i, double *dp);
double orc_poisson_residual(int ni, int nj, int nk, double rho, double dxinv, double dyinv, double
    dzinv, const double *div, const double *dp, double *err_abs);
void orc_solp_fft_tdma4(const int *ipara, const int *ID, const int *ndiv, const int *ni, const int
    *nj, const int *nk, const int *nstep, const int *imon_t, const double *rho, const double *dxinv,
    const double *dyinv, const double *dzinv, const double *div, const double *dp);
void orc_corunp_explicit(const int *nID, const int *ni_, const int *nj_, const int *nk_, const
    double *rho_, const double *dxinv_, const double *dyinv_, const double *dzinv_, const double
    *dt_, const double *dp, const double *phat, double *un, double *vn, double *wn, double *pn);
void orc_cal_coef_prs(const int *nj_, const double *rho_, const double *dxinv_, const double
    *dyinv_, const double *dzinv_, double *aw, double *ae, double *as, double *an, double *ab,
    double *at, double *ap);
void orc_cal_coef_solp_fft(const int *nj_, const double *as_p, const double *an_p, double *as_fft,
    double *an_fft, double *ap_fft);
void orc_gnbc(const int *nID, const int *ni_, const int *nj_, const int *nk_, const double *uk,
    const double *wk, const double *uwall_, const double *theta_0_array, const double
    *surface_tension, const double *zeta_array, double *theta_array);
void orc_bnd_contact_angle(const int *nID, const int *ni_, const int *nj_, const int *nk_, double
    *qk, const double *theta_array, const double *dx_, const double *dy_, const double *dz_);
void orc_cal_center(const int *ipara, const int *ID, const int *svall, const int *ni_, const int
    *nj_, const int *nk_, const double *xl_, const double *yl_, const double *zl_, const double
    *dt_, const double *rhog, const double *phi, double *center, double *center_pre1, double
    *center_pre2, int *area_pre1, int *area_pre2, int *lap, double *velocity);
void orc_cal_vel(const int *ipara, const int *ni_, const int *nj_, const int *nk_, const double
    *radius, const double *rhol, const double *rmul, const double *u, const double *v, const double
    *w, const double *phi, double *vel, double *re);
void orc_caldt(const int *ipara, const int *nID, const int *ID, const int *ndiv, const int *ni_,
    const int *nj_, const int *nk_, const int *nstep, const int *imon_t, const double *dxinv,
    const double *dyinv, const double *dzinv, const double *cfl, const double *rhol, const double
    *rhog, const double *rmul, const double *rmug, const double *surface_tension, const double *u,
    const double *v, const double *w, double *dt, double *time, double *lim);
#endif
