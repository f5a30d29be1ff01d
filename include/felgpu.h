/* felgpu.h -- C ABI of the gfx950 (MI355X) timestep library libfelgpu.so.
 *
 * Every solver entry point keeps the argument list of the Fortran subroutine
 * it replaces and its external name as the Fortran compilers used by the
 * reference emit it (lower case plus one trailing underscore), so the
 * reference's implicit-interface `call name(...)` statements link against
 * this library without source changes.  All arguments are passed by
 * reference, as Fortran passes them.
 *
 * Arrays.  A "padded field" is a REAL(8) array (-2:ni+3,-2:nj+3,-2:nk+3),
 * i fastest, as allocated in srcs/allocate.h:7-165; "padded x3" is the same
 * with a trailing dimension 1:3, "padded x6" with 1:6.  j is the wall-normal
 * direction; i and k are periodic.
 *
 * Residency.  A host array registered with felgpu_register_ lives on the
 * device: entry points resolve any pointer inside a registered range
 * (interior and slice pointers included) to its device copy and move no
 * data.  An unregistered array is copied to the device for the call and its
 * written part copied back, so every entry point also works on plain host
 * memory (slowly).  Host code that reads or writes a registered array
 * between calls uses felgpu_download_ / felgpu_upload_.
 *
 * Errors.  A HIP failure, an argument the one-GPU library does not support
 * (ipara != 0, a second MPI rank) or an inconsistent size prints the entry
 * point's name and the reason to stderr and aborts the process; there is no
 * host fallback.
 *
 * Streams.  All work is queued on one library stream (felgpu_stream_) in
 * call order; entry points that return host scalars (cal_av_, caldiv_'s
 * err_div, cal_center_, cal_vel_, felgpu_poisson_residual_) synchronise it.
 */
#ifndef FELGPU_H
#define FELGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Library control (no counterpart in the reference: it runs on host memory).
 * ---------------------------------------------------------------------- */

/* Selects the HIP device and creates the library stream.  Optional: every
 * entry point initialises on first use (device 0). */
void felgpu_init_(const int *device);

/* Makes [host, host + *nbytes) device-resident.  *upload != 0 copies the
 * host contents; *upload == 0 zero-fills the device copy.  Call once per
 * allocatable after `allocate`, before the first solver call. */
void felgpu_register_(const void *host, const int64_t *nbytes, const int *upload);

/* As felgpu_register_, with an FP64 device copy in either precision mode.
 * Used for the arrays that stay FP64 in the FP32 mode: phi, phin, the
 * MTHINC fluxes, the contact-line maps (theta, theta_0, l1, l2, zeta), the
 * divergence and the pressure correction dp. */
void felgpu_register_fp64_(const void *host, const int64_t *nbytes, const int *upload);

/* Device precision of arrays registered with felgpu_register_: *bits = 64
 * (default; every result equals the reference's FP64 arithmetic) or 32 (the
 * FP32 mode: flow fields, stencils, Gauss-Seidel and FFT+TDMA in FP32, with
 * FP64 volume-fraction fluxes and mass sums, FP64 contact-line velocity and
 * FP64 iterative refinement of the pressure correction).  Only while
 * nothing is registered; FELGPU_PRECISION=fp32 sets the same at start-up.
 * felgpu_precision_ returns the current setting. */
void felgpu_set_precision_(const int *bits);
int felgpu_precision_(void);

/* Tangent mode (no counterpart in the reference, which has no derivatives):
 * *on = 1 makes every registered device copy a pair (FP64 value, derivative
 * along one direction of the inputs) and runs every entry point on it, so a
 * run of the unchanged timestep computes the values (bitwise equal to the
 * FP64 mode) and their directional derivative (forward-mode
 * differentiation of the discrete step).  Only while nothing is registered,
 * in FP64.  felgpu_tangent_ returns 1 when on.
 *   felgpu_tangent_seed_  sets the derivative part of the whole registered
 *                         array containing `host` from seed[] (one double
 *                         per element; e.g. 1 on the wall cells of a
 *                         contact-angle stripe in theta_0, or 1 everywhere
 *                         in l1 and l2 for the slip length);
 *   felgpu_tangent_fetch_ copies the derivative part into out[].
 * Uploads and registrations set the derivative to zero; downloads return
 * the values. */
void felgpu_set_tangent_(const int *on);
int felgpu_tangent_(void);
void felgpu_tangent_seed_(const void *host, const double *seed);
void felgpu_tangent_fetch_(const void *host, double *out);

/* Reuse of results already computed (no counterpart in the reference):
 * when an entry point is asked for a result whose inputs have exactly the
 * contents of an earlier call's (tracked by content numbers; e.g. the
 * surface-tension chain of phi at step n+1 repeats the chain of phin at
 * step n), the result is taken from the arrays that still hold it, by a
 * buffer exchange, instead of being recomputed.  *on = 0 turns this off
 * (as FELGPU_NO_RESULT_REUSE=1 does at start-up); results are bitwise the
 * same either way.  felgpu_result_reuse_hits_ counts the calls answered so. */
void felgpu_set_result_reuse_(const int *on);
/* The current setting: 1 on, 0 off. */
int felgpu_result_reuse_(void);
int64_t felgpu_result_reuse_hits_(void);
/* Calls answered from a result a previous call's pass formed alongside its
 * own (the other staggered normals of calrn_, the curvature of calrkap_);
 * no counterpart in the reference, for tests and diagnostics. */
int64_t felgpu_stash_hits_(void);

/* Host -> device copy of the whole registered range containing `host`. */
void felgpu_upload_(const void *host);

/* Device -> host copy of the whole registered range containing `host`. */
void felgpu_download_(void *host);

/* Releases the device copy of one registered range / of all of them. */
void felgpu_unregister_(const void *host);
void felgpu_unregister_all_(void);

/* Device address backing a registered host address (an unregistered one is
 * an error).  The caller may write through it, so the range's contents
 * count as changed. */
void *felgpu_device_ptr_(const void *host);

/* Waits for all queued work; returns the library's hipStream_t. */
void felgpu_sync_(void);
void *felgpu_stream_(void);

/* Event timer of the colour-pair Gauss-Seidel update (id 3): every launch's
 * own start and end while enabled; read gives the total ms and launches. */
void felgpu_prof_enable_(const int *on);
void felgpu_prof_read_(const int *id, double *ms_total, int64_t *count);
void felgpu_prof_reset_(void);

/* ------------------------------------------------------------------------
 * VOF transport (MTHINC).  Call sites: srcs/main.f90:491-509.
 * ---------------------------------------------------------------------- */

/* srcs/cal_grad_p2a.f:1 -- vertex-centred gradient of q on [-2,n+2]^3.
 * dyinv_array(-2:svall+3) is a host array; svall = global nj. */
void cal_grad_p2a_(const int *ID, const int *svall, const int *ni, const int *nj, const int *nk,
                   const double *dxinv, const double *dyinv_array, const double *dzinv,
                   const double *q, double *qx, double *qy, double *qz);

/* srcs/solphi_mthinc1.f:1 -- MTHINC face fluxes flphix/y/z (padded). */
void solphi_mthinc1_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                     const double *dyinv, const double *dzinv, const double *dt,
                     const double *bet_mthinc, const double *phi, const double *phix,
                     const double *phiy, const double *phiz, const double *u, const double *v,
                     const double *w, double *flphix, double *flphiy, double *flphiz);

/* srcs/solphi_mthinc2.f:1 -- conservative update phin = phi - div(flux) + phi div(u) dt. */
void solphi_mthinc2_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                     const double *dyinv, const double *dzinv, const double *dt, const double *u,
                     const double *v, const double *w, const double *flphix,
                     const double *flphiy, const double *flphiz, const double *phi, double *phin);

/* srcs/solphi_mthinc3.f:1 -- truncation of flat cells, clamp to [0,1] and
 * mass-preserving redistribution.  *ipara must be 0. */
void solphi_mthinc3_(const int *ipara, const int *ni, const int *nj, const int *nk,
                     const double *dxinv, const double *dyinv, const double *dzinv,
                     const double *bet_mthinc, const double *phix, const double *phiy,
                     const double *phiz, const double *phi, double *phin);

/* srcs/summation.f:1 -- phi(:,:,:,0) = clamp(sum of slices 1..nbub);
 * qk is the (padded, 0:nbub) array. */
void summation_(const int *ni, const int *nj, const int *nk, double *qk, const int *nbub);

/* ------------------------------------------------------------------------
 * Momentum right-hand side.  Call sites: srcs/main.f90:520-579.
 * ---------------------------------------------------------------------- */

/* srcs/cal_advu_weno5.f:1 -- WENO5 advection terms at u, v, w points. */
void cal_advu_weno5_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                     const double *dyinv, const double *dzinv, const double *u, const double *v,
                     const double *w, double *adv_u, double *adv_v, double *adv_w);

/* srcs/cal_arith_mean.f:1 -- qk = phik*qg + (1-phik)*ql on the padded image. */
void cal_arith_mean_(const int *ni, const int *nj, const int *nk, const double *phik,
                     const double *ql, const double *qg, double *qk);

/* srcs/calphat.f:1 -- pressure predictor phat = 2p - po. */
void calphat_(const int *ni, const int *nj, const int *nk, const double *p, const double *po,
              double *phat);

/* srcs/cal_grad_prs.f:1 -- pressure gradient at u, v, w points. */
void cal_grad_prs_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                   const double *dyinv, const double *dzinv, const double *q, double *qx_u,
                   double *qy_v, double *qz_w);

/* srcs/calsij.f:5 -- strain-rate tensor sk (padded x6). */
void calsij_(const int *ni, const int *nj, const int *nk, const double *dxinv,
             const double *dyinv, const double *dzinv, const double *uk, const double *vk,
             const double *wk, double *sk);

/* srcs/cal_arith_tau.f:7 -- stress tau = 2 mu S with face-averaged mu (padded x6). */
void cal_arith_tau_(const int *ni, const int *nj, const int *nk, const double *s,
                    const double *rmu, double *tau);

/* srcs/cal_div_tensor.f:1 -- divergence of the stress at u, v, w points. */
void cal_div_tensor_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                     const double *dyinv, const double *dzinv, const double *tauk, double *uk,
                     double *vk, double *wk);

/* srcs/init_q.f:1 -- zero three padded fields. */
void init_q_(const int *ni, const int *nj, const int *nk, double *q1, double *q2, double *q3);

/* srcs/cal_grad_p2uvw.f:1 -- face gradients of q at u, v, w points. */
void cal_grad_p2uvw_(const int *ID, const int *svall, const int *ni, const int *nj,
                     const int *nk, const double *dxinv, const double *dyinv_array,
                     const double *dzinv, const double *q, double *qx_u, double *qy_v,
                     double *qz_w);

/* srcs/calrn.f:1 -- unit normals averaged onto the (ip,jp,kp)-shifted
 * staggered position (rn padded x3). */
void calrn_(const int *ni, const int *nj, const int *nk, const double *dxinv, const double *dyinv,
            const double *dzinv, const int *ip, const int *jp, const int *kp, const double *phix,
            const double *phiy, const double *phiz, double *rn);

/* srcs/calrkap.f:1 -- curvature at u, v, w points (rkap padded x3). */
void calrkap_(const int *ni, const int *nj, const int *nk, const double *dxinv,
              const double *dyinv, const double *dzinv, const double *rn_pp, const double *rn_12,
              const double *rn_13, const double *rn_23, double *rkap);

/* srcs/calfst.f:7 -- CSF force -2 sigma/(rho_l+rho_g) kappa grad(phi). */
void calfst_(const int *ni, const int *nj, const int *nk, const double *rhol, const double *rhog,
             const double *surface_tension, const double *phix_u, const double *phiy_v,
             const double *phiz_w, const double *rkap, double *fst_u, double *fst_v,
             double *fst_w);

/* srcs/sum_fst.f:1 -- sum_* += fst_* over the padded image. */
void sum_fst_(const int *ni, const int *nj, const int *nk, const double *fst_u,
              const double *fst_v, const double *fst_w, double *sum_u, double *sum_v,
              double *sum_w);

/* srcs/cal_av.f:1 -- interior mean of q into the host scalar *av.  *ipara must be 0. */
void cal_av_(const int *ipara, const int *ndiv, const int *ni, const int *nj, const int *nk,
             const double *q, double *av);

/* srcs/cal_srcu.f:1 -- explicit right-hand side src_u/v/w of the momentum step. */
void cal_srcu_(const int *nID, const int *ni, const int *nj, const int *nk, const double *rho_av,
               const double *dt, const double *grvb, const double *grvp, const double *u,
               const double *v, const double *w, const double *rho, const double *rhon,
               const double *adv_u, const double *adv_v, const double *adv_w,
               const double *adv_uo, const double *adv_vo, const double *adv_wo,
               const double *prs_u, const double *prs_v, const double *prs_w,
               const double *vis_u, const double *vis_v, const double *vis_w,
               const double *fst_u, const double *fst_v, const double *fst_w,
               const double *fst_un, const double *fst_vn, const double *fst_wn, double *src_u,
               double *src_v, double *src_w);

/* ------------------------------------------------------------------------
 * Viscous Crank-Nicolson solve.  Call sites: srcs/main.f90:590, 609.
 * ---------------------------------------------------------------------- */

/* srcs/cal_arith_coef_vis.f:6 -- the 45 stencil coefficient arrays, in the
 * reference's argument order (au_w_uk ... av_tn_wk).  When rmuk and all 45
 * outputs are registered the arrays are not written: the library records
 * rmuk (and keeps a device snapshot if rmuk is later overwritten) and
 * solu_sor4_ recomputes the coefficients on the fly; the arrays are
 * materialised the first time any other entry point or felgpu_download_
 * touches one of them, so they always read back exactly as the reference
 * computes them. */
void cal_arith_coef_vis_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                         const double *dyinv, const double *dzinv, const double *rmuk,
                         double *au_w_uk, double *au_e_uk, double *au_s_uk, double *au_n_uk,
                         double *au_b_uk, double *au_t_uk, double *au_p_uk, double *av_ws_uk,
                         double *av_es_uk, double *av_wn_uk, double *av_en_uk, double *aw_wb_uk,
                         double *aw_eb_uk, double *aw_wt_uk, double *aw_et_uk, double *av_w_vk,
                         double *av_e_vk, double *av_s_vk, double *av_n_vk, double *av_b_vk,
                         double *av_t_vk, double *av_p_vk, double *au_sw_vk, double *au_nw_vk,
                         double *au_se_vk, double *au_ne_vk, double *aw_sb_vk, double *aw_nb_vk,
                         double *aw_st_vk, double *aw_nt_vk, double *aw_w_wk, double *aw_e_wk,
                         double *aw_s_wk, double *aw_n_wk, double *aw_b_wk, double *aw_t_wk,
                         double *aw_p_wk, double *au_bw_wk, double *au_tw_wk, double *au_be_wk,
                         double *au_te_wk, double *av_bs_wk, double *av_ts_wk, double *av_bn_wk,
                         double *av_tn_wk);

/* srcs/solu_sor4.f:1 -- 8-colour Gauss-Seidel for (un,vn,wn), u -> v -> w
 * per sweep, ghost refresh (bndu + bnd_periodic) after each colour pair,
 * stop at err/err0 <= 1e-3 or after 21 sweeps.  *ipara must be 0;
 * sendjb/recvjb are unused on one GPU.  phi is the single phase slice the
 * wall condition reads. */
void solu_sor4_(const int *ipara, const int *ID, const int *nID, const int *ndiv, const int *ni,
                const int *nj, const int *nk, const int *key, double *sendjb, double *recvjb,
                const int *nstep, const int *imon_t, const double *dt, const double *rho,
                const double *rhon, double *au_w_u, double *au_e_u, double *au_s_u,
                double *au_n_u, double *au_b_u, double *au_t_u, double *au_p_u,
                double *av_ws_u, double *av_es_u, double *av_wn_u, double *av_en_u,
                double *aw_wb_u, double *aw_eb_u, double *aw_wt_u, double *aw_et_u,
                double *av_w_v, double *av_e_v, double *av_s_v, double *av_n_v, double *av_b_v,
                double *av_t_v, double *av_p_v, double *au_sw_v, double *au_nw_v,
                double *au_se_v, double *au_ne_v, double *aw_sb_v, double *aw_nb_v,
                double *aw_st_v, double *aw_nt_v, double *aw_w_w, double *aw_e_w, double *aw_s_w,
                double *aw_n_w, double *aw_b_w, double *aw_t_w, double *aw_p_w,
                double *au_bw_w, double *au_tw_w, double *au_be_w, double *au_te_w,
                double *av_bs_w, double *av_ts_w, double *av_bn_w, double *av_tn_w,
                const double *src_u, const double *src_v, const double *src_w, double *un,
                double *vn, double *wn, const double *uwall, const double *dy, const double *l1,
                const double *l2, const double *phi);

/* The sweep index lp (0..20, 21 when the cap is reached) and err/err0 of the
 * last solu_sor4_ call, the values srcs/solu_sor4.f:298-305 prints.
 * No counterpart: the reference only prints them. */
void felgpu_sor_last_(int *lp, double *ratio);

/* Allocation-free form of solu_sor4_.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * 8-colour Gauss-Seidel, ghost refreshes and stop rule as solu_sor4_
 * (srcs/solu_sor4.f:1-308), FP64 only, bit-identical to solu_sor4_ with the
 * coefficients of cal_arith_coef_vis_.  Every array argument is a DEVICE
 * pointer to a padded (-2:n+3) column-major image: rho, rhon, rmun (the
 * viscosity the 45 coefficients are formed from), src_u..w, l1, l2, phi in;
 * un, vn, wn in/out.  Scalars are host pointers as elsewhere.  The work
 * memory (residual partials, slip-wall coefficients; the face densities are
 * formed in the sweep kernels) lives in a static 16 MiB device array per
 * GPU, located once per device at the library's initialisation: the call
 * allocates nothing.  Limit: grids whose work memory exceeds 16 MiB
 * (8 (12 nb + 257) + 32 (ni+6)(nk+6) bytes, nb the blocks of one launch;
 * 5.9 MB at 512x128x256, 14.9 MB at 512x256x512) are refused.  Calls that
 * overlap on one GPU race on that array.  *lp, *ratio: as felgpu_sor_last_.
 * Returns 0; 1 when the grid needs more work memory than the array holds;
 * 2 for an invalid grid or a non-FP64 mode (both before any launch). */
int felgpu_solu_sor4_noalloc_(const int *nID, const int *ni, const int *nj, const int *nk,
                              const double *dt, const double *dxinv, const double *dyinv,
                              const double *dzinv, const double *rho, const double *rhon,
                              const double *rmun, const double *src_u, const double *src_v,
                              const double *src_w, double *un, double *vn, double *wn,
                              const double *uwall, const double *dy, const double *l1,
                              const double *l2, const double *phi, int *lp, double *ratio);

/* Number of solu_sor4_ calls and of colour sweeps (lp+1, or 21 at the cap)
 * they made since the last reset; *reset != 0 clears the counts after
 * reading them.  No counterpart: a benchmark statistic. */
void felgpu_sor_stats_(const int *reset, long long *calls, long long *sweeps);

/* ------------------------------------------------------------------------
 * Boundary conditions.  Call sites: srcs/main.f90:324-422, 497-517, 625-661.
 * ---------------------------------------------------------------------- */

/* srcs/bndu.f:6 -- Navier-slip / no-penetration ghosts of u, v, w at the
 * j walls (nID(Y_MINUS/Y_PLUS) < 0) with phase-weighted slip length. */
void bndu_(const int *nID, const int *ni, const int *nj, const int *nk, double *uk, double *vk,
           double *wk, const double *uwall, const double *dy, const double *l1, const double *l2,
           const double *phi);

/* srcs/bnd_periodic.f:4 -- periodic ghosts in i and k. */
void bnd_periodic_(const int *ni, const int *nj, const int *nk, double *qk);

/* srcs/bnd_neumann.f:4 -- mirrored (zero-gradient) ghosts at the j walls. */
void bnd_neumann_(const int *nID, const int *ni, const int *nj, const int *nk, double *qk);

/* srcs/bnd_comm.f:1 -- j-halo exchange between MPI ranks: nothing to do on
 * one GPU; *ipara != 0 is rejected. */
void bnd_comm_(const int *ipara, const int *nID, const int *ni, const int *nj, const int *nk,
               const int *key, double *sendjb, double *recvjb, double *qk);

/* srcs/gnbc.f90:1 -- dynamic contact angle theta_array(i,k,side) from the
 * contact-line slip velocity. */
void gnbc_(const int *nID, const int *ni, const int *nj, const int *nk, const double *uk,
           const double *wk, const double *uwall, const double *theta_0_array,
           const double *surface_tension, const double *zeta_array, double *theta_array);

/* srcs/bnd_contact_angle.f90:3 -- contact-angle ghosts of phi at the j walls. */
void bnd_contact_angle_(const int *nID, const int *ni, const int *nj, const int *nk, double *qk,
                        const double *theta_array, const double *dx, const double *dy,
                        const double *dz);

/* Cross-check entry points (no counterpart): bndu with its three ordered
 * passes as separate launches, and the single composed launch of bndu plus
 * bnd_periodic of component *swept (0 u, 1 v, 2 w; -1 none) that the
 * Gauss-Seidel ghost refresh uses.  Same results as bndu_ followed by
 * bnd_periodic_. */
void felgpu_bndu_phased_(const int *nID, const int *ni, const int *nj, const int *nk, double *uk,
                         double *vk, double *wk, const double *uwall, const double *dy,
                         const double *l1, const double *l2, const double *phi,
                         const int *swept);
void felgpu_bc_fused_(const int *nID, const int *ni, const int *nj, const int *nk, double *uk,
                      double *vk, double *wk, const double *uwall, const double *dy,
                      const double *l1, const double *l2, const double *phi, const int *swept);

/* ------------------------------------------------------------------------
 * Pressure projection (FFT in i,k + Thomas in j).  Call sites:
 * srcs/main.f90:467-468 (set-up), 633-661.
 * ---------------------------------------------------------------------- */

/* srcs/caldiv.f:1 -- div = div(u)/dt and *err_div = its max norm.  *ipara must be 0. */
void caldiv_(const int *ipara, const int *ndiv, const int *ni, const int *nj, const int *nk,
             const double *dxinv, const double *dyinv, const double *dzinv, const double *dt,
             const double *uk, const double *vk, const double *wk, double *div, double *err_div);

/* srcs/solp_fft_tdma1.f:4 -- forward 2-D DFT (i,k) of div per j plane into
 * phir_r/phir_i (centred mode layout).  fftdata (work array) is unused. */
void solp_fft_tdma1_(const int *ni, const int *nj, const int *nk, double *fftdata,
                     const double *div, double *phir_r, double *phir_i);

/* srcs/trans_r2w.f:1, srcs/trans_w2r.f:1 -- slab transposes between MPI
 * ranks; on one GPU the transposed array has the same memory image, so these
 * copy only when r and w are different arrays. */
void trans_r2w_(const int *ipara, const int *ID, const int *ndiv, const int *svall,
                const int *key, double *sendbuf, double *recvbuf, const double *r, double *w);
void trans_w2r_(const int *ipara, const int *ID, const int *ndiv, const int *svall,
                const int *key, double *sendbuf, double *recvbuf, const double *w, double *r);

/* srcs/solp_fft_tdma2.f:4 -- one Thomas solve along j per (kx,kz) mode.
 * as/an/ap_p_fft(-2:nj+3) are host arrays; atdma/btdma_r/btdma_i are unused. */
void solp_fft_tdma2_(const int *ID, const int *ndiv, const int *svall, const double *rho,
                     const double *dxinv, const double *dzinv, const double *as_p_fft,
                     const double *an_p_fft, const double *ap_p_fft, double *atdma,
                     double *btdma_r, double *btdma_i, double *phiw_r, double *phiw_i);

/* srcs/solp_fft_tdma3.f:4 -- inverse 2-D DFT per j plane into dp (interior).
 * fftdata is unused. */
void solp_fft_tdma3_(const int *ni, const int *nj, const int *nk, double *fftdata,
                     const double *phir_r, const double *phir_i, double *dp);

/* srcs/solp_fft_tdma4.f:4 -- residual monitor: computes and prints the
 * relative residual every imon_t steps (on ID 0). */
void solp_fft_tdma4_(const int *ipara, const int *ID, const int *ndiv, const int *ni,
                     const int *nj, const int *nk, const int *nstep, const int *imon_t,
                     const double *rho, const double *dxinv, const double *dyinv,
                     const double *dzinv, const double *div, const double *dp);

/* The monitor's relative residual ||L dp/rho - div|| / ||div|| as a return
 * value (no counterpart; srcs/solp_fft_tdma4.f:28-72 prints it). */
double felgpu_poisson_residual_(const int *ni, const int *nj, const int *nk, const double *rho,
                                const double *dxinv, const double *dyinv, const double *dzinv,
                                const double *div, const double *dp);

/* Limits of the FP32 mode's iterative refinement of the pressure
 * correction: at most *max_steps refinement steps (0..8; 0 also turns the
 * warm start off), stopping once |div - L dp| / |div| <= *tol.  A negative
 * *max_steps or a non-positive *tol keeps that limit.  Defaults: 3 and 1e-7
 * (FELGPU_REFINE, FELGPU_REFINE_TOL).  No counterpart: the reference solves
 * in FP64 without refinement. */
void felgpu_set_refinement_(const int *max_steps, const double *tol);

/* The FP32 mode's refined pressure solves since the last reset: solves,
 * warm-started solves, refinement steps, and the relative residual
 * |div - L dp| / |div| after the last solve's first FP32 solve and at its
 * end; *reset != 0 clears the counts after reading them.  No counterpart:
 * the reference solves in FP64 without refinement. */
void felgpu_poisson_stats_(const int *reset, long long *solves, long long *warm_solves,
                           long long *refine_steps, double *first_rel, double *final_rel);

/* srcs/bnd_avzero.f:1 -- subtract the interior mean from q (padded image).
 * *ipara must be 0. */
void bnd_avzero_(const int *ipara, const int *ndiv, const int *ni, const int *nj, const int *nk,
                 double *q);

/* srcs/corunp_explicit.f:1 -- projection un -= dt/rho grad(dp), pn = phat + dp. */
void corunp_explicit_(const int *nID, const int *ni, const int *nj, const int *nk,
                      const double *rho, const double *dxinv, const double *dyinv,
                      const double *dzinv, const double *dt, const double *dp,
                      const double *phat, double *un, double *vn, double *wn, double *pn);

/* srcs/cal_coef_prs.f:5 -- set-up: 1-D pressure coefficients (-2:nj+3), host arrays. */
void cal_coef_prs_(const int *nj, const double *rho, const double *dxinv, const double *dyinv,
                   const double *dzinv, double *aw, double *ae, double *as, double *an,
                   double *ab, double *at, double *ap);

/* srcs/cal_coef_solp_fft.f:5 -- set-up: Thomas coefficients with the wall
 * rows folded in (-2:nj+3), host arrays. */
void cal_coef_solp_fft_(const int *nj, const double *as_p, const double *an_p, double *as_fft,
                        double *an_fft, double *ap_fft);

/* srcs/cpy.f:1 -- q2 = q1 over the padded image. */
void cpy_(const int *ni, const int *nj, const int *nk, const double *q1, double *q2);

/* ------------------------------------------------------------------------
 * Monitors.  Call sites: srcs/main.f90:664, 666.
 * ---------------------------------------------------------------------- */

/* srcs/cal_center.f:1 -- centroid of the liquid with periodic lap tracking
 * and its velocity (host scalars in/out).  *ipara must be 0. */
void cal_center_(const int *ipara, const int *ID, const int *svall, const int *ni, const int *nj,
                 const int *nk, const double *xl, const double *yl, const double *zl,
                 const double *dt, const double *rhog, const double *phi, double *center,
                 double *center_pre1, double *center_pre2, int *area_pre1, int *area_pre2,
                 int *lap, double *velocity);

/* srcs/cal_vel.f:1 -- liquid-weighted mean velocity vel(3) and Reynolds
 * numbers re(3).  *ipara must be 0. */
void cal_vel_(const int *ipara, const int *ni, const int *nj, const int *nk, const double *radius,
              const double *rhol, const double *rmul, const double *u, const double *v,
              const double *w, const double *phi, double *vel, double *re);

/* srcs/caldt.f:1 -- CFL time step: dt = min(cfl/max(|u|dxinv, |v|dyinv
 * (j <= jen), |w|dzinv), cfl*0.5*dcell^2/nu*100, cfl*sqrt(dcell^3 (rho_l +
 * rho_g)/(2 sigma))) and time += dt; prints the reference's monitor line
 * every imon_t steps on ID 0.  The reference's time loop has the call
 * commented out (srcs/main.f90:479); the Fortran host calls it when
 * fel_cfg.cfl > 0.  ipara must be 0. */
void caldt_(const int *ipara, const int *nID, const int *ID, const int *ndiv, const int *ni,
            const int *nj, const int *nk, const int *nstep, const int *imon_t, const double *dxinv,
            const double *dyinv, const double *dzinv, const double *cfl, const double *rhol,
            const double *rhog, const double *rmul, const double *rmug,
            const double *surface_tension, const double *u, const double *v, const double *w,
            double *dt, double *time);

/* ------------------------------------------------------------------------
 * Output cadences (srcs/main.f90:706-743): the device computes the reduced
 * data, the host formats the files (host/fel_driver.f90).
 * ---------------------------------------------------------------------- */

/* srcs/mk_slice_j.f:1 -- 8-point average of qk at plane j with offsets
 * (ip,jp,kp) into the (0:ni,0:nk) host slices.  ipara must be 0. */
void mk_slice_j_(const int *ipara, const int *ID, const int *nID, const int *ndiv, const int *ni,
                 const int *nj, const int *nk, const int *j, const int *ip, const int *jp,
                 const int *kp, double *wrk_slice, const double *qk, double *qk_slice);

/* Vertex values written by srcs/mkvtk_phi.f (*kind = 0: ghost fills x
 * periodic, y and z Neumann) and srcs/mkvtk_p.f (*kind = 1: x and z
 * periodic, y Neumann): out[(ni+1)(nj+1)(nk+1)], i fastest, the 8-point
 * average at each vertex (0..ni, 0..nj, 0..nk).  No counterpart as a
 * separate routine in the reference (its writers compute them inline). */
void felgpu_vtk_points_(const int *ni, const int *nj, const int *nk, const double *q,
                        const int *kind, double *out);

/* Interface crossings phi(i) > 0.5 > phi(i+1) along row *jrow of each
 * k-plane, at x = (i-1/2)dx + (1/2-phi(i))dx/(phi(i+1)-phi(i)) - 3/4 xl, as
 * srcs/find_interface_positions_upper.f90 (every crossing) and _lower.f90
 * (*first_only != 0: the first) compute them: x[(k-1)*ni + m], m < cnt[k-1]. */
void felgpu_interface_crossings_(const int *ni, const int *nj, const int *nk, const double *phi,
                                 const int *jrow, const double *dx, const double *xl,
                                 const int *first_only, double *x, int *cnt);

/* ----------------------------------------------------------------------
 * Adjoints (reverse mode) of the routines of the timestep.  No counterpart
 * in the reference, which has no adjoint: <name>_b_ applies the transpose
 * of the derivative of <name>_ with respect to its active arrays.  Argument
 * list: the forward's, with every active array replaced by its adjoint
 * (same shape and layout); arrays that only parameterise the map
 * (densities, coefficients, slip lengths, phi) are passed as in the
 * forward, and nonlinear routines take each primal array the derivative
 * needs followed by its adjoint.  Adjoints of the forward's inputs are
 * incremented, adjoints of what the forward overwrites are consumed (set to
 * zero), in/out adjoints are updated in place.  <name>_bp_ (further below)
 * also forms the adjoints of the parameterising arrays.  FP64 mode only
 * (abort otherwise).  A reverse sweep calls them in the reverse of the
 * forward's order.
 * ---------------------------------------------------------------------- */

/* srcs/caldiv.f:27-38 transposed: uk/vk/wkb += from divb on [-1,n+3]^3,
 * then divb = 0 there.  The RMS output has no adjoint. */
void caldiv_b_(const int *ipara, const int *ndiv, const int *ni, const int *nj, const int *nk,
               const double *dxinv, const double *dyinv, const double *dzinv, const double *dt,
               double *ukb, double *vkb, double *wkb, double *divb);

/* srcs/solp_fft_tdma1.f:25-47 transposed: divb(interior) += Re(unnormalised
 * inverse DFT of phir_rb + i phir_ib) / (ni nk); phir_rb = phir_ib = 0. */
void solp_fft_tdma1_b_(const int *ni, const int *nj, const int *nk, double *fftdata, double *divb,
                       double *phir_rb, double *phir_ib);

/* srcs/trans_r2w.f, srcs/trans_w2r.f (one rank: a copy) transposed:
 * source_b += dest_b, dest_b = 0 (nothing when the arrays coincide). */
void trans_r2w_b_(const int *ipara, const int *ID, const int *ndiv, const int *svall,
                  const int *key, double *sendbuf, double *recvbuf, double *rb, double *wb);
void trans_w2r_b_(const int *ipara, const int *ID, const int *ndiv, const int *svall,
                  const int *key, double *sendbuf, double *recvbuf, double *wb, double *rb);

/* srcs/solp_fft_tdma2.f:46-87 transposed: Thomas solves of the transposed
 * j systems, in place on phiw_rb, phiw_ib; the singular (0,0) mode solved
 * pinned (x(nj) = 0), the transpose of the pinned solve. */
void solp_fft_tdma2_b_(const int *ID, const int *ndiv, const int *svall, const double *rho,
                       const double *dxinv, const double *dzinv, const double *as_p_fft,
                       const double *an_p_fft, const double *ap_p_fft, double *atdma,
                       double *btdma_r, double *btdma_i, double *phiw_rb, double *phiw_ib);

/* srcs/solp_fft_tdma3.f:25-44 transposed: phir_rb + i phir_ib +=
 * unnormalised forward DFT of dpb(interior); dpb(interior) = 0. */
void solp_fft_tdma3_b_(const int *ni, const int *nj, const int *nk, double *fftdata,
                       double *phir_rb, double *phir_ib, double *dpb);

/* srcs/bnd_avzero.f:10-23 transposed: qb(interior) -= sum(qb over the
 * padded array) / (ni nj nk ndiv). */
void bnd_avzero_b_(const int *ipara, const int *ndiv, const int *ni, const int *nj, const int *nk,
                   double *qb);

/* srcs/bnd_neumann.f:14-41 and srcs/bnd_periodic.f:14-43 transposed: each
 * ghost's adjoint added to its source and cleared, in reverse order. */
void bnd_neumann_b_(const int *nID, const int *ni, const int *nj, const int *nk, double *qb);
void bnd_periodic_b_(const int *ni, const int *nj, const int *nk, double *qb);

/* srcs/corunp_explicit.f:26-74 transposed: unb, vnb, wnb unchanged (read);
 * dpb += the transposed gradients of them and pnb(interior); phatb +=
 * pnb(interior); pnb(interior) = 0. */
void corunp_explicit_b_(const int *nID, const int *ni, const int *nj, const int *nk,
                        const double *rho, const double *dxinv, const double *dyinv,
                        const double *dzinv, const double *dt, double *dpb, double *phatb,
                        double *unb, double *vnb, double *wnb, double *pnb);

/* srcs/bndu.f:29-158 transposed: the z fills, the x fills (shifted u and w
 * sources included), the wall rows; ukb/vkb/wkb in place, *uwallb +=
 * sum(-c1 ghost_b) at the bottom wall and sum(+c1 ghost_b) at the top. */
void bndu_b_(const int *nID, const int *ni, const int *nj, const int *nk, double *ukb,
             double *vkb, double *wkb, double *uwallb, const double *dy, const double *l1,
             const double *l2, const double *phi);

/* srcs/solu_sor4.f:131-293 transposed with the forward call's sweep count
 * frozen: *lp is that call's loop variable on exit (felgpu_sor_last_), and
 * min(lp+1, 21) sweeps are reversed.  src_ub/vb/wb incremented; unb, vnb,
 * wnb hold the adjoint of the solve's final state on entry and of its
 * initial state on return; *uwallb incremented.  Coefficients formed on the
 * fly from the viscosity when cal_arith_coef_vis_ deferred them, as in
 * solu_sor4_; otherwise the 45 arrays are read. */
void solu_sor4_b_(const int *ipara, const int *ID, const int *nID, const int *ndiv, const int *ni,
                  const int *nj, const int *nk, const int *key, double *sendjb, double *recvjb,
                  const int *nstep, const int *imon_t, const double *dt, const double *rho,
                  const double *rhon, double *c0, double *c1, double *c2, double *c3, double *c4,
                  double *c5, double *c6, double *c7, double *c8, double *c9, double *c10,
                  double *c11, double *c12, double *c13, double *c14, double *c15, double *c16,
                  double *c17, double *c18, double *c19, double *c20, double *c21, double *c22,
                  double *c23, double *c24, double *c25, double *c26, double *c27, double *c28,
                  double *c29, double *c30, double *c31, double *c32, double *c33, double *c34,
                  double *c35, double *c36, double *c37, double *c38, double *c39, double *c40,
                  double *c41, double *c42, double *c43, double *c44, double *src_ub,
                  double *src_vb, double *src_wb, double *unb, double *vnb, double *wnb,
                  double *uwallb, const double *dy, const double *l1, const double *l2,
                  const double *phi, const int *lp);

/* srcs/cpy.f:14-20 transposed: q1b += q2b, q2b = 0. */
void cpy_b_(const int *ni, const int *nj, const int *nk, double *q1b, double *q2b);

/* srcs/calphat.f:16-22 transposed: pb += 2 phatb, pob -= phatb, phatb = 0
 * on the box the forward writes (i >= -1). */
void calphat_b_(const int *ni, const int *nj, const int *nk, double *pb, double *pob, double *phatb);

/* srcs/cal_grad_prs.f:24-37 transposed: qb gathers the face-gradient
 * adjoints; they are consumed on [-2,n+2]^3. */
void cal_grad_prs_b_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                     const double *dyinv, const double *dzinv, double *qb, double *qx_ub,
                     double *qy_vb, double *qz_wb);

/* srcs/calsij.f:28-139 transposed: ukb, vkb, wkb gather the adjoints of
 * the six strain-rate components skb(..,1..6), consumed on their boxes. */
void calsij_b_(const int *ni, const int *nj, const int *nk, const double *dxinv,
               const double *dyinv, const double *dzinv, double *ukb, double *vkb, double *wkb,
               double *skb);

/* srcs/cal_arith_tau.f:21-118 transposed, viscosity rmu fixed. */
void cal_arith_tau_b_(const int *ni, const int *nj, const int *nk, double *sb, const double *rmu,
                      double *taub);

/* srcs/cal_div_tensor.f:20-86 transposed: taukb gathers; ukb, vkb, wkb
 * consumed (the forward overwrites them whole). */
void cal_div_tensor_b_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                       const double *dyinv, const double *dzinv, double *taukb, double *ukb,
                       double *vkb, double *wkb);

/* srcs/cal_srcu.f:51-122 transposed, densities fixed: every array term's
 * adjoint incremented from src_*b (interior, v to jen), which is consumed.
 * The gravity terms are constants of the map. */
void cal_srcu_b_(const int *nID, const int *ni, const int *nj, const int *nk, const double *rho_av,
                 const double *dt, const double *grvb, const double *grvp, double *ub, double *vb,
                 double *wb, const double *rho, const double *rhon, double *adv_ub, double *adv_vb,
                 double *adv_wb, double *adv_uob, double *adv_vob, double *adv_wob, double *prs_ub,
                 double *prs_vb, double *prs_wb, double *vis_ub, double *vis_vb, double *vis_wb,
                 double *fst_ub, double *fst_vb, double *fst_wb, double *fst_unb, double *fst_vnb,
                 double *fst_wnb, double *src_ub, double *src_vb, double *src_wb);

/* srcs/cal_advu_weno5.f:1-257 reversed at the state (u, v, w), which is not
 * a linear map: each velocity is followed by its adjoint (incremented); the
 * adv adjoints are consumed.  The derivative of the upwind switch |vel| is
 * taken as 0 at vel = 0. */
void cal_advu_weno5_b_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                       const double *dyinv, const double *dzinv, const double *u, double *ub,
                       const double *v, double *vb, const double *w, double *wb, double *adv_ub,
                       double *adv_vb, double *adv_wb);

/* srcs/cal_grad_p2a.f:30-71 transposed: qb gathers the vertex-gradient
 * adjoints, which are consumed on [-2,n+2]^3. */
void cal_grad_p2a_b_(const int *ID, const int *svall, const int *ni, const int *nj, const int *nk,
                     const double *dxinv, const double *dyinv_array, const double *dzinv, double *qb,
                     double *qxb, double *qyb, double *qzb);

/* srcs/solphi_mthinc1.f:1-292 reversed at its inputs: each of phi, phix,
 * phiy, phiz, u, v, w followed by its adjoint (incremented); the flux
 * adjoints consumed on [-1,n+1]^3.  Branches and upwind directions are the
 * forward's; the Newton offset is differentiated by the implicit-function
 * theorem at the forward's iterate. */
void solphi_mthinc1_b_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                       const double *dyinv, const double *dzinv, const double *dt,
                       const double *bet_mthinc, const double *phi, double *phib, const double *phix,
                       double *phixb, const double *phiy, double *phiyb, const double *phiz,
                       double *phizb, const double *u, double *ub, const double *v, double *vb,
                       const double *w, double *wb, double *flphixb, double *flphiyb,
                       double *flphizb);

/* srcs/solphi_mthinc2.f:30-41 reversed: each of u, v, w, flphix, flphiy,
 * flphiz, phi followed by its adjoint; phinb consumed on the interior (the
 * clamp passes the derivative where the value lies in [0,1]). */
void solphi_mthinc2_b_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                       const double *dyinv, const double *dzinv, const double *dt, const double *u,
                       double *ub, const double *v, double *vb, const double *w, double *wb,
                       const double *flphix, double *flphixb, const double *flphiy,
                       double *flphiyb, const double *flphiz, double *flphizb, const double *phi,
                       double *phib, double *phinb);

/* srcs/solphi_mthinc3.f:43-160 reversed: phix, phiy, phiz select the flat
 * cells; phi followed by its adjoint; phin is the forward's input phin and
 * phinb the adjoint of its output on entry, of its input on return. */
void solphi_mthinc3_b_(const int *ipara, const int *ni, const int *nj, const int *nk,
                       const double *dxinv, const double *dyinv, const double *dzinv,
                       const double *bet_mthinc, const double *phix, const double *phiy,
                       const double *phiz, const double *phi, double *phib, const double *phin,
                       double *phinb);

/* srcs/summation.f:27-50 reversed: qk the forward's input slices
 * (nbub+1 padded arrays), qkb(.,0) consumed into qkb(.,1..nbub). */
void summation_b_(const int *ni, const int *nj, const int *nk, const double *qk, double *qkb,
                  const int *nbub);

/* srcs/gnbc.f90:22-95 reversed: uk, wk, uwall, theta_0_array, zeta_array
 * each followed by its adjoint (incremented); theta_arrayb consumed on the
 * wall rows.  The clamp and acos pass the derivative strictly inside
 * (-1,1). */
void gnbc_b_(const int *nID, const int *ni, const int *nj, const int *nk, const double *uk,
             double *ukb, const double *wk, double *wkb, const double *uwall, double *uwallb,
             const double *theta_0_array, double *theta_0b, const double *surface_tension,
             const double *zeta_array, double *zetab, double *theta_arrayb);

/* srcs/bnd_contact_angle.f90:17-72 reversed: qk the forward's input, qkb the
 * adjoint of its output on entry and of its input on return; theta_array
 * followed by its adjoint (incremented). */
void bnd_contact_angle_b_(const int *nID, const int *ni, const int *nj, const int *nk,
                          const double *qk, double *qkb, const double *theta_array,
                          double *theta_arrayb, const double *dx, const double *dy,
                          const double *dz);

/* ---- Material properties, surface tension and parameter adjoints ----
 * <name>_bp_ is the adjoint with respect to every real argument: what
 * <name>_b_ does, plus the adjoints of the arrays <name>_b_ takes as fixed
 * (densities, viscosity, coefficients, slip lengths, phi), incremented.
 * Each array whose primal value the derivative needs is followed by its
 * adjoint.  No counterpart in the reference (it has no adjoint). */

/* srcs/bndu.f:1-161, all arguments: bndu_b_ plus l1b, l2b, phib
 * incremented.  uk, wk: the call's input velocities (rows 1 and nj read). */
void bndu_bp_(const int *nID, const int *ni, const int *nj, const int *nk, const double *uk,
              double *ukb, double *vkb, const double *wk, double *wkb, const double *uwall,
              double *uwallb, const double *dy, const double *l1, double *l1b, const double *l2,
              double *l2b, const double *phi, double *phib);

/* srcs/solu_sor4.f:1-308, all arguments, sweep count frozen at *lp (as in
 * solu_sor4_b_): adds the adjoints of rho, rhon, the 45 coefficients (c0..c44
 * then their adjoints c0b..c44b), l1, l2 and phi.  src_*, un, vn, wn are the
 * call's inputs, each followed by its adjoint.  Coefficients are read from
 * the arrays.  Device work memory, nsw = min(lp+1,21): 3 padded arrays plus
 * an undo log of 12 nsw (a quarter array + three ghost shells). */
void solu_sor4_bp_(const int *ipara, const int *ID, const int *nID, const int *ndiv, const int *ni,
                   const int *nj, const int *nk, const int *key, double *sendjb, double *recvjb,
                   const int *nstep, const int *imon_t, const double *dt, const double *rho,
                   double *rhob, const double *rhon, double *rhonb,
                   double *c0, double *c1, double *c2, double *c3, double *c4, double *c5,
                   double *c6, double *c7, double *c8, double *c9,
                   double *c10, double *c11, double *c12, double *c13,
                   double *c14, double *c15, double *c16, double *c17,
                   double *c18, double *c19, double *c20, double *c21,
                   double *c22, double *c23, double *c24, double *c25,
                   double *c26, double *c27, double *c28, double *c29,
                   double *c30, double *c31, double *c32, double *c33,
                   double *c34, double *c35, double *c36, double *c37,
                   double *c38, double *c39, double *c40, double *c41,
                   double *c42, double *c43, double *c44,
                   double *c0b, double *c1b, double *c2b, double *c3b, double *c4b, double *c5b,
                   double *c6b, double *c7b, double *c8b, double *c9b,
                   double *c10b, double *c11b, double *c12b, double *c13b,
                   double *c14b, double *c15b, double *c16b, double *c17b,
                   double *c18b, double *c19b, double *c20b, double *c21b,
                   double *c22b, double *c23b, double *c24b, double *c25b,
                   double *c26b, double *c27b, double *c28b, double *c29b,
                   double *c30b, double *c31b, double *c32b, double *c33b,
                   double *c34b, double *c35b, double *c36b, double *c37b,
                   double *c38b, double *c39b, double *c40b, double *c41b,
                   double *c42b, double *c43b, double *c44b,
                   const double *src_u, double *src_ub, const double *src_v, double *src_vb,
                   const double *src_w, double *src_wb, const double *un, double *unb,
                   const double *vn, double *vnb, const double *wn, double *wnb,
                   const double *uwall, double *uwallb, const double *dy, const double *l1,
                   double *l1b, const double *l2, double *l2b, const double *phi, double *phib,
                   const int *lp);

/* srcs/cal_arith_coef_vis.f then srcs/solu_sor4.f, all arguments, the 45
 * coefficients formed from the viscosity rmu (dxinv, dyinv, dzinv): what
 * solu_sor4_bp_ and then cal_arith_coef_vis_b_ give, without coefficient
 * arrays; rmub incremented.  Work memory: solu_sor4_bp_'s plus 18 arrays. */
void solu_sor4_bpv_(const int *ipara, const int *ID, const int *nID, const int *ndiv, const int *ni,
                    const int *nj, const int *nk, const int *key, double *sendjb, double *recvjb,
                    const int *nstep, const int *imon_t, const double *dt, const double *rho,
                    double *rhob, const double *rhon, double *rhonb, const double *dxinv,
                    const double *dyinv, const double *dzinv, const double *rmu, double *rmub,
                    const double *src_u, double *src_ub, const double *src_v, double *src_vb,
                    const double *src_w, double *src_wb, const double *un, double *unb,
                    const double *vn, double *vnb, const double *wn, double *wnb,
                    const double *uwall, double *uwallb, const double *dy, const double *l1,
                    double *l1b, const double *l2, double *l2b, const double *phi, double *phib,
                    const int *lp);

/* srcs/cal_arith_tau.f:21-118, all arguments: cal_arith_tau_b_ plus rmub. */
void cal_arith_tau_bp_(const int *ni, const int *nj, const int *nk, const double *s, double *sb,
                       const double *rmu, double *rmub, double *taub);

/* srcs/cal_srcu.f:51-122, all arguments: cal_srcu_b_ plus rhob, rhonb
 * (r_b = src_b (prs - 0.5 vis)/r^2 spread over the face's four cells). */
void cal_srcu_bp_(const int *nID, const int *ni, const int *nj, const int *nk, const double *rho_av,
                  const double *dt, const double *grvb, const double *grvp, double *ub, double *vb,
                  double *wb, const double *rho, double *rhob, const double *rhon, double *rhonb,
                  double *adv_ub, double *adv_vb, double *adv_wb, double *adv_uob, double *adv_vob,
                  double *adv_wob, const double *prs_u, double *prs_ub, const double *prs_v,
                  double *prs_vb, const double *prs_w, double *prs_wb, const double *vis_u,
                  double *vis_ub, const double *vis_v, double *vis_vb, const double *vis_w,
                  double *vis_wb, double *fst_ub, double *fst_vb, double *fst_wb, double *fst_unb,
                  double *fst_vnb, double *fst_wnb, double *src_ub, double *src_vb,
                  double *src_wb);

/* srcs/cal_arith_coef_vis.f:1-411 transposed (linear in the viscosity):
 * rmukb incremented, the 45 coefficient adjoints consumed. */
void cal_arith_coef_vis_b_(const int *ni, const int *nj, const int *nk, const double *dxinv,
                           const double *dyinv, const double *dzinv, double *rmukb,
                           double *c0b, double *c1b, double *c2b, double *c3b, double *c4b, double *c5b,
                   double *c6b, double *c7b, double *c8b, double *c9b,
                   double *c10b, double *c11b, double *c12b, double *c13b,
                   double *c14b, double *c15b, double *c16b, double *c17b,
                   double *c18b, double *c19b, double *c20b, double *c21b,
                   double *c22b, double *c23b, double *c24b, double *c25b,
                   double *c26b, double *c27b, double *c28b, double *c29b,
                   double *c30b, double *c31b, double *c32b, double *c33b,
                   double *c34b, double *c35b, double *c36b, double *c37b,
                   double *c38b, double *c39b, double *c40b, double *c41b,
                   double *c42b, double *c43b, double *c44b);

/* srcs/cal_arith_mean.f:18-26 transposed: phikb += (qg - ql) qkb; qkb
 * consumed. */
void cal_arith_mean_b_(const int *ni, const int *nj, const int *nk, double *phikb, const double *ql,
                       const double *qg, double *qkb);

/* srcs/calrn.f:1-95 reversed (nonlinear): phix, phiy, phiz each followed by
 * its adjoint (incremented); rnb (three components) consumed.  d|r|/dr is
 * taken as 0 at r = 0. */
void calrn_b_(const int *ni, const int *nj, const int *nk, const double *dxinv, const double *dyinv,
              const double *dzinv, const int *ip, const int *jp, const int *kp, const double *phix,
              double *phixb, const double *phiy, double *phiyb, const double *phiz, double *phizb,
              double *rnb);

/* srcs/calrkap.f:1-102 transposed: the four normal adjoints incremented,
 * rkapb consumed. */
void calrkap_b_(const int *ni, const int *nj, const int *nk, const double *dxinv, const double *dyinv,
                const double *dzinv, double *rn_ppb, double *rn_12b, double *rn_13b, double *rn_23b,
                double *rkapb);

/* srcs/cal_grad_p2uvw.f:1-47 transposed: qb incremented, the face-gradient
 * adjoints consumed on [-2,n+2]^3. */
void cal_grad_p2uvw_b_(const int *ID, const int *svall, const int *ni, const int *nj, const int *nk,
                       const double *dxinv, const double *dyinv_array, const double *dzinv,
                       double *qb, double *qx_ub, double *qy_vb, double *qz_wb);

/* srcs/calfst.f:1-50 reversed (bilinear): the face gradients and rkap each
 * followed by its adjoint (incremented); the force adjoints consumed. */
void calfst_b_(const int *ni, const int *nj, const int *nk, const double *rhol, const double *rhog,
               const double *surface_tension, const double *phix_u, double *phix_ub,
               const double *phiy_v, double *phiy_vb, const double *phiz_w, double *phiz_wb,
               const double *rkap, double *rkapb, double *fst_ub, double *fst_vb, double *fst_wb);

/* srcs/sum_fst.f:1-37 transposed: f*b += s*b (s*b unchanged). */
void sum_fst_b_(const int *ni, const int *nj, const int *nk, double *fub, double *fvb, double *fwb,
                const double *sub, const double *svb, const double *swb);

/* How the last solu_sor4_b_ call obtained its coefficients: 1 formed on the
 * fly from the viscosity, 0 read from the 45 arrays, -1 before any call. */
int felgpu_sor_b_onfly_(void);

#ifdef __cplusplus
}
#endif

#endif /* FELGPU_H */
