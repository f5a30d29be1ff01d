// k_visc.hip -- mixture properties, pressure predictor, explicit viscous
// stress, momentum source and the viscous stencil coefficients.
//
// Reference: srcs/cal_arith_mean.f, srcs/calphat.f, srcs/cal_grad_prs.f,
// srcs/calsij.f, srcs/cal_arith_tau.f, srcs/cal_div_tensor.f,
// srcs/cal_srcu.f, srcs/cal_arith_coef_vis.f, srcs/cpy.f, srcs/init_q.f,
// srcs/sum_fst.f.  Each kernel writes exactly the index ranges the reference
// loops write and leaves every other element untouched.
#include "felgpu.hpp"
#include "felgpu_visc.hpp"

namespace fel {

__device__ __forceinline__ bool inb(int i, int a, int b) { return i >= a && i <= b; }

void arith_mean(const Dims d, const double* phik, double ql, double qg, double* qk) {
  launch_flat(d.n3, [=] __device__(int m) {
    const double pl = 1.0 - phik[m], pg = phik[m];
    qk[m] = ql * pl + qg * pg;
  });
}

void calphat(const Dims d, const double* p, const double* po, double* phat) {
  launch_box(Box{-1, d.ni + 3, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    phat[c] = p[c] * 2.0 - po[c];
  });
}

void grad_faces(const Dims d, double dxinv, double dyinv, double dzinv, const double* q, double* qx,
                double* qy, double* qz) {
  launch_box(Box{-2, d.ni + 2, -2, d.nj + 2, -2, d.nk + 2}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const double q0 = q[c];
    qx[c] = (-q0 + q[c + 1]) * dxinv;
    qy[c] = (-q0 + q[c + d.sx]) * dyinv;
    qz[c] = (-q0 + q[c + d.sx * d.sy]) * dzinv;
  });
}

// srcs/calsij.f:28-139
void calsij(const Dims d, double dxinv, double dyinv, double dzinv, const double* u, const double* v,
            const double* w, double* s) {
  const int n3 = d.n3, ni = d.ni, nj = d.nj, nk = d.nk;
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    if (i >= -1) s[c] = (-u[c - X] + u[c]) * dxinv;
    if (j >= -1) s[n3 + c] = (-v[c - Y] + v[c]) * dyinv;
    if (i >= -1 && j >= -1 && k >= -1) s[2 * n3 + c] = (-w[c - Z] + w[c]) * dzinv;
    if (i <= ni + 2 && j <= nj + 2)
      s[3 * n3 + c] = (+(-u[c] + u[c + Y]) * dyinv + (-v[c] + v[c + X]) * dxinv) * 0.5;
    if (i <= ni + 2 && k <= nk + 2)
      s[4 * n3 + c] = (+(-u[c] + u[c + Z]) * dzinv + (-w[c] + w[c + X]) * dxinv) * 0.5;
    if (j <= nj + 2 && k <= nk + 2)
      s[5 * n3 + c] = ((-v[c] + v[c + Z]) * dzinv + (-w[c] + w[c + Y]) * dyinv) * 0.5;
  });
}

// srcs/cal_arith_tau.f:21-117
void arith_tau(const Dims d, const double* s, const double* rmu, double* tau) {
  const int n3 = d.n3, ni = d.ni, nj = d.nj, nk = d.nk;
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    const double r = rmu[c];
    tau[c] = s[c] * r * 2.0;
    tau[n3 + c] = s[n3 + c] * r * 2.0;
    tau[2 * n3 + c] = s[2 * n3 + c] * r * 2.0;
    if (i <= ni + 2 && j <= nj + 2) {
      const double m = +0.25 * r + 0.25 * rmu[c + X] + 0.25 * rmu[c + Y] + 0.25 * rmu[c + X + Y];
      tau[3 * n3 + c] = s[3 * n3 + c] * m * 2.0;
    }
    if (i <= ni + 2 && k <= nk + 2) {
      const double m = +0.25 * r + 0.25 * rmu[c + X] + 0.25 * rmu[c + Z] + 0.25 * rmu[c + X + Z];
      tau[4 * n3 + c] = s[4 * n3 + c] * m * 2.0;
    }
    if (j <= nj + 2 && k <= nk + 2) {
      const double m = +0.25 * r + 0.25 * rmu[c + Y] + 0.25 * rmu[c + Z] + 0.25 * rmu[c + Y + Z];
      tau[5 * n3 + c] = s[5 * n3 + c] * m * 2.0;
    }
  });
}

// srcs/cal_div_tensor.f:20-86 (clearing folded into the same pass)
void div_tensor(const Dims d, double dxinv, double dyinv, double dzinv, const double* tau, double* vu,
                double* vv, double* vw) {
  const int n3 = d.n3, ni = d.ni, nj = d.nj, nk = d.nk;
  const double *t1 = tau, *t2 = tau + n3, *t3 = tau + 2 * n3, *t4 = tau + 3 * n3, *t5 = tau + 4 * n3,
               *t6 = tau + 5 * n3;
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    double a = 0.0, b = 0.0, e = 0.0;
    if (inb(k, -1, nk + 3) && inb(j, -1, nj + 3) && inb(i, -2, ni + 2))
      a = +(-t1[c] + t1[c + X]) * dxinv + (-t4[c - Y] + t4[c]) * dyinv + (-t5[c - Z] + t5[c]) * dzinv;
    if (inb(k, -1, nk + 3) && inb(j, -2, nj + 2) && inb(i, -1, ni + 3))
      b = +(-t4[c - X] + t4[c]) * dxinv + (-t2[c] + t2[c + Y]) * dyinv + (-t6[c - Z] + t6[c]) * dzinv;
    if (inb(k, -2, nk + 2) && inb(j, -1, nj + 3) && inb(i, -1, ni + 3))
      e = +(-t5[c - X] + t5[c]) * dxinv + (-t6[c - Y] + t6[c]) * dyinv + (-t3[c] + t3[c + Z]) * dzinv;
    vu[c] = a;
    vv[c] = b;
    vw[c] = e;
  });
}

// srcs/cal_srcu.f:59-136
void srcu(const Dims d, int jen, double dtinv, double grvb, double grvp, SrcArgs a) {
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    {
      const double r = (a.rho[c] + a.rho[c + X] + a.rhon[c] + a.rhon[c + X]) * 0.25;
      a.src_u[c] = +a.u[c] * dtinv - a.adv_u[c] * 1.5 + a.adv_uo[c] * 0.5 - a.prs_u[c] / r +
                   a.vis_u[c] / r * 0.5 + a.fst_u[c] * 0.5 + a.fst_un[c] * 0.5 + grvb;
    }
    if (j <= jen) {
      const double r = (a.rho[c] + a.rho[c + Y] + a.rhon[c] + a.rhon[c + Y]) * 0.25;
      a.src_v[c] = +a.v[c] * dtinv - a.adv_v[c] * 1.5 + a.adv_vo[c] * 0.5 - a.prs_v[c] / r +
                   a.vis_v[c] / r * 0.5 + a.fst_v[c] * 0.5 + a.fst_vn[c] * 0.5 + grvp;
    }
    {
      const double r = (a.rho[c] + a.rho[c + Z] + a.rhon[c] + a.rhon[c + Z]) * 0.25;
      a.src_w[c] = +a.w[c] * dtinv - a.adv_w[c] * 1.5 + a.adv_wo[c] * 0.5 - a.prs_w[c] / r +
                   a.vis_w[c] / r * 0.5 + a.fst_w[c] * 0.5 + a.fst_wn[c] * 0.5;
    }
  });
}

// srcs/cal_arith_coef_vis.f:99-407: zero on the padded image, the 45
// coefficients on [-1,n+2]^3.
void coef_vis(const Dims d, double dxinv, double dyinv, double dzinv, const double* rmu, Coef45 a) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const VisScales sc = vis_scales(dxinv, dyinv, dzinv);
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    double v[45];
    if (inb(i, -1, ni + 2) && inb(j, -1, nj + 2) && inb(k, -1, nk + 2)) {
      coef_u(d, sc, rmu, c, v);
      coef_v(d, sc, rmu, c, v + 15);
      coef_w(d, sc, rmu, c, v + 30);
    } else {
#pragma unroll
      for (int m = 0; m < 45; m++) v[m] = 0.0;
    }
#pragma unroll
    for (int m = 0; m < 45; m++) a.p[m][c] = v[m];
  });
}

void coef_vis(const Dims d, double dxinv, double dyinv, double dzinv, const double* rmu,
              double* const* out45) {
  Coef45 a;
  for (int m = 0; m < 45; m++) a.p[m] = out45[m];
  coef_vis(d, dxinv, dyinv, dzinv, rmu, a);
}

void copy(const Dims d, const double* q1, double* q2) {
  FEL_CHECK(hipMemcpyAsync(q2, q1, sizeof(double) * (size_t)d.n3, hipMemcpyDeviceToDevice, stream()));
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/cal_arith_mean.f:1
void cal_arith_mean_(const int* ni, const int* nj, const int* nk, const double* phik,
                     const double* ql, const double* qg, double* qk) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_arith_mean_");
  const double* p = c.in(phik, d.n3);
  double* q = c.out(qk, d.n3);
  arith_mean(d, p, *ql, *qg, q);
  c.finish();
}

// srcs/calphat.f:1
void calphat_(const int* ni, const int* nj, const int* nk, const double* p, const double* po,
              double* phat) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("calphat_");
  const double* dp = c.in(p, d.n3);
  const double* dpo = c.in(po, d.n3);
  double* dph = c.io(phat, d.n3);
  calphat(d, dp, dpo, dph);
  c.finish();
}

// srcs/cal_grad_prs.f:1-3
void cal_grad_prs_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                   const double* dyinv, const double* dzinv, const double* q, double* qx_u,
                   double* qy_v, double* qz_w) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_grad_prs_");
  const double* dq = c.in(q, d.n3);
  double* x = c.io(qx_u, d.n3);
  double* y = c.io(qy_v, d.n3);
  double* z = c.io(qz_w, d.n3);
  grad_faces(d, *dxinv, *dyinv, *dzinv, dq, x, y, z);
  c.finish();
}

// srcs/calsij.f:5-6
void calsij_(const int* ni, const int* nj, const int* nk, const double* dxinv, const double* dyinv,
             const double* dzinv, const double* uk, const double* vk, const double* wk, double* sk) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("calsij_");
  const double* u = c.in(uk, d.n3);
  const double* v = c.in(vk, d.n3);
  const double* w = c.in(wk, d.n3);
  double* s = c.io(sk, 6 * (size_t)d.n3);
  calsij(d, *dxinv, *dyinv, *dzinv, u, v, w, s);
  c.finish();
}

// srcs/cal_arith_tau.f:7
void cal_arith_tau_(const int* ni, const int* nj, const int* nk, const double* s, const double* rmu,
                    double* tau) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_arith_tau_");
  const double* ds = c.in(s, 6 * (size_t)d.n3);
  const double* dr = c.in(rmu, d.n3);
  double* dt = c.io(tau, 6 * (size_t)d.n3);
  arith_tau(d, ds, dr, dt);
  c.finish();
}

// srcs/cal_div_tensor.f:1-2
void cal_div_tensor_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                     const double* dyinv, const double* dzinv, const double* tauk, double* uk,
                     double* vk, double* wk) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_div_tensor_");
  const double* t = c.in(tauk, 6 * (size_t)d.n3);
  double* u = c.out(uk, d.n3);
  double* v = c.out(vk, d.n3);
  double* w = c.out(wk, d.n3);
  div_tensor(d, *dxinv, *dyinv, *dzinv, t, u, v, w);
  c.finish();
}

// srcs/cal_srcu.f:1-10
void cal_srcu_(const int* nID, const int* ni, const int* nj, const int* nk, const double* rho_av,
               const double* dt, const double* grvb, const double* grvp, const double* u,
               const double* v, const double* w, const double* rho, const double* rhon,
               const double* adv_u, const double* adv_v, const double* adv_w, const double* adv_uo,
               const double* adv_vo, const double* adv_wo, const double* prs_u, const double* prs_v,
               const double* prs_w, const double* vis_u, const double* vis_v, const double* vis_w,
               const double* fst_u, const double* fst_v, const double* fst_w, const double* fst_un,
               const double* fst_vn, const double* fst_wn, double* src_u, double* src_v,
               double* src_w) {
  (void)rho_av;
  const Dims d = make_dims(*ni, *nj, *nk);
  const int n = d.n3;
  Call c("cal_srcu_");
  SrcArgs a;
  a.u = c.in(u, n); a.v = c.in(v, n); a.w = c.in(w, n);
  a.rho = c.in(rho, n); a.rhon = c.in(rhon, n);
  a.adv_u = c.in(adv_u, n); a.adv_v = c.in(adv_v, n); a.adv_w = c.in(adv_w, n);
  a.adv_uo = c.in(adv_uo, n); a.adv_vo = c.in(adv_vo, n); a.adv_wo = c.in(adv_wo, n);
  a.prs_u = c.in(prs_u, n); a.prs_v = c.in(prs_v, n); a.prs_w = c.in(prs_w, n);
  a.vis_u = c.in(vis_u, n); a.vis_v = c.in(vis_v, n); a.vis_w = c.in(vis_w, n);
  a.fst_u = c.in(fst_u, n); a.fst_v = c.in(fst_v, n); a.fst_w = c.in(fst_w, n);
  a.fst_un = c.in(fst_un, n); a.fst_vn = c.in(fst_vn, n); a.fst_wn = c.in(fst_wn, n);
  a.src_u = c.io(src_u, n); a.src_v = c.io(src_v, n); a.src_w = c.io(src_w, n);
  const int jen = (nID[kYPlus] < 0) ? *nj - 1 : *nj;
  srcu(d, jen, 1.0 / *dt, *grvb, *grvp, a);
  c.finish();
}

// srcs/cal_arith_coef_vis.f:6-19
void cal_arith_coef_vis_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                         const double* dyinv, const double* dzinv, const double* rmuk, double* a0,
                         double* a1, double* a2, double* a3, double* a4, double* a5, double* a6,
                         double* a7, double* a8, double* a9, double* a10, double* a11, double* a12,
                         double* a13, double* a14, double* a15, double* a16, double* a17,
                         double* a18, double* a19, double* a20, double* a21, double* a22,
                         double* a23, double* a24, double* a25, double* a26, double* a27,
                         double* a28, double* a29, double* a30, double* a31, double* a32,
                         double* a33, double* a34, double* a35, double* a36, double* a37,
                         double* a38, double* a39, double* a40, double* a41, double* a42,
                         double* a43, double* a44) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_arith_coef_vis_");
  double* h[45] = {a0,  a1,  a2,  a3,  a4,  a5,  a6,  a7,  a8,  a9,  a10, a11, a12, a13, a14,
                   a15, a16, a17, a18, a19, a20, a21, a22, a23, a24, a25, a26, a27, a28, a29,
                   a30, a31, a32, a33, a34, a35, a36, a37, a38, a39, a40, a41, a42, a43, a44};
  // a pending record for other output arrays is materialised before it is replaced
  DeferredCoef& dc = deferred_coef();
  if (dc.active) {
    // the same output arrays are about to be redefined: drop the pending
    // record; other outputs still owe their values
    bool same = true;
    for (int m = 0; m < 45 && same; m++) same = dc.out_host[m] == (const void*)h[m];
    if (!same) deferred_coef_flush();
    dc.active = false;
  }
  bool resident = is_registered(rmuk, d.n3);
  for (int m = 0; m < 45 && resident; m++) resident = is_registered(h[m], d.n3);
  const double* r = c.in(rmuk, d.n3);
  if (resident && !deferred_coef_disabled()) {
    dc.d = d;
    dc.dxinv = *dxinv; dc.dyinv = *dyinv; dc.dzinv = *dzinv;
    dc.src_host = rmuk;
    dc.src_dev = r;
    for (int m = 0; m < 45; m++) {
      dc.out_host[m] = h[m];
      dc.out_dev[m] = c.out(h[m], d.n3);
    }
    dc.active = true;
  } else {
    Coef45 a;
    for (int m = 0; m < 45; m++) a.p[m] = c.out(h[m], d.n3);
    coef_vis(d, *dxinv, *dyinv, *dzinv, r, a);
  }
  c.finish();
}

// srcs/cpy.f:1
void cpy_(const int* ni, const int* nj, const int* nk, const double* q1, double* q2) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cpy_");
  const double* a = c.in(q1, d.n3);
  double* b = c.out(q2, d.n3);
  copy(d, a, b);
  c.finish();
  copy_content_version(q1, q2, d.n3);
}

// srcs/init_q.f:1
void init_q_(const int* ni, const int* nj, const int* nk, double* q1, double* q2, double* q3) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("init_q_");
  double* p[3] = {c.out(q1, d.n3), c.out(q2, d.n3), c.out(q3, d.n3)};
  for (int m = 0; m < 3; m++)
    FEL_CHECK(hipMemsetAsync(p[m], 0, sizeof(double) * (size_t)d.n3, stream()));
  c.finish();
}

// srcs/sum_fst.f:1-3
void sum_fst_(const int* ni, const int* nj, const int* nk, const double* fst_u, const double* fst_v,
              const double* fst_w, double* sum_u, double* sum_v, double* sum_w) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("sum_fst_");
  const double *fu = c.in(fst_u, d.n3), *fv = c.in(fst_v, d.n3), *fw = c.in(fst_w, d.n3);
  double *su = c.io(sum_u, d.n3), *sv = c.io(sum_v, d.n3), *sw = c.io(sum_w, d.n3);
  launch_flat(d.n3, [=] __device__(int m) {
    su[m] = su[m] + fu[m];
    sv[m] = sv[m] + fv[m];
    sw[m] = sw[m] + fw[m];
  });
  c.finish();
}

}  // extern "C"
