// k_visc.hip -- mixture properties, pressure predictor, explicit viscous
// stress, momentum source and the viscous stencil coefficients.
//
// Reference: srcs/cal_arith_mean.f, srcs/calphat.f, srcs/cal_grad_prs.f,
// srcs/calsij.f, srcs/cal_arith_tau.f, srcs/cal_div_tensor.f,
// srcs/cal_srcu.f, srcs/cal_arith_coef_vis.f, srcs/cpy.f, srcs/init_q.f,
// srcs/sum_fst.f.  Each kernel writes exactly the index ranges the reference
// loops write and leaves every other element untouched.
#include <cstdlib>

#include "felgpu.hpp"
#include "felgpu_visc.hpp"

namespace fel {

__device__ __forceinline__ bool inb(int i, int a, int b) { return i >= a && i <= b; }

// T is the flow element type (double, float in the FP32 mode, Dual in the
// tangent mode); the volume fraction is in the phase type PT<T> (FP64 unless
// in the tangent mode).  With T = double every expression is the
// reference's FP64 one.

template <class T, class P = PT<T>>
void arith_mean(const Dims d, const P* phik, double ql, double qg, T* qk) {
  launch_flat(d.n3, [=] __device__(int m) {
    const P pl = 1.0 - phik[m], pg = phik[m];
    qk[m] = T(ql * pl + qg * pg);
  });
}

template <class T>
void calphat(const Dims d, const T* p, const T* po, T* phat) {
  using A = acc_t<T>;  // FP64 for FP32 storage, rounded once at the store
  launch_box(Box{-1, d.ni + 3, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    phat[c] = T(A(p[c]) * A(2.0) - A(po[c]));
  });
}

template <class T>
void grad_faces(const Dims d, double dxinv_, double dyinv_, double dzinv_, const T* q, T* qx, T* qy,
                T* qz) {
  const T dxinv = T(dxinv_), dyinv = T(dyinv_), dzinv = T(dzinv_);
  launch_box(Box{-2, d.ni + 2, -2, d.nj + 2, -2, d.nk + 2}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const T q0 = q[c];
    qx[c] = (-q0 + q[c + 1]) * dxinv;
    qy[c] = (-q0 + q[c + d.sx]) * dyinv;
    qz[c] = (-q0 + q[c + d.sx * d.sy]) * dzinv;
  });
}

// srcs/calsij.f:28-139
template <class T>
void calsij(const Dims d, double dxinv_, double dyinv_, double dzinv_, const T* u, const T* v,
            const T* w, T* s) {
  const int n3 = d.n3, ni = d.ni, nj = d.nj, nk = d.nk;
  const T dxinv = T(dxinv_), dyinv = T(dyinv_), dzinv = T(dzinv_);
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    if (i >= -1) s[c] = (-u[c - X] + u[c]) * dxinv;
    if (j >= -1) s[n3 + c] = (-v[c - Y] + v[c]) * dyinv;
    if (i >= -1 && j >= -1 && k >= -1) s[2 * n3 + c] = (-w[c - Z] + w[c]) * dzinv;
    if (i <= ni + 2 && j <= nj + 2)
      s[3 * n3 + c] = (+(-u[c] + u[c + Y]) * dyinv + (-v[c] + v[c + X]) * dxinv) * T(0.5);
    if (i <= ni + 2 && k <= nk + 2)
      s[4 * n3 + c] = (+(-u[c] + u[c + Z]) * dzinv + (-w[c] + w[c + X]) * dxinv) * T(0.5);
    if (j <= nj + 2 && k <= nk + 2)
      s[5 * n3 + c] = ((-v[c] + v[c + Z]) * dzinv + (-w[c] + w[c + Y]) * dyinv) * T(0.5);
  });
}

// srcs/cal_arith_tau.f:21-117
template <class T>
void arith_tau(const Dims d, const T* s, const T* rmu, T* tau) {
  const int n3 = d.n3, ni = d.ni, nj = d.nj, nk = d.nk;
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    const T r = rmu[c], q = T(0.25);
    tau[c] = s[c] * r * T(2.0);
    tau[n3 + c] = s[n3 + c] * r * T(2.0);
    tau[2 * n3 + c] = s[2 * n3 + c] * r * T(2.0);
    if (i <= ni + 2 && j <= nj + 2) {
      const T m = +q * r + q * rmu[c + X] + q * rmu[c + Y] + q * rmu[c + X + Y];
      tau[3 * n3 + c] = s[3 * n3 + c] * m * T(2.0);
    }
    if (i <= ni + 2 && k <= nk + 2) {
      const T m = +q * r + q * rmu[c + X] + q * rmu[c + Z] + q * rmu[c + X + Z];
      tau[4 * n3 + c] = s[4 * n3 + c] * m * T(2.0);
    }
    if (j <= nj + 2 && k <= nk + 2) {
      const T m = +q * r + q * rmu[c + Y] + q * rmu[c + Z] + q * rmu[c + Y + Z];
      tau[5 * n3 + c] = s[5 * n3 + c] * m * T(2.0);
    }
  });
}

// srcs/cal_div_tensor.f:20-86 (clearing folded into the same pass).  One
// thread per (i,j) column of the padded image and kDivChunk planes: tau13
// and tau23 at k-1 and tau33 at k+1 are the column's own values of the
// neighbouring plane, kept in registers instead of read again.  Each output
// is the expression of the one-thread-per-cell form on the same operands.
// Measured at 512x256x128 (FP64): one cell per thread 345 us, chunks of 2
// planes 328 us, 4: 334, 8: 344, 16: 361 (longer chunks leave too few
// threads to cover the load latency).
constexpr int kDivChunk = 2;

template <class T>
__global__ void __launch_bounds__(256) div_tensor_kernel(Dims d, T dxinv, T dyinv, T dzinv,
                                                         const T* __restrict__ tau,
                                                         T* __restrict__ vu, T* __restrict__ vv,
                                                         T* __restrict__ vw, int gx, int gy) {
  const int L = (int)blockIdx.x;
  const int bx = L % gx, tt = L / gx, by = tt % gy, bz = tt / gy;
  const int ni = d.ni, nj = d.nj, nk = d.nk, n3 = d.n3;
  const int i = -2 + bx * 64 + (int)threadIdx.x, j = -2 + by * 4 + (int)threadIdx.y;
  if (i > ni + 3 || j > nj + 3) return;
  const int k0 = -2 + bz * kDivChunk, k1 = min(nk + 3, k0 + kDivChunk - 1);
  const T *t1 = tau, *t2 = tau + n3, *t3 = tau + 2 * n3, *t4 = tau + 3 * n3, *t5 = tau + 4 * n3,
          *t6 = tau + 5 * n3;
  const int X = 1, Y = d.sx, Z = d.sx * d.sy;
  const bool ia = inb(i, -2, ni + 2), ib = inb(i, -1, ni + 3);
  const bool ja = inb(j, -1, nj + 3), jb = inb(j, -2, nj + 2);
  int c = d.ix(i, j, k0);
  T t5m = k0 - 1 >= -2 ? t5[c - Z] : T(0.0);  // tau13, tau23 at k-1
  T t6m = k0 - 1 >= -2 ? t6[c - Z] : T(0.0);
  T t3c = t3[c];                              // tau33 at k
  for (int k = k0; k <= k1; k++, c += Z) {
    const T t5c = t5[c], t6c = t6[c];
    const T t3p = k + 1 <= nk + 3 ? t3[c + Z] : T(0.0);
    T a = T(0.0), b = T(0.0), e = T(0.0);
    if (inb(k, -1, nk + 3) && ja && ia)
      a = +(-t1[c] + t1[c + X]) * dxinv + (-t4[c - Y] + t4[c]) * dyinv + (-t5m + t5c) * dzinv;
    if (inb(k, -1, nk + 3) && jb && ib)
      b = +(-t4[c - X] + t4[c]) * dxinv + (-t2[c] + t2[c + Y]) * dyinv + (-t6m + t6c) * dzinv;
    if (inb(k, -2, nk + 2) && ja && ib)
      e = +(-t5[c - X] + t5c) * dxinv + (-t6[c - Y] + t6c) * dyinv + (-t3c + t3p) * dzinv;
    vu[c] = a;
    vv[c] = b;
    vw[c] = e;
    t5m = t5c;
    t6m = t6c;
    t3c = t3p;
  }
}

template <class T>
void div_tensor(const Dims d, double dxinv, double dyinv, double dzinv, const T* tau, T* vu,
                T* vv, T* vw) {
  const int gx = (d.sx + 63) / 64, gy = (d.sy + 3) / 4, gz = (d.sz + kDivChunk - 1) / kDivChunk;
  div_tensor_kernel<T><<<dim3(gx * gy * gz), dim3(64, 4, 1), 0, stream()>>>(
      d, T(dxinv), T(dyinv), T(dzinv), tau, vu, vv, vw, gx, gy);
  FEL_CHECK(hipGetLastError());
}

// calsij followed by arith_tau in one pass (the chain calsij_ ->
// cal_arith_tau_ recorded by the runtime): every thread computes s at its
// cell, writes it where calsij writes, and writes tau where arith_tau
// writes, from that s and the viscosity -- the same operations in the same
// order, so the values are bitwise those of the two kernels, without the
// round trip of s through memory.  Where calsij leaves s untouched, tau is
// computed from the untouched contents (a.s[c] is never written there, so
// the read does not race).
template <class T>
struct SijTau {
  const T *u, *v, *w, *rmu;
  T *s, *tau;
  T dxinv, dyinv, dzinv;
};

// One thread per (i,j) column of the padded image, marching over a chunk of
// kSijChunk planes: the column's u, v (at k+1), w (at k) and viscosity (at
// k+1, i+1 / j+1 beside it) loaded for one plane are the values the next
// plane reads at k (and w at k-1), so each plane of the inputs is fetched
// once per column instead of once per plane touching it.  Per cell the
// operations are calsij's (srcs/calsij.f:28-139, where it writes) and
// arith_tau's (srcs/cal_arith_tau.f:21-118), in the same order.
// (measured at 512x256x128 in FP64: 674 us with 2 planes, 691 with 4, 723
// with 8, 752 with 16: short chunks leave more threads to cover latency)
constexpr int kSijChunk = 2;

template <class T>
__global__ void __launch_bounds__(256) sij_tau_kernel(Dims d, SijTau<T> a, int gx, int gy) {
  const int L = (int)blockIdx.x;
  const int bx = L % gx, t = L / gx, by = t % gy, bz = t / gy;
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int i = -2 + bx * 64 + (int)threadIdx.x;
  const int j = -2 + by * 4 + (int)threadIdx.y;
  if (i > ni + 3 || j > nj + 3) return;
  const int n3 = d.n3, X = 1, Y = d.sx, Z = d.sx * d.sy;
  const int k0 = -2 + bz * kSijChunk, k1 = min(nk + 3, k0 + kSijChunk - 1);
  const T *u = a.u, *v = a.v, *w = a.w, *rmu = a.rmu;
  const bool xl = i >= -1, xr = i <= ni + 2, yl = j >= -1, yr = j <= nj + 2;
  const T q = T(0.25), h = T(0.5), two = T(2.0);
  // the column at plane k0 (and w at k0-1)
  int c = d.ix(i, j, k0);
  T u_c = u[c], v_c = v[c], r_c = rmu[c];
  T r_x = xr ? rmu[c + X] : T(0.0), r_y = yr ? rmu[c + Y] : T(0.0);
  T w_m = k0 >= -1 ? w[c - Z] : T(0.0);
  for (int k = k0; k <= k1; k++, c += Z) {
    const bool kz = k <= nk + 2;
    const T u_xm = xl ? u[c - X] : T(0.0), u_yp = yr ? u[c + Y] : T(0.0),
            u_zp = kz ? u[c + Z] : T(0.0);
    const T v_ym = yl ? v[c - Y] : T(0.0), v_xp = xr ? v[c + X] : T(0.0),
            v_zp = kz ? v[c + Z] : T(0.0);
    const T w_c = w[c], w_xp = xr ? w[c + X] : T(0.0), w_yp = yr ? w[c + Y] : T(0.0);
    const T r_xy = (xr && yr) ? rmu[c + X + Y] : T(0.0), r_z = kz ? rmu[c + Z] : T(0.0);
    const T r_xz = (xr && kz) ? rmu[c + X + Z] : T(0.0), r_yz = (yr && kz) ? rmu[c + Y + Z] : T(0.0);
    T sv[6];
    bool sw[6];
    sw[0] = xl;
    sv[0] = sw[0] ? (-u_xm + u_c) * a.dxinv : a.s[c];
    sw[1] = yl;
    sv[1] = sw[1] ? (-v_ym + v_c) * a.dyinv : a.s[n3 + c];
    sw[2] = xl && yl && k >= -1;
    sv[2] = sw[2] ? (-w_m + w_c) * a.dzinv : a.s[2 * n3 + c];
    sw[3] = xr && yr;
    sv[3] = sw[3] ? (+(-u_c + u_yp) * a.dyinv + (-v_c + v_xp) * a.dxinv) * h : a.s[3 * n3 + c];
    sw[4] = xr && kz;
    sv[4] = sw[4] ? (+(-u_c + u_zp) * a.dzinv + (-w_c + w_xp) * a.dxinv) * h : a.s[4 * n3 + c];
    sw[5] = yr && kz;
    sv[5] = sw[5] ? ((-v_c + v_zp) * a.dzinv + (-w_c + w_yp) * a.dyinv) * h : a.s[5 * n3 + c];
#pragma unroll
    for (int m = 0; m < 6; m++)
      if (sw[m]) a.s[(size_t)m * n3 + c] = sv[m];
#pragma unroll
    for (int m = 0; m < 3; m++) a.tau[(size_t)m * n3 + c] = sv[m] * r_c * two;
    // tau where arith_tau writes it: the 4-point viscosity of the edge
    if (sw[3]) a.tau[3 * (size_t)n3 + c] = sv[3] * (+q * r_c + q * r_x + q * r_y + q * r_xy) * two;
    if (sw[4]) a.tau[4 * (size_t)n3 + c] = sv[4] * (+q * r_c + q * r_x + q * r_z + q * r_xz) * two;
    if (sw[5]) a.tau[5 * (size_t)n3 + c] = sv[5] * (+q * r_c + q * r_y + q * r_z + q * r_yz) * two;
    u_c = u_zp;
    v_c = v_zp;
    w_m = w_c;
    r_c = r_z;
    r_x = r_xz;
    r_y = r_yz;
  }
}

template <class T>
void sij_tau(const Dims d, SijTau<T> a) {
  const int gx = (d.sx + 63) / 64, gy = (d.sy + 3) / 4, gz = (d.sz + kSijChunk - 1) / kSijChunk;
  sij_tau_kernel<T><<<dim3(gx * gy * gz), dim3(64, 4, 1), 0, stream()>>>(d, a, gx, gy);
  FEL_CHECK(hipGetLastError());
}

// srcs/cal_srcu.f:59-136
template <class T>
void srcu(const Dims d, int jen, double dtinv_, double grvb_, double grvp_, SrcArgs<T> a) {
  // A: FP64 for FP32 storage, whose source (u/dt plus increments) is
  // rounded once at the store; FP64 and the tangent mode evaluate in T
  using A = acc_t<T>;
  const A dtinv = A(T(dtinv_)), grvb = A(T(grvb_)), grvp = A(T(grvp_));
  const A q = A(0.25), h = A(0.5), t15 = A(1.5);
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    auto sum = [&](const T* f, int bit) { return (a.addz & bit) ? A(T(0.0) + f[c]) : A(f[c]); };
    auto ld = [&](const T* f, int o_) { return A(f[o_]); };
    const size_t slab4 = 4 * r00_slab(d);
    const size_t o = pair_of(i, j, k) * r00_slab(d) + r00_index(d, (i - 1) >> 1, (j - 1) >> 1, k);
    if (a.r00d[0]) {
      // solu_sor4_'s face densities, in its arithmetic (srcs/solu_sor4.f:150-152)
      const T q4 = T(0.25);
      a.r00d[0][o] = (a.rho[c] + a.rho[c + X] + a.rhon[c] + a.rhon[c + X]) * q4;
      a.r00d[1][o] = (a.rho[c] + a.rho[c + Y] + a.rhon[c] + a.rhon[c + Y]) * q4;
      a.r00d[2][o] = (a.rho[c] + a.rho[c + Z] + a.rhon[c] + a.rhon[c + Z]) * q4;
    }
    {
      const A r = (ld(a.rho, c) + ld(a.rho, c + X) + ld(a.rhon, c) + ld(a.rhon, c + X)) * q;
      const A sv = +ld(a.u, c) * dtinv - ld(a.adv_u, c) * t15 + ld(a.adv_uo, c) * h -
                   ld(a.prs_u, c) / r + ld(a.vis_u, c) / r * h + sum(a.fst_u, 1) * h +
                   sum(a.fst_un, 8) * h + grvb;
      const T s = T(sv);
      a.src_u[c] = s;
      if (a.dense) a.dense[o] = s;
    }
    if (j <= jen) {
      const A r = (ld(a.rho, c) + ld(a.rho, c + Y) + ld(a.rhon, c) + ld(a.rhon, c + Y)) * q;
      const A sv = +ld(a.v, c) * dtinv - ld(a.adv_v, c) * t15 + ld(a.adv_vo, c) * h -
                   ld(a.prs_v, c) / r + ld(a.vis_v, c) / r * h + sum(a.fst_v, 2) * h +
                   sum(a.fst_vn, 16) * h + grvp;
      const T s = T(sv);
      a.src_v[c] = s;
      if (a.dense) a.dense[slab4 + o] = s;
    }
    {
      const A r = (ld(a.rho, c) + ld(a.rho, c + Z) + ld(a.rhon, c) + ld(a.rhon, c + Z)) * q;
      const A sv = +ld(a.w, c) * dtinv - ld(a.adv_w, c) * t15 + ld(a.adv_wo, c) * h -
                   ld(a.prs_w, c) / r + ld(a.vis_w, c) / r * h + sum(a.fst_w, 4) * h +
                   sum(a.fst_wn, 32) * h;
      const T s = T(sv);
      a.src_w[c] = s;
      if (a.dense) a.dense[2 * slab4 + o] = s;
    }
  });
}

// srcs/cal_arith_coef_vis.f:99-407: zero on the padded image, the 45
// coefficients on [-1,n+2]^3.
template <class T>
void coef_vis(const Dims d, double dxinv, double dyinv, double dzinv, const T* rmu, Coef45<T> a) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const VisScales<T> sc = vis_scales<T>(dxinv, dyinv, dzinv);
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    T v[45];
    if (inb(i, -1, ni + 2) && inb(j, -1, nj + 2) && inb(k, -1, nk + 2)) {
      coef_u(d, sc, rmu, c, v);
      coef_v(d, sc, rmu, c, v + 15);
      coef_w(d, sc, rmu, c, v + 30);
    } else {
#pragma unroll
      for (int m = 0; m < 45; m++) v[m] = T(0.0);
    }
#pragma unroll
    for (int m = 0; m < 45; m++) a.p[m][c] = v[m];
  });
}

void coef_vis_any(const Dims d, double dxinv, double dyinv, double dzinv, const void* rmu,
                  void* const* out45, int esize) {
  with_elem(esize, [&](auto t) {
    using T = decltype(t);
    Coef45<T> a;
    for (int m = 0; m < 45; m++) a.p[m] = static_cast<T*>(out45[m]);
    coef_vis<T>(d, dxinv, dyinv, dzinv, static_cast<const T*>(rmu), a);
  });
}

// ---------------------------------------------------------------------------
// Drop-in bodies, instantiated for the flow element type T.
// ---------------------------------------------------------------------------

template <class T>
void cal_arith_mean_t(const Dims d, const double* phik, double ql, double qg, double* qk) {
  Call c("cal_arith_mean_");
  const PT<T>* p = c.in_t<PT<T>>(phik, d.n3, "phik");
  T* q = c.all_t<T>(qk, d.n3, "qk");  // every element is written
  arith_mean<T>(d, p, ql, qg, q);
  c.finish();
}

template <class T>
void calphat_t(const Dims d, const double* p, const double* po, double* phat) {
  Call c("calphat_");
  const T* dp = c.in_t<T>(p, d.n3, "p");
  const T* dpo = c.in_t<T>(po, d.n3, "po");
  T* dph = c.io_t<T>(phat, d.n3, "phat");
  calphat<T>(d, dp, dpo, dph);
  c.finish();
}

template <class T>
void cal_grad_prs_t(const Dims d, double dxinv, double dyinv, double dzinv, const double* q,
                    double* qx_u, double* qy_v, double* qz_w) {
  Call c("cal_grad_prs_");
  const T* dq = c.in_t<T>(q, d.n3, "q");
  T* x = c.io_t<T>(qx_u, d.n3, "qx_u");
  T* y = c.io_t<T>(qy_v, d.n3, "qy_v");
  T* z = c.io_t<T>(qz_w, d.n3, "qz_w");
  grad_faces<T>(d, dxinv, dyinv, dzinv, dq, x, y, z);
  c.finish();
}

// The viscous-stress chain of the step (srcs/main.f90:537-540): calsij_
// records s, cal_arith_tau_ records tau, cal_div_tensor_ runs sij_tau and
// the divergence.  Any other access to s or tau, or a write to u, v, w or the
// viscosity, in between runs what has been recorded (calsij, then
// arith_tau) at that point.
constexpr int kChainVisc = 1;

struct ViscRecord {
  Dims d;
  int esize, stage;  // stage 1: s recorded, 2: s and tau recorded
  double dxinv, dyinv, dzinv;
  const void *u, *v, *w, *rmu;  // device
  void *s, *tau;                // device
  const void* s_host;
  const void* tau_host;
};
static ViscRecord g_visc;

static bool whole_registered(const void* host, size_t count, int esize) {
  return is_registered(host, count) && content_version(host, count) != 0 && elem_size(host) == esize;
}

template <class T>
static void visc_record_flush() {
  const ViscRecord r = g_visc;
  if (r.stage == 2) {
    sij_tau<T>(r.d, SijTau<T>{static_cast<const T*>(r.u), static_cast<const T*>(r.v),
                              static_cast<const T*>(r.w), static_cast<const T*>(r.rmu),
                              static_cast<T*>(r.s), static_cast<T*>(r.tau), T(r.dxinv),
                              T(r.dyinv), T(r.dzinv)});
    return;
  }
  calsij<T>(r.d, r.dxinv, r.dyinv, r.dzinv, static_cast<const T*>(r.u), static_cast<const T*>(r.v),
            static_cast<const T*>(r.w), static_cast<T*>(r.s));
}

template <class T>
void calsij_t(const Dims d, double dxinv, double dyinv, double dzinv, const double* uk,
              const double* vk, const double* wk, double* sk) {
  const size_t n = (size_t)d.n3, n6 = 6 * n;
  const int es = (int)sizeof(T);
  const bool defer = chain_deferral_enabled() && whole_registered(uk, n, es) &&
                     whole_registered(vk, n, es) && whole_registered(wk, n, es) &&
                     whole_registered(sk, n6, es);
  Call c("calsij_");
  const T* u = c.in_t<T>(uk, n, "uk");
  const T* v = c.in_t<T>(vk, n, "vk");
  const T* w = c.in_t<T>(wk, n, "wk");
  T* s = c.io_t<T>(sk, n6, "sk");
  if (defer && sk != uk && sk != vk && sk != wk) {
    deferred_chain_flush();
    g_visc = ViscRecord{d, es, 1, dxinv, dyinv, dzinv, u, v, w, nullptr, s, nullptr, sk, nullptr};
    DeferredChain& ch = deferred_chain();
    ch.active = true;
    ch.kind = kChainVisc;
    ch.outs = {sk};
    ch.srcs = {uk, vk, wk};
    ch.flush = [] { visc_record_flush<T>(); };
    c.finish();
    return;
  }
  calsij<T>(d, dxinv, dyinv, dzinv, u, v, w, s);
  c.finish();
}

template <class T>
void cal_arith_tau_t(const Dims d, const double* s, const double* rmu, double* tau) {
  const size_t n = (size_t)d.n3, n6 = 6 * n;
  DeferredChain& ch = deferred_chain();
  const bool chained = ch.active && ch.kind == kChainVisc && g_visc.stage == 1 &&
                       g_visc.s_host == s && g_visc.esize == (int)sizeof(T) &&
                       g_visc.d.ni == d.ni && g_visc.d.nj == d.nj && g_visc.d.nk == d.nk &&
                       whole_registered(rmu, n, (int)sizeof(T)) &&
                       whole_registered(tau, n6, (int)sizeof(T));
  if (chained) {
    // an argument that is one of the recorded arrays runs the chain here
    Call c("cal_arith_tau_");
    const T* dr = c.in_t<T>(rmu, n, "rmu");
    T* dt = c.io_t<T>(tau, n6, "tau");
    c.finish();
    if (ch.active && ch.kind == kChainVisc && g_visc.stage == 1) {
      g_visc.stage = 2;
      g_visc.rmu = dr;
      g_visc.tau = dt;
      g_visc.tau_host = tau;
      ch.outs = {s, tau};
      ch.srcs.push_back(rmu);
      return;
    }
  }
  Call c("cal_arith_tau_");
  const T* ds = c.in_t<T>(s, n6, "s");
  const T* dr = c.in_t<T>(rmu, n, "rmu");
  T* dt = c.io_t<T>(tau, n6, "tau");
  arith_tau<T>(d, ds, dr, dt);
  c.finish();
}

template <class T>
void cal_div_tensor_t(const Dims d, double dxinv, double dyinv, double dzinv, const double* tauk,
                      double* uk, double* vk, double* wk) {
  DeferredChain& ch = deferred_chain();
  if (ch.active && ch.kind == kChainVisc && g_visc.stage == 2 && g_visc.tau_host == tauk &&
      g_visc.esize == (int)sizeof(T) && g_visc.dxinv == dxinv && g_visc.dyinv == dyinv &&
      g_visc.dzinv == dzinv && g_visc.d.ni == d.ni && g_visc.d.nj == d.nj && g_visc.d.nk == d.nk) {
    Call c("cal_div_tensor_");
    T* u = c.out_t<T>(uk, d.n3, "uk");  // a recorded array among these runs the chain first
    T* v = c.out_t<T>(vk, d.n3, "vk");
    T* w = c.out_t<T>(wk, d.n3, "wk");
    if (ch.active) {
      ch.active = false;
      ch.flush = nullptr;
      ch.outs.clear();
      ch.srcs.clear();
      const ViscRecord& r = g_visc;
      sij_tau<T>(d, SijTau<T>{static_cast<const T*>(r.u), static_cast<const T*>(r.v),
                              static_cast<const T*>(r.w), static_cast<const T*>(r.rmu),
                              static_cast<T*>(r.s), static_cast<T*>(r.tau), T(r.dxinv),
                              T(r.dyinv), T(r.dzinv)});
      div_tensor<T>(d, dxinv, dyinv, dzinv, static_cast<const T*>(r.tau), u, v, w);
      c.finish();
      return;
    }
    c.finish();
  }
  Call c("cal_div_tensor_");
  const T* t = c.in_t<T>(tauk, 6 * (size_t)d.n3, "tauk");
  T* u = c.out_t<T>(uk, d.n3, "uk");
  T* v = c.out_t<T>(vk, d.n3, "vk");
  T* w = c.out_t<T>(wk, d.n3, "wk");
  div_tensor<T>(d, dxinv, dyinv, dzinv, t, u, v, w);
  c.finish();
}

template <class T>
void cal_srcu_t(const Dims d, int jen, double dt, double grvb, double grvp, const double* const* in,
                double* const* out) {
  Call c("cal_srcu_");
  const int n = d.n3;
  const T* p[23];
  int addz = 0;
  for (int m = 0; m < 23; m++) {
    // a force sum that is a deferred 0.0 + force: read the force instead
    const void* f = m >= 17 ? deferred_add_zero_source(in[m], n) : nullptr;
    if (f) {
      p[m] = static_cast<const T*>(f);
      addz |= 1 << (m - 17);
    } else {
      p[m] = c.in_t<T>(in[m], n, "cal_srcu_ input");
    }
  }
  SrcArgs<T> a;
  a.addz = addz;
  a.u = p[0]; a.v = p[1]; a.w = p[2]; a.rho = p[3]; a.rhon = p[4];
  a.adv_u = p[5]; a.adv_v = p[6]; a.adv_w = p[7]; a.adv_uo = p[8]; a.adv_vo = p[9]; a.adv_wo = p[10];
  a.prs_u = p[11]; a.prs_v = p[12]; a.prs_w = p[13]; a.vis_u = p[14]; a.vis_v = p[15]; a.vis_w = p[16];
  a.fst_u = p[17]; a.fst_v = p[18]; a.fst_w = p[19]; a.fst_un = p[20]; a.fst_vn = p[21];
  a.fst_wn = p[22];
  a.src_u = c.io_t<T>(out[0], n, "src_u");
  a.src_v = c.io_t<T>(out[1], n, "src_v");
  a.src_w = c.io_t<T>(out[2], n, "src_w");
  // registered sources are also stored pair by pair for solu_sor4_
  DenseSrc& ds = dense_src();
  ds.valid = false;
  bool reg = true;
  for (int m = 0; m < 3; m++) reg = reg && content_key(out[m], n) != 0;
  a.dense = reg ? static_cast<T*>(scratch(17, sizeof(T) * 12 * r00_slab(d))) : nullptr;
  // and, from registered densities, solu_sor4_'s face-density table
  const bool rreg = reg && content_key(in[3], n) != 0 && content_key(in[4], n) != 0;
  for (int m = 0; m < 3; m++)
    a.r00d[m] = rreg ? static_cast<T*>(scratch(10 + m, sizeof(T) * 4 * r00_slab(d))) : nullptr;
  srcu<T>(d, jen, 1.0 / dt, grvb, grvp, a);
  c.finish();
  if (reg) {
    ds.valid = true;
    for (int m = 0; m < 3; m++) {
      ds.host[m] = out[m];
      ds.key[m] = content_key(out[m], n);
    }
    ds.ni = d.ni;
    ds.nj = d.nj;
    ds.nk = d.nk;
    ds.esize = (int)sizeof(T);
    ds.buf = a.dense;
    ds.r00 = rreg;
    if (rreg) {
      ds.rho_host = in[3];
      ds.rhon_host = in[4];
      ds.rho_key = content_key(in[3], n);
      ds.rhon_key = content_key(in[4], n);
    }
  }
}

DenseSrc& dense_src() {
  static DenseSrc s;
  return s;
}

void dense_src_release() { dense_src() = DenseSrc{}; }

template <class T>
void cal_arith_coef_vis_t(const Dims d, double dxinv, double dyinv, double dzinv, const double* rmuk,
                          double* const* h) {
  Call c("cal_arith_coef_vis_");
  // a pending record for other output arrays is materialised before it is replaced
  DeferredCoef& dc = deferred_coef();
  if (dc.active) {
    // the same output arrays are about to be redefined: drop the pending
    // record; other outputs still owe their values
    bool same = true;
    for (int m = 0; m < 45 && same; m++) same = dc.out_host[m] == (const void*)h[m];
    if (!same) deferred_coef_flush();
    deferred_coef_drop();
  }
  bool resident = is_registered(rmuk, d.n3);
  for (int m = 0; m < 45 && resident; m++) resident = is_registered(h[m], d.n3);
  const T* r = c.in_t<T>(rmuk, d.n3, "rmuk");
  if (resident && !deferred_coef_disabled()) {
    dc.d = d;
    dc.esize = (int)sizeof(T);
    dc.dxinv = dxinv; dc.dyinv = dyinv; dc.dzinv = dzinv;
    dc.src_host = rmuk;
    dc.src_dev = r;
    for (int m = 0; m < 45; m++) {
      dc.out_host[m] = h[m];
      dc.out_dev[m] = c.out_t<T>(h[m], d.n3, "coefficient");
    }
    dc.active = true;
  } else {
    Coef45<T> a;
    for (int m = 0; m < 45; m++) a.p[m] = c.out_t<T>(h[m], d.n3, "coefficient");
    coef_vis<T>(d, dxinv, dyinv, dzinv, r, a);
  }
  c.finish();
}

template <class T>
void cpy_t(const Dims d, const double* q1, double* q2) {
  if (q1 == q2) return;
  if (deferred_copy(q1, q2, d.n3)) {
    copy_content_version(q1, q2, d.n3);
    return;
  }
  Call c("cpy_");
  const T* a = c.in_t<T>(q1, d.n3, "q1");
  T* b = c.out_t<T>(q2, d.n3, "q2");
  trace_event("array copy", sizeof(T) * (size_t)d.n3);
  FEL_CHECK(hipMemcpyAsync(b, a, sizeof(T) * (size_t)d.n3, hipMemcpyDeviceToDevice, stream()));
  c.finish();
  copy_content_version(q1, q2, d.n3);
}

template <class T>
void init_q_t(const Dims d, double* q1, double* q2, double* q3) {
  const size_t n = (size_t)d.n3;
  double* q[3] = {q1, q2, q3};
  for (int m = 0; m < 3; m++) {
    // device-resident: a deferred fill (the accumulation of sum_fst_ that
    // follows in the timestep usually makes it unnecessary)
    if (defer_zero(q[m], n)) continue;
    Call c("init_q_");
    T* p = c.all_t<T>(q[m], n, "q");
    FEL_CHECK(hipMemsetAsync(p, 0, sizeof(T) * n, stream()));
    c.finish();
  }
}

template <class T>
void sum_fst_t(const Dims d, const double* fst_u, const double* fst_v, const double* fst_w,
               double* sum_u, double* sum_v, double* sum_w) {
  const size_t n = (size_t)d.n3;
  // sums still holding a deferred zero fill become deferred 0.0 + fst
  // (srcs/sum_fst.f:28-30 on zeros), read as such by cal_srcu_
  const void* f[3] = {fst_u, fst_v, fst_w};
  const void* s[3] = {sum_u, sum_v, sum_w};
  if (defer_add_zero(s, f, 3, n)) return;
  Call c("sum_fst_");
  const T *fu = c.in_t<T>(fst_u, d.n3, "fst_u"), *fv = c.in_t<T>(fst_v, d.n3, "fst_v"),
          *fw = c.in_t<T>(fst_w, d.n3, "fst_w");
  T *su = c.io_t<T>(sum_u, d.n3, "sum_u"), *sv = c.io_t<T>(sum_v, d.n3, "sum_v"),
    *sw = c.io_t<T>(sum_w, d.n3, "sum_w");
  launch_flat(d.n3, [=] __device__(int m) {
    su[m] = su[m] + fu[m];
    sv[m] = sv[m] + fv[m];
    sw[m] = sw[m] + fw[m];
  });
  c.finish();
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/cal_arith_mean.f:1
void cal_arith_mean_(const int* ni, const int* nj, const int* nk, const double* phik,
                     const double* ql, const double* qg, double* qk) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) { cal_arith_mean_t<decltype(t)>(d, phik, *ql, *qg, qk); });
}

// srcs/calphat.f:1
void calphat_(const int* ni, const int* nj, const int* nk, const double* p, const double* po,
              double* phat) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) { calphat_t<decltype(t)>(d, p, po, phat); });
}

// srcs/cal_grad_prs.f:1-3
void cal_grad_prs_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                   const double* dyinv, const double* dzinv, const double* q, double* qx_u,
                   double* qy_v, double* qz_w) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) {
    cal_grad_prs_t<decltype(t)>(d, *dxinv, *dyinv, *dzinv, q, qx_u, qy_v, qz_w);
  });
}

// srcs/calsij.f:5-6
void calsij_(const int* ni, const int* nj, const int* nk, const double* dxinv, const double* dyinv,
             const double* dzinv, const double* uk, const double* vk, const double* wk, double* sk) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) { calsij_t<decltype(t)>(d, *dxinv, *dyinv, *dzinv, uk, vk, wk, sk); });
}

// srcs/cal_arith_tau.f:7
void cal_arith_tau_(const int* ni, const int* nj, const int* nk, const double* s, const double* rmu,
                    double* tau) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) { cal_arith_tau_t<decltype(t)>(d, s, rmu, tau); });
}

// srcs/cal_div_tensor.f:1-2
void cal_div_tensor_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                     const double* dyinv, const double* dzinv, const double* tauk, double* uk,
                     double* vk, double* wk) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) {
    cal_div_tensor_t<decltype(t)>(d, *dxinv, *dyinv, *dzinv, tauk, uk, vk, wk);
  });
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
  const double* in[23] = {u,      v,      w,      rho,    rhon,   adv_u,  adv_v,  adv_w,
                          adv_uo, adv_vo, adv_wo, prs_u,  prs_v,  prs_w,  vis_u,  vis_v,
                          vis_w,  fst_u,  fst_v,  fst_w,  fst_un, fst_vn, fst_wn};
  double* out[3] = {src_u, src_v, src_w};
  const int jen = (nID[kYPlus] < 0) ? *nj - 1 : *nj;
  with_flow([&](auto t) { cal_srcu_t<decltype(t)>(d, jen, *dt, *grvb, *grvp, in, out); });
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
  double* h[45] = {a0,  a1,  a2,  a3,  a4,  a5,  a6,  a7,  a8,  a9,  a10, a11, a12, a13, a14,
                   a15, a16, a17, a18, a19, a20, a21, a22, a23, a24, a25, a26, a27, a28, a29,
                   a30, a31, a32, a33, a34, a35, a36, a37, a38, a39, a40, a41, a42, a43, a44};
  with_flow([&](auto t) { cal_arith_coef_vis_t<decltype(t)>(d, *dxinv, *dyinv, *dzinv, rmuk, h); });
}

// srcs/cpy.f:1 (either precision class: copies the device copy as it is held)
void cpy_(const int* ni, const int* nj, const int* nk, const double* q1, double* q2) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_elem(elem_size(q1), [&](auto t) { cpy_t<decltype(t)>(d, q1, q2); });
}

// srcs/init_q.f:1
void init_q_(const int* ni, const int* nj, const int* nk, double* q1, double* q2, double* q3) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_elem(elem_size(q1), [&](auto t) { init_q_t<decltype(t)>(d, q1, q2, q3); });
}

// srcs/sum_fst.f:1-3
void sum_fst_(const int* ni, const int* nj, const int* nk, const double* fst_u, const double* fst_v,
              const double* fst_w, double* sum_u, double* sum_v, double* sum_w) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) { sum_fst_t<decltype(t)>(d, fst_u, fst_v, fst_w, sum_u, sum_v, sum_w); });
}

}  // extern "C"
