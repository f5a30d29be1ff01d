// felgpu_visc.hpp -- viscous stencil coefficients as device functions.
//
// srcs/cal_arith_coef_vis.f:184-407 builds 45 coefficient arrays from the
// viscosity; the same expressions are evaluated here per cell, so the
// coefficient kernel and the Gauss-Seidel kernel (which recomputes them on
// the fly instead of streaming 15 arrays per sweep) produce identical bits.
#pragma once

#include "felgpu.hpp"

namespace fel {

constexpr int kYMinus = 1, kYPlus = 4;  // nID(Y_MINUS), nID(Y_PLUS) as 0-based (srcs/param.h)

// T is the element type of the flow fields: double (FP64 mode) or float
// (FP32 mode); with T = double every expression is the FP64 one.
template <class T>
struct Coef45 {
  T* p[45];
};

template <class T>
struct SrcArgs {
  const T *u, *v, *w, *rho, *rhon;
  const T *adv_u, *adv_v, *adv_w, *adv_uo, *adv_vo, *adv_wo;
  const T *prs_u, *prs_v, *prs_w, *vis_u, *vis_v, *vis_w;
  const T *fst_u, *fst_v, *fst_w, *fst_un, *fst_vn, *fst_wn;
  T *src_u, *src_v, *src_w;
  // bit m set: the m-th force sum (fst_u, fst_v, fst_w, fst_un, fst_vn,
  // fst_wn) points at the force it is 0.0 + of, element by element
  int addz;
  // when not null: the sources also stored pair by pair for the
  // Gauss-Seidel colour pairs (component q, pair p at
  // dense + (4 q + p) r00_slab + r00_index)
  T* dense;
  // when not null: the face densities of the three components, pair by pair
  // (component q, pair p at r00d[q] + p r00_slab + r00_index), the table
  // solu_sor4_ forms for its solve
  T* r00d[3];
};

// Layout of values stored pair by pair for the Gauss-Seidel colour pairs
// (srcs/solu_sor4.f:94-120).  Dense index of a pair's cell: ii-th cell of its
// parity in row jj of its parity, plane k (every cell (i,j,k) of a pair has
// ii = (i-1)/2, jj = (j-1)/2).  One pair's slab holds
// nk * ceil(nj/2) * ceil(ni/2) values.
__host__ __device__ __forceinline__ int r00_index(const Dims& d, int ii, int jj, int k) {
  const int hx = (d.ni + 1) / 2, hy = (d.nj + 1) / 2;
  return ((k - 1) * hy + jj) * hx + ii;
}
__host__ __device__ __forceinline__ size_t r00_slab(const Dims& d) {
  return (size_t)d.nk * ((d.nj + 1) / 2) * ((d.ni + 1) / 2);
}

// The colour pair (0..3) that cell (i,j,k) belongs to.
__host__ __device__ __forceinline__ int pair_of(int i, int j, int k) {
  const int ci = 2 - (i & 1), cj = 2 - (j & 1), ck = 2 - (k & 1);  // 1 odd, 2 even
  // pairs (111,222), (112,221), (121,212), (211,122): one parity differs
  // from the other two, or none does
  if (ci == cj && cj == ck) return 0;
  if (ci == cj) return 1;
  if (ci == ck) return 2;
  return 3;
}

// The pair-by-pair copy of the last cal_srcu_ sources: valid for a
// solu_sor4_ whose source arrays still carry the content numbers recorded
// here (the arrays are registered and unchanged since).
struct DenseSrc {
  bool valid = false;
  const void* host[3] = {nullptr, nullptr, nullptr};
  uint64_t key[3] = {0, 0, 0};
  int ni = 0, nj = 0, nk = 0, esize = 0;
  void* buf = nullptr;
  // the face-density table made from these densities (valid while they
  // carry these content numbers)
  bool r00 = false;
  const void* rho_host = nullptr;
  const void* rhon_host = nullptr;
  uint64_t rho_key = 0, rhon_key = 0;
};
DenseSrc& dense_src();

template <class T>
struct VisScales {
  T dx2, dy2, dz2, dxy, dxz, dyz;
};

// The products are formed in FP64 and rounded once to T.
template <class T>
inline VisScales<T> vis_scales(double dxinv, double dyinv, double dzinv) {
  return VisScales<T>{T(dxinv * dxinv), T(dyinv * dyinv), T(dzinv * dzinv),
                      T(dxinv * dyinv), T(dxinv * dzinv), T(dyinv * dzinv)};
}

// Coefficients of the u equation in the reference's argument order:
// w,e,s,n,b,t,p, v_ws,v_es,v_wn,v_en, w_wb,w_eb,w_wt,w_et.
template <class T>
__device__ __forceinline__ void coef_u(const Dims& d, const VisScales<T>& s, const T* r, int c,
                                       T* v) {
  const int X = 1, Y = d.sx, Z = d.sx * d.sy;
  const T xm = r[c], xp = r[c + X];
  const T ym = T(0.25) * (+r[c - Y] + r[c + X - Y] + r[c] + r[c + X]);
  const T yp = T(0.25) * (+r[c] + r[c + X] + r[c + Y] + r[c + X + Y]);
  const T zm = T(0.25) * (+r[c - Z] + r[c + X - Z] + r[c] + r[c + X]);
  const T zp = T(0.25) * (+r[c] + r[c + X] + r[c + Z] + r[c + X + Z]);
  const T b = s.dz2 * zm, ss = s.dy2 * ym, w = s.dx2 * xm * T(2.0), e = s.dx2 * xp * T(2.0);
  const T n = s.dy2 * yp, t = s.dz2 * zp;
  v[0] = w; v[1] = e; v[2] = ss; v[3] = n; v[4] = b; v[5] = t;
  v[6] = +b + ss + w + e + n + t;
  v[7] = s.dxy * ym; v[8] = -(s.dxy * ym); v[9] = -(s.dxy * yp); v[10] = s.dxy * yp;
  v[11] = s.dxz * zm; v[12] = -(s.dxz * zm); v[13] = -(s.dxz * zp); v[14] = s.dxz * zp;
}

// v equation: w,e,s,n,b,t,p, u_sw,u_nw,u_se,u_ne, w_sb,w_nb,w_st,w_nt.
template <class T>
__device__ __forceinline__ void coef_v(const Dims& d, const VisScales<T>& s, const T* r, int c,
                                       T* v) {
  const int X = 1, Y = d.sx, Z = d.sx * d.sy;
  const T xm = T(0.25) * (+r[c - X] + r[c] + r[c - X + Y] + r[c + Y]);
  const T xp = T(0.25) * (+r[c] + r[c + X] + r[c + Y] + r[c + X + Y]);
  const T ym = r[c], yp = r[c + Y];
  const T zm = T(0.25) * (+r[c - Z] + r[c + Y - Z] + r[c] + r[c + Y]);
  const T zp = T(0.25) * (+r[c] + r[c + Y] + r[c + Z] + r[c + Y + Z]);
  const T b = s.dz2 * zm, ss = s.dy2 * ym * T(2.0), w = s.dx2 * xm, e = s.dx2 * xp;
  const T n = s.dy2 * yp * T(2.0), t = s.dz2 * zp;
  v[0] = w; v[1] = e; v[2] = ss; v[3] = n; v[4] = b; v[5] = t;
  v[6] = +b + ss + w + e + n + t;
  v[7] = s.dxy * xm; v[8] = -(s.dxy * xm); v[9] = -(s.dxy * xp); v[10] = s.dxy * xp;
  v[11] = s.dyz * zm; v[12] = -(s.dyz * zm); v[13] = -(s.dyz * zp); v[14] = s.dyz * zp;
}

// w equation: w,e,s,n,b,t,p, u_bw,u_tw,u_be,u_te, v_bs,v_ts,v_bn,v_tn.
template <class T>
__device__ __forceinline__ void coef_w(const Dims& d, const VisScales<T>& s, const T* r, int c,
                                       T* v) {
  const int X = 1, Y = d.sx, Z = d.sx * d.sy;
  const T xm = T(0.25) * (+r[c - X] + r[c] + r[c - X + Z] + r[c + Z]);
  const T xp = T(0.25) * (+r[c] + r[c + X] + r[c + Z] + r[c + X + Z]);
  const T ym = T(0.25) * (+r[c - Y] + r[c] + r[c - Y + Z] + r[c + Z]);
  const T yp = T(0.25) * (+r[c] + r[c + Y] + r[c + Z] + r[c + Y + Z]);
  const T zm = r[c], zp = r[c + Z];
  const T b = s.dz2 * zm * T(2.0), ss = s.dy2 * ym, w = s.dx2 * xm, e = s.dx2 * xp;
  const T n = s.dy2 * yp, t = s.dz2 * zp * T(2.0);
  v[0] = w; v[1] = e; v[2] = ss; v[3] = n; v[4] = b; v[5] = t;
  v[6] = +b + ss + w + e + n + t;
  v[7] = s.dxz * xm; v[8] = -(s.dxz * xm); v[9] = -(s.dxz * xp); v[10] = s.dxz * xp;
  v[11] = s.dyz * ym; v[12] = -(s.dyz * ym); v[13] = -(s.dyz * yp); v[14] = s.dyz * yp;
}

}  // namespace fel
