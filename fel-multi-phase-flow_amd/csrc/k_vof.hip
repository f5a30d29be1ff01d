// k_vof.hip -- volume-fraction transport: vertex gradients, MTHINC face
// fluxes, conservative update, truncation and global mass correction.
//
// Reference: srcs/cal_grad_p2a.f, srcs/solphi_mthinc1.f (+ mthinc_x.h,
// mthinc_tanh.h), srcs/solphi_mthinc2.f, srcs/solphi_mthinc3.f,
// srcs/summation.f.  All FP64, expressions evaluated in the reference's
// order (the library is compiled without contraction).
#include <cmath>

#include <cstdlib>
#include "felgpu.hpp"

namespace fel {

// Per-j inverse spacing: the reference's dyinv_array is uniform
// (srcs/main.f90:215-224); a non-uniform table is uploaded.
struct DyTable {
  double uniform;
  const double* table;  // device, indexed [dis+2], or null when uniform
  __device__ __forceinline__ double at(int dis) const { return table ? table[dis + 2] : uniform; }
};

static DyTable dy_table(const double* host_arr, int count, int slot) {
  DyTable t{host_arr[2], nullptr};
  bool uni = true;
  for (int m = 0; m < count; m++)
    if (host_arr[m] != host_arr[0]) uni = false;
  if (uni) {
    t.uniform = host_arr[0];
    return t;
  }
  double* dev = static_cast<double*>(scratch(slot, sizeof(double) * count));
  FEL_CHECK(hipMemcpyAsync(dev, host_arr, sizeof(double) * count, hipMemcpyHostToDevice, stream()));
  FEL_CHECK(hipStreamSynchronize(stream()));
  t.table = dev;
  return t;
}

// Vertex gradient, srcs/cal_grad_p2a.f:30-71, box [-2,n+2]^3, formed in the
// phase type P (FP64) from the volume fraction and stored in the gradient
// element type T.
template <class T, class P = PT<T>>
void grad_p2a(const Dims d, int ID, double dxinv, DyTable dyt, double dzinv, const P* q, T* qx,
              T* qy, T* qz) {
  launch_box(Box{-2, d.ni + 2, -2, d.nj + 2, -2, d.nk + 2}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const int X = 1, Y = d.sx, Z = d.sx * d.sy;
    const P q000 = q[c], qp00 = q[c + X], q0p0 = q[c + Y], qpp0 = q[c + X + Y];
    const P q00p = q[c + Z], qp0p = q[c + X + Z], q0pp = q[c + Y + Z], qppp = q[c + X + Y + Z];
    const P gx = -q000 + qp00 - q0p0 + qpp0 - q00p + qp0p - q0pp + qppp;
    const P gy = -q000 - qp00 + q0p0 + qpp0 - q00p - qp0p + q0pp + qppp;
    const P gz = -q000 - qp00 - q0p0 - qpp0 + q00p + qp0p + q0pp + qppp;
    const double dyinv = dyt.at(j + d.nj * ID);
    qx[c] = T(gx * 0.25 * dxinv);
    qy[c] = T(gy * 0.25 * dyinv);
    qz[c] = T(gz * 0.25 * dzinv);
  });
}

// Cell normal from the 8 surrounding vertex gradients
// (srcs/solphi_mthinc1.f:81-107).
template <class T>
__device__ __forceinline__ T vavg8(const T* p, int c, int X, int Y, int Z) {
  return (+p[c - X - Y - Z] + p[c - Y - Z] + p[c - X - Z] + p[c - Z] + p[c - X - Y] + p[c - Y] +
          p[c - X] + p[c]) *
         T(0.125);
}

// srcs/mthinc_x.h + srcs/mthinc_tanh.h: 8-point Gauss sum of the tanh
// profile; xs receives the 8 projected Gauss coordinates.
template <class T>
__device__ __forceinline__ T tanh_sum(T nx, T ny, T nz, T xg1, T xg2, T yg1, T yg2, T zg1, T zg2,
                                      T wght, T bet, T ddd, T xs[8]) {
  const T x1 = nx * xg1, x2 = nx * xg2, y1 = ny * yg1, y2 = ny * yg2;
  const T z1 = nz * zg1, z2 = nz * zg2;
  xs[0] = x1 + y1 + z1;
  xs[1] = x2 + y1 + z1;
  xs[2] = x1 + y2 + z1;
  xs[3] = x2 + y2 + z1;
  xs[4] = x1 + y1 + z2;
  xs[5] = x2 + y1 + z2;
  xs[6] = x1 + y2 + z2;
  xs[7] = x2 + y2 + z2;
  T s = wght / (T(1.0) + exp(-(T(2.0) * bet * (xs[0] + ddd))));
#pragma unroll
  for (int m = 1; m < 8; m++) s = s + wght / (T(1.0) + exp(-(T(2.0) * bet * (xs[m] + ddd))));
  return s;
}

struct MthincConst {
  double bet, gp1, gp2, dcellinv, eps_mthinc, eps_trunc;
};

// The constants in the element type of the MTHINC arithmetic.
template <class T>
struct MthincC {
  T bet, gp1, gp2, dcellinv, eps_mthinc, eps_trunc;
  explicit MthincC(const MthincConst& m)
      : bet(T(m.bet)), gp1(T(m.gp1)), gp2(T(m.gp2)), dcellinv(T(m.dcellinv)),
        eps_mthinc(T(m.eps_mthinc)), eps_trunc(T(m.eps_trunc)) {}
};

static MthincConst mthinc_const(double dxinv, double dyinv, double dzinv, double bet) {
  // srcs/solphi_mthinc1.f:33-35,55-60, evaluated on the host with the same
  // libm as the CPU reference so that the branch thresholds are identical.
  MthincConst m;
  const double dx = 1.0 / dxinv, dy = 1.0 / dyinv, dz = 1.0 / dzinv;
  m.bet = bet;
  m.gp1 = -(0.5 / std::sqrt(3.0));
  m.gp2 = 0.5 / std::sqrt(3.0);
  m.dcellinv = 1.0 / std::sqrt((dx * dx + dy * dy + dz * dz) / 3.0);
  const double ch = std::cosh(-(bet * 1.5));
  m.eps_mthinc = bet / (2.0 * (ch * ch));
  m.eps_trunc = 0.5 * (std::tanh(-(bet * 2.5)) + 1.0);
  return m;
}

// srcs/solphi_mthinc1.f:38-288.  One thread per cell of [-1,n+1]^3.  A cell
// of [0,n+1]^3 writes its outflow faces (right face when the face velocity
// is positive, left face when negative), exactly the faces the reference
// cell writes with a non-zero value.  Every face of [-1,n+1]^3 that no cell
// writes with a non-zero value -- zero velocity, or an upwind cell outside
// [0,n+1]^3 -- is zeroed by the cell to its left, which replaces the
// reference's zero-fill pass (solphi_mthinc1.f:44-52).  Each face therefore
// has exactly one writer.
// T: element type of the gradients, the velocities and of the MTHINC
// arithmetic (double, or float in the FP32 mode); the fluxes are stored in
// the phase type P (FP64).
template <class T, class P = PT<T>>
void mthinc_flux(const Dims d, double dxinv_, double dyinv_, double dzinv_, double dt_,
                 MthincConst mc64, const P* phi, const T* phix, const T* phiy, const T* phiz,
                 const T* u, const T* v, const T* w, P* flx, P* fly, P* flz) {
  const int n1 = d.ni + 1, n2 = d.nj + 1, n3 = d.nk + 1;
  const MthincC<T> mc(mc64);
  const T dxinv = T(dxinv_), dyinv = T(dyinv_), dzinv = T(dzinv_), dt = T(dt_);
  launch_box(Box{-1, d.ni + 1, -1, d.nj + 1, -1, d.nk + 1}, [=] __device__(int i, int j, int k) {
    const int X = 1, Y = d.sx, Z = d.sx * d.sy;
    const int c = d.ix(i, j, k);
    const bool in_i = i >= 0, in_j = j >= 0, in_k = k >= 0;      // cell inside [0,n+1]^3 ?
    const bool self = in_i && in_j && in_k;
    const T up = u[c], vp = v[c], wp = w[c];
    const T zero = T(0.0);
    // zero the right faces that have no outflow writer
    {
      const bool rx = (i + 1 <= n1) && in_j && in_k;  // right neighbour in [0,n+1]^3
      if (up == zero || (up > zero && !self) || (up < zero && !rx) || !in_j || !in_k) flx[c] = P(0.0);
      const bool ry = (j + 1 <= n2) && in_i && in_k;
      if (vp == zero || (vp > zero && !self) || (vp < zero && !ry) || !in_i || !in_k) fly[c] = P(0.0);
      const bool rz = (k + 1 <= n3) && in_i && in_j;
      if (wp == zero || (wp > zero && !self) || (wp < zero && !rz) || !in_i || !in_j) flz[c] = P(0.0);
    }
    if (!self) return;
    const T um = u[c - X], vm = v[c - Y], wm = w[c - Z];
    T rnx = vavg8(phix, c, X, Y, Z);
    T rny = vavg8(phiy, c, X, Y, Z);
    T rnz = vavg8(phiz, c, X, Y, Z);
    const T q00 = sqrt(rnx * rnx + rny * rny + rnz * rnz);
    if (q00 <= mc.eps_mthinc * mc.dcellinv) {
      // first-order upwind with the truncated cell value (solphi_mthinc1.f:110-135)
      T phi00 = T(phi[c]);
      if (phi00 <= mc.eps_trunc) phi00 = zero;
      if (T(1.0) - phi00 <= mc.eps_trunc) phi00 = T(1.0);
      if (um < zero) flx[c - X] = um * phi00 * dt * dxinv;
      if (up > zero) flx[c] = up * phi00 * dt * dxinv;
      if (vm < zero) fly[c - Y] = vm * phi00 * dt * dyinv;
      if (vp > zero) fly[c] = vp * phi00 * dt * dyinv;
      if (wm < zero) flz[c - Z] = wm * phi00 * dt * dzinv;
      if (wp > zero) flz[c] = wp * phi00 * dt * dzinv;
      return;
    }
    const T coef = T(1.0) / q00;
    rnx = rnx * coef;
    rny = rny * coef;
    rnz = rnz * coef;
    const T bet = mc.bet, gp1 = mc.gp1, gp2 = mc.gp2;
    const T phic = T(phi[c]);
    const T h = T(0.5), e8 = T(0.125);
    T xs[8];
    T ddd = zero;
    // Newton iterations for the interface offset (solphi_mthinc1.f:161-176)
    for (int inr = 0; inr < 10; inr++) {
      const T wt = tanh_sum(rnx, rny, rnz, gp1, gp2, gp1, gp2, gp1, gp2, e8, bet, ddd, xs);
      const T f = wt - phic;
      T ch = cosh(bet * (xs[0] + ddd));
      T s = e8 / (ch * ch);
#pragma unroll
      for (int m = 1; m < 8; m++) {
        ch = cosh(bet * (xs[m] + ddd));
        s = s + e8 / (ch * ch);
      }
      const T df = h * bet * s;
      ddd = ddd - f * df / (df * df + guard<T>(1.0e-99));
    }
    // fluxes through the outflow faces (solphi_mthinc1.f:182-282)
    if (um < zero) {
      const T vd = um * dt * dxinv, st = -h, en = -h - vd;
      const T g1 = (st + en) * h + (-st + en) * gp1, g2 = (st + en) * h + (-st + en) * gp2;
      flx[c - X] = tanh_sum(rnx, rny, rnz, g1, g2, gp1, gp2, gp1, gp2, e8 * vd, bet, ddd, xs);
    }
    if (up > zero) {
      const T vd = up * dt * dxinv, st = h - vd, en = h;
      const T g1 = (st + en) * h + (-st + en) * gp1, g2 = (st + en) * h + (-st + en) * gp2;
      flx[c] = tanh_sum(rnx, rny, rnz, g1, g2, gp1, gp2, gp1, gp2, e8 * vd, bet, ddd, xs);
    }
    if (vm < zero) {
      const T vd = vm * dt * dyinv, st = -h, en = -h - vd;
      const T g1 = (st + en) * h + (-st + en) * gp1, g2 = (st + en) * h + (-st + en) * gp2;
      fly[c - Y] = tanh_sum(rnx, rny, rnz, gp1, gp2, g1, g2, gp1, gp2, e8 * vd, bet, ddd, xs);
    }
    if (vp > zero) {
      const T vd = vp * dt * dyinv, st = h - vd, en = h;
      const T g1 = (st + en) * h + (-st + en) * gp1, g2 = (st + en) * h + (-st + en) * gp2;
      fly[c] = tanh_sum(rnx, rny, rnz, gp1, gp2, g1, g2, gp1, gp2, e8 * vd, bet, ddd, xs);
    }
    if (wm < zero) {
      const T vd = wm * dt * dzinv, st = -h, en = -h - vd;
      const T g1 = (st + en) * h + (-st + en) * gp1, g2 = (st + en) * h + (-st + en) * gp2;
      flz[c - Z] = tanh_sum(rnx, rny, rnz, gp1, gp2, gp1, gp2, g1, g2, e8 * vd, bet, ddd, xs);
    }
    if (wp > zero) {
      const T vd = wp * dt * dzinv, st = h - vd, en = h;
      const T g1 = (st + en) * h + (-st + en) * gp1, g2 = (st + en) * h + (-st + en) * gp2;
      flz[c] = tanh_sum(rnx, rny, rnz, gp1, gp2, gp1, gp2, g1, g2, e8 * vd, bet, ddd, xs);
    }
  });
}

// srcs/solphi_mthinc2.f:27-44.
// The update is FP64: velocity differences are formed from the velocities
// (FP32 values are exact in FP64) and the fluxes are FP64.
template <class T, class P = PT<T>>
void mthinc_update(const Dims d, double dxinv, double dyinv, double dzinv, double dt, const T* u,
                   const T* v, const T* w, const P* flx, const P* fly, const P* flz, const P* phi,
                   P* phin) {
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    const int X = 1, Y = d.sx, Z = d.sx * d.sy;
    const int c = d.ix(i, j, k);
    const P div00 = +(-P(u[c - X]) + P(u[c])) * dxinv + (-P(v[c - Y]) + P(v[c])) * dyinv +
                    (-P(w[c - Z]) + P(w[c])) * dzinv;
    P p = phi[c] * (1.0 + dt * div00) + flx[c - X] - flx[c] + fly[c - Y] - fly[c] + flz[c - Z] -
          flz[c];
    p = fmin(p, P(1.0));
    p = fmax(p, P(0.0));
    phin[c] = p;
  });
}

// srcs/solphi_mthinc3.f:43-159: truncation on [0,n+1]^3 (clamped in the
// interior in the same pass, the reference's second loop), the three global
// sums, and the mass correction.  The offset stays on the device.
// The truncation pass of srcs/solphi_mthinc3.f:43-83 on [0,n+1]^3.  One
// thread per (i,j) column, marching over kMassChunk planes: the 8-point
// averages of the gradients take their k-1 plane from the values the column
// loaded for the previous plane, so each plane of phix, phiy, phiz is
// fetched once per column.  The sums are vavg8's, in its order.
// (measured at 512x256x128 in FP64: 177 us with 8 planes, 183 with 2, 176
// with 4, 180 with 16)
constexpr int kMassChunk = 8;

template <class T, class P>
__global__ void __launch_bounds__(256) mthinc_trunc_kernel(Dims d, MthincC<T> mc, double eps_trunc,
                                                           const T* __restrict__ phix,
                                                           const T* __restrict__ phiy,
                                                           const T* __restrict__ phiz,
                                                           P* __restrict__ phin, int gx, int gy) {
  const int L = (int)blockIdx.x;
  const int bx = L % gx, tt = L / gx, by = tt % gy, bz = tt / gy;
  const int i = bx * 64 + (int)threadIdx.x, j = by * 4 + (int)threadIdx.y;
  if (i > d.ni + 1 || j > d.nj + 1) return;
  const int k0 = bz * kMassChunk, k1 = min(d.nk + 1, k0 + kMassChunk - 1);
  const int X = 1, Y = d.sx, Z = d.sx * d.sy;
  const T* g[3] = {phix, phiy, phiz};
  // q[m][0..3]: array m at (i-1,j-1), (i,j-1), (i-1,j), (i,j) of plane k-1
  T q[3][4];
  int c = d.ix(i, j, k0);
#pragma unroll
  for (int m = 0; m < 3; m++) {
    const int b = c - Z;
    q[m][0] = g[m][b - X - Y];
    q[m][1] = g[m][b - Y];
    q[m][2] = g[m][b - X];
    q[m][3] = g[m][b];
  }
  const bool interior = i >= 1 && i <= d.ni && j >= 1 && j <= d.nj;
  for (int k = k0; k <= k1; k++, c += Z) {
    T r[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const T a0 = g[m][c - X - Y], a1 = g[m][c - Y], a2 = g[m][c - X], a3 = g[m][c];
      r[m] = (+q[m][0] + q[m][1] + q[m][2] + q[m][3] + a0 + a1 + a2 + a3) * T(0.125);
      q[m][0] = a0;
      q[m][1] = a1;
      q[m][2] = a2;
      q[m][3] = a3;
    }
    const T q00 = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    P p = phin[c];
    const P p0 = p;
    if (q00 <= mc.eps_mthinc * mc.dcellinv) {
      if (p <= eps_trunc) p = P(0.0);
      if (1.0 - p <= eps_trunc) p = P(1.0);
    }
    if (interior && k >= 1 && k <= d.nk) {
      p = fmin(p, P(1.0));
      p = fmax(p, P(0.0));
    }
    // a truncated or clamped cell also loses its derivative
    if (p != p0 || (is_dual<P> && !(p0 > 0.0 && p0 < 1.0))) phin[c] = p;
  }
}

template <class T, class P = PT<T>>
void mthinc_mass(const Dims d, MthincConst mc64, const T* phix, const T* phiy, const T* phiz,
                 const P* phi, P* phin) {
  const MthincC<T> mc(mc64);
  {
    const int gx = (d.ni + 2 + 63) / 64, gy = (d.nj + 2 + 3) / 4;
    const int gz = (d.nk + 2 + kMassChunk - 1) / kMassChunk;
    mthinc_trunc_kernel<T, P><<<dim3(gx * gy * gz), dim3(64, 4, 1), 0, stream()>>>(
        d, mc, mc64.eps_trunc, phix, phiy, phiz, phin, gx, gy);
    FEL_CHECK(hipGetLastError());
  }
  // the sums and the offset, in P, stay on the device (slot 0 holds 4 P's)
  double* sums = dev_scalars(0);
  interior_reduce<3, P>(
      d,
      [=] __device__(int i, int j, int k, P* v) {
        const int c = d.ix(i, j, k);
        const P p = phin[c];
        v[0] = phi[c];
        v[1] = p;
        v[2] = p * p * (1.0 - p);
      },
      [=] __device__(const P* t, double* out) {
        P* o = reinterpret_cast<P*>(out);
        o[0] = t[0];
        o[1] = t[1];
        o[2] = t[2];
        o[3] = (t[0] - t[1]) / (t[2] + 1.0e-99);
      },
      sums);
  const P* offp = reinterpret_cast<const P*>(sums) + 3;
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const P offset = *offp;
    P p = phin[c];
    p = p + offset * (p * p) * (1.0 - p);
    p = fmin(p, P(1.0));
    phin[c] = fmax(p, P(0.0));
  });
}

// srcs/summation.f:14-50 for nbub slices: slice 0 = sum of slices 1..nbub
// over the padded image, clamped to [0,1] in the interior.  One launch per
// added slice (first: sum = slice 1; last: clamp), which keeps the
// reference's summation order; the slices need not be adjacent on the device.
template <class P>
void summation(const Dims d, P* dst, const P* src, bool first, bool last) {
  launch_flat(d.n3, [=] __device__(int m) {
    P s = first ? src[m] : dst[m] + src[m];
    const int i = m % d.sx - 2, r = m / d.sx, j = r % d.sy - 2, k = r / d.sy - 2;
    if (last && i >= 1 && i <= d.ni && j >= 1 && j <= d.nj && k >= 1 && k <= d.nk)
      s = fmin(fmax(s, P(0.0)), P(1.0));
    dst[m] = s;
  });
}

template <class T>
void cal_grad_p2a_t(const Dims d, int ID, int svall, double dxinv, const double* dyinv_array,
                    double dzinv, const double* q, double* qx, double* qy, double* qz) {
  const size_t n = (size_t)d.n3;
  const void* outs[3] = {qx, qy, qz};
  uint64_t sh[3];
  const MemoKey k = gradient_key(kOpGradP2a, (int)sizeof(T), d, ID, svall, dxinv, dyinv_array,
                                 dzinv, q, outs, sh);
  // the arrays, or parked buffers, still hold this very result
  if (memo_reuse(k, {{qx, n}, {qy, n}, {qz, n}}, false)) {
    for (int m = 0; m < 3; m++) keep_shell(outs[m], n, sh[m]);
    return;
  }
  memo_park(d, {{qx, n}, {qy, n}, {qz, n}}, true);
  Call c("cal_grad_p2a_");
  const PT<T>* dq = c.in_t<PT<T>>(q, d.n3, "q");
  T* dx = c.io_t<T>(qx, d.n3, "qx");
  T* dy = c.io_t<T>(qy, d.n3, "qy");
  T* dz = c.io_t<T>(qz, d.n3, "qz");
  grad_p2a<T>(d, ID, dxinv, dy_table(dyinv_array, svall + 6, 1), dzinv, dq, dx, dy, dz);
  c.finish();
  memo_record(k, {{qx, n}, {qy, n}, {qz, n}});
  for (int m = 0; m < 3; m++) keep_shell(outs[m], n, sh[m]);  // written: the box only
}

template <class T>
void solphi_mthinc1_t(const Dims d, double dxinv, double dyinv, double dzinv, double dt, double bet,
                      const double* phi, const double* phix, const double* phiy, const double* phiz,
                      const double* u, const double* v, const double* w, double* flphix,
                      double* flphiy, double* flphiz) {
  using P = PT<T>;
  Call c("solphi_mthinc1_");
  const P* dphi = c.in_t<P>(phi, d.n3, "phi");
  const T* gx = c.in_t<T>(phix, d.n3, "phix");
  const T* gy = c.in_t<T>(phiy, d.n3, "phiy");
  const T* gz = c.in_t<T>(phiz, d.n3, "phiz");
  const T* du = c.in_t<T>(u, d.n3, "u");
  const T* dv = c.in_t<T>(v, d.n3, "v");
  const T* dw = c.in_t<T>(w, d.n3, "w");
  P* fx = c.io_t<P>(flphix, d.n3, "flphix");
  P* fy = c.io_t<P>(flphiy, d.n3, "flphiy");
  P* fz = c.io_t<P>(flphiz, d.n3, "flphiz");
  mthinc_flux<T>(d, dxinv, dyinv, dzinv, dt, mthinc_const(dxinv, dyinv, dzinv, bet), dphi, gx, gy,
                 gz, du, dv, dw, fx, fy, fz);
  c.finish();
}

template <class T>
void solphi_mthinc2_t(const Dims d, double dxinv, double dyinv, double dzinv, double dt,
                      const double* u, const double* v, const double* w, const double* flphix,
                      const double* flphiy, const double* flphiz, const double* phi, double* phin) {
  using P = PT<T>;
  Call c("solphi_mthinc2_");
  const T* du = c.in_t<T>(u, d.n3, "u");
  const T* dv = c.in_t<T>(v, d.n3, "v");
  const T* dw = c.in_t<T>(w, d.n3, "w");
  const P* fx = c.in_t<P>(flphix, d.n3, "flphix");
  const P* fy = c.in_t<P>(flphiy, d.n3, "flphiy");
  const P* fz = c.in_t<P>(flphiz, d.n3, "flphiz");
  const P* dphi = c.in_t<P>(phi, d.n3, "phi");
  P* dphin = c.interior_t<P>(phin, d, "phin");  // written on the interior only
  mthinc_update<T>(d, dxinv, dyinv, dzinv, dt, du, dv, dw, fx, fy, fz, dphi, dphin);
  c.finish();
}

template <class T>
void solphi_mthinc3_t(const Dims d, double dxinv, double dyinv, double dzinv, double bet,
                      const double* phix, const double* phiy, const double* phiz, const double* phi,
                      double* phin) {
  using P = PT<T>;
  Call c("solphi_mthinc3_");
  const T* gx = c.in_t<T>(phix, d.n3, "phix");
  const T* gy = c.in_t<T>(phiy, d.n3, "phiy");
  const T* gz = c.in_t<T>(phiz, d.n3, "phiz");
  const P* dphi = c.in_t<P>(phi, d.n3, "phi");
  P* dphin = c.io_t<P>(phin, d.n3, "phin");
  mthinc_mass<T>(d, mthinc_const(dxinv, dyinv, dzinv, bet), gx, gy, gz, dphi, dphin);
  c.finish();
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/cal_grad_p2a.f:1-2
void cal_grad_p2a_(const int* ID, const int* svall, const int* ni, const int* nj, const int* nk,
                   const double* dxinv, const double* dyinv_array, const double* dzinv,
                   const double* q, double* qx, double* qy, double* qz) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) {
    cal_grad_p2a_t<decltype(t)>(d, *ID, *svall, *dxinv, dyinv_array, *dzinv, q, qx, qy, qz);
  });
}

// srcs/solphi_mthinc1.f:1-6
void solphi_mthinc1_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                     const double* dyinv, const double* dzinv, const double* dt,
                     const double* bet_mthinc, const double* phi, const double* phix,
                     const double* phiy, const double* phiz, const double* u, const double* v,
                     const double* w, double* flphix, double* flphiy, double* flphiz) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) {
    solphi_mthinc1_t<decltype(t)>(d, *dxinv, *dyinv, *dzinv, *dt, *bet_mthinc, phi, phix, phiy,
                                  phiz, u, v, w, flphix, flphiy, flphiz);
  });
}

// srcs/solphi_mthinc2.f:1-2
void solphi_mthinc2_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                     const double* dyinv, const double* dzinv, const double* dt, const double* u,
                     const double* v, const double* w, const double* flphix, const double* flphiy,
                     const double* flphiz, const double* phi, double* phin) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) {
    solphi_mthinc2_t<decltype(t)>(d, *dxinv, *dyinv, *dzinv, *dt, u, v, w, flphix, flphiy, flphiz,
                                  phi, phin);
  });
}

// srcs/solphi_mthinc3.f:1-5 (ipara must be 0: one GPU, no MPI)
void solphi_mthinc3_(const int* ipara, const int* ni, const int* nj, const int* nk,
                     const double* dxinv, const double* dyinv, const double* dzinv,
                     const double* bet_mthinc, const double* phix, const double* phiy,
                     const double* phiz, const double* phi, double* phin) {
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: solphi_mthinc3_: ipara=%d, only the single-domain case (0) is supported\n", *ipara);
    std::abort();
  }
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) {
    solphi_mthinc3_t<decltype(t)>(d, *dxinv, *dyinv, *dzinv, *bet_mthinc, phix, phiy, phiz, phi,
                                  phin);
  });
}

// srcs/summation.f:1
void summation_(const int* ni, const int* nj, const int* nk, double* qk, const int* nbub) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) {
    using P = PT<decltype(t)>;
    Call c("summation_");
    // slice by slice: a caller may register each slice as a range of its own
    const size_t n = (size_t)d.n3;
    const int nb = *nbub;
    if (nb < 1) {
      std::fprintf(stderr, "felgpu: summation_: nbub = %d, at least one slice is summed\n", nb);
      std::abort();
    }
    std::vector<const P*> src(nb);
    for (int l = 1; l <= nb; l++) src[l - 1] = c.in_t<P>(qk + (size_t)l * n, n, "qk(l)");
    P* dst = c.all_t<P>(qk, n, "qk(0)");
    for (int l = 1; l <= nb; l++) summation<P>(d, dst, src[l - 1], l == 1, l == nb);
    c.finish();
  });
}

}  // extern "C"
