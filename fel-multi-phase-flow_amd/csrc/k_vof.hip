// k_vof.hip -- volume-fraction transport: vertex gradients, MTHINC face
// fluxes, conservative update, truncation and global mass correction.
//
// Reference: srcs/cal_grad_p2a.f, srcs/solphi_mthinc1.f (+ mthinc_x.h,
// mthinc_tanh.h), srcs/solphi_mthinc2.f, srcs/solphi_mthinc3.f,
// srcs/summation.f.  All FP64, expressions evaluated in the reference's
// order (the library is compiled without contraction).
#include <cmath>

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

// Vertex gradient, srcs/cal_grad_p2a.f:30-71, box [-2,n+2]^3.
void grad_p2a(const Dims d, int ID, double dxinv, DyTable dyt, double dzinv, const double* q,
              double* qx, double* qy, double* qz) {
  launch_box(Box{-2, d.ni + 2, -2, d.nj + 2, -2, d.nk + 2}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const int X = 1, Y = d.sx, Z = d.sx * d.sy;
    const double q000 = q[c], qp00 = q[c + X], q0p0 = q[c + Y], qpp0 = q[c + X + Y];
    const double q00p = q[c + Z], qp0p = q[c + X + Z], q0pp = q[c + Y + Z], qppp = q[c + X + Y + Z];
    const double gx = -q000 + qp00 - q0p0 + qpp0 - q00p + qp0p - q0pp + qppp;
    const double gy = -q000 - qp00 + q0p0 + qpp0 - q00p - qp0p + q0pp + qppp;
    const double gz = -q000 - qp00 - q0p0 - qpp0 + q00p + qp0p + q0pp + qppp;
    const double dyinv = dyt.at(j + d.nj * ID);
    qx[c] = gx * 0.25 * dxinv;
    qy[c] = gy * 0.25 * dyinv;
    qz[c] = gz * 0.25 * dzinv;
  });
}

// Cell normal from the 8 surrounding vertex gradients
// (srcs/solphi_mthinc1.f:81-107).
__device__ __forceinline__ double vavg8(const double* p, int c, int X, int Y, int Z) {
  return (+p[c - X - Y - Z] + p[c - Y - Z] + p[c - X - Z] + p[c - Z] + p[c - X - Y] + p[c - Y] +
          p[c - X] + p[c]) *
         0.125;
}

// srcs/mthinc_x.h + srcs/mthinc_tanh.h: 8-point Gauss sum of the tanh
// profile; xs receives the 8 projected Gauss coordinates.
__device__ __forceinline__ double tanh_sum(double nx, double ny, double nz, double xg1, double xg2,
                                           double yg1, double yg2, double zg1, double zg2,
                                           double wght, double bet, double ddd, double xs[8]) {
  const double x1 = nx * xg1, x2 = nx * xg2, y1 = ny * yg1, y2 = ny * yg2;
  const double z1 = nz * zg1, z2 = nz * zg2;
  xs[0] = x1 + y1 + z1;
  xs[1] = x2 + y1 + z1;
  xs[2] = x1 + y2 + z1;
  xs[3] = x2 + y2 + z1;
  xs[4] = x1 + y1 + z2;
  xs[5] = x2 + y1 + z2;
  xs[6] = x1 + y2 + z2;
  xs[7] = x2 + y2 + z2;
  double s = wght / (1.0 + exp(-(2.0 * bet * (xs[0] + ddd))));
#pragma unroll
  for (int m = 1; m < 8; m++) s = s + wght / (1.0 + exp(-(2.0 * bet * (xs[m] + ddd))));
  return s;
}

struct MthincConst {
  double bet, gp1, gp2, dcellinv, eps_mthinc, eps_trunc;
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
void mthinc_flux(const Dims d, double dxinv, double dyinv, double dzinv, double dt, MthincConst mc,
                 const double* phi, const double* phix, const double* phiy, const double* phiz,
                 const double* u, const double* v, const double* w, double* flx, double* fly,
                 double* flz) {
  const int n1 = d.ni + 1, n2 = d.nj + 1, n3 = d.nk + 1;
  launch_box(Box{-1, d.ni + 1, -1, d.nj + 1, -1, d.nk + 1}, [=] __device__(int i, int j, int k) {
    const int X = 1, Y = d.sx, Z = d.sx * d.sy;
    const int c = d.ix(i, j, k);
    const bool in_i = i >= 0, in_j = j >= 0, in_k = k >= 0;      // cell inside [0,n+1]^3 ?
    const bool self = in_i && in_j && in_k;
    const double up = u[c], vp = v[c], wp = w[c];
    // zero the right faces that have no outflow writer
    {
      const bool rx = (i + 1 <= n1) && in_j && in_k;  // right neighbour in [0,n+1]^3
      if (up == 0.0 || (up > 0.0 && !self) || (up < 0.0 && !rx) || !in_j || !in_k) flx[c] = 0.0;
      const bool ry = (j + 1 <= n2) && in_i && in_k;
      if (vp == 0.0 || (vp > 0.0 && !self) || (vp < 0.0 && !ry) || !in_i || !in_k) fly[c] = 0.0;
      const bool rz = (k + 1 <= n3) && in_i && in_j;
      if (wp == 0.0 || (wp > 0.0 && !self) || (wp < 0.0 && !rz) || !in_i || !in_j) flz[c] = 0.0;
    }
    if (!self) return;
    const double um = u[c - X], vm = v[c - Y], wm = w[c - Z];
    double rnx = vavg8(phix, c, X, Y, Z);
    double rny = vavg8(phiy, c, X, Y, Z);
    double rnz = vavg8(phiz, c, X, Y, Z);
    const double q00 = sqrt(rnx * rnx + rny * rny + rnz * rnz);
    if (q00 <= mc.eps_mthinc * mc.dcellinv) {
      // first-order upwind with the truncated cell value (solphi_mthinc1.f:110-135)
      double phi00 = phi[c];
      if (phi00 <= mc.eps_trunc) phi00 = 0.0;
      if (1.0 - phi00 <= mc.eps_trunc) phi00 = 1.0;
      if (um < 0.0) flx[c - X] = um * phi00 * dt * dxinv;
      if (up > 0.0) flx[c] = up * phi00 * dt * dxinv;
      if (vm < 0.0) fly[c - Y] = vm * phi00 * dt * dyinv;
      if (vp > 0.0) fly[c] = vp * phi00 * dt * dyinv;
      if (wm < 0.0) flz[c - Z] = wm * phi00 * dt * dzinv;
      if (wp > 0.0) flz[c] = wp * phi00 * dt * dzinv;
      return;
    }
    const double coef = 1.0 / q00;
    rnx = rnx * coef;
    rny = rny * coef;
    rnz = rnz * coef;
    const double bet = mc.bet, gp1 = mc.gp1, gp2 = mc.gp2;
    const double phic = phi[c];
    double xs[8];
    double ddd = 0.0;
    // Newton iterations for the interface offset (solphi_mthinc1.f:161-176)
    for (int inr = 0; inr < 10; inr++) {
      const double wt = tanh_sum(rnx, rny, rnz, gp1, gp2, gp1, gp2, gp1, gp2, 0.125, bet, ddd, xs);
      const double f = wt - phic;
      double ch = cosh(bet * (xs[0] + ddd));
      double s = 0.125 / (ch * ch);
#pragma unroll
      for (int m = 1; m < 8; m++) {
        ch = cosh(bet * (xs[m] + ddd));
        s = s + 0.125 / (ch * ch);
      }
      const double df = 0.5 * bet * s;
      ddd = ddd - f * df / (df * df + 1.0e-99);
    }
    // fluxes through the outflow faces (solphi_mthinc1.f:182-282)
    if (um < 0.0) {
      const double vd = um * dt * dxinv, st = -0.5, en = -0.5 - vd;
      const double g1 = (st + en) * 0.5 + (-st + en) * gp1, g2 = (st + en) * 0.5 + (-st + en) * gp2;
      flx[c - X] = tanh_sum(rnx, rny, rnz, g1, g2, gp1, gp2, gp1, gp2, 0.125 * vd, bet, ddd, xs);
    }
    if (up > 0.0) {
      const double vd = up * dt * dxinv, st = 0.5 - vd, en = 0.5;
      const double g1 = (st + en) * 0.5 + (-st + en) * gp1, g2 = (st + en) * 0.5 + (-st + en) * gp2;
      flx[c] = tanh_sum(rnx, rny, rnz, g1, g2, gp1, gp2, gp1, gp2, 0.125 * vd, bet, ddd, xs);
    }
    if (vm < 0.0) {
      const double vd = vm * dt * dyinv, st = -0.5, en = -0.5 - vd;
      const double g1 = (st + en) * 0.5 + (-st + en) * gp1, g2 = (st + en) * 0.5 + (-st + en) * gp2;
      fly[c - Y] = tanh_sum(rnx, rny, rnz, gp1, gp2, g1, g2, gp1, gp2, 0.125 * vd, bet, ddd, xs);
    }
    if (vp > 0.0) {
      const double vd = vp * dt * dyinv, st = 0.5 - vd, en = 0.5;
      const double g1 = (st + en) * 0.5 + (-st + en) * gp1, g2 = (st + en) * 0.5 + (-st + en) * gp2;
      fly[c] = tanh_sum(rnx, rny, rnz, gp1, gp2, g1, g2, gp1, gp2, 0.125 * vd, bet, ddd, xs);
    }
    if (wm < 0.0) {
      const double vd = wm * dt * dzinv, st = -0.5, en = -0.5 - vd;
      const double g1 = (st + en) * 0.5 + (-st + en) * gp1, g2 = (st + en) * 0.5 + (-st + en) * gp2;
      flz[c - Z] = tanh_sum(rnx, rny, rnz, gp1, gp2, gp1, gp2, g1, g2, 0.125 * vd, bet, ddd, xs);
    }
    if (wp > 0.0) {
      const double vd = wp * dt * dzinv, st = 0.5 - vd, en = 0.5;
      const double g1 = (st + en) * 0.5 + (-st + en) * gp1, g2 = (st + en) * 0.5 + (-st + en) * gp2;
      flz[c] = tanh_sum(rnx, rny, rnz, gp1, gp2, gp1, gp2, g1, g2, 0.125 * vd, bet, ddd, xs);
    }
  });
}

// srcs/solphi_mthinc2.f:27-44.
void mthinc_update(const Dims d, double dxinv, double dyinv, double dzinv, double dt,
                   const double* u, const double* v, const double* w, const double* flx,
                   const double* fly, const double* flz, const double* phi, double* phin) {
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    const int X = 1, Y = d.sx, Z = d.sx * d.sy;
    const int c = d.ix(i, j, k);
    const double div00 = +(-u[c - X] + u[c]) * dxinv + (-v[c - Y] + v[c]) * dyinv +
                         (-w[c - Z] + w[c]) * dzinv;
    double p = phi[c] * (1.0 + dt * div00) + flx[c - X] - flx[c] + fly[c - Y] - fly[c] +
               flz[c - Z] - flz[c];
    p = fmin(p, 1.0);
    p = fmax(p, 0.0);
    phin[c] = p;
  });
}

// srcs/solphi_mthinc3.f:43-159: truncation on [0,n+1]^3 (clamped in the
// interior in the same pass, the reference's second loop), the three global
// sums, and the mass correction.  The offset stays on the device.
void mthinc_mass(const Dims d, MthincConst mc, const double* phix, const double* phiy,
                 const double* phiz, const double* phi, double* phin) {
  launch_box(Box{0, d.ni + 1, 0, d.nj + 1, 0, d.nk + 1}, [=] __device__(int i, int j, int k) {
    const int X = 1, Y = d.sx, Z = d.sx * d.sy;
    const int c = d.ix(i, j, k);
    const double rnx = vavg8(phix, c, X, Y, Z);
    const double rny = vavg8(phiy, c, X, Y, Z);
    const double rnz = vavg8(phiz, c, X, Y, Z);
    const double q00 = sqrt(rnx * rnx + rny * rny + rnz * rnz);
    double p = phin[c];
    const double p0 = p;
    if (q00 <= mc.eps_mthinc * mc.dcellinv) {
      if (p <= mc.eps_trunc) p = 0.0;
      if (1.0 - p <= mc.eps_trunc) p = 1.0;
    }
    if (i >= 1 && i <= d.ni && j >= 1 && j <= d.nj && k >= 1 && k <= d.nk) {
      p = fmin(p, 1.0);
      p = fmax(p, 0.0);
    }
    if (p != p0) phin[c] = p;
  });
  double* sums = dev_scalars(0);
  interior_reduce<3>(
      d,
      [=] __device__(int i, int j, int k, double* v) {
        const int c = d.ix(i, j, k);
        const double p = phin[c];
        v[0] = phi[c];
        v[1] = p;
        v[2] = p * p * (1.0 - p);
      },
      [=] __device__(const double* t, double* out) {
        out[0] = t[0];
        out[1] = t[1];
        out[2] = t[2];
        out[3] = (t[0] - t[1]) / (t[2] + 1.0e-99);
      },
      sums);
  const double* offp = sums + 3;
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const double offset = *offp;
    double p = phin[c];
    p = p + offset * (p * p) * (1.0 - p);
    p = fmin(p, 1.0);
    phin[c] = fmax(p, 0.0);
  });
}

// srcs/summation.f:14-50 for nbub slices: slice 0 = sum of slices 1..nbub
// over the padded image, clamped to [0,1] in the interior.
void summation(const Dims d, double* qk, int nbub) {
  const int n3 = d.n3;
  launch_flat(n3, [=] __device__(int m) {
    double s = qk[n3 + m];
    for (int l = 2; l <= nbub; l++) s = s + qk[(size_t)l * n3 + m];
    const int i = m % d.sx - 2, r = m / d.sx, j = r % d.sy - 2, k = r / d.sy - 2;
    if (i >= 1 && i <= d.ni && j >= 1 && j <= d.nj && k >= 1 && k <= d.nk) s = fmin(fmax(s, 0.0), 1.0);
    qk[m] = s;
  });
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/cal_grad_p2a.f:1-2
void cal_grad_p2a_(const int* ID, const int* svall, const int* ni, const int* nj, const int* nk,
                   const double* dxinv, const double* dyinv_array, const double* dzinv,
                   const double* q, double* qx, double* qy, double* qz) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_grad_p2a_");
  const double* dq = c.in(q, d.n3, "q");
  double* dx = c.io(qx, d.n3, "qx");
  double* dy = c.io(qy, d.n3, "qy");
  double* dz = c.io(qz, d.n3, "qz");
  grad_p2a(d, *ID, *dxinv, dy_table(dyinv_array, *svall + 6, 1), *dzinv, dq, dx, dy, dz);
  c.finish();
}

// srcs/solphi_mthinc1.f:1-6
void solphi_mthinc1_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                     const double* dyinv, const double* dzinv, const double* dt,
                     const double* bet_mthinc, const double* phi, const double* phix,
                     const double* phiy, const double* phiz, const double* u, const double* v,
                     const double* w, double* flphix, double* flphiy, double* flphiz) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("solphi_mthinc1_");
  const double* dphi = c.in(phi, d.n3, "phi");
  const double* gx = c.in(phix, d.n3, "phix");
  const double* gy = c.in(phiy, d.n3, "phiy");
  const double* gz = c.in(phiz, d.n3, "phiz");
  const double* du = c.in(u, d.n3, "u");
  const double* dv = c.in(v, d.n3, "v");
  const double* dw = c.in(w, d.n3, "w");
  double* fx = c.io(flphix, d.n3, "flphix");
  double* fy = c.io(flphiy, d.n3, "flphiy");
  double* fz = c.io(flphiz, d.n3, "flphiz");
  prof_begin(1);
  mthinc_flux(d, *dxinv, *dyinv, *dzinv, *dt, mthinc_const(*dxinv, *dyinv, *dzinv, *bet_mthinc), dphi,
              gx, gy, gz, du, dv, dw, fx, fy, fz);
  prof_end(1);
  c.finish();
}

// srcs/solphi_mthinc2.f:1-2
void solphi_mthinc2_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                     const double* dyinv, const double* dzinv, const double* dt, const double* u,
                     const double* v, const double* w, const double* flphix, const double* flphiy,
                     const double* flphiz, const double* phi, double* phin) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("solphi_mthinc2_");
  const double* du = c.in(u, d.n3);
  const double* dv = c.in(v, d.n3);
  const double* dw = c.in(w, d.n3);
  const double* fx = c.in(flphix, d.n3);
  const double* fy = c.in(flphiy, d.n3);
  const double* fz = c.in(flphiz, d.n3);
  const double* dphi = c.in(phi, d.n3);
  double* dphin = c.io(phin, d.n3);
  mthinc_update(d, *dxinv, *dyinv, *dzinv, *dt, du, dv, dw, fx, fy, fz, dphi, dphin);
  c.finish();
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
  Call c("solphi_mthinc3_");
  const double* gx = c.in(phix, d.n3);
  const double* gy = c.in(phiy, d.n3);
  const double* gz = c.in(phiz, d.n3);
  const double* dphi = c.in(phi, d.n3);
  double* dphin = c.io(phin, d.n3);
  mthinc_mass(d, mthinc_const(*dxinv, *dyinv, *dzinv, *bet_mthinc), gx, gy, gz, dphi, dphin);
  c.finish();
}

// srcs/summation.f:1
void summation_(const int* ni, const int* nj, const int* nk, double* qk, const int* nbub) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("summation_");
  double* dq = c.io(qk, (size_t)d.n3 * (*nbub + 1));
  summation(d, dq, *nbub);
  c.finish();
}

}  // extern "C"
