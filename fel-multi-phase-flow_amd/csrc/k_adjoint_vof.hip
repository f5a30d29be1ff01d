// k_adjoint_vof.hip -- adjoints (reverse mode) of the volume-fraction
// transport: vertex gradients, MTHINC face fluxes, conservative update,
// truncation and mass correction, summation of the slices.
//
// Conventions as for the other <name>_b_ entry points: the forward's
// argument list with every active array replaced by its adjoint; where the
// routine is not linear, the primal array (the forward's input) comes first,
// followed by its adjoint.  Branches -- flat or interface cell, upwind
// direction, truncation, the clamps (which pass the derivative where the
// value lies in [0,1], as fmin/fmax select their first operand on ties) --
// are those of the forward at its inputs.  The MTHINC offset d solves
// sum_g s(n.x_g + d)/8 = phi (s(z) = 1/(1 + exp(-2 bet z)), 8 Gauss points)
// by 10 Newton steps in the forward; its derivative is taken by the
// implicit-function theorem at the Newton iterate: dd = (dphi - dn.G_n)/G_d
// with G_d = sum s'/8, G_n = sum s' x_g/8.  FP64 mode only.
//
// Each kernel gathers what one element receives.  Where a cell's results
// go to elements shared with other cells (the eight vertex gradients of a
// cell normal, the two faces of a divergence), a first launch stores
// per-cell values in scratch and a second gathers them.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "felgpu.hpp"

namespace fel {

// Constants of srcs/solphi_mthinc1.f:33-35,55-60, on the host with the
// libm the forward uses (the branch thresholds are the forward's bits).
struct VofConst {
  double bet, gp1, gp2, dcellinv, eps_mthinc, eps_trunc;
};

static VofConst vof_const(double dxinv, double dyinv, double dzinv, double bet) {
  VofConst m;
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

// Mean of the 8 vertex values around cell c (srcs/solphi_mthinc1.f:81-107).
__device__ __forceinline__ double vavg(const double* p, int c, int X, int Y, int Z) {
  return (+p[c - X - Y - Z] + p[c - Y - Z] + p[c - X - Z] + p[c - Z] + p[c - X - Y] + p[c - Y] +
          p[c - X] + p[c]) *
         0.125;
}

__device__ __forceinline__ bool is_flat(const VofConst& m, const double* gx, const double* gy,
                                        const double* gz, int c, int X, int Y, int Z) {
  const double rx = vavg(gx, c, X, Y, Z), ry = vavg(gy, c, X, Y, Z), rz = vavg(gz, c, X, Y, Z);
  return sqrt(rx * rx + ry * ry + rz * rz) <= m.eps_mthinc * m.dcellinv;
}

__device__ __forceinline__ double sig2(double bet, double z) {
  return 1.0 / (1.0 + exp(-(2.0 * bet * z)));
}

// ---------------------------------------------------------------------------
// srcs/cal_grad_p2a.f:30-71 transposed (linear): q_b at each cell gathers
// the signed gradient adjoints of the 8 vertex cells [-2,n+2]^3 that read
// it; then the gradient adjoints are cleared there.
// ---------------------------------------------------------------------------
static void grad_p2a_adj(const Dims d, int ID, double dxinv, const double* dyt, double dzinv,
                         double* qb, double* gx, double* gy, double* gz) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int y = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    double t = 0.0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
      const int a = m & 1, b = (m >> 1) & 1, e = (m >> 2) & 1;  // y = c + (a,b,e)
      const int ci = i - a, cj = j - b, ck = k - e;
      if (ci < -2 || ci > ni + 2 || cj < -2 || cj > nj + 2 || ck < -2 || ck > nk + 2) continue;
      const int c = y - a * X - b * Y - e * Z;
      const double dyinv = dyt[cj + nj * ID + 2];
      t = t + (a ? 1.0 : -1.0) * (gx[c] * 0.25 * dxinv) + (b ? 1.0 : -1.0) * (gy[c] * 0.25 * dyinv) +
          (e ? 1.0 : -1.0) * (gz[c] * 0.25 * dzinv);
    }
    qb[y] = qb[y] + t;
  });
  launch_box(Box{-2, ni + 2, -2, nj + 2, -2, nk + 2}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    gx[c] = 0.0;
    gy[c] = 0.0;
    gz[c] = 0.0;
  });
}

// The 8 Gauss-point coordinates n.x_g of the cell's sub-cell (the points
// (gp1|gp2) in each direction), in the order of srcs/mthinc_x.h.
__device__ __forceinline__ void gauss_proj(const double* n, double gp1, double gp2, double* xs) {
  const double x1 = n[0] * gp1, x2 = n[0] * gp2, y1 = n[1] * gp1, y2 = n[1] * gp2;
  const double z1 = n[2] * gp1, z2 = n[2] * gp2;
  xs[0] = x1 + y1 + z1;
  xs[1] = x2 + y1 + z1;
  xs[2] = x1 + y2 + z1;
  xs[3] = x2 + y2 + z1;
  xs[4] = x1 + y1 + z2;
  xs[5] = x2 + y1 + z2;
  xs[6] = x1 + y2 + z2;
  xs[7] = x2 + y2 + z2;
}

// The forward's Newton iterate for the offset d (srcs/solphi_mthinc1.f:
// 161-176): 10 steps from 0 in the forward's operation order.
__device__ __noinline__ double newton_offset(const double* n, double phic, const VofConst m) {
  double xs[8];
  gauss_proj(n, m.gp1, m.gp2, xs);
  const double bet = m.bet;
  double ddd = 0.0;
  for (int inr = 0; inr < 10; inr++) {
    double wt = 0.125 / (1.0 + exp(-(2.0 * bet * (xs[0] + ddd))));
    for (int q = 1; q < 8; q++) wt = wt + 0.125 / (1.0 + exp(-(2.0 * bet * (xs[q] + ddd))));
    const double f = wt - phic;
    double ch = cosh(bet * (xs[0] + ddd));
    double s = 0.125 / (ch * ch);
    for (int q = 1; q < 8; q++) {
      ch = cosh(bet * (xs[q] + ddd));
      s = s + 0.125 / (ch * ch);
    }
    const double df = 0.5 * bet * s;
    ddd = ddd - f * df / (df * df + 1.0e-99);
  }
  return ddd;
}

// ---------------------------------------------------------------------------
// srcs/solphi_mthinc1.f:38-288 reversed.
// ---------------------------------------------------------------------------
struct Mthinc1Adj {
  VofConst m;
  double dinv[3], dt;
  const double *phi, *g[3], *vel[3];
  double *phib, *velb[3], *flb[3];
  double* rb[3];  // per-cell adjoint of the unnormalised normal r (scratch)
};

// Does the cell at c write face (c - E on its minus side, or c on its plus
// side) of axis a?  Flat cells write the face when the velocity points out
// or is zero, interface cells when it points out.
__device__ __forceinline__ bool writes_face(bool flat, bool plus, double vface) {
  if (plus) return flat ? vface >= 0.0 : vface > 0.0;
  return flat ? vface <= 0.0 : vface < 0.0;
}

// First launch, one thread per cell of [0,n+1]^3: the adjoints of the faces
// the cell owns go to the face velocities and to phi (flat cells) or to d
// and n; d's to phi and n; n's to r (stored).  A face both neighbours write
// (flat cells, zero velocity) belongs to the one with the larger index, the
// last writer of the forward's loop.
static void mthinc1_cells(const Dims d, const Mthinc1Adj A) {
  launch_box(Box{0, d.ni + 1, 0, d.nj + 1, 0, d.nk + 1}, [=] __device__(int i, int j, int k) {
    const int X = 1, Y = d.sx, Z = d.sx * d.sy;
    const int E[3] = {X, Y, Z};
    const int c = d.ix(i, j, k);
    const VofConst& m = A.m;
    double r[3];
#pragma unroll
    for (int q = 0; q < 3; q++) r[q] = vavg(A.g[q], c, X, Y, Z);
    const double q00 = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const bool flat = q00 <= m.eps_mthinc * m.dcellinv;
    const int pos[3] = {i, j, k};
    const int top[3] = {d.ni + 1, d.nj + 1, d.nk + 1};
    // which of the cell's 6 faces it owns
    bool own[3][2];
#pragma unroll 1
    for (int a = 0; a < 3; a++) {
      const double vm = A.vel[a][c - E[a]], vp = A.vel[a][c];
      own[a][0] = writes_face(flat, false, vm);
      bool pw = writes_face(flat, true, vp);
      // the right neighbour, when it also writes this face, owns it
      if (pw && flat && vp == 0.0 && pos[a] + 1 <= top[a] &&
          is_flat(m, A.g[0], A.g[1], A.g[2], c + E[a], X, Y, Z))
        pw = false;
      own[a][1] = pw;
    }
    if (flat) {
      double phi00 = A.phi[c], dphi = 1.0;
      if (phi00 <= m.eps_trunc) { phi00 = 0.0; dphi = 0.0; }
      if (1.0 - phi00 <= m.eps_trunc) { phi00 = 1.0; dphi = 0.0; }
      double pb = 0.0;
#pragma unroll 1
      for (int a = 0; a < 3; a++)
#pragma unroll 1
        for (int side = 0; side < 2; side++) {
          if (!own[a][side]) continue;
          const int fc = side ? c : c - E[a];
          const double fbv = A.flb[a][fc], vf = A.vel[a][fc];
          A.velb[a][fc] = A.velb[a][fc] + fbv * phi00 * A.dt * A.dinv[a];
          pb = pb + fbv * vf * A.dt * A.dinv[a] * dphi;
        }
      A.phib[c] = A.phib[c] + pb;
#pragma unroll
      for (int q = 0; q < 3; q++) A.rb[q][c] = 0.0;
      return;
    }
    const double coef = 1.0 / q00;
    const double n[3] = {r[0] * coef, r[1] * coef, r[2] * coef};
    const double gp[2] = {m.gp1, m.gp2};
    const double bet = m.bet;
    // the forward's Newton iterate (10 steps from 0)
    const double ddd = newton_offset(n, A.phi[c], m);
    double Gd = 0.0, Gn[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const double Xg[3] = {gp[q & 1], gp[(q >> 1) & 1], gp[(q >> 2) & 1]};
      const double sg = sig2(bet, n[0] * Xg[0] + n[1] * Xg[1] + n[2] * Xg[2] + ddd);
      const double sp = 0.125 * 2.0 * bet * sg * (1.0 - sg);
      Gd = Gd + sp;
#pragma unroll
      for (int a = 0; a < 3; a++) Gn[a] = Gn[a] + sp * Xg[a];
    }
    double nb[3] = {0.0, 0.0, 0.0}, db = 0.0;
#pragma unroll 1
    for (int a = 0; a < 3; a++)
#pragma unroll 1
      for (int side = 0; side < 2; side++) {
        if (!own[a][side]) continue;
        const int fc = side ? c : c - E[a];
        const double fbv = A.flb[a][fc];
        const double vd = A.vel[a][fc] * A.dt * A.dinv[a];
        // sub-cell limits st, en and their derivatives in vd
        const double st = side ? 0.5 - vd : -0.5, en = side ? 0.5 : -0.5 - vd;
        const double dst = side ? -1.0 : 0.0, den = side ? 0.0 : -1.0;
        const double g[2] = {(st + en) * 0.5 + (-st + en) * m.gp1, (st + en) * 0.5 + (-st + en) * m.gp2};
        const double dg[2] = {(dst + den) * 0.5 + (-dst + den) * m.gp1,
                              (dst + den) * 0.5 + (-dst + den) * m.gp2};
        const double W = 0.125 * vd;
        double S = 0.0, Sp = 0.0, Spx[3] = {0.0, 0.0, 0.0}, Svd = 0.0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int ch[3] = {q & 1, (q >> 1) & 1, (q >> 2) & 1};
          double Xg[3];
#pragma unroll
          for (int b = 0; b < 3; b++) Xg[b] = b == a ? g[ch[b]] : gp[ch[b]];
          const double sg = sig2(bet, n[0] * Xg[0] + n[1] * Xg[1] + n[2] * Xg[2] + ddd);
          const double sp = 2.0 * bet * sg * (1.0 - sg);
          S = S + sg;
          Sp = Sp + sp;
#pragma unroll
          for (int b = 0; b < 3; b++) Spx[b] = Spx[b] + sp * Xg[b];
          Svd = Svd + sp * n[a] * dg[ch[a]];
        }
        A.velb[a][fc] = A.velb[a][fc] + fbv * (0.125 * S + W * Svd) * A.dt * A.dinv[a];
#pragma unroll
        for (int b = 0; b < 3; b++) nb[b] = nb[b] + fbv * W * Spx[b];
        db = db + fbv * W * Sp;
      }
    A.phib[c] = A.phib[c] + db / Gd;
#pragma unroll
    for (int b = 0; b < 3; b++) nb[b] = nb[b] - db * Gn[b] / Gd;
    const double ndn = n[0] * nb[0] + n[1] * nb[1] + n[2] * nb[2];
#pragma unroll
    for (int b = 0; b < 3; b++) A.rb[b][c] = (nb[b] - ndn * n[b]) * coef;
  });
}

// Second launch: each vertex gradient (box [-1,n+1]^3) gathers a eighth of
// r's adjoint of the 8 cells of [0,n+1]^3 whose normal averages it; third:
// the flux adjoints are cleared on [-1,n+1]^3.
static void mthinc1_vertices(const Dims d, const Mthinc1Adj A, double* gbx, double* gby,
                             double* gbz) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  launch_box(Box{-1, ni + 1, -1, nj + 1, -1, nk + 1}, [=] __device__(int i, int j, int k) {
    const int y = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    double t[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < 8; m++) {
      const int a = m & 1, b = (m >> 1) & 1, e = (m >> 2) & 1;  // cell = y + (a,b,e)
      const int ci = i + a, cj = j + b, ck = k + e;
      if (ci < 0 || ci > ni + 1 || cj < 0 || cj > nj + 1 || ck < 0 || ck > nk + 1) continue;
      const int c = y + a * X + b * Y + e * Z;
#pragma unroll
      for (int q = 0; q < 3; q++) t[q] = t[q] + A.rb[q][c];
    }
    gbx[y] = gbx[y] + t[0] * 0.125;
    gby[y] = gby[y] + t[1] * 0.125;
    gbz[y] = gbz[y] + t[2] * 0.125;
  });
  launch_box(Box{-1, ni + 1, -1, nj + 1, -1, nk + 1}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
#pragma unroll
    for (int q = 0; q < 3; q++) A.flb[q][c] = 0.0;
  });
}

// ---------------------------------------------------------------------------
// srcs/solphi_mthinc2.f:30-41 reversed: g = phin_b where the unclamped value
// lies in [0,1]; phi_b += g (1 + dt div), and, through the two faces of each
// direction, u_b += (h(c) - h(c+x)) dxinv with h = g phi dt and flux_b +=
// g(c+x) - g(c) (first launch: g, h per interior cell; second: gathers).
// ---------------------------------------------------------------------------
struct Mthinc2Adj {
  double dinv[3], dt;
  const double *vel[3], *fl[3], *phi;
  double *velb[3], *flb[3], *phib, *phinb;
  double *G, *H;  // scratch
};

static void mthinc2_adj(const Dims d, const Mthinc2Adj A) {
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    const double *u = A.vel[0], *v = A.vel[1], *w = A.vel[2];
    const double div00 = +(-u[c - X] + u[c]) * A.dinv[0] + (-v[c - Y] + v[c]) * A.dinv[1] +
                         (-w[c - Z] + w[c]) * A.dinv[2];
    const double p = A.phi[c] * (1.0 + A.dt * div00) + A.fl[0][c - X] - A.fl[0][c] +
                     A.fl[1][c - Y] - A.fl[1][c] + A.fl[2][c - Z] - A.fl[2][c];
    const double g = (p <= 1.0 && p >= 0.0) ? A.phinb[c] : 0.0;
    A.phinb[c] = 0.0;
    A.phib[c] = A.phib[c] + g * (1.0 + A.dt * div00);
    A.G[c] = g;
    A.H[c] = g * A.phi[c] * A.dt;
  });
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  launch_box(Box{0, ni, 0, nj, 0, nk}, [=] __device__(int i, int j, int k) {
    const int y = d.ix(i, j, k);
    const int E[3] = {1, d.sx, d.sx * d.sy};
    const int pos[3] = {i, j, k}, top[3] = {ni, nj, nk};
    const bool in0 = i >= 1 && j >= 1 && k >= 1;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      // the cell y + e_a is interior when y's other indices are and y_a < n_a
      bool inp = pos[a] + 1 <= top[a];
#pragma unroll
      for (int b = 0; b < 3; b++)
        if (b != a) inp = inp && pos[b] >= 1;
      const double g0 = in0 ? A.G[y] : 0.0, h0 = in0 ? A.H[y] : 0.0;
      const double g1 = inp ? A.G[y + E[a]] : 0.0, h1 = inp ? A.H[y + E[a]] : 0.0;
      A.velb[a][y] = A.velb[a][y] + (h0 - h1) * A.dinv[a];
      A.flb[a][y] = A.flb[a][y] + (g1 - g0);
    }
  });
}

// ---------------------------------------------------------------------------
// srcs/solphi_mthinc3.f:43-160 reversed: the forward is replayed (flat-cell
// truncation on [0,n+1]^3, clamp, the three interior sums and the offset),
// the adjoint of the offset is summed, and each cell's adjoint is formed.
// ---------------------------------------------------------------------------
static void mthinc3_adj(const Dims d, const VofConst m, const double* gx, const double* gy,
                        const double* gz, const double* phi, double* phib, const double* phin,
                        double* phinb) {
  double* p = static_cast<double*>(scratch(33, sizeof(double) * (size_t)d.n3));
  double* mask = static_cast<double*>(scratch(34, sizeof(double) * (size_t)d.n3));
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  // p = phin after the truncation; mask = 0 where truncated, 2 where the
  // interior clamp takes the bound, else 1
  launch_flat(d.n3, [=] __device__(int q) {
    p[q] = phin[q];
    mask[q] = 1.0;
  });
  launch_box(Box{0, ni + 1, 0, nj + 1, 0, nk + 1}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    if (!is_flat(m, gx, gy, gz, c, X, Y, Z)) return;
    double v = p[c];
    double mk = 1.0;
    if (v <= m.eps_trunc) { v = 0.0; mk = 0.0; }
    if (1.0 - v <= m.eps_trunc) { v = 1.0; mk = 0.0; }
    p[c] = v;
    mask[c] = mk;
  });
  launch_box(Box{1, ni, 1, nj, 1, nk}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const double v = p[c];
    if (!(v <= 1.0 && v >= 0.0) && mask[c] != 0.0) mask[c] = 2.0;
    p[c] = fmax(fmin(v, 1.0), 0.0);
  });
  // sums: [sum phi, sum p, sum p^2 (1-p)] -> s[0..2], offset -> s[3],
  // 1/(den + 1e-99) -> s[4]
  double* s = dev_scalars(22);
  interior_reduce<3>(
      d,
      [=] __device__(int i, int j, int k, double* o) {
        const int c = d.ix(i, j, k);
        const double v = p[c];
        o[0] = phi[c];
        o[1] = v;
        o[2] = v * v * (1.0 - v);
      },
      [=] __device__(const double* t, double* out) {
        const double den = t[2] + 1.0e-99;
        out[0] = t[0];
        out[1] = t[1];
        out[2] = t[2];
        out[3] = (t[0] - t[1]) / den;
        out[4] = 1.0 / den;
      },
      s);
  // adjoint of the offset: sum over the cells whose final clamp passes
  double* sb = dev_scalars(23);
  interior_reduce<1>(
      d,
      [=] __device__(int i, int j, int k, double* o) {
        const int c = d.ix(i, j, k);
        const double v = p[c], off = s[3];
        const double q = v + off * (v * v) * (1.0 - v);
        o[0] = (q <= 1.0 && q >= 0.0) ? phinb[c] * (v * v) * (1.0 - v) : 0.0;
      },
      [=] __device__(const double* t, double* out) { out[0] = t[0]; }, sb);
  launch_box(Box{1, ni, 1, nj, 1, nk}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const double off = s[3], rden = s[4], offb = sb[0];
    const double s1b = offb * rden, s2b = -offb * rden, denb = -offb * off * rden;
    const double v = p[c], gp = 2.0 * v - 3.0 * v * v;
    const double q = v + off * (v * v) * (1.0 - v);
    double l = (q <= 1.0 && q >= 0.0) ? phinb[c] * (1.0 + off * gp) : 0.0;
    l = l + s2b + denb * gp;
    phib[c] = phib[c] + s1b;
    phinb[c] = mask[c] == 1.0 ? l : 0.0;
  });
  // truncated cells outside the interior
  launch_box(Box{0, ni + 1, 0, nj + 1, 0, nk + 1}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const bool interior = i >= 1 && i <= ni && j >= 1 && j <= nj && k >= 1 && k <= nk;
    if (!interior && mask[c] == 0.0) phinb[c] = 0.0;
  });
}

}  // namespace fel

using namespace fel;

extern "C" {

// Adjoint of cal_grad_p2a_ (srcs/cal_grad_p2a.f:30-71).
void cal_grad_p2a_b_(const int* ID, const int* svall, const int* ni, const int* nj, const int* nk,
                     const double* dxinv, const double* dyinv_array, const double* dzinv, double* qb,
                     double* qxb, double* qyb, double* qzb) {
  adjoint_fp64_only("cal_grad_p2a_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const int count = *svall + 6;  // dyinv_array(-2:svall+3)
  double* dyt = static_cast<double*>(scratch(35, sizeof(double) * (size_t)count));
  FEL_CHECK(hipMemcpyAsync(dyt, dyinv_array, sizeof(double) * count, hipMemcpyHostToDevice, stream()));
  Call c("cal_grad_p2a_b_");
  double* q = c.io(qb, d.n3, "qb");
  double* gx = c.io(qxb, d.n3, "qxb");
  double* gy = c.io(qyb, d.n3, "qyb");
  double* gz = c.io(qzb, d.n3, "qzb");
  grad_p2a_adj(d, *ID, *dxinv, dyt, *dzinv, q, gx, gy, gz);
  c.finish();
  FEL_CHECK(hipStreamSynchronize(stream()));  // dyinv_array is a host array read asynchronously
}

// Adjoint of solphi_mthinc1_ (srcs/solphi_mthinc1.f:1-292): each of phi,
// phix, phiy, phiz, u, v, w is followed by its adjoint (incremented); the
// flux adjoints are consumed on [-1,n+1]^3.
void solphi_mthinc1_b_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                       const double* dyinv, const double* dzinv, const double* dt,
                       const double* bet_mthinc, const double* phi, double* phib, const double* phix,
                       double* phixb, const double* phiy, double* phiyb, const double* phiz,
                       double* phizb, const double* u, double* ub, const double* v, double* vb,
                       const double* w, double* wb, double* flphixb, double* flphiyb,
                       double* flphizb) {
  adjoint_fp64_only("solphi_mthinc1_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("solphi_mthinc1_b_");
  Mthinc1Adj A;
  A.m = vof_const(*dxinv, *dyinv, *dzinv, *bet_mthinc);
  A.dinv[0] = *dxinv;
  A.dinv[1] = *dyinv;
  A.dinv[2] = *dzinv;
  A.dt = *dt;
  A.phi = c.in(phi, d.n3, "phi");
  A.g[0] = c.in(phix, d.n3, "phix");
  A.g[1] = c.in(phiy, d.n3, "phiy");
  A.g[2] = c.in(phiz, d.n3, "phiz");
  A.vel[0] = c.in(u, d.n3, "u");
  A.vel[1] = c.in(v, d.n3, "v");
  A.vel[2] = c.in(w, d.n3, "w");
  A.phib = c.io(phib, d.n3, "phib");
  double* gbx = c.io(phixb, d.n3, "phixb");
  double* gby = c.io(phiyb, d.n3, "phiyb");
  double* gbz = c.io(phizb, d.n3, "phizb");
  A.velb[0] = c.io(ub, d.n3, "ub");
  A.velb[1] = c.io(vb, d.n3, "vb");
  A.velb[2] = c.io(wb, d.n3, "wb");
  A.flb[0] = c.io(flphixb, d.n3, "flphixb");
  A.flb[1] = c.io(flphiyb, d.n3, "flphiyb");
  A.flb[2] = c.io(flphizb, d.n3, "flphizb");
  for (int q = 0; q < 3; q++)
    A.rb[q] = static_cast<double*>(scratch(36 + q, sizeof(double) * (size_t)d.n3));
  mthinc1_cells(d, A);
  mthinc1_vertices(d, A, gbx, gby, gbz);
  c.finish();
}

// Adjoint of solphi_mthinc2_ (srcs/solphi_mthinc2.f:1-48): each of u, v,
// w, flphix, flphiy, flphiz, phi followed by its adjoint; phinb consumed on
// the interior.
void solphi_mthinc2_b_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                       const double* dyinv, const double* dzinv, const double* dt, const double* u,
                       double* ub, const double* v, double* vb, const double* w, double* wb,
                       const double* flphix, double* flphixb, const double* flphiy,
                       double* flphiyb, const double* flphiz, double* flphizb, const double* phi,
                       double* phib, double* phinb) {
  adjoint_fp64_only("solphi_mthinc2_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("solphi_mthinc2_b_");
  Mthinc2Adj A;
  A.dinv[0] = *dxinv;
  A.dinv[1] = *dyinv;
  A.dinv[2] = *dzinv;
  A.dt = *dt;
  A.vel[0] = c.in(u, d.n3, "u");
  A.vel[1] = c.in(v, d.n3, "v");
  A.vel[2] = c.in(w, d.n3, "w");
  A.fl[0] = c.in(flphix, d.n3, "flphix");
  A.fl[1] = c.in(flphiy, d.n3, "flphiy");
  A.fl[2] = c.in(flphiz, d.n3, "flphiz");
  A.phi = c.in(phi, d.n3, "phi");
  A.velb[0] = c.io(ub, d.n3, "ub");
  A.velb[1] = c.io(vb, d.n3, "vb");
  A.velb[2] = c.io(wb, d.n3, "wb");
  A.flb[0] = c.io(flphixb, d.n3, "flphixb");
  A.flb[1] = c.io(flphiyb, d.n3, "flphiyb");
  A.flb[2] = c.io(flphizb, d.n3, "flphizb");
  A.phib = c.io(phib, d.n3, "phib");
  A.phinb = c.io(phinb, d.n3, "phinb");
  A.G = static_cast<double*>(scratch(33, sizeof(double) * (size_t)d.n3));
  A.H = static_cast<double*>(scratch(34, sizeof(double) * (size_t)d.n3));
  mthinc2_adj(d, A);
  c.finish();
}

// Adjoint of solphi_mthinc3_ (srcs/solphi_mthinc3.f:1-163): phix, phiy,
// phiz select the flat cells (parameters); phi followed by its adjoint;
// phin is the forward's input phin, phinb the adjoint of its output on
// entry and of its input on return.
void solphi_mthinc3_b_(const int* ipara, const int* ni, const int* nj, const int* nk,
                       const double* dxinv, const double* dyinv, const double* dzinv,
                       const double* bet_mthinc, const double* phix, const double* phiy,
                       const double* phiz, const double* phi, double* phib, const double* phin,
                       double* phinb) {
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: solphi_mthinc3_b_: only the single-domain case is supported\n");
    std::abort();
  }
  adjoint_fp64_only("solphi_mthinc3_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("solphi_mthinc3_b_");
  const double* gx = c.in(phix, d.n3, "phix");
  const double* gy = c.in(phiy, d.n3, "phiy");
  const double* gz = c.in(phiz, d.n3, "phiz");
  const double* p0 = c.in(phi, d.n3, "phi");
  double* pb = c.io(phib, d.n3, "phib");
  const double* pn = c.in(phin, d.n3, "phin");
  double* pnb = c.io(phinb, d.n3, "phinb");
  mthinc3_adj(d, vof_const(*dxinv, *dyinv, *dzinv, *bet_mthinc), gx, gy, gz, p0, pb, pn, pnb);
  c.finish();
}

// Adjoint of summation_ (srcs/summation.f:27-50): qk holds the forward's
// input slices (nbub+1 padded arrays), qkb their adjoints; qkb(.,0) is
// consumed into qkb(.,1..nbub).
void summation_b_(const int* ni, const int* nj, const int* nk, const double* qk, double* qkb,
                  const int* nbub) {
  adjoint_fp64_only("summation_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const int nb = *nbub;
  Call c("summation_b_");
  const double* q = c.in(qk, (size_t)d.n3 * (nb + 1), "qk");
  double* b = c.io(qkb, (size_t)d.n3 * (nb + 1), "qkb");
  const size_t n3 = (size_t)d.n3;
  launch_box(Box{-2, d.ni + 3, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c0 = d.ix(i, j, k);
    double s = q[n3 + c0];
    for (int l = 2; l <= nb; l++) s = s + q[(size_t)l * n3 + c0];
    const bool inner = i >= 1 && i <= d.ni && j >= 1 && j <= d.nj && k >= 1 && k <= d.nk;
    const double g = (!inner || (s <= 1.0 && s >= 0.0)) ? b[c0] : 0.0;
    b[c0] = 0.0;
    for (int l = 1; l <= nb; l++) b[(size_t)l * n3 + c0] = b[(size_t)l * n3 + c0] + g;
  });
  c.finish();
}

}  // extern "C"
