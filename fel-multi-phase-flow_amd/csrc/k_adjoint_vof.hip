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
#include "felgpu_visc.hpp"

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
    // Gd = 0: every sigmoid of the cell saturated (phi at 0 or 1 to
    // rounding), the Newton step vanished and the fluxes do not vary with
    // d: no term through d
    if (Gd > 0.0) {
      A.phib[c] = A.phib[c] + db / Gd;
#pragma unroll
      for (int b = 0; b < 3; b++) nb[b] = nb[b] - db * Gn[b] / Gd;
    }
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

// ---------------------------------------------------------------------------
// Contact line.
// ---------------------------------------------------------------------------

// srcs/gnbc.f90:22-95 reversed, first launch, one thread per wall column of
// each wall: the adjoints of the contact angle's inputs theta_0 and zeta
// (its own cell) and of the column's averaged velocities ux, uz (stored:
// plane [wall][ux|uz][column]); theta_b consumed.  The clamp and acos pass
// the derivative only strictly inside (-1,1).
struct GnbcAdj {
  const double *u, *w, *th0, *zeta;
  double *th0b, *zetab, *thb;
  double uwall, sigma;
  double* UV;  // scratch: 4 planes of sx*sz
  bool lo, hi;
};

static void gnbc_adj(const Dims d, const GnbcAdj a) {
  const int ni = d.ni, nj = d.nj, plane = d.sx * d.sz;
  launch_flat(2 * plane, [=] __device__(int m) {
    const int side = m / plane, col = m - side * plane;
    double* uxb = a.UV + (2 * side) * plane;
    double* uzb = a.UV + (2 * side + 1) * plane;
    uxb[col] = 0.0;
    uzb[col] = 0.0;
    if (side == 0 ? !a.lo : !a.hi) return;
    const int i = col % d.sx - 2, k = col / d.sx - 2;
    const int jw = side ? nj : 1, ja = side ? nj : 0, jb = side ? nj + 1 : 1;
    const int mi = (i + ni + 2) % (ni + 3);
    const double pi = acos(-1.0), eps = 1.0e-12;
    const double ux = (a.u[d.ix(mi, ja, k)] + a.u[d.ix(mi, jb, k)] + a.u[d.ix(i, ja, k)] +
                       a.u[d.ix(i, jb, k)]) / 4.0 + (side ? -a.uwall : a.uwall);
    const double uz = (a.w[d.ix(mi, ja, k)] + a.w[d.ix(mi, jb, k)] + a.w[d.ix(i, ja, k)] +
                       a.w[d.ix(i, jb, k)]) / 4.0;
    const double R = sqrt(ux * ux + uz * uz + eps);
    double sgn = ux >= 0.0 ? 1.0 : -1.0;
    if (i <= ni / 2) sgn = -sgn;
    const double ucl = sgn * R;
    const int cw = d.ix(i, jw, k);
    const double th0 = a.th0[cw] * (pi / 180.0), zeta = a.zeta[cw];
    const double ct = cos(th0) - zeta * ucl / a.sigma;
    const double tb = a.thb[cw];
    a.thb[cw] = 0.0;
    if (!(ct > -1.0 && ct < 1.0)) return;
    const double ctb = tb * (180.0 / pi) * (-1.0 / sqrt(1.0 - ct * ct));
    a.th0b[cw] = a.th0b[cw] + ctb * (-sin(th0)) * (pi / 180.0);
    a.zetab[cw] = a.zetab[cw] + ctb * (-ucl / a.sigma);
    const double uclb = ctb * (-zeta / a.sigma);
    uxb[col] = uclb * sgn * ux / R;
    uzb[col] = uclb * sgn * uz / R;
  });
}

// Second launch: each velocity of the two rows either side of a wall
// gathers a quarter of ux_b, uz_b of the columns whose averages read it
// (its own column i, and the columns i' with mod(i'+ni+2, ni+3) = i).
static void gnbc_gather(const Dims d, const GnbcAdj a, double* ub, double* wb) {
  const int ni = d.ni, nj = d.nj, sx = d.sx, plane = d.sx * d.sz;
  launch_flat(4 * plane, [=] __device__(int m) {
    const int side = m / (2 * plane), r = (m / plane) % 2, col = m % plane;
    if (side == 0 ? !a.lo : !a.hi) return;
    const int x = col % sx - 2, k = col / sx - 2;
    const int j = side ? (r ? nj + 1 : nj) : (r ? 1 : 0);
    const double* uxb = a.UV + (2 * side) * plane;
    const double* uzb = a.UV + (2 * side + 1) * plane;
    double su = uxb[col], sw = uzb[col];
    const int cand[2] = {x + 1, x - ni - 2};
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int i = cand[t];
      if (i < -2 || i > ni + 3 || (i + ni + 2) % (ni + 3) != x) continue;
      const int ci = (i + 2) + sx * (k + 2);
      su = su + uxb[ci];
      sw = sw + uzb[ci];
    }
    const int c = d.ix(x, j, k);
    ub[c] = ub[c] + su / 4.0;
    wb[c] = wb[c] + sw / 4.0;
  });
}

// R = n_y/(n_t + 1e-20) of the contact-angle normal and dR/dtheta (theta in
// degrees), t = x (comp 0) or z (comp 2): n = (s sin t ax, cos t,
// s sin t az)/sqrt(|.|^2 + 1e-20) (srcs/bnd_contact_angle.f90:25-35).
__device__ __forceinline__ double ca_ratio(double th_deg, double sg, double ax, double az, int comp,
                                           double* dr) {
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

// srcs/bnd_contact_angle.f90:17-72 reversed, first launch, one thread per
// wall column: the ghost rows' adjoints (consumed) folded into the wall
// value's own part Q and the neighbour weights A (x) and B (z), stored per
// column; theta_b gathered in place.  Second launch: each wall-row value
// gathers Q of its column and A, B of its x and z neighbours (values outside
// the padded range are read as 0 by the forward and receive nothing).
static void contact_angle_adj(const Dims d, bool lo, bool hi, double ax, double az,
                              const double* q, double* qb, const double* th, double* thb,
                              double* S) {
  const int ni = d.ni, nj = d.nj, nk = d.nk, sx = d.sx, plane = d.sx * d.sz;
  launch_flat(2 * plane, [=] __device__(int m) {
    const int side = m / plane, col = m - side * plane;
    double* Q = S + (3 * side) * plane;
    double* Aw = S + (3 * side + 1) * plane;
    double* Bw = S + (3 * side + 2) * plane;
    Q[col] = 0.0;
    Aw[col] = 0.0;
    Bw[col] = 0.0;
    if (side == 0 ? !lo : !hi) return;
    const int i = col % sx - 2, k = col / sx - 2;
    const int jw = side == 0 ? 1 : nj, jg = side == 0 ? 0 : nj + 1, js = side == 0 ? -1 : 1;
    const double sg = copysign(1.0, (double)(i - ni / 2));
    const int cw = d.ix(i, jw, k);
    double dRx, dRz;
    const double Rx = ca_ratio(th[cw], sg, ax, az, 0, &dRx);
    const double Rz = ca_ratio(th[cw], sg, ax, az, 2, &dRz);
    const double qp = i + 1 > ni + 3 ? 0.0 : q[d.ix(i + 1, jw, k)];
    const double qm = i - 1 < -2 ? 0.0 : q[d.ix(i - 1, jw, k)];
    const double qt = k + 1 > nk + 3 ? 0.0 : q[d.ix(i, jw, k + 1)];
    const double qd = k - 1 < -2 ? 0.0 : q[d.ix(i, jw, k - 1)];
    const double Dx = (qp - qm) * 0.5 * ax, Dz = (qt - qd) * 0.5 * az;
    const int g0 = d.ix(i, jg, k), g1 = d.ix(i, jg + js, k), g2 = d.ix(i, jg + 2 * js, k);
    double l2 = qb[g2], l1 = qb[g1], l0 = qb[g0];
    qb[g2] = 0.0;
    qb[g1] = 0.0;
    qb[g0] = 0.0;
    l1 = l1 + 2.0 * l2;
    l0 = l0 - l2;
    l0 = l0 + 2.0 * l1;
    Q[col] = l0 - l1;
    Aw[col] = l0 * Rx * 0.5 * ax;
    Bw[col] = l0 * Rz * 0.5 * az;
    thb[cw] = thb[cw] + l0 * (-Dx * dRx - Dz * dRz);
  });
  launch_flat(2 * plane, [=] __device__(int m) {
    const int side = m / plane, col = m - side * plane;
    if (side == 0 ? !lo : !hi) return;
    const double* Q = S + (3 * side) * plane;
    const double* Aw = S + (3 * side + 1) * plane;
    const double* Bw = S + (3 * side + 2) * plane;
    const int x = col % sx - 2, k = col / sx - 2;
    const int jw = side == 0 ? 1 : nj;
    double t = Q[col];
    if (x - 1 >= -2) t = t - Aw[col - 1];
    if (x + 1 <= ni + 3) t = t + Aw[col + 1];
    if (k - 1 >= -2) t = t - Bw[col - sx];
    if (k + 1 <= nk + 3) t = t + Bw[col + sx];
    const int c = d.ix(x, jw, k);
    qb[c] = qb[c] + t;
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

// Adjoint of gnbc_ (srcs/gnbc.f90:22-95): u, w, uwall, theta_0_array and
// zeta_array each followed by its adjoint (incremented); theta_arrayb
// consumed on the wall rows.
void gnbc_b_(const int* nID, const int* ni, const int* nj, const int* nk, const double* uk,
             double* ukb, const double* wk, double* wkb, const double* uwall, double* uwallb,
             const double* theta_0_array, double* theta_0b, const double* surface_tension,
             const double* zeta_array, double* zetab, double* theta_arrayb) {
  adjoint_fp64_only("gnbc_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("gnbc_b_");
  GnbcAdj a;
  a.u = c.in(uk, d.n3, "uk");
  a.w = c.in(wk, d.n3, "wk");
  a.th0 = c.in(theta_0_array, d.n3, "theta_0_array");
  a.zeta = c.in(zeta_array, d.n3, "zeta_array");
  double* ub = c.io(ukb, d.n3, "ukb");
  double* wb = c.io(wkb, d.n3, "wkb");
  a.th0b = c.io(theta_0b, d.n3, "theta_0b");
  a.zetab = c.io(zetab, d.n3, "zetab");
  a.thb = c.io(theta_arrayb, d.n3, "theta_arrayb");
  a.uwall = *uwall;
  a.sigma = *surface_tension;
  a.lo = nID[kYMinus] < 0;
  a.hi = nID[kYPlus] < 0;
  const int plane = d.sx * d.sz;
  a.UV = static_cast<double*>(scratch(39, sizeof(double) * 4 * (size_t)plane));
  gnbc_adj(d, a);
  gnbc_gather(d, a, ub, wb);
  // uwall_b: + the bottom wall's ux_b, - the top wall's
  double* t = static_cast<double*>(scratch(40, sizeof(double) * (size_t)plane));
  const double* U = a.UV;
  launch_flat(plane, [=] __device__(int m) { t[m] = U[m] - U[2 * plane + m]; });
  adjoint_add_sum(t, plane, uwallb);
  c.finish();
}

// Adjoint of bnd_contact_angle_ (srcs/bnd_contact_angle.f90:17-72): qk is
// the forward's input, qkb the adjoint of its output on entry and of its
// input on return; theta_array followed by its adjoint (incremented).
void bnd_contact_angle_b_(const int* nID, const int* ni, const int* nj, const int* nk,
                          const double* qk, double* qkb, const double* theta_array,
                          double* theta_arrayb, const double* dx, const double* dy,
                          const double* dz) {
  adjoint_fp64_only("bnd_contact_angle_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("bnd_contact_angle_b_");
  const double* q = c.in(qk, d.n3, "qk");
  double* qb = c.io(qkb, d.n3, "qkb");
  const double* th = c.in(theta_array, d.n3, "theta_array");
  double* thb = c.io(theta_arrayb, d.n3, "theta_arrayb");
  double* S = static_cast<double*>(scratch(41, sizeof(double) * 6 * (size_t)d.sx * d.sz));
  contact_angle_adj(d, nID[kYMinus] < 0, nID[kYPlus] < 0, *dy / *dx, *dy / *dz, q, qb, th, thb, S);
  c.finish();
}

}  // extern "C"
