// k_adjoint.hip -- adjoints (reverse mode) of the ghost fills and of the
// semi-implicit viscous solve.
//
// The reference has no adjoint.  <name>_b_ applies the transpose of the
// linear map the drop-in routine <name>_ applies to its active arrays: the
// forward's argument list with every active array replaced by its adjoint,
// of the same shape (arrays that only parameterise the map -- densities,
// coefficients, slip lengths, the volume fraction -- as in the forward).
// Adjoints of the forward's inputs are incremented, adjoints of what it
// overwrites are consumed (set to zero), in/out adjoints updated in place.
//
// * bnd_periodic_b_, bnd_neumann_b_, bndu_b_: the forward's ghost writes
//   (ghost = weights x earlier values) taken in reverse order, one thread per
//   ghost line, so each line's dependent writes (the slip-wall
//   extrapolations) stay in order and no two threads touch one element.
// * solu_sor4_b_: srcs/solu_sor4.f:131-293 with the forward call's sweep
//   count frozen.  For fixed densities and coefficients the sweeps are
//   affine in (un, vn, wn, src_u, src_v, src_w, uwall); the adjoint runs the
//   sweeps, components and colours in reverse, with the refresh after every
//   second colour (bndu, then bnd_periodic of the component) reversed before
//   its colour.  One colour's update x = (0.5 S/r00 + src)/den (S = sum of
//   a_m x neighbour_m, den = dtinv + a_p/r00 0.5) is reversed in two
//   launches, for both colours of a pair at once (they read no cell of each
//   other): the first, over the pair's cells, adds x_b/den to src_b and
//   stores G = x_b/den 0.5/r00; the second, over every cell that a pair
//   cell reads, gathers sum a_m G of the pair cells that read it into its
//   adjoint and clears the pair cells' own adjoints (a cell never reads a
//   cell of its own pair, so the gather reads no adjoint it writes).
//   With coefficients formed from the viscosity on the fly, the weight a
//   colour cell x gives cell y equals the coefficient of y's own equation
//   for x, bit for bit (cal_arith_coef_vis forms both from the same four
//   viscosities in the same order), so the gather evaluates each cell's
//   own three coefficient sets once; with coefficient arrays it reads the
//   colour cells' coefficients.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "felgpu.hpp"
#include "felgpu_bnd.hpp"
#include "felgpu_visc.hpp"

namespace fel {

// q[s] += q[g]; q[g] = 0: the transpose of the copy q[g] = q[s].
__device__ __forceinline__ void take(double* q, int g, int s) {
  q[s] = q[s] + q[g];
  q[g] = 0.0;
}

// srcs/bnd_periodic.f:14-43 transposed: the z pass (every (i,j) line), then
// the x pass (every (j,k) line), each in reverse.
static void periodic_adj(const Dims d, double* q) {
  const int ni = d.ni, nk = d.nk, sx = d.sx, sy = d.sy, sz = d.sz, Z = d.sx * d.sy;
  launch_flat(sx * sy, [=] __device__(int m) {
    const int c = m;  // (i, j, k = -2) is m + 0 * Z
    const int g[6] = {0, 1, 2, nk + 3, nk + 4, nk + 5};   // k = -2..0, nk+1..nk+3
    const int s[6] = {nk, nk + 1, nk + 2, 3, 4, 5};       // k = nk-2..nk, 1..3
    for (int t = 5; t >= 0; t--) take(q, c + g[t] * Z, c + s[t] * Z);
  });
  launch_flat(sy * sz, [=] __device__(int m) {
    const int c = m * sx;  // (i = -2, j, k)
    const int g[6] = {0, 1, 2, ni + 3, ni + 4, ni + 5};
    const int s[6] = {ni, ni + 1, ni + 2, 3, 4, 5};
    for (int t = 5; t >= 0; t--) take(q, c + g[t], c + s[t]);
  });
  (void)sz;
}

// srcs/bnd_neumann.f:14-41 transposed (mirror ghosts of the j walls).
static void neumann_adj(const Dims d, double* q, bool lo, bool hi) {
  const int nj = d.nj, sx = d.sx, Y = d.sx;
  const int plane = d.sx * d.sy;
  launch_flat(d.sx * d.sz, [=] __device__(int m) {
    const int i = m % sx, k = m / sx;
    const int c = i + k * plane;  // (i-2, j = -2, k-2)
    if (hi)
      for (int g = 3; g >= 1; g--) take(q, c + (nj + 2 + g) * Y, c + (nj + 3 - g) * Y);
    if (lo)
      for (int g = 0; g >= -2; g--) take(q, c + (g + 2) * Y, c + (3 - g) * Y);
  });
}

// srcs/bndu.f:29-158 transposed: the z fills, the x fills (with the shifted
// sources of u in x and of w in z), then the wall rows, top first.  `wacc`
// (one value per (i,k) wall column) gathers the uwall weights: -c1 x u_b of
// the bottom ghost and +c1 x u_b of the top one.
static void bndu_adj(const Dims d, double* u, double* v, double* w, double* wacc, double dy,
                     const double* l1, const double* l2, const double* phi, bool lo, bool hi) {
  const int ni = d.ni, nj = d.nj, nk = d.nk, sx = d.sx, sy = d.sy, Z = d.sx * d.sy;
  // z fills (bndu.f:130-155), every (i,j) line, statements reversed
  launch_flat(sx * sy, [=] __device__(int c) {
    auto kz = [=](int k) { return c + (k + 2) * Z; };
    take(w, kz(nk + 3), kz(3)); take(w, kz(nk + 2), kz(2));
    take(v, kz(nk + 3), kz(3)); take(v, kz(nk + 2), kz(2)); take(v, kz(nk + 1), kz(1));
    take(u, kz(nk + 3), kz(3)); take(u, kz(nk + 2), kz(2)); take(u, kz(nk + 1), kz(1));
    take(w, kz(-2), kz(nk - 1)); take(w, kz(-1), kz(nk)); take(w, kz(0), kz(nk + 1));
    take(v, kz(-2), kz(nk - 2)); take(v, kz(-1), kz(nk - 1)); take(v, kz(0), kz(nk));
    take(u, kz(-2), kz(nk - 2)); take(u, kz(-1), kz(nk - 1)); take(u, kz(0), kz(nk));
  });
  // x fills (bndu.f:104-127), every (j,k) line, statements reversed
  launch_flat(sy * d.sz, [=] __device__(int m) {
    const int c = m * sx;
    auto ix = [=](int i) { return c + i + 2; };
    take(w, ix(ni + 3), ix(3)); take(w, ix(ni + 2), ix(2)); take(w, ix(ni + 1), ix(1));
    take(v, ix(ni + 3), ix(3)); take(v, ix(ni + 2), ix(2)); take(v, ix(ni + 1), ix(1));
    take(u, ix(ni + 3), ix(3)); take(u, ix(ni + 2), ix(2));
    take(w, ix(-2), ix(ni - 2)); take(w, ix(-1), ix(ni - 1)); take(w, ix(0), ix(ni));
    take(v, ix(-2), ix(ni - 2)); take(v, ix(-1), ix(ni - 1)); take(v, ix(0), ix(ni));
    take(u, ix(-2), ix(ni - 1)); take(u, ix(-1), ix(ni)); take(u, ix(0), ix(ni + 1));
  });
  if (!lo && !hi) return;
  // wall rows (bndu.f:29-98), every (i,k) column, top wall first
  launch_flat(sx * d.sz, [=] __device__(int m) {
    const int i = m % sx - 2, k = m / sx - 2;
    auto jy = [=](int j) { return d.ix(i, j, k); };
    double acc = 0.0;
    if (hi) {
      const WallCoef<double> cf = wall_coef(d, i, nj, k, dy, l1, l2, phi);
      // u, w: g1 = c1 (+uwall) + c2 q(nj); g2 = 2 g1 - q(nj); g3 = 2 g2 - g1
      double* qs[2] = {w, u};
      for (int t = 0; t < 2; t++) {
        double* q = qs[t];
        const double b3 = q[jy(nj + 3)];
        q[jy(nj + 2)] = q[jy(nj + 2)] + 2.0 * b3;
        q[jy(nj + 1)] = q[jy(nj + 1)] - b3;
        q[jy(nj + 3)] = 0.0;
        const double b2 = q[jy(nj + 2)];
        q[jy(nj + 1)] = q[jy(nj + 1)] + 2.0 * b2;
        q[jy(nj)] = q[jy(nj)] - b2;
        q[jy(nj + 2)] = 0.0;
        const double b1 = q[jy(nj + 1)];
        q[jy(nj)] = q[jy(nj)] + cf.c2 * b1;
        if (t == 1) acc = acc + cf.c1 * b1;
        q[jy(nj + 1)] = 0.0;
      }
      // v: v(nj) = 0, v(nj+g) = -v(nj-g)
      for (int g = 3; g >= 1; g--) {
        v[jy(nj - g)] = v[jy(nj - g)] - v[jy(nj + g)];
        v[jy(nj + g)] = 0.0;
      }
      v[jy(nj)] = 0.0;
    }
    if (lo) {
      const WallCoef<double> cf = wall_coef(d, i, 1, k, dy, l1, l2, phi);
      double* qs[2] = {w, u};
      for (int t = 0; t < 2; t++) {
        double* q = qs[t];
        const double b3 = q[jy(-2)];
        q[jy(-1)] = q[jy(-1)] + 2.0 * b3;
        q[jy(0)] = q[jy(0)] - b3;
        q[jy(-2)] = 0.0;
        const double b2 = q[jy(-1)];
        q[jy(0)] = q[jy(0)] + 2.0 * b2;
        q[jy(1)] = q[jy(1)] - b2;
        q[jy(-1)] = 0.0;
        const double b1 = q[jy(0)];
        q[jy(1)] = q[jy(1)] + cf.c2 * b1;
        if (t == 1) acc = acc - cf.c1 * b1;
        q[jy(0)] = 0.0;
      }
      // v: v(-2) = -v(2), v(-1) = -v(1), v(0) = 0
      v[jy(0)] = 0.0;
      v[jy(1)] = v[jy(1)] - v[jy(-1)];
      v[jy(-1)] = 0.0;
      v[jy(2)] = v[jy(2)] - v[jy(-2)];
      v[jy(-2)] = 0.0;
    }
    if (wacc) wacc[m] = wacc[m] + acc;
  });
}

// Sum of the n values of a device array, added to *out on the host.
void adjoint_add_sum(const double* a, int n, double* out) {
  double* s = dev_scalars(21);
  Dims p{};
  p.ni = n;
  p.nj = 1;
  p.nk = 1;
  interior_reduce<1>(
      p, [=] __device__(int i, int, int, double* o) { o[0] = a[i - 1]; },
      [=] __device__(const double* t, double* o) { o[0] = t[0]; }, s);
  *out += read_scalars(s, 1)[0];
}

// ---------------------------------------------------------------------------
// Gauss-Seidel solve, reversed.
// ---------------------------------------------------------------------------
struct SorAdj {
  const double *rho, *rhon, *rmu;
  const double* cf[45];
  VisScales<double> sc;
  double dtinv;
  double* lb[3];    // adjoints of un, vn, wn
  double* srcb[3];  // adjoints of src_u, src_v, src_w
  double* G;        // x_b/den 0.5/r00 at the current colour's cells
  int jen;
};

// Colour parities of srcs/solu_sor4.f:94-120 in sweep order.
static const int kIsor[8] = {1, 2, 1, 2, 1, 2, 2, 1};
static const int kJsor[8] = {1, 2, 1, 2, 2, 1, 1, 2};
static const int kKsor[8] = {1, 2, 2, 1, 1, 2, 1, 2};

template <int P>
__device__ __forceinline__ void coef_of(const Dims& d, const VisScales<double>& s, const double* r,
                                        int c, double* v) {
  if (P == 0) coef_u(d, s, r, c, v);
  if (P == 1) coef_v(d, s, r, c, v);
  if (P == 2) coef_w(d, s, r, c, v);
}

// First launch of a colour pair: src_b += x_b/den and G = x_b/den 0.5/r00
// at the cells of component P of both colours: the colour with base
// parities (is,js,ks) and its complement (the pair's other colour).  The two
// colours of a pair differ in the parity of every index, so no cell of one
// reads a cell of the other and the forward's two updates commute: their
// reverse is one pass.
template <int P, bool ONFLY>
static void colour_scale(const Dims d, const SorAdj a, int is, int js, int ks) {
  const int jend = P == 1 ? a.jen : d.nj;
  int h[2][3], base[2][3], cnt[2];
  for (int t = 0; t < 2; t++) {
    const int ps[3] = {t ? 3 - is : is, t ? 3 - js : js, t ? 3 - ks : ks};
    const int top[3] = {d.ni, jend, d.nk};
    cnt[t] = 1;
    for (int q = 0; q < 3; q++) {
      base[t][q] = ps[q];
      h[t][q] = top[q] >= ps[q] ? (top[q] - ps[q]) / 2 + 1 : 0;
      cnt[t] *= h[t][q];
    }
  }
  const int n0 = cnt[0];
  const int hx0 = h[0][0], hy0 = h[0][1], hx1 = h[1][0], hy1 = h[1][1];
  const int b00 = base[0][0], b01 = base[0][1], b02 = base[0][2];
  const int b10 = base[1][0], b11 = base[1][1], b12 = base[1][2];
  launch_flat(cnt[0] + cnt[1], [=] __device__(int m) {
    const bool second = m >= n0;
    const int mm = second ? m - n0 : m;
    const int hx = second ? hx1 : hx0, hy = second ? hy1 : hy0;
    const int ii = mm % hx, t = mm / hx, jj = t % hy, kk = t / hy;
    const int c = second ? d.ix(b10 + 2 * ii, b11 + 2 * jj, b12 + 2 * kk)
                         : d.ix(b00 + 2 * ii, b01 + 2 * jj, b02 + 2 * kk);
    const int F = P == 0 ? 1 : (P == 1 ? d.sx : d.sx * d.sy);
    const double r00 = (a.rho[c] + a.rho[c + F] + a.rhon[c] + a.rhon[c + F]) * 0.25;
    double ap;
    if (ONFLY) {
      double cf[15];
      coef_of<P>(d, a.sc, a.rmu, c, cf);
      ap = cf[6];
    } else {
      ap = a.cf[15 * P + 6][c];
    }
    const double den = a.dtinv + ap / r00 * 0.5;
    const double g = a.lb[P][c] / den;
    a.srcb[P][c] = a.srcb[P][c] + g;
    a.G[c] = g * 0.5 / r00;
  });
}

// Second launch of a colour pair: every cell y in [0,n+1]^3 gathers the
// weights of the pair's cells of component P that read it, into the
// adjoints of u, v and w; the pair cells' own adjoints are cleared.
template <int P, bool ONFLY>
static void colour_gather(const Dims d, const SorAdj a, int is, int js, int ks) {
  const int ni = d.ni, nk = d.nk, jend = P == 1 ? a.jen : d.nj;
  launch_box(Box{0, d.ni + 1, 0, d.nj + 1, 0, d.nk + 1}, [=] __device__(int i, int j, int k) {
    const int y = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    // G of the pair cell at offset (di,dj,dk), 0 when that cell is not one:
    // its parities all match (is,js,ks) or all differ from them
    auto in_c = [=](int di, int dj, int dk) {
      const int xi = i + di, xj = j + dj, xk = k + dk;
      const int bi = (xi ^ is) & 1, bj = (xj ^ js) & 1, bk = (xk ^ ks) & 1;
      return bi == bj && bj == bk && xi >= 1 && xi <= ni && xj >= 1 && xj <= jend && xk >= 1 &&
             xk <= nk;
    };
    auto g_at = [=](int di, int dj, int dk) {
      return in_c(di, dj, dk) ? a.G[y + di * X + dj * Y + dk * Z] : 0.0;
    };
    // weight array m of the colour cell at offset o (array path)
    auto w_at = [=](int m, int di, int dj, int dk) {
      return in_c(di, dj, dk) ? a.cf[15 * P + m][y + di * X + dj * Y + dk * Z] *
                                    a.G[y + di * X + dj * Y + dk * Z]
                              : 0.0;
    };
    // own component
    double* self = a.lb[P];
    if (in_c(0, 0, 0)) {
      self[y] = 0.0;
    } else {
      double s;
      if (ONFLY) {
        double cf[15];
        coef_of<P>(d, a.sc, a.rmu, y, cf);
        s = cf[0] * g_at(-1, 0, 0) + cf[1] * g_at(1, 0, 0) + cf[2] * g_at(0, -1, 0) +
            cf[3] * g_at(0, 1, 0) + cf[4] * g_at(0, 0, -1) + cf[5] * g_at(0, 0, 1);
      } else {
        s = w_at(1, -1, 0, 0) + w_at(0, 1, 0, 0) + w_at(3, 0, -1, 0) + w_at(2, 0, 1, 0) +
            w_at(5, 0, 0, -1) + w_at(4, 0, 0, 1);
      }
      self[y] = self[y] + s;
    }
    // the other two components
#pragma unroll
    for (int Q = 0; Q < 3; Q++) {
      if (Q == P) continue;
      double s;
      if (ONFLY) {
        // terms of Q's own equation at y that read component P
        double cf[15];
        if (Q == 0) coef_u(d, a.sc, a.rmu, y, cf);
        if (Q == 1) coef_v(d, a.sc, a.rmu, y, cf);
        if (Q == 2) coef_w(d, a.sc, a.rmu, y, cf);
        if (Q == 0 && P == 1)
          s = cf[7] * g_at(0, -1, 0) + cf[8] * g_at(1, -1, 0) + cf[9] * g_at(0, 0, 0) +
              cf[10] * g_at(1, 0, 0);
        else if (Q == 0 && P == 2)
          s = cf[11] * g_at(0, 0, -1) + cf[12] * g_at(1, 0, -1) + cf[13] * g_at(0, 0, 0) +
              cf[14] * g_at(1, 0, 0);
        else if (Q == 1 && P == 0)
          s = cf[7] * g_at(-1, 0, 0) + cf[9] * g_at(0, 0, 0) + cf[8] * g_at(-1, 1, 0) +
              cf[10] * g_at(0, 1, 0);
        else if (Q == 1 && P == 2)
          s = cf[11] * g_at(0, 0, -1) + cf[12] * g_at(0, 1, -1) + cf[13] * g_at(0, 0, 0) +
              cf[14] * g_at(0, 1, 0);
        else if (Q == 2 && P == 0)
          s = cf[7] * g_at(-1, 0, 0) + cf[9] * g_at(0, 0, 0) + cf[8] * g_at(-1, 0, 1) +
              cf[10] * g_at(0, 0, 1);
        else  // Q == 2 && P == 1
          s = cf[11] * g_at(0, -1, 0) + cf[13] * g_at(0, 0, 0) + cf[12] * g_at(0, -1, 1) +
              cf[14] * g_at(0, 0, 1);
      } else {
        // terms of P's equation at the colour cells that read Q at y
        if (P == 0 && Q == 1)
          s = w_at(7, 0, 1, 0) + w_at(8, -1, 1, 0) + w_at(9, 0, 0, 0) + w_at(10, -1, 0, 0);
        else if (P == 0 && Q == 2)
          s = w_at(11, 0, 0, 1) + w_at(12, -1, 0, 1) + w_at(13, 0, 0, 0) + w_at(14, -1, 0, 0);
        else if (P == 1 && Q == 0)
          s = w_at(7, 1, 0, 0) + w_at(9, 0, 0, 0) + w_at(8, 1, -1, 0) + w_at(10, 0, -1, 0);
        else if (P == 1 && Q == 2)
          s = w_at(11, 0, 0, 1) + w_at(12, 0, -1, 1) + w_at(13, 0, 0, 0) + w_at(14, 0, -1, 0);
        else if (P == 2 && Q == 0)
          s = w_at(7, 1, 0, 0) + w_at(9, 0, 0, 0) + w_at(8, 1, 0, -1) + w_at(10, 0, 0, -1);
        else  // P == 2 && Q == 1
          s = w_at(11, 0, 1, 0) + w_at(13, 0, 0, 0) + w_at(12, 0, 1, -1) + w_at(14, 0, 0, -1);
      }
      a.lb[Q][y] = a.lb[Q][y] + s;
    }
  });
}

// Colour pair p (colours 2p and 2p+1 of the sweep order) of component P.
template <int P>
static void pair_adj(bool onfly, const Dims d, const SorAdj& a, int p) {
  const int is = kIsor[2 * p], js = kJsor[2 * p], ks = kKsor[2 * p];
  if (onfly) {
    colour_scale<P, true>(d, a, is, js, ks);
    colour_gather<P, true>(d, a, is, js, ks);
  } else {
    colour_scale<P, false>(d, a, is, js, ks);
    colour_gather<P, false>(d, a, is, js, ks);
  }
}

// ---------------------------------------------------------------------------
// Explicit momentum terms, reversed.  Every kernel gathers what one cell
// receives; where a cell's adjoint is also read by its neighbours, the
// consumed outputs are cleared by a second launch.
// ---------------------------------------------------------------------------

// Membership of (i,j,k) in the box [i0,i1]x[j0,j1]x[k0,k1].
struct IBox {
  int i0, i1, j0, j1, k0, k1;
  __device__ __forceinline__ bool has(int i, int j, int k) const {
    return i >= i0 && i <= i1 && j >= j0 && j <= j1 && k >= k0 && k <= k1;
  }
};

// srcs/calphat.f:16-22 reversed (phat on i in [-1,ni+3], j, k in [-2,n+3]).
static void calphat_adj(const Dims d, double* pb, double* pob, double* phatb) {
  launch_box(Box{-1, d.ni + 3, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const double g = phatb[c];
    pb[c] = pb[c] + g * 2.0;
    pob[c] = pob[c] - g;
    phatb[c] = 0.0;
  });
}

// srcs/cal_grad_prs.f:24-37 reversed (gradients on [-2,n+2]^3).
static void grad_prs_adj(const Dims d, double dxinv, double dyinv, double dzinv, double* qb,
                         double* gx, double* gy, double* gz) {
  const IBox B{-2, d.ni + 2, -2, d.nj + 2, -2, d.nk + 2};
  launch_box(Box{-2, d.ni + 3, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    double t = 0.0;
    if (B.has(i, j, k)) t = t - gx[c] * dxinv - gy[c] * dyinv - gz[c] * dzinv;
    if (B.has(i - 1, j, k)) t = t + gx[c - X] * dxinv;
    if (B.has(i, j - 1, k)) t = t + gy[c - Y] * dyinv;
    if (B.has(i, j, k - 1)) t = t + gz[c - Z] * dzinv;
    qb[c] = qb[c] + t;
  });
  launch_box(Box{-2, d.ni + 2, -2, d.nj + 2, -2, d.nk + 2}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    gx[c] = 0.0;
    gy[c] = 0.0;
    gz[c] = 0.0;
  });
}

// Boxes of the six strain-rate components written by srcs/calsij.f:28-139
// (the shear boxes are also those of the edge viscosities of
// srcs/cal_arith_tau.f:40-118).
struct SijBoxes {
  IBox b[6];
};
static SijBoxes sij_boxes(const Dims& d) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  return SijBoxes{{IBox{-1, ni + 3, -2, nj + 3, -2, nk + 3}, IBox{-2, ni + 3, -1, nj + 3, -2, nk + 3},
                   IBox{-1, ni + 3, -1, nj + 3, -1, nk + 3}, IBox{-2, ni + 2, -2, nj + 2, -2, nk + 3},
                   IBox{-2, ni + 2, -2, nj + 3, -2, nk + 2}, IBox{-2, ni + 3, -2, nj + 2, -2, nk + 2}}};
}

// srcs/calsij.f:28-139 reversed: velocity adjoints gather the strain-rate
// adjoints of the components that read them; then the strain-rate adjoints
// are cleared on their boxes.
static void calsij_adj(const Dims d, double dxinv, double dyinv, double dzinv, double* ub,
                       double* vb, double* wb, double* sb) {
  const SijBoxes S = sij_boxes(d);
  const size_t n3 = (size_t)d.n3;
  launch_box(Box{-2, d.ni + 3, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    const double *s1 = sb, *s2 = sb + n3, *s3 = sb + 2 * n3, *s4 = sb + 3 * n3,
                 *s5 = sb + 4 * n3, *s6 = sb + 5 * n3;
    auto at = [&](const double* s, int l, int di, int dj, int dk, int o) {
      return S.b[l].has(i + di, j + dj, k + dk) ? s[c + o] : 0.0;
    };
    const double tu = (at(s1, 0, 0, 0, 0, 0) - at(s1, 0, 1, 0, 0, X)) * dxinv +
                      (at(s4, 3, 0, -1, 0, -Y) - at(s4, 3, 0, 0, 0, 0)) * 0.5 * dyinv +
                      (at(s5, 4, 0, 0, -1, -Z) - at(s5, 4, 0, 0, 0, 0)) * 0.5 * dzinv;
    const double tv = (at(s2, 1, 0, 0, 0, 0) - at(s2, 1, 0, 1, 0, Y)) * dyinv +
                      (at(s4, 3, -1, 0, 0, -X) - at(s4, 3, 0, 0, 0, 0)) * 0.5 * dxinv +
                      (at(s6, 5, 0, 0, -1, -Z) - at(s6, 5, 0, 0, 0, 0)) * 0.5 * dzinv;
    const double tw = (at(s3, 2, 0, 0, 0, 0) - at(s3, 2, 0, 0, 1, Z)) * dzinv +
                      (at(s5, 4, -1, 0, 0, -X) - at(s5, 4, 0, 0, 0, 0)) * 0.5 * dxinv +
                      (at(s6, 5, 0, -1, 0, -Y) - at(s6, 5, 0, 0, 0, 0)) * 0.5 * dyinv;
    ub[c] = ub[c] + tu;
    vb[c] = vb[c] + tv;
    wb[c] = wb[c] + tw;
  });
  launch_box(Box{-2, d.ni + 3, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
#pragma unroll
    for (int l = 0; l < 6; l++)
      if (S.b[l].has(i, j, k)) sb[l * n3 + c] = 0.0;
  });
}

// srcs/cal_arith_tau.f:21-118 reversed (pointwise; the viscosity fixed).
static void arith_tau_adj(const Dims d, double* sb, const double* rmu, double* taub) {
  const SijBoxes S = sij_boxes(d);
  const size_t n3 = (size_t)d.n3;
  launch_box(Box{-2, d.ni + 3, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
#pragma unroll
    for (int l = 0; l < 3; l++) {
      sb[l * n3 + c] = sb[l * n3 + c] + taub[l * n3 + c] * rmu[c] * 2.0;
      taub[l * n3 + c] = 0.0;
    }
    const int o1[3] = {X, X, Y}, o2[3] = {Y, Z, Z};
#pragma unroll
    for (int l = 3; l < 6; l++) {
      if (!S.b[l].has(i, j, k)) continue;
      const int a = o1[l - 3], b = o2[l - 3];
      const double m = +0.25 * rmu[c] + 0.25 * rmu[c + a] + 0.25 * rmu[c + b] + 0.25 * rmu[c + a + b];
      sb[l * n3 + c] = sb[l * n3 + c] + taub[l * n3 + c] * m * 2.0;
      taub[l * n3 + c] = 0.0;
    }
  });
}

// srcs/cal_div_tensor.f:20-86 reversed: the stress adjoints gather the
// velocity-point adjoints on the boxes the forward writes; then the vis
// adjoints are cleared everywhere (the forward overwrites the arrays whole).
static void div_tensor_adj(const Dims d, double dxinv, double dyinv, double dzinv, double* tb,
                           double* ub, double* vb, double* wb) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const IBox Bu{-2, ni + 2, -1, nj + 3, -1, nk + 3}, Bv{-1, ni + 3, -2, nj + 2, -1, nk + 3},
      Bw{-1, ni + 3, -1, nj + 3, -2, nk + 2};
  const size_t n3 = (size_t)d.n3;
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    auto U = [&](int di, int dj, int dk, int o) { return Bu.has(i + di, j + dj, k + dk) ? ub[c + o] : 0.0; };
    auto V = [&](int di, int dj, int dk, int o) { return Bv.has(i + di, j + dj, k + dk) ? vb[c + o] : 0.0; };
    auto W = [&](int di, int dj, int dk, int o) { return Bw.has(i + di, j + dj, k + dk) ? wb[c + o] : 0.0; };
    const double u0 = U(0, 0, 0, 0), v0 = V(0, 0, 0, 0), w0 = W(0, 0, 0, 0);
    double* t = tb + c;
    t[0] = t[0] + (U(-1, 0, 0, -X) - u0) * dxinv;
    t[n3] = t[n3] + (V(0, -1, 0, -Y) - v0) * dyinv;
    t[2 * n3] = t[2 * n3] + (W(0, 0, -1, -Z) - w0) * dzinv;
    t[3 * n3] = t[3 * n3] + (u0 - U(0, 1, 0, Y)) * dyinv + (v0 - V(1, 0, 0, X)) * dxinv;
    t[4 * n3] = t[4 * n3] + (u0 - U(0, 0, 1, Z)) * dzinv + (w0 - W(1, 0, 0, X)) * dxinv;
    t[5 * n3] = t[5 * n3] + (v0 - V(0, 0, 1, Z)) * dzinv + (w0 - W(0, 1, 0, Y)) * dyinv;
  });
  launch_flat(d.n3, [=] __device__(int m) {
    ub[m] = 0.0;
    vb[m] = 0.0;
    wb[m] = 0.0;
  });
}

// srcs/cal_srcu.f:51-122 reversed (pointwise, densities fixed).
struct SrcuAdj {
  double* vel[3];
  double *adv[3], *advo[3], *prs[3], *vis[3], *fst[3], *fstn[3], *src[3];
  const double *rho, *rhon;
};

static void srcu_adj(const Dims d, int jen, double dtinv, const SrcuAdj a) {
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const int F[3] = {1, d.sx, d.sx * d.sy};
#pragma unroll
    for (int q = 0; q < 3; q++) {
      if (q == 1 && j > jen) continue;
      const double r = (a.rho[c] + a.rho[c + F[q]] + a.rhon[c] + a.rhon[c + F[q]]) * 0.25;
      const double s = a.src[q][c];
      a.vel[q][c] = a.vel[q][c] + s * dtinv;
      a.adv[q][c] = a.adv[q][c] - s * 1.5;
      a.advo[q][c] = a.advo[q][c] + s * 0.5;
      a.prs[q][c] = a.prs[q][c] - s / r;
      a.vis[q][c] = a.vis[q][c] + s / r * 0.5;
      a.fst[q][c] = a.fst[q][c] + s * 0.5;
      a.fstn[q][c] = a.fstn[q][c] + s * 0.5;
      a.src[q][c] = 0.0;
    }
  });
}

// Reverse of the WENO5 derivative (srcs/weno5.h:24-60): from the six
// differences df (dfm5, dfm3, dfm1, dfp1, dfp3, dfp5), the transport
// velocity and the adjoint rb of the result, the adjoints of the
// differences (db) and, returned, of the velocity; d|vel|/dvel = 0 at 0.
__device__ __forceinline__ double weno5_adj(const double* df, double vel00, double rb, double* db) {
  const double verysmall = 1.0e-80, d1_12 = 1.0 / 12.0;
  const double dfm5 = df[0], dfm3 = df[1], dfm1 = df[2], dfp1 = df[3], dfp3 = df[4], dfp5 = df[5];
  const double f[6] = {dfm5 * 2.0 - dfm3 * 7.0 + dfm1 * 11.0, -dfm3 + dfm1 * 5.0 + dfp1 * 2.0,
                       dfm1 * 2.0 + dfp1 * 5.0 - dfp3, dfp5 * 2.0 - dfp3 * 7.0 + dfp1 * 11.0,
                       -dfp3 + dfp1 * 5.0 + dfm1 * 2.0, dfp1 * 2.0 + dfm1 * 5.0 - dfm3};
  double a[6], b[6], sm[6];
  a[0] = dfm5 - dfm3 * 2.0 + dfm1; b[0] = dfm5 - dfm3 * 4.0 + dfm1 * 3.0;
  a[1] = dfm3 - dfm1 * 2.0 + dfp1; b[1] = dfm3 - dfp1;
  a[2] = dfm1 - dfp1 * 2.0 + dfp3; b[2] = dfm1 * 3.0 - dfp1 * 4.0 + dfp3;
  a[3] = dfp5 - dfp3 * 2.0 + dfp1; b[3] = dfp5 - dfp3 * 4.0 + dfp1 * 3.0;
  a[4] = dfp3 - dfp1 * 2.0 + dfm1; b[4] = dfp3 - dfm1;
  a[5] = dfp1 - dfm1 * 2.0 + dfm3; b[5] = dfp1 * 3.0 - dfm1 * 4.0 + dfm3;
#pragma unroll
  for (int t = 0; t < 6; t++) sm[t] = a[t] * a[t] * 13.0 + b[t] * b[t] * 3.0;
  const double velab = fabs(vel00);
  const double vfac[2] = {vel00 + velab, vel00 - velab};
  const double sg = vel00 > 0.0 ? 1.0 : (vel00 < 0.0 ? -1.0 : 0.0);
  double velb = 0.0, smb[6], cb[6];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const double* s = sm + 3 * side;
    const double* fc = f + 3 * side;
    const double P1 = s[1] * s[2], P2 = s[0] * s[2], P3 = s[0] * s[1];
    const double am1 = P1 * P1, am2 = P2 * P2 * 6.0, am3 = P3 * P3 * 3.0;
    const double deninv = 1.0 / (am1 + am2 + am3 + verysmall);
    const double w1 = am1 * deninv, w3 = am3 * deninv, w2 = 1.0 - w1 - w3;
    const double R = w1 * fc[0] + w2 * fc[1] + w3 * fc[2];
    velb = velb + R * d1_12 * rb * (side ? (1.0 - sg) : (1.0 + sg));
    const double Rb = vfac[side] * d1_12 * rb;
    cb[3 * side] = Rb * w1;
    cb[3 * side + 1] = Rb * w2;
    cb[3 * side + 2] = Rb * w3;
    const double w1b = Rb * (fc[0] - fc[1]), w3b = Rb * (fc[2] - fc[1]);
    const double denb = -(w1b * am1 + w3b * am3) * deninv * deninv;
    const double am1b = w1b * deninv + denb, am2b = denb, am3b = w3b * deninv + denb;
    const double P1b = 2.0 * P1 * am1b, P2b = 12.0 * P2 * am2b, P3b = 6.0 * P3 * am3b;
    smb[3 * side] = P2b * s[2] + P3b * s[1];
    smb[3 * side + 1] = P1b * s[2] + P3b * s[0];
    smb[3 * side + 2] = P1b * s[1] + P2b * s[0];
  }
  // candidate stencils (dfm_k, dfp_k) and smoothness indicators
  db[0] = cb[0] * 2.0;
  db[1] = -cb[0] * 7.0 - cb[1] - cb[5];
  db[2] = cb[0] * 11.0 + cb[1] * 5.0 + cb[2] * 2.0 + cb[4] * 2.0 + cb[5] * 5.0;
  db[3] = cb[1] * 2.0 + cb[2] * 5.0 + cb[3] * 11.0 + cb[4] * 5.0 + cb[5] * 2.0;
  db[4] = -cb[2] - cb[3] * 7.0 - cb[4];
  db[5] = cb[3] * 2.0;
  double ab[6], bb[6];
#pragma unroll
  for (int t = 0; t < 6; t++) {
    ab[t] = 26.0 * a[t] * smb[t];
    bb[t] = 6.0 * b[t] * smb[t];
  }
  // d a_t / d df and d b_t / d df, row by row as a[t], b[t] above
  db[0] += ab[0] + bb[0];
  db[1] += -2.0 * ab[0] - 4.0 * bb[0] + ab[1] + bb[1] + ab[5] + bb[5];
  db[2] += ab[0] + 3.0 * bb[0] - 2.0 * ab[1] + ab[2] + 3.0 * bb[2] + ab[4] - bb[4] - 2.0 * ab[5] -
           4.0 * bb[5];
  db[3] += ab[1] - bb[1] - 2.0 * ab[2] - 4.0 * bb[2] + ab[3] + 3.0 * bb[3] - 2.0 * ab[4] + ab[5] +
           3.0 * bb[5];
  db[4] += ab[2] + bb[2] - 2.0 * ab[3] - 4.0 * bb[3] + ab[4] + bb[4];
  db[5] += ab[3] + bb[3];
  return velb;
}

// The WENO5 derivative of component field f at interior cell c along the
// direction of stride st (srcs/cal_advu_weno5.f:86-92), reversed: from the
// adjoint rb of the scaled derivative, the adjoints of its six differences
// (P[0..5], dfm5 .. dfp5) and a quarter of its transport velocity's
// adjoint (V, the weight of each of the four averaged values, which lie at
// c - st, c - st + eq, c, c + eq of the transporting component tv).
__device__ __forceinline__ void weno_cell_adj(const double* f, const double* tv, int c, int st,
                                              int eq, double rb, double* P, double* V) {
  const double cm = (tv[c - st] + tv[c - st + eq]) * 0.5, cp = (tv[c] + tv[c + eq]) * 0.5;
  const double vel00 = (cm + cp) * 0.5;
  double df[6];
#pragma unroll
  for (int m = 0; m < 6; m++) df[m] = -f[c + (m - 3) * st] + f[c + (m - 2) * st];
  *V = weno5_adj(df, vel00, rb, P) * 0.25;
}

// Difference adjoints gathered into the component's adjoint along x, one
// block per interior (j,k) row: the six adjoints of every cell of the row go
// to LDS, then each position y in [-2, ni+3] adds
//   sum over cells c = y + o (o in -3..3) of P[2-o](c) - P[3-o](c)
// (difference m of cell c reads f at c + (m-3) st with -1 and c + (m-2) st
// with +1).  Needs 48 ni bytes of LDS.
__global__ void __launch_bounds__(256) weno_adj_x_kernel(Dims d, const double* f, const double* tv,
                                                         int eq, const double* advb, double dinv,
                                                         double* fb, double* V) {
  extern __shared__ double sh[];
  const int ni = d.ni;
  const int j = (int)blockIdx.x % d.nj + 1, k = (int)blockIdx.x / d.nj + 1;
  for (int i = (int)threadIdx.x + 1; i <= ni; i += 256) {
    const int c = d.ix(i, j, k);
    double P[6], v;
    weno_cell_adj(f, tv, c, 1, eq, advb[c] * dinv, P, &v);
#pragma unroll
    for (int m = 0; m < 6; m++) sh[m * ni + i - 1] = P[m];
    V[c] = v;
  }
  __syncthreads();
  for (int y = (int)threadIdx.x - 2; y <= ni + 3; y += 256) {
    double g = 0.0;
#pragma unroll
    for (int o = -3; o <= 3; o++) {
      const int x = y + o;
      if (x < 1 || x > ni) continue;
      if (2 - o >= 0 && 2 - o <= 5) g = g + sh[(2 - o) * ni + x - 1];
      if (3 - o >= 0 && 3 - o <= 5) g = g - sh[(3 - o) * ni + x - 1];
    }
    double* o = fb + d.ix(y, j, k);
    *o = *o + g;
  }
}

// The same gather along y or z (or x for rows too long for LDS): one thread
// per segment of kWenoAdjSeg positions of an interior line, marching along
// it.  Cell c's adjoints are added to the running sums of the seven
// positions c-3..c+3, after which position c-3 is complete; a segment
// [y0, y0+kWenoAdjSeg) takes the cells c in [y0-3, y0+kWenoAdjSeg+2] (the
// three cells either side are formed by both neighbouring segments).  Each
// cell's difference adjoints go to its segment's thread only, and V is
// stored by the segment that owns the cell.
constexpr int kWenoAdjSeg = 32;

__global__ void __launch_bounds__(256) weno_adj_march_kernel(Dims d, int dd, const double* f,
                                                             const double* tv, int eq,
                                                             const double* advb, double dinv,
                                                             double* fb, double* V, int nseg) {
  const int n[3] = {d.ni, d.nj, d.nk};
  const int stv[3] = {1, d.sx, d.sx * d.sy};
  const int a0 = dd == 0 ? 1 : 0, a1 = dd == 2 ? 1 : 2;  // the two line axes
  const int gid = (int)(blockIdx.x * 256 + threadIdx.x);
  const int l0 = gid % n[a0], r = gid / n[a0], seg = r % nseg, l1 = r / nseg;
  if (l1 >= n[a1]) return;
  int pos[3];
  pos[a0] = l0 + 1;
  pos[a1] = l1 + 1;
  pos[dd] = 0;
  const int base = d.ix(pos[0], pos[1], pos[2]);
  const int st = stv[dd], nl = n[dd];
  const int y0 = -2 + seg * kWenoAdjSeg;
  const int y1 = min(y0 + kWenoAdjSeg - 1, nl + 3);
  double acc[7];
#pragma unroll
  for (int t = 0; t < 7; t++) acc[t] = 0.0;
  for (int c = y0 - 3; c <= y1 + 3; c++) {
    if (c >= 1 && c <= nl) {
      const int cc = base + c * st;
      double P[6], v;
      weno_cell_adj(f, tv, cc, st, eq, advb[cc] * dinv, P, &v);
      if (c >= y0 && c <= y1) V[cc] = v;
      // position c + o receives P[2+o] - P[3+o] (o = -3..3)
      acc[0] = acc[0] - P[0];
#pragma unroll
      for (int o = -2; o <= 2; o++) acc[o + 3] = acc[o + 3] + (P[2 + o] - P[3 + o]);
      acc[6] = acc[6] + P[5];
    }
    if (c - 3 >= y0) {
      double* out = fb + base + (c - 3) * st;
      *out = *out + acc[0];
    }
#pragma unroll
    for (int t = 0; t < 6; t++) acc[t] = acc[t + 1];
    acc[6] = 0.0;
  }
}

// srcs/cal_advu_weno5.f:1-257 reversed for component q and direction dd:
// the difference adjoints gathered along the direction (LDS rows along x,
// marching lines along y and z), then the transport velocity's adjoints
// (V, one scratch array) gathered into the transporting component's
// adjoint: cell y receives V of the cells y + e_d, y + e_d - e_q, y, y - e_q.
static void weno_adj_dir(const Dims d, int q, int dd, double dinv, const double* const* vel,
                         double* const* velb, const double* advb, double* V) {
  const int E[3] = {1, d.sx, d.sx * d.sy};
  const int st = E[dd], eq = E[q];
  const double* f = vel[q];
  const double* tv = vel[dd];
  double* fb = velb[q];
  double* tb = velb[dd];
  if (dd == 0 && d.ni <= 1024) {
    weno_adj_x_kernel<<<dim3(d.nj * d.nk), dim3(256), 6 * d.ni * sizeof(double), stream()>>>(
        d, f, tv, eq, advb, dinv, fb, V);
  } else {
    const int n[3] = {d.ni, d.nj, d.nk};
    const int lines = dd == 0 ? n[1] * n[2] : (dd == 1 ? n[0] * n[2] : n[0] * n[1]);
    const int nseg = (n[dd] + 6 + kWenoAdjSeg - 1) / kWenoAdjSeg;
    const long long threads = (long long)lines * nseg;
    weno_adj_march_kernel<<<dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream()>>>(
        d, dd, f, tv, eq, advb, dinv, fb, V, nseg);
  }
  FEL_CHECK(hipGetLastError());
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int di[3] = {dd == 0, dd == 1, dd == 2}, qi[3] = {q == 0, q == 1, q == 2};
  const int sdi = di[0], sdj = di[1], sdk = di[2], sqi = qi[0], sqj = qi[1], sqk = qi[2];
  launch_box(Box{-1, ni + 1, -1, nj + 1, -1, nk + 1}, [=] __device__(int i, int j, int k) {
    const int y = d.ix(i, j, k);
    // interior cell at y + s*(direction) + t*(component q's direction)
    auto in_at = [=](int s, int t) {
      const int xi = i + s * sdi + t * sqi, xj = j + s * sdj + t * sqj, xk = k + s * sdk + t * sqk;
      return xi >= 1 && xi <= ni && xj >= 1 && xj <= nj && xk >= 1 && xk <= nk;
    };
    double t = 0.0;
    if (in_at(1, 0)) t = t + V[y + st];
    if (in_at(1, -1)) t = t + V[y + st - eq];
    if (in_at(0, 0)) t = t + V[y];
    if (in_at(0, -1)) t = t + V[y - eq];
    tb[y] = tb[y] + t;
  });
}

}  // namespace fel

using namespace fel;

// 1 when the last solu_sor4_b_ formed its coefficients from the viscosity
static int g_last_adj_onfly = -1;

extern "C" {

// How the last solu_sor4_b_ call got its coefficients: 1 formed on the fly
// from the viscosity, 0 read from the 45 arrays, -1 no call yet.
int felgpu_sor_b_onfly_(void) { return g_last_adj_onfly; }

// Adjoint of bnd_periodic_ (srcs/bnd_periodic.f:4-46).
void bnd_periodic_b_(const int* ni, const int* nj, const int* nk, double* qb) {
  adjoint_fp64_only("bnd_periodic_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("bnd_periodic_b_");
  periodic_adj(d, c.io(qb, d.n3, "qb"));
  c.finish();
}

// Adjoint of bnd_neumann_ (srcs/bnd_neumann.f:4-46).
void bnd_neumann_b_(const int* nID, const int* ni, const int* nj, const int* nk, double* qb) {
  adjoint_fp64_only("bnd_neumann_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("bnd_neumann_b_");
  neumann_adj(d, c.io(qb, d.n3, "qb"), nID[kYMinus] < 0, nID[kYPlus] < 0);
  c.finish();
}

// Adjoint of bndu_ (srcs/bndu.f:1-161): the velocity adjoints in place and
// *uwallb incremented.  The slip lengths and phi parameterise the map.
void bndu_b_(const int* nID, const int* ni, const int* nj, const int* nk, double* ukb, double* vkb,
             double* wkb, double* uwallb, const double* dy, const double* l1, const double* l2,
             const double* phi) {
  adjoint_fp64_only("bndu_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const int nw = d.sx * d.sz;
  double* wacc = static_cast<double*>(scratch(25, sizeof(double) * nw));
  launch_flat(nw, [=] __device__(int m) { wacc[m] = 0.0; });
  Call c("bndu_b_");
  double* ub = c.io(ukb, d.n3, "ukb");
  double* vb = c.io(vkb, d.n3, "vkb");
  double* wb = c.io(wkb, d.n3, "wkb");
  const double* dl1 = c.in(l1, d.n3, "l1");
  const double* dl2 = c.in(l2, d.n3, "l2");
  const double* dphi = c.in(phi, d.n3, "phi");
  bndu_adj(d, ub, vb, wb, wacc, *dy, dl1, dl2, dphi, nID[kYMinus] < 0, nID[kYPlus] < 0);
  adjoint_add_sum(wacc, nw, uwallb);
  c.finish();
}

// Adjoint of solu_sor4_ (srcs/solu_sor4.f:1-308): the forward's argument
// list with src_u, src_v, src_w, un, vn, wn replaced by their adjoints
// (src_*b incremented; un/vn/wnb hold the adjoint of the final state on
// entry and of the initial state on return), *uwallb incremented after the
// velocities, and lp -- the forward call's loop variable on exit, from
// felgpu_sor_last_ -- appended: min(lp+1, 21) sweeps are reversed.  The
// densities, coefficients, slip lengths and phi parameterise the map.
void solu_sor4_b_(const int* ipara, const int* ID, const int* nID, const int* ndiv, const int* ni,
                  const int* nj, const int* nk, const int* key, double* sendjb, double* recvjb,
                  const int* nstep, const int* imon_t, const double* dt, const double* rho,
                  const double* rhon, double* c0, double* c1, double* c2, double* c3, double* c4,
                  double* c5, double* c6, double* c7, double* c8, double* c9, double* c10,
                  double* c11, double* c12, double* c13, double* c14, double* c15, double* c16,
                  double* c17, double* c18, double* c19, double* c20, double* c21, double* c22,
                  double* c23, double* c24, double* c25, double* c26, double* c27, double* c28,
                  double* c29, double* c30, double* c31, double* c32, double* c33, double* c34,
                  double* c35, double* c36, double* c37, double* c38, double* c39, double* c40,
                  double* c41, double* c42, double* c43, double* c44, double* src_ub,
                  double* src_vb, double* src_wb, double* unb, double* vnb, double* wnb,
                  double* uwallb, const double* dy, const double* l1, const double* l2,
                  const double* phi, const int* lp) {
  (void)ID; (void)ndiv; (void)key; (void)sendjb; (void)recvjb; (void)nstep; (void)imon_t;
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: solu_sor4_b_: only the single-domain case is supported\n");
    std::abort();
  }
  if (*lp < 0 || *lp > 21) {
    std::fprintf(stderr, "felgpu: solu_sor4_b_: lp = %d is not a sweep index of solu_sor4_ (0..21)\n", *lp);
    std::abort();
  }
  adjoint_fp64_only("solu_sor4_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  double* h[45] = {c0,  c1,  c2,  c3,  c4,  c5,  c6,  c7,  c8,  c9,  c10, c11, c12, c13, c14,
                   c15, c16, c17, c18, c19, c20, c21, c22, c23, c24, c25, c26, c27, c28, c29,
                   c30, c31, c32, c33, c34, c35, c36, c37, c38, c39, c40, c41, c42, c43, c44};
  // on-the-fly coefficients when they are a pending function of the viscosity
  DeferredCoef& dc = deferred_coef();
  bool onfly = dc.active && dc.d.ni == d.ni && dc.d.nj == d.nj && dc.d.nk == d.nk && dc.esize == 8;
  for (int m = 0; m < 45 && onfly; m++) onfly = dc.out_host[m] == (const void*)h[m];
  Call c("solu_sor4_b_");
  SorAdj a;
  a.rho = c.in(rho, d.n3, "rho");
  a.rhon = c.in(rhon, d.n3, "rhon");
  if (onfly) {
    a.rmu = static_cast<const double*>(dc.src_dev);
    a.sc = vis_scales<double>(dc.dxinv, dc.dyinv, dc.dzinv);
    for (int m = 0; m < 45; m++) a.cf[m] = nullptr;
  } else {
    a.rmu = nullptr;
    a.sc = vis_scales<double>(0.0, 0.0, 0.0);
    for (int m = 0; m < 45; m++) a.cf[m] = c.in(h[m], d.n3, "coefficient");
  }
  a.srcb[0] = c.io(src_ub, d.n3, "src_ub");
  a.srcb[1] = c.io(src_vb, d.n3, "src_vb");
  a.srcb[2] = c.io(src_wb, d.n3, "src_wb");
  a.lb[0] = c.io(unb, d.n3, "unb");
  a.lb[1] = c.io(vnb, d.n3, "vnb");
  a.lb[2] = c.io(wnb, d.n3, "wnb");
  const double* dl1 = c.in(l1, d.n3, "l1");
  const double* dl2 = c.in(l2, d.n3, "l2");
  const double* dphi = c.in(phi, d.n3, "phi");
  a.G = static_cast<double*>(scratch(24, sizeof(double) * (size_t)d.n3));
  a.dtinv = 1.0 / *dt;
  const bool lo = nID[kYMinus] < 0, hi = nID[kYPlus] < 0;
  a.jen = hi ? d.nj - 1 : d.nj;
  const int nw = d.sx * d.sz;
  double* wacc = static_cast<double*>(scratch(25, sizeof(double) * nw));
  launch_flat(nw, [=] __device__(int m) { wacc[m] = 0.0; });
  g_last_adj_onfly = onfly ? 1 : 0;
  const int nsw = *lp < 21 ? *lp + 1 : 21;
  // the refresh follows the second colour of each pair in the forward
  for (int s = nsw - 1; s >= 0; s--)
    for (int q = 2; q >= 0; q--)
      for (int p = 3; p >= 0; p--) {
        periodic_adj(d, a.lb[q]);
        bndu_adj(d, a.lb[0], a.lb[1], a.lb[2], wacc, *dy, dl1, dl2, dphi, lo, hi);
        if (q == 0) pair_adj<0>(onfly, d, a, p);
        if (q == 1) pair_adj<1>(onfly, d, a, p);
        if (q == 2) pair_adj<2>(onfly, d, a, p);
      }
  adjoint_add_sum(wacc, nw, uwallb);
  c.finish();
}

// ---------------------------------------------------------------------------
// Explicit momentum terms (adjoint entry points).
// ---------------------------------------------------------------------------

// Adjoint of cpy_ (srcs/cpy.f:14-20): q1b += q2b, q2b = 0.
void cpy_b_(const int* ni, const int* nj, const int* nk, double* q1b, double* q2b) {
  adjoint_fp64_only("cpy_b_");
  if (q1b == q2b) return;
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cpy_b_");
  double* a = c.io(q1b, d.n3, "q1b");
  double* b = c.io(q2b, d.n3, "q2b");
  launch_flat(d.n3, [=] __device__(int m) {
    a[m] = a[m] + b[m];
    b[m] = 0.0;
  });
  c.finish();
}

// Adjoint of calphat_ (srcs/calphat.f:16-22).
void calphat_b_(const int* ni, const int* nj, const int* nk, double* pb, double* pob, double* phatb) {
  adjoint_fp64_only("calphat_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("calphat_b_");
  double* a = c.io(pb, d.n3, "pb");
  double* b = c.io(pob, d.n3, "pob");
  double* h = c.io(phatb, d.n3, "phatb");
  calphat_adj(d, a, b, h);
  c.finish();
}

// Adjoint of cal_grad_prs_ (srcs/cal_grad_prs.f:24-37).
void cal_grad_prs_b_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                     const double* dyinv, const double* dzinv, double* qb, double* qx_ub,
                     double* qy_vb, double* qz_wb) {
  adjoint_fp64_only("cal_grad_prs_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_grad_prs_b_");
  double* q = c.io(qb, d.n3, "qb");
  double* gx = c.io(qx_ub, d.n3, "qx_ub");
  double* gy = c.io(qy_vb, d.n3, "qy_vb");
  double* gz = c.io(qz_wb, d.n3, "qz_wb");
  grad_prs_adj(d, *dxinv, *dyinv, *dzinv, q, gx, gy, gz);
  c.finish();
}

// Adjoint of calsij_ (srcs/calsij.f:28-139); skb holds the six components.
void calsij_b_(const int* ni, const int* nj, const int* nk, const double* dxinv, const double* dyinv,
               const double* dzinv, double* ukb, double* vkb, double* wkb, double* skb) {
  adjoint_fp64_only("calsij_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("calsij_b_");
  double* u = c.io(ukb, d.n3, "ukb");
  double* v = c.io(vkb, d.n3, "vkb");
  double* w = c.io(wkb, d.n3, "wkb");
  double* s = c.io(skb, 6 * (size_t)d.n3, "skb");
  calsij_adj(d, *dxinv, *dyinv, *dzinv, u, v, w, s);
  c.finish();
}

// Adjoint of cal_arith_tau_ (srcs/cal_arith_tau.f:21-118), viscosity fixed.
void cal_arith_tau_b_(const int* ni, const int* nj, const int* nk, double* sb, const double* rmu,
                      double* taub) {
  adjoint_fp64_only("cal_arith_tau_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_arith_tau_b_");
  double* s = c.io(sb, 6 * (size_t)d.n3, "sb");
  const double* m = c.in(rmu, d.n3, "rmu");
  double* t = c.io(taub, 6 * (size_t)d.n3, "taub");
  arith_tau_adj(d, s, m, t);
  c.finish();
}

// Adjoint of cal_div_tensor_ (srcs/cal_div_tensor.f:20-86).
void cal_div_tensor_b_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                       const double* dyinv, const double* dzinv, double* taukb, double* ukb,
                       double* vkb, double* wkb) {
  adjoint_fp64_only("cal_div_tensor_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_div_tensor_b_");
  double* t = c.io(taukb, 6 * (size_t)d.n3, "taukb");
  double* u = c.io(ukb, d.n3, "ukb");
  double* v = c.io(vkb, d.n3, "vkb");
  double* w = c.io(wkb, d.n3, "wkb");
  div_tensor_adj(d, *dxinv, *dyinv, *dzinv, t, u, v, w);
  c.finish();
}

// Adjoint of cal_srcu_ (srcs/cal_srcu.f:51-122), densities fixed; the
// gravity terms are constants of the map.
void cal_srcu_b_(const int* nID, const int* ni, const int* nj, const int* nk, const double* rho_av,
                 const double* dt, const double* grvb, const double* grvp, double* ub, double* vb,
                 double* wb, const double* rho, const double* rhon, double* adv_ub, double* adv_vb,
                 double* adv_wb, double* adv_uob, double* adv_vob, double* adv_wob, double* prs_ub,
                 double* prs_vb, double* prs_wb, double* vis_ub, double* vis_vb, double* vis_wb,
                 double* fst_ub, double* fst_vb, double* fst_wb, double* fst_unb, double* fst_vnb,
                 double* fst_wnb, double* src_ub, double* src_vb, double* src_wb) {
  (void)rho_av; (void)grvb; (void)grvp;
  adjoint_fp64_only("cal_srcu_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const int jen = nID[kYPlus] < 0 ? d.nj - 1 : d.nj;
  Call c("cal_srcu_b_");
  SrcuAdj a;
  double* hv[3] = {ub, vb, wb};
  double* ha[3] = {adv_ub, adv_vb, adv_wb};
  double* hao[3] = {adv_uob, adv_vob, adv_wob};
  double* hp[3] = {prs_ub, prs_vb, prs_wb};
  double* hvis[3] = {vis_ub, vis_vb, vis_wb};
  double* hf[3] = {fst_ub, fst_vb, fst_wb};
  double* hfn[3] = {fst_unb, fst_vnb, fst_wnb};
  double* hs[3] = {src_ub, src_vb, src_wb};
  for (int q = 0; q < 3; q++) {
    a.vel[q] = c.io(hv[q], d.n3, "velocity adjoint");
    a.adv[q] = c.io(ha[q], d.n3, "adv adjoint");
    a.advo[q] = c.io(hao[q], d.n3, "adv_o adjoint");
    a.prs[q] = c.io(hp[q], d.n3, "prs adjoint");
    a.vis[q] = c.io(hvis[q], d.n3, "vis adjoint");
    a.fst[q] = c.io(hf[q], d.n3, "fst adjoint");
    a.fstn[q] = c.io(hfn[q], d.n3, "fst_n adjoint");
    a.src[q] = c.io(hs[q], d.n3, "src adjoint");
  }
  a.rho = c.in(rho, d.n3, "rho");
  a.rhon = c.in(rhon, d.n3, "rhon");
  srcu_adj(d, jen, 1.0 / *dt, a);
  c.finish();
}

// Adjoint of cal_advu_weno5_ (srcs/cal_advu_weno5.f:1-257).  Not linear:
// u, v, w are the forward's inputs, each followed by its adjoint
// (incremented); the adv adjoints are consumed.
void cal_advu_weno5_b_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                       const double* dyinv, const double* dzinv, const double* u, double* ub,
                       const double* v, double* vb, const double* w, double* wb, double* adv_ub,
                       double* adv_vb, double* adv_wb) {
  adjoint_fp64_only("cal_advu_weno5_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_advu_weno5_b_");
  const double* vel[3] = {c.in(u, d.n3, "u"), c.in(v, d.n3, "v"), c.in(w, d.n3, "w")};
  double* velb[3] = {c.io(ub, d.n3, "ub"), c.io(vb, d.n3, "vb"), c.io(wb, d.n3, "wb")};
  double* advb[3] = {c.io(adv_ub, d.n3, "adv_ub"), c.io(adv_vb, d.n3, "adv_vb"),
                     c.io(adv_wb, d.n3, "adv_wb")};
  double* V = static_cast<double*>(scratch(26, sizeof(double) * (size_t)d.n3));
  const double dinv[3] = {*dxinv, *dyinv, *dzinv};
  for (int q = 2; q >= 0; q--)
    for (int dd = 2; dd >= 0; dd--) weno_adj_dir(d, q, dd, dinv[dd], vel, velb, advb[q], V);
  for (int q = 0; q < 3; q++) {
    double* a = advb[q];
    launch_flat(d.n3, [=] __device__(int m) { a[m] = 0.0; });
  }
  c.finish();
}

}  // extern "C"
