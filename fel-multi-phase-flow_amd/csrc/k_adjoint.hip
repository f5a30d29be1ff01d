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
//   launches: the first, over the colour's cells, adds x_b/den to src_b and
//   stores G = x_b/den 0.5/r00; the second, over every cell that a colour
//   cell reads, gathers sum a_m G of the colour cells that read it into its
//   adjoint and clears the colour cells' own adjoints (a cell never reads a
//   cell of its own colour, so the gather reads no adjoint it writes).
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
static void add_sum(const double* a, int n, double* out) {
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

// First launch of a colour: src_b += x_b/den and G = x_b/den 0.5/r00 at the
// colour's cells of component P.
template <int P, bool ONFLY>
static void colour_scale(const Dims d, const SorAdj a, int is, int js, int ks) {
  const int jend = P == 1 ? a.jen : d.nj;
  const int hx = (d.ni - is) / 2 + 1, hy = jend >= js ? (jend - js) / 2 + 1 : 0,
            hz = (d.nk - ks) / 2 + 1;
  if (hy <= 0) return;
  launch_flat(hx * hy * hz, [=] __device__(int m) {
    const int ii = m % hx, t = m / hx, jj = t % hy, kk = t / hy;
    const int c = d.ix(is + 2 * ii, js + 2 * jj, ks + 2 * kk);
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

// Second launch of a colour: every cell y in [0,n+1]^3 gathers the weights
// of the colour's cells of component P that read it, into the adjoints of
// u, v and w; the colour cells' own adjoints are cleared.
template <int P, bool ONFLY>
static void colour_gather(const Dims d, const SorAdj a, int is, int js, int ks) {
  const int ni = d.ni, nk = d.nk, jend = P == 1 ? a.jen : d.nj;
  launch_box(Box{0, d.ni + 1, 0, d.nj + 1, 0, d.nk + 1}, [=] __device__(int i, int j, int k) {
    const int y = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    // G of the colour cell at offset (di,dj,dk), 0 when that cell is not one
    auto in_c = [=](int di, int dj, int dk) {
      const int xi = i + di, xj = j + dj, xk = k + dk;
      return ((xi ^ is) & 1) == 0 && ((xj ^ js) & 1) == 0 && ((xk ^ ks) & 1) == 0 && xi >= 1 &&
             xi <= ni && xj >= 1 && xj <= jend && xk >= 1 && xk <= nk;
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

template <int P>
static void colour_adj(bool onfly, const Dims d, const SorAdj& a, int col) {
  const int is = kIsor[col], js = kJsor[col], ks = kKsor[col];
  if (onfly) {
    colour_scale<P, true>(d, a, is, js, ks);
    colour_gather<P, true>(d, a, is, js, ks);
  } else {
    colour_scale<P, false>(d, a, is, js, ks);
    colour_gather<P, false>(d, a, is, js, ks);
  }
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
  add_sum(wacc, nw, uwallb);
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
  for (int s = nsw - 1; s >= 0; s--)
    for (int q = 2; q >= 0; q--)
      for (int col = 7; col >= 0; col--) {
        if (col % 2 == 1) {
          periodic_adj(d, a.lb[q]);
          bndu_adj(d, a.lb[0], a.lb[1], a.lb[2], wacc, *dy, dl1, dl2, dphi, lo, hi);
        }
        if (q == 0) colour_adj<0>(onfly, d, a, col);
        if (q == 1) colour_adj<1>(onfly, d, a, col);
        if (q == 2) colour_adj<2>(onfly, d, a, col);
      }
  add_sum(wacc, nw, uwallb);
  c.finish();
}

}  // extern "C"
