// k_adjoint.hip -- adjoints (reverse mode) of the ghost fills, the
// semi-implicit viscous solve, the explicit momentum terms, WENO5
// advection, the material properties and the surface-tension chain, and
// the all-argument adjoints (<name>_bp_) that also form the adjoints of the
// densities, viscosity, coefficients, slip lengths and phi.  (The phase
// transport's are in k_adjoint_vof.hip, the projection's in k_poisson.hip.)
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
// * solu_sor4_bp_ / solu_sor4_bpv_: the same reverse with the parameter
//   adjoints: the sweeps are recomputed once into an undo log (the values
//   each colour pair and each refresh overwrite), restored pair by pair in
//   the reverse, so each pair's coefficient and face-density adjoints and
//   each refresh's slip-length and phi adjoints are formed at the state
//   the forward saw.
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

// Parameter adjoints of bndu (bndu_bp_): the call's input velocities on
// the rows next to the walls (rows: u at j = 1, u at j = nj, w at j = 1, w
// at j = nj, one value per (i,k) column each), uwall, and the adjoints of
// the slip lengths and phi; pa holds the phi_av adjoint of each wall cell
// (bottom then top, one value per (i,k) column).  l1b == nullptr: none
// formed.
struct BnduParam {
  const double* rows;
  double uwall;
  double *l1b, *l2b, *phib, *pa;
};

// Adjoint of the slip coefficients of wall cell (i, jw, k) given c1_b and
// c2_b: ls = 1/(phi_av/l1 + (1-phi_av)/l2), c1 = 2dy/(2ls+dy), c2 =
// (2ls-dy)/(2ls+dy).  l1_b, l2_b of the cell incremented; phi_av's adjoint
// returned (it belongs to two cells, gathered by bndu_phi_gather).
__device__ __forceinline__ double wall_coef_b(const Dims& d, int i, int jw, int k, double dy,
                                              const double* l1, const double* l2,
                                              const double* phi, double c1b, double c2b,
                                              double* l1b, double* l2b) {
  const int ip = (i + 1) % (d.ni + 3);
  const int cw = d.ix(i, jw, k);
  const double phi_av = (phi[cw] + phi[d.ix(ip, jw, k)]) / 2.0;
  const double inv1 = phi_av / l1[cw], inv2 = (1.0 - phi_av) / l2[cw];
  const double ls = 1.0 / (inv1 + inv2);
  const double dd = 2.0 * ls + dy;
  const double lsb = c1b * (-4.0 * dy / (dd * dd)) + c2b * (4.0 * dy / (dd * dd));
  const double sb = -lsb * ls * ls;
  l1b[cw] = l1b[cw] - sb * phi_av / (l1[cw] * l1[cw]);
  l2b[cw] = l2b[cw] - sb * (1.0 - phi_av) / (l2[cw] * l2[cw]);
  return sb / l1[cw] - sb / l2[cw];
}

// phi_b of the wall rows from the phi_av adjoints: phi_av of column i is
// the mean of phi at i and (i+1) mod (ni+3), so cell i receives half of its
// own column's and of the columns i' with (i'+1) mod (ni+3) = i (i-1 for
// i in [-1, ni+2]; also ni+2 for i = 0 and ni+3 for i = 1).
static void bndu_phi_gather(const Dims d, const BnduParam bp, bool lo, bool hi) {
  const int ni = d.ni, nj = d.nj, sx = d.sx, nw = d.sx * d.sz;
  launch_flat(nw, [=] __device__(int m) {
    const int i = m % sx - 2, k = m / sx - 2;
    for (int t = 0; t < 2; t++) {
      if ((t == 0 && !lo) || (t == 1 && !hi)) continue;
      const double* pa = bp.pa + t * nw;
      double s = pa[m];
      if (i >= -1) s = s + pa[m - 1];
      if (i == 0) s = s + pa[(ni + 4) + (k + 2) * sx];
      if (i == 1) s = s + pa[(ni + 5) + (k + 2) * sx];
      const int c = d.ix(i, t == 0 ? 1 : nj, k);
      bp.phib[c] = bp.phib[c] + 0.5 * s;
    }
  });
}

// srcs/bndu.f:29-158 transposed: the z fills, the x fills (with the shifted
// sources of u in x and of w in z), then the wall rows, top first.  `wacc`
// (one value per (i,k) wall column) gathers the uwall weights: -c1 x u_b of
// the bottom ghost and +c1 x u_b of the top one.  With bp.l1b set, the
// adjoints of the slip lengths and phi are formed too: a wall row's ghost
// u0 = c1 (-+uwall) + c2 u1 (w0 = c2 w1), reversed with adjoint A (the rows
// beyond folded in), gives c1_b = A (-+uwall) and c2_b = A_u u1 + A_w w1.
static void bndu_adj(const Dims d, double* u, double* v, double* w, double* wacc, double dy,
                     const double* l1, const double* l2, const double* phi, bool lo, bool hi,
                     const BnduParam bp = BnduParam{nullptr, 0.0, nullptr, nullptr, nullptr,
                                                    nullptr}) {
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
  const int nw = sx * d.sz;
  // wall rows (bndu.f:29-98), every (i,k) column, top wall first
  launch_flat(nw, [=] __device__(int m) {
    const int i = m % sx - 2, k = m / sx - 2;
    auto jy = [=](int j) { return d.ix(i, j, k); };
    double acc = 0.0;
    if (hi) {
      const WallCoef<double> cf = wall_coef(d, i, nj, k, dy, l1, l2, phi);
      double c1b = 0.0, c2b = 0.0;
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
        if (bp.l1b) {
          if (t == 0) {
            c2b = c2b + b1 * bp.rows[3 * nw + m];
          } else {
            c1b = c1b + b1 * bp.uwall;
            c2b = c2b + b1 * bp.rows[nw + m];
          }
        }
        q[jy(nj + 1)] = 0.0;
      }
      if (bp.l1b)
        bp.pa[m + nw] = wall_coef_b(d, i, nj, k, dy, l1, l2, phi, c1b, c2b, bp.l1b, bp.l2b);
      // v: v(nj) = 0, v(nj+g) = -v(nj-g)
      for (int g = 3; g >= 1; g--) {
        v[jy(nj - g)] = v[jy(nj - g)] - v[jy(nj + g)];
        v[jy(nj + g)] = 0.0;
      }
      v[jy(nj)] = 0.0;
    }
    if (lo) {
      const WallCoef<double> cf = wall_coef(d, i, 1, k, dy, l1, l2, phi);
      double c1b = 0.0, c2b = 0.0;
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
        if (bp.l1b) {
          if (t == 0) {
            c2b = c2b + b1 * bp.rows[2 * nw + m];
          } else {
            c1b = c1b - b1 * bp.uwall;
            c2b = c2b + b1 * bp.rows[m];
          }
        }
        q[jy(0)] = 0.0;
      }
      if (bp.l1b) bp.pa[m] = wall_coef_b(d, i, 1, k, dy, l1, l2, phi, c1b, c2b, bp.l1b, bp.l2b);
      // v: v(-2) = -v(2), v(-1) = -v(1), v(0) = 0
      v[jy(0)] = 0.0;
      v[jy(1)] = v[jy(1)] - v[jy(-1)];
      v[jy(-1)] = 0.0;
      v[jy(2)] = v[jy(2)] - v[jy(-2)];
      v[jy(-2)] = 0.0;
    }
    if (wacc) wacc[m] = wacc[m] + acc;
  });
  if (bp.l1b) bndu_phi_gather(d, bp, lo, hi);
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


// ---------------------------------------------------------------------------
// Parameter adjoints (<name>_bp_): what <name>_b_ does, plus the adjoints of
// the arrays it takes as fixed (densities, viscosity, coefficients, slip
// lengths, phi), incremented.
// ---------------------------------------------------------------------------

// Terms of component P's Gauss-Seidel update (srcs/solu_sor4.f:137-152,
// 186-201, 236-251) in the order the forward sums them: coefficient index
// within the component's 15, neighbour component, offset.
template <int P>
__device__ __forceinline__ void sor_terms(const Dims& d, int* idx, int* comp, int* off) {
  const int X = 1, Y = d.sx, Z = d.sx * d.sy;
  const int own_i[6] = {4, 2, 0, 1, 3, 5};
  const int own_o[6] = {-Z, -Y, -X, X, Y, Z};
#pragma unroll
  for (int t = 0; t < 6; t++) {
    idx[t] = own_i[t];
    comp[t] = P;
    off[t] = own_o[t];
  }
  if (P == 0) {
    const int i8[8] = {7, 8, 9, 10, 11, 12, 13, 14}, c8[8] = {1, 1, 1, 1, 2, 2, 2, 2};
    const int o8[8] = {-Y, X - Y, 0, X, -Z, X - Z, 0, X};
#pragma unroll
    for (int t = 0; t < 8; t++) { idx[6 + t] = i8[t]; comp[6 + t] = c8[t]; off[6 + t] = o8[t]; }
  } else if (P == 1) {
    const int i8[8] = {7, 9, 8, 10, 11, 12, 13, 14}, c8[8] = {0, 0, 0, 0, 2, 2, 2, 2};
    const int o8[8] = {-X, 0, -X + Y, Y, -Z, Y - Z, 0, Y};
#pragma unroll
    for (int t = 0; t < 8; t++) { idx[6 + t] = i8[t]; comp[6 + t] = c8[t]; off[6 + t] = o8[t]; }
  } else {
    const int i8[8] = {7, 9, 8, 10, 11, 13, 12, 14}, c8[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    const int o8[8] = {-X, 0, -X + Z, Z, -Y, 0, -Y + Z, Z};
#pragma unroll
    for (int t = 0; t < 8; t++) { idx[6 + t] = i8[t]; comp[6 + t] = c8[t]; off[6 + t] = o8[t]; }
  }
}

// The cells of component P of colour pair p (both colours), as in
// colour_scale: f(m, c) for each, m the cell's place in the pair's
// enumeration (0 .. count-1, the same in every launch).
template <int P, class F>
static void for_pair_cells(const Dims d, int jen, int p, F f) {
  const int is = kIsor[2 * p], js = kJsor[2 * p], ks = kKsor[2 * p];
  const int jend = P == 1 ? jen : d.nj;
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
    f(m, c);
  });
}

// The rows next to the walls of u and w (j = 1 and j = nj, every (i,k)
// column) into rows[4 nw], the layout BnduParam reads.
static void wall_rows(const Dims d, const double* u, const double* w, double* rows) {
  const int nj = d.nj, sx = d.sx, nw = d.sx * d.sz;
  launch_flat(nw, [=] __device__(int m) {
    const int i = m % sx - 2, k = m / sx - 2;
    rows[m] = u[d.ix(i, 1, k)];
    rows[nw + m] = u[d.ix(i, nj, k)];
    rows[2 * nw + m] = w[d.ix(i, 1, k)];
    rows[3 * nw + m] = w[d.ix(i, nj, k)];
  });
}

// Arrays of the all-argument Gauss-Seidel adjoint.
struct SorParam {
  const double* src[3];  // primal sources
  double* cfb[45];       // coefficient adjoints (coefficient arrays)
  double* fb[18];        // face-viscosity adjoints, 6 per component (viscosity form)
  double *rhob, *rhonb;
  double dinv[3];        // dxinv, dyinv, dzinv (viscosity form)
};

// The 15 coefficients of component P at cell c: read from the arrays, or
// (V) formed from the viscosity with the expressions of
// cal_arith_coef_vis (the same bits).
template <int P, bool V>
__device__ __forceinline__ void coefs_at(const Dims& d, const SorAdj& A, int c, double* cf) {
  if (V) {
    coef_of<P>(d, A.sc, A.rmu, c, cf);
  } else {
#pragma unroll
    for (int t = 0; t < 15; t++) cf[t] = A.cf[15 * P + t][c];
  }
}

// The value the Gauss-Seidel update (srcs/solu_sor4.f:131-175) gives cell c
// of component P from the state x0, x1, x2.
template <int P, bool V>
__device__ __forceinline__ double pair_value(const Dims& d, const SorAdj& A, const double* src,
                                             const double* x0, const double* x1,
                                             const double* x2, int c) {
  int idx[14], comp[14], off[14];
  sor_terms<P>(d, idx, comp, off);
  double cf[15];
  coefs_at<P, V>(d, A, c, cf);
  const double* xs[3] = {x0, x1, x2};
  const int F = P == 0 ? 1 : (P == 1 ? d.sx : d.sx * d.sy);
  const double r00 = (A.rho[c] + A.rho[c + F] + A.rhon[c] + A.rhon[c + F]) * 0.25;
  double sum = 0.0;
#pragma unroll
  for (int t = 0; t < 14; t++) sum = sum + cf[idx[t]] * xs[comp[t]][c + off[t]];
  const double visn = sum / r00;
  const double xo = xs[P][c];
  const double du = (visn * 0.5 + src[c]) / (A.dtinv + cf[6] / r00 * 0.5) - xo;
  return xo + du;
}

// Forward update of colour pair p of component P on x: the pair's cells
// read no cell of the pair, so one pass.
template <int P, bool V>
static void pair_fwd(const Dims d, const SorAdj& a, const SorParam& sp, double* const* x, int p) {
  const double* src = sp.src[P];
  double* x0 = x[0];
  double* x1 = x[1];
  double* x2 = x[2];
  const SorAdj A = a;
  for_pair_cells<P>(d, a.jen, p, [=] __device__(int, int c) {
    const double v = pair_value<P, V>(d, A, src, x0, x1, x2, c);
    double* xp = P == 0 ? x0 : (P == 1 ? x1 : x2);
    xp[c] = v;
  });
}

// The undo log of the all-argument solve adjoint: the forward pass keeps,
// for every colour pair, the values its update overwrites (the pair's cells
// of the swept component, in for_pair_cells order) and the values the
// refresh after it overwrites (every cell outside [1,ni] x [1,nj-1] x
// [1,nk] of u, v and w: the ghosts, and v's top wall row); the reverse
// restores them in reverse order, so the working state is, in turn, the
// state before each refresh and before each pair -- exactly, bit for bit.
template <int P>
static void pair_save(const Dims d, int jen, int p, const double* x, double* buf) {
  for_pair_cells<P>(d, jen, p, [=] __device__(int m, int c) { buf[m] = x[c]; });
}
template <int P>
static void pair_restore(const Dims d, int jen, int p, double* x, const double* buf) {
  for_pair_cells<P>(d, jen, p, [=] __device__(int m, int c) { x[c] = buf[m]; });
}

// Offsets of the cells outside [1,ni] x [1,nj-1] x [1,nk] (every cell a
// refresh writes, and a few it does not), on the device; count in *n.
static const int* shell_cells(const Dims& d, int* n) {
  static int cached[3] = {-1, -1, -1};
  static int count = 0;
  int* dev = static_cast<int*>(scratch(47, sizeof(int) * (size_t)d.n3));
  if (cached[0] != d.ni || cached[1] != d.nj || cached[2] != d.nk) {
    std::vector<int> h;
    h.reserve((size_t)d.n3 - (size_t)d.ni * (size_t)(d.nj - 1) * (size_t)d.nk);
    for (int k = -2; k <= d.nk + 3; k++)
      for (int j = -2; j <= d.nj + 3; j++)
        for (int i = -2; i <= d.ni + 3; i++)
          if (i < 1 || i > d.ni || j < 1 || j > d.nj - 1 || k < 1 || k > d.nk) h.push_back(d.ix(i, j, k));
    count = (int)h.size();
    FEL_CHECK(hipMemcpyAsync(dev, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, stream()));
    FEL_CHECK(hipStreamSynchronize(stream()));
    cached[0] = d.ni;
    cached[1] = d.nj;
    cached[2] = d.nk;
  }
  *n = count;
  return dev;
}

static void shell_save(const int* cells, int n, double* const* x, double* buf) {
  double *x0 = x[0], *x1 = x[1], *x2 = x[2];
  launch_flat(3 * n, [=] __device__(int m) {
    const int q = m / n, t = m - q * n;
    const double* xq = q == 0 ? x0 : (q == 1 ? x1 : x2);
    buf[m] = xq[cells[t]];
  });
}
static void shell_restore(const int* cells, int n, double* const* x, const double* buf) {
  double *x0 = x[0], *x1 = x[1], *x2 = x[2];
  launch_flat(3 * n, [=] __device__(int m) {
    const int q = m / n, t = m - q * n;
    double* xq = q == 0 ? x0 : (q == 1 ? x1 : x2);
    xq[cells[t]] = buf[m];
  });
}

// The adjoints of component P's six face viscosities (f = xm, xp, ym, yp,
// zm, zp) from its 15 coefficient adjoints cb, the per-cell map
// coef_vis_adj applies: d2_a (cb_f + cb_p), doubled on P's own axis, plus
// off it dinv_P dinv_a (cb_A - cb_B) or (-cb_A + cb_B) for the n-th cross
// pair (7+2n, 8+2n).
template <int P>
__device__ __forceinline__ void face_adj_of(const double* dv, const double* cb, double* fa) {
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const int ax = f / 2;
    const double base = dv[ax] * dv[ax] * (cb[f] + cb[6]);
    if (ax == P) {
      fa[f] = 2.0 * base;
    } else {
      const int n = f < 2 * P ? f : f - 2;
      const double cr = (n % 2 == 0) ? (cb[7 + 2 * n] - cb[8 + 2 * n]) : (-cb[7 + 2 * n] + cb[8 + 2 * n]);
      fa[f] = base + dv[P] * dv[ax] * cr;
    }
  }
}

// First launch of colour pair p of component P in the all-argument
// reverse: colour_scale's src_b += x_b/den and G = x_b/den 0.5/r00, and the
// parameter adjoints (the pair cells' state adjoints still in place): for
// x = (0.5 S/r00 + src)/den, den = dtinv + 0.5 a_p/r00, S = sum a_t y_t over
// the state xs before the pair, g = x_b/den and h = 0.5 g/r00:
// a_t_b += h y_t, a_p_b -= h x, r00_b = g (0.5 x a_p - 0.5 S)/r00^2 spread
// over rho and rhon at c and c+F.  V: the coefficient adjoints go straight
// into the cell's six face-viscosity adjoints (face_adj_of) instead of the
// 15 coefficient arrays.  No two cells of a pair are F apart, so the
// scatters do not collide.
template <int P, bool V>
static void pair_param(const Dims d, const SorAdj& a, const SorParam& sp,
                       const double* const* xs, int p) {
  const double *rho = a.rho, *rhon = a.rhon, *src = sp.src[P];
  const double* lb = a.lb[P];
  const double dtinv = a.dtinv;
  const double *s0 = xs[0], *s1 = xs[1], *s2 = xs[2];
  const SorAdj A = a;
  const SorParam S = sp;
  for_pair_cells<P>(d, a.jen, p, [=] __device__(int, int c) {
    int idx[14], comp[14], off[14];
    sor_terms<P>(d, idx, comp, off);
    double cf[15];
    coefs_at<P, V>(d, A, c, cf);
    const double* st[3] = {s0, s1, s2};
    const int F = P == 0 ? 1 : (P == 1 ? d.sx : d.sx * d.sy);
    const double r00 = (rho[c] + rho[c + F] + rhon[c] + rhon[c + F]) * 0.25;
    const double ap = cf[6];
    const double den = dtinv + ap / r00 * 0.5;
    double y[14];
    double sum = 0.0;
#pragma unroll
    for (int t = 0; t < 14; t++) {
      y[t] = st[comp[t]][c + off[t]];
      sum = sum + cf[idx[t]] * y[t];
    }
    const double x = (sum / r00 * 0.5 + src[c]) / den;
    const double g = lb[c] / den;
    const double h = g * 0.5 / r00;
    // colour_scale's work, fused: src_b += g, G = h
    A.srcb[P][c] = A.srcb[P][c] + g;
    A.G[c] = h;
    if (V) {
      double cb[15];
#pragma unroll
      for (int t = 0; t < 14; t++) cb[idx[t]] = h * y[t];
      cb[6] = -(h * x);
      double fa[6];
      face_adj_of<P>(S.dinv, cb, fa);
#pragma unroll
      for (int f = 0; f < 6; f++) {
        double* fb = S.fb[6 * P + f];
        fb[c] = fb[c] + fa[f];
      }
    } else {
#pragma unroll
      for (int t = 0; t < 14; t++) {
        double* cb = S.cfb[15 * P + idx[t]];
        cb[c] = cb[c] + h * y[t];
      }
      double* cbp = S.cfb[15 * P + 6];
      cbp[c] = cbp[c] - h * x;
    }
    const double rb = 0.25 * g * (0.5 * x * ap - 0.5 * sum) / (r00 * r00);
    S.rhob[c] = S.rhob[c] + rb;
    S.rhob[c + F] = S.rhob[c + F] + rb;
    S.rhonb[c] = S.rhonb[c] + rb;
    S.rhonb[c + F] = S.rhonb[c + F] + rb;
  });
}

// The viscosity adjoint from face-viscosity adjoints fb (6 per component,
// nonzero only where the coefficients exist): each cell gathers the face
// adjoints of the cells whose faces include it (the geometry of
// coef_vis_adj).
static void face_gather(const Dims d, const double* const* fb, double* rmub) {
  struct Fb { const double* a[18]; };
  Fb B;
  for (int m = 0; m < 18; m++) B.a[m] = fb[m];
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    auto fa = [&](int q, int f, int ci, int cj, int ck) -> double {
      if (ci < -1 || ci > ni + 2 || cj < -1 || cj > nj + 2 || ck < -1 || ck > nk + 2) return 0.0;
      return B.a[6 * q + f][d.ix(ci, cj, ck)];
    };
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
      for (int f = 0; f < 6; f++) {
        const int ax = f / 2, side = f % 2;
        int e_a[3] = {0, 0, 0}, e_q[3] = {0, 0, 0};
        e_a[ax] = 1;
        e_q[q] = 1;
        if (ax == q) {
          acc = acc + fa(q, f, i - side * e_a[0], j - side * e_a[1], k - side * e_a[2]);
        } else {
          const int o[3] = {side ? 0 : -e_a[0], side ? 0 : -e_a[1], side ? 0 : -e_a[2]};
#pragma unroll
          for (int al = 0; al < 2; al++)
#pragma unroll
            for (int be = 0; be < 2; be++)
              acc = acc + 0.25 * fa(q, f, i - o[0] - al * e_q[0] - be * e_a[0],
                                    j - o[1] - al * e_q[1] - be * e_a[1],
                                    k - o[2] - al * e_q[2] - be * e_a[2]);
        }
      }
    const int c = d.ix(i, j, k);
    rmub[c] = rmub[c] + acc;
  });
}

static void copy_dev(double* dst, const double* src, int n) {
  launch_flat(n, [=] __device__(int m) { dst[m] = src[m]; });
}

// Viscosity adjoint of the 45 coefficients (srcs/cal_arith_coef_vis.f):
// each coefficient of component q at cell c is a multiple of one of six
// face viscosities (face f: axis f/2, side f%2), the cell value of c or of
// its neighbour along q's own axis, or along the other axes the mean of
// the four cells {c+o, c+o+e_q, c+o+e_a, c+o+e_q+e_a} (o = -e_a on the
// minus side); a_p is the weighted sum of the six.  The face adjoint of
// (q, f) at c is d2_a (a_f + a_p) (doubled on q's own axis) plus, off the
// own axis, dinv_q dinv_a (a_A - a_B) or (-a_A + a_B) for the n-th
// cross pair (7+2n, 8+2n).  Every cell gathers the face adjoints of the
// cells whose faces include it (coefficients exist on [-1, n+2]^3); the
// coefficient adjoints are then cleared whole.
static void coef_vis_adj(const Dims d, double dxinv, double dyinv, double dzinv, double* rmub,
                         double* const* cb) {
  struct Cb { const double* a[45]; };
  Cb C;
  for (int m = 0; m < 45; m++) C.a[m] = cb[m];
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const double dinv[3] = {dxinv, dyinv, dzinv};
  const double di0 = dinv[0], di1 = dinv[1], di2 = dinv[2];
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const double dv[3] = {di0, di1, di2};
    auto fa = [&](int q, int f, int ci, int cj, int ck) -> double {
      if (ci < -1 || ci > ni + 2 || cj < -1 || cj > nj + 2 || ck < -1 || ck > nk + 2) return 0.0;
      const int c = d.ix(ci, cj, ck);
      const double* const* a = C.a + 15 * q;
      const int ax = f / 2;
      const double base = dv[ax] * dv[ax] * (a[f][c] + a[6][c]);
      if (ax == q) return 2.0 * base;
      const int n = f < 2 * q ? f : f - 2;
      const int ia = 7 + 2 * n, ib = 8 + 2 * n;
      const double cr = (n % 2 == 0) ? (a[ia][c] - a[ib][c]) : (-a[ia][c] + a[ib][c]);
      return base + dv[q] * dv[ax] * cr;
    };
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
      for (int f = 0; f < 6; f++) {
        const int ax = f / 2, side = f % 2;
        int e_a[3] = {0, 0, 0}, e_q[3] = {0, 0, 0};
        e_a[ax] = 1;
        e_q[q] = 1;
        if (ax == q) {
          // single cell c + side e_a
          acc = acc + fa(q, f, i - side * e_a[0], j - side * e_a[1], k - side * e_a[2]);
        } else {
          const int o[3] = {side ? 0 : -e_a[0], side ? 0 : -e_a[1], side ? 0 : -e_a[2]};
#pragma unroll
          for (int al = 0; al < 2; al++)
#pragma unroll
            for (int be = 0; be < 2; be++)
              acc = acc + 0.25 * fa(q, f, i - o[0] - al * e_q[0] - be * e_a[0],
                                    j - o[1] - al * e_q[1] - be * e_a[1],
                                    k - o[2] - al * e_q[2] - be * e_a[2]);
        }
      }
    const int c = d.ix(i, j, k);
    rmub[c] = rmub[c] + acc;
  });
  for (int m = 0; m < 45; m++) {
    double* a = cb[m];
    launch_flat(d.n3, [=] __device__(int e) { a[e] = 0.0; });
  }
}

// Viscosity adjoint of cal_arith_tau (before arith_tau_adj consumes tau_b):
// the normal stresses read rmu at their cell, each shear stress the mean of
// four cells; every cell gathers 2 s tau_b of the stresses that read it.
static void arith_tau_param(const Dims d, const double* s, double* rmub, const double* taub) {
  const SijBoxes S = sij_boxes(d);
  const size_t n3 = (size_t)d.n3;
  launch_box(Box{-2, d.ni + 3, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    double acc = 0.0;
#pragma unroll
    for (int l = 0; l < 3; l++) acc = acc + taub[l * n3 + c] * s[l * n3 + c] * 2.0;
    // shear l reads (c, c+a, c+b, c+a+b): sources c, c-a, c-b, c-a-b
    const int ai[3] = {1, 1, 0}, aj[3] = {0, 0, 1}, ak[3] = {0, 0, 0};
    const int bi[3] = {0, 0, 0}, bj[3] = {1, 0, 0}, bk[3] = {0, 1, 1};
    const int ao[3] = {X, X, Y}, bo[3] = {Y, Z, Z};
#pragma unroll
    for (int l = 3; l < 6; l++) {
      const int t = l - 3;
#pragma unroll
      for (int al = 0; al < 2; al++)
#pragma unroll
        for (int be = 0; be < 2; be++) {
          const int si = i - al * ai[t] - be * bi[t], sj = j - al * aj[t] - be * bj[t],
                    sk = k - al * ak[t] - be * bk[t];
          if (!S.b[l].has(si, sj, sk)) continue;
          const int sc = c - al * ao[t] - be * bo[t];
          acc = acc + 0.25 * taub[l * n3 + sc] * s[l * n3 + sc] * 2.0;
        }
    }
    rmub[c] = rmub[c] + acc;
  });
}

// Density adjoints of cal_srcu (before srcu_adj consumes src_b): the face
// density r = (rho + rho(+F) + rhon + rhon(+F))/4 of component q enters as
// -prs/r + 0.5 vis/r, so r_b = src_b (prs - 0.5 vis)/r^2; every cell
// gathers a quarter of r_b of the faces at c and c-F of each component.
static void srcu_param(const Dims d, int jen, const double* const* prs, const double* const* vis,
                       const double* const* srcb, const double* rho, const double* rhon,
                       double* rhob, double* rhonb) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const double *p0 = prs[0], *p1 = prs[1], *p2 = prs[2];
  const double *v0 = vis[0], *v1 = vis[1], *v2 = vis[2];
  const double *s0 = srcb[0], *s1 = srcb[1], *s2 = srcb[2];
  launch_box(Box{1, ni + 1, 1, nj + 1, 1, nk + 1}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const int F[3] = {1, d.sx, d.sx * d.sy};
    const double* P[3] = {p0, p1, p2};
    const double* V[3] = {v0, v1, v2};
    const double* Sb[3] = {s0, s1, s2};
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int jend = q == 1 ? jen : nj;
      const int di[3] = {q == 0, q == 1, q == 2};
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const int ci = i - t * di[0], cj = j - t * di[1], ck = k - t * di[2];
        if (ci < 1 || ci > ni || cj < 1 || cj > jend || ck < 1 || ck > nk) continue;
        const int cc = c - t * F[q];
        const double r = (rho[cc] + rho[cc + F[q]] + rhon[cc] + rhon[cc + F[q]]) * 0.25;
        acc = acc + 0.25 * Sb[q][cc] * (P[q][cc] - 0.5 * V[q][cc]) / (r * r);
      }
    }
    rhob[c] = rhob[c] + acc;
    rhonb[c] = rhonb[c] + acc;
  });
}

// srcs/calrn.f:1-95 reversed (normals rn = r/(|r| + verysmall) of the mean
// r of eight vertex gradients, on the box of staggering (ip,jp,kp), zero
// elsewhere): first r_b per box cell into R (d|r|/dr taken as 0 at r = 0),
// then every vertex gathers r_b/8 of the box cells whose eight vertices
// include it (a vertex counted twice when a staggering offset is 0, as the
// forward sums it twice); rn_b is consumed whole.
static void calrn_adj(const Dims d, double dxinv, double dyinv, double dzinv, int ip, int jp,
                      int kp, const double* const* g, double* const* gb, double* rnb, double* R) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const size_t n3 = (size_t)d.n3;
  const double verysmall = sqrt(dxinv * dxinv + dyinv * dyinv + dzinv * dzinv) * 1.0e-30;
  const double *g0 = g[0], *g1 = g[1], *g2 = g[2];
  launch_box(Box{-2, ni + 3 - ip, -2, nj + 3 - jp, -2, nk + 3 - kp},
             [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = ip, Y = jp * d.sx, Z = kp * d.sx * d.sy;
    const int v[8] = {c, c + X, c + Y, c + X + Y, c + Z, c + X + Z, c + Y + Z, c + X + Y + Z};
    const double* a[3] = {g0, g1, g2};
    double r[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const double* p = a[q];
      r[q] = (+p[v[0]] + p[v[1]] + p[v[2]] + p[v[3]] + p[v[4]] + p[v[5]] + p[v[6]] + p[v[7]]) *
             0.125;
    }
    const double nr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const double inv = 1.0 / (nr + verysmall);
    const int o = c + X + Y + Z;
    const double b0 = rnb[o], b1 = rnb[n3 + o], b2 = rnb[2 * n3 + o];
    const double nrb = -(b0 * r[0] + b1 * r[1] + b2 * r[2]) * inv * inv;
    const double gr = nr > 0.0 ? nrb / nr : 0.0;
    R[c] = b0 * inv + gr * r[0];
    R[n3 + c] = b1 * inv + gr * r[1];
    R[2 * n3 + c] = b2 * inv + gr * r[2];
  });
  double *h0 = gb[0], *h1 = gb[1], *h2 = gb[2];
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int e = 0; e < 2; e++) {
          const int si = i - a * ip, sj = j - b * jp, sk = k - e * kp;
          if (si < -2 || si > ni + 3 - ip || sj < -2 || sj > nj + 3 - jp || sk < -2 ||
              sk > nk + 3 - kp)
            continue;
          const int s = d.ix(si, sj, sk);
#pragma unroll
          for (int q = 0; q < 3; q++) acc[q] = acc[q] + R[q * n3 + s] * 0.125;
        }
    h0[c] = h0[c] + acc[0];
    h1[c] = h1[c] + acc[1];
    h2[c] = h2[c] + acc[2];
    rnb[c] = 0.0;
    rnb[n3 + c] = 0.0;
    rnb[2 * n3 + c] = 0.0;
  });
}

// srcs/calrkap.f:1-102 reversed: every normal component gathers the
// curvature adjoints of the (at most two) box cells that difference it;
// then the curvature adjoints are cleared whole.
static void calrkap_adj(const Dims d, double dxinv, double dyinv, double dzinv, double* pp,
                        double* r12, double* r13, double* r23, double* kb) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const size_t n3 = (size_t)d.n3;
  const IBox B1{-2, ni + 2, -1, nj + 3, -1, nk + 3}, B2{-1, ni + 3, -2, nj + 2, -1, nk + 3},
      B3{-1, ni + 3, -1, nj + 3, -2, nk + 2};
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    // curvature adjoint of component t at offset (di,dj,dk), 0 off its box
    auto t1 = [&](int di, int dj, int dk, int o) { return B1.has(i + di, j + dj, k + dk) ? kb[c + o] : 0.0; };
    auto t2 = [&](int di, int dj, int dk, int o) { return B2.has(i + di, j + dj, k + dk) ? kb[n3 + c + o] : 0.0; };
    auto t3 = [&](int di, int dj, int dk, int o) { return B3.has(i + di, j + dj, k + dk) ? kb[2 * n3 + c + o] : 0.0; };
    pp[c] = pp[c] + (t1(-1, 0, 0, -X) - t1(0, 0, 0, 0)) * dxinv;
    pp[n3 + c] = pp[n3 + c] + (t2(0, -1, 0, -Y) - t2(0, 0, 0, 0)) * dyinv;
    pp[2 * n3 + c] = pp[2 * n3 + c] + (t3(0, 0, -1, -Z) - t3(0, 0, 0, 0)) * dzinv;
    r12[c] = r12[c] + (t2(0, 0, 0, 0) - t2(1, 0, 0, X)) * dxinv;
    r12[n3 + c] = r12[n3 + c] + (t1(0, 0, 0, 0) - t1(0, 1, 0, Y)) * dyinv;
    r13[c] = r13[c] + (t3(0, 0, 0, 0) - t3(1, 0, 0, X)) * dxinv;
    r13[2 * n3 + c] = r13[2 * n3 + c] + (t1(0, 0, 0, 0) - t1(0, 0, 1, Z)) * dzinv;
    r23[n3 + c] = r23[n3 + c] + (t3(0, 0, 0, 0) - t3(0, 1, 0, Y)) * dyinv;
    r23[2 * n3 + c] = r23[2 * n3 + c] + (t2(0, 0, 0, 0) - t2(0, 0, 1, Z)) * dzinv;
  });
  launch_flat(3 * d.n3, [=] __device__(int m) { kb[m] = 0.0; });
}

// srcs/cal_grad_p2uvw.f:1-47 reversed (forward differences on [-2,n+2]^3):
// q_b gathers the face adjoints at c (minus) and at c-X, c-Y, c-Z (plus);
// then the face adjoints are cleared on the box.
static void grad_p2uvw_adj(const Dims d, double dxinv, const double* dyinv_row, double dzinv,
                           double* qb, double* gx, double* gy, double* gz) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const IBox B{-2, ni + 2, -2, nj + 2, -2, nk + 2};
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    double acc = 0.0;
    if (B.has(i, j, k)) acc = acc - gx[c] * dxinv - gy[c] * dyinv_row[j + 2] - gz[c] * dzinv;
    if (B.has(i - 1, j, k)) acc = acc + gx[c - X] * dxinv;
    if (B.has(i, j - 1, k)) acc = acc + gy[c - Y] * dyinv_row[j + 1];
    if (B.has(i, j, k - 1)) acc = acc + gz[c - Z] * dzinv;
    qb[c] = qb[c] + acc;
  });
  launch_box(Box{-2, ni + 2, -2, nj + 2, -2, nk + 2}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    gx[c] = 0.0;
    gy[c] = 0.0;
    gz[c] = 0.0;
  });
}

// The reverse sweeps of solu_sor4_bp_ (V false: coefficient arrays and
// their adjoints) and solu_sor4_bpv_ (V true: coefficients from the
// viscosity, face-viscosity adjoints), a, sp and bp set up by the entry.
template <bool V>
static void sor_bp_run(const Dims d, const int* nID, SorAdj& a, SorParam& sp, BnduParam& bp,
                       const double* const* x0, const double* dl1, const double* dl2,
                       const double* dphi, double uwall, double dy, int lp, double* uwallb) {
  a.G = static_cast<double*>(scratch(24, sizeof(double) * (size_t)d.n3));
  const bool lo = nID[kYMinus] < 0, hi = nID[kYPlus] < 0;
  a.jen = hi ? d.nj - 1 : d.nj;
  const int nw = d.sx * d.sz;
  double* wacc = static_cast<double*>(scratch(25, sizeof(double) * nw));
  bp.pa = static_cast<double*>(scratch(43, sizeof(double) * 2 * nw));
  launch_flat(nw, [=] __device__(int m) { wacc[m] = 0.0; });
  const int nsw = lp < 21 ? lp + 1 : 21;
  const size_t n3 = (size_t)d.n3;
  int ns = 0;
  const int* shell = shell_cells(d, &ns);
  // undo-log slot of one pair: its cells' old values, then the shell's
  const size_t np = 2 * (size_t)((d.ni + 1) / 2) * (size_t)((d.nj + 1) / 2) * (size_t)((d.nk + 1) / 2);
  const size_t slot = np + 3 * (size_t)ns;
  double* work = static_cast<double*>(scratch(42, sizeof(double) * (3 * n3 + slot * 12 * (size_t)nsw)));
  double* X[3] = {work, work + n3, work + 2 * n3};
  double* undo = work + 3 * n3;
  double* rows = static_cast<double*>(scratch(45, sizeof(double) * 4 * nw));
  bp.rows = rows;
  const double uw = uwall, dyv = dy;
  const int jen = a.jen;
  for (int r = 0; r < 3; r++) copy_dev(X[r], x0[r], d.n3);
  for (int s = 0; s < nsw; s++)
    for (int q = 0; q < 3; q++)
      for (int p = 0; p < 4; p++) {
        double* u = undo + slot * (size_t)((s * 3 + q) * 4 + p);
        if (q == 0) { pair_save<0>(d, jen, p, X[0], u); pair_fwd<0, V>(d, a, sp, X, p); }
        if (q == 1) { pair_save<1>(d, jen, p, X[1], u); pair_fwd<1, V>(d, a, sp, X, p); }
        if (q == 2) { pair_save<2>(d, jen, p, X[2], u); pair_fwd<2, V>(d, a, sp, X, p); }
        shell_save(shell, ns, X, u + np);
        bc_fused<double, double>(d, X[0], X[1], X[2], uw, dyv, dl1, dl2, dphi, lo, hi, q, 7);
      }
  for (int s = nsw - 1; s >= 0; s--)
    for (int q = 2; q >= 0; q--)
      for (int p = 3; p >= 0; p--) {
        double* u = undo + slot * (size_t)((s * 3 + q) * 4 + p);
        // X: the state before the refresh, then before the pair
        shell_restore(shell, ns, X, u + np);
        wall_rows(d, X[0], X[2], rows);
        periodic_adj(d, a.lb[q]);
        bndu_adj(d, a.lb[0], a.lb[1], a.lb[2], wacc, dyv, dl1, dl2, dphi, lo, hi, bp);
        const int is = kIsor[2 * p], js = kJsor[2 * p], ks = kKsor[2 * p];
        if (q == 0) {
          pair_restore<0>(d, jen, p, X[0], u);
          pair_param<0, V>(d, a, sp, X, p);
          colour_gather<0, V>(d, a, is, js, ks);
        } else if (q == 1) {
          pair_restore<1>(d, jen, p, X[1], u);
          pair_param<1, V>(d, a, sp, X, p);
          colour_gather<1, V>(d, a, is, js, ks);
        } else {
          pair_restore<2>(d, jen, p, X[2], u);
          pair_param<2, V>(d, a, sp, X, p);
          colour_gather<2, V>(d, a, is, js, ks);
        }
      }
  adjoint_add_sum(wacc, nw, uwallb);
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


// ---------------------------------------------------------------------------
// Parameter (all-argument) adjoints and the adjoints of the material
// properties and the surface-tension chain (entry points).
// ---------------------------------------------------------------------------

// All-argument adjoint of bndu_: bndu_b_ plus the adjoints of l1, l2 and phi
// (incremented).  uk, wk are the call's input velocities (rows 1 and nj are
// read), each followed by its adjoint; uwall is followed by *uwallb.
void bndu_bp_(const int* nID, const int* ni, const int* nj, const int* nk, const double* uk,
              double* ukb, double* vkb, const double* wk, double* wkb, const double* uwall,
              double* uwallb, const double* dy, const double* l1, double* l1b, const double* l2,
              double* l2b, const double* phi, double* phib) {
  adjoint_fp64_only("bndu_bp_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const int nw = d.sx * d.sz;
  double* wacc = static_cast<double*>(scratch(25, sizeof(double) * nw));
  double* pa = static_cast<double*>(scratch(43, sizeof(double) * 2 * nw));
  launch_flat(nw, [=] __device__(int m) { wacc[m] = 0.0; });
  Call c("bndu_bp_");
  BnduParam bp;
  double* rows = static_cast<double*>(scratch(45, sizeof(double) * 4 * nw));
  wall_rows(d, c.in(uk, d.n3, "uk"), c.in(wk, d.n3, "wk"), rows);
  bp.rows = rows;
  bp.uwall = *uwall;
  double* ub = c.io(ukb, d.n3, "ukb");
  double* vb = c.io(vkb, d.n3, "vkb");
  double* wb = c.io(wkb, d.n3, "wkb");
  const double* dl1 = c.in(l1, d.n3, "l1");
  const double* dl2 = c.in(l2, d.n3, "l2");
  const double* dphi = c.in(phi, d.n3, "phi");
  bp.l1b = c.io(l1b, d.n3, "l1b");
  bp.l2b = c.io(l2b, d.n3, "l2b");
  bp.phib = c.io(phib, d.n3, "phib");
  bp.pa = pa;
  bndu_adj(d, ub, vb, wb, wacc, *dy, dl1, dl2, dphi, nID[kYMinus] < 0, nID[kYPlus] < 0, bp);
  adjoint_add_sum(wacc, nw, uwallb);
  c.finish();
}

// All-argument adjoint of solu_sor4_ with the forward call's sweep count
// lp frozen: solu_sor4_b_ plus the adjoints of rho, rhon, the 45
// coefficients, l1, l2 and phi (incremented).  Each of rho, rhon, the
// coefficients (all 45, then their 45 adjoints), src_*, un, vn, wn (the
// call's inputs), uwall, l1, l2, phi is followed by its adjoint.  The
// sweeps are recomputed from the inputs, one sweep at a time before it is
// reversed from an undo log of the forward pass (the values each pair and
// each refresh overwrite), restored in reverse order: each pair's parameter
// adjoints are formed at the state before it (pair_param), and each
// refresh's bndu is reversed with its parameter adjoints at the state
// before the refresh.  Device work memory, nsw = min(lp+1, 21) sweeps:
// 3 padded arrays plus 12 nsw undo slots of a quarter array and three
// padded arrays' ghost shells each.
void solu_sor4_bp_(const int* ipara, const int* ID, const int* nID, const int* ndiv, const int* ni,
                   const int* nj, const int* nk, const int* key, double* sendjb, double* recvjb,
                   const int* nstep, const int* imon_t, const double* dt, const double* rho,
                   double* rhob, const double* rhon, double* rhonb, double* c0, double* c1, double* c2, double* c3, double* c4, double* c5, double* c6, double* c7, double* c8, double* c9, double* c10, double* c11, double* c12, double* c13, double* c14, double* c15, double* c16, double* c17, double* c18, double* c19, double* c20, double* c21, double* c22, double* c23, double* c24, double* c25, double* c26, double* c27, double* c28, double* c29, double* c30, double* c31, double* c32, double* c33, double* c34, double* c35, double* c36, double* c37, double* c38, double* c39, double* c40, double* c41, double* c42, double* c43, double* c44, double* c0b, double* c1b, double* c2b, double* c3b, double* c4b, double* c5b, double* c6b, double* c7b, double* c8b, double* c9b, double* c10b, double* c11b, double* c12b, double* c13b, double* c14b, double* c15b, double* c16b, double* c17b, double* c18b, double* c19b, double* c20b, double* c21b, double* c22b, double* c23b, double* c24b, double* c25b, double* c26b, double* c27b, double* c28b, double* c29b, double* c30b, double* c31b, double* c32b, double* c33b, double* c34b, double* c35b, double* c36b, double* c37b, double* c38b, double* c39b, double* c40b, double* c41b, double* c42b, double* c43b, double* c44b,
                   const double* src_u, double* src_ub, const double* src_v, double* src_vb,
                   const double* src_w, double* src_wb, const double* un, double* unb,
                   const double* vn, double* vnb, const double* wn, double* wnb,
                   const double* uwall, double* uwallb, const double* dy, const double* l1,
                   double* l1b, const double* l2, double* l2b, const double* phi, double* phib,
                   const int* lp) {
  (void)ID; (void)ndiv; (void)key; (void)sendjb; (void)recvjb; (void)nstep; (void)imon_t;
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: solu_sor4_bp_: only the single-domain case is supported\n");
    std::abort();
  }
  if (*lp < 0 || *lp > 21) {
    std::fprintf(stderr, "felgpu: solu_sor4_bp_: lp = %d is not a sweep index of solu_sor4_ (0..21)\n", *lp);
    std::abort();
  }
  adjoint_fp64_only("solu_sor4_bp_");
  const Dims d = make_dims(*ni, *nj, *nk);
  double* h[45] = {c0, c1, c2, c3, c4, c5, c6, c7, c8, c9, c10, c11, c12, c13, c14, c15, c16, c17, c18, c19, c20, c21, c22, c23, c24, c25, c26, c27, c28, c29, c30, c31, c32, c33, c34, c35, c36, c37, c38, c39, c40, c41, c42, c43, c44};
  double* hb[45] = {c0b, c1b, c2b, c3b, c4b, c5b, c6b, c7b, c8b, c9b, c10b, c11b, c12b, c13b, c14b, c15b, c16b, c17b, c18b, c19b, c20b, c21b, c22b, c23b, c24b, c25b, c26b, c27b, c28b, c29b, c30b, c31b, c32b, c33b, c34b, c35b, c36b, c37b, c38b, c39b, c40b, c41b, c42b, c43b, c44b};
  Call c("solu_sor4_bp_");
  SorAdj a;
  SorParam sp;
  a.rho = c.in(rho, d.n3, "rho");
  a.rhon = c.in(rhon, d.n3, "rhon");
  sp.rhob = c.io(rhob, d.n3, "rhob");
  sp.rhonb = c.io(rhonb, d.n3, "rhonb");
  a.rmu = nullptr;
  a.sc = vis_scales<double>(0.0, 0.0, 0.0);
  for (int m = 0; m < 45; m++) a.cf[m] = c.in(h[m], d.n3, "coefficient");
  for (int m = 0; m < 45; m++) sp.cfb[m] = c.io(hb[m], d.n3, "coefficient adjoint");
  sp.src[0] = c.in(src_u, d.n3, "src_u");
  sp.src[1] = c.in(src_v, d.n3, "src_v");
  sp.src[2] = c.in(src_w, d.n3, "src_w");
  a.srcb[0] = c.io(src_ub, d.n3, "src_ub");
  a.srcb[1] = c.io(src_vb, d.n3, "src_vb");
  a.srcb[2] = c.io(src_wb, d.n3, "src_wb");
  const double* x0[3] = {c.in(un, d.n3, "un"), c.in(vn, d.n3, "vn"), c.in(wn, d.n3, "wn")};
  a.lb[0] = c.io(unb, d.n3, "unb");
  a.lb[1] = c.io(vnb, d.n3, "vnb");
  a.lb[2] = c.io(wnb, d.n3, "wnb");
  const double* dl1 = c.in(l1, d.n3, "l1");
  const double* dl2 = c.in(l2, d.n3, "l2");
  const double* dphi = c.in(phi, d.n3, "phi");
  BnduParam bp;
  bp.uwall = *uwall;
  bp.l1b = c.io(l1b, d.n3, "l1b");
  bp.l2b = c.io(l2b, d.n3, "l2b");
  bp.phib = c.io(phib, d.n3, "phib");
  a.dtinv = 1.0 / *dt;
  g_last_adj_onfly = 0;
  sor_bp_run<false>(d, nID, a, sp, bp, x0, dl1, dl2, dphi, *uwall, *dy, *lp, uwallb);
  c.finish();
}

// All-argument adjoint of cal_arith_coef_vis_ followed by solu_sor4_ (the
// 45 coefficients formed from the viscosity rmu with dxinv, dyinv, dzinv),
// sweep count frozen at *lp: what solu_sor4_bp_ and then
// cal_arith_coef_vis_b_ give, without the coefficient arrays -- the
// coefficients are formed on the fly (the same bits) and each pair cell's
// coefficient adjoints go straight into its six face-viscosity adjoints,
// gathered into rmub at the end.  The forward's argument list with the 45
// coefficients replaced by dxinv, dyinv, dzinv, rmu, rmub; each of rho,
// rhon, rmu, src_*, un, vn, wn, uwall, l1, l2, phi followed by its adjoint.
// Device work memory as solu_sor4_bp_'s plus 18 arrays of the padded grid.
void solu_sor4_bpv_(const int* ipara, const int* ID, const int* nID, const int* ndiv, const int* ni,
                    const int* nj, const int* nk, const int* key, double* sendjb, double* recvjb,
                    const int* nstep, const int* imon_t, const double* dt, const double* rho,
                    double* rhob, const double* rhon, double* rhonb, const double* dxinv,
                    const double* dyinv, const double* dzinv, const double* rmu, double* rmub,
                    const double* src_u, double* src_ub, const double* src_v, double* src_vb,
                    const double* src_w, double* src_wb, const double* un, double* unb,
                    const double* vn, double* vnb, const double* wn, double* wnb,
                    const double* uwall, double* uwallb, const double* dy, const double* l1,
                    double* l1b, const double* l2, double* l2b, const double* phi, double* phib,
                    const int* lp) {
  (void)ID; (void)ndiv; (void)key; (void)sendjb; (void)recvjb; (void)nstep; (void)imon_t;
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: solu_sor4_bpv_: only the single-domain case is supported\n");
    std::abort();
  }
  if (*lp < 0 || *lp > 21) {
    std::fprintf(stderr, "felgpu: solu_sor4_bpv_: lp = %d is not a sweep index of solu_sor4_ (0..21)\n", *lp);
    std::abort();
  }
  adjoint_fp64_only("solu_sor4_bpv_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("solu_sor4_bpv_");
  SorAdj a;
  SorParam sp;
  a.rho = c.in(rho, d.n3, "rho");
  a.rhon = c.in(rhon, d.n3, "rhon");
  sp.rhob = c.io(rhob, d.n3, "rhob");
  sp.rhonb = c.io(rhonb, d.n3, "rhonb");
  a.rmu = c.in(rmu, d.n3, "rmu");
  double* mb = c.io(rmub, d.n3, "rmub");
  a.sc = vis_scales<double>(*dxinv, *dyinv, *dzinv);
  for (int m = 0; m < 45; m++) a.cf[m] = nullptr;
  for (int m = 0; m < 45; m++) sp.cfb[m] = nullptr;
  sp.dinv[0] = *dxinv;
  sp.dinv[1] = *dyinv;
  sp.dinv[2] = *dzinv;
  double* faces = static_cast<double*>(scratch(46, sizeof(double) * 18 * (size_t)d.n3));
  for (int m = 0; m < 18; m++) {
    double* f = faces + (size_t)m * d.n3;
    sp.fb[m] = f;
    launch_flat(d.n3, [=] __device__(int e) { f[e] = 0.0; });
  }
  sp.src[0] = c.in(src_u, d.n3, "src_u");
  sp.src[1] = c.in(src_v, d.n3, "src_v");
  sp.src[2] = c.in(src_w, d.n3, "src_w");
  a.srcb[0] = c.io(src_ub, d.n3, "src_ub");
  a.srcb[1] = c.io(src_vb, d.n3, "src_vb");
  a.srcb[2] = c.io(src_wb, d.n3, "src_wb");
  const double* x0[3] = {c.in(un, d.n3, "un"), c.in(vn, d.n3, "vn"), c.in(wn, d.n3, "wn")};
  a.lb[0] = c.io(unb, d.n3, "unb");
  a.lb[1] = c.io(vnb, d.n3, "vnb");
  a.lb[2] = c.io(wnb, d.n3, "wnb");
  const double* dl1 = c.in(l1, d.n3, "l1");
  const double* dl2 = c.in(l2, d.n3, "l2");
  const double* dphi = c.in(phi, d.n3, "phi");
  BnduParam bp;
  bp.uwall = *uwall;
  bp.l1b = c.io(l1b, d.n3, "l1b");
  bp.l2b = c.io(l2b, d.n3, "l2b");
  bp.phib = c.io(phib, d.n3, "phib");
  a.dtinv = 1.0 / *dt;
  g_last_adj_onfly = 1;
  sor_bp_run<true>(d, nID, a, sp, bp, x0, dl1, dl2, dphi, *uwall, *dy, *lp, uwallb);
  const double* fb[18];
  for (int m = 0; m < 18; m++) fb[m] = sp.fb[m];
  face_gather(d, fb, mb);
  c.finish();
}

// All-argument adjoint of cal_arith_tau_: cal_arith_tau_b_ plus the
// viscosity's adjoint (incremented); s and rmu each followed by its adjoint.
void cal_arith_tau_bp_(const int* ni, const int* nj, const int* nk, const double* s, double* sb,
                       const double* rmu, double* rmub, double* taub) {
  adjoint_fp64_only("cal_arith_tau_bp_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_arith_tau_bp_");
  const double* sv = c.in(s, 6 * (size_t)d.n3, "s");
  double* sa = c.io(sb, 6 * (size_t)d.n3, "sb");
  const double* m = c.in(rmu, d.n3, "rmu");
  double* mb = c.io(rmub, d.n3, "rmub");
  double* t = c.io(taub, 6 * (size_t)d.n3, "taub");
  arith_tau_param(d, sv, mb, t);
  arith_tau_adj(d, sa, m, t);
  c.finish();
}

// All-argument adjoint of cal_srcu_: cal_srcu_b_ plus the densities'
// adjoints (incremented).  rho, rhon, prs_*, vis_* are each followed by
// their adjoints (the primal pressure gradients and viscous terms are what
// the density derivative needs).
void cal_srcu_bp_(const int* nID, const int* ni, const int* nj, const int* nk, const double* rho_av,
                  const double* dt, const double* grvb, const double* grvp, double* ub, double* vb,
                  double* wb, const double* rho, double* rhob, const double* rhon, double* rhonb,
                  double* adv_ub, double* adv_vb, double* adv_wb, double* adv_uob, double* adv_vob,
                  double* adv_wob, const double* prs_u, double* prs_ub, const double* prs_v,
                  double* prs_vb, const double* prs_w, double* prs_wb, const double* vis_u,
                  double* vis_ub, const double* vis_v, double* vis_vb, const double* vis_w,
                  double* vis_wb, double* fst_ub, double* fst_vb, double* fst_wb, double* fst_unb,
                  double* fst_vnb, double* fst_wnb, double* src_ub, double* src_vb,
                  double* src_wb) {
  (void)rho_av; (void)grvb; (void)grvp;
  adjoint_fp64_only("cal_srcu_bp_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const int jen = nID[kYPlus] < 0 ? d.nj - 1 : d.nj;
  Call c("cal_srcu_bp_");
  SrcuAdj a;
  double* hv[3] = {ub, vb, wb};
  double* ha[3] = {adv_ub, adv_vb, adv_wb};
  double* hao[3] = {adv_uob, adv_vob, adv_wob};
  double* hp[3] = {prs_ub, prs_vb, prs_wb};
  double* hvis[3] = {vis_ub, vis_vb, vis_wb};
  double* hf[3] = {fst_ub, fst_vb, fst_wb};
  double* hfn[3] = {fst_unb, fst_vnb, fst_wnb};
  double* hs[3] = {src_ub, src_vb, src_wb};
  const double* pp[3] = {prs_u, prs_v, prs_w};
  const double* pv[3] = {vis_u, vis_v, vis_w};
  const double *P[3], *V[3];
  for (int q = 0; q < 3; q++) {
    a.vel[q] = c.io(hv[q], d.n3, "velocity adjoint");
    a.adv[q] = c.io(ha[q], d.n3, "adv adjoint");
    a.advo[q] = c.io(hao[q], d.n3, "adv_o adjoint");
    a.prs[q] = c.io(hp[q], d.n3, "prs adjoint");
    a.vis[q] = c.io(hvis[q], d.n3, "vis adjoint");
    a.fst[q] = c.io(hf[q], d.n3, "fst adjoint");
    a.fstn[q] = c.io(hfn[q], d.n3, "fst_n adjoint");
    a.src[q] = c.io(hs[q], d.n3, "src adjoint");
    P[q] = c.in(pp[q], d.n3, "prs");
    V[q] = c.in(pv[q], d.n3, "vis");
  }
  a.rho = c.in(rho, d.n3, "rho");
  a.rhon = c.in(rhon, d.n3, "rhon");
  double* rb = c.io(rhob, d.n3, "rhob");
  double* rnb = c.io(rhonb, d.n3, "rhonb");
  const double* sb[3] = {a.src[0], a.src[1], a.src[2]};
  srcu_param(d, jen, P, V, sb, a.rho, a.rhon, rb, rnb);
  srcu_adj(d, jen, 1.0 / *dt, a);
  c.finish();
}

// Adjoint of cal_arith_coef_vis_ (srcs/cal_arith_coef_vis.f:1-411), linear
// in the viscosity: rmukb incremented, the 45 coefficient adjoints consumed.
void cal_arith_coef_vis_b_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                           const double* dyinv, const double* dzinv, double* rmukb, double* c0b, double* c1b, double* c2b, double* c3b, double* c4b, double* c5b, double* c6b, double* c7b, double* c8b, double* c9b, double* c10b, double* c11b, double* c12b, double* c13b, double* c14b, double* c15b, double* c16b, double* c17b, double* c18b, double* c19b, double* c20b, double* c21b, double* c22b, double* c23b, double* c24b, double* c25b, double* c26b, double* c27b, double* c28b, double* c29b, double* c30b, double* c31b, double* c32b, double* c33b, double* c34b, double* c35b, double* c36b, double* c37b, double* c38b, double* c39b, double* c40b, double* c41b, double* c42b, double* c43b, double* c44b) {
  adjoint_fp64_only("cal_arith_coef_vis_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  double* hb[45] = {c0b, c1b, c2b, c3b, c4b, c5b, c6b, c7b, c8b, c9b, c10b, c11b, c12b, c13b, c14b, c15b, c16b, c17b, c18b, c19b, c20b, c21b, c22b, c23b, c24b, c25b, c26b, c27b, c28b, c29b, c30b, c31b, c32b, c33b, c34b, c35b, c36b, c37b, c38b, c39b, c40b, c41b, c42b, c43b, c44b};
  Call c("cal_arith_coef_vis_b_");
  double* mb = c.io(rmukb, d.n3, "rmukb");
  double* cb[45];
  for (int m = 0; m < 45; m++) cb[m] = c.io(hb[m], d.n3, "coefficient adjoint");
  coef_vis_adj(d, *dxinv, *dyinv, *dzinv, mb, cb);
  c.finish();
}

// Adjoint of cal_arith_mean_ (srcs/cal_arith_mean.f:18-26): phikb += (qg -
// ql) qkb, qkb consumed.
void cal_arith_mean_b_(const int* ni, const int* nj, const int* nk, double* phikb, const double* ql,
                       const double* qg, double* qkb) {
  adjoint_fp64_only("cal_arith_mean_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_arith_mean_b_");
  double* pb = c.io(phikb, d.n3, "phikb");
  double* qb = c.io(qkb, d.n3, "qkb");
  const double l = *ql, g = *qg;
  launch_flat(d.n3, [=] __device__(int m) {
    const double t = qb[m];
    pb[m] = pb[m] + (-(l * t) + g * t);
    qb[m] = 0.0;
  });
  c.finish();
}

// Adjoint of calrn_ (srcs/calrn.f:1-95), nonlinear: phix, phiy, phiz each
// followed by its adjoint (incremented); rnb (three components) consumed.
void calrn_b_(const int* ni, const int* nj, const int* nk, const double* dxinv, const double* dyinv,
              const double* dzinv, const int* ip, const int* jp, const int* kp, const double* phix,
              double* phixb, const double* phiy, double* phiyb, const double* phiz, double* phizb,
              double* rnb) {
  adjoint_fp64_only("calrn_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("calrn_b_");
  const double* g[3] = {c.in(phix, d.n3, "phix"), c.in(phiy, d.n3, "phiy"), c.in(phiz, d.n3, "phiz")};
  double* gb[3] = {c.io(phixb, d.n3, "phixb"), c.io(phiyb, d.n3, "phiyb"), c.io(phizb, d.n3, "phizb")};
  double* r = c.io(rnb, 3 * (size_t)d.n3, "rnb");
  double* R = static_cast<double*>(scratch(44, sizeof(double) * 3 * (size_t)d.n3));
  calrn_adj(d, *dxinv, *dyinv, *dzinv, *ip, *jp, *kp, g, gb, r, R);
  c.finish();
}

// Adjoint of calrkap_ (srcs/calrkap.f:1-102), linear: the four normal
// adjoints (three components each) incremented, rkapb consumed.
void calrkap_b_(const int* ni, const int* nj, const int* nk, const double* dxinv, const double* dyinv,
                const double* dzinv, double* rn_ppb, double* rn_12b, double* rn_13b, double* rn_23b,
                double* rkapb) {
  adjoint_fp64_only("calrkap_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("calrkap_b_");
  const size_t n = 3 * (size_t)d.n3;
  double* pp = c.io(rn_ppb, n, "rn_ppb");
  double* a12 = c.io(rn_12b, n, "rn_12b");
  double* a13 = c.io(rn_13b, n, "rn_13b");
  double* a23 = c.io(rn_23b, n, "rn_23b");
  double* kb = c.io(rkapb, n, "rkapb");
  calrkap_adj(d, *dxinv, *dyinv, *dzinv, pp, a12, a13, a23, kb);
  c.finish();
}

// Adjoint of cal_grad_p2uvw_ (srcs/cal_grad_p2uvw.f:1-47), linear: qb
// incremented, the face-gradient adjoints consumed on [-2,n+2]^3.
void cal_grad_p2uvw_b_(const int* ID, const int* svall, const int* ni, const int* nj, const int* nk,
                       const double* dxinv, const double* dyinv_array, const double* dzinv,
                       double* qb, double* qx_ub, double* qy_vb, double* qz_wb) {
  adjoint_fp64_only("cal_grad_p2uvw_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const int count = *svall + 6;  // dyinv_array(-2:svall+3)
  double* dyt = static_cast<double*>(scratch(35, sizeof(double) * (size_t)count));
  FEL_CHECK(hipMemcpyAsync(dyt, dyinv_array, sizeof(double) * count, hipMemcpyHostToDevice, stream()));
  Call c("cal_grad_p2uvw_b_");
  double* q = c.io(qb, d.n3, "qb");
  double* gx = c.io(qx_ub, d.n3, "qx_ub");
  double* gy = c.io(qy_vb, d.n3, "qy_vb");
  double* gz = c.io(qz_wb, d.n3, "qz_wb");
  grad_p2uvw_adj(d, *dxinv, dyt + (size_t)(*nj) * (size_t)(*ID), *dzinv, q, gx, gy, gz);
  c.finish();
  FEL_CHECK(hipStreamSynchronize(stream()));  // dyinv_array is a host array read asynchronously
}

// Adjoint of calfst_ (srcs/calfst.f:1-50), bilinear: the face gradients and
// the curvature (three components) each followed by its adjoint
// (incremented); the force adjoints consumed.
void calfst_b_(const int* ni, const int* nj, const int* nk, const double* rhol, const double* rhog,
               const double* surface_tension, const double* phix_u, double* phix_ub,
               const double* phiy_v, double* phiy_vb, const double* phiz_w, double* phiz_wb,
               const double* rkap, double* rkapb, double* fst_ub, double* fst_vb, double* fst_wb) {
  adjoint_fp64_only("calfst_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("calfst_b_");
  const double *gx = c.in(phix_u, d.n3, "phix_u"), *gy = c.in(phiy_v, d.n3, "phiy_v"),
               *gz = c.in(phiz_w, d.n3, "phiz_w");
  double *gxb = c.io(phix_ub, d.n3, "phix_ub"), *gyb = c.io(phiy_vb, d.n3, "phiy_vb"),
         *gzb = c.io(phiz_wb, d.n3, "phiz_wb");
  const double* k = c.in(rkap, 3 * (size_t)d.n3, "rkap");
  double* kb = c.io(rkapb, 3 * (size_t)d.n3, "rkapb");
  double *fu = c.io(fst_ub, d.n3, "fst_ub"), *fv = c.io(fst_vb, d.n3, "fst_vb"),
         *fw = c.io(fst_wb, d.n3, "fst_wb");
  const double st = 2.0 * (*surface_tension) / (*rhol + *rhog);
  const size_t n3 = (size_t)d.n3;
  launch_flat(d.n3, [=] __device__(int m) {
    const double* g[3] = {gx, gy, gz};
    double* gb[3] = {gxb, gyb, gzb};
    double* fb[3] = {fu, fv, fw};
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const double t = fb[q][m];
      gb[q][m] = gb[q][m] + -(st * k[q * n3 + m] * t);
      kb[q * n3 + m] = kb[q * n3 + m] + -(st * g[q][m] * t);
      fb[q][m] = 0.0;
    }
  });
  c.finish();
}

// Adjoint of sum_fst_ (srcs/sum_fst.f:1-37): f_b += s_b (s_b unchanged).
void sum_fst_b_(const int* ni, const int* nj, const int* nk, double* fub, double* fvb, double* fwb,
                const double* sub, const double* svb, const double* swb) {
  adjoint_fp64_only("sum_fst_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("sum_fst_b_");
  double *fu = c.io(fub, d.n3, "fub"), *fv = c.io(fvb, d.n3, "fvb"), *fw = c.io(fwb, d.n3, "fwb");
  const double *su = c.in(sub, d.n3, "sub"), *sv = c.in(svb, d.n3, "svb"),
               *sw = c.in(swb, d.n3, "swb");
  launch_flat(d.n3, [=] __device__(int m) {
    fu[m] = fu[m] + su[m];
    fv[m] = fv[m] + sv[m];
    fw[m] = fw[m] + sw[m];
  });
  c.finish();
}

}  // extern "C"
