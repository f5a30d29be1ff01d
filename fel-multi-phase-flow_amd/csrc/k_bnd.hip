// k_bnd.hip -- ghost fills: Navier slip walls, periodic x/z, Neumann walls,
// dynamic contact angle (GNBC) and the contact-angle ghosts of phi.
//
// Reference: srcs/bndu.f, srcs/bnd_periodic.f (bnd_periodic, bnd_neumann),
// srcs/bnd_comm.f, srcs/gnbc.f90, srcs/bnd_contact_angle.f90.
#include <cmath>

#include "felgpu.hpp"
#include "felgpu_bnd.hpp"
#include "felgpu_visc.hpp"

namespace fel {

// srcs/bndu.f:29-96: wall rows, one thread per (i,k) of the padded plane.
static void bndu_walls(const Dims d, double* uk, double* vk, double* wk, double uwall, double dy,
                       const double* l1, const double* l2, const double* phi, bool lo, bool hi) {
  const int ni = d.ni, nj = d.nj;
  launch_box(Box{-2, d.ni + 3, 0, 0, -2, d.nk + 3}, [=] __device__(int i, int, int k) {
    const int ip = (i + 1) % (ni + 3);  // Fortran mod(i+1,ni+3), sign of the dividend
    if (lo) {
      const int c1 = d.ix(i, 1, k);
      const double phi_av = (phi[c1] + phi[d.ix(ip, 1, k)]) / 2.0;
      const double inv1 = phi_av / l1[c1];
      const double inv2 = (1.0 - phi_av) / l2[c1];
      const double ls = 1.0 / (inv1 + inv2);
      const double coef1 = (2.0 * dy) / (2.0 * ls + dy);
      const double coef2 = (2.0 * ls - dy) / (2.0 * ls + dy);
      const int c0 = d.ix(i, 0, k), cm = d.ix(i, -1, k), cmm = d.ix(i, -2, k), c2 = d.ix(i, 2, k);
      const double u1 = uk[c1];
      const double u0 = coef1 * (-uwall) + coef2 * u1;
      const double um = 2.0 * u0 - u1;
      uk[c0] = u0;
      uk[cm] = um;
      uk[cmm] = 2.0 * um - u0;
      vk[cmm] = -vk[c2];
      vk[cm] = -vk[c1];
      vk[c0] = 0.0;
      const double w1 = wk[c1];
      const double w0 = coef2 * w1;
      const double wm = 2.0 * w0 - w1;
      wk[c0] = w0;
      wk[cm] = wm;
      wk[cmm] = 2.0 * wm - w0;
    }
    if (hi) {
      const int cn = d.ix(i, nj, k);
      const double phi_av = (phi[cn] + phi[d.ix(ip, nj, k)]) / 2.0;
      const double inv1 = phi_av / l1[cn];
      const double inv2 = (1.0 - phi_av) / l2[cn];
      const double ls = 1.0 / (inv1 + inv2);
      const double coef1 = (2.0 * dy) / (2.0 * ls + dy);
      const double coef2 = (2.0 * ls - dy) / (2.0 * ls + dy);
      const int Y = d.sx;
      const double un = uk[cn];
      const double u1 = coef1 * uwall + coef2 * un;
      const double u2 = 2.0 * u1 - un;
      uk[cn + Y] = u1;
      uk[cn + 2 * Y] = u2;
      uk[cn + 3 * Y] = 2.0 * u2 - u1;
      vk[cn] = 0.0;
      vk[cn + Y] = -vk[cn - Y];
      vk[cn + 2 * Y] = -vk[cn - 2 * Y];
      vk[cn + 3 * Y] = -vk[cn - 3 * Y];
      const double wn = wk[cn];
      const double w1 = coef2 * wn;
      const double w2 = 2.0 * w1 - wn;
      wk[cn + Y] = w1;
      wk[cn + 2 * Y] = w2;
      wk[cn + 3 * Y] = 2.0 * w2 - w1;
    }
  });
}

// srcs/bndu.f:105-126: x ghosts, including the reference's shifted source
// for u (u(0)<-u(ni+1), u(-1)<-u(ni), u(-2)<-u(ni-1); u(ni+1) untouched).
static void bndu_x(const Dims d, double* uk, double* vk, double* wk) {
  const int ni = d.ni;
  launch_box(Box{0, 0, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int, int j, int k) {
    const int r = d.ix(0, j, k);  // offset of i=0 in this row
#define AT(q, i) q[r + (i)]
    AT(uk, 0) = AT(uk, ni + 1);
    AT(uk, -1) = AT(uk, ni);
    AT(uk, -2) = AT(uk, ni - 1);
    AT(vk, 0) = AT(vk, ni);
    AT(vk, -1) = AT(vk, ni - 1);
    AT(vk, -2) = AT(vk, ni - 2);
    AT(wk, 0) = AT(wk, ni);
    AT(wk, -1) = AT(wk, ni - 1);
    AT(wk, -2) = AT(wk, ni - 2);
    AT(uk, ni + 2) = AT(uk, 2);
    AT(uk, ni + 3) = AT(uk, 3);
    AT(vk, ni + 1) = AT(vk, 1);
    AT(vk, ni + 2) = AT(vk, 2);
    AT(vk, ni + 3) = AT(vk, 3);
    AT(wk, ni + 1) = AT(wk, 1);
    AT(wk, ni + 2) = AT(wk, 2);
    AT(wk, ni + 3) = AT(wk, 3);
#undef AT
  });
}

// srcs/bndu.f:136-157: z ghosts, with the shifted source for w
// (w(0)<-w(nk+1), w(-1)<-w(nk), w(-2)<-w(nk-1); w(nk+1) untouched).
static void bndu_z(const Dims d, double* uk, double* vk, double* wk) {
  const int nk = d.nk;
  launch_box(Box{-2, d.ni + 3, -2, d.nj + 3, 0, 0}, [=] __device__(int i, int j, int) {
    const int r = d.ix(i, j, 0), Z = d.sx * d.sy;
#define AT(q, k) q[r + (k) * Z]
    AT(uk, 0) = AT(uk, nk);
    AT(uk, -1) = AT(uk, nk - 1);
    AT(uk, -2) = AT(uk, nk - 2);
    AT(vk, 0) = AT(vk, nk);
    AT(vk, -1) = AT(vk, nk - 1);
    AT(vk, -2) = AT(vk, nk - 2);
    AT(wk, 0) = AT(wk, nk + 1);
    AT(wk, -1) = AT(wk, nk);
    AT(wk, -2) = AT(wk, nk - 1);
    AT(uk, nk + 1) = AT(uk, 1);
    AT(uk, nk + 2) = AT(uk, 2);
    AT(uk, nk + 3) = AT(uk, 3);
    AT(vk, nk + 1) = AT(vk, 1);
    AT(vk, nk + 2) = AT(vk, 2);
    AT(vk, nk + 3) = AT(vk, 3);
    AT(wk, nk + 2) = AT(wk, 2);
    AT(wk, nk + 3) = AT(wk, 3);
#undef AT
  });
}

void bndu(const Dims d, double* u, double* v, double* w, double uwall, double dy, const double* l1,
          const double* l2, const double* phi, bool lo, bool hi) {
  bndu_walls(d, u, v, w, uwall, dy, l1, l2, phi, lo, hi);
  bndu_x(d, u, v, w);
  bndu_z(d, u, v, w);
}

// srcs/bnd_periodic.f:16-43, x then z, in one pass: a ghost cell takes the
// value at (mx(i), j, mz(k)), where mx, mz map ghosts to their interior
// images and are the identity inside.  Sources are never ghosts in x or z,
// so the in-place pass has no read-after-write hazard and equals the
// reference's two passes, corners included.
__device__ __forceinline__ int pmap(int i, int n) {
  return i < 1 ? i + n : (i > n ? i - n : i);
}

void periodic(const Dims d, double* q) {
  // ghost cells only: the x-ghost columns of the interior planes, then the
  // six z-ghost planes (all i, all j); the two sets are disjoint
  const int ni = d.ni, nk = d.nk, sx = d.sx, sy = d.sy;
  const int na = 6 * sy * nk, nb = 6 * sx * sy;
  launch_flat(na + nb, [=] __device__(int m) {
    int i, j, k;
    if (m < na) {
      const int g = m % 6, r = m / 6;
      j = r % sy - 2;
      k = r / sy + 1;
      i = g < 3 ? g - 2 : ni + g - 2;  // -2,-1,0, ni+1,ni+2,ni+3
    } else {
      const int t = m - na;
      i = t % sx - 2;
      const int r = t / sx;
      j = r % sy - 2;
      const int g = r / sy;
      k = g < 3 ? g - 2 : nk + g - 2;
    }
    q[d.ix(i, j, k)] = q[d.ix(pmap(i, ni), j, pmap(k, nk))];
  });
}

// srcs/bnd_periodic.f:60-90 (bnd_neumann)
void neumann(const Dims d, double* q, bool lo, bool hi) {
  const int nj = d.nj;
  launch_box(Box{-2, d.ni + 3, 0, 0, -2, d.nk + 3}, [=] __device__(int i, int, int k) {
    const int Y = d.sx;
    if (lo) {
      const int c1 = d.ix(i, 1, k);
      q[c1 - 3 * Y] = q[c1 + 2 * Y];
      q[c1 - 2 * Y] = q[c1 + Y];
      q[c1 - Y] = q[c1];
    }
    if (hi) {
      const int cn = d.ix(i, nj, k);
      q[cn + Y] = q[cn];
      q[cn + 2 * Y] = q[cn - Y];
      q[cn + 3 * Y] = q[cn - 2 * Y];
    }
  });
}

// srcs/gnbc.f90:25-97
void gnbc(const Dims d, const double* uk, const double* wk, double uwall, const double* th0,
          double st, const double* zeta, double* theta, bool lo, bool hi, double pi) {
  const int ni = d.ni, nj = d.nj;
  launch_box(Box{-2, d.ni + 3, 0, 1, -2, d.nk + 3}, [=] __device__(int i, int side, int k) {
    if ((side == 0 && !lo) || (side == 1 && !hi)) return;
    const int j0 = side == 0 ? 0 : nj, j1 = j0 + 1, jw = side == 0 ? 1 : nj;
    const double uw = side == 0 ? uwall : -uwall;
    const int mi = (i + ni + 2) % (ni + 3);
    const double ux =
        (uk[d.ix(mi, j0, k)] + uk[d.ix(mi, j1, k)] + uk[d.ix(i, j0, k)] + uk[d.ix(i, j1, k)]) / 4.0 + uw;
    const double uz =
        (wk[d.ix(mi, j0, k)] + wk[d.ix(mi, j1, k)] + wk[d.ix(i, j0, k)] + wk[d.ix(i, j1, k)]) / 4.0;
    double ucl = copysign(sqrt(ux * ux + uz * uz + 1.0e-12), ux);
    if (i <= ni / 2) ucl = -ucl;
    const int cw = d.ix(i, jw, k);
    const double th0r = th0[cw] * (pi / 180.0);
    double ct = cos(th0r) - zeta[cw] * ucl / st;
    ct = fmin(1.0, fmax(-1.0, ct));
    theta[cw] = acos(ct) * (180.0 / pi);
  });
}

// srcs/bnd_contact_angle.f90:25-85.  Reads along the wall row that fall
// outside -2..n+3 in i or k (the outermost ghost column and plane, which the
// reference reads past its row or array) are taken as 0; those ghost values
// are overwritten by the following periodic fill or never reach the
// interior.
void contact_angle(const Dims d, double* q, const double* theta, double ax, double az, bool lo,
                   bool hi, double pi) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  launch_box(Box{-2, d.ni + 3, 0, 1, -2, d.nk + 3}, [=] __device__(int i, int side, int k) {
    if ((side == 0 && !lo) || (side == 1 && !hi)) return;
    const int jw = side == 0 ? 1 : nj, jg = side == 0 ? 0 : nj + 1, js = side == 0 ? -1 : 1;
    const int cw = d.ix(i, jw, k);
    const double th = theta[cw] * pi / 180.0;
    const double ct = cos(th), sn = sin(th);
    const double sg = copysign(1.0, (double)(i - ni / 2));
    double nx = sg * sn * ax, nz = sg * sn * az, ny = ct;
    const double nrm = sqrt(nx * nx + ny * ny + nz * nz + 1.0e-20);
    nx = nx / nrm;
    ny = ny / nrm;
    nz = nz / nrm;
    const double q1 = q[cw];
    const double qe = i + 1 <= ni + 3 ? q[cw + 1] : 0.0;
    const double qw = i - 1 >= -2 ? q[cw - 1] : 0.0;
    const double qt = k + 1 <= nk + 3 ? q[cw + d.sx * d.sy] : 0.0;
    const double qb = k - 1 >= -2 ? q[cw - d.sx * d.sy] : 0.0;
    const double g = q1 - (ny / (nx + 1.0e-20)) * (qe - qw) * 0.5 * ax -
                     (ny / (nz + 1.0e-20)) * (qt - qb) * 0.5 * az;
    const int Y = d.sx;
    const int c0 = d.ix(i, jg, k);
    const double g1 = 2.0 * g - q1;
    q[c0] = g;
    q[c0 + js * Y] = g1;
    q[c0 + 2 * js * Y] = 2.0 * g1 - g;
  });
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/bndu.f:6
void bndu_(const int* nID, const int* ni, const int* nj, const int* nk, double* uk, double* vk,
           double* wk, const double* uwall, const double* dy, const double* l1, const double* l2,
           const double* phi) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("bndu_");
  double *u = c.io(uk, d.n3), *v = c.io(vk, d.n3), *w = c.io(wk, d.n3);
  const double *a = c.in(l1, d.n3), *b = c.in(l2, d.n3), *p = c.in(phi, d.n3);
  bndu(d, u, v, w, *uwall, *dy, a, b, p, nID[kYMinus] < 0, nID[kYPlus] < 0);
  c.finish();
}

// srcs/bnd_periodic.f:4
void bnd_periodic_(const int* ni, const int* nj, const int* nk, double* qk) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("bnd_periodic_");
  periodic(d, c.io(qk, d.n3));
  c.finish();
}

// srcs/bnd_periodic.f:51
void bnd_neumann_(const int* nID, const int* ni, const int* nj, const int* nk, double* qk) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("bnd_neumann_");
  neumann(d, c.io(qk, d.n3), nID[kYMinus] < 0, nID[kYPlus] < 0);
  c.finish();
}

// srcs/bnd_comm.f:1-18: halo exchange between j-slabs; one GPU holds the
// whole domain, so there is nothing to exchange (the reference skips it for
// ipara=0 as well).  Multi-rank use is rejected.
void bnd_comm_(const int* ipara, const int* nID, const int* ni, const int* nj, const int* nk,
               const int* key, double* sendjb, double* recvjb, double* qk) {
  (void)nID; (void)ni; (void)nj; (void)nk; (void)key; (void)sendjb; (void)recvjb; (void)qk;
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: bnd_comm_: ipara=%d, only the single-domain case (0) is supported\n", *ipara);
    std::abort();
  }
}

// srcs/gnbc.f90:1-2
void gnbc_(const int* nID, const int* ni, const int* nj, const int* nk, const double* uk,
           const double* wk, const double* uwall, const double* theta_0_array,
           const double* surface_tension, const double* zeta_array, double* theta_array) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("gnbc_");
  const double *u = c.in(uk, d.n3), *w = c.in(wk, d.n3);
  const double *t0 = c.in(theta_0_array, d.n3), *z = c.in(zeta_array, d.n3);
  double* th = c.io(theta_array, d.n3);
  gnbc(d, u, w, *uwall, t0, *surface_tension, z, th, nID[kYMinus] < 0, nID[kYPlus] < 0, std::acos(-1.0));
  c.finish();
}

// srcs/bnd_contact_angle.f90:3
void bnd_contact_angle_(const int* nID, const int* ni, const int* nj, const int* nk, double* qk,
                        const double* theta_array, const double* dx, const double* dy,
                        const double* dz) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("bnd_contact_angle_");
  double* q = c.io(qk, d.n3);
  const double* th = c.in(theta_array, d.n3);
  contact_angle(d, q, th, *dy / *dx, *dy / *dz, nID[kYMinus] < 0, nID[kYPlus] < 0, std::acos(-1.0));
  c.finish();
}

}  // extern "C"
