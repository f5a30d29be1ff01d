// k_weno.hip -- momentum advection, WENO5 in derivative form.
//
// Reference: srcs/cal_advu_weno5.f:1-257 with srcs/weno5.h:1-71.  One
// thread per padded cell: interior cells get the three upwind-weighted
// derivatives of each velocity component, every other cell is zeroed (the
// reference's clearing pass, cal_advu_weno5.f:44-52), so the outputs are
// written exactly once.
#include "felgpu.hpp"

namespace fel {

// srcs/weno5.h:24-71 (nonlinear weights with the 1e-80 floor of
// cal_advu_weno5.f:31), expressions in the reference's order.
__device__ __forceinline__ double weno5(double dfm5, double dfm3, double dfm1, double dfp1,
                                        double dfp3, double dfp5, double vel00) {
  const double dfm_1 = dfm5 * 2.0 - dfm3 * 7.0 + dfm1 * 11.0;
  const double dfm_2 = -dfm3 + dfm1 * 5.0 + dfp1 * 2.0;
  const double dfm_3 = dfm1 * 2.0 + dfp1 * 5.0 - dfp3;
  const double dfp_1 = dfp5 * 2.0 - dfp3 * 7.0 + dfp1 * 11.0;
  const double dfp_2 = -dfp3 + dfp1 * 5.0 + dfm1 * 2.0;
  const double dfp_3 = dfp1 * 2.0 + dfm1 * 5.0 - dfm3;
  double a, b;
  a = dfm5 - dfm3 * 2.0 + dfm1;
  b = dfm5 - dfm3 * 4.0 + dfm1 * 3.0;
  const double sm_1 = a * a * 13.0 + b * b * 3.0;
  a = dfm3 - dfm1 * 2.0 + dfp1;
  b = dfm3 - dfp1;
  const double sm_2 = a * a * 13.0 + b * b * 3.0;
  a = dfm1 - dfp1 * 2.0 + dfp3;
  b = dfm1 * 3.0 - dfp1 * 4.0 + dfp3;
  const double sm_3 = a * a * 13.0 + b * b * 3.0;
  a = dfp5 - dfp3 * 2.0 + dfp1;
  b = dfp5 - dfp3 * 4.0 + dfp1 * 3.0;
  const double sp_1 = a * a * 13.0 + b * b * 3.0;
  a = dfp3 - dfp1 * 2.0 + dfm1;
  b = dfp3 - dfm1;
  const double sp_2 = a * a * 13.0 + b * b * 3.0;
  a = dfp1 - dfm1 * 2.0 + dfm3;
  b = dfp1 * 3.0 - dfm1 * 4.0 + dfm3;
  const double sp_3 = a * a * 13.0 + b * b * 3.0;
  double am_1 = sm_2 * sm_3, am_2 = sm_1 * sm_3, am_3 = sm_1 * sm_2;
  double ap_1 = sp_2 * sp_3, ap_2 = sp_1 * sp_3, ap_3 = sp_1 * sp_2;
  am_1 = am_1 * am_1;
  am_2 = am_2 * am_2;
  am_3 = am_3 * am_3;
  ap_1 = ap_1 * ap_1;
  ap_2 = ap_2 * ap_2;
  ap_3 = ap_3 * ap_3;
  am_2 = am_2 * 6.0;
  am_3 = am_3 * 3.0;
  ap_2 = ap_2 * 6.0;
  ap_3 = ap_3 * 3.0;
  const double deninv_am = 1.0 / (am_1 + am_2 + am_3 + 1.0e-80);
  const double deninv_ap = 1.0 / (ap_1 + ap_2 + ap_3 + 1.0e-80);
  const double wgm_1 = am_1 * deninv_am, wgm_3 = am_3 * deninv_am;
  const double wgp_1 = ap_1 * deninv_ap, wgp_3 = ap_3 * deninv_ap;
  const double wgm_2 = 1.0 - wgm_1 - wgm_3;
  const double wgp_2 = 1.0 - wgp_1 - wgp_3;
  const double velab = fabs(vel00);
  const double velm1 = vel00 + velab, velp1 = vel00 - velab;
  return (velm1 * (wgm_1 * dfm_1 + wgm_2 * dfm_2 + wgm_3 * dfm_3) +
          velp1 * (wgp_1 * dfp_1 + wgp_2 * dfp_2 + wgp_3 * dfp_3)) *
         (1.0 / 12.0);
}

__device__ __forceinline__ double weno_dir(const double* __restrict__ f, int c, int st, double vel00) {
  const double fm3 = f[c - 3 * st], fm2 = f[c - 2 * st], fm1 = f[c - st], f0 = f[c];
  const double fp1 = f[c + st], fp2 = f[c + 2 * st], fp3 = f[c + 3 * st];
  return weno5(-fm3 + fm2, -fm2 + fm1, -fm1 + f0, -f0 + fp1, -fp1 + fp2, -fp2 + fp3, vel00);
}

// All three components in one pass over the padded image.
void advect_weno5(const Dims d, double dxinv, double dyinv, double dzinv, const double* __restrict__ u,
                  const double* __restrict__ v, const double* __restrict__ w, double* __restrict__ au,
                  double* __restrict__ av, double* __restrict__ aw) {
  launch_box(Box{-2, d.ni + 3, -2, d.nj + 3, -2, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    if (i < 1 || i > d.ni || j < 1 || j > d.nj || k < 1 || k > d.nk) {
      au[c] = 0.0;
      av[c] = 0.0;
      aw[c] = 0.0;
      return;
    }
    const int X = 1, Y = d.sx, Z = d.sx * d.sy;
    double a, cm, cp;
    // u (cal_advu_weno5.f:80-120)
    cm = (u[c - X] + u[c]) * 0.5;
    cp = (u[c] + u[c + X]) * 0.5;
    a = weno_dir(u, c, X, (cm + cp) * 0.5) * dxinv;
    cm = (v[c - Y] + v[c + X - Y]) * 0.5;
    cp = (v[c] + v[c + X]) * 0.5;
    a = a + weno_dir(u, c, Y, (cm + cp) * 0.5) * dyinv;
    cm = (w[c - Z] + w[c + X - Z]) * 0.5;
    cp = (w[c] + w[c + X]) * 0.5;
    au[c] = a + weno_dir(u, c, Z, (cm + cp) * 0.5) * dzinv;
    // v (cal_advu_weno5.f:147-187)
    cm = (u[c - X] + u[c - X + Y]) * 0.5;
    cp = (u[c] + u[c + Y]) * 0.5;
    a = weno_dir(v, c, X, (cm + cp) * 0.5) * dxinv;
    cm = (v[c - Y] + v[c]) * 0.5;
    cp = (v[c] + v[c + Y]) * 0.5;
    a = a + weno_dir(v, c, Y, (cm + cp) * 0.5) * dyinv;
    cm = (w[c - Z] + w[c + Y - Z]) * 0.5;
    cp = (w[c] + w[c + Y]) * 0.5;
    av[c] = a + weno_dir(v, c, Z, (cm + cp) * 0.5) * dzinv;
    // w (cal_advu_weno5.f:214-254)
    cm = (u[c - X] + u[c - X + Z]) * 0.5;
    cp = (u[c] + u[c + Z]) * 0.5;
    a = weno_dir(w, c, X, (cm + cp) * 0.5) * dxinv;
    cm = (v[c - Y] + v[c - Y + Z]) * 0.5;
    cp = (v[c] + v[c + Z]) * 0.5;
    a = a + weno_dir(w, c, Y, (cm + cp) * 0.5) * dyinv;
    cm = (w[c - Z] + w[c]) * 0.5;
    cp = (w[c] + w[c + Z]) * 0.5;
    aw[c] = a + weno_dir(w, c, Z, (cm + cp) * 0.5) * dzinv;
  });
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/cal_advu_weno5.f:1-2
void cal_advu_weno5_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                     const double* dyinv, const double* dzinv, const double* u, const double* v,
                     const double* w, double* adv_u, double* adv_v, double* adv_w) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_advu_weno5_");
  const double* du = c.in(u, d.n3, "u");
  const double* dv = c.in(v, d.n3, "v");
  const double* dw = c.in(w, d.n3, "w");
  double* au = c.out(adv_u, d.n3, "adv_u");
  double* av = c.out(adv_v, d.n3, "adv_v");
  double* aw = c.out(adv_w, d.n3, "adv_w");
  prof_begin(2);
  advect_weno5(d, *dxinv, *dyinv, *dzinv, du, dv, dw, au, av, aw);
  prof_end(2);
  c.finish();
}

}  // extern "C"
