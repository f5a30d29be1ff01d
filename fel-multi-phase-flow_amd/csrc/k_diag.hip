// k_diag.hip -- per-step diagnostics: phase centroid with lap tracking and
// phase-weighted velocity (device reductions, scalars returned to the host).
//
// Reference: srcs/cal_center.f:1-120, srcs/cal_vel.f:1-83 (called every step
// from srcs/main.f90:663-666).
#include <cmath>

#include "felgpu.hpp"

namespace fel {

// Phase-weighted velocity sums (FP64, of the values) of
// srcs/cal_vel.f:43-53; T is the velocity element type.  Returns the pinned
// read-back.
template <class T>
const double* vel_sums(const Dims d, const double* phi, const double* u, const double* v,
                       const double* w) {
  Call c("cal_vel_");
  const PT<T>* p = c.in_t<PT<T>>(phi, d.n3, "phi");
  const T *du = c.in_t<T>(u, d.n3, "u"), *dv = c.in_t<T>(v, d.n3, "v"), *dw = c.in_t<T>(w, d.n3, "w");
  double* s = dev_scalars(7);
  interior_reduce<4>(
      d,
      [=] __device__(int i, int j, int k, double* o) {
        const int m = d.ix(i, j, k);
        const double ph = val(p[m]);
        o[0] = ph;
        o[1] = ph * val(du[m]);
        o[2] = ph * val(dv[m]);
        o[3] = ph * val(dw[m]);
      },
      [=] __device__(const double* t, double* out) {
        for (int m = 0; m < 4; m++) out[m] = t[m];
      },
      s);
  const double* r = read_scalars(s, 4);
  c.finish();
  return r;
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/cal_center.f:1-3 (ipara must be 0)
void cal_center_(const int* ipara, const int* ID, const int* svall, const int* ni, const int* nj,
                 const int* nk, const double* xl_, const double* yl_, const double* zl_,
                 const double* dt, const double* rhog, const double* phi, double* center,
                 double* center_pre1, double* center_pre2, int* area_pre1, int* area_pre2, int* lap,
                 double* velocity) {
  (void)ipara; (void)zl_; (void)rhog;
  const Dims d = make_dims(*ni, *nj, *nk);
  const double xl = *xl_, yl = *yl_;
  const double delta = yl / (double)(*svall);
  const double s1 = xl * 0.25, s2 = xl * 0.75;
  if (*area_pre1 == 1 && *area_pre2 == 3) *lap = *lap + 1;
  const int ap1 = *area_pre1, lp = *lap, id = *ID, njl = *nj;
  double* s = dev_scalars(6);
  const double* r = nullptr;
  with_flow([&](auto t) {
    using P = PT<decltype(t)>;
    Call c("cal_center_");
    const P* p = c.in_t<P>(phi, d.n3, "phi");
    interior_reduce<4>(
        d,
        [=] __device__(int i, int j, int k, double* o) {
          const double ph = val(p[d.ix(i, j, k)]);
          double lx = delta * ((double)i - 0.5);
          const double ly = delta * ((double)(j + id * njl) - 0.5);
          const double lz = delta * ((double)k - 0.5);
          if (ap1 == 3 && lx < s1) lx = lx + xl;
          if (ap1 == 1 && lx > s2) lx = lx - xl;
          lx = lx + xl * (double)lp;
          o[0] = ph;
          o[1] = ph * lx;
          o[2] = ph * ly;
          o[3] = ph * lz;
        },
        [=] __device__(const double* t, double* out) {
          for (int m = 0; m < 4; m++) out[m] = t[m];
        },
        s);
    r = read_scalars(s, 4);
    c.finish();
  });
  center[0] = r[1] / r[0];
  center[1] = r[2] / r[0];
  center[2] = r[3] / r[0];
  *velocity = (center[0] - *center_pre1) / *dt;
  *center_pre2 = *center_pre1;
  *center_pre1 = center[0];
  double cc = center[0];
  *area_pre2 = *area_pre1;
  if (cc > xl) {
    do {
      cc = cc - xl;
    } while (!(cc < xl));
  }
  if (cc < s1) *area_pre1 = 1;
  if (cc > s2) *area_pre1 = 3;
  if (s1 <= cc && cc <= s2) *area_pre1 = 2;
}

// srcs/cal_vel.f:1-2 (ipara must be 0)
void cal_vel_(const int* ipara, const int* ni, const int* nj, const int* nk, const double* radius,
              const double* rhol, const double* rmul, const double* u, const double* v,
              const double* w, const double* phi, double* vel, double* re) {
  (void)ipara;
  const Dims d = make_dims(*ni, *nj, *nk);
  const double* r = nullptr;
  with_flow([&](auto t) { r = vel_sums<decltype(t)>(d, phi, u, v, w); });
  const double diameter = *radius * 2.0;
  for (int m = 0; m < 3; m++) {
    vel[m] = r[m + 1] / r[0];
    re[m] = *rhol * vel[m] * diameter / *rmul;
  }
}

}  // extern "C"
