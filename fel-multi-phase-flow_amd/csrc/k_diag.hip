// k_diag.hip -- per-step diagnostics: phase centroid with lap tracking and
// phase-weighted velocity (device reductions, scalars returned to the host).
//
// Reference: srcs/cal_center.f:1-120, srcs/cal_vel.f:1-83 (called every step
// from srcs/main.f90:663-666).
#include <cmath>

#include "felgpu.hpp"
#include "felgpu_visc.hpp"

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

// Largest advective rate of srcs/caldt.f:33-84 over the interior:
// max(|u| dxinv, |v| dyinv on j <= jen, |w| dzinv, 1e-20).  A maximum does
// not depend on the order of evaluation, so the two-stage reduction gives
// the reference's value exactly.  T: velocity element type (values only).
template <class T>
__global__ void __launch_bounds__(256) advmax_stage1(Dims d, int jen, double dxinv, double dyinv,
                                                     double dzinv, const T* u, const T* v,
                                                     const T* w, double* part) {
  __shared__ double sh[256];
  double m = 1.0e-20;
  const int nxy = d.ni * d.nj, n = nxy * d.nk;
  for (int q = (int)(blockIdx.x * 256 + threadIdx.x); q < n; q += (int)(gridDim.x * 256)) {
    const int k = q / nxy, r = q - k * nxy, j = r / d.ni, i = r - j * d.ni;
    const int c = d.ix(i + 1, j + 1, k + 1);
    m = fmax(fabs(val(u[c]) * dxinv), m);
    if (j + 1 <= jen) m = fmax(fabs(val(v[c]) * dyinv), m);
    m = fmax(fabs(val(w[c]) * dzinv), m);
  }
  const int t = (int)threadIdx.x;
  sh[t] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) advmax_stage2(int nb, const double* part, double* out) {
  __shared__ double sh[256];
  const int t = (int)threadIdx.x;
  double m = 1.0e-20;
  for (int b = t; b < nb; b += 256) m = fmax(m, part[b]);
  sh[t] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  if (t == 0) out[0] = sh[0];
}

template <class T>
double advmax_t(const Dims d, int jen, double dxinv, double dyinv, double dzinv, const double* u,
                const double* v, const double* w) {
  Call c("caldt_");
  const T *du = c.in_t<T>(u, d.n3, "u"), *dv = c.in_t<T>(v, d.n3, "v"), *dw = c.in_t<T>(w, d.n3, "w");
  double* part = static_cast<double*>(scratch(15, sizeof(double) * kRedBlocks));
  double* out = dev_scalars(8);
  advmax_stage1<T><<<kRedBlocks, 256, 0, stream()>>>(d, jen, dxinv, dyinv, dzinv, du, dv, dw, part);
  FEL_CHECK(hipGetLastError());
  advmax_stage2<<<1, 256, 0, stream()>>>(kRedBlocks, part, out);
  FEL_CHECK(hipGetLastError());
  const double r = read_scalars(out, 1)[0];
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

// srcs/caldt.f:1-116 (ipara must be 0): the CFL time step.  The largest
// advective rate is a device reduction; the three limits and the update of
// time are evaluated on the host with the reference's expressions.
void caldt_(const int* ipara, const int* nID, const int* ID, const int* ndiv, const int* ni,
            const int* nj, const int* nk, const int* nstep, const int* imon_t, const double* dxinv,
            const double* dyinv, const double* dzinv, const double* cfl, const double* rhol,
            const double* rhog, const double* rmul, const double* rmug,
            const double* surface_tension, const double* u, const double* v, const double* w,
            double* dt, double* time) {
  (void)ndiv;
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: caldt_: ipara=%d, only the single-domain case (0) is supported\n", *ipara);
    std::abort();
  }
  const Dims d = make_dims(*ni, *nj, *nk);
  const int jen = nID[kYPlus] < 0 ? *nj - 1 : *nj;
  double advmax = 0.0;
  with_flow([&](auto t) {
    advmax = advmax_t<decltype(t)>(d, jen, *dxinv, *dyinv, *dzinv, u, v, w);
  });
  const double dcell = std::pow(*dxinv * *dyinv * *dzinv, -1.0 / 3.0);
  const double rnu = std::fmax(*rmul / *rhol, *rmug / *rhog);
  const double st_rhominv = 2.0 * *surface_tension / (*rhol + *rhog);
  const double dt1 = *cfl / advmax;
  const double dt2 = *cfl * 0.5 * (dcell * dcell) / rnu * 1.0e2;
  const double dt3 = *cfl * std::sqrt((dcell * dcell * dcell) / st_rhominv);
  double step = std::fmin(dt1, dt2);
  step = std::fmin(step, dt3);
  *dt = step;
  *time = *time + step;
  if (*imon_t > 0 && *nstep % *imon_t == 0 && *ID == 0) {
    char a[5][32];
    std::printf("time=%s dt=%s dt_adv=%s dt_vis=%s dt_srf=%s\n", fortran_e(a[0], *time, 17, 10),
                fortran_e(a[1], step, 17, 10), fortran_e(a[2], dt1, 17, 10),
                fortran_e(a[3], dt2, 17, 10), fortran_e(a[4], dt3, 17, 10));
  }
}

}  // extern "C"
