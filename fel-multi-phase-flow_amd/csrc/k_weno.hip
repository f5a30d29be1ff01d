// k_weno.hip -- momentum advection, WENO5 in derivative form.
//
// Reference: srcs/cal_advu_weno5.f:1-257 with srcs/weno5.h:1-71.  Interior
// cells get the three upwind-weighted derivatives of each velocity
// component; the ghost shell is zeroed (the reference's clearing pass,
// cal_advu_weno5.f:44-52), so the outputs are written exactly once.
#include <cstdlib>
#include <vector>

#include "felgpu.hpp"

namespace fel {

// srcs/weno5.h:24-71 (nonlinear weights with the 1e-80 floor of
// cal_advu_weno5.f:31), expressions in the reference's order.
//
// T = double is the reference's FP64 arithmetic.  T = float (the FP32 mode)
// evaluates the same weights from differences scaled by the power of two
// that brings their largest magnitude to [1, 2): the weights are homogeneous
// of degree 0 in that scale, so the scaling changes nothing but keeps the
// degree-8 products of the smoothness indicators inside the FP32 range; the
// stencil derivatives use the unscaled differences.
// 1/x for the weights: IEEE division in FP64 (and the tangent mode); in the
// FP32 mode the hardware reciprocal (v_rcp_f32, 1 ulp), whose rounding is
// far below the FP32 mode's accuracy bar and which replaces a ten-instruction
// division sequence in a compute-bound kernel.  The weight sums are floored
// at 1e-30 and, after the scaling, lie far above the smallest normal float.
template <class T>
__device__ __forceinline__ T weight_rcp(T x) {
  if constexpr (is_f32<T>) return __builtin_amdgcn_rcpf(x);
  else return T(1.0) / x;
}

// One upwind-biased side of srcs/weno5.h: w1*df1 + w2*df2 + w3*df3 of the
// side whose differences are (d5, d3, d1, e1, e3) -- the m side (dfm5,
// dfm3, dfm1, dfp1, dfp3), the p side mirrored (dfp5, dfp3, dfp1, dfm1,
// dfm3) -- with the candidate derivatives from those differences and the
// smoothness indicators from their scaled copies (s5, s3, s1, t1, t3);
// expressions in the reference's order.
template <class T>
__device__ __forceinline__ T weno_side(T d5, T d3, T d1, T e1, T e3, T s5, T s3, T s1, T t1,
                                       T t3) {
  const T c2 = T(2.0), c3 = T(3.0), c4 = T(4.0), c5 = T(5.0), c7 = T(7.0), c11 = T(11.0),
          c13 = T(13.0);
  const T df_1 = d5 * c2 - d3 * c7 + d1 * c11;
  const T df_2 = -d3 + d1 * c5 + e1 * c2;
  const T df_3 = d1 * c2 + e1 * c5 - e3;
  T a, b;
  a = s5 - s3 * c2 + s1;
  b = s5 - s3 * c4 + s1 * c3;
  const T sm_1 = a * a * c13 + b * b * c3;
  a = s3 - s1 * c2 + t1;
  b = s3 - t1;
  const T sm_2 = a * a * c13 + b * b * c3;
  a = s1 - t1 * c2 + t3;
  b = s1 * c3 - t1 * c4 + t3;
  const T sm_3 = a * a * c13 + b * b * c3;
  T a_1 = sm_2 * sm_3, a_2 = sm_1 * sm_3, a_3 = sm_1 * sm_2;
  a_1 = a_1 * a_1;
  a_2 = a_2 * a_2;
  a_3 = a_3 * a_3;
  a_2 = a_2 * T(6.0);
  a_3 = a_3 * c3;
  const T deninv = weight_rcp(a_1 + a_2 + a_3 + guard<T>(1.0e-80));
  const T w_1 = a_1 * deninv, w_3 = a_3 * deninv;
  const T w_2 = T(1.0) - w_1 - w_3;
  return w_1 * df_1 + w_2 * df_2 + w_3 * df_3;
}

// srcs/weno5.h:24-71: (velm1 * m side + velp1 * p side) / 12 with velm1 =
// vel + |vel| and velp1 = vel - |vel|.  For vel != 0 one of the two factors
// is an exact zero and its side is not evaluated: the sum is the other
// term's value, the same bits (up to the sign of a zero result), and in the
// tangent mode the zero factor's derivative is zero too.  vel = 0 (or NaN)
// evaluates both sides as written.
template <class T>
__device__ __forceinline__ T weno5(T dfm5, T dfm3, T dfm1, T dfp1, T dfp3, T dfp5, T vel00) {
  T m5 = dfm5, m3 = dfm3, m1 = dfm1, p1 = dfp1, p3 = dfp3, p5 = dfp5;
  if constexpr (is_f32<T>) {
    const T sc = fmax(fmax(fmax(fabs(m5), fabs(m3)), fmax(fabs(m1), fabs(p1))),
                      fmax(fabs(p3), fabs(p5)));
    // scale by the exact power of two that brings the largest difference to
    // [1, 2): the weights are invariant to it, and unlike 1/sc (which
    // overflows for sc below ~2.9e-39) it stays finite for subnormal sc
    const int e = sc > T(0.0) ? ilogbf(sc) : 0;
    m5 = ldexpf(m5, -e); m3 = ldexpf(m3, -e); m1 = ldexpf(m1, -e);
    p1 = ldexpf(p1, -e); p3 = ldexpf(p3, -e); p5 = ldexpf(p5, -e);
  }
  const T velab = fabs(vel00);
  const T velm1 = vel00 + velab, velp1 = vel00 - velab;
  const T r12 = T(1.0) / T(12.0);
  if (vel00 > T(0.0))
    return (velm1 * weno_side(dfm5, dfm3, dfm1, dfp1, dfp3, m5, m3, m1, p1, p3)) * r12;
  if (vel00 < T(0.0))
    return (velp1 * weno_side(dfp5, dfp3, dfp1, dfm1, dfm3, p5, p3, p1, m1, m3)) * r12;
  return (velm1 * weno_side(dfm5, dfm3, dfm1, dfp1, dfp3, m5, m3, m1, p1, p3) +
          velp1 * weno_side(dfp5, dfp3, dfp1, dfm1, dfm3, p5, p3, p1, m1, m3)) *
         r12;
}

template <class T>
__device__ __forceinline__ T weno_dir(const T* __restrict__ f, int c, int st, T vel00) {
  const T fm3 = f[c - 3 * st], fm2 = f[c - 2 * st], fm1 = f[c - st], f0 = f[c];
  const T fp1 = f[c + st], fp2 = f[c + 2 * st], fp3 = f[c + 3 * st];
  return weno5(-fm3 + fm2, -fm2 + fm1, -fm1 + f0, -f0 + fp1, -fp1 + fp2, -fp2 + fp3, vel00);
}

// The WENO5 derivative from a register queue f(k-3..k+3) of the cell's own
// column (the z direction), same differences as weno_dir.
template <class T>
__device__ __forceinline__ T weno_q(const T* q, T vel00) {
  return weno5(-q[0] + q[1], -q[1] + q[2], -q[2] + q[3], -q[3] + q[4], -q[4] + q[5],
               -q[5] + q[6], vel00);
}

// Interior cells, k-marching: one thread per (i,j) column over kWenoK planes.
// The own column's u, v, w at k-3..k+3 live in register queues, so the
// z-direction stencils load one new plane per step instead of seven; the
// in-plane (x, y) stencils and the neighbour-column transport velocities are
// loaded from the same plane, which the block's other lanes share in cache.
// Every value entering the arithmetic is the same double as in the
// one-thread-per-cell form, so the results are identical.
// Planes per thread (measured at 512x256x128: 745 us in FP64 and 407 us in
// FP32 with 4, against 759 and 430 with 16; 2 and 8 within 3%).
constexpr int kWenoK = 4;

template <class T>
__global__ void __launch_bounds__(256) weno_march_kernel(Dims d, T dxinv, T dyinv, T dzinv,
                                                         const T* __restrict__ u,
                                                         const T* __restrict__ v,
                                                         const T* __restrict__ w,
                                                         T* __restrict__ au, T* __restrict__ av,
                                                         T* __restrict__ aw, int gx, int gy) {
  const int L = (int)blockIdx.x;
  const int bx = L % gx, t = L / gx, by = t % gy, bz = t / gy;
  const int i = 1 + bx * 64 + (int)threadIdx.x, j = 1 + by * 4 + (int)threadIdx.y;
  if (i > d.ni || j > d.nj) return;
  const int k0 = 1 + bz * kWenoK, k1 = min(d.nk, k0 + kWenoK - 1);
  const int X = 1, Y = d.sx, Z = d.sx * d.sy;
  int c = d.ix(i, j, k0);
  T qu[7], qv[7], qw[7];
  const T h = T(0.5);
#pragma unroll
  for (int m = 0; m < 6; m++) {
    qu[m + 1] = u[c + (m - 3) * Z];
    qv[m + 1] = v[c + (m - 3) * Z];
    qw[m + 1] = w[c + (m - 3) * Z];
  }
  for (int k = k0; k <= k1; k++, c += Z) {
#pragma unroll
    for (int m = 0; m < 6; m++) {
      qu[m] = qu[m + 1];
      qv[m] = qv[m + 1];
      qw[m] = qw[m + 1];
    }
    qu[6] = u[c + 3 * Z];
    qv[6] = v[c + 3 * Z];
    qw[6] = w[c + 3 * Z];
    T a, cm, cp;
    // u (cal_advu_weno5.f:80-120)
    cm = (u[c - X] + qu[3]) * h;
    cp = (qu[3] + u[c + X]) * h;
    a = weno_dir(u, c, X, (cm + cp) * h) * dxinv;
    cm = (v[c - Y] + v[c + X - Y]) * h;
    cp = (qv[3] + v[c + X]) * h;
    a = a + weno_dir(u, c, Y, (cm + cp) * h) * dyinv;
    cm = (qw[2] + w[c + X - Z]) * h;
    cp = (qw[3] + w[c + X]) * h;
    au[c] = a + weno_q(qu, (cm + cp) * h) * dzinv;
    // v (cal_advu_weno5.f:147-187)
    cm = (u[c - X] + u[c - X + Y]) * h;
    cp = (qu[3] + u[c + Y]) * h;
    a = weno_dir(v, c, X, (cm + cp) * h) * dxinv;
    cm = (v[c - Y] + qv[3]) * h;
    cp = (qv[3] + v[c + Y]) * h;
    a = a + weno_dir(v, c, Y, (cm + cp) * h) * dyinv;
    cm = (qw[2] + w[c + Y - Z]) * h;
    cp = (qw[3] + w[c + Y]) * h;
    av[c] = a + weno_q(qv, (cm + cp) * h) * dzinv;
    // w (cal_advu_weno5.f:214-254)
    cm = (u[c - X] + u[c - X + Z]) * h;
    cp = (qu[3] + qu[4]) * h;
    a = weno_dir(w, c, X, (cm + cp) * h) * dxinv;
    cm = (v[c - Y] + v[c - Y + Z]) * h;
    cp = (qv[3] + qv[4]) * h;
    a = a + weno_dir(w, c, Y, (cm + cp) * h) * dyinv;
    cm = (qw[2] + qw[3]) * h;
    cp = (qw[3] + qw[4]) * h;
    aw[c] = a + weno_q(qw, (cm + cp) * h) * dzinv;
  }
}

// The reference clears the outputs over the padded image and fills the
// interior (cal_advu_weno5.f:44-52): interior by the marching kernel, the
// ghost shell zeroed separately, so every output element is written once.
template <class T>
void advect_weno5(const Dims d, double dxinv, double dyinv, double dzinv, const T* __restrict__ u,
                  const T* __restrict__ v, const T* __restrict__ w, T* __restrict__ au,
                  T* __restrict__ av, T* __restrict__ aw) {
  const int gx = (d.ni + 63) / 64, gy = (d.nj + 3) / 4, gz = (d.nk + kWenoK - 1) / kWenoK;
  weno_march_kernel<T><<<dim3(gx * gy * gz), dim3(64, 4, 1), 0, stream()>>>(
      d, T(dxinv), T(dyinv), T(dzinv), u, v, w, au, av, aw, gx, gy);
  FEL_CHECK(hipGetLastError());
  launch_shell(d, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    au[c] = T(0.0);
    av[c] = T(0.0);
    aw[c] = T(0.0);
  });
}

// Result reuse.  The timestep asks for the advection of (u,v,w) and then of
// (uo,vo,wo) (srcs/main.f90:532-533); uo..wo are copies of the previous
// step's u..w (main.f90:686-688), whose advection the previous step already
// computed into adv_u..adv_w.  Content versions identify that case exactly:
// before adv_u..adv_w are overwritten, their buffers (holding the advection
// of the inputs whose versions they were computed from) are parked in a
// stash, and a later request whose inputs carry those same versions gets the
// parked buffers swapped in instead of a recomputation -- the same bits, no
// kernel.  Anything not provably identical is recomputed.
namespace {

struct WenoKey {
  uint64_t ver[3];
  int ni, nj, nk;
  double dx, dy, dz;
  bool operator==(const WenoKey& o) const {
    return ver[0] == o.ver[0] && ver[1] == o.ver[1] && ver[2] == o.ver[2] && ni == o.ni &&
           nj == o.nj && nk == o.nk && dx == o.dx && dy == o.dy && dz == o.dz;
  }
};

struct WenoHeld {  // what one output triple holds, valid while its versions are unchanged
  const void* out[3];
  uint64_t out_ver[3];
  WenoKey key;
};

struct WenoStash {
  bool valid = false;
  WenoKey key{};
  void* buf[3] = {nullptr, nullptr, nullptr};
  size_t bytes = 0;
};

std::vector<WenoHeld> g_held;
WenoStash g_stash;

bool reuse_disabled() { return !result_reuse(); }

WenoHeld* held_for(const void* const* out) {
  for (auto& h : g_held)
    if (h.out[0] == out[0] && h.out[1] == out[1] && h.out[2] == out[2]) return &h;
  return nullptr;
}

void record_held(const void* const* out, size_t n, const WenoKey& k) {
  WenoHeld* h = held_for(out);
  if (!h) {
    g_held.push_back(WenoHeld{});
    h = &g_held.back();
    for (int m = 0; m < 3; m++) h->out[m] = out[m];
  }
  for (int m = 0; m < 3; m++) h->out_ver[m] = content_version(out[m], n);
  h->key = k;
}

}  // namespace

template <class T>
void weno_compute(const Dims d, double dxinv, double dyinv, double dzinv, const double* u,
                  const double* v, const double* w, double* adv_u, double* adv_v, double* adv_w) {
  Call c("cal_advu_weno5_");
  const T* du = c.in_t<T>(u, d.n3, "u");
  const T* dv = c.in_t<T>(v, d.n3, "v");
  const T* dw = c.in_t<T>(w, d.n3, "w");
  T* au = c.out_t<T>(adv_u, d.n3, "adv_u");
  T* av = c.out_t<T>(adv_v, d.n3, "adv_v");
  T* aw = c.out_t<T>(adv_w, d.n3, "adv_w");
  advect_weno5<T>(d, dxinv, dyinv, dzinv, du, dv, dw, au, av, aw);
  c.finish();
}

void weno_release() {
  for (int m = 0; m < 3; m++)
    if (g_stash.buf[m]) FEL_CHECK(hipFree(g_stash.buf[m]));
  g_stash = WenoStash{};
  g_held.clear();
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/cal_advu_weno5.f:1-2
void cal_advu_weno5_(const int* ni, const int* nj, const int* nk, const double* dxinv,
                     const double* dyinv, const double* dzinv, const double* u, const double* v,
                     const double* w, double* adv_u, double* adv_v, double* adv_w) {
  const Dims d = make_dims(*ni, *nj, *nk);
  const size_t n = (size_t)d.n3;
  const void* out[3] = {adv_u, adv_v, adv_w};
  const void* in[3] = {u, v, w};
  WenoKey key{{content_version(u, n), content_version(v, n), content_version(w, n)},
              *ni, *nj, *nk, *dxinv, *dyinv, *dzinv};
  bool track = !reuse_disabled();
  for (int m = 0; m < 3 && track; m++) {
    if (key.ver[m] == 0 || content_version(out[m], n) == 0) track = false;
    for (int q = 0; q < 3; q++)
      if (out[m] == in[q] || (q != m && out[m] == out[q])) track = false;
  }
  if (track) {
    WenoHeld* h = held_for(out);
    bool h_valid = h != nullptr;
    for (int m = 0; m < 3 && h_valid; m++)
      if (h->out_ver[m] != content_version(out[m], n)) h_valid = false;
    if (g_stash.valid && g_stash.key == key) {
      // the parked buffers hold exactly this result: swap them in
      for (int m = 0; m < 3; m++) g_stash.buf[m] = exchange_device_buffer(out[m], n, g_stash.buf[m]);
      g_stash.valid = h_valid;
      if (h_valid) g_stash.key = h->key;
      Call c("cal_advu_weno5_");  // the outputs change: fresh versions
      with_flow([&](auto t) {
        for (int m = 0; m < 3; m++) c.out_t<decltype(t)>(out[m], n, "adv");
      });
      c.finish();
      record_held(out, n, key);
      return;
    }
    if (h_valid) {
      // park what the outputs hold before overwriting them
      const size_t bytes = n * (size_t)elem_size(out[0]);
      if (g_stash.bytes != bytes) {
        for (int m = 0; m < 3; m++) {
          if (g_stash.buf[m]) FEL_CHECK(hipFree(g_stash.buf[m]));
          FEL_CHECK(hipMalloc(&g_stash.buf[m], bytes));
        }
        g_stash.bytes = bytes;
      }
      for (int m = 0; m < 3; m++) g_stash.buf[m] = exchange_device_buffer(out[m], n, g_stash.buf[m]);
      g_stash.valid = true;
      g_stash.key = h->key;
    }
  }
  with_flow([&](auto t) {
    weno_compute<decltype(t)>(d, *dxinv, *dyinv, *dzinv, u, v, w, adv_u, adv_v, adv_w);
  });
  if (track) record_held(out, n, key);
}

}  // extern "C"
