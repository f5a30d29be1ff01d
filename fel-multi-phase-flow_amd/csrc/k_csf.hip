// k_csf.hip -- continuum-surface-force surface tension: face gradients,
// staggered unit normals, curvature and force.
//
// Reference: srcs/cal_grad_p2uvw.f, srcs/calrn.f, srcs/calrkap.f,
// srcs/calfst.f (called from srcs/main.f90:545-570).
#include <cmath>
#include <cstdlib>

#include "felgpu.hpp"

namespace fel {

// T: element type of the gradients, normals, curvature and forces (double,
// float in the FP32 mode, Dual in the tangent mode); the volume fraction is
// in the phase type PT<T>.

// srcs/cal_grad_p2uvw.f:25-43 (per-j dyinv from dyinv_array)
template <class T, class P = PT<T>>
void grad_p2uvw(const Dims d, int ID, double dxinv, const double* dyt_dev, double dyu, double dzinv,
                const P* q, T* qx, T* qy, T* qz) {
  launch_box(Box{-2, d.ni + 2, -2, d.nj + 2, -2, d.nk + 2}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k);
    const double dyinv = dyt_dev ? dyt_dev[j + d.nj * ID + 2] : dyu;
    const P q0 = q[c];
    qx[c] = T((-q0 + q[c + 1]) * dxinv);
    qy[c] = T((-q0 + q[c + d.sx]) * dyinv);
    qz[c] = T((-q0 + q[c + d.sx * d.sy]) * dzinv);
  });
}

// srcs/calrn.f:23-87.  One thread per target column (ii,jj), marching over
// a chunk of kk planes: the four source values of plane kk are kept for the
// next plane, so each plane is loaded once per array instead of twice, and
// the 8-term sums keep the reference's order (plane kk-kp first).  Targets
// outside the shifted loop range keep the cleared value 0.
constexpr int kRnChunk = 16;

template <class T>
__global__ void __launch_bounds__(256) calrn_kernel(Dims d, T verysmall, int ip, int jp, int kp,
                                                    const T* __restrict__ px,
                                                    const T* __restrict__ py,
                                                    const T* __restrict__ pz, T* __restrict__ rn,
                                                    int gx, int gy, int ntiles) {
  const int L = xcd_tile((int)blockIdx.x, ntiles);
  const int bx = L % gx, t = L / gx, by = t % gy, bz = t / gy;
  const int ii = -2 + bx * 64 + (int)threadIdx.x;
  const int jj = -2 + by * 4 + (int)threadIdx.y;
  if (ii > d.ni + 3 || jj > d.nj + 3) return;
  const int n3 = d.n3, Z = d.sx * d.sy;
  const int k0 = -2 + bz * kRnChunk, k1 = min(d.nk + 3, k0 + kRnChunk - 1);
  // tangent mode: gradients below 1e-10 |dxinv| count as flat (no interface)
  const double flat2 = (val(verysmall) * 1.0e20) * (val(verysmall) * 1.0e20);
  const bool col_ok = ii >= -2 + ip && jj >= -2 + jp;
  const int X = ip, Y = jp * d.sx;
  T pa[3][4];  // previous plane: (i,j), (ii,j), (i,jj), (ii,jj)
  bool have_prev = false;
  for (int kk = k0; kk <= k1; kk++) {
    const int tgt = d.ix(ii, jj, kk);
    if (!col_ok || kk < -2 + kp) {
      rn[tgt] = T(0.0);
      rn[n3 + tgt] = T(0.0);
      rn[2 * n3 + tgt] = T(0.0);
      have_prev = false;
      continue;
    }
    const int c = tgt - X - Y;  // (i, j, kk) with i = ii-ip, j = jj-jp
    T cur[3][4];
    const T* a3[3] = {px, py, pz};
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const T* p = a3[m];
      cur[m][0] = p[c];
      cur[m][1] = p[c + X];
      cur[m][2] = p[c + Y];
      cur[m][3] = p[c + X + Y];
    }
    if (kp == 0 || !have_prev) {
      if (kp == 0) {
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
          for (int q = 0; q < 4; q++) pa[m][q] = cur[m][q];
      } else {
        const int cb = c - Z;
#pragma unroll
        for (int m = 0; m < 3; m++) {
          const T* p = a3[m];
          pa[m][0] = p[cb];
          pa[m][1] = p[cb + X];
          pa[m][2] = p[cb + Y];
          pa[m][3] = p[cb + X + Y];
        }
      }
    }
    T r[3];
#pragma unroll
    for (int m = 0; m < 3; m++)
      r[m] = (+pa[m][0] + pa[m][1] + pa[m][2] + pa[m][3] + cur[m][0] + cur[m][1] + cur[m][2] +
              cur[m][3]) *
             T(0.125);
    flat_zero(r, flat2);
    const T arn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + verysmall;
    const T inv = T(1.0) / arn;
    rn[tgt] = r[0] * inv;
    rn[n3 + tgt] = r[1] * inv;
    rn[2 * n3 + tgt] = r[2] * inv;
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int q = 0; q < 4; q++) pa[m][q] = cur[m][q];
    have_prev = true;
  }
}

template <class T>
void calrn(const Dims d, double verysmall, int ip, int jp, int kp, const T* px, const T* py,
           const T* pz, T* rn) {
  const int gx = (d.sx + 63) / 64, gy = (d.sy + 3) / 4, gz = (d.sz + kRnChunk - 1) / kRnChunk;
  const int n = gx * gy * gz;
  calrn_kernel<T><<<dim3(n), dim3(64, 4, 1), 0, stream()>>>(d, T(verysmall), ip, jp, kp, px, py, pz,
                                                            rn, gx, gy, xcd_arg(2, n));
  FEL_CHECK(hipGetLastError());
}

// The four staggerings the timestep asks for (srcs/main.f90:549-552 and
// 555-558) in one pass: (ip,jp,kp) = (1,1,1), (0,0,1), (0,1,0), (1,0,0).
// Every staggering's 8 source points lie in the 2x2x2 block of cells
// (ii-1..ii, jj-1..jj, kk-1..kk) of the target (ii,jj,kk); the block's
// k-1 plane is the previous iteration's k plane.  Each staggering sums its
// 8 values in calrn_kernel's order, so the four outputs are identical to
// four separate calls.
constexpr int kStag[4][3] = {{1, 1, 1}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}};

template <class T>
__global__ void __launch_bounds__(256) calrn4_kernel(Dims d, T verysmall, const T* __restrict__ px,
                                                     const T* __restrict__ py,
                                                     const T* __restrict__ pz, T* __restrict__ r0,
                                                     T* __restrict__ r1, T* __restrict__ r2,
                                                     T* __restrict__ r3, int gx, int gy) {
  const int L = (int)blockIdx.x;
  const int bx = L % gx, t = L / gx, by = t % gy, bz = t / gy;
  const int ii = -2 + bx * 64 + (int)threadIdx.x;
  const int jj = -2 + by * 4 + (int)threadIdx.y;
  if (ii > d.ni + 3 || jj > d.nj + 3) return;
  const int n3 = d.n3, X = 1, Y = d.sx;
  const int k0 = -2 + bz * kRnChunk, k1 = min(d.nk + 3, k0 + kRnChunk - 1);
  // tangent mode: gradients below 1e-10 |dxinv| count as flat (no interface)
  const double flat2 = (val(verysmall) * 1.0e20) * (val(verysmall) * 1.0e20);
  const bool xm = ii >= -1, ym = jj >= -1;  // the i-1 / j-1 neighbours exist
  T* out[4] = {r0, r1, r2, r3};
  const T* a3[3] = {px, py, pz};
  // V[m][a][b][c]: array m at (ii-1+a, jj-1+b, kk-1+c)
  T V[3][2][2][2];
  auto load_plane = [&](int kk, int cpl) {
    const int c = d.ix(ii, jj, kk);
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const T* p = a3[m];
      V[m][1][1][cpl] = p[c];
      V[m][0][1][cpl] = xm ? p[c - X] : T(0.0);
      V[m][1][0][cpl] = ym ? p[c - Y] : T(0.0);
      V[m][0][0][cpl] = (xm && ym) ? p[c - X - Y] : T(0.0);
    }
  };
  if (k0 - 1 >= -2) {
    load_plane(k0 - 1, 1);
  }
  for (int kk = k0; kk <= k1; kk++) {
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) V[m][a][b][0] = V[m][a][b][1];
    load_plane(kk, 1);
    const int tgt = d.ix(ii, jj, kk);
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int ip = kStag[s][0], jp = kStag[s][1], kp = kStag[s][2];
      T* rn = out[s];
      if (ii < -2 + ip || jj < -2 + jp || kk < -2 + kp) {
        rn[tgt] = T(0.0);
        rn[n3 + tgt] = T(0.0);
        rn[2 * n3 + tgt] = T(0.0);
        continue;
      }
      const int ai = 1 - ip, bj = 1 - jp, cp = 1 - kp;
      T r[3];
#pragma unroll
      for (int m = 0; m < 3; m++)
        r[m] = (+V[m][ai][bj][cp] + V[m][1][bj][cp] + V[m][ai][1][cp] + V[m][1][1][cp] +
                V[m][ai][bj][1] + V[m][1][bj][1] + V[m][ai][1][1] + V[m][1][1][1]) *
               T(0.125);
      flat_zero(r, flat2);
      const T arn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + verysmall;
      const T inv = T(1.0) / arn;
      rn[tgt] = r[0] * inv;
      rn[n3 + tgt] = r[1] * inv;
      rn[2 * n3 + tgt] = r[2] * inv;
    }
  }
}

template <class T>
void calrn4(const Dims d, double verysmall, const T* px, const T* py, const T* pz, T* const* rn) {
  const int gx = (d.sx + 63) / 64, gy = (d.sy + 3) / 4, gz = (d.sz + kRnChunk - 1) / kRnChunk;
  calrn4_kernel<T><<<dim3(gx * gy * gz), dim3(64, 4, 1), 0, stream()>>>(
      d, T(verysmall), px, py, pz, rn[0], rn[1], rn[2], rn[3], gx, gy);
  FEL_CHECK(hipGetLastError());
}

// The four staggered normals of calrn4_kernel and, from them, the
// curvature of srcs/calrkap.f:42-98 in one pass.  A 64 x 8 block covers
// 62 x 6 target columns plus a one-column / one-row halo on each side: the
// halo threads compute the normals their neighbours' curvature reads, and
// the six x / y neighbour values go through LDS.  Marching in k, the
// curvature of plane kk-1 is formed once the normals of plane kk (its +z
// neighbour) exist.  The normals are calrn4_kernel's values and the
// curvature is calrkap's expression on them, so the outputs equal those of
// the two separate passes bit for bit.
constexpr int kKapTx = 62, kKapTy = 6;

template <class T>
__global__ void __launch_bounds__(512) calrn4_kap_kernel(
    Dims d, T verysmall, T dxinv, T dyinv, T dzinv, const T* __restrict__ px,
    const T* __restrict__ py, const T* __restrict__ pz, T* __restrict__ r0, T* __restrict__ r1,
    T* __restrict__ r2, T* __restrict__ r3, T* __restrict__ rk, int gx, int gy) {
  // [plane parity][rpp x, r12 x, r13 x, rpp y, r12 y, r23 y][row][lane]
  __shared__ T xs[2][6][kKapTy + 2][64];
  const int L = (int)blockIdx.x;
  const int bx = L % gx, t = L / gx, by = t % gy, bz = t / gy;
  const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int ii = -3 + bx * kKapTx + tx;
  const int jj = -3 + by * kKapTy + ty;
  const bool vt = ii >= -2 && ii <= ni + 3 && jj >= -2 && jj <= nj + 3;
  const bool own = vt && tx >= 1 && tx <= kKapTx && ty >= 1 && ty <= kKapTy;
  const int n3 = d.n3, X = 1, Y = d.sx;
  const int k0 = -2 + bz * kRnChunk, k1 = min(nk + 3, k0 + kRnChunk - 1);
  const double flat2 = (val(verysmall) * 1.0e20) * (val(verysmall) * 1.0e20);
  const bool xm = ii >= -1, ym = jj >= -1;
  T* out[4] = {r0, r1, r2, r3};
  const T* a3[3] = {px, py, pz};
  T V[3][2][2][2];
#pragma unroll
  for (int m = 0; m < 3; m++)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) V[m][a][b][0] = V[m][a][b][1] = T(0.0);
  auto load_plane = [&](int kk) {
    const int c = d.ix(ii, jj, kk);
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const T* p = a3[m];
      V[m][1][1][1] = p[c];
      V[m][0][1][1] = xm ? p[c - X] : T(0.0);
      V[m][1][0][1] = ym ? p[c - Y] : T(0.0);
      V[m][0][0][1] = (xm && ym) ? p[c - X - Y] : T(0.0);
    }
  };
  // the gradient plane below the first normal plane (k0 - 1)
  if (vt && k0 - 2 >= -2) load_plane(k0 - 2);
  // own normals of the previous plane (kk-1) that its curvature reads, and
  // the z components of rn_13, rn_23 two planes back (kk-2)
  T o_pp0 = T(0.0), o_pp1 = T(0.0), o_pp2 = T(0.0), o_120 = T(0.0), o_121 = T(0.0);
  T o_132 = T(0.0), o_130 = T(0.0), o_231 = T(0.0), o_232 = T(0.0);
  T z_132 = T(0.0), z_232 = T(0.0);
  for (int kk = k0 - 1; kk <= k1 + 1; kk++) {
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) V[m][a][b][0] = V[m][a][b][1];
    const bool vp = vt && kk >= -2 && kk <= nk + 3;
    if (vp) load_plane(kk);
    T nr[4][3];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int ip = kStag[s][0], jp = kStag[s][1], kp = kStag[s][2];
      nr[s][0] = nr[s][1] = nr[s][2] = T(0.0);
      if (!vp || ii < -2 + ip || jj < -2 + jp || kk < -2 + kp) continue;
      const int ai = 1 - ip, bj = 1 - jp, cp = 1 - kp;
      T r[3];
#pragma unroll
      for (int m = 0; m < 3; m++)
        r[m] = (+V[m][ai][bj][cp] + V[m][1][bj][cp] + V[m][ai][1][cp] + V[m][1][1][cp] +
                V[m][ai][bj][1] + V[m][1][bj][1] + V[m][ai][1][1] + V[m][1][1][1]) *
               T(0.125);
      flat_zero(r, flat2);
      const T arn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + verysmall;
      const T inv = T(1.0) / arn;
      nr[s][0] = r[0] * inv;
      nr[s][1] = r[1] * inv;
      nr[s][2] = r[2] * inv;
    }
    if (own && kk >= k0 && kk <= k1) {
      const int tgt = d.ix(ii, jj, kk);
#pragma unroll
      for (int s = 0; s < 4; s++) {
        out[s][tgt] = nr[s][0];
        out[s][n3 + tgt] = nr[s][1];
        out[s][2 * n3 + tgt] = nr[s][2];
      }
    }
    const int w = kk & 1;
    xs[w][0][ty][tx] = nr[0][0];
    xs[w][1][ty][tx] = nr[1][0];
    xs[w][2][ty][tx] = nr[2][0];
    xs[w][3][ty][tx] = nr[0][1];
    xs[w][4][ty][tx] = nr[1][1];
    xs[w][5][ty][tx] = nr[3][1];
    __syncthreads();
    // curvature of plane kc = kk - 1 (srcs/calrkap.f:42-98)
    const int kc = kk - 1;
    if (own && kc >= k0 && kc <= k1) {
      const int q = kc & 1;
      T a = T(0.0), b = T(0.0), e = T(0.0);
      if (kc >= -1 && jj >= -1 && ii <= ni + 2)
        a = +(-o_pp0 + xs[q][0][ty][tx + 1]) * dxinv + (-xs[q][4][ty - 1][tx] + o_121) * dyinv +
            (-z_132 + o_132) * dzinv;
      if (kc >= -1 && jj <= nj + 2 && ii >= -1)
        b = +(-xs[q][1][ty][tx - 1] + o_120) * dxinv + (-o_pp1 + xs[q][3][ty + 1][tx]) * dyinv +
            (-z_232 + o_232) * dzinv;
      if (kc <= nk + 2 && jj >= -1 && ii >= -1)
        e = +(-xs[q][2][ty][tx - 1] + o_130) * dxinv + (-xs[q][5][ty - 1][tx] + o_231) * dyinv +
            (-o_pp2 + nr[0][2]) * dzinv;
      const int c = d.ix(ii, jj, kc);
      rk[c] = a;
      rk[n3 + c] = b;
      rk[2 * n3 + c] = e;
    }
    z_132 = o_132;
    z_232 = o_232;
    o_pp0 = nr[0][0]; o_pp1 = nr[0][1]; o_pp2 = nr[0][2];
    o_120 = nr[1][0]; o_121 = nr[1][1];
    o_130 = nr[2][0]; o_132 = nr[2][2];
    o_231 = nr[3][1]; o_232 = nr[3][2];
    __syncthreads();  // the buffer of plane kk-1 is rewritten for plane kk+1
  }
}

template <class T>
void calrn4_kap(const Dims d, double verysmall, double dxinv, double dyinv, double dzinv,
                const T* px, const T* py, const T* pz, T* const* rn, T* rk) {
  const int gx = (d.sx + kKapTx - 1) / kKapTx, gy = (d.sy + kKapTy - 1) / kKapTy;
  const int gz = (d.sz + kRnChunk - 1) / kRnChunk;
  calrn4_kap_kernel<T><<<dim3(gx * gy * gz), dim3(64, kKapTy + 2, 1), 0, stream()>>>(
      d, T(verysmall), T(dxinv), T(dyinv), T(dzinv), px, py, pz, rn[0], rn[1], rn[2], rn[3], rk,
      gx, gy);
  FEL_CHECK(hipGetLastError());
}

// Normals of the other three staggerings, computed by the pass that
// answered the first request and parked until they are asked for: a request
// whose inputs carry the same content versions gets the parked buffer
// swapped in (the same bits, no kernel).
namespace {
struct RnStash {
  bool valid[4] = {false, false, false, false};
  uint64_t ver[3] = {0, 0, 0};
  int ni = 0, nj = 0, nk = 0;
  double vs = 0.0;
  void* buf[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t bytes = 0;
  // the curvature of the four normals, formed by the same pass
  // (calrn4_kap): valid for a calrkap_ request with these scales whose four
  // inputs still hold the normals as delivered (host array, content number)
  bool kap_valid = false;
  double kdx = 0.0, kdy = 0.0, kdz = 0.0;
  void* kap = nullptr;
  const void* dhost[4] = {nullptr, nullptr, nullptr, nullptr};
  uint64_t dver[4] = {0, 0, 0, 0};
};
RnStash g_rn;

int stagger_index(int ip, int jp, int kp) {
  for (int s = 0; s < 4; s++)
    if (kStag[s][0] == ip && kStag[s][1] == jp && kStag[s][2] == kp) return s;
  return -1;
}
}  // namespace

void rn_release() {
  for (int s = 0; s < 4; s++)
    if (g_rn.buf[s]) FEL_CHECK(hipFree(g_rn.buf[s]));
  if (g_rn.kap) FEL_CHECK(hipFree(g_rn.kap));
  g_rn = RnStash{};
}

// srcs/calrkap.f:20-97
template <class T>
void calrkap(const Dims d, double dxinv_, double dyinv_, double dzinv_, const T* rpp, const T* r12,
             const T* r13, const T* r23, T* rk) {
  const int n3 = d.n3, ni = d.ni, nj = d.nj, nk = d.nk;
  const T dxinv = T(dxinv_), dyinv = T(dyinv_), dzinv = T(dzinv_);
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    const int C2 = n3, C3 = 2 * n3;
    T a = T(0.0), b = T(0.0), e = T(0.0);
    if (k >= -1 && j >= -1 && i <= ni + 2)
      a = +(-rpp[c] + rpp[c + X]) * dxinv + (-r12[C2 + c - Y] + r12[C2 + c]) * dyinv +
          (-r13[C3 + c - Z] + r13[C3 + c]) * dzinv;
    if (k >= -1 && j <= nj + 2 && i >= -1)
      b = +(-r12[c - X] + r12[c]) * dxinv + (-rpp[C2 + c] + rpp[C2 + c + Y]) * dyinv +
          (-r23[C3 + c - Z] + r23[C3 + c]) * dzinv;
    if (k <= nk + 2 && j >= -1 && i >= -1)
      e = +(-r13[c - X] + r13[c]) * dxinv + (-r23[C2 + c - Y] + r23[C2 + c]) * dyinv +
          (-rpp[C3 + c] + rpp[C3 + c + Z]) * dzinv;
    rk[c] = a;
    rk[C2 + c] = b;
    rk[C3 + c] = e;
  });
}

// srcs/calfst.f:37-45
template <class T>
void calfst(const Dims d, double st_, const T* px, const T* py, const T* pz, const T* rk, T* fu,
            T* fv, T* fw) {
  const int n3 = d.n3;
  const T st = T(st_);
  launch_flat(n3, [=] __device__(int m) {
    fu[m] = -(st * rk[m] * px[m]);
    fv[m] = -(st * rk[n3 + m] * py[m]);
    fw[m] = -(st * rk[2 * n3 + m] * pz[m]);
  });
}

template <class T>
void cal_grad_p2uvw_t(const Dims d, int ID, int svall, double dxinv, const double* dyinv_array,
                      double dzinv, const double* q, double* qx_u, double* qy_v, double* qz_w) {
  const size_t n = (size_t)d.n3;
  const void* outs[3] = {qx_u, qy_v, qz_w};
  uint64_t sh[3];
  const MemoKey k = gradient_key(kOpGradP2uvw, (int)sizeof(T), d, ID, svall, dxinv, dyinv_array,
                                 dzinv, q, outs, sh);
  if (memo_reuse(k, {{qx_u, n}, {qy_v, n}, {qz_w, n}}, false)) {
    for (int m = 0; m < 3; m++) keep_shell(outs[m], n, sh[m]);
    return;
  }
  memo_park(d, {{qx_u, n}, {qy_v, n}, {qz_w, n}}, true);
  Call c("cal_grad_p2uvw_");
  bool uni = true;
  const int cnt = svall + 6;
  for (int m = 0; m < cnt; m++)
    if (dyinv_array[m] != dyinv_array[0]) uni = false;
  const double* table = nullptr;
  if (!uni) {
    double* dev = static_cast<double*>(scratch(2, sizeof(double) * cnt));
    FEL_CHECK(hipMemcpyAsync(dev, dyinv_array, sizeof(double) * cnt, hipMemcpyHostToDevice, stream()));
    FEL_CHECK(hipStreamSynchronize(stream()));
    table = dev;
  }
  const PT<T>* dq = c.in_t<PT<T>>(q, d.n3, "q");
  T* x = c.io_t<T>(qx_u, d.n3, "qx_u");
  T* y = c.io_t<T>(qy_v, d.n3, "qy_v");
  T* z = c.io_t<T>(qz_w, d.n3, "qz_w");
  grad_p2uvw<T>(d, ID, dxinv, table, dyinv_array[0], dzinv, dq, x, y, z);
  c.finish();
  memo_record(k, {{qx_u, n}, {qy_v, n}, {qz_w, n}});
  for (int m = 0; m < 3; m++) keep_shell(outs[m], n, sh[m]);  // written: the box only
}

template <class T>
void calrn_t(const Dims d, double vs, double dxinv, double dyinv, double dzinv, int ip, int jp,
             int kp, const double* phix, const double* phiy, const double* phiz, double* rn) {
  const size_t n = (size_t)d.n3, n3x = 3 * n;
  MemoKey k(kOpCalrn);
  k.num((int64_t)sizeof(T)).dims(d).num(vs).num((int64_t)ip).num((int64_t)jp).num((int64_t)kp);
  k.content(phix, n).content(phiy, n).content(phiz, n);
  if (memo_reuse(k, {{rn, n3x}}, true)) return;
  const int sreq = stagger_index(ip, jp, kp);
  const uint64_t ver[3] = {content_version(phix, n), content_version(phiy, n),
                           content_version(phiz, n)};
  const bool track = sreq >= 0 && ver[0] && ver[1] && ver[2] && content_version(rn, n3x) != 0 &&
                     result_reuse();
  if (track) {
    const bool same = g_rn.ver[0] == ver[0] && g_rn.ver[1] == ver[1] && g_rn.ver[2] == ver[2] &&
                      g_rn.ni == d.ni && g_rn.nj == d.nj && g_rn.nk == d.nk && g_rn.vs == vs;
    if (same && g_rn.valid[sreq]) {
      g_rn.buf[sreq] = exchange_device_buffer(rn, n3x, g_rn.buf[sreq]);
      g_rn.valid[sreq] = false;
      Call c("calrn_");  // the output changes: a fresh version
      c.all_t<T>(rn, n3x, "rn");
      c.finish();
      memo_record(k, {{rn, n3x}});
      g_rn.dhost[sreq] = rn;
      g_rn.dver[sreq] = content_version(rn, n3x);
      count_stash_hit();
      return;
    }
    const size_t bytes = n3x * sizeof(T);
    if (g_rn.bytes != bytes) {
      for (int s = 0; s < 4; s++) {
        if (g_rn.buf[s]) FEL_CHECK(hipFree(g_rn.buf[s]));
        FEL_CHECK(hipMalloc(&g_rn.buf[s], bytes));
      }
      if (g_rn.kap) FEL_CHECK(hipFree(g_rn.kap));
      g_rn.kap = nullptr;
      g_rn.bytes = bytes;
    }
    Call c("calrn_");
    const T* x = c.in_t<T>(phix, n, "phix");
    const T* y = c.in_t<T>(phiy, n, "phiy");
    const T* z = c.in_t<T>(phiz, n, "phiz");
    T* out[4];
    for (int s = 0; s < 4; s++)
      out[s] = s == sreq ? c.all_t<T>(rn, n3x, "rn") : static_cast<T*>(g_rn.buf[s]);
    g_rn.kap_valid = false;
    if constexpr (!is_dual<T>) {
      // the curvature calrkap_ will ask of these normals, in the same pass
      if (!g_rn.kap) FEL_CHECK(hipMalloc(&g_rn.kap, bytes));
      calrn4_kap<T>(d, vs, dxinv, dyinv, dzinv, x, y, z, out, static_cast<T*>(g_rn.kap));
      g_rn.kap_valid = true;
      g_rn.kdx = dxinv; g_rn.kdy = dyinv; g_rn.kdz = dzinv;
    } else {
      calrn4<T>(d, vs, x, y, z, out);
    }
    c.finish();
    for (int s = 0; s < 4; s++) g_rn.valid[s] = s != sreq;
    for (int m = 0; m < 3; m++) g_rn.ver[m] = ver[m];
    g_rn.ni = d.ni; g_rn.nj = d.nj; g_rn.nk = d.nk;
    g_rn.vs = vs;
    memo_record(k, {{rn, n3x}});
    for (int s = 0; s < 4; s++) {
      g_rn.dhost[s] = nullptr;
      g_rn.dver[s] = 0;
    }
    g_rn.dhost[sreq] = rn;
    g_rn.dver[sreq] = content_version(rn, n3x);
    return;
  }
  Call c("calrn_");
  const T* x = c.in_t<T>(phix, d.n3, "phix");
  const T* y = c.in_t<T>(phiy, d.n3, "phiy");
  const T* z = c.in_t<T>(phiz, d.n3, "phiz");
  T* r = c.all_t<T>(rn, 3 * (size_t)d.n3, "rn");
  calrn<T>(d, vs, ip, jp, kp, x, y, z, r);
  c.finish();
  memo_record(k, {{rn, n3x}});
}

template <class T>
void calrkap_t(const Dims d, double dxinv, double dyinv, double dzinv, const double* rn_pp,
               const double* rn_12, const double* rn_13, const double* rn_23, double* rkap) {
  const size_t n = 3 * (size_t)d.n3;
  MemoKey k(kOpCalrkap);
  k.num((int64_t)sizeof(T)).dims(d).num(dxinv).num(dyinv).num(dzinv);
  k.content(rn_pp, n).content(rn_12, n).content(rn_13, n).content(rn_23, n);
  if (memo_reuse(k, {{rkap, n}}, true)) return;
  {
    // the curvature formed with the normals by calrn4_kap, while the four
    // inputs hold those normals
    const void* in4[4] = {rn_pp, rn_12, rn_13, rn_23};
    bool ok = g_rn.kap_valid && result_reuse() && g_rn.ni == d.ni && g_rn.nj == d.nj &&
              g_rn.nk == d.nk && g_rn.kdx == dxinv && g_rn.kdy == dyinv && g_rn.kdz == dzinv &&
              g_rn.bytes == n * sizeof(T) && content_version(rkap, n) != 0;
    for (int s = 0; s < 4 && ok; s++)
      ok = g_rn.dhost[s] == in4[s] && g_rn.dver[s] != 0 && content_version(in4[s], n) == g_rn.dver[s];
    if (ok) {
      g_rn.kap = exchange_device_buffer(rkap, n, g_rn.kap);
      g_rn.kap_valid = false;
      Call c("calrkap_");  // the output changes: a fresh version
      c.all_t<T>(rkap, n, "rkap");
      c.finish();
      memo_record(k, {{rkap, n}});
      count_stash_hit();
      return;
    }
  }
  Call c("calrkap_");
  const T *a = c.in_t<T>(rn_pp, n, "rn_pp"), *b = c.in_t<T>(rn_12, n, "rn_12");
  const T *e = c.in_t<T>(rn_13, n, "rn_13"), *f = c.in_t<T>(rn_23, n, "rn_23");
  T* r = c.all_t<T>(rkap, n, "rkap");
  calrkap<T>(d, dxinv, dyinv, dzinv, a, b, e, f, r);
  c.finish();
  memo_record(k, {{rkap, n}});
}

template <class T>
void calfst_t(const Dims d, double st, const double* phix_u, const double* phiy_v,
              const double* phiz_w, const double* rkap, double* fst_u, double* fst_v,
              double* fst_w) {
  const size_t n = (size_t)d.n3;
  MemoKey k(kOpCalfst);
  k.num((int64_t)sizeof(T)).dims(d).num(st);
  k.content(phix_u, n).content(phiy_v, n).content(phiz_w, n).content(rkap, 3 * n);
  if (memo_reuse(k, {{fst_u, n}, {fst_v, n}, {fst_w, n}}, true)) return;
  Call c("calfst_");
  const T *x = c.in_t<T>(phix_u, d.n3, "phix_u"), *y = c.in_t<T>(phiy_v, d.n3, "phiy_v"),
          *z = c.in_t<T>(phiz_w, d.n3, "phiz_w");
  const T* rk = c.in_t<T>(rkap, 3 * (size_t)d.n3, "rkap");
  T *fu = c.all_t<T>(fst_u, d.n3, "fst_u"), *fv = c.all_t<T>(fst_v, d.n3, "fst_v"),
    *fw = c.all_t<T>(fst_w, d.n3, "fst_w");
  calfst<T>(d, st, x, y, z, rk, fu, fv, fw);
  c.finish();
  memo_record(k, {{fst_u, n}, {fst_v, n}, {fst_w, n}});
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/cal_grad_p2uvw.f:1-3
void cal_grad_p2uvw_(const int* ID, const int* svall, const int* ni, const int* nj, const int* nk,
                     const double* dxinv, const double* dyinv_array, const double* dzinv,
                     const double* q, double* qx_u, double* qy_v, double* qz_w) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) {
    cal_grad_p2uvw_t<decltype(t)>(d, *ID, *svall, *dxinv, dyinv_array, *dzinv, q, qx_u, qy_v, qz_w);
  });
}

// srcs/calrn.f:1-2
void calrn_(const int* ni, const int* nj, const int* nk, const double* dxinv, const double* dyinv,
            const double* dzinv, const int* ip, const int* jp, const int* kp, const double* phix,
            const double* phiy, const double* phiz, double* rn) {
  const Dims d = make_dims(*ni, *nj, *nk);
  const double vs = std::sqrt(*dxinv * *dxinv + *dyinv * *dyinv + *dzinv * *dzinv) * 1.0e-30;
  with_flow([&](auto t) {
    calrn_t<decltype(t)>(d, vs, *dxinv, *dyinv, *dzinv, *ip, *jp, *kp, phix, phiy, phiz, rn);
  });
}

// srcs/calrkap.f:1-2
void calrkap_(const int* ni, const int* nj, const int* nk, const double* dxinv, const double* dyinv,
              const double* dzinv, const double* rn_pp, const double* rn_12, const double* rn_13,
              const double* rn_23, double* rkap) {
  const Dims d = make_dims(*ni, *nj, *nk);
  with_flow([&](auto t) {
    calrkap_t<decltype(t)>(d, *dxinv, *dyinv, *dzinv, rn_pp, rn_12, rn_13, rn_23, rkap);
  });
}

// srcs/calfst.f:1-3
void calfst_(const int* ni, const int* nj, const int* nk, const double* rhol, const double* rhog,
             const double* surface_tension, const double* phix_u, const double* phiy_v,
             const double* phiz_w, const double* rkap, double* fst_u, double* fst_v, double* fst_w) {
  const Dims d = make_dims(*ni, *nj, *nk);
  const double st = 2.0 * (*surface_tension) / (*rhol + *rhog);
  with_flow([&](auto t) {
    calfst_t<decltype(t)>(d, st, phix_u, phiy_v, phiz_w, rkap, fst_u, fst_v, fst_w);
  });
}

}  // extern "C"
