// k_poisson.hip -- pressure projection: divergence, 2-D FFT in the periodic
// (i,k) planes, Thomas solves along the wall-normal j, inverse FFT, mean
// removal, velocity correction.
//
// Reference: srcs/caldiv.f, srcs/solp_fft_tdma1.f, srcs/trans_r2w.f,
// srcs/solp_fft_tdma2.f, srcs/trans_w2r.f, srcs/solp_fft_tdma3.f,
// srcs/solp_fft_tdma4.f, srcs/bnd_avzero.f, srcs/cal_av.f,
// srcs/corunp_explicit.f, srcs/cal_coef_prs.f, srcs/cal_coef_solp_fft.f.
//
// The transforms are this library's own batched mixed-radix (radix-4 with a
// final radix-2 stage) Stockham FFTs run in LDS (the reference's
// srcs/fourn.f is not used): the x lines are
// contiguous and one workgroup transforms whole lines; the k lines are
// strided, so one workgroup transforms a tile of lines that are adjacent in
// memory, making every global access a contiguous row of the tile.  They
// compute the same DFT as fourn (exp(-2 pi i ...) for the forward transform,
// unnormalised inverse), to rounding.
#include <cmath>
#include <map>
#include <utility>
#include <vector>

#include "felgpu.hpp"
#include "felgpu_visc.hpp"

namespace fel {

// Complex element of the transforms: double2 in FP64, float2 in the FP32
// mode, a pair of Duals in the tangent mode (the transforms are linear and
// the twiddles constant, so the derivative part is the transform of the
// input's derivative part).
struct Dual2 {
  Dual x, y;
};
template <class T>
struct Cplx;
template <>
struct Cplx<Dual> {
  using type = Dual2;
  __host__ __device__ static Dual2 make(Dual a, Dual b) { return Dual2{a, b}; }
};
template <>
struct Cplx<double> {
  using type = double2;
  __host__ __device__ static double2 make(double a, double b) { return make_double2(a, b); }
};
template <>
struct Cplx<float> {
  using type = float2;
  __host__ __device__ static float2 make(float a, float b) { return make_float2(a, b); }
};

// Twiddles exp(+2 pi i t / N), t in [0, N), from the host libm (FP64,
// rounded once to T).
template <class T>
static const typename Cplx<T>::type* twiddles(int n) {
  using C = typename Cplx<T>::type;
  static std::map<int, C*> cache;
  auto it = cache.find(n);
  if (it != cache.end()) return it->second;
  std::vector<C> h(n > 0 ? n : 1);
  const double pi = std::acos(-1.0);
  for (int t = 0; t < n; t++) {
    const double a = 2.0 * pi * (double)t / (double)n;
    h[t] = Cplx<T>::make(T(std::cos(a)), T(std::sin(a)));
  }
  C* d = nullptr;
  FEL_CHECK(hipMalloc((void**)&d, sizeof(C) * h.size()));
  FEL_CHECK(hipMemcpy(d, h.data(), sizeof(C) * h.size(), hipMemcpyHostToDevice));
  cache[n] = d;
  return d;
}

// (a) * conj-or-not twiddle w: w = (wr, sign * wi) of exp(+2 pi i t / N).
template <class T, class C>
__device__ __forceinline__ C twiddle_mul(C a, C w0, T sign) {
  const T wr = w0.x, wi = sign * w0.y;
  return Cplx<T>::make(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
}

// Mixed-radix Stockham autosort DIF transform (radix-4 stages, a final
// radix-2 stage when log2 N is odd) of G-thread groups, one line per group,
// ping-ponging between x and y in LDS.  At stride s with radix r and
// m = N/(r s), butterfly b = (p, q) = (b / s, b % s) reads
// x[q + s (p + j m)], j < r, and writes y[q + s (r p + k)] = (sum_j a_j
// e^{sign 2 pi i j k / r}) e^{sign 2 pi i k p s / N}.  Half the passes (and
// barriers) of radix 2.  sign = -1 forward.  Returns the buffer that holds
// the result.
template <class T, class C>
__device__ __forceinline__ C* stockham(C* x, C* y, int n, int lane, int g, const C* __restrict__ tw,
                                       T sign) {
  int s = 1;
  for (; n / s >= 4; s *= 4) {
    const int m = n / (4 * s);
    for (int b = lane; b < n / 4; b += g) {
      const int p = b / s, q = b - p * s;
      const C a0 = x[q + s * p], a1 = x[q + s * (p + m)], a2 = x[q + s * (p + 2 * m)],
              a3 = x[q + s * (p + 3 * m)];
      const C t0 = Cplx<T>::make(a0.x + a2.x, a0.y + a2.y);
      const C t1 = Cplx<T>::make(a0.x - a2.x, a0.y - a2.y);
      const C t2 = Cplx<T>::make(a1.x + a3.x, a1.y + a3.y);
      const T dx = a1.x - a3.x, dy = a1.y - a3.y;
      const C t3 = Cplx<T>::make(-(sign * dy), sign * dx);  // (a1 - a3) * (sign i)
      const int o = q + s * (4 * p), ps = p * s;
      y[o] = Cplx<T>::make(t0.x + t2.x, t0.y + t2.y);
      y[o + s] = twiddle_mul<T, C>(Cplx<T>::make(t1.x + t3.x, t1.y + t3.y), tw[ps], sign);
      y[o + 2 * s] = twiddle_mul<T, C>(Cplx<T>::make(t0.x - t2.x, t0.y - t2.y), tw[2 * ps], sign);
      y[o + 3 * s] = twiddle_mul<T, C>(Cplx<T>::make(t1.x - t3.x, t1.y - t3.y), tw[3 * ps], sign);
    }
    __syncthreads();
    C* t = x;
    x = y;
    y = t;
  }
  if (s < n) {  // n / s == 2: the last radix-2 stage (m = 1, p = 0: no twiddles)
    for (int q = lane; q < s; q += g) {
      const C a = x[q], c = x[q + s];
      y[q] = Cplx<T>::make(a.x + c.x, a.y + c.y);
      y[q + s] = Cplx<T>::make(a.x - c.x, a.y - c.y);
    }
    __syncthreads();
    C* t = x;
    x = y;
    y = t;
  }
  return x;
}

// x lines: N = line length (contiguous), LPB lines per 256-thread block.
template <class T, class Load, class Store>
__global__ void __launch_bounds__(256) fft_x_kernel(int n, int nlines, T sign,
                                                    const typename Cplx<T>::type* __restrict__ tw,
                                                    Load ld, Store st) {
  using C = typename Cplx<T>::type;
  extern __shared__ unsigned char smem_x[];
  C* sm = reinterpret_cast<C*>(smem_x);
  const int lpb = n >= 256 ? 1 : 256 / (n / 2 > 0 ? n / 2 : 1);  // lines per block
  const int g = 256 / lpb;                                        // threads per line
  const int li = (int)threadIdx.x / g, lane = (int)threadIdx.x - li * g;
  const int line = (int)blockIdx.x * lpb + li;
  C* x = sm + li * 2 * n;
  C* y = x + n;
  const bool act = line < nlines;
  for (int e = lane; e < n; e += g) x[e] = act ? ld(line, e) : Cplx<T>::make(T(0.0), T(0.0));
  __syncthreads();
  C* r = stockham<T, C>(x, y, n, lane, g, tw, sign);
  if (act)
    for (int e = lane; e < n; e += g) st(line, e, r[e]);
}

// Strided lines: a block owns TL lines that are adjacent in memory (tile
// index t = 0..TL-1) and transforms them together; element e of line l is
// loaded/stored by the functors.  512 threads: a tile's two LDS buffers fill
// 64 KiB, so two blocks share a CU, and 512-thread blocks keep 16 waves per
// CU in flight for the strided loads (measured: 256-thread blocks reached
// 2.0-2.6 TB/s).
constexpr int kZThreads = 512;

template <class T, class Load, class Store>
__global__ void __launch_bounds__(kZThreads) fft_z_kernel(int n, int tiles, int TL, T sign,
                                                    const typename Cplx<T>::type* __restrict__ tw,
                                                    Load ld, Store st) {
  using C = typename Cplx<T>::type;
  extern __shared__ unsigned char smem_z[];
  C* sm = reinterpret_cast<C*>(smem_z);
  const int g = kZThreads / TL;           // threads per line
  const int tile = (int)blockIdx.x;
  // cooperative, row-major loads: consecutive threads take consecutive lines
  for (int idx = (int)threadIdx.x; idx < TL * n; idx += kZThreads) {
    const int e = idx / TL, l = idx - e * TL;
    sm[l * 2 * n + e] = ld(tile, l, e);
  }
  __syncthreads();
  const int l = (int)threadIdx.x / g, lane = (int)threadIdx.x - l * g;
  C* x = sm + l * 2 * n;
  C* r = stockham<T, C>(x, x + n, n, lane, g, tw, sign);
  const int off = (int)(r - x);  // 0 or n, the same for every line
  for (int idx = (int)threadIdx.x; idx < TL * n; idx += kZThreads) {
    const int e = idx / TL, ll = idx - e * TL;
    st(tile, ll, e, sm[ll * 2 * n + off + e]);
  }
  (void)tiles;
}

template <class T, class Load, class Store>
static void fft_x(int n, int nlines, T sign, Load ld, Store st) {
  using C = typename Cplx<T>::type;
  const int lpb = n >= 256 ? 1 : 256 / (n / 2 > 0 ? n / 2 : 1);
  const size_t shm = sizeof(C) * 2 * n * lpb;
  fft_x_kernel<T, Load, Store><<<(nlines + lpb - 1) / lpb, 256, shm, stream()>>>(
      n, nlines, sign, twiddles<T>(n), ld, st);
  FEL_CHECK(hipGetLastError());
}

// Tile width: TL lines of n points, two LDS buffers each, <= 64 KiB.
template <class T>
static int z_tile(int n, int nlines_fast) {
  using C = typename Cplx<T>::type;
  int TL = 16;
  while (TL > 1 && (size_t)TL * 2 * n * sizeof(C) > 65536) TL /= 2;
  while (TL > 1 && nlines_fast % TL != 0) TL /= 2;
  return TL;
}

template <class T, class Load, class Store>
static void fft_z(int n, int tiles, int TL, T sign, Load ld, Store st) {
  using C = typename Cplx<T>::type;
  const size_t shm = sizeof(C) * 2 * n * TL;
  fft_z_kernel<T, Load, Store><<<tiles, kZThreads, shm, stream()>>>(n, tiles, TL, sign,
                                                                    twiddles<T>(n), ld, st);
  FEL_CHECK(hipGetLastError());
}

static void check_pow2(int n, const char* who) {
  if (n < 2 || (n & (n - 1)) != 0) {
    std::fprintf(stderr, "felgpu: %s: FFT extent %d is not a power of two (srcs/main.f90:117-126)\n", who, n);
    std::abort();
  }
}

// srcs/solp_fft_tdma1.f:25-47: phir(i,j,k) = DFT(div(1:ni,j,1:nk))/(ni nk)
// at centred mode indices; the scratch holds (nk, nj, ni) complex values.
// T: element type of the modes; the divergence is in the phase type PT<T>.
// `rhs(i,j,k)` gives the right-hand side at interior cell (i,j,k).
// ACC: the transform divided by `nrm` is added to the mode arrays instead of
// stored (the adjoint of solp_fft_tdma3_, unnormalised, nrm = 1).
template <class T, bool ACC, class Rhs>
void fft_forward_core(const Dims d, Rhs rhs, T* pr, T* pi, T nrm) {
  using C = typename Cplx<T>::type;
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  check_pow2(ni, "solp_fft_tdma1_");
  check_pow2(nk, "solp_fft_tdma1_");
  C* TS = static_cast<C*>(scratch(4, sizeof(C) * (size_t)ni * nj * nk));
  // x transforms of the real rows (line = j + nj*k)
  fft_x<T>(ni, nj * nk, T(-1.0),
           [=] __device__(int line, int e) {
             const int j = line % nj, k = line / nj;
             return Cplx<T>::make(T(rhs(e + 1, j + 1, k + 1)), T(0.0));
           },
           [=] __device__(int line, int e, C v) { TS[(size_t)line * ni + e] = v; });
  // k transforms; tile = (j, kx block), lines adjacent in kx
  const int Tw = z_tile<T>(nk, ni);
  const int tiles_per_j = ni / Tw;
  fft_z<T>(nk, nj * tiles_per_j, Tw, T(-1.0),
           [=] __device__(int tile, int l, int e) {
             const int j = tile % nj, kx = (tile / nj) * Tw + l;
             return TS[((size_t)e * nj + j) * ni + kx];
           },
           [=] __device__(int tile, int l, int e, C v) {
             const int j = tile % nj, kx = (tile / nj) * Tw + l;
             const int ic = (kx + ni / 2) % ni, kc = (e + nk / 2) % nk;
             const size_t o = (size_t)ic + (size_t)ni * ((size_t)j + (size_t)nj * kc);
             if constexpr (ACC) {
               pr[o] = pr[o] + v.x / nrm;
               pi[o] = pi[o] + v.y / nrm;
             } else {
               pr[o] = v.x / nrm;
               pi[o] = v.y / nrm;
             }
           });
}

template <class T, class Rhs>
void fft_forward_of(const Dims d, Rhs rhs, T* pr, T* pi) {
  fft_forward_core<T, false>(d, rhs, pr, pi, T((double)(d.ni * d.nk)));
}

template <class T>
void fft_forward(const Dims d, const PT<T>* div, T* pr, T* pi) {
  fft_forward_of<T>(d, [=] __device__(int i, int j, int k) { return div[d.ix(i, j, k)]; }, pr, pi);
}

// srcs/solp_fft_tdma3.f:25-44: dp(1:ni,j,1:nk) = Re(unnormalised inverse DFT)
// (added to dp instead when `add`: the refinement corrections).
template <class T, class P = PT<T>>
void fft_inverse(const Dims d, const T* pr, const T* pi, P* dp, bool add) {
  using C = typename Cplx<T>::type;
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  check_pow2(ni, "solp_fft_tdma3_");
  check_pow2(nk, "solp_fft_tdma3_");
  C* TS = static_cast<C*>(scratch(4, sizeof(C) * (size_t)ni * nj * nk));
  const int Tw = z_tile<T>(nk, ni);
  const int tiles_per_j = ni / Tw;
  fft_z<T>(nk, nj * tiles_per_j, Tw, T(1.0),
           [=] __device__(int tile, int l, int e) {
             const int j = tile % nj, kx = (tile / nj) * Tw + l;
             const int ic = (kx + ni / 2) % ni, kc = (e + nk / 2) % nk;
             const size_t o = (size_t)ic + (size_t)ni * ((size_t)j + (size_t)nj * kc);
             return Cplx<T>::make(pr[o], pi[o]);
           },
           [=] __device__(int tile, int l, int e, C v) {
             const int j = tile % nj, kx = (tile / nj) * Tw + l;
             TS[((size_t)e * nj + j) * ni + kx] = v;
           });
  fft_x<T>(ni, nj * nk, T(1.0),
           [=] __device__(int line, int e) { return TS[(size_t)line * ni + e]; },
           [=] __device__(int line, int e, C v) {
             const int j = line % nj, k = line / nj;
             P* o = dp + d.ix(e + 1, j + 1, k + 1);
             *o = add ? *o + P(v.x) : P(v.x);
           });
}

// srcs/solp_fft_tdma2.f:46-87: Thomas solve along j for every (kx,kz) mode
// of the centred layout (i fastest, j, k).  c1x/c1z tables come from the
// host libm; a(j) is kept in device scratch for the back substitution.
// In FP32 the last pivot of the singular (0,0) mode, zero in exact
// arithmetic, is set to pin that mode's x(nj) to 0 exactly, which the FP64
// expression -d/(d^2+1e-99) does when the pivot rounds to zero.
//
// TRANS: the solve of the transposed systems (sub-diagonal an(j-1),
// super-diagonal as(j+1)), the adjoint of the forward solve; its singular
// mode is pinned in FP64 too (x(nj) = 0 and the leading nj-1 rows: the
// transpose of the pinned solve).
template <class T, bool TRANS = false>
void tdma_modes(int ni, int nj, int nk, const double* cx, const double* cz, const double* coef,
                double dx2r, double dz2r, T* pr, T* pi) {
  T* at = static_cast<T*>(scratch(5, sizeof(T) * (size_t)ni * nj * nk));
  const double *as = coef, *an = coef + (nj + 6), *ap0 = coef + 2 * (nj + 6);
  launch_flat(ni * nk, [=] __device__(int m) {
    const int i = m % ni, k = m / ni;
    const double c1x = cx[i], c1z = cz[k];
    const bool singular = (is_f32<T> || TRANS) && c1x == 2.0 && c1z == 2.0;
    T a_prev = T(0.0), br = T(0.0), bi = T(0.0);
    const size_t base = (size_t)i + (size_t)ni * nj * k;
    for (int j = 1; j <= nj; j++) {
      const size_t o = base + (size_t)ni * (j - 1);
      const T ap = T(ap0[j + 2] + dx2r * (2.0 - c1x) + dz2r * (2.0 - c1z));
      const T asj = T(TRANS ? (j > 1 ? an[j + 1] : 0.0) : as[j + 2]);
      const T d00 = -ap + asj * a_prev;
      T apinv = -d00 / (d00 * d00 + guard<T>(1.0e-99));
      if (singular && j == nj) apinv = T(0.0);
      a_prev = apinv * T(TRANS ? (j < nj ? as[j + 3] : 0.0) : an[j + 2]);
      br = apinv * (-pr[o] + asj * br);
      bi = apinv * (-pi[o] + asj * bi);
      at[o] = a_prev;
      pr[o] = br;
      pi[o] = bi;
    }
    T xr = br, xi = bi;
    for (int j = nj - 1; j >= 1; j--) {
      const size_t o = base + (size_t)ni * (j - 1);
      xr = pr[o] + at[o] * xr;
      xi = pi[o] + at[o] * xi;
      pr[o] = xr;
      pi[o] = xi;
    }
  });
}

// srcs/caldiv.f:24-41 on [-1,n+3]^3; the RMS (caldiv.f:43-67) is returned
// through dsum (device, sum of squares of the values over the interior).
// The divergence is in the phase type P (velocity values of either
// precision are exact in FP64).
template <class T, class P = PT<T>>
void divergence(const Dims d, double dxinv, double dyinv, double dzinv, double dtinv, const T* u,
                const T* v, const T* w, P* div, double* dsum) {
  launch_box(Box{-1, d.ni + 3, -1, d.nj + 3, -1, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    div[c] = ((-P(u[c - X]) + P(u[c])) * dxinv + (-P(v[c - Y]) + P(v[c])) * dyinv +
              (-P(w[c - Z]) + P(w[c])) * dzinv) *
             dtinv;
  });
  interior_reduce<1>(
      d,
      [=] __device__(int i, int j, int k, double* o) {
        const double x = val(div[d.ix(i, j, k)]);
        o[0] = x * x;
      },
      [=] __device__(const double* t, double* out) { out[0] = t[0]; }, dsum);
}

// srcs/bnd_avzero.f:10-23 with srcs/cal_av.f:20-37: the mean stays on the
// device.  Q: element type of the field; the mean is FP64 (Dual in the
// tangent mode, whose mean removal also removes the derivative's mean).
template <class Q>
void avzero(const Dims d, Q* q, int ndiv) {
  using A = PT<Q>;
  double* s = dev_scalars(2);
  const double nn = (double)(d.ni * d.nj * d.nk * ndiv);
  interior_reduce<1, A>(
      d, [=] __device__(int i, int j, int k, A* o) { o[0] = A(q[d.ix(i, j, k)]); },
      [=] __device__(const A* t, double* out) { reinterpret_cast<A*>(out)[0] = t[0] / nn; }, s);
  const A* av = reinterpret_cast<const A*>(s);
  launch_flat(d.n3, [=] __device__(int m) { q[m] = Q(A(q[m]) - *av); });
}

// srcs/corunp_explicit.f:26-74; dp is in the phase type P, the velocities
// and pressure T.
template <class T, class P = PT<T>>
void correct(const Dims d, int jen, double rho, double dxinv, double dyinv, double dzinv, double dt,
             const P* dp, const T* phat, T* un, T* vn, T* wn, T* pn) {
  // the sums are formed in P (FP64 for FP32 storage) and rounded once
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    const P p0 = dp[c];
    un[c] = T(P(un[c]) + (p0 - dp[c + X]) / rho * dxinv * dt);
    if (j <= jen) vn[c] = T(P(vn[c]) + (p0 - dp[c + Y]) / rho * dyinv * dt);
    wn[c] = T(P(wn[c]) + (p0 - dp[c + Z]) / rho * dzinv * dt);
    pn[c] = T(P(phat[c]) + p0);
  });
}

// Sums for srcs/solp_fft_tdma4.f:28-72: [sum div^2, sum (L dp/rho - div)^2],
// of the values.
template <class P>
void poisson_residual_sums(const Dims d, double rho, double dxinv, double dyinv, double dzinv,
                           const P* div_, const P* dp_, double* dsum) {
  const double* div = reinterpret_cast<const double*>(div_);
  const double* dp = reinterpret_cast<const double*>(dp_);
  const int st = (int)(sizeof(P) / sizeof(double));  // value stride: 1, or 2 for Dual
  const double aw = dxinv * dxinv / rho, ae = dxinv * dxinv / rho;
  const double as = dyinv * dyinv / rho, an = dyinv * dyinv / rho;
  const double ab = dzinv * dzinv / rho, at = dzinv * dzinv / rho;
  const double ap = aw + ae + as + an + ab + at;
  interior_reduce_stencil<2>(
      d,
      [=] __device__(int i, int j, int k, double* o) {
        const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
        const double dv = div[st * c];
        const double r = +dp[st * (c - Z)] * ab + dp[st * (c - Y)] * as + dp[st * (c - X)] * aw -
                         dp[st * c] * ap + dp[st * (c + X)] * ae + dp[st * (c + Y)] * an +
                         dp[st * (c + Z)] * at - dv;
        o[0] = dv * dv;
        o[1] = r * r;
      },
      [=] __device__(const double* t, double* out) {
        out[0] = t[0];
        out[1] = t[1];
      },
      dsum);
}

// ---------------------------------------------------------------------------
// FP32 mode: iterative refinement of the pressure correction.
//
// The modes and the Thomas solves run in FP32; the divergence (right-hand
// side) and the correction dp are FP64.  Each refinement step forms the FP64
// residual r = div - L dp of the operator the FFT+TDMA inverts (periodic in
// i and k, the wall rows' zero fluxes folded into as(1) = an(nj) = 0),
// solves L c = r in FP32 and adds c to dp; the residual shrinks by the FP32
// solve's relative accuracy (~1e-6) per step.  solp_fft_tdma1_ records the
// right-hand side, solp_fft_tdma2_ the modal coefficients; solp_fft_tdma3_
// refines until the relative residual |r|/|div| (L2 over the interior) is
// at most FELGPU_REFINE_TOL (default 1e-7: the relative rounding of the FP32
// velocities the correction updates is 6e-8), with at most FELGPU_REFINE
// steps (default 3; 0 turns refinement and the warm start off); the
// felgpu_set_refinement_ entry point overrides both.  Each check is one
// FP64 reduction read back on the host.
// ---------------------------------------------------------------------------
struct PoissonRecord {
  bool rhs = false, coef = false;
  int ni = 0, nj = 0, nk = 0;
  uint64_t div_ver = 0;
  const void* div_host = nullptr;
  double dx2r = 0.0, dz2r = 0.0;
  const double* tab = nullptr;   // device: c1x[ni], c1z[nk], as, an, ap (nj+6 each)
  uint64_t tab_gen = 0;          // counts uploads of a changed table
  // warm start: the correction array of the last solve, and whether (and
  // from which contents of it and which table) solp_fft_tdma1_ transformed
  // the residual div - L dp instead of div
  const void* dp_host = nullptr;
  bool warm = false;
  uint64_t dp0_ver = 0, warm_gen = 0;
  // statistics (felgpu_poisson_stats_): solves, warm starts, refinement
  // steps, and the relative residuals after the first FP32 solve and at the
  // end of the last solve
  uint64_t solves = 0, warm_solves = 0, refine_total = 0;
  double first_rel = 0.0, final_rel = 0.0;
};

static PoissonRecord& prec() {
  static PoissonRecord r;
  return r;
}

// Refinement limits: at most `steps` steps, stopping at relative residual
// `tol`; from FELGPU_REFINE / FELGPU_REFINE_TOL, or felgpu_set_refinement_.
struct RefineLimits {
  int steps;
  double tol;
};

static RefineLimits& refine_limits() {
  static RefineLimits r = [] {
    const char* e = std::getenv("FELGPU_REFINE");
    const int v = e ? std::atoi(e) : 3;
    const char* et = std::getenv("FELGPU_REFINE_TOL");
    const double t = et ? std::atof(et) : 1.0e-7;
    return RefineLimits{v < 0 ? 0 : (v > 8 ? 8 : v), t > 0.0 ? t : 1.0e-7};
  }();
  return r;
}

static int refine_steps() { return refine_limits().steps; }
static double refine_tol() { return refine_limits().tol; }

// The FP64 residual r = div - L dp of the operator the FFT+TDMA inverts
// (periodic in i and k; the wall rows' zero fluxes are as(1) = an(nj) = 0).
struct PoissonResidual {
  Dims d;
  const double *div, *dp, *as, *an;
  double dx2r, dz2r;
};

static PoissonResidual poisson_residual_args(const Dims d, const double* div, const double* dp) {
  const PoissonRecord& pr = prec();
  const double* as = pr.tab + d.ni + d.nk;
  return PoissonResidual{d, div, dp, as, as + (d.nj + 6), pr.dx2r, pr.dz2r};
}

// One pass of the residual: r = div - L dp in FP64 at every interior cell,
// stored rounded to float into rc (compact, i fastest, then j, then k: the
// real input of the x transforms), and per block the sums of div^2 and r^2.
// A block covers 64 x 4 cells of kResPlanes planes; each thread marches in
// k keeping dp(k-1), dp(k), dp(k+1) of its column in registers.
constexpr int kResPlanes = 8;

__global__ void __launch_bounds__(256) poisson_residual_kernel(PoissonResidual a, int gx, int gy,
                                                               float* rc, double* part) {
  __shared__ double sh[2][256];
  const Dims& d = a.d;
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int L = (int)blockIdx.x;
  const int bx = L % gx, tt = L / gx, by = tt % gy, bz = tt / gy;
  const int i = 1 + bx * 64 + (int)threadIdx.x, j = 1 + by * 4 + (int)threadIdx.y;
  double s0 = 0.0, s1 = 0.0;
  if (i <= ni && j <= nj) {
    const int Y = d.sx, Z = d.sx * d.sy;
    const int dim = (i == 1 ? ni : i - 1) - i, dip = (i == ni ? 1 : i + 1) - i;
    const double a_s = a.as[j + 2], a_n = a.an[j + 2];
    const int k0 = 1 + bz * kResPlanes;
    const int k1 = k0 + kResPlanes - 1 < nk ? k0 + kResPlanes - 1 : nk;
    int c = d.ix(i, j, k0);
    double pm = a.dp[k0 == 1 ? d.ix(i, j, nk) : c - Z];
    double p0 = a.dp[c];
    size_t o = (size_t)(i - 1) + (size_t)ni * ((size_t)(j - 1) + (size_t)nj * (size_t)(k0 - 1));
    const size_t plane = (size_t)ni * nj;
    for (int k = k0; k <= k1; k++) {
      const double pp = a.dp[k == nk ? d.ix(i, j, 1) : c + Z];
      double l = a.dx2r * (a.dp[c + dim] - 2.0 * p0 + a.dp[c + dip]) +
                 a.dz2r * (pm - 2.0 * p0 + pp);
      if (a_s != 0.0) l += a_s * (a.dp[c - Y] - p0);
      if (a_n != 0.0) l += a_n * (a.dp[c + Y] - p0);
      const double dv = a.div[c];
      const double r = dv - l;
      rc[o] = (float)r;
      s0 += dv * dv;
      s1 += r * r;
      pm = p0;
      p0 = pp;
      c += Z;
      o += plane;
    }
  }
  const int t = (int)threadIdx.x + 64 * (int)threadIdx.y;
  sh[0][t] = s0;
  sh[1][t] = s1;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) {
      sh[0][t] += sh[0][t + st];
      sh[1][t] += sh[1][t + st];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[2 * (size_t)L] = sh[0][0];
    part[2 * (size_t)L + 1] = sh[1][0];
  }
}

// The residual of dp into the compact float array rc; the sums of div^2 and
// r^2 go to the device scalars `sums` (two doubles).
static void poisson_residual_pass(const Dims d, const double* div, const double* dp, float* rc,
                                  double* sums) {
  const int gx = (d.ni + 63) / 64, gy = (d.nj + 3) / 4, gz = (d.nk + kResPlanes - 1) / kResPlanes;
  const int nb = gx * gy * gz;
  double* part = static_cast<double*>(scratch(18, sizeof(double) * 2 * (size_t)nb));
  poisson_residual_kernel<<<dim3(nb), dim3(64, 4, 1), 0, stream()>>>(
      poisson_residual_args(d, div, dp), gx, gy, rc, part);
  FEL_CHECK(hipGetLastError());
  auto g = [=] __device__(const double* t, double* out) {
    out[0] = t[0];
    out[1] = t[1];
  };
  sum_stage2<2, double, decltype(g)><<<dim3(1), dim3(256), 0, stream()>>>(nb, part, sums, g);
  FEL_CHECK(hipGetLastError());
}

// |r| / |div| from the sums (0 when both vanish); waits for them.
static double residual_ratio(const double* sums) {
  const double* r = read_scalars(sums, 2);
  if (r[0] > 0.0) return std::sqrt(r[1] / r[0]);
  return r[1] > 0.0 ? HUGE_VAL : 0.0;
}

// The x transforms' real input from the compact residual.
static auto compact_rhs(const Dims d, const float* rc) {
  const size_t ni = (size_t)d.ni, nj = (size_t)d.nj;
  return [=] __device__(int i, int j, int k) {
    return rc[(size_t)(i - 1) + ni * ((size_t)(j - 1) + nj * (size_t)(k - 1))];
  };
}

static float* residual_buffer(const Dims d) {
  return static_cast<float*>(scratch(19, sizeof(float) * (size_t)d.ni * d.nj * d.nk));
}

// Refines dp (each step one FP32 solve of the FP64 residual, added in FP64)
// until the relative residual is at most refine_tol(), with at most
// refine_steps() steps; `warm` counts the solve as a warm start.  Each
// residual pass checks the last iterate and is the next step's right-hand
// side.
static void poisson_refine(const Dims d, double* dp, bool warm) {
  PoissonRecord& pr = prec();
  const int steps = refine_steps();
  if (steps <= 0 || !pr.rhs || !pr.coef || pr.ni != d.ni || pr.nj != d.nj || pr.nk != d.nk) return;
  if (content_version(pr.div_host, d.n3) != pr.div_ver) return;  // right-hand side changed
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const size_t nm = (size_t)ni * nj * nk;
  float* mr = static_cast<float*>(scratch(9, sizeof(float) * 2 * nm));
  float* mi = mr + nm;
  float* rc = residual_buffer(d);
  double* sums = dev_scalars(9);
  // the right-hand side's current device copy (its buffer may have moved)
  Call c("solp_fft_tdma3_");
  const double* div = c.in_t<double>(pr.div_host, d.n3, "div");
  const double tol = refine_tol();
  poisson_residual_pass(d, div, dp, rc, sums);
  double rel = residual_ratio(sums);
  pr.first_rel = rel;
  int it = 0;
  for (; it < steps && rel > tol; it++) {
    fft_forward_of<float>(d, compact_rhs(d, rc), mr, mi);
    tdma_modes<float>(ni, nj, nk, pr.tab, pr.tab + ni, pr.tab + ni + nk, pr.dx2r, pr.dz2r, mr, mi);
    fft_inverse<float>(d, mr, mi, dp, true);
    poisson_residual_pass(d, div, dp, rc, sums);
    rel = residual_ratio(sums);
  }
  pr.solves++;
  pr.warm_solves += warm ? 1 : 0;
  pr.refine_total += (uint64_t)it;
  pr.final_rel = rel;
}

template <class T>
void caldiv_t(const Dims d, double dxinv, double dyinv, double dzinv, double dt, const double* uk,
              const double* vk, const double* wk, double* div, double* sum) {
  Call c("caldiv_");
  const T *u = c.in_t<T>(uk, d.n3, "uk"), *v = c.in_t<T>(vk, d.n3, "vk"),
          *w = c.in_t<T>(wk, d.n3, "wk");
  PT<T>* dv = c.io_t<PT<T>>(div, d.n3, "div");
  double* s = dev_scalars(3);
  divergence<T>(d, dxinv, dyinv, dzinv, 1.0 / dt, u, v, w, dv, s);
  *sum = read_scalars(s, 1)[0];
  c.finish();
}

// Warm start of the FP32 solve (FP32 modes, FELGPU_REFINE >= 1): from the
// second solve on, solp_fft_tdma1_ transforms the FP64 residual
// r = div - L dp0 of the correction dp0 the last solve left in its array
// instead of div, and solp_fft_tdma3_ adds the solved correction to dp0:
// one refinement step from dp0, after which the residual check of
// poisson_refine decides on further steps.  The mode arrays then hold the
// residual's transform, not div's.  If dp0 or the modal table changed
// between the two calls, tdma3 refines from dp = 0 instead.
template <class T>
void tdma1_t(const Dims d, const double* div, double* phir_r, double* phir_i) {
  const size_t nm = (size_t)d.ni * d.nj * d.nk;
  PoissonRecord& rec = prec();
  bool warm = false;
  if constexpr (is_f32<T>)
    warm = refine_steps() > 0 && rec.coef && rec.dp_host && rec.ni == d.ni && rec.nj == d.nj &&
           rec.nk == d.nk && content_version(div, d.n3) != 0 &&
           content_version(rec.dp_host, d.n3) != 0;
  Call c("solp_fft_tdma1_");
  const PT<T>* dv = c.in_t<PT<T>>(div, d.n3, "div");
  T* pr = c.out_t<T>(phir_r, nm, "phir_r");
  T* pi = c.out_t<T>(phir_i, nm, "phir_i");
  if constexpr (is_f32<T>) {
    if (warm) {
      const double* dp0 = c.in_t<double>(rec.dp_host, d.n3, "dp");
      float* rc = residual_buffer(d);
      poisson_residual_pass(d, dv, dp0, rc, dev_scalars(9));
      fft_forward_of<T>(d, compact_rhs(d, rc), pr, pi);
    }
  }
  if (!warm) fft_forward<T>(d, dv, pr, pi);
  c.finish();
  if constexpr (!is_f32<T>) return;  // the record serves the FP32 refinement
  rec.rhs = true;
  rec.ni = d.ni; rec.nj = d.nj; rec.nk = d.nk;
  rec.div_host = div;
  rec.div_ver = content_version(div, d.n3);
  if (rec.div_ver == 0) rec.rhs = false;  // unregistered: no refinement
  rec.warm = warm;
  if (warm) {
    rec.dp0_ver = content_version(rec.dp_host, d.n3);
    rec.warm_gen = rec.tab_gen;
  }
}

template <class T>
void tdma3_t(const Dims d, const double* phir_r, const double* phir_i, double* dp) {
  const size_t nm = (size_t)d.ni * d.nj * d.nk;
  PoissonRecord& rec = prec();
  bool warm = false;
  if constexpr (is_f32<T>)
    warm = rec.warm && rec.dp_host == (const void*)dp && rec.rhs &&
           content_version(dp, d.n3) == rec.dp0_ver && rec.warm_gen == rec.tab_gen;
  const bool stale = is_f32<T> && rec.warm && !warm;  // modes of a residual no longer valid
  Call c("solp_fft_tdma3_");
  const T* pr = c.in_t<T>(phir_r, nm, "phir_r");
  const T* pi = c.in_t<T>(phir_i, nm, "phir_i");
  PT<T>* ddp = c.io_t<PT<T>>(dp, d.n3, "dp");
  if constexpr (is_f32<T>) {
    if (stale) {
      // refine from dp = 0 (its residual is div)
      launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk},
                 [=] __device__(int i, int j, int k) { ddp[d.ix(i, j, k)] = 0.0; });
    } else {
      fft_inverse<T>(d, pr, pi, ddp, warm);
    }
    poisson_refine(d, ddp, warm);
    rec.warm = false;
    rec.dp_host = dp;
  } else {
    fft_inverse<T>(d, pr, pi, ddp, false);
  }
  c.finish();
}

template <class Q>
void avzero_t(const Dims d, double* q, int ndiv) {
  Call c("bnd_avzero_");
  avzero<Q>(d, c.io_t<Q>(q, d.n3, "q"), ndiv);
  c.finish();
}

template <class Q>
double cal_av_t(const Dims d, const double* q) {
  Call c("cal_av_");
  const Q* dq = c.in_t<Q>(q, d.n3, "q");
  double* s = dev_scalars(5);
  interior_reduce<1>(
      d, [=] __device__(int i, int j, int k, double* o) { o[0] = val(dq[d.ix(i, j, k)]); },
      [=] __device__(const double* t, double* out) { out[0] = t[0]; }, s);
  const double sum = read_scalars(s, 1)[0];
  c.finish();
  return sum;
}

template <class T>
void corunp_t(const Dims d, int jen, double rho, double dxinv, double dyinv, double dzinv, double dt,
              const double* dp, const double* phat, double* un, double* vn, double* wn, double* pn) {
  Call c("corunp_explicit_");
  const PT<T>* ddp = c.in_t<PT<T>>(dp, d.n3, "dp");
  const T* ph = c.in_t<T>(phat, d.n3, "phat");
  T *u = c.io_t<T>(un, d.n3, "un"), *v = c.io_t<T>(vn, d.n3, "vn"), *w = c.io_t<T>(wn, d.n3, "wn"),
    *p = c.interior_t<T>(pn, d, "pn");  // written on the interior only
  correct<T>(d, jen, rho, dxinv, dyinv, dzinv, dt, ddp, ph, u, v, w, p);
  c.finish();
}

// ---------------------------------------------------------------------------
// Adjoints (reverse mode) of the projection.  The reference has no adjoint:
// <name>_b_ applies the transpose of the linear map <name>_ applies to its
// active arrays.  Every active array argument of the forward list is
// replaced by its adjoint, of the same shape; adjoints of the forward's
// inputs are incremented, adjoints of what it overwrites are consumed (set
// to zero), in/out adjoints are updated in place.  Each kernel gathers what
// one cell receives, so no two threads write one element.  FP64 mode only.
// ---------------------------------------------------------------------------
void adjoint_fp64_only(const char* who) {
  if (flow_fp32() || tangent_mode()) {
    std::fprintf(stderr, "felgpu: %s: the adjoint routines run in the FP64 mode only\n", who);
    std::abort();
  }
}

// srcs/caldiv.f:27-38 transposed: div(c) on [-1,n+3]^3 reads u(c) with +fx
// and u(c-x) with -fx (fx = dxinv/dt), so u_b(c) += (div_b(c) - div_b(c+x))
// fx over the cells of the box, likewise v, w; then div_b = 0 on the box.
static void caldiv_adj(const Dims d, double fx, double fy, double fz, double* ub, double* vb,
                       double* wb, double* divb) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  launch_box(Box{-2, ni + 3, -2, nj + 3, -2, nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    const bool b0 = i >= -1 && j >= -1 && k >= -1;
    const double g0 = b0 ? divb[c] : 0.0;
    const double gx = (j >= -1 && k >= -1 && i + 1 <= ni + 3) ? divb[c + X] : 0.0;
    const double gy = (i >= -1 && k >= -1 && j + 1 <= nj + 3) ? divb[c + Y] : 0.0;
    const double gz = (i >= -1 && j >= -1 && k + 1 <= nk + 3) ? divb[c + Z] : 0.0;
    ub[c] = ub[c] + (g0 - gx) * fx;
    vb[c] = vb[c] + (g0 - gy) * fy;
    wb[c] = wb[c] + (g0 - gz) * fz;
  });
  launch_box(Box{-1, ni + 3, -1, nj + 3, -1, nk + 3},
             [=] __device__(int i, int j, int k) { divb[d.ix(i, j, k)] = 0.0; });
}

// srcs/bnd_avzero.f:10-23 transposed: q_b(interior) -= (sum of q_b over the
// padded array) / (ni nj nk ndiv), the sum formed on the device.
static void avzero_adj(const Dims d, double* qb, int ndiv) {
  double* s = dev_scalars(20);
  const double nn = (double)(d.ni * d.nj * d.nk * ndiv);
  Dims p = d;  // the padded array as the reduction's index range
  p.ni = d.sx;
  p.nj = d.sy;
  p.nk = d.sz;
  const int sx = d.sx, sy = d.sy;
  interior_reduce<1>(
      p,
      [=] __device__(int i, int j, int k, double* o) {
        o[0] = qb[(i - 1) + sx * ((j - 1) + sy * (k - 1))];
      },
      [=] __device__(const double* t, double* out) { out[0] = t[0] / nn; }, s);
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    double* q = qb + d.ix(i, j, k);
    *q = *q - *s;
  });
}

// srcs/corunp_explicit.f:26-74 transposed: the velocity adjoints pass
// through unchanged; dp_b(c) gathers +f u_b(c) - f u_b(c-x) (likewise v up
// to jen, w) and pn_b(c) over the interior; phat_b += pn_b and pn_b = 0 on
// the interior.
static void corunp_adj(const Dims d, int jen, double fx, double fy, double fz, double* dpb,
                       double* phatb, const double* unb, const double* vnb, const double* wnb,
                       double* pnb) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  launch_box(Box{1, ni + 1, 1, nj + 1, 1, nk + 1}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    double t = 0.0;
    if (i <= ni && j <= nj && k <= nk) {
      t = t + unb[c] * fx;
      if (j <= jen) t = t + vnb[c] * fy;
      t = t + wnb[c] * fz;
      const double pb = pnb[c];
      t = t + pb;
      phatb[c] = phatb[c] + pb;
      pnb[c] = 0.0;
    }
    if (i >= 2 && j <= nj && k <= nk) t = t - unb[c - X] * fx;
    if (j >= 2 && j - 1 <= jen && i <= ni && k <= nk) t = t - vnb[c - Y] * fy;
    if (k >= 2 && i <= ni && j <= nj) t = t - wnb[c - Z] * fz;
    dpb[c] = dpb[c] + t;
  });
}

// Device table of the Thomas solves of solp_fft_tdma2_ (srcs/solp_fft_tdma2.f:
// 46-87): the modal factors 2 cos(theta_x) (ni), 2 cos(theta_z) (nk) from the
// host libm, as the reference's cos, then as, an, ap (nj+6 each).  Uploaded
// only when it changes (the coefficients are fixed for a run), so a solve
// has no host synchronisation here; a change counts as a new operator for
// the FP32 warm start.
static const double* mode_table(int ID, const int* svall, const double* as_p_fft,
                                const double* an_p_fft, const double* ap_p_fft) {
  const int ni = svall[0], nj = svall[1], nk = svall[2];
  const double pi_ = std::atan(1.0) * 4.0;
  const double pi2_ni = pi_ * 2.0 / (double)ni, pi2_nk = pi_ * 2.0 / (double)nk;
  std::vector<double> h((size_t)ni + nk + 3 * (nj + 6));
  for (int i = 0; i < ni; i++) h[i] = std::cos((double)(i - ni / 2) * pi2_ni) * 2.0;
  for (int k = 0; k < nk; k++) h[ni + k] = std::cos((double)(k + 1 + ID * nk - nk / 2 - 1) * pi2_nk) * 2.0;
  for (int j = 0; j < nj + 6; j++) {
    h[ni + nk + j] = as_p_fft[j];
    h[ni + nk + (nj + 6) + j] = an_p_fft[j];
    h[ni + nk + 2 * (nj + 6) + j] = ap_p_fft[j];
  }
  static std::vector<double> up;  // what the device table holds
  static double* tab = nullptr;
  static size_t cap = 0;
  if (h.size() > cap) {
    if (tab) {
      FEL_CHECK(hipStreamSynchronize(stream()));
      FEL_CHECK(hipFree(tab));
    }
    FEL_CHECK(hipMalloc(&tab, sizeof(double) * h.size()));
    cap = h.size();
    up.clear();
  }
  if (h != up) {
    up = h;
    FEL_CHECK(hipMemcpyAsync(tab, up.data(), sizeof(double) * up.size(), hipMemcpyHostToDevice,
                             stream()));
    FEL_CHECK(hipStreamSynchronize(stream()));
    prec().tab_gen++;
  }
  return tab;
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/caldiv.f:1-2 (ipara must be 0)
void caldiv_(const int* ipara, const int* ndiv, const int* ni, const int* nj, const int* nk,
             const double* dxinv, const double* dyinv, const double* dzinv, const double* dt,
             const double* uk, const double* vk, const double* wk, double* div, double* err_div) {
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: caldiv_: ipara=%d, only the single-domain case (0) is supported\n", *ipara);
    std::abort();
  }
  const Dims d = make_dims(*ni, *nj, *nk);
  double sum = 0.0;
  with_flow([&](auto t) {
    caldiv_t<decltype(t)>(d, *dxinv, *dyinv, *dzinv, *dt, uk, vk, wk, div, &sum);
  });
  *err_div = std::sqrt(sum / (double)(*ni * *nj * *nk * *ndiv)) + 1.0e-99;
}

// srcs/solp_fft_tdma1.f:4-5.  fftdata (the reference's work array) is not used.
void solp_fft_tdma1_(const int* ni, const int* nj, const int* nk, double* fftdata, const double* div,
                     double* phir_r, double* phir_i) {
  (void)fftdata;
  const Dims d = make_dims(*ni, *nj, *nk);
  // the modes' precision is the modal arrays' (FP32 in the FP32 mode when
  // they are registered with the flow class: FP32 solve + FP64 refinement)
  with_elem(elem_size(phir_r), [&](auto t) { tdma1_t<decltype(t)>(d, div, phir_r, phir_i); });
}

// srcs/trans_r2w.f:1-38 on one rank: the relabelled array has the same
// memory image; copy only when the caller passes distinct arrays.
void trans_r2w_(const int* ipara, const int* ID, const int* ndiv, const int* svall, const int* key,
                double* sendbuf, double* recvbuf, const double* r, double* w) {
  (void)ID; (void)key; (void)sendbuf; (void)recvbuf;
  if (*ipara != 0 || *ndiv != 1) {
    std::fprintf(stderr, "felgpu: trans_r2w_: only the single-domain case is supported\n");
    std::abort();
  }
  if ((const void*)r == (const void*)w) return;
  const size_t n = (size_t)svall[0] * svall[1] * svall[2];
  Call c("trans_r2w_");
  with_elem(elem_size(r), [&](auto t) {
    using T = decltype(t);
    const T* a = c.in_t<T>(r, n, "r");
    T* b = c.out_t<T>(w, n, "w");
    FEL_CHECK(hipMemcpyAsync(b, a, sizeof(T) * n, hipMemcpyDeviceToDevice, stream()));
  });
  c.finish();
}

void trans_w2r_(const int* ipara, const int* ID, const int* ndiv, const int* svall, const int* key,
                double* sendbuf, double* recvbuf, const double* w, double* r) {
  trans_r2w_(ipara, ID, ndiv, svall, key, sendbuf, recvbuf, w, r);
}

// srcs/solp_fft_tdma2.f:4-5.  The 1-D coefficient arrays (as/an/ap_p_fft,
// declared (-2:nj+3)) are host arrays read here; atdma/btdma_r/btdma_i (the
// reference's per-line work arrays) are not used.
void solp_fft_tdma2_(const int* ID, const int* ndiv, const int* svall, const double* rho,
                     const double* dxinv, const double* dzinv, const double* as_p_fft,
                     const double* an_p_fft, const double* ap_p_fft, double* atdma, double* btdma_r,
                     double* btdma_i, double* phiw_r, double* phiw_i) {
  (void)atdma; (void)btdma_r; (void)btdma_i;
  if (*ndiv != 1) {
    std::fprintf(stderr, "felgpu: solp_fft_tdma2_: only the single-domain case is supported\n");
    std::abort();
  }
  const int ni = svall[0], nj = svall[1], nk = svall[2];
  const size_t nm = (size_t)ni * nj * nk;
  const double* tab = mode_table(*ID, svall, as_p_fft, an_p_fft, ap_p_fft);
  const double dx2r = (*dxinv * *dxinv) / *rho, dz2r = (*dzinv * *dzinv) / *rho;
  {
    PoissonRecord& rec = prec();
    if (rec.dx2r != dx2r || rec.dz2r != dz2r) rec.tab_gen++;  // a warm start's operator
  }
  Call c("solp_fft_tdma2_");
  with_elem(elem_size(phiw_r), [&](auto t) {
    using T = decltype(t);
    T* pr = c.io_t<T>(phiw_r, nm, "phiw_r");
    T* pi = c.io_t<T>(phiw_i, nm, "phiw_i");
    tdma_modes<T>(ni, nj, nk, tab, tab + ni, tab + ni + nk, dx2r, dz2r, pr, pi);
  });
  c.finish();
  PoissonRecord& rec = prec();
  rec.coef = true;
  rec.dx2r = dx2r;
  rec.dz2r = dz2r;
  rec.tab = tab;
}

// srcs/solp_fft_tdma3.f:4-5
void solp_fft_tdma3_(const int* ni, const int* nj, const int* nk, double* fftdata,
                     const double* phir_r, const double* phir_i, double* dp) {
  (void)fftdata;
  const Dims d = make_dims(*ni, *nj, *nk);
  with_elem(elem_size(phir_r), [&](auto t) { tdma3_t<decltype(t)>(d, phir_r, phir_i, dp); });
}

// srcs/solp_fft_tdma4.f:4-5: residual monitor every imon_t steps.
void solp_fft_tdma4_(const int* ipara, const int* ID, const int* ndiv, const int* ni, const int* nj,
                     const int* nk, const int* nstep, const int* imon_t, const double* rho,
                     const double* dxinv, const double* dyinv, const double* dzinv, const double* div,
                     const double* dp) {
  (void)ipara;
  if (!(*imon_t > 0 && (*nstep % *imon_t) == 0)) return;
  const Dims d = make_dims(*ni, *nj, *nk);
  double* s = dev_scalars(4);
  const double* r = nullptr;
  with_flow([&](auto t) {
    using P = PT<decltype(t)>;
    Call c("solp_fft_tdma4_");
    const P* dv = c.in_t<P>(div, d.n3, "div");
    const P* dd = c.in_t<P>(dp, d.n3, "dp");
    poisson_residual_sums<P>(d, *rho, *dxinv, *dyinv, *dzinv, dv, dd, s);
    r = read_scalars(s, 2);
    c.finish();
  });
  const double nn = (double)(*ni * *nj * *nk * *ndiv);
  const double e0 = std::sqrt(r[0] / nn) + 1.0e-99, e = std::sqrt(r[1] / nn) + 1.0e-99;
  if (*ID == 0) {
    char a[3][32];
    std::printf("Err_solp %10d%s%s%s\n", *nstep, fortran_e(a[0], e / e0, 20, 10),
                fortran_e(a[1], e, 20, 10), fortran_e(a[2], e0, 20, 10));
  }
}

// Relative Poisson residual (the monitor's value), for tests and drivers.
double felgpu_poisson_residual_(const int* ni, const int* nj, const int* nk, const double* rho,
                                const double* dxinv, const double* dyinv, const double* dzinv,
                                const double* div, const double* dp) {
  const Dims d = make_dims(*ni, *nj, *nk);
  double* s = dev_scalars(4);
  const double* r = nullptr;
  with_flow([&](auto t) {
    using P = PT<decltype(t)>;
    Call c("felgpu_poisson_residual_");
    const P* dv = c.in_t<P>(div, d.n3, "div");
    const P* dd = c.in_t<P>(dp, d.n3, "dp");
    poisson_residual_sums<P>(d, *rho, *dxinv, *dyinv, *dzinv, dv, dd, s);
    r = read_scalars(s, 2);
    c.finish();
  });
  const double nn = (double)(*ni * *nj * *nk);
  return (std::sqrt(r[1] / nn) + 1.0e-99) / (std::sqrt(r[0] / nn) + 1.0e-99);
}

// Refinement limits of the FP32 pressure solve: *max_steps < 0 or
// *tol <= 0 keeps the current value.
void felgpu_set_refinement_(const int* max_steps, const double* tol) {
  RefineLimits& r = refine_limits();
  if (*max_steps >= 0) r.steps = *max_steps > 8 ? 8 : *max_steps;
  if (*tol > 0.0) r.tol = *tol;
}

// Statistics of the FP32 mode's refined pressure solves (see
// poisson_refine); *reset != 0 clears the counts after reading them.
void felgpu_poisson_stats_(const int* reset, long long* solves, long long* warm_solves,
                           long long* refine_steps, double* first_rel, double* final_rel) {
  PoissonRecord& pr = prec();
  *solves = (long long)pr.solves;
  *warm_solves = (long long)pr.warm_solves;
  *refine_steps = (long long)pr.refine_total;
  *first_rel = pr.first_rel;
  *final_rel = pr.final_rel;
  if (*reset) pr.solves = pr.warm_solves = pr.refine_total = 0;
}

// srcs/bnd_avzero.f:1 (ipara must be 0)
void bnd_avzero_(const int* ipara, const int* ndiv, const int* ni, const int* nj, const int* nk,
                 double* q) {
  (void)ipara;
  const Dims d = make_dims(*ni, *nj, *nk);
  with_elem(elem_size(q), [&](auto t) { avzero_t<decltype(t)>(d, q, *ndiv); });
}

// srcs/cal_av.f:1
void cal_av_(const int* ipara, const int* ndiv, const int* ni, const int* nj, const int* nk,
             const double* q, double* av) {
  (void)ipara;
  const Dims d = make_dims(*ni, *nj, *nk);
  double sum = 0.0;
  with_elem(elem_size(q), [&](auto t) { sum = cal_av_t<decltype(t)>(d, q); });
  *av = sum / (double)(*ni * *nj * *nk * *ndiv);
}

// srcs/corunp_explicit.f:1-2
void corunp_explicit_(const int* nID, const int* ni, const int* nj, const int* nk, const double* rho,
                      const double* dxinv, const double* dyinv, const double* dzinv, const double* dt,
                      const double* dp, const double* phat, double* un, double* vn, double* wn,
                      double* pn) {
  const Dims d = make_dims(*ni, *nj, *nk);
  const int jen = (nID[kYPlus] < 0) ? *nj - 1 : *nj;
  with_flow([&](auto t) {
    corunp_t<decltype(t)>(d, jen, *rho, *dxinv, *dyinv, *dzinv, *dt, dp, phat, un, vn, wn, pn);
  });
}

// srcs/cal_coef_prs.f:5-57 (set-up, host arrays declared (-2:nj+3))
void cal_coef_prs_(const int* nj_, const double* rho_, const double* dxinv_, const double* dyinv_,
                   const double* dzinv_, double* aw, double* ae, double* as, double* an, double* ab,
                   double* at, double* ap) {
  const int nj = *nj_;
  const double rho = *rho_, dxinv = *dxinv_, dyinv = *dyinv_, dzinv = *dzinv_;
  aw += 2; ae += 2; as += 2; an += 2; ab += 2; at += 2; ap += 2;
  for (int j = 1; j <= nj; j++) {
    aw[j] = dxinv * dxinv / rho;
    ae[j] = dxinv * dxinv / rho;
    as[j] = dyinv * dyinv / rho;
    an[j] = dyinv * dyinv / rho;
    ab[j] = dzinv * dzinv / rho;
    at[j] = dzinv * dzinv / rho;
    ap[j] = +aw[j] + ae[j] + as[j] + an[j] + ab[j] + at[j];
  }
  ap[1] = ap[1] - as[1];
  as[1] = 0.0;
  ap[nj] = ap[nj] - an[nj];
  an[nj] = 0.0;
}

// srcs/cal_coef_solp_fft.f:5-42 (set-up, host arrays)
void cal_coef_solp_fft_(const int* nj_, const double* as_p, const double* an_p, double* as_fft,
                        double* an_fft, double* ap_fft) {
  const int nj = *nj_;
  as_p += 2; an_p += 2; as_fft += 2; an_fft += 2; ap_fft += 2;
  for (int j = 1; j <= nj; j++) {
    as_fft[j] = as_p[j];
    an_fft[j] = an_p[j];
    ap_fft[j] = as_fft[j] + an_fft[j];
  }
  ap_fft[1] = ap_fft[1] - as_fft[1];
  as_fft[1] = 0.0;
  ap_fft[nj] = ap_fft[nj] - an_fft[nj];
  an_fft[nj] = 0.0;
}

// ---------------------------------------------------------------------------
// Adjoint entry points (the forward's argument list with the active arrays
// replaced by their adjoints; see the adjoint section above).
// ---------------------------------------------------------------------------

// Adjoint of caldiv_ (srcs/caldiv.f:27-38).  The RMS output is a
// diagnostic and has no adjoint argument.
void caldiv_b_(const int* ipara, const int* ndiv, const int* ni, const int* nj, const int* nk,
               const double* dxinv, const double* dyinv, const double* dzinv, const double* dt,
               double* ukb, double* vkb, double* wkb, double* divb) {
  (void)ndiv;
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: caldiv_b_: only the single-domain case is supported\n");
    std::abort();
  }
  adjoint_fp64_only("caldiv_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const double dtinv = 1.0 / *dt;
  Call c("caldiv_b_");
  double* ub = c.io(ukb, d.n3, "ukb");
  double* vb = c.io(vkb, d.n3, "vkb");
  double* wb = c.io(wkb, d.n3, "wkb");
  double* gb = c.io(divb, d.n3, "divb");
  caldiv_adj(d, *dxinv * dtinv, *dyinv * dtinv, *dzinv * dtinv, ub, vb, wb, gb);
  c.finish();
}

// Adjoint of solp_fft_tdma1_ (srcs/solp_fft_tdma1.f:25-47): div_b(interior)
// += Re(unnormalised inverse DFT of phir_b) / (ni nk); phir_b = 0.
void solp_fft_tdma1_b_(const int* ni, const int* nj, const int* nk, double* fftdata, double* divb,
                       double* phir_rb, double* phir_ib) {
  (void)fftdata;
  adjoint_fp64_only("solp_fft_tdma1_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const int nm = d.ni * d.nj * d.nk;
  Call c("solp_fft_tdma1_b_");
  double* gb = c.io(divb, d.n3, "divb");
  double* pr = c.io(phir_rb, nm, "phir_rb");
  double* pi = c.io(phir_ib, nm, "phir_ib");
  const double s = 1.0 / (double)(d.ni * d.nk);  // a power of two: exact scaling
  launch_flat(nm, [=] __device__(int m) {
    pr[m] = pr[m] * s;
    pi[m] = pi[m] * s;
  });
  fft_inverse<double, double>(d, pr, pi, gb, true);
  launch_flat(nm, [=] __device__(int m) {
    pr[m] = 0.0;
    pi[m] = 0.0;
  });
  c.finish();
}

// Adjoints of trans_r2w_ / trans_w2r_ on one rank (a copy, none when the
// arrays coincide): source_b += dest_b, dest_b = 0.
void trans_r2w_b_(const int* ipara, const int* ID, const int* ndiv, const int* svall, const int* key,
                  double* sendbuf, double* recvbuf, double* rb, double* wb) {
  (void)ID; (void)key; (void)sendbuf; (void)recvbuf;
  if (*ipara != 0 || *ndiv != 1) {
    std::fprintf(stderr, "felgpu: trans_r2w_b_: only the single-domain case is supported\n");
    std::abort();
  }
  adjoint_fp64_only("trans_r2w_b_");
  if (rb == wb) return;
  const int n = svall[0] * svall[1] * svall[2];
  Call c("trans_r2w_b_");
  double* a = c.io(rb, n, "rb");
  double* b = c.io(wb, n, "wb");
  launch_flat(n, [=] __device__(int m) {
    a[m] = a[m] + b[m];
    b[m] = 0.0;
  });
  c.finish();
}

void trans_w2r_b_(const int* ipara, const int* ID, const int* ndiv, const int* svall, const int* key,
                  double* sendbuf, double* recvbuf, double* wb, double* rb) {
  trans_r2w_b_(ipara, ID, ndiv, svall, key, sendbuf, recvbuf, wb, rb);
}

// Adjoint of solp_fft_tdma2_ (srcs/solp_fft_tdma2.f:46-87): the Thomas
// solve of the transposed systems, in place on the adjoints; the singular
// (0,0) mode solved pinned (x(nj) = 0), the transpose of the pinned solve.
void solp_fft_tdma2_b_(const int* ID, const int* ndiv, const int* svall, const double* rho,
                       const double* dxinv, const double* dzinv, const double* as_p_fft,
                       const double* an_p_fft, const double* ap_p_fft, double* atdma,
                       double* btdma_r, double* btdma_i, double* phiw_rb, double* phiw_ib) {
  (void)atdma; (void)btdma_r; (void)btdma_i;
  if (*ndiv != 1) {
    std::fprintf(stderr, "felgpu: solp_fft_tdma2_b_: only the single-domain case is supported\n");
    std::abort();
  }
  adjoint_fp64_only("solp_fft_tdma2_b_");
  const int ni = svall[0], nj = svall[1], nk = svall[2];
  const size_t nm = (size_t)ni * nj * nk;
  const double* tab = mode_table(*ID, svall, as_p_fft, an_p_fft, ap_p_fft);
  const double dx2r = (*dxinv * *dxinv) / *rho, dz2r = (*dzinv * *dzinv) / *rho;
  Call c("solp_fft_tdma2_b_");
  double* pr = c.io(phiw_rb, nm, "phiw_rb");
  double* pi = c.io(phiw_ib, nm, "phiw_ib");
  tdma_modes<double, true>(ni, nj, nk, tab, tab + ni, tab + ni + nk, dx2r, dz2r, pr, pi);
  c.finish();
}

// Adjoint of solp_fft_tdma3_ (srcs/solp_fft_tdma3.f:25-44): phir_b +=
// unnormalised forward DFT of dp_b(interior); dp_b(interior) = 0.
void solp_fft_tdma3_b_(const int* ni, const int* nj, const int* nk, double* fftdata, double* phir_rb,
                       double* phir_ib, double* dpb) {
  (void)fftdata;
  adjoint_fp64_only("solp_fft_tdma3_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const size_t nm = (size_t)d.ni * d.nj * d.nk;
  Call c("solp_fft_tdma3_b_");
  double* pr = c.io(phir_rb, nm, "phir_rb");
  double* pi = c.io(phir_ib, nm, "phir_ib");
  double* g = c.io(dpb, d.n3, "dpb");
  fft_forward_core<double, true>(
      d, [=] __device__(int i, int j, int k) { return g[d.ix(i, j, k)]; }, pr, pi, 1.0);
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk},
             [=] __device__(int i, int j, int k) { g[d.ix(i, j, k)] = 0.0; });
  c.finish();
}

// Adjoint of bnd_avzero_ (srcs/bnd_avzero.f:10-23).
void bnd_avzero_b_(const int* ipara, const int* ndiv, const int* ni, const int* nj, const int* nk,
                   double* qb) {
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: bnd_avzero_b_: only the single-domain case is supported\n");
    std::abort();
  }
  adjoint_fp64_only("bnd_avzero_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("bnd_avzero_b_");
  avzero_adj(d, c.io(qb, d.n3, "qb"), *ndiv);
  c.finish();
}

// Adjoint of corunp_explicit_ (srcs/corunp_explicit.f:26-74).
void corunp_explicit_b_(const int* nID, const int* ni, const int* nj, const int* nk,
                        const double* rho, const double* dxinv, const double* dyinv,
                        const double* dzinv, const double* dt, double* dpb, double* phatb,
                        double* unb, double* vnb, double* wnb, double* pnb) {
  adjoint_fp64_only("corunp_explicit_b_");
  const Dims d = make_dims(*ni, *nj, *nk);
  const int jen = nID[kYPlus] < 0 ? d.nj - 1 : d.nj;
  const double fx = 1.0 / *rho * *dxinv * *dt, fy = 1.0 / *rho * *dyinv * *dt,
               fz = 1.0 / *rho * *dzinv * *dt;
  Call c("corunp_explicit_b_");
  double* gdp = c.io(dpb, d.n3, "dpb");
  double* gph = c.io(phatb, d.n3, "phatb");
  const double* gu = c.in(unb, d.n3, "unb");
  const double* gv = c.in(vnb, d.n3, "vnb");
  const double* gw = c.in(wnb, d.n3, "wnb");
  double* gpn = c.io(pnb, d.n3, "pnb");
  corunp_adj(d, jen, fx, fy, fz, gdp, gph, gu, gv, gw, gpn);
  c.finish();
}

}  // extern "C"
