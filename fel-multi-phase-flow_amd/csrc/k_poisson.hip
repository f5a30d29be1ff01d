// k_poisson.hip -- pressure projection: divergence, 2-D FFT in the periodic
// (i,k) planes, Thomas solves along the wall-normal j, inverse FFT, mean
// removal, velocity correction.
//
// Reference: srcs/caldiv.f, srcs/solp_fft_tdma1.f, srcs/trans_r2w.f,
// srcs/solp_fft_tdma2.f, srcs/trans_w2r.f, srcs/solp_fft_tdma3.f,
// srcs/solp_fft_tdma4.f, srcs/bnd_avzero.f, srcs/cal_av.f,
// srcs/corunp_explicit.f, srcs/cal_coef_prs.f, srcs/cal_coef_solp_fft.f.
//
// The transforms are this library's own batched radix-2 Stockham FFTs run in
// LDS (the reference's srcs/fourn.f is not used): the x lines are
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

// Twiddles exp(+2 pi i t / N), t in [0, N/2), from the host libm.
static const double2* twiddles(int n) {
  static std::map<int, double2*> cache;
  auto it = cache.find(n);
  if (it != cache.end()) return it->second;
  std::vector<double2> h(n / 2 > 0 ? n / 2 : 1);
  const double pi = std::acos(-1.0);
  for (int t = 0; t < n / 2; t++) {
    const double a = 2.0 * pi * (double)t / (double)n;
    h[t] = make_double2(std::cos(a), std::sin(a));
  }
  double2* d = nullptr;
  FEL_CHECK(hipMalloc((void**)&d, sizeof(double2) * h.size()));
  FEL_CHECK(hipMemcpy(d, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice));
  cache[n] = d;
  return d;
}

// Stockham radix-2 DIF transform of G-thread groups, one line per group,
// ping-ponging between x and y (LDS, line stride ls).  sign = -1 forward.
// Returns the buffer that holds the result.
__device__ __forceinline__ double2* stockham(double2* x, double2* y, int n, int lane, int g,
                                              const double2* __restrict__ tw, double sign) {
  for (int s = 1; s < n; s *= 2) {
    const int m = n / (2 * s);
    for (int b = lane; b < n / 2; b += g) {
      const int p = b / s, q = b - p * s;
      const double2 w0 = tw[p * s];
      const double wr = w0.x, wi = sign * w0.y;
      const double2 a = x[q + s * p], c = x[q + s * (p + m)];
      const double dr = a.x - c.x, di = a.y - c.y;
      y[q + s * (2 * p)] = make_double2(a.x + c.x, a.y + c.y);
      y[q + s * (2 * p + 1)] = make_double2(dr * wr - di * wi, dr * wi + di * wr);
    }
    __syncthreads();
    double2* t = x;
    x = y;
    y = t;
  }
  return x;
}

// x lines: N = line length (contiguous), LPB lines per 256-thread block.
template <class Load, class Store>
__global__ void __launch_bounds__(256) fft_x_kernel(int n, int nlines, double sign,
                                                    const double2* __restrict__ tw, Load ld, Store st) {
  extern __shared__ double2 sm[];
  const int lpb = n >= 256 ? 1 : 256 / (n / 2 > 0 ? n / 2 : 1);  // lines per block
  const int g = 256 / lpb;                                        // threads per line
  const int li = (int)threadIdx.x / g, lane = (int)threadIdx.x - li * g;
  const int line = (int)blockIdx.x * lpb + li;
  double2* x = sm + li * 2 * n;
  double2* y = x + n;
  const bool act = line < nlines;
  for (int e = lane; e < n; e += g) x[e] = act ? ld(line, e) : make_double2(0.0, 0.0);
  __syncthreads();
  double2* r = stockham(x, y, n, lane, g, tw, sign);
  if (act)
    for (int e = lane; e < n; e += g) st(line, e, r[e]);
}

// Strided lines: a block owns T lines that are adjacent in memory (tile
// index t = 0..T-1) and transforms them together; element e of line l is
// loaded/stored by the functors.
template <class Load, class Store>
__global__ void __launch_bounds__(256) fft_z_kernel(int n, int tiles, int T, double sign,
                                                    const double2* __restrict__ tw, Load ld, Store st) {
  extern __shared__ double2 sm[];
  const int g = 256 / T;                  // threads per line
  const int tile = (int)blockIdx.x;
  // cooperative, row-major loads: consecutive threads take consecutive lines
  for (int idx = (int)threadIdx.x; idx < T * n; idx += 256) {
    const int e = idx / T, l = idx - e * T;
    sm[l * 2 * n + e] = ld(tile, l, e);
  }
  __syncthreads();
  const int l = (int)threadIdx.x / g, lane = (int)threadIdx.x - l * g;
  double2* x = sm + l * 2 * n;
  double2* r = stockham(x, x + n, n, lane, g, tw, sign);
  const int off = (int)(r - x);  // 0 or n, the same for every line
  for (int idx = (int)threadIdx.x; idx < T * n; idx += 256) {
    const int e = idx / T, ll = idx - e * T;
    st(tile, ll, e, sm[ll * 2 * n + off + e]);
  }
  (void)tiles;
}

template <class Load, class Store>
static void fft_x(int n, int nlines, double sign, Load ld, Store st) {
  const int lpb = n >= 256 ? 1 : 256 / (n / 2 > 0 ? n / 2 : 1);
  const size_t shm = sizeof(double2) * 2 * n * lpb;
  fft_x_kernel<Load, Store><<<(nlines + lpb - 1) / lpb, 256, shm, stream()>>>(n, nlines, sign,
                                                                           twiddles(n), ld, st);
  FEL_CHECK(hipGetLastError());
}

// Tile width: T lines of n points, two LDS buffers each, <= 64 KiB.
static int z_tile(int n, int nlines_fast) {
  int T = 16;
  while (T > 1 && (size_t)T * 2 * n * sizeof(double2) > 65536) T /= 2;
  while (T > 1 && nlines_fast % T != 0) T /= 2;
  return T;
}

template <class Load, class Store>
static void fft_z(int n, int tiles, int T, double sign, Load ld, Store st) {
  const size_t shm = sizeof(double2) * 2 * n * T;
  fft_z_kernel<Load, Store><<<tiles, 256, shm, stream()>>>(n, tiles, T, sign, twiddles(n), ld, st);
  FEL_CHECK(hipGetLastError());
}

static void check_pow2(int n, const char* who) {
  if (n < 2 || (n & (n - 1)) != 0) {
    std::fprintf(stderr, "felgpu: %s: FFT extent %d is not a power of two (srcs/main.f90:117-126)\n", who, n);
    std::abort();
  }
}

// srcs/solp_fft_tdma1.f:25-47: phir(i,j,k) = DFT(div(1:ni,j,1:nk))/(ni nk)
// at centred mode indices; T is (nk, nj, ni) complex scratch.
void fft_forward(const Dims d, const double* div, double* pr, double* pi) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  check_pow2(ni, "solp_fft_tdma1_");
  check_pow2(nk, "solp_fft_tdma1_");
  double2* T = static_cast<double2*>(scratch(4, sizeof(double2) * (size_t)ni * nj * nk));
  // x transforms of the real rows (line = j + nj*k)
  fft_x(ni, nj * nk, -1.0,
        [=] __device__(int line, int e) {
          const int j = line % nj, k = line / nj;
          return make_double2(div[d.ix(e + 1, j + 1, k + 1)], 0.0);
        },
        [=] __device__(int line, int e, double2 v) { T[(size_t)line * ni + e] = v; });
  // k transforms; tile = (j, kx block), lines adjacent in kx
  const int Tw = z_tile(nk, ni);
  const int tiles_per_j = ni / Tw;
  const double nrm = (double)(ni * nk);
  fft_z(nk, nj * tiles_per_j, Tw, -1.0,
        [=] __device__(int tile, int l, int e) {
          const int j = tile % nj, kx = (tile / nj) * Tw + l;
          return T[((size_t)e * nj + j) * ni + kx];
        },
        [=] __device__(int tile, int l, int e, double2 v) {
          const int j = tile % nj, kx = (tile / nj) * Tw + l;
          const int ic = (kx + ni / 2) % ni, kc = (e + nk / 2) % nk;
          const size_t o = (size_t)ic + (size_t)ni * ((size_t)j + (size_t)nj * kc);
          pr[o] = v.x / nrm;
          pi[o] = v.y / nrm;
        });
}

// srcs/solp_fft_tdma3.f:25-44: dp(1:ni,j,1:nk) = Re(unnormalised inverse DFT).
void fft_inverse(const Dims d, const double* pr, const double* pi, double* dp) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  check_pow2(ni, "solp_fft_tdma3_");
  check_pow2(nk, "solp_fft_tdma3_");
  double2* T = static_cast<double2*>(scratch(4, sizeof(double2) * (size_t)ni * nj * nk));
  const int Tw = z_tile(nk, ni);
  const int tiles_per_j = ni / Tw;
  fft_z(nk, nj * tiles_per_j, Tw, +1.0,
        [=] __device__(int tile, int l, int e) {
          const int j = tile % nj, kx = (tile / nj) * Tw + l;
          const int ic = (kx + ni / 2) % ni, kc = (e + nk / 2) % nk;
          const size_t o = (size_t)ic + (size_t)ni * ((size_t)j + (size_t)nj * kc);
          return make_double2(pr[o], pi[o]);
        },
        [=] __device__(int tile, int l, int e, double2 v) {
          const int j = tile % nj, kx = (tile / nj) * Tw + l;
          T[((size_t)e * nj + j) * ni + kx] = v;
        });
  fft_x(ni, nj * nk, +1.0,
        [=] __device__(int line, int e) { return T[(size_t)line * ni + e]; },
        [=] __device__(int line, int e, double2 v) {
          const int j = line % nj, k = line / nj;
          dp[d.ix(e + 1, j + 1, k + 1)] = v.x;
        });
}

// srcs/solp_fft_tdma2.f:46-87: Thomas solve along j for every (kx,kz) mode
// of the centred layout (i fastest, j, k).  c1x/c1z tables come from the
// host libm; a(j) is kept in device scratch for the back substitution.
void tdma_modes(int ni, int nj, int nk, const double* cx, const double* cz, const double* coef,
                double dx2r, double dz2r, double* pr, double* pi) {
  double* at = static_cast<double*>(scratch(5, sizeof(double) * (size_t)ni * nj * nk));
  const double *as = coef, *an = coef + (nj + 6), *ap0 = coef + 2 * (nj + 6);
  launch_flat(ni * nk, [=] __device__(int m) {
    const int i = m % ni, k = m / ni;
    const double c1x = cx[i], c1z = cz[k];
    double a_prev = 0.0, br = 0.0, bi = 0.0;
    const size_t base = (size_t)i + (size_t)ni * nj * k;
    for (int j = 1; j <= nj; j++) {
      const size_t o = base + (size_t)ni * (j - 1);
      const double ap = ap0[j + 2] + dx2r * (2.0 - c1x) + dz2r * (2.0 - c1z);
      const double d00 = -ap + as[j + 2] * a_prev;
      const double apinv = -d00 / (d00 * d00 + 1.0e-99);
      a_prev = apinv * an[j + 2];
      br = apinv * (-pr[o] + as[j + 2] * br);
      bi = apinv * (-pi[o] + as[j + 2] * bi);
      at[o] = a_prev;
      pr[o] = br;
      pi[o] = bi;
    }
    double xr = br, xi = bi;
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
// through dsum (device, sum of squares over the interior).
void divergence(const Dims d, double dxinv, double dyinv, double dzinv, double dtinv,
                const double* u, const double* v, const double* w, double* div, double* dsum) {
  launch_box(Box{-1, d.ni + 3, -1, d.nj + 3, -1, d.nk + 3}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    div[c] = ((-u[c - X] + u[c]) * dxinv + (-v[c - Y] + v[c]) * dyinv + (-w[c - Z] + w[c]) * dzinv) * dtinv;
  });
  interior_reduce<1>(
      d,
      [=] __device__(int i, int j, int k, double* o) {
        const double x = div[d.ix(i, j, k)];
        o[0] = x * x;
      },
      [=] __device__(const double* t, double* out) { out[0] = t[0]; }, dsum);
}

// srcs/bnd_avzero.f:10-23 with srcs/cal_av.f:20-37: the mean stays on
// the device.
void avzero(const Dims d, double* q, int ndiv) {
  double* s = dev_scalars(2);
  const double nn = (double)(d.ni * d.nj * d.nk * ndiv);
  interior_reduce<1>(
      d, [=] __device__(int i, int j, int k, double* o) { o[0] = q[d.ix(i, j, k)]; },
      [=] __device__(const double* t, double* out) { out[0] = t[0] / nn; }, s);
  const double* av = s;
  launch_flat(d.n3, [=] __device__(int m) { q[m] = q[m] - *av; });
}

// srcs/corunp_explicit.f:26-74
void correct(const Dims d, int jen, double rho, double dxinv, double dyinv, double dzinv, double dt,
             const double* dp, const double* phat, double* un, double* vn, double* wn, double* pn) {
  launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
    const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
    const double p0 = dp[c];
    un[c] = un[c] + (p0 - dp[c + X]) / rho * dxinv * dt;
    if (j <= jen) vn[c] = vn[c] + (p0 - dp[c + Y]) / rho * dyinv * dt;
    wn[c] = wn[c] + (p0 - dp[c + Z]) / rho * dzinv * dt;
    pn[c] = phat[c] + p0;
  });
}

// Sums for srcs/solp_fft_tdma4.f:28-72: [sum div^2, sum (L dp/rho - div)^2].
void poisson_residual_sums(const Dims d, double rho, double dxinv, double dyinv, double dzinv,
                           const double* div, const double* dp, double* dsum) {
  const double aw = dxinv * dxinv / rho, ae = dxinv * dxinv / rho;
  const double as = dyinv * dyinv / rho, an = dyinv * dyinv / rho;
  const double ab = dzinv * dzinv / rho, at = dzinv * dzinv / rho;
  const double ap = aw + ae + as + an + ab + at;
  interior_reduce<2>(
      d,
      [=] __device__(int i, int j, int k, double* o) {
        const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
        const double dv = div[c];
        const double r = +dp[c - Z] * ab + dp[c - Y] * as + dp[c - X] * aw - dp[c] * ap +
                         dp[c + X] * ae + dp[c + Y] * an + dp[c + Z] * at - dv;
        o[0] = dv * dv;
        o[1] = r * r;
      },
      [=] __device__(const double* t, double* out) {
        out[0] = t[0];
        out[1] = t[1];
      },
      dsum);
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
  Call c("caldiv_");
  const double *u = c.in(uk, d.n3), *v = c.in(vk, d.n3), *w = c.in(wk, d.n3);
  double* dv = c.io(div, d.n3);
  double* s = dev_scalars(3);
  divergence(d, *dxinv, *dyinv, *dzinv, 1.0 / *dt, u, v, w, dv, s);
  const double sum = read_scalars(s, 1)[0];
  c.finish();
  *err_div = std::sqrt(sum / (double)(*ni * *nj * *nk * *ndiv)) + 1.0e-99;
}

// srcs/solp_fft_tdma1.f:4-5.  fftdata (the reference's work array) is not used.
void solp_fft_tdma1_(const int* ni, const int* nj, const int* nk, double* fftdata, const double* div,
                     double* phir_r, double* phir_i) {
  (void)fftdata;
  const Dims d = make_dims(*ni, *nj, *nk);
  const size_t nm = (size_t)d.ni * d.nj * d.nk;
  Call c("solp_fft_tdma1_");
  const double* dv = c.in(div, d.n3);
  double* pr = c.out(phir_r, nm);
  double* pi = c.out(phir_i, nm);
  prof_begin(5);
  fft_forward(d, dv, pr, pi);
  prof_end(5);
  c.finish();
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
  const double* a = c.in(r, n);
  double* b = c.out(w, n);
  FEL_CHECK(hipMemcpyAsync(b, a, sizeof(double) * n, hipMemcpyDeviceToDevice, stream()));
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
  const double pi_ = std::atan(1.0) * 4.0;
  const double pi2_ni = pi_ * 2.0 / (double)ni, pi2_nk = pi_ * 2.0 / (double)nk;
  // mode tables (host libm, as the reference's cos)
  std::vector<double> h((size_t)ni + nk + 3 * (nj + 6));
  for (int i = 0; i < ni; i++) h[i] = std::cos((double)(i - ni / 2) * pi2_ni) * 2.0;
  for (int k = 0; k < nk; k++) h[ni + k] = std::cos((double)(k + 1 + (*ID) * nk - nk / 2 - 1) * pi2_nk) * 2.0;
  for (int j = 0; j < nj + 6; j++) {
    h[ni + nk + j] = as_p_fft[j];
    h[ni + nk + (nj + 6) + j] = an_p_fft[j];
    h[ni + nk + 2 * (nj + 6) + j] = ap_p_fft[j];
  }
  double* tab = static_cast<double*>(scratch(6, sizeof(double) * h.size()));
  FEL_CHECK(hipMemcpyAsync(tab, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, stream()));
  Call c("solp_fft_tdma2_");
  double* pr = c.io(phiw_r, nm);
  double* pi = c.io(phiw_i, nm);
  prof_begin(6);
  tdma_modes(ni, nj, nk, tab, tab + ni, tab + ni + nk, (*dxinv * *dxinv) / *rho,
             (*dzinv * *dzinv) / *rho, pr, pi);
  prof_end(6);
  c.finish();
  FEL_CHECK(hipStreamSynchronize(stream()));  // h is released on return
}

// srcs/solp_fft_tdma3.f:4-5
void solp_fft_tdma3_(const int* ni, const int* nj, const int* nk, double* fftdata,
                     const double* phir_r, const double* phir_i, double* dp) {
  (void)fftdata;
  const Dims d = make_dims(*ni, *nj, *nk);
  const size_t nm = (size_t)d.ni * d.nj * d.nk;
  Call c("solp_fft_tdma3_");
  const double* pr = c.in(phir_r, nm);
  const double* pi = c.in(phir_i, nm);
  double* ddp = c.io(dp, d.n3);
  prof_begin(7);
  fft_inverse(d, pr, pi, ddp);
  prof_end(7);
  c.finish();
}

// srcs/solp_fft_tdma4.f:4-5: residual monitor every imon_t steps.
void solp_fft_tdma4_(const int* ipara, const int* ID, const int* ndiv, const int* ni, const int* nj,
                     const int* nk, const int* nstep, const int* imon_t, const double* rho,
                     const double* dxinv, const double* dyinv, const double* dzinv, const double* div,
                     const double* dp) {
  (void)ipara;
  if (!(*imon_t > 0 && (*nstep % *imon_t) == 0)) return;
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("solp_fft_tdma4_");
  const double* dv = c.in(div, d.n3);
  const double* dd = c.in(dp, d.n3);
  double* s = dev_scalars(4);
  poisson_residual_sums(d, *rho, *dxinv, *dyinv, *dzinv, dv, dd, s);
  const double* r = read_scalars(s, 2);
  c.finish();
  const double nn = (double)(*ni * *nj * *nk * *ndiv);
  const double e0 = std::sqrt(r[0] / nn) + 1.0e-99, e = std::sqrt(r[1] / nn) + 1.0e-99;
  if (*ID == 0) std::printf("Err_solp %10d%20.10E%20.10E%20.10E\n", *nstep, e / e0, e, e0);
}

// Relative Poisson residual (the monitor's value), for tests and drivers.
double felgpu_poisson_residual_(const int* ni, const int* nj, const int* nk, const double* rho,
                                const double* dxinv, const double* dyinv, const double* dzinv,
                                const double* div, const double* dp) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("felgpu_poisson_residual_");
  const double* dv = c.in(div, d.n3);
  const double* dd = c.in(dp, d.n3);
  double* s = dev_scalars(4);
  poisson_residual_sums(d, *rho, *dxinv, *dyinv, *dzinv, dv, dd, s);
  const double* r = read_scalars(s, 2);
  c.finish();
  const double nn = (double)(*ni * *nj * *nk);
  return (std::sqrt(r[1] / nn) + 1.0e-99) / (std::sqrt(r[0] / nn) + 1.0e-99);
}

// srcs/bnd_avzero.f:1 (ipara must be 0)
void bnd_avzero_(const int* ipara, const int* ndiv, const int* ni, const int* nj, const int* nk,
                 double* q) {
  (void)ipara;
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("bnd_avzero_");
  avzero(d, c.io(q, d.n3), *ndiv);
  c.finish();
}

// srcs/cal_av.f:1
void cal_av_(const int* ipara, const int* ndiv, const int* ni, const int* nj, const int* nk,
             const double* q, double* av) {
  (void)ipara;
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("cal_av_");
  const double* dq = c.in(q, d.n3);
  double* s = dev_scalars(5);
  interior_reduce<1>(
      d, [=] __device__(int i, int j, int k, double* o) { o[0] = dq[d.ix(i, j, k)]; },
      [=] __device__(const double* t, double* out) { out[0] = t[0]; }, s);
  const double sum = read_scalars(s, 1)[0];
  c.finish();
  *av = sum / (double)(*ni * *nj * *nk * *ndiv);
}

// srcs/corunp_explicit.f:1-2
void corunp_explicit_(const int* nID, const int* ni, const int* nj, const int* nk, const double* rho,
                      const double* dxinv, const double* dyinv, const double* dzinv, const double* dt,
                      const double* dp, const double* phat, double* un, double* vn, double* wn,
                      double* pn) {
  const Dims d = make_dims(*ni, *nj, *nk);
  Call c("corunp_explicit_");
  const double *ddp = c.in(dp, d.n3), *ph = c.in(phat, d.n3);
  double *u = c.io(un, d.n3), *v = c.io(vn, d.n3), *w = c.io(wn, d.n3), *p = c.io(pn, d.n3);
  const int jen = (nID[kYPlus] < 0) ? *nj - 1 : *nj;
  correct(d, jen, *rho, *dxinv, *dyinv, *dzinv, *dt, ddp, ph, u, v, w, p);
  c.finish();
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

}  // extern "C"
