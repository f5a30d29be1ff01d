// k_sor.hip -- semi-implicit (Crank-Nicolson) viscous solve.
//
// Reference: srcs/solu_sor4.f:1-308.  Block Gauss-Seidel u -> v -> w; each
// component is swept over 8 parity colours in the order
// (111,222), (112,221), (121,212), (211,122), omega = 1, with the ghost
// refresh bndu + bnd_periodic(component) after every second colour.  The two
// colours of each pair differ in the parity of all three indices, so no cell
// of one is a face neighbour of the other and updating both in one launch
// gives the reference's values exactly; the refresh points are unchanged.
// The stop test (err/err0 <= 1e-3, at most 21 sweeps) is evaluated on the
// host from a deterministic device sum of du^2 per sweep.
//
// Coefficients: when cal_arith_coef_vis_ deferred them (device-resident
// arguments), each update recomputes its 15 coefficients from the viscosity
// with the same expressions (felgpu_visc.hpp), reading one array instead of
// fifteen; otherwise the 45 arrays are read.
#include <hip/hip_ext.h>

#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "felgpu.hpp"
#include "felgpu_bnd.hpp"
#include "felgpu_visc.hpp"

namespace fel {

// T: element type of the velocities, densities, viscosity and sources
// (double, float in the FP32 mode, Dual in the tangent mode).  The residual
// of the stop test is summed from the values.
template <class T>
struct SorArgs {
  const T *rho, *rhon, *src[3], *rmu;
  // face densities (rho+rho'+rhon+rhon')/4 of u, v, w, computed once per
  // solve and stored pair by pair in the launch's thread order (r00_index),
  // so a launch reads its N/4 values densely; the launch's pointer is
  // offset to its pair
  const T* r00[3];
  // the sources stored the same way when cal_srcu_ made them (else null;
  // the launch's pointer is offset to its pair)
  const T* srcd[3];
  const T* cf[45];
  T *un, *vn, *wn;
  T dtinv;
  VisScales<T> sc;
  // ghost refresh of the swept component inside the pair launch (see
  // write_images): slip-wall data and which walls exist
  const PT<T> *l1, *l2, *phi;
  double uwall, dy;
  bool lo, hi;
  // slip-wall coefficients c1, c2 of wall_coef at every (i,k) of the
  // bottom (j = 1) and top (j = nj) walls: [c1 lo | c2 lo | c1 hi | c2 hi]
  const PT<T>* wcoef;
  int wplane;  // (ni+6)(nk+6)
};

// The refresh after a colour pair (bndu + bnd_periodic of the swept
// component, srcs/solu_sor4.f:176-180) rewrites the ghost cells whose
// source -- the periodic image, or the first interior row of a slip wall --
// was updated by the pair; every other ghost keeps its value.  The updating
// thread can write those ghosts itself: a ghost of an updated cell is never
// read by another cell of the same pair (the face neighbour of the ghost
// (0,j,k) is (1,j,k), whose parity differs from the source (ni,j,k) in i
// only when ni is even, so the two are never in one colour pair; likewise
// in k; a wall ghost's only interior neighbour is its source).
// Images of source (i,j,k) of component q with new value f written here: x
// ghosts i-ni (i >= ni-2) and i+ni (i <= 3), z likewise, and for u, w the
// first wall ghost row (j = 0 from row 1, j = nj+1 from row nj) of bndu's
// slip formula.  The deeper wall rows (u, w: j = -1, -2 and nj+2, nj+3; v:
// the mirrored rows) are read by no stencil of the solve; they are written
// once after the last pair, with the refreshes that leave every ghost as the
// per-pair refreshes do (each is a function of the final interiors).

// Stores v at the cell c0 (when `self`) and at its periodic images: x
// (c0 - ni when xa, c0 + ni when xb), z (c0 -+ nk*sx*sy when za / zb) and
// the xz corners.  Straight-line conditional stores keep the hot kernel's
// register count unchanged.
template <class T>
__device__ __forceinline__ void put_row(T* F, int c0, int ni, int nkz, bool xa, bool xb, bool za,
                                        bool zb, bool self, T v) {
  if (self) F[c0] = v;
  if (xa) F[c0 - ni] = v;
  if (xb) F[c0 + ni] = v;
  if (za) {
    F[c0 - nkz] = v;
    if (xa) F[c0 - nkz - ni] = v;
    if (xb) F[c0 - nkz + ni] = v;
  }
  if (zb) {
    F[c0 + nkz] = v;
    if (xa) F[c0 + nkz - ni] = v;
    if (xb) F[c0 + nkz + ni] = v;
  }
}

template <int COMP, class T>
__device__ __forceinline__ void write_images(const Dims& d, const SorArgs<T>& a, T* F, int i, int j,
                                             int k, T f) {
  using P = PT<T>;
  const int ni = d.ni, nj = d.nj, nk = d.nk, Y = d.sx, nkz = nk * d.sx * d.sy;
  const bool xa = i >= ni - 2, xb = i <= 3, za = k >= nk - 2, zb = k <= 3;
  const bool wall_src = COMP != 1 && ((a.lo && j == 1) || (a.hi && j == nj));
  if (!(xa || xb || za || zb || wall_src)) return;
  const int c0 = d.ix(i, j, k);
  put_row(F, c0, ni, nkz, xa, xb, za, zb, false, f);
  if (COMP == 1 || !wall_src) return;
  // g0 = c1 (-+uwall) + c2 f with the wall coefficients precomputed for the
  // solve (wall_coef of this cell)
  const bool low = j == 1;
  const int wc = (i + 2) + d.sx * (k + 2) + (low ? 0 : 2 * a.wplane);
  const P c1 = a.wcoef[wc], c2 = a.wcoef[wc + a.wplane];
  const T g0 = T(COMP == 0 ? c1 * (low ? -a.uwall : a.uwall) + c2 * P(f) : c2 * P(f));
  put_row(F, c0 + (low ? -Y : Y), ni, nkz, xa, xb, za, zb, true, g0);
}

// The updated value's store: non-temporal for double and float (the next
// launch reads the array from memory anyway; measured 0.9% faster per
// colour-pair launch at 512x256x128 in FP64), plain for Dual.
template <class T>
__device__ __forceinline__ void put_update(T* F, int c, T nv) {
  if constexpr (std::is_same_v<T, double> || std::is_same_v<T, float>)
    __builtin_nontemporal_store(nv, F + c);
  else
    F[c] = nv;
}

// Colour-pair base parities (srcs/solu_sor4.f:94-120); the partner is the
// complement in every index.
static const int kIs[4] = {1, 1, 1, 2}, kJs[4] = {1, 1, 2, 1}, kKs[4] = {1, 2, 1, 1};

// Rows of one parity per block (8 measured no faster: the neighbouring
// rows a block re-reads at its edges come from L2).
static int sor_rows() { return 4; }

// k planes per block: one block sums du^2 over kKB planes, which keeps the
// number of partial sums (and the final reduction) small.
constexpr int kKB = 1;

// R00I: the face density of each update is formed in the launch from rho and
// rhon with the expression of the per-solve table (the same bits), for the
// allocation-free entry point, which keeps no table.
// One cell's update of component COMP (srcs/solu_sor4.f:131-175) at c, with
// the swept component's x neighbours xm = F[c-X], xp = F[c+X] supplied by
// the caller; returns the new value (not stored) and adds du^2 to e.
template <int COMP, bool ONFLY, class T, bool R00I>
__device__ __forceinline__ T sor_cell(const Dims& d, const SorArgs<T>& a, int c, int ridx, T xm,
                                      T xp, double& e) {
  const int X = 1, Y = d.sx, Z = d.sx * d.sy;
  T cf[15];
  if (ONFLY) {
    if (COMP == 0) coef_u(d, a.sc, a.rmu, c, cf);
    if (COMP == 1) coef_v(d, a.sc, a.rmu, c, cf);
    if (COMP == 2) coef_w(d, a.sc, a.rmu, c, cf);
  } else {
#pragma unroll
    for (int m = 0; m < 15; m++) cf[m] = a.cf[15 * COMP + m][c];
  }
  const T *un = a.un, *vn = a.vn, *wn = a.wn;
  const T h = T(0.5);
  using A = acc_t<T>;  // FP64 for FP32 storage (rounded once, at the store), else T
  T r00;
  if (R00I) {
    const int F = COMP == 0 ? X : (COMP == 1 ? Y : Z);
    const T q = T(0.25);
    r00 = (a.rho[c] + a.rho[c + F] + a.rhon[c] + a.rhon[c + F]) * q;
  } else {
    r00 = a.r00[COMP][ridx];
  }
  const T srcv = a.srcd[COMP] ? a.srcd[COMP][ridx] : a.src[COMP][c];
  const T* F = COMP == 0 ? un : (COMP == 1 ? vn : wn);
  A visn;
  if (COMP == 0) {
    visn = (+A(cf[4]) * A(un[c - Z]) + A(cf[2]) * A(un[c - Y]) + A(cf[0]) * A(xm) +
                 A(cf[1]) * A(xp) + A(cf[3]) * A(un[c + Y]) + A(cf[5]) * A(un[c + Z]) +
                 A(cf[7]) * A(vn[c - Y]) + A(cf[8]) * A(vn[c + X - Y]) + A(cf[9]) * A(vn[c]) +
                 A(cf[10]) * A(vn[c + X]) + A(cf[11]) * A(wn[c - Z]) + A(cf[12]) * A(wn[c + X - Z]) +
                 A(cf[13]) * A(wn[c]) + A(cf[14]) * A(wn[c + X])) /
                A(r00);
  } else if (COMP == 1) {
    visn = (+A(cf[4]) * A(vn[c - Z]) + A(cf[2]) * A(vn[c - Y]) + A(cf[0]) * A(xm) +
                 A(cf[1]) * A(xp) + A(cf[3]) * A(vn[c + Y]) + A(cf[5]) * A(vn[c + Z]) +
                 A(cf[7]) * A(un[c - X]) + A(cf[9]) * A(un[c]) + A(cf[8]) * A(un[c - X + Y]) +
                 A(cf[10]) * A(un[c + Y]) + A(cf[11]) * A(wn[c - Z]) + A(cf[12]) * A(wn[c + Y - Z]) +
                 A(cf[13]) * A(wn[c]) + A(cf[14]) * A(wn[c + Y])) /
                A(r00);
  } else {
    visn = (+A(cf[4]) * A(wn[c - Z]) + A(cf[2]) * A(wn[c - Y]) + A(cf[0]) * A(xm) +
                 A(cf[1]) * A(xp) + A(cf[3]) * A(wn[c + Y]) + A(cf[5]) * A(wn[c + Z]) +
                 A(cf[7]) * A(un[c - X]) + A(cf[9]) * A(un[c]) + A(cf[8]) * A(un[c - X + Z]) +
                 A(cf[10]) * A(un[c + Z]) + A(cf[11]) * A(vn[c - Y]) + A(cf[13]) * A(vn[c]) +
                 A(cf[12]) * A(vn[c - Y + Z]) + A(cf[14]) * A(vn[c + Z])) /
                A(r00);
  }
  const A f = (visn * A(h) + A(srcv)) / (A(a.dtinv) + A(cf[6]) / A(r00) * A(h));
  T du, nv;
  if constexpr (is_f32<T>) {
    nv = T(f);
    du = nv - F[c];
  } else {
    du = T(f) - F[c];
    nv = F[c] + du;
  }
  e += val(du) * val(du);
  return nv;
}

template <int COMP, bool ONFLY, bool FUSE, class T, bool R00I = false>
__device__ __forceinline__ void sor_pair_body(const Dims& d, const SorArgs<T>& a, int is, int js,
                                              int ks, int jend, int gx, int gy, int ntiles,
                                              double* part);


template <int COMP, bool ONFLY, bool FUSE, class T, bool R00I = false>
__global__ void __launch_bounds__(512) sor_pair_kernel(Dims d, SorArgs<T> a, int is, int js, int ks,
                                                       int jend, int gx, int gy, int ntiles,
                                                       double* part) {
  sor_pair_body<COMP, ONFLY, FUSE, T, R00I>(d, a, is, js, ks, jend, gx, gy, ntiles, part);
}

template <int COMP, bool ONFLY, bool FUSE, class T, bool R00I>
__device__ __forceinline__ void sor_pair_body(const Dims& d, const SorArgs<T>& a, int is, int js,
                                              int ks, int jend, int gx, int gy, int ntiles,
                                              double* part) {
  const int L = xcd_tile((int)blockIdx.x, ntiles);
  const int bx = L % gx, tt = L / gx, by = tt % gy, bz = tt / gy;
  double e = 0.0;
  for (int kz = 0; kz < kKB; kz++) {
  const int k = 1 + bz * kKB + kz;
  if (k > d.nk) break;
  const int q = (k - ks) & 1;
  const int ii = bx * 64 + (int)threadIdx.x;
  const int i = 1 + ((is + q - 1) & 1) + 2 * ii;
  const int j = 1 + ((js + q - 1) & 1) + 2 * (by * (int)blockDim.y + (int)threadIdx.y);
  if (i <= d.ni && j <= jend) {
    const int X = 1;
    const int c = d.ix(i, j, k);
    const int ridx = r00_index(d, ii, by * (int)blockDim.y + (int)threadIdx.y, k);
    T* F = COMP == 0 ? a.un : (COMP == 1 ? a.vn : a.wn);
    const T nv = sor_cell<COMP, ONFLY, T, R00I>(d, a, c, ridx, F[c - X], F[c + X], e);
    put_update(F, c, nv);
    if (FUSE) write_images<COMP>(d, a, F, i, j, k, nv);
  }
  }
  // block sum of du^2 in a fixed order
  __shared__ double sh[512];
  const int t = (int)(threadIdx.y * 64 + threadIdx.x);
  sh[t] = e;
  __syncthreads();
  for (int s = (int)(blockDim.y * 32); s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  if (t == 0) part[L] = sh[0];
}

// Colour pairs 1 and 2 of a component's sweep in one launch.  Pair 1 =
// (112,221) and pair 2 = (121,212): on plane k both lie on the rows of one j
// parity (k even: odd rows, pair 1 at odd i and pair 2 at even i; k odd:
// even rows, pair 1 at even i and pair 2 at odd i).  A pair-2 cell's face
// neighbours are pair-1 cells only along x (its y and z neighbours belong
// to pairs 0 and 3, which neither pair updates), and a pair-1 cell's x
// neighbours are pair-2 cells.  So a block that owns whole rows (one row,
// thread t: the pair-1 and pair-2 cells of the t-th x parity pair) updates
// its pair-1 cells from the old values, stores them, passes them through
// LDS, and then updates its pair-2 cells with the new pair-1 values as x
// neighbours -- the periodic x ghost of a row end is the other end's cell,
// which the same block owns.  Every operand is the one the two separate
// launches (with the in-launch ghost images between them) would read, so
// the values are identical.  Needs even ni and nk (the images' parity,
// as for the fused images) and ni / 2 <= 512 threads.
template <int COMP, bool ONFLY, class T>
__global__ void __launch_bounds__(512) sor_pair12_kernel(Dims d, SorArgs<T> a1, SorArgs<T> a2,
                                                         int jend, double* part1,
                                                         double* part2, int ntiles, int band) {
  __shared__ T s1_all[512];
  __shared__ double sh[2][512];
  const int hx = (int)blockDim.x, R = (int)blockDim.y, hy = (d.nj + 1) / 2;
  const int gyr = (hy + R - 1) / R;  // blocks per plane
  const int t = (int)threadIdx.x, L = xcd_tile((int)blockIdx.x, ntiles);
  int jj, k, slot = L;
  if (band > 0) {  // bands of `band` rows, planes in descending order within a band
    const int per = band * d.nk;
    const int bnd = L / per, rem = L - bnd * per;
    k = d.nk - rem / band;
    jj = bnd * band + rem % band;
    slot = bnd * per + (k - 1) * band + rem % band;  // the ascending order's slot
  } else {
    jj = (L % gyr) * R + (int)threadIdx.y;
    k = 1 + L / gyr;
  }
  T* s1 = s1_all + threadIdx.y * hx;
  const int q1 = (k - 2) & 1, q2 = (k - 1) & 1;  // pair 1 (1,1,2), pair 2 (1,2,1)
  const int j = 1 + (q1 & 1) + 2 * jj;           // the row of both pairs
  const int i1 = 1 + (q1 & 1) + 2 * t, i2 = 1 + (q2 & 1) + 2 * t;
  const bool row = jj < hy && j <= jend;
  const int X = 1;
  T* F = COMP == 0 ? a1.un : (COMP == 1 ? a1.vn : a1.wn);
  const int ridx = r00_index(d, t, jj, k);
  double e1 = 0.0, e2 = 0.0;
  T nv1 = T(0.0);
  if (row) {
    const int c = d.ix(i1, j, k);
    nv1 = sor_cell<COMP, ONFLY, T, false>(d, a1, c, ridx, F[c - X], F[c + X], e1);
    put_update(F, c, nv1);
    write_images<COMP>(d, a1, F, i1, j, k, nv1);
  }
  s1[t] = nv1;
  __syncthreads();
  if (row) {
    const int c = d.ix(i2, j, k);
    T xm, xp;
    if (i2 > i1) {  // pair-1 neighbours: own (left) and the next thread's (right)
      xm = s1[t];
      xp = s1[t + 1 < hx ? t + 1 : 0];
    } else {        // the previous thread's (left) and own (right)
      xm = s1[t > 0 ? t - 1 : hx - 1];
      xp = s1[t];
    }
    const T nv2 = sor_cell<COMP, ONFLY, T, false>(d, a2, c, ridx, xm, xp, e2);
    put_update(F, c, nv2);
    write_images<COMP>(d, a2, F, i2, j, k, nv2);
  }
  // block sums of du^2 of each pair in a fixed order
  const int tb = (int)threadIdx.y * hx + t, nt = R * hx;
  sh[0][tb] = e1;
  sh[1][tb] = e2;
  __syncthreads();
  int p2 = 1;
  while (p2 < nt) p2 <<= 1;
  for (int st = p2 >> 1; st > 0; st >>= 1) {
    if (tb < st && tb + st < nt) {
      sh[0][tb] += sh[0][tb + st];
      sh[1][tb] += sh[1][tb + st];
    }
    __syncthreads();
  }
  if (tb == 0) {
    part1[slot] = sh[0][0];
    part2[slot] = sh[1][0];
  }
}

struct PartSlots {
  int off[12], n[12];
};

// Two-stage fixed-order sum of the 12 launches' block partials: stage 1
// reduces contiguous chunks (one block each), stage 2 the chunk sums.
constexpr int kErrChunks = 256;

__global__ void __launch_bounds__(256) sor_err_stage1(const double* part, int n, double* chunk) {
  __shared__ double sh[256];
  const int t = (int)threadIdx.x;
  const int per = (n + kErrChunks - 1) / kErrChunks;
  const int b0 = (int)blockIdx.x * per, b1 = min(n, b0 + per);
  double acc = 0.0;
  for (int b = b0 + t; b < b1; b += 256) acc += part[b];
  sh[t] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  if (t == 0) chunk[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) sor_err_stage2(const double* chunk, double* out) {
  __shared__ double sh[256];
  const int t = (int)threadIdx.x;
  sh[t] = chunk[t];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  if (t == 0) out[0] = sh[0];
}

struct SorResult {
  int lp;
  double err, err0;
};

template <int COMP, class T>
void launch_pair(bool onfly, bool fuse, dim3 grid, dim3 block, const Dims& d, const SorArgs<T>& a,
                 int p, int jend, int gx, int gy, int nt, double* pp, hipEvent_t e0,
                 hipEvent_t e1) {
  hipStream_t s = stream();
  const bool r00i = a.r00[COMP] == nullptr;  // no per-solve table: formed in the launch
  SorArgs<T> a_ = a;  // the face densities of this pair
  if (!r00i) a_.r00[COMP] = a.r00[COMP] + (size_t)p * r00_slab(d);
  if (a.srcd[COMP]) a_.srcd[COMP] = a.srcd[COMP] + (size_t)p * r00_slab(d);
  const SorArgs<T>& ap = a_;
  auto go = [&](auto kernel) {
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, e0, e1, 0, d, ap, kIs[p], kJs[p], kKs[p],
                          jend, gx, gy, nt, pp);
  };
  if (r00i) {
    // the allocation-free form: coefficients from the viscosity, FP64 only
    if constexpr (std::is_same_v<T, double>) {
      if (fuse) go(sor_pair_kernel<COMP, true, true, T, true>);
      else go(sor_pair_kernel<COMP, true, false, T, true>);
    }
    return;
  }
  if (onfly && fuse) go(sor_pair_kernel<COMP, true, true, T>);
  else if (onfly) go(sor_pair_kernel<COMP, true, false, T>);
  else if (fuse) go(sor_pair_kernel<COMP, false, true, T>);
  else go(sor_pair_kernel<COMP, false, false, T>);
}

// Whether pairs 1 and 2 run as one launch (sor_pair12_kernel): even ni and
// nk, and a row of ni/2 threads in one block.
inline bool sor_pair12_ok(const Dims& d) {
  return d.ni % 2 == 0 && d.nk % 2 == 0 && d.ni / 2 <= 512 && d.ni >= 2;
}

// Rows per block of the pair-1/2 launch (1; 2 rows of 256 threads measured
// 4% slower at 512x256x128 with the XCD-contiguous block order below).
template <int COMP, class T>
void launch_pair12(bool onfly, const Dims& d, const SorArgs<T>& a, int jend, double* part1,
                   double* part2, hipEvent_t e0, hipEvent_t e1) {
  SorArgs<T> a1 = a, a2 = a;  // the face densities and sources of pairs 1 and 2
  a1.r00[COMP] = a.r00[COMP] + r00_slab(d);
  a2.r00[COMP] = a.r00[COMP] + 2 * r00_slab(d);
  if (a.srcd[COMP]) {
    a1.srcd[COMP] = a.srcd[COMP] + r00_slab(d);
    a2.srcd[COMP] = a.srcd[COMP] + 2 * r00_slab(d);
  }
  // one row per block (two rows of 256 threads measured 4% slower; one
  // 64-lane wave per row with four cells per lane, which needs no block
  // barrier, 17-28% slower)
  const int R = 1, hy = (d.nj + 1) / 2;
  const int nb = hy * d.nk;
  // Block order: bands of 8 rows, each swept through all planes before the
  // next, the bands split over the XCDs in contiguous ranges, so a row's
  // neighbour rows and planes are read by blocks that run together on one
  // XCD (measured at 512x256x128, FP64: 201 us per launch, against 220 with
  // whole planes per XCD range and 228 in dispatch order; bands of 4, 16 and
  // 32 rows 207, 206 and 209 us; the same banding made the single-pair
  // launch slower).  The planes of a band run in descending order, so the
  // launch ends on the low planes where the pair-3 launch that follows
  // starts (same box: pair-3 launch 134 -> 126 us, this launch 202 -> 196
  // us); each block's partial keeps the ascending order's slot, so the
  // residual sum is unchanged.
  const int band = hy % 8 == 0 ? 8 : 0;
  const int nt = nb;
  const dim3 grid(nb), block(d.ni / 2, R, 1);
  if (onfly)
    hipExtLaunchKernelGGL(sor_pair12_kernel<COMP, true, T>, grid, block, 0, stream(), e0, e1, 0, d,
                          a1, a2, jend, part1, part2, nt, band);
  else
    hipExtLaunchKernelGGL(sor_pair12_kernel<COMP, false, T>, grid, block, 0, stream(), e0, e1, 0,
                          d, a1, a2, jend, part1, part2, nt, band);
}

// Launch blocks of one colour-pair launch.
inline int sor_blocks(const Dims& d) {
  const int gx = ((d.ni + 1) / 2 + 63) / 64;
  const int gy = ((d.nj + 1) / 2 + sor_rows() - 1) / sor_rows();
  const int gz = (d.nk + kKB - 1) / kKB;
  return gx * gy * gz;
}

// Entries per launch slot in the block-partial array: the larger of the
// pair launch's and the pair-1/2 launch's block counts (the unused tail of a
// slot stays zero).
inline int sor_slot(const Dims& d) {
  const int n12 = sor_pair12_ok(d) ? ((d.nj + 1) / 2) * d.nk : 0;
  const int nb = sor_blocks(d);
  return n12 > nb ? n12 : nb;
}

// Device work memory of one solve: the block partials of the 12 launches of
// a sweep and their chunk sums (12 nb + kErrChunks doubles), the residual
// sum (1 double) and the slip-wall coefficients (4 (ni+6)(nk+6) elements).
template <class P>
struct SorWork {
  double* part;
  double* derr;
  P* wcoef;
};

template <class P>
size_t sor_work_bytes(const Dims& d) {
  const size_t a = sizeof(double) * ((size_t)sor_slot(d) * 12 + kErrChunks + 1);
  const size_t wb = sizeof(P) * 4 * (size_t)d.sx * d.sz;
  return ((a + 255) / 256) * 256 + wb;
}

// Lays the work memory out in one buffer of sor_work_bytes<P>(d) bytes.
template <class P>
SorWork<P> sor_work_in(const Dims& d, void* buf) {
  const size_t a = sizeof(double) * ((size_t)sor_slot(d) * 12 + kErrChunks + 1);
  SorWork<P> w;
  w.part = static_cast<double*>(buf);
  w.derr = w.part + (size_t)sor_slot(d) * 12 + kErrChunks;
  w.wcoef = reinterpret_cast<P*>(static_cast<char*>(buf) + ((a + 255) / 256) * 256);
  return w;
}

template <class T, class P = PT<T>>
SorResult sor4(const Dims d, SorArgs<T> a, bool onfly, bool lo, bool hi, double uwall, double dy,
               const P* l1, const P* l2, const P* phi, int ndiv, const SorWork<P>& work) {
  const int jen = hi ? d.nj - 1 : d.nj;
  const int gx = ((d.ni + 1) / 2 + 63) / 64;  // cells of one parity per row
  const int rows = sor_rows();                // rows of one parity per block
  const int gy = ((d.nj + 1) / 2 + rows - 1) / rows;
  const int nb = sor_blocks(d);
  const int slotn = sor_slot(d);
  double* part = work.part;
  double* chunk = part + (size_t)slotn * 12;
  PartSlots ps;
  for (int s = 0; s < 12; s++) {
    ps.off[s] = s * slotn;
    ps.n[s] = nb;
  }
  // the slots' unused tails add zero to the stop test's sum
  if (slotn > nb || sor_pair12_ok(d)) {
    const int n = 12 * slotn;
    launch_flat(n, [=] __device__(int m) { part[m] = 0.0; });
  }
  double* derr = work.derr;
  // in-launch ghost refresh needs even ni, nk (colour parity of periodic images)
  const bool fuse_ok = d.ni % 2 == 0 && d.nk % 2 == 0;
  // pairs 1 and 2 in one launch (with the per-solve face-density table)
  const bool p12 = fuse_ok && sor_pair12_ok(d) && a.r00[0] != nullptr;
  a.l1 = l1;
  a.l2 = l2;
  a.phi = phi;
  a.uwall = uwall;
  a.dy = dy;
  a.lo = lo;
  a.hi = hi;
  a.wplane = d.sx * d.sz;
  {
    P* wc = work.wcoef;
    const int wp = a.wplane, nj = d.nj;
    launch_flat(wp, [=] __device__(int m) {
      const int i = m % d.sx - 2, k = m / d.sx - 2;
      if (lo) {
        const WallCoef<P> cf = wall_coef(d, i, 1, k, dy, l1, l2, phi);
        wc[m] = cf.c1;
        wc[m + wp] = cf.c2;
      }
      if (hi) {
        const WallCoef<P> cf = wall_coef(d, i, nj, k, dy, l1, l2, phi);
        wc[m + 2 * wp] = cf.c1;
        wc[m + 3 * wp] = cf.c2;
      }
    });
    a.wcoef = wc;
  }
  SorResult res{0, 0.0, 1.0};
  double err0 = 1.0, err = 0.0;
  int lp;
  for (lp = 0; lp <= 20; lp++) {
    int slot = 0;
    for (int comp = 0; comp < 3; comp++) {
      const int jend = comp == 1 ? jen : d.nj;
      for (int p = 0; p < 4; p++, slot++) {
        dim3 grid(nb), block(64, rows, 1);
        double* pp = part + ps.off[slot];
        if (p12 && p == 2) continue;  // ran with pair 1
        // The benchmark's timers stamp the launches of the first sweep only:
        // each stamped launch leaves a gap of ~5 us after it, and every sweep
        // runs the same launches.
        const bool stamp = lp == 0;
        if (p12 && p == 1) {
          hipEvent_t e0 = nullptr, e1 = nullptr;
          if (stamp) prof_launch_events(4, true, true, 1, &e0, &e1);
          if (comp == 0) launch_pair12<0, T>(onfly, d, a, jend, pp, part + ps.off[slot + 1], e0, e1);
          if (comp == 1) launch_pair12<1, T>(onfly, d, a, jend, pp, part + ps.off[slot + 1], e0, e1);
          if (comp == 2) launch_pair12<2, T>(onfly, d, a, jend, pp, part + ps.off[slot + 1], e0, e1);
          FEL_CHECK(hipGetLastError());
          continue;
        }
        // Pairs after the first write the swept component's ghost images
        // themselves (write_images); the refresh after the first pair also
        // rewrites the other components' ghosts, so it stays a launch.
        const bool fuse = fuse_ok && p > 0;
        const int nt = xcd_arg(1, nb);
        // The colour-pair timer: with pairs 1 and 2 joined, pair 3 (pair 0
        // is followed by a refresh launch and not timed); with fused images
        // only, pairs 1-3 back to back as one span of three launches;
        // without them every pair is a span of its own.
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (!stamp) {
        } else if (p12) {
          if (p == 3) prof_launch_events(3, true, true, 1, &e0, &e1);
        } else if (fuse_ok) {
          if (p > 0) prof_launch_events(3, p == 1, p == 3, 3, &e0, &e1);
        } else {
          prof_launch_events(3, true, true, 1, &e0, &e1);
        }
        if (comp == 0) launch_pair<0, T>(onfly, fuse, grid, block, d, a, p, jend, gx, gy, nt, pp, e0, e1);
        if (comp == 1) launch_pair<1, T>(onfly, fuse, grid, block, d, a, p, jend, gx, gy, nt, pp, e0, e1);
        if (comp == 2) launch_pair<2, T>(onfly, fuse, grid, block, d, a, p, jend, gx, gy, nt, pp, e0, e1);
        FEL_CHECK(hipGetLastError());
        if (!fuse) {
          // After the first pair of a component's sweep only that component
          // changes, and bndu's ghosts of the other two are functions of their
          // unchanged interiors: rewriting them would store the same values.
          // With fused images the deep wall rows, read by no stencil of the
          // solve, are left to the refreshes after the last sweep.
          bc_fused<T, P>(d, a.un, a.vn, a.wn, uwall, dy, l1, l2, phi, lo, hi, comp,
                         p == 0 ? 7 : (1 << comp), fuse_ok);
        }
      }
    }
    sor_err_stage1<<<kErrChunks, 256, 0, stream()>>>(part, 12 * slotn, chunk);
    sor_err_stage2<<<1, 256, 0, stream()>>>(chunk, derr);
    FEL_CHECK(hipGetLastError());
    const double sum = read_scalars(derr, 1)[0];
    err = std::sqrt(sum / (double)(d.ni * d.nj * d.nk * ndiv)) + 1.0e-99;
    if (lp == 0) err0 = err;
    if (err / err0 <= 1.0e-3) break;
  }
  if (fuse_ok) {
    // The deeper wall rows the fused pairs leave behind (write_images), in
    // the state the per-pair refreshes leave: each component's ghosts in the
    // swept form of its last pair, then bndu's rows of u and v as rewritten
    // by the refreshes of the later components' pairs.
    bc_fused<T, P>(d, a.un, a.vn, a.wn, uwall, dy, l1, l2, phi, lo, hi, 0, 1);
    bc_fused<T, P>(d, a.un, a.vn, a.wn, uwall, dy, l1, l2, phi, lo, hi, 1, 2);
    bc_fused<T, P>(d, a.un, a.vn, a.wn, uwall, dy, l1, l2, phi, lo, hi, 2, 7);
  }
  // lp is the loop variable of srcs/solu_sor4.f:131 on exit: the sweep
  // index that met the tolerance, or 21 after 21 sweeps without it
  res.lp = lp;
  res.err = err;
  res.err0 = err0;
  return res;
}

static int g_last_lp = -1;
static double g_last_ratio = 0.0;
// solu_sor4_ calls and colour sweeps since the last felgpu_sor_stats_ reset
static long long g_sor_calls = 0, g_sor_sweeps = 0;

template <class T>
SorResult solu_sor4_t(const Dims d, double dt, const double* rho, const double* rhon,
                      double* const* h, const double* src_u, const double* src_v,
                      const double* src_w, double* un, double* vn, double* wn, double uwall,
                      double dy, const double* l1, const double* l2, const double* phi, int jm,
                      int jp, int ndiv) {
  // on-the-fly coefficients when they are a pending function of the viscosity
  DeferredCoef& dc = deferred_coef();
  bool onfly = dc.active && dc.d.ni == d.ni && dc.d.nj == d.nj && dc.d.nk == d.nk &&
               dc.esize == (int)sizeof(T);
  for (int m = 0; m < 45 && onfly; m++) onfly = dc.out_host[m] == (const void*)h[m];
  Call c("solu_sor4_");
  SorArgs<T> a;
  a.rho = c.in_t<T>(rho, d.n3, "rho");
  a.rhon = c.in_t<T>(rhon, d.n3, "rhon");
  if (onfly) {
    a.rmu = static_cast<const T*>(dc.src_dev);
    a.sc = vis_scales<T>(dc.dxinv, dc.dyinv, dc.dzinv);
    for (int m = 0; m < 45; m++) a.cf[m] = nullptr;
  } else {
    a.rmu = nullptr;
    a.sc = vis_scales<T>(0.0, 0.0, 0.0);
    for (int m = 0; m < 45; m++) a.cf[m] = c.in_t<T>(h[m], d.n3, "coefficient");
  }
  a.src[0] = c.in_t<T>(src_u, d.n3, "src_u");
  a.src[1] = c.in_t<T>(src_v, d.n3, "src_v");
  a.src[2] = c.in_t<T>(src_w, d.n3, "src_w");
  {
    // the pair-by-pair sources of cal_srcu_, if they are still these arrays'
    const DenseSrc& ds = dense_src();
    const void* sh[3] = {src_u, src_v, src_w};
    bool ok = ds.valid && ds.ni == d.ni && ds.nj == d.nj && ds.nk == d.nk &&
              ds.esize == (int)sizeof(T);
    for (int m = 0; m < 3 && ok; m++)
      ok = ds.host[m] == sh[m] && ds.key[m] == content_key(sh[m], d.n3);
    for (int m = 0; m < 3; m++)
      a.srcd[m] = ok ? static_cast<const T*>(ds.buf) + (size_t)4 * m * r00_slab(d) : nullptr;
  }
  a.un = c.io_t<T>(un, d.n3, "un");
  a.vn = c.io_t<T>(vn, d.n3, "vn");
  a.wn = c.io_t<T>(wn, d.n3, "wn");
  a.dtinv = T(1.0 / dt);
  // srcs/solu_sor4.f:150-152 (and the v, w forms): the face density of each
  // update, the same expression as the sweeps evaluate, once for the solve
  // -- or the table cal_srcu_ formed from these very densities
  {
    T* r00[3];
    const size_t slab = r00_slab(d);
    for (int m = 0; m < 3; m++) {
      r00[m] = static_cast<T*>(scratch(10 + m, sizeof(T) * 4 * slab));
      a.r00[m] = r00[m];
    }
    const DenseSrc& ds = dense_src();
    const bool have = ds.valid && ds.r00 && ds.ni == d.ni && ds.nj == d.nj && ds.nk == d.nk &&
                      ds.esize == (int)sizeof(T) && ds.rho_host == (const void*)rho &&
                      ds.rhon_host == (const void*)rhon &&
                      ds.rho_key == content_key(rho, d.n3) &&
                      ds.rhon_key == content_key(rhon, d.n3);
    const T *rho_d = a.rho, *rhon_d = a.rhon;
    if (!have) launch_box(Box{1, d.ni, 1, d.nj, 1, d.nk}, [=] __device__(int i, int j, int k) {
      const int c = d.ix(i, j, k), X = 1, Y = d.sx, Z = d.sx * d.sy;
      const size_t o = pair_of(i, j, k) * slab + r00_index(d, (i - 1) >> 1, (j - 1) >> 1, k);
      const T q = T(0.25);
      r00[0][o] = (rho_d[c] + rho_d[c + X] + rhon_d[c] + rhon_d[c + X]) * q;
      r00[1][o] = (rho_d[c] + rho_d[c + Y] + rhon_d[c] + rhon_d[c + Y]) * q;
      r00[2][o] = (rho_d[c] + rho_d[c + Z] + rhon_d[c] + rhon_d[c + Z]) * q;
    });
  }
  using P = PT<T>;
  const P* dl1 = c.in_t<P>(l1, d.n3, "l1");
  const P* dl2 = c.in_t<P>(l2, d.n3, "l2");
  const P* dphi = c.in_t<P>(phi, d.n3, "phi");
  const SorWork<P> work = sor_work_in<P>(d, scratch(3, sor_work_bytes<P>(d)));
  const SorResult r = sor4<T>(d, a, onfly, jm < 0, jp < 0, uwall, dy, dl1, dl2, dphi, ndiv, work);
  c.finish();
  return r;
}

}  // namespace fel

// ---------------------------------------------------------------------------
// Allocation-free solve (felgpu_solu_sor4_noalloc_): its work memory -- the
// residual partials, the residual sum and the slip-wall coefficients -- lives
// in this static device array instead of library-allocated buffers.  Each
// device holds its own copy at its own address, found once per device with
// hipGetSymbolAddress (fel::sor_noalloc_setup, called from the library's
// initialisation) and used through that address only: the array's name in
// host code is not a device address.  Every call on a device uses the whole
// array from its start, so calls that overlap on one device race; nothing in
// it needs to start at zero (every element is written before it is read).
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = size_t(16) << 20;  // 16 MiB per device
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) char scratch[kNoAllocBytes];
static void* g_noalloc_ptr[kNoAllocDevices];

namespace fel {

void sor_noalloc_setup() {
  int dev = 0;
  FEL_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices || g_noalloc_ptr[dev]) return;
  void* ptr = nullptr;
  FEL_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch)));
  g_noalloc_ptr[dev] = ptr;
}

}  // namespace fel

using namespace fel;

extern "C" {

// Allocation-free form of solu_sor4_ (srcs/solu_sor4.f:1-308): the same
// 8-colour Gauss-Seidel, refreshes and stop rule on device arrays, FP64,
// coefficients formed from the viscosity as cal_arith_coef_vis forms them
// and face densities from rho and rhon as solu_sor4 forms them.  Returns 0,
// 1 when the grid needs more work memory than the static array holds, 2 for
// invalid arguments (both before any launch).
int felgpu_solu_sor4_noalloc_(const int* nID, const int* ni, const int* nj, const int* nk,
                              const double* dt, const double* dxinv, const double* dyinv,
                              const double* dzinv, const double* rho, const double* rhon,
                              const double* rmun, const double* src_u, const double* src_v,
                              const double* src_w, double* un, double* vn, double* wn,
                              const double* uwall, const double* dy, const double* l1,
                              const double* l2, const double* phi, int* lp, double* ratio) {
  if (*ni < 4 || *nj < 4 || *nk < 4 ||
      (long long)(*ni + 6) * (*nj + 6) * (*nk + 6) * 6 >= (1LL << 31)) {
    std::fprintf(stderr, "felgpu: felgpu_solu_sor4_noalloc_: invalid grid %dx%dx%d\n", *ni, *nj, *nk);
    return 2;
  }
  // the size checks are host arithmetic only: a rejected call touches no device
  const Dims d = make_dims(*ni, *nj, *nk);
  const size_t need = sor_work_bytes<double>(d);
  if (need > kNoAllocBytes) {
    std::fprintf(stderr,
                 "felgpu: felgpu_solu_sor4_noalloc_: grid %dx%dx%d needs %zu B of work memory, "
                 "more than the %zu B static array\n", *ni, *nj, *nk, need, kNoAllocBytes);
    return 1;
  }
  if (flow_fp32() || tangent_mode()) {
    std::fprintf(stderr, "felgpu: felgpu_solu_sor4_noalloc_: FP64 mode only\n");
    return 2;
  }
  sor_noalloc_setup();  // no-op after the library's initialisation on this device
  int dev = 0;
  FEL_CHECK(hipGetDevice(&dev));
  void* buf = dev >= 0 && dev < kNoAllocDevices ? g_noalloc_ptr[dev] : nullptr;
  if (!buf) {
    std::fprintf(stderr, "felgpu: felgpu_solu_sor4_noalloc_: device %d has no work array\n", dev);
    return 2;
  }
  SorArgs<double> a;
  a.rho = rho;
  a.rhon = rhon;
  a.rmu = rmun;
  a.sc = vis_scales<double>(*dxinv, *dyinv, *dzinv);
  for (int m = 0; m < 45; m++) a.cf[m] = nullptr;
  a.src[0] = src_u;
  a.src[1] = src_v;
  a.src[2] = src_w;
  for (int m = 0; m < 3; m++) {
    a.r00[m] = nullptr;
    a.srcd[m] = nullptr;
  }
  a.un = un;
  a.vn = vn;
  a.wn = wn;
  a.dtinv = 1.0 / *dt;
  const int jm = nID[kYMinus], jp = nID[kYPlus];
  const SorResult r = sor4<double>(d, a, true, jm < 0, jp < 0, *uwall, *dy, l1, l2, phi, 1,
                                   sor_work_in<double>(d, buf));
  *lp = r.lp;
  *ratio = r.err / r.err0;
  return 0;
}

// Iteration count and err/err0 of the last solu_sor4_ call (the reference
// prints them every imon_t steps, srcs/solu_sor4.f:298-305).
void felgpu_sor_last_(int* lp, double* ratio) {
  *lp = g_last_lp;
  *ratio = g_last_ratio;
}

// Number of solu_sor4_ calls and of colour sweeps they made since the last
// reset (*reset != 0 clears the counts after reading them).
void felgpu_sor_stats_(const int* reset, long long* calls, long long* sweeps) {
  *calls = g_sor_calls;
  *sweeps = g_sor_sweeps;
  if (*reset) g_sor_calls = g_sor_sweeps = 0;
}

// srcs/solu_sor4.f:1-17 (ipara must be 0)
void solu_sor4_(const int* ipara, const int* ID, const int* nID, const int* ndiv, const int* ni,
                const int* nj, const int* nk, const int* key, double* sendjb, double* recvjb,
                const int* nstep, const int* imon_t, const double* dt, const double* rho,
                const double* rhon, double* c0, double* c1, double* c2, double* c3, double* c4,
                double* c5, double* c6, double* c7, double* c8, double* c9, double* c10,
                double* c11, double* c12, double* c13, double* c14, double* c15, double* c16,
                double* c17, double* c18, double* c19, double* c20, double* c21, double* c22,
                double* c23, double* c24, double* c25, double* c26, double* c27, double* c28,
                double* c29, double* c30, double* c31, double* c32, double* c33, double* c34,
                double* c35, double* c36, double* c37, double* c38, double* c39, double* c40,
                double* c41, double* c42, double* c43, double* c44, const double* src_u,
                const double* src_v, const double* src_w, double* un, double* vn, double* wn,
                const double* uwall, const double* dy, const double* l1, const double* l2,
                const double* phi) {
  (void)key; (void)sendjb; (void)recvjb;
  if (*ipara != 0) {
    std::fprintf(stderr, "felgpu: solu_sor4_: ipara=%d, only the single-domain case (0) is supported\n", *ipara);
    std::abort();
  }
  const Dims d = make_dims(*ni, *nj, *nk);
  double* h[45] = {c0,  c1,  c2,  c3,  c4,  c5,  c6,  c7,  c8,  c9,  c10, c11, c12, c13, c14,
                   c15, c16, c17, c18, c19, c20, c21, c22, c23, c24, c25, c26, c27, c28, c29,
                   c30, c31, c32, c33, c34, c35, c36, c37, c38, c39, c40, c41, c42, c43, c44};
  const int jm = nID[kYMinus], jp = nID[kYPlus];
  SorResult r{0, 0.0, 1.0};
  with_flow([&](auto t) {
    r = solu_sor4_t<decltype(t)>(d, *dt, rho, rhon, h, src_u, src_v, src_w, un, vn, wn, *uwall, *dy,
                                 l1, l2, phi, jm, jp, *ndiv);
  });
  g_last_lp = r.lp;
  g_last_ratio = r.err / r.err0;
  g_sor_calls += 1;
  g_sor_sweeps += r.lp < 21 ? r.lp + 1 : 21;
  if (*imon_t > 0 && (*nstep % *imon_t) == 0 && *ID == 0)
  {
    char a[3][32];
    std::printf("Err_solu %10d%10d%s%s%s\n", *nstep, r.lp, fortran_e(a[0], r.err / r.err0, 20, 10),
                fortran_e(a[1], r.err, 20, 10), fortran_e(a[2], r.err0, 20, 10));
  }
}

}  // extern "C"
