// felgpu.hpp -- shared definitions of the gfx950 implementation of the
// per-timestep solver of toruneko1129/fel-multi-phase-flow.
//
// Arrays keep the reference's image: (-2:ni+3,-2:nj+3,-2:nk+3), x fastest,
// FP64 (srcs/allocate.h:7-165).  Offsets are 32-bit: the largest grid that
// fits (512x256x512 plus ghosts, 6 components) stays below 2^31 elements and
// make_dims() rejects anything larger.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <type_traits>
#include <vector>

// The public C ABI: every extern "C" definition is checked against it.
#include "felgpu.h"
#include "felgpu_dual.hpp"

#define FEL_CHECK(call)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      std::fprintf(stderr, "felgpu: %s failed at %s:%d: %s\n", #call, __FILE__, __LINE__, \
                   hipGetErrorString(e_));                                              \
      std::abort();                                                                     \
    }                                                                                   \
  } while (0)

namespace fel {

// Grid extents and flat offsets of the padded image.
struct Dims {
  int ni, nj, nk;
  int sx, sy, sz;  // padded extents ni+6, nj+6, nk+6
  int n3;          // sx*sy*sz
  __host__ __device__ __forceinline__ int ix(int i, int j, int k) const {
    return (i + 2) + sx * ((j + 2) + sy * (k + 2));
  }
};

Dims make_dims(int ni, int nj, int nk);

// Inclusive index box [i0,i1]x[j0,j1]x[k0,k1].
struct Box {
  int i0, i1, j0, j1, k0, k1;
  __host__ __device__ int nx() const { return i1 - i0 + 1; }
  __host__ __device__ int ny() const { return j1 - j0 + 1; }
  __host__ __device__ int nz() const { return k1 - k0 + 1; }
};

// ---------------------------------------------------------------------------
// Runtime: stream, device mirrors of host arrays, scalar read-back.
// ---------------------------------------------------------------------------
hipStream_t stream();
void ensure_init();

// How an argument array is used by one drop-in call.
// kOverwrite: every element of the argument is written (no element of the
// previous contents survives the call).
// kInterior: the call writes exactly the interior cells [1,n]^3 of a
// padded array and reads none of it.
enum Access { kRead = 1, kWrite = 2, kReadWrite = 3, kOverwrite = 6, kInterior = 18 };

// Precision of the device copies.  In the default FP64 mode every device
// copy is FP64, like the host arrays.  In the FP32 mode (FELGPU_PRECISION=fp32
// or felgpu_set_precision_ before the first registration) arrays registered
// with felgpu_register_ -- the flow fields and their work arrays -- are
// stored and computed in FP32 on the device, while arrays registered with
// felgpu_register_fp64_ -- the volume fraction, its fluxes, the contact-line
// maps, the divergence and the pressure correction -- stay FP64.  Host
// arrays are FP64 in both modes; uploads and downloads convert.
bool flow_fp32();
// Reuse of results already computed (parked results, the result memo): on
// unless FELGPU_NO_RESULT_REUSE=1 or felgpu_set_result_reuse_(0).
bool result_reuse();
// The tangent mode (felgpu_set_tangent_ before the first registration):
// every device copy holds an FP64 value and its derivative along one input
// direction (Dual, 16 bytes); the kernels run with T = PT<T> = Dual.
bool tangent_mode();
// Element size (4, 8 or 16) of the device copy of the array containing
// `host`; 8 for an unregistered array.
int elem_size(const void* host);

// Runs f(T{}) with T the flow element type of the current mode: double,
// float (FP32 mode) or Dual (tangent mode).
template <class F>
inline void with_flow(F&& f) {
  if (tangent_mode()) f(Dual{});
  else if (flow_fp32()) f(float{});
  else f(double{});
}
// Runs f(T{}) with T the element type of the given element size.
template <class F>
inline void with_elem(int esize, F&& f) {
  if (esize == 16) f(Dual{});
  else if (esize == 4) f(float{});
  else f(double{});
}

// Resolves the host array arguments of one drop-in call to device memory.
// Arrays registered with felgpu_register_ are device-resident: the device
// copy is the current one and no transfer happens.  Any other host pointer is
// mirrored for the duration of the call: copied to the device before the
// kernels (every access, so partial writes keep the untouched elements) and
// back afterwards when written, converted to the element type the kernel
// reads.  finish() completes those copies.  A registered array whose device
// precision differs from the element type asked for is an error.
class Call {
 public:
  explicit Call(const char* name);
  ~Call();
  template <class T>
  T* arr(const void* host, size_t count, Access a, const char* what) {
    return static_cast<T*>(resolve(host, count, a, what, (int)sizeof(T)));
  }
  double* in(const void* host, size_t count, const char* what = "") {
    return arr<double>(host, count, kRead, what);
  }
  double* out(const void* host, size_t count, const char* what = "") {
    return arr<double>(host, count, kWrite, what);
  }
  double* io(const void* host, size_t count, const char* what = "") {
    return arr<double>(host, count, kReadWrite, what);
  }
  template <class T>
  T* in_t(const void* host, size_t count, const char* what = "") {
    return arr<T>(host, count, kRead, what);
  }
  template <class T>
  T* out_t(const void* host, size_t count, const char* what = "") {
    return arr<T>(host, count, kWrite, what);
  }
  template <class T>
  T* io_t(const void* host, size_t count, const char* what = "") {
    return arr<T>(host, count, kReadWrite, what);
  }
  // an output whose every element the call writes
  template <class T>
  T* all_t(const void* host, size_t count, const char* what = "") {
    return arr<T>(host, count, kOverwrite, what);
  }
  // an output of which the call writes exactly the interior cells of the
  // padded grid d (deferred contents are then made for the ghost cells only)
  template <class T>
  T* interior_t(const void* host, const Dims& d, const char* what = "") {
    box_ = d;
    return arr<T>(host, (size_t)d.n3, kInterior, what);
  }
  void finish();

 private:
  void* resolve(const void* host, size_t count, Access a, const char* what, int esize);
  const char* name_;
  Dims box_{};
  struct Pending {
    void* host;
    void* dev;
    size_t count;
    int esize;
  };
  Pending pend_[64];
  int npend_ = 0;
  bool finished_ = false;
};

// Content versions of registered arrays.  Every write through an entry
// point, a registration or an upload gives the array a fresh number; cpy_
// hands the source's number to the destination.  Equal numbers therefore
// mean equal contents, which lets an entry point recognise that it is asked
// for a result it has already computed.  content_version returns 0 unless
// `host` is exactly one whole registered array of `count` doubles.
uint64_t content_version(const void* host, size_t count);
// As content_version, also for a sub-range of a registered array (a slice
// such as phi(:,:,:,l)): a number that stays the same until a write
// overlaps the sub-range.  0 for memory that is not registered.
uint64_t content_key(const void* host, size_t count);
// After a copy of count doubles from src to dst (whole arrays or slices):
// dst's contents are numbered as src's.
void copy_content_version(const void* src, const void* dst, size_t count);
// Content number of the elements of a whole padded array outside the box
// (-2..n+2)^3, and its restoration by an entry point that writes exactly
// that box (the number survives such a write).
uint64_t shell_key(const void* host, size_t count);
void keep_shell(const void* host, size_t count, uint64_t key);

// Result memo.  An entry point whose outputs are a pure function of its
// arguments describes the call by a key: an operation number, the scalar
// arguments (bit patterns) and the content numbers of the input arrays (and
// of whatever part of an output the call does not overwrite).  memo_record
// notes which arrays hold the outputs of a key; a later call with the same
// key finds them with memo_reuse while they still hold those contents and
// gets them without a kernel: in place when the output array is the holder,
// else (allow_copy) by a deferred copy from the holder.  A key computed
// again gives its outputs the same content numbers as before, so a chain of
// calls repeated on equal inputs is recognised link by link.
// FELGPU_NO_RESULT_REUSE=1 turns reuse off.
struct MemoKey {
  std::vector<uint64_t> w;
  bool ok = true;  // false when an input is not a registered array
  explicit MemoKey(int op) { w.push_back((uint64_t)op); }
  MemoKey& num(int64_t x) {
    w.push_back((uint64_t)x);
    return *this;
  }
  MemoKey& num(double x) {
    uint64_t b;
    std::memcpy(&b, &x, sizeof b);
    w.push_back(b);
    return *this;
  }
  MemoKey& dims(const Dims& d) { return num((int64_t)d.ni).num((int64_t)d.nj).num((int64_t)d.nk); }
  MemoKey& content(const void* host, size_t count);
  MemoKey& shell(const void* host, size_t count);
};
struct MemoOut {
  const void* host;
  size_t count;
};
bool memo_reuse(const MemoKey& k, std::initializer_list<MemoOut> outs, bool allow_copy);
void memo_record(const MemoKey& k, std::initializer_list<MemoOut> outs);
// Before a call overwrites outputs that still hold a memoised result: the
// result is parked -- the buffer that holds it is kept by the memo and the
// array gets another (with the contents outside the box (-2..n+2)^3 copied
// over when `box`, the call writing the box only) -- so a later call with
// that result's key gets the buffer back instead of recomputing it.  At
// most a fixed number of bytes is parked; the least recently used results
// are dropped first.
void memo_park(const Dims& d, std::initializer_list<MemoOut> outs, bool box);
void memo_release();
// Key of a face-gradient call (cal_grad_p2a_, cal_grad_p2uvw_): the
// arguments, the input's contents and the shells of the three outputs,
// which the call leaves as they are; sh receives those shell numbers.
inline MemoKey gradient_key(int op, int esize, const Dims& d, int ID, int svall, double dxinv,
                            const double* dyinv_array, double dzinv, const void* q,
                            const void* const* out, uint64_t* sh) {
  MemoKey k(op);
  k.num((int64_t)esize).dims(d).num((int64_t)ID).num((int64_t)svall).num(dxinv).num(dzinv);
  for (int m = 0; m < svall + 6; m++) k.num(dyinv_array[m]);
  k.content(q, (size_t)d.n3);
  for (int m = 0; m < 3; m++) {
    k.shell(out[m], (size_t)d.n3);
    sh[m] = k.w.back();
  }
  return k;
}
// Operation numbers of the memoised entry points.
enum MemoOp {
  kOpGradP2a = 1,
  kOpGradP2uvw,
  kOpCalrn,
  kOpCalrkap,
  kOpCalfst,
  kOpSumFst,
  kOpZero
};
// cpy_ between two whole registered arrays of one element type: moves the
// source's device buffer to the destination and defers the source's copy
// until the source is accessed or the destination changes.  Returns false
// (nothing done) for any other pair of arrays.
bool deferred_copy(const void* src, const void* dst, size_t count);
// init_q_ on a whole registered array: its contents become zeros without a
// fill; the fill is made when the array is accessed (a whole overwrite needs
// none).  False (nothing done) for any other array.
bool defer_zero(const void* host, size_t count);
// sum_fst_ of src[m] onto dst[m] (m < narr) whose contents are deferred
// zero fills: each dst's contents become 0.0 + src element by element, made
// when dst is accessed or before src changes.  All or nothing: false
// (nothing done) unless every pair is two distinct whole registered arrays
// of one element type, dst a deferred zero, and no dst is another pair's
// array.
bool defer_add_zero(const void* const* dst, const void* const* src, int narr, size_t count);
// The device buffer a deferred 0.0 + src array is made from (src's
// contents), or null when `host` is not such an array; a consumer may read
// it and add 0.0 itself.
const void* deferred_add_zero_source(const void* host, size_t count);
// Replaces the device buffer behind a whole registered array (same size and
// element type) and returns the previous one; the caller owns it.
void* exchange_device_buffer(const void* host, size_t count, void* buf);
// Free the result stashes (called when all arrays are unregistered).
void weno_release();
void rn_release();
void dense_src_release();
// Counts a call answered from a pass of an earlier call (felgpu_stash_hits_).
void count_stash_hit();

// A deferred producer: the 45 viscous coefficient arrays of
// cal_arith_coef_vis_ are a pure function of the viscosity.  When all of
// them are device-resident the entry point only records where they come
// from; the Gauss-Seidel solve, which is their only consumer in the
// reference's step, recomputes the values on the fly from the viscosity.
// Any other access to one of the arrays (an argument of another entry point,
// a download), or a write to the viscosity, materialises them first.
struct DeferredCoef {
  bool active = false;
  int esize = 8;  // element size of the viscosity and the coefficients
  const void* out_host[45];
  void* out_dev[45];
  const void* src_host;
  const void* src_dev;
  Dims d;
  double dxinv, dyinv, dzinv;
  // the viscosity's former buffer, taken over when the viscosity array was
  // overwritten as a whole (returned when the record ends)
  void* owned = nullptr;
  size_t owned_bytes = 0;
};
DeferredCoef& deferred_coef();
bool deferred_coef_disabled();                    // FELGPU_EAGER_COEF=1 turns deferral off
bool is_registered(const void* host, size_t count);  // whole range device-resident?
void deferred_coef_flush();                       // materialise now, if pending
// Flush if host conflicts.  `whole`: the access overwrites the whole array
// containing host, so a source about to change can hand over its buffer
// instead of being copied.
void deferred_coef_check(const void* host, bool write, bool whole = false);
// Ends the pending record without materialising it (the caller redefines
// every one of its outputs).
void deferred_coef_drop();
// The 45 coefficient arrays from the viscosity, in the element type esize.
void coef_vis_any(const Dims d, double dxinv, double dyinv, double dzinv, const void* rmu,
                  void* const* out45, int esize);

// A deferred chain of entry points: consecutive calls of the reference's
// step whose intermediate arrays the next call consumes (calsij_ ->
// cal_arith_tau_ -> cal_div_tensor_) are recorded instead of run, and the
// last one runs the whole chain in one kernel that writes every output of
// every call.  `outs` are the arrays whose contents are pending, `srcs` the
// arrays the pending contents are computed from; any access to an output,
// or a write to a source, through any entry point runs `flush` first (the
// chain's own kernels, writing what the calls so far would have written).
// One chain is pending at a time.  FELGPU_EAGER_CHAINS=1 turns this off.
struct DeferredChain {
  bool active = false;
  int kind = 0;
  std::vector<const void*> outs, srcs;  // host base addresses of whole registered arrays
  std::function<void()> flush;
};
DeferredChain& deferred_chain();
bool chain_deferral_enabled();
void deferred_chain_flush();
void deferred_chain_check(const void* host, bool write);

// Pinned host scratch for small device->host read-backs (reductions).
double* pinned_scalars();  // 64 doubles
void sync();

// Device scratch reused across calls (grows, never shrinks).
void* scratch(int slot, size_t bytes);

// A positive guard constant of the reference (1e-99, 1e-80) in element type
// T: the value itself in FP64 (and in the tangent mode); in FP32, where it
// would flush to zero, 1e-30 (still far below every quantity it guards).
template <class T>
__host__ __device__ constexpr T guard(double v) {
  return is_f32<T> ? T(v < 1.0e-30 ? 1.0e-30 : v) : T(v);
}

// Launch helpers --------------------------------------------------------------

// Kernel over an index box, one thread per cell, 64x4 thread blocks.
//
// XCD-aware block order: the 8 XCDs of the chip take workgroups round-robin
// by dispatch id and each has its own L2.  Dispatch id n is mapped to the
// logical tile L so that XCD x owns one contiguous range of tiles (x fastest
// within a row of 64, then j, then k): the tiles an XCD runs back to back
// are neighbours in j and k, and their stencil halos hit its own L2.
__device__ __forceinline__ int xcd_tile(int n, int ntiles) {
  if (ntiles < 0) return n;  // plain dispatch order (the launch passes -ntiles)
  const int x = n & 7, pos = n >> 3;
  const int q = ntiles >> 3, r = ntiles & 7;
  return x * q + (x < r ? x : r) + pos;
}

template <class F>
__global__ void __launch_bounds__(256) box_kernel(Box b, int gx, int gy, int ntiles, F f) {
  const int L = xcd_tile((int)blockIdx.x, ntiles);  // ntiles < 0: dispatch order
  const int bx = L % gx, t = L / gx, by = t % gy, bz = t / gy;
  const int i = b.i0 + bx * 64 + (int)threadIdx.x;
  const int j = b.j0 + by * 4 + (int)threadIdx.y;
  const int k = b.k0 + bz;
  if (i <= b.i1 && j <= b.j1) f(i, j, k);
}

// Tile-count argument for xcd_tile: n with XCD-contiguous order, -n without.
// Order per kernel class: 0 = stencil kernels (launch_box), 1 = colour-pair
// Gauss-Seidel, 2 = CSF normals.  Default: dispatch order for all three
// (measured on MI355X at 512x128x256: contiguous XCD ranges gave no gain --
// halo re-reads across XCDs are served by the 256 MB last-level cache --
// and unbalanced the interface-heavy MTHINC kernel); FELGPU_XCD_ORDER
// (three digits, e.g. "001") overrides.
int xcd_arg(int kernel_class, int ntiles);

template <class F>
inline void launch_box(const Box& b, F f) {
  if (b.nx() <= 0 || b.ny() <= 0 || b.nz() <= 0) return;
  const int gx = (b.nx() + 63) / 64, gy = (b.ny() + 3) / 4, n = gx * gy * b.nz();
  box_kernel<F><<<dim3(n), dim3(64, 4, 1), 0, stream()>>>(b, gx, gy, xcd_arg(0, n), f);
  FEL_CHECK(hipGetLastError());
}

// Kernel over a flat range [0,n), grid-stride.
template <class F>
__global__ void __launch_bounds__(256) flat_kernel(int n, F f) {
  for (int m = (int)(blockIdx.x * 256 + threadIdx.x); m < n; m += (int)(gridDim.x * 256)) f(m);
}

template <class F>
inline void launch_flat(int n, F f) {
  if (n <= 0) return;
  int blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  flat_kernel<F><<<dim3(blocks), dim3(256), 0, stream()>>>(n, f);
  FEL_CHECK(hipGetLastError());
}

// Kernel over the ghost shell of the padded image (every cell outside the
// interior [1,ni]x[1,nj]x[1,nk]): the 6 k-ghost planes, the 6 j-ghost rows
// of each interior plane, and the 6 i-ghost cells of each interior row.
template <class F>
inline void launch_shell(const Dims& d, F f) {
  const int planes = 6 * d.sx * d.sy, rows = d.nk * 6 * d.sx, cols = d.nk * d.nj * 6;
  launch_flat(planes + rows + cols, [=] __device__(int m) {
    int i, j, k;
    if (m < planes) {
      const int p = m / (d.sx * d.sy), r = m - p * (d.sx * d.sy);
      k = p < 3 ? p - 2 : d.nk + p - 2;
      j = r / d.sx - 2;
      i = r % d.sx - 2;
    } else if (m < planes + rows) {
      const int q = m - planes, kk = q / (6 * d.sx), r = q - kk * (6 * d.sx), jr = r / d.sx;
      k = kk + 1;
      j = jr < 3 ? jr - 2 : d.nj + jr - 2;
      i = r % d.sx - 2;
    } else {
      const int q = m - planes - rows, kk = q / (6 * d.nj), r = q - kk * (6 * d.nj), jj = r / 6,
                ir = r % 6;
      k = kk + 1;
      j = jj + 1;
      i = ir < 3 ? ir - 2 : d.ni + ir - 2;
    }
    f(i, j, k);
  });
}

// Deterministic reductions over the interior cells [1,n]^3, accumulated in
// A (double; Dual in the tangent mode, whose value part then adds in the
// same order as the FP64 sum).  Stage 1: fixed grid, each thread sums a
// fixed strided subset, fixed LDS tree per block.  Stage 2: one block sums
// the block partials in fixed order and hands the NV totals to g(total,
// out), which runs on the device, so a following kernel can consume derived
// scalars without a host round trip.
constexpr int kRedBlocks = 1024;

template <int NV, class A, class F>
__global__ void __launch_bounds__(256) sum_stage1(Dims d, F f, A* part) {
  __shared__ A sh[NV][256];
  A acc[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) acc[c] = A(0.0);
  const int nxy = d.ni * d.nj, n = nxy * d.nk;
  // thread (b, t) adds the cells m = b*256 + t + q*G, q = 0, 1, ..., in
  // order; (i, j, k) of m advance by the decomposition of G with carries
  // instead of two integer divisions per cell
  const int G = (int)(gridDim.x * 256);
  const int gk = G / nxy, gr = G - gk * nxy, gj = gr / d.ni, gi = gr - gj * d.ni;
  int m = (int)(blockIdx.x * 256 + threadIdx.x);
  int k = m / nxy, i = m - k * nxy, j = i / d.ni;
  i -= j * d.ni;
#pragma unroll 4
  for (; m < n; m += G) {
    A v[NV];
    f(i + 1, j + 1, k + 1, v);
#pragma unroll
    for (int c = 0; c < NV; c++) acc[c] = acc[c] + v[c];
    i += gi;
    j += gj;
    k += gk;
    if (i >= d.ni) {
      i -= d.ni;
      j++;
    }
    if (j >= d.nj) {
      j -= d.nj;
      k++;
    }
  }
  const int t = (int)threadIdx.x;
#pragma unroll
  for (int c = 0; c < NV; c++) sh[c][t] = acc[c];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
#pragma unroll
      for (int c = 0; c < NV; c++) sh[c][t] = sh[c][t] + sh[c][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int c = 0; c < NV; c++) part[blockIdx.x * NV + c] = sh[c][0];
  }
}

template <int NV, class A, class G>
__global__ void __launch_bounds__(256) sum_stage2(int nb, const A* part, double* out, G g) {
  __shared__ A sh[NV][256];
  const int t = (int)threadIdx.x;
  A acc[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) acc[c] = A(0.0);
  for (int b = t; b < nb; b += 256) {
#pragma unroll
    for (int c = 0; c < NV; c++) acc[c] = acc[c] + part[b * NV + c];
  }
#pragma unroll
  for (int c = 0; c < NV; c++) sh[c][t] = acc[c];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
#pragma unroll
      for (int c = 0; c < NV; c++) sh[c][t] = sh[c][t] + sh[c][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    A tot[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) tot[c] = sh[c][0];
    g(tot, out);
  }
}

// Runs both stages; `out` is device memory.  f(i, j, k, A* v) gives the NV
// addends of cell (i,j,k); g(const A* totals, double* out) stores results.
template <int NV, class A = double, class F, class G>
inline void interior_reduce(const Dims& d, F f, G g, double* out) {
  A* part = static_cast<A*>(scratch(15, sizeof(A) * NV * kRedBlocks));
  sum_stage1<NV, A, F><<<dim3(kRedBlocks), dim3(256), 0, stream()>>>(d, f, part);
  FEL_CHECK(hipGetLastError());
  sum_stage2<NV, A, G><<<dim3(1), dim3(256), 0, stream()>>>(kRedBlocks, (const A*)part, out, g);
  FEL_CHECK(hipGetLastError());
}

// Reduction over the interior in tiles of launch_box's shape (64 x 4 cells,
// kBoxRedPlanes consecutive planes per 256-thread block), for functors that
// read neighbour cells: neighbouring rows are read by the same block and a
// thread's next plane is its last plane's upper neighbour.  Deterministic
// like interior_reduce: each thread adds its cells plane by plane, a fixed
// LDS tree per block, and sum_stage2 adds the block partials in order.
constexpr int kBoxRedPlanes = 8;

template <int NV, class A, class F>
__global__ void __launch_bounds__(256) box_sum_stage1(Dims d, int gx, int gy, F f, A* part) {
  __shared__ A sh[NV][256];
  const int L = (int)blockIdx.x;
  const int bx = L % gx, tt = L / gx, by = tt % gy, bz = tt / gy;
  const int i = 1 + bx * 64 + (int)threadIdx.x, j = 1 + by * 4 + (int)threadIdx.y;
  A acc[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) acc[c] = A(0.0);
  if (i <= d.ni && j <= d.nj) {
    const int k0 = 1 + bz * kBoxRedPlanes;
    const int k1 = k0 + kBoxRedPlanes - 1 < d.nk ? k0 + kBoxRedPlanes - 1 : d.nk;
    for (int k = k0; k <= k1; k++) {
      A v[NV];
      f(i, j, k, v);
#pragma unroll
      for (int c = 0; c < NV; c++) acc[c] = acc[c] + v[c];
    }
  }
  const int t = (int)threadIdx.x + 64 * (int)threadIdx.y;
#pragma unroll
  for (int c = 0; c < NV; c++) sh[c][t] = acc[c];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
#pragma unroll
      for (int c = 0; c < NV; c++) sh[c][t] = sh[c][t] + sh[c][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int c = 0; c < NV; c++) part[(size_t)L * NV + c] = sh[c][0];
  }
}

// interior_reduce for stencil functors (box_sum_stage1, then sum_stage2).
template <int NV, class A = double, class F, class G>
inline void interior_reduce_stencil(const Dims& d, F f, G g, double* out) {
  const int gx = (d.ni + 63) / 64, gy = (d.nj + 3) / 4;
  const int gz = (d.nk + kBoxRedPlanes - 1) / kBoxRedPlanes;
  const int nb = gx * gy * gz;
  A* part = static_cast<A*>(scratch(18, sizeof(A) * NV * (size_t)nb));
  box_sum_stage1<NV, A, F><<<dim3(nb), dim3(64, 4, 1), 0, stream()>>>(d, gx, gy, f, part);
  FEL_CHECK(hipGetLastError());
  sum_stage2<NV, A, G><<<dim3(1), dim3(256), 0, stream()>>>(nb, (const A*)part, out, g);
  FEL_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// FP32 mode: updates of the prognostic state.
//
// In FP32 arithmetic an update u + du loses the part of du below half an ulp
// of u at every operation, and it loses it in the same direction step after
// step wherever the flow changes slowly, so the difference from FP64 grows
// with the step count.  The velocity and pressure updates of the FP32 mode
// are therefore formed in FP64 from the FP32 inputs and rounded to FP32 once,
// at the store (acc_t is their arithmetic type).  FP64 and the tangent mode
// evaluate in their element type as the reference does.
// ---------------------------------------------------------------------------

// Arithmetic type of those updates: FP64 for FP32 storage, else the element.
template <class T>
using acc_t = std::conditional_t<is_f32<T>, double, T>;

// Small device buffer for reduction results (slot-indexed, 16 doubles each).
double* dev_scalars(int slot);
// Copies n doubles from device to the pinned buffer and waits for them.
const double* read_scalars(const double* dev, int n);

// Aborts with a message naming `who` unless the library runs in the FP64
// mode (the adjoint routines have no FP32 or tangent form).
void adjoint_fp64_only(const char* who);
// Adds the sum of the n values of device array a to the host scalar *out
// (a deterministic two-stage reduction; waits for it).
void adjoint_add_sum(const double* a, int n, double* out);

// Finds this device's copy of the allocation-free solve's static work array
// (k_sor.hip); once per device, with the device current.
void sor_noalloc_setup();

// Fortran list of edit descriptor Ew.d: v right-justified in w characters
// as 0.<d digits>E+xx (E dropped for 3-digit exponents), the form the
// reference's monitor lines print (e.g. srcs/solu_sor4.f:299).  Returns out.
char* fortran_e(char* out, double v, int w, int d);

// With FELGPU_TRACE_COPIES=1, notes a copy or transfer on stderr with the
// entry point and argument it serves (a diagnostic).
void trace_event(const char* kind, size_t bytes);

// Launch event timer used by the benchmark (felgpu_prof_*), for
// hipExtLaunchKernelGGL, which stamps an event with its launch's own start
// or end (no marker packets between launches).  One timed span covers
// `launches` back-to-back launches on the library stream: the first gets
// *start (first = true), the last *stop (last = true), the ones between
// neither; a single launch passes both flags and launches = 1.  Both
// events are null while the timers are off.
void prof_launch_events(int id, bool first, bool last, int launches, hipEvent_t* start,
                        hipEvent_t* stop);

}  // namespace fel
