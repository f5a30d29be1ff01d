// runtime.hip -- stream, device mirrors of the host arrays the Fortran host
// passes by address, scratch memory, scalar read-back and the event timers.
//
// The reference allocates every field once on the host (srcs/allocate.h) and
// passes base addresses (or slice addresses such as phi(-2,-2,-2,l)) to each
// subroutine.  The drop-in entry points receive the same host addresses and
// resolve them here.  Registered ranges (felgpu_register_) live on the device
// for the whole run; a pointer anywhere inside a registered range resolves to
// the same offset of its device copy.  Unregistered pointers are mirrored per
// call (copy in, run, copy out), so every entry point also works, slowly, on
// plain host memory.
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "felgpu.hpp"

namespace fel {

namespace {

struct Range {
  uintptr_t host;
  size_t bytes;  // host bytes (FP64 elements)
  void* dev;
  int esize;     // device element size: 8, 4 for an FP32 copy, 16 for a Dual copy
  // content version: a fresh number on every write (entry point output,
  // registration, upload); a whole-array copy hands the source's number to
  // the destination, so equal numbers mean equal contents
  uint64_t ver;
  // deferred copy: this range's contents equal those of `lazy` (which holds
  // them in its buffer); ndeps counts the ranges deferred on this one
  struct Range* lazy;
  int ndeps;
  // what the deferred contents are: kCopy (lazy's contents) or kAddZero
  // (0.0 + lazy's contents, element by element: sum_fst_ onto zeros)
  int lkind = 0;
  // deferred zero fill (init_q_): the contents are zeros, the buffer is not
  bool zero = false;
  // content number of the elements outside the box (-2..n+2)^3 of a padded
  // array (0: none assigned yet); kept by the entry points that write
  // exactly that box, renewed by every other write
  uint64_t shell = 0;
  // content numbers of sub-ranges (e.g. one slice phi(:,:,:,l)), assigned
  // on request and dropped by any overlapping write
  struct Sub {
    size_t off, count;  // in elements
    uint64_t key;
  };
  std::vector<Sub> subs;
};

uint64_t g_version = 0;
uint64_t next_version() { return ++g_version; }

// A write to [p, p+bytes) of g: fresh content number for the whole range,
// unknown shell, and the sub-range numbers the write overlaps dropped.
void note_write(Range& g, uintptr_t p, size_t bytes) {
  g.ver = next_version();
  g.shell = 0;
  if (p == g.host && bytes == g.bytes) {
    g.subs.clear();
    return;
  }
  const size_t a = (p - g.host) / sizeof(double), b = a + bytes / sizeof(double);
  std::vector<Range::Sub> keep;
  for (const Range::Sub& x : g.subs)
    if (x.off + x.count <= a || x.off >= b) keep.push_back(x);
  g.subs.swap(keep);
}
void note_write(Range& g) { note_write(g, g.host, g.bytes); }

struct Runtime {
  bool init = false;
  int device = 0;
  int flow_esize = 8;  // element size of arrays registered with felgpu_register_
  bool tangent = false;  // every device copy is a Dual (value, derivative)
  hipStream_t s = nullptr;
  std::map<uintptr_t, Range> reg;         // registered, keyed by host start
  std::map<uintptr_t, Range> transient;   // per-call mirrors, cached
  std::vector<void*> scratch_ptr;
  std::vector<size_t> scratch_size;
  double* pinned = nullptr;
  double* dscal = nullptr;
  // event timers: per id, a list of recorded (begin, end) pairs that are
  // only read back (and recycled) when the caller asks for the totals, so
  // timing adds no host synchronisation to the timed region
  struct Timer {
    std::vector<hipEvent_t> b, e;
    std::vector<int> span;  // launches between b[m]'s launch and e[m]'s, inclusive
    int used = 0;
    double ms = 0.0;
    long n = 0;
  };
  std::vector<Timer> timers;
  bool prof = false;
};

Runtime& R() {
  static Runtime r;
  return r;
}

Range* find_range(std::map<uintptr_t, Range>& m, uintptr_t p, size_t bytes) {
  auto it = m.upper_bound(p);
  if (it == m.begin()) return nullptr;
  --it;
  Range& r = it->second;
  if (p >= r.host && p + bytes <= r.host + r.bytes) return &r;
  return nullptr;
}

// The registered range that is exactly [host, host + count doubles).
Range* whole_range(const void* host, size_t count) {
  Runtime& r = R();
  auto it = r.reg.find(reinterpret_cast<uintptr_t>(host));
  if (it == r.reg.end() || it->second.bytes != count * sizeof(double)) return nullptr;
  return &it->second;
}

// ---------------------------------------------------------------------------
// Deferred whole-array copies.  cpy_ between two whole registered arrays
// moves the source's buffer to the destination and leaves the source
// "deferred": its contents are defined as the destination's.  The copy is
// made only when the deferred array is accessed, or before the array it
// defers to changes.  The timestep's rotation (uo <- u, u <- un, and
// un <- u before the viscous solve) then costs one copy per component
// instead of three.
// ---------------------------------------------------------------------------
void stream_copy(void* dst, const void* src, size_t bytes) {
  FEL_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, R().s));
}

constexpr int kCopy = 0, kAddZero = 1;

template <class T>
__global__ void __launch_bounds__(256) add_zero_kernel(T* dst, const T* src, size_t n) {
  const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) dst[m] = T(0.0) + src[m];
}

void add_zero(void* dst, const void* src, size_t n, int esize) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipStream_t s = R().s;
  if (esize == 16)
    add_zero_kernel<Dual><<<nb, 256, 0, s>>>(static_cast<Dual*>(dst), static_cast<const Dual*>(src), n);
  else if (esize == 4)
    add_zero_kernel<float><<<nb, 256, 0, s>>>(static_cast<float*>(dst), static_cast<const float*>(src), n);
  else
    add_zero_kernel<double><<<nb, 256, 0, s>>>(static_cast<double*>(dst), static_cast<const double*>(src), n);
  FEL_CHECK(hipGetLastError());
}

// Copies the cells of a padded array outside the interior [1,n]^3: one
// thread per ghost cell (launch_shell), B bytes each.
template <int B>
void copy_outside_t(const Dims& d, char* dst, const char* src) {
  launch_shell(d, [=] __device__(int i, int j, int k) {
    const size_t o = (size_t)d.ix(i, j, k) * B;
    for (int b = 0; b < B; b += 4)
      *reinterpret_cast<uint32_t*>(dst + o + b) = *reinterpret_cast<const uint32_t*>(src + o + b);
  });
}

void copy_outside_interior(const Dims& d, void* dst, const void* src, int esize) {
  char* o = static_cast<char*>(dst);
  const char* i = static_cast<const char*>(src);
  if (esize == 16) copy_outside_t<16>(d, o, i);
  else if (esize == 4) copy_outside_t<4>(d, o, i);
  else copy_outside_t<8>(d, o, i);
}

// FELGPU_TRACE_COPIES=1: every deferred copy, sum or fill made, with the
// entry point and argument that needed it, on stderr (a diagnostic).
const char* g_trace_call = "";
const char* g_trace_what = "";
bool trace_copies() {
  static const int on = [] {
    const char* e = std::getenv("FELGPU_TRACE_COPIES");
    return e && *e == '1' ? 1 : 0;
  }();
  return on != 0;
}
}  // namespace

void trace_event(const char* kind, size_t bytes) {
  if (trace_copies())
    std::fprintf(stderr, "felgpu: %s %zu B, in %s (%s)\n", kind, bytes, g_trace_call, g_trace_what);
}

namespace {
void trace_copy(const char* kind, const Range& g, size_t bytes) {
  if (trace_copies())
    std::fprintf(stderr, "felgpu: %s %zu B into host range %p, in %s (%s)\n", kind, bytes,
                 reinterpret_cast<void*>(g.host), g_trace_call, g_trace_what);
}

void materialize(Range& g) {
  if (g.zero) {
    trace_copy("zero fill", g, g.bytes / sizeof(double) * (size_t)g.esize);
    FEL_CHECK(hipMemsetAsync(g.dev, 0, g.bytes / sizeof(double) * (size_t)g.esize, R().s));
    g.zero = false;
  }
  if (!g.lazy) return;
  const size_t n = g.bytes / sizeof(double);
  trace_copy(g.lkind == kAddZero ? "deferred sum" : "deferred copy", g, n * (size_t)g.esize);
  if (g.lkind == kAddZero)
    add_zero(g.dev, g.lazy->dev, n, g.esize);
  else
    stream_copy(g.dev, g.lazy->dev, n * (size_t)g.esize);
  g.lazy->ndeps--;
  g.lazy = nullptr;
  g.lkind = kCopy;
}

void materialize_dependents(Range& g) {
  if (g.ndeps == 0) return;
  for (auto& kv : R().reg)
    if (kv.second.lazy == &g) materialize(kv.second);
}

// The contents are about to be replaced as a whole: nothing deferred needs
// to be made.
void drop_lazy(Range& g) {
  g.zero = false;
  if (!g.lazy) return;
  g.lazy->ndeps--;
  g.lazy = nullptr;
  g.lkind = kCopy;
}

// Everything that reads or writes the device copy of g directly.
void settle(Range& g, bool write) {
  materialize(g);
  if (write) materialize_dependents(g);
}

}  // namespace

void ensure_init() {
  Runtime& r = R();
  if (r.init) return;
  const char* dev = std::getenv("FELGPU_DEVICE");
  r.device = dev ? std::atoi(dev) : 0;
  FEL_CHECK(hipSetDevice(r.device));
  sor_noalloc_setup();
  FEL_CHECK(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
  FEL_CHECK(hipHostMalloc((void**)&r.pinned, 64 * sizeof(double), hipHostMallocDefault));
  FEL_CHECK(hipMalloc((void**)&r.dscal, 64 * 16 * sizeof(double)));
  FEL_CHECK(hipMemsetAsync(r.dscal, 0, 64 * 16 * sizeof(double), r.s));
  const char* pr = std::getenv("FELGPU_PRECISION");
  if (pr && (std::strcmp(pr, "fp32") == 0 || std::strcmp(pr, "32") == 0)) r.flow_esize = 4;
  r.init = true;
}

bool flow_fp32() {
  ensure_init();
  return R().flow_esize == 4;
}

namespace {
int g_reuse = -1;  // -1: not yet read from the environment
}

bool result_reuse() {
  if (g_reuse < 0) {
    const char* e = std::getenv("FELGPU_NO_RESULT_REUSE");
    g_reuse = (e && std::atoi(e) != 0) ? 0 : 1;
  }
  return g_reuse != 0;
}

bool tangent_mode() {
  ensure_init();
  return R().tangent;
}

int elem_size(const void* host) {
  const Range* g = find_range(R().reg, reinterpret_cast<uintptr_t>(host), 1);
  return g ? g->esize : 8;
}

namespace {

// Host FP64 <-> device copies of n elements of size esize.  A Dual copy
// takes the host values with a zero derivative, and gives back its values.
void h2d(void* dev, const void* host, size_t n, int esize) {
  Runtime& r = R();
  trace_event("upload", n * (size_t)esize);
  if (esize == 8) {
    FEL_CHECK(hipMemcpyAsync(dev, host, n * 8, hipMemcpyHostToDevice, r.s));
    return;
  }
  if (esize == 16) {
    std::vector<Dual> tmp(n);
    const double* h = static_cast<const double*>(host);
    for (size_t m = 0; m < n; m++) tmp[m] = Dual(h[m], 0.0);
    FEL_CHECK(hipMemcpyAsync(dev, tmp.data(), n * 16, hipMemcpyHostToDevice, r.s));
    FEL_CHECK(hipStreamSynchronize(r.s));
    return;
  }
  std::vector<float> tmp(n);
  const double* h = static_cast<const double*>(host);
  for (size_t m = 0; m < n; m++) tmp[m] = (float)h[m];
  FEL_CHECK(hipMemcpyAsync(dev, tmp.data(), n * 4, hipMemcpyHostToDevice, r.s));
  FEL_CHECK(hipStreamSynchronize(r.s));
}

void d2h(void* host, const void* dev, size_t n, int esize) {
  Runtime& r = R();
  trace_event("download", n * (size_t)esize);
  if (esize == 8) {
    FEL_CHECK(hipMemcpyAsync(host, dev, n * 8, hipMemcpyDeviceToHost, r.s));
    return;
  }
  if (esize == 16) {
    std::vector<Dual> tmp(n);
    FEL_CHECK(hipMemcpyAsync(tmp.data(), dev, n * 16, hipMemcpyDeviceToHost, r.s));
    FEL_CHECK(hipStreamSynchronize(r.s));
    double* h = static_cast<double*>(host);
    for (size_t m = 0; m < n; m++) h[m] = tmp[m].v;
    return;
  }
  std::vector<float> tmp(n);
  FEL_CHECK(hipMemcpyAsync(tmp.data(), dev, n * 4, hipMemcpyDeviceToHost, r.s));
  FEL_CHECK(hipStreamSynchronize(r.s));
  double* h = static_cast<double*>(host);
  for (size_t m = 0; m < n; m++) h[m] = (double)tmp[m];
}

}  // namespace

hipStream_t stream() {
  ensure_init();
  return R().s;
}

void sync() { FEL_CHECK(hipStreamSynchronize(stream())); }

Dims make_dims(int ni, int nj, int nk) {
  if (ni < 4 || nj < 4 || nk < 4) {
    std::fprintf(stderr, "felgpu: grid %dx%dx%d too small (each extent must be >= 4)\n", ni, nj, nk);
    std::abort();
  }
  const long long n3 = (long long)(ni + 6) * (nj + 6) * (nk + 6);
  if (n3 * 6 >= (1LL << 31)) {
    std::fprintf(stderr, "felgpu: grid %dx%dx%d exceeds the 32-bit offset range\n", ni, nj, nk);
    std::abort();
  }
  Dims d;
  d.ni = ni; d.nj = nj; d.nk = nk;
  d.sx = ni + 6; d.sy = nj + 6; d.sz = nk + 6;
  d.n3 = (int)n3;
  return d;
}

void* scratch(int slot, size_t bytes) {
  Runtime& r = R();
  ensure_init();
  if ((int)r.scratch_ptr.size() <= slot) {
    r.scratch_ptr.resize(slot + 1, nullptr);
    r.scratch_size.resize(slot + 1, 0);
  }
  if (r.scratch_size[slot] < bytes) {
    if (r.scratch_ptr[slot]) {
      FEL_CHECK(hipStreamSynchronize(r.s));
      FEL_CHECK(hipFree(r.scratch_ptr[slot]));
    }
    FEL_CHECK(hipMalloc(&r.scratch_ptr[slot], bytes));
    r.scratch_size[slot] = bytes;
  }
  return r.scratch_ptr[slot];
}

double* pinned_scalars() {
  ensure_init();
  return R().pinned;
}

double* dev_scalars(int slot) {
  ensure_init();
  return R().dscal + 16 * slot;
}

char* fortran_e(char* out, double v, int w, int d) {
  char body[64];
  if (v == 0.0 || !std::isfinite(v)) {
    if (v == 0.0)
      std::snprintf(body, sizeof body, "%s0.%0*dE+00", std::signbit(v) ? "-" : "", d, 0);
    else
      std::snprintf(body, sizeof body, "%s", std::isnan(v) ? "NaN" : (v > 0 ? "Infinity" : "-Infinity"));
  } else {
    // d significant digits, correctly rounded, then shifted behind "0."
    char m[64];
    std::snprintf(m, sizeof m, "%.*E", d - 1, std::fabs(v));
    const char* e = std::strchr(m, 'E');
    const int ex = std::atoi(e + 1) + 1;
    char digits[64];
    int nd = 0;
    for (const char* c = m; c < e; c++)
      if (*c >= '0' && *c <= '9') digits[nd++] = *c;
    digits[nd] = 0;
    const char* sg = v < 0 ? "-" : "";
    if (std::abs(ex) < 100)
      std::snprintf(body, sizeof body, "%s0.%sE%c%02d", sg, digits, ex < 0 ? '-' : '+', std::abs(ex));
    else
      std::snprintf(body, sizeof body, "%s0.%s%c%03d", sg, digits, ex < 0 ? '-' : '+', std::abs(ex));
  }
  std::snprintf(out, (size_t)w + 1, "%*s", w, body);
  return out;
}

const double* read_scalars(const double* dev, int n) {
  Runtime& r = R();
  FEL_CHECK(hipMemcpyAsync(r.pinned, dev, sizeof(double) * n, hipMemcpyDeviceToHost, r.s));
  FEL_CHECK(hipStreamSynchronize(r.s));
  return r.pinned;
}

// --------------------------------------------------------------------------
// Deferred viscous coefficients
// --------------------------------------------------------------------------

DeferredCoef& deferred_coef() {
  static DeferredCoef dc;
  return dc;
}

bool deferred_coef_disabled() {
  static const int off = [] {
    const char* e = std::getenv("FELGPU_EAGER_COEF");
    return (e && std::atoi(e) != 0) ? 1 : 0;
  }();
  return off != 0;
}

int xcd_arg(int kernel_class, int ntiles) {
  static char order[4] = {0, 0, 0, 0};
  static bool init = false;
  if (!init) {
    const char* e = std::getenv("FELGPU_XCD_ORDER");
    const char* def = "000";
    for (int m = 0; m < 3; m++) order[m] = (e && (int)std::strlen(e) > m) ? e[m] : def[m];
    init = true;
  }
  return order[kernel_class] == '1' ? ntiles : -ntiles;
}

bool is_registered(const void* host, size_t count) {
  return find_range(R().reg, reinterpret_cast<uintptr_t>(host), count * sizeof(double)) != nullptr;
}

bool deferred_copy(const void* src_host, const void* dst_host, size_t count) {
  Range* src = whole_range(src_host, count);
  Range* dst = whole_range(dst_host, count);
  if (!src || !dst || src == dst || src->esize != dst->esize) return false;
  deferred_coef_check(src_host, true);
  deferred_coef_check(dst_host, true);
  if (src->lazy == dst && src->lkind == kCopy) return true;  // the contents are already equal
  // a source whose contents are not its holder's buffer as it is: make them
  if (src->zero || (src->lazy && src->lkind != kCopy)) materialize(*src);
  materialize_dependents(*dst);       // the destination's contents change
  drop_lazy(*dst);
  Range* holder = src->lazy ? src->lazy : src;
  if (holder != src) {
    // the holder's buffer changes hands: a record that reads it (the
    // viscosity of deferred coefficients, a source of a deferred chain)
    // settles first, as for a write to the holder
    deferred_coef_check(reinterpret_cast<const void*>(holder->host), true);
  }
  std::swap(holder->dev, dst->dev);   // dst now holds the contents
  // everything deferred on the old holder, and the holder itself, now defers to dst
  for (auto& kv : R().reg)
    if (kv.second.lazy == holder) {
      kv.second.lazy = dst;
      dst->ndeps++;
    }
  holder->ndeps = 0;
  holder->lazy = dst;
  holder->lkind = kCopy;
  dst->ndeps++;
  return true;
}

bool defer_zero(const void* host, size_t count) {
  Range* g = whole_range(host, count);
  if (!g) return false;
  deferred_coef_check(host, true);
  materialize_dependents(*g);
  drop_lazy(*g);
  g->zero = true;
  note_write(*g);
  return true;
}

bool defer_add_zero(const void* const* dst_host, const void* const* src_host, int narr,
                    size_t count) {
  for (int m = 0; m < narr; m++) {
    Range* dst = whole_range(dst_host[m], count);
    Range* src = whole_range(src_host[m], count);
    if (!dst || !src || dst == src || !dst->zero || dst->ndeps != 0 || dst->esize != src->esize)
      return false;
    for (int q = 0; q < narr; q++)
      if (q != m && (dst_host[q] == dst_host[m] || src_host[q] == dst_host[m])) return false;
  }
  for (int m = 0; m < narr; m++) {
    Range* dst = whole_range(dst_host[m], count);
    Range* src = whole_range(src_host[m], count);
    deferred_coef_check(src_host[m], false);
    deferred_coef_check(dst_host[m], true);
    if (src->zero || (src->lazy && src->lkind != kCopy)) materialize(*src);
    Range* holder = src->lazy ? src->lazy : src;
    dst->zero = false;
    dst->lazy = holder;
    dst->lkind = kAddZero;
    holder->ndeps++;
    note_write(*dst);
  }
  return true;
}

const void* deferred_add_zero_source(const void* host, size_t count) {
  Range* g = whole_range(host, count);
  if (!g || !g->lazy || g->lkind != kAddZero) return nullptr;
  deferred_coef_check(reinterpret_cast<const void*>(g->lazy->host), false);
  return g->lazy->dev;
}

void pool_put_buffer(void* p, size_t bytes);

void deferred_coef_drop() {
  DeferredCoef& dc = deferred_coef();
  dc.active = false;
  if (dc.owned) pool_put_buffer(dc.owned, dc.owned_bytes);
  dc.owned = nullptr;
  dc.owned_bytes = 0;
}

void deferred_coef_flush() {
  DeferredCoef& dc = deferred_coef();
  if (!dc.active) return;
  dc.active = false;
  for (int m = 0; m < 45; m++)  // every element is rewritten
    if (Range* g = find_range(R().reg, reinterpret_cast<uintptr_t>(dc.out_host[m]), 1)) {
      materialize_dependents(*g);
      drop_lazy(*g);
    }
  coef_vis_any(dc.d, dc.dxinv, dc.dyinv, dc.dzinv, dc.src_dev, dc.out_dev, dc.esize);
  for (int m = 0; m < 45; m++)
    if (Range* g = find_range(R().reg, reinterpret_cast<uintptr_t>(dc.out_host[m]), 1))
      note_write(*g);
  deferred_coef_drop();  // the kernel is queued before the buffer can be reused
}

uint64_t content_version(const void* host, size_t count) {
  const Range* g = whole_range(host, count);
  return g ? g->ver : 0;
}

uint64_t content_key(const void* host, size_t count) {
  const uintptr_t p = reinterpret_cast<uintptr_t>(host);
  Range* g = find_range(R().reg, p, count * sizeof(double));
  if (!g) return 0;
  if (p == g->host && count * sizeof(double) == g->bytes) return g->ver;
  const size_t off = (p - g->host) / sizeof(double);
  for (const Range::Sub& x : g->subs)
    if (x.off == off && x.count == count) return x.key;
  g->subs.push_back(Range::Sub{off, count, next_version()});
  return g->subs.back().key;
}

void copy_content_version(const void* src, const void* dst, size_t count) {
  const uintptr_t q = reinterpret_cast<uintptr_t>(dst);
  Range* b = find_range(R().reg, q, count * sizeof(double));
  const uint64_t key = content_key(src, count);
  if (!b || key == 0) return;
  const Range* a = whole_range(src, count);
  if (q == b->host && count * sizeof(double) == b->bytes) {
    b->ver = key;
    b->subs = a ? a->subs : std::vector<Range::Sub>{};
    b->shell = a ? a->shell : 0;
    return;
  }
  // a slice: the copy's write has dropped what it overlapped
  b->subs.push_back(Range::Sub{(q - b->host) / sizeof(double), count, key});
}

uint64_t shell_key(const void* host, size_t count) {
  Range* g = whole_range(host, count);
  if (!g) return 0;
  if (g->shell == 0) g->shell = next_version();
  return g->shell;
}

void keep_shell(const void* host, size_t count, uint64_t key) {
  if (Range* g = whole_range(host, count)) g->shell = key;
}

// --------------------------------------------------------------------------
// Result memo
// --------------------------------------------------------------------------

namespace {
struct MemoEntry {
  std::vector<uint64_t> key;
  int nout = 0;
  uintptr_t host[4];
  size_t count[4];
  uint64_t ver[4];
  // a result whose array was about to be overwritten, kept in the buffer
  // that held it (owned by the memo; null: none)
  void* parked[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t pbytes[4] = {0, 0, 0, 0};
  uint64_t used = 0;
};
std::vector<MemoEntry> g_memo;
uint64_t g_memo_tick = 0;
int64_t g_memo_hits = 0;
int64_t g_stash_hits = 0;
constexpr size_t kMemoEntries = 64;
constexpr size_t kParkedMax = (size_t)4 << 30;  // bytes of parked results at most
size_t g_parked_total = 0;

// Device buffers of whole-array size, recycled between the ranges and the
// parked results (all work is on one stream, so a recycled buffer is only
// written after every earlier use of it).
std::multimap<size_t, void*> g_pool;

void* pool_get(size_t bytes) {
  auto it = g_pool.find(bytes);
  if (it != g_pool.end()) {
    void* p = it->second;
    g_pool.erase(it);
    return p;
  }
  void* p = nullptr;
  FEL_CHECK(hipMalloc(&p, bytes));
  return p;
}

void pool_put(void* p, size_t bytes) { g_pool.emplace(bytes, p); }
}  // namespace

void* pool_get_buffer(size_t bytes) { return pool_get(bytes); }
void pool_put_buffer(void* p, size_t bytes) { pool_put(p, bytes); }

namespace {

void unpark(MemoEntry& e, int m) {
  if (!e.parked[m]) return;
  pool_put(e.parked[m], e.pbytes[m]);
  g_parked_total -= e.pbytes[m];
  e.parked[m] = nullptr;
  e.pbytes[m] = 0;
}

void unpark_all(MemoEntry& e) {
  for (int m = 0; m < 4; m++) unpark(e, m);
}

// Parks g's buffer with the memo entry whose result it holds (not `keep`'s
// entries' room), evicting the least recently used parked results to stay
// within the limit; true when parked (g->dev is then the memo's).
bool park_buffer(Range& g, const MemoEntry* keep) {
  if (g.lazy || g.zero) return false;
  MemoEntry* e = nullptr;
  int m = -1;
  for (MemoEntry& x : g_memo)
    for (int q = 0; q < x.nout; q++)
      if (x.host[q] == g.host && x.ver[q] == g.ver && !x.parked[q]) {
        e = &x;
        m = q;
      }
  if (!e) return false;
  const size_t bytes = g.bytes / sizeof(double) * (size_t)g.esize;
  while (g_parked_total + bytes > kParkedMax) {  // the least recently used go first
    MemoEntry* v = nullptr;
    for (MemoEntry& x : g_memo)
      for (int q = 0; q < 4; q++)
        if (x.parked[q] && &x != e && &x != keep && (!v || x.used < v->used)) v = &x;
    if (!v) break;
    unpark_all(*v);
  }
  if (g_parked_total + bytes > kParkedMax) return false;
  e->parked[m] = g.dev;
  e->pbytes[m] = bytes;
  g_parked_total += bytes;
  return true;
}

// Cells of a padded array outside the box (-2..n+2)^3: the faces i = ni+3,
// j = nj+3 (i <= ni+2) and k = nk+3 (i <= ni+2, j <= nj+2), each cell once.
template <int B>
__global__ void __launch_bounds__(256) copy_shell_kernel(Dims d, char* dst, const char* src) {
  const int fi = d.sy * d.sz, fj = (d.sx - 1) * d.sz, fk = (d.sx - 1) * (d.sy - 1);
  const int m = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  int i, j, k;
  if (m < fi) {
    i = d.ni + 3;
    j = m % d.sy - 2;
    k = m / d.sy - 2;
  } else if (m < fi + fj) {
    const int t = m - fi;
    j = d.nj + 3;
    i = t % (d.sx - 1) - 2;
    k = t / (d.sx - 1) - 2;
  } else if (m < fi + fj + fk) {
    const int t = m - fi - fj;
    k = d.nk + 3;
    i = t % (d.sx - 1) - 2;
    j = t / (d.sx - 1) - 2;
  } else {
    return;
  }
  const size_t o = (size_t)d.ix(i, j, k) * B;
  for (int b = 0; b < B; b += 4)
    *reinterpret_cast<uint32_t*>(dst + o + b) = *reinterpret_cast<const uint32_t*>(src + o + b);
}

void copy_shell(const Dims& d, void* dst, const void* src, int esize) {
  const int n = d.sy * d.sz + (d.sx - 1) * d.sz + (d.sx - 1) * (d.sy - 1);
  const unsigned nb = (unsigned)((n + 255) / 256);
  char* o = static_cast<char*>(dst);
  const char* i = static_cast<const char*>(src);
  if (esize == 16) copy_shell_kernel<16><<<nb, 256, 0, R().s>>>(d, o, i);
  else if (esize == 4) copy_shell_kernel<4><<<nb, 256, 0, R().s>>>(d, o, i);
  else copy_shell_kernel<8><<<nb, 256, 0, R().s>>>(d, o, i);
  FEL_CHECK(hipGetLastError());
}

bool memo_enabled() { return result_reuse(); }

MemoEntry* memo_find(const MemoKey& k) {
  for (MemoEntry& e : g_memo)
    if (e.key == k.w) return &e;
  return nullptr;
}

bool outs_registered(std::initializer_list<MemoOut> outs) {
  if (outs.size() > 4) return false;
  for (const MemoOut& o : outs)
    if (!whole_range(o.host, o.count)) return false;
  return true;
}
}  // namespace

MemoKey& MemoKey::content(const void* host, size_t count) {
  const uint64_t v = content_key(host, count);
  if (v == 0) ok = false;
  w.push_back(v);
  return *this;
}

MemoKey& MemoKey::shell(const void* host, size_t count) {
  const uint64_t v = shell_key(host, count);
  if (v == 0) ok = false;
  w.push_back(v);
  return *this;
}

bool memo_reuse(const MemoKey& k, std::initializer_list<MemoOut> outs, bool allow_copy) {
  if (!k.ok || !memo_enabled() || !outs_registered(outs)) return false;
  MemoEntry* e = memo_find(k);
  if (!e || e->nout != (int)outs.size()) return false;
  bool held[4];
  int m = 0;
  for (const MemoOut& o : outs) {
    Range* h = whole_range(reinterpret_cast<const void*>(e->host[m]), e->count[m]);
    Range* d = whole_range(o.host, o.count);
    held[m] = h && h->ver == e->ver[m] && e->count[m] == o.count && h->esize == d->esize &&
              (h == d || allow_copy);
    const bool parked = e->parked[m] && e->count[m] == o.count &&
                        e->pbytes[m] == d->bytes / sizeof(double) * (size_t)d->esize;
    if (!held[m] && !parked) return false;
    m++;
  }
  m = 0;
  for (const MemoOut& o : outs) {
    const void* src = reinterpret_cast<const void*>(e->host[m]);
    if (!held[m]) {
      // the parked buffer becomes the output's; the output's own contents
      // are parked in turn when they are a memoised result
      Range* d = whole_range(o.host, o.count);
      deferred_coef_check(o.host, true);
      materialize_dependents(*d);
      drop_lazy(*d);
      void* mine = e->parked[m];
      const size_t pb = e->pbytes[m];
      e->parked[m] = nullptr;
      e->pbytes[m] = 0;
      g_parked_total -= pb;
      if (!park_buffer(*d, e)) pool_put(d->dev, pb);
      d->dev = mine;
      d->ver = e->ver[m];
      d->subs.clear();
      d->shell = 0;
      e->host[m] = reinterpret_cast<uintptr_t>(o.host);
    } else if (src != o.host) {
      if (!deferred_copy(src, o.host, o.count)) {
        std::fprintf(stderr, "felgpu: internal: memo copy refused\n");
        std::abort();
      }
      copy_content_version(src, o.host, o.count);
      e->host[m] = reinterpret_cast<uintptr_t>(o.host);
    }
    m++;
  }
  e->used = ++g_memo_tick;
  g_memo_hits++;
  return true;
}

void memo_record(const MemoKey& k, std::initializer_list<MemoOut> outs) {
  if (!k.ok || !memo_enabled() || !outs_registered(outs)) return;
  MemoEntry* e = memo_find(k);
  if (e && e->nout == (int)outs.size()) {
    // the same function of the same inputs: the outputs carry the contents
    // numbers given to this key before
    unpark_all(*e);
    int m = 0;
    for (const MemoOut& o : outs) {
      whole_range(o.host, o.count)->ver = e->ver[m];
      e->host[m] = reinterpret_cast<uintptr_t>(o.host);
      e->count[m] = o.count;
      m++;
    }
    e->used = ++g_memo_tick;
    return;
  }
  if (!e) {
    if (g_memo.size() < kMemoEntries) {
      g_memo.emplace_back();
      e = &g_memo.back();
    } else {
      e = &g_memo[0];
      for (MemoEntry& x : g_memo)
        if (x.used < e->used) e = &x;
    }
  }
  unpark_all(*e);
  e->key = k.w;
  e->nout = (int)outs.size();
  int m = 0;
  for (const MemoOut& o : outs) {
    e->host[m] = reinterpret_cast<uintptr_t>(o.host);
    e->count[m] = o.count;
    e->ver[m] = whole_range(o.host, o.count)->ver;
    m++;
  }
  e->used = ++g_memo_tick;
}

void memo_park(const Dims& d, std::initializer_list<MemoOut> outs, bool box) {
  if (!memo_enabled()) return;
  for (const MemoOut& o : outs) {
    Range* g = whole_range(o.host, o.count);
    if (!g || g->lazy || g->zero) continue;
    deferred_coef_check(o.host, true);  // a recorded chain reading this buffer runs now
    materialize_dependents(*g);         // copies deferred on this buffer are made now
    void* old = g->dev;
    if (!park_buffer(*g, nullptr)) continue;
    const size_t bytes = g->bytes / sizeof(double) * (size_t)g->esize;
    g->dev = pool_get(bytes);
    if (box) copy_shell(d, g->dev, old, g->esize);  // the call writes the box only
  }
}

void memo_release() {
  for (MemoEntry& e : g_memo) unpark_all(e);
  g_memo.clear();
  for (auto& kv : g_pool) FEL_CHECK(hipFree(kv.second));
  g_pool.clear();
  g_parked_total = 0;
}

void* exchange_device_buffer(const void* host, size_t count, void* buf) {
  Range* g = whole_range(host, count);
  if (!g) {
    std::fprintf(stderr, "felgpu: internal: buffer exchange on an unregistered range\n");
    std::abort();
  }
  deferred_coef_check(host, true);
  // the whole contents are replaced
  materialize_dependents(*g);
  drop_lazy(*g);
  void* old = g->dev;
  g->dev = buf;
  return old;
}

// --------------------------------------------------------------------------
// Deferred chain of entry points
// --------------------------------------------------------------------------

DeferredChain& deferred_chain() {
  static DeferredChain w;
  return w;
}

bool chain_deferral_enabled() {
  static const int on = [] {
    const char* e = std::getenv("FELGPU_EAGER_CHAINS");
    return (e && std::atoi(e) != 0) ? 0 : 1;
  }();
  return on != 0;
}

static bool chain_holds(const std::vector<const void*>& v, const Range* g) {
  for (const void* h : v)
    if (find_range(R().reg, reinterpret_cast<uintptr_t>(h), 1) == g) return true;
  return false;
}

void deferred_chain_flush() {
  DeferredChain& w = deferred_chain();
  if (!w.active) return;
  w.active = false;
  std::function<void()> f;
  f.swap(w.flush);
  w.outs.clear();
  w.srcs.clear();
  if (f) f();
}

void deferred_chain_check(const void* host, bool write) {
  DeferredChain& w = deferred_chain();
  if (!w.active) return;
  const Range* g = find_range(R().reg, reinterpret_cast<uintptr_t>(host), 1);
  if (!g) return;
  if (chain_holds(w.outs, g) || (write && chain_holds(w.srcs, g))) deferred_chain_flush();
}

void deferred_coef_check(const void* host, bool write, bool whole) {
  deferred_chain_check(host, write);
  DeferredCoef& dc = deferred_coef();
  if (!dc.active) return;
  const uintptr_t p = reinterpret_cast<uintptr_t>(host);
  const size_t bytes = sizeof(double) * (size_t)dc.d.n3;
  for (int m = 0; m < 45; m++) {
    const uintptr_t o = reinterpret_cast<uintptr_t>(dc.out_host[m]);
    if (p >= o && p < o + bytes) {
      deferred_coef_flush();  // a coefficient array is accessed: materialise
      return;
    }
  }
  const uintptr_t s = reinterpret_cast<uintptr_t>(dc.src_host);
  if (write && whole && dc.src_host && p == s) {
    // the viscosity is overwritten as a whole: keep its buffer, give the
    // array another
    Range* g = whole_range(dc.src_host, (size_t)dc.d.n3);
    if (g && !g->lazy && !g->zero && g->dev == dc.src_dev) {
      const size_t b = (size_t)g->esize * (size_t)dc.d.n3;
      dc.owned = g->dev;
      dc.owned_bytes = b;
      g->dev = pool_get_buffer(b);
      dc.src_host = nullptr;
      return;
    }
  }
  if (write && dc.src_host && p >= s && p < s + bytes) {
    // the viscosity is about to change: keep the values the coefficients
    // are a function of (16 B/cell) instead of materialising 45 arrays
    const size_t dbytes = (size_t)dc.esize * (size_t)dc.d.n3;
    void* snap = scratch(7, dbytes);
    if (trace_copies())
      std::fprintf(stderr, "felgpu: viscosity snapshot %zu B, in %s (%s)\n", dbytes, g_trace_call,
                   g_trace_what);
    FEL_CHECK(hipMemcpyAsync(snap, dc.src_dev, dbytes, hipMemcpyDeviceToDevice, stream()));
    dc.src_dev = snap;
    dc.src_host = nullptr;
  }
}

// --------------------------------------------------------------------------
// Call: argument resolution
// --------------------------------------------------------------------------

Call::Call(const char* name) : name_(name) {
  ensure_init();
  g_trace_call = name;
  g_trace_what = "";
}

Call::~Call() {
  if (!finished_) finish();
}

void* Call::resolve(const void* host, size_t count, Access a, const char* what, int esize) {
  Runtime& r = R();
  g_trace_what = what;
  const uintptr_t p = reinterpret_cast<uintptr_t>(host);
  const size_t bytes = count * sizeof(double);
  if (host == nullptr) {
    std::fprintf(stderr, "felgpu: %s: null array argument %s\n", name_, what);
    std::abort();
  }
  {
    const Range* w = find_range(r.reg, p, bytes);
    const bool whole = (a & kOverwrite) == kOverwrite && w && p == w->host && bytes == w->bytes;
    deferred_coef_check(host, (a & kWrite) != 0, whole);
  }
  if (Range* g = find_range(r.reg, p, bytes)) {
    if (a == kInterior && p == g->host && bytes == g->bytes && g->lazy && g->lkind == kCopy) {
      // the interior is about to be rewritten: only the ghost cells of the
      // deferred copy are made
      trace_copy("ghost copy", *g, 0);
      copy_outside_interior(box_, g->dev, g->lazy->dev, g->esize);
      drop_lazy(*g);
      materialize_dependents(*g);
    } else if ((a & kOverwrite) == kOverwrite && p == g->host && bytes == g->bytes) {
      // every element is replaced: a deferred copy of the old contents is
      // not needed, only the ranges deferred on this one
      materialize_dependents(*g);
      drop_lazy(*g);
    } else {
      settle(*g, (a & kWrite) != 0);
    }
    if (g->esize != esize) {
      std::fprintf(stderr,
                   "felgpu: %s: array %s is held with %d-byte elements on the device, this call "
                   "needs %d-byte elements (register it with felgpu_register%s_)\n",
                   name_, what, g->esize, esize, esize == 8 ? "_fp64" : "");
      std::abort();
    }
    if (a & kWrite) note_write(*g, p, bytes);
    return static_cast<char*>(g->dev) + (p - g->host) / sizeof(double) * (size_t)esize;
  }
  // an argument that overlaps registered memory without lying inside one
  // range (e.g. a 4-D array whose slices are registered one by one): a
  // mirror of the host copy would read stale values
  {
    auto it = r.reg.upper_bound(p + bytes - 1);
    if (it != r.reg.begin() && (--it)->second.host + it->second.bytes > p) {
      std::fprintf(stderr,
                   "felgpu: %s: array %s (%zu bytes at %p) overlaps the registered range at %p "
                   "without lying inside it; pass each registered range on its own\n",
                   name_, what, bytes, host, reinterpret_cast<void*>(it->second.host));
      std::abort();
    }
  }
  // per-call mirror, in the element type the kernel reads
  const size_t dbytes = count * (size_t)esize;
  void* dev = nullptr;
  auto it = r.transient.find(p);
  if (it != r.transient.end() && it->second.bytes >= dbytes) {
    dev = it->second.dev;
  } else {
    if (it != r.transient.end()) {
      FEL_CHECK(hipStreamSynchronize(r.s));
      FEL_CHECK(hipFree(it->second.dev));
      r.transient.erase(it);
    }
    FEL_CHECK(hipMalloc(&dev, dbytes));
    r.transient[p] = Range{p, dbytes, dev, esize, 0, nullptr, 0};
  }
  h2d(dev, host, count, esize);
  if (a & kWrite) {
    if (npend_ >= 64) {
      std::fprintf(stderr, "felgpu: %s: too many unregistered output arrays\n", name_);
      std::abort();
    }
    pend_[npend_++] = Pending{const_cast<void*>(host), dev, count, esize};
  }
  return dev;
}

void Call::finish() {
  Runtime& r = R();
  finished_ = true;
  if (npend_ == 0) return;
  for (int m = 0; m < npend_; m++) d2h(pend_[m].host, pend_[m].dev, pend_[m].count, pend_[m].esize);
  FEL_CHECK(hipStreamSynchronize(r.s));
  npend_ = 0;
}

// --------------------------------------------------------------------------
// Event timers (opt-in; the benchmark reads them)
// --------------------------------------------------------------------------

static Runtime::Timer& prof_slot(int id) {
  Runtime& r = R();
  if ((int)r.timers.size() <= id) r.timers.resize(id + 1);
  Runtime::Timer& t = r.timers[id];
  if (t.used == (int)t.b.size()) {
    hipEvent_t a, e;
    FEL_CHECK(hipEventCreate(&a));
    FEL_CHECK(hipEventCreate(&e));
    t.b.push_back(a);
    t.e.push_back(e);
  }
  return t;
}

void prof_launch_events(int id, bool first, bool last, int launches, hipEvent_t* start,
                        hipEvent_t* stop) {
  *start = *stop = nullptr;
  if (!R().prof) return;
  Runtime::Timer& t = prof_slot(id);
  if (first) *start = t.b[t.used];
  if (last) {
    *stop = t.e[t.used];
    if ((int)t.span.size() <= t.used) t.span.resize(t.used + 1);
    t.span[t.used] = launches;
    t.used++;
  }
}

static void prof_collect() {
  Runtime& r = R();
  if (!r.init) return;
  FEL_CHECK(hipStreamSynchronize(r.s));
  for (auto& t : r.timers) {
    for (int m = 0; m < t.used; m++) {
      float ms = 0.f;
      FEL_CHECK(hipEventElapsedTime(&ms, t.b[m], t.e[m]));
      t.ms += ms;
      t.n += t.span[m];
    }
    t.used = 0;
  }
}

void count_stash_hit() { g_stash_hits++; }

}  // namespace fel

using namespace fel;

extern "C" {

// Selects the device and creates the library's stream.  Optional: every
// entry point initialises on first use.
void felgpu_init_(const int* device) {
  if (device && *device >= 0 && !R().init) {
    char buf[16];
    std::snprintf(buf, sizeof buf, "%d", *device);
    setenv("FELGPU_DEVICE", buf, 1);
  }
  ensure_init();
}

static void register_range(const void* host, const int64_t* nbytes, const int* upload, int esize) {
  ensure_init();
  Runtime& r = R();
  const uintptr_t p = reinterpret_cast<uintptr_t>(host);
  const size_t bytes = (size_t)*nbytes;
  if (bytes % sizeof(double) != 0) {
    std::fprintf(stderr, "felgpu: host range %p: size %zu is not a whole number of doubles\n",
                 host, bytes);
    std::abort();
  }
  const size_t n = bytes / sizeof(double);
  auto it = r.reg.find(p);
  if (it != r.reg.end()) {
    if (it->second.bytes != bytes || it->second.esize != esize) {
      std::fprintf(stderr, "felgpu: host range %p registered twice with a different size or "
                           "precision\n", host);
      std::abort();
    }
    // registered again: the contents are replaced, as by felgpu_upload_ --
    // records reading the range settle, ranges deferring to it take their
    // contents, and its own deferred state is dropped
    deferred_coef_check(host, true);
    materialize_dependents(it->second);
    drop_lazy(it->second);
  } else {
    // overlap check
    auto nx = r.reg.lower_bound(p);
    if (nx != r.reg.end() && nx->second.host < p + bytes) {
      std::fprintf(stderr, "felgpu: host range %p overlaps a registered range\n", host);
      std::abort();
    }
    if (nx != r.reg.begin()) {
      auto pv = std::prev(nx);
      if (pv->second.host + pv->second.bytes > p) {
        std::fprintf(stderr, "felgpu: host range %p overlaps a registered range\n", host);
        std::abort();
      }
    }
    void* dev = nullptr;
    FEL_CHECK(hipMalloc(&dev, n * (size_t)esize));
    it = r.reg.emplace(p, Range{p, bytes, dev, esize, 0, nullptr, 0}).first;
  }
  note_write(it->second);
  if (*upload)
    h2d(it->second.dev, host, n, esize);
  else
    FEL_CHECK(hipMemsetAsync(it->second.dev, 0, n * (size_t)esize, r.s));
  FEL_CHECK(hipStreamSynchronize(r.s));
}

// Makes [host, host+nbytes) device-resident.  upload != 0 copies the host
// contents; upload == 0 zero-fills the device copy (the state of a freshly
// allocated, never-written host array on this platform: large allocations
// come from zero pages).  In the FP32 mode the device copy is FP32.
void felgpu_register_(const void* host, const int64_t* nbytes, const int* upload) {
  ensure_init();
  register_range(host, nbytes, upload, R().tangent ? 16 : R().flow_esize);
}

// As felgpu_register_, with an FP64 device copy in either precision (a Dual
// copy in the tangent mode).
void felgpu_register_fp64_(const void* host, const int64_t* nbytes, const int* upload) {
  ensure_init();
  register_range(host, nbytes, upload, R().tangent ? 16 : 8);
}

static Range& must_find(const void* host, const char* who) {
  Range* g = find_range(R().reg, reinterpret_cast<uintptr_t>(host), 1);
  if (!g) {
    std::fprintf(stderr, "felgpu: %s: %p is not inside a registered range\n", who, host);
    std::abort();
  }
  return *g;
}

// Selects the device precision of arrays registered with felgpu_register_
// (64 or 32); only while nothing is registered.
void felgpu_set_precision_(const int* bits) {
  ensure_init();
  Runtime& r = R();
  if (*bits != 32 && *bits != 64) {
    std::fprintf(stderr, "felgpu: felgpu_set_precision_: %d bits is not 32 or 64\n", *bits);
    std::abort();
  }
  if (!r.reg.empty()) {
    std::fprintf(stderr, "felgpu: felgpu_set_precision_: arrays are registered\n");
    std::abort();
  }
  if (*bits == 32 && r.tangent) {
    std::fprintf(stderr, "felgpu: felgpu_set_precision_: the tangent mode is FP64\n");
    std::abort();
  }
  r.flow_esize = *bits / 8;
}

int felgpu_precision_() { return flow_fp32() ? 32 : 64; }

// Switches the tangent mode on (1) or off (0); only while nothing is
// registered, and only in FP64.
void felgpu_set_tangent_(const int* on) {
  ensure_init();
  Runtime& r = R();
  if (!r.reg.empty()) {
    std::fprintf(stderr, "felgpu: felgpu_set_tangent_: arrays are registered\n");
    std::abort();
  }
  if (*on && r.flow_esize != 8) {
    std::fprintf(stderr, "felgpu: felgpu_set_tangent_: the tangent mode needs FP64 precision\n");
    std::abort();
  }
  r.tangent = *on != 0;
}

int felgpu_tangent_() { return tangent_mode() ? 1 : 0; }

// Turns the reuse of results already computed (parked results, the result
// memo) on or off; the default is on unless FELGPU_NO_RESULT_REUSE=1.
// Results are the same either way.
void felgpu_set_result_reuse_(const int* on) {
  result_reuse();
  g_reuse = *on != 0 ? 1 : 0;
}

// The current reuse setting (1 on, 0 off).
int felgpu_result_reuse_() { return result_reuse() ? 1 : 0; }

// Number of entry-point calls answered from the result memo so far.
int64_t felgpu_result_reuse_hits_() { return g_memo_hits; }

// Number of entry-point calls answered from a pass of an earlier call.
int64_t felgpu_stash_hits_() { return g_stash_hits; }

// Sets the derivative part of a whole registered array from the host array
// `seed` (same number of elements); the values are kept.
void felgpu_tangent_seed_(const void* host, const double* seed) {
  deferred_coef_check(host, true);
  Range& g = must_find(host, "felgpu_tangent_seed_");
  settle(g, true);
  if (g.esize != 16) {
    std::fprintf(stderr, "felgpu: felgpu_tangent_seed_: not in the tangent mode\n");
    std::abort();
  }
  const size_t n = g.bytes / sizeof(double);
  std::vector<Dual> tmp(n);
  FEL_CHECK(hipMemcpyAsync(tmp.data(), g.dev, n * 16, hipMemcpyDeviceToHost, stream()));
  FEL_CHECK(hipStreamSynchronize(stream()));
  for (size_t m = 0; m < n; m++) tmp[m].d = seed[m];
  FEL_CHECK(hipMemcpyAsync(g.dev, tmp.data(), n * 16, hipMemcpyHostToDevice, stream()));
  FEL_CHECK(hipStreamSynchronize(stream()));
  note_write(g);
}

// Copies the derivative part of a whole registered array into `out`.
void felgpu_tangent_fetch_(const void* host, double* out) {
  deferred_coef_check(host, false);
  Range& g = must_find(host, "felgpu_tangent_fetch_");
  settle(g, false);
  if (g.esize != 16) {
    std::fprintf(stderr, "felgpu: felgpu_tangent_fetch_: not in the tangent mode\n");
    std::abort();
  }
  const size_t n = g.bytes / sizeof(double);
  std::vector<Dual> tmp(n);
  FEL_CHECK(hipMemcpyAsync(tmp.data(), g.dev, n * 16, hipMemcpyDeviceToHost, stream()));
  FEL_CHECK(hipStreamSynchronize(stream()));
  for (size_t m = 0; m < n; m++) out[m] = tmp[m].d;
}

// Host -> device copy of a whole registered range (after host-side writes).
void felgpu_upload_(const void* host) {
  deferred_coef_check(host, true);
  Range& g = must_find(host, "felgpu_upload_");
  materialize_dependents(g);
  drop_lazy(g);  // the whole contents are replaced
  note_write(g);
  h2d(g.dev, (const void*)g.host, g.bytes / sizeof(double), g.esize);
  FEL_CHECK(hipStreamSynchronize(stream()));
}

// Device -> host copy of a whole registered range (before host-side reads).
void felgpu_download_(void* host) {
  deferred_coef_check(host, false);
  Range& g = must_find(host, "felgpu_download_");
  settle(g, false);
  d2h((void*)g.host, g.dev, g.bytes / sizeof(double), g.esize);
  FEL_CHECK(hipStreamSynchronize(stream()));
}

void felgpu_unregister_(const void* host) {
  deferred_coef_check(host, true);
  Runtime& r = R();
  auto it = r.reg.find(reinterpret_cast<uintptr_t>(host));
  if (it == r.reg.end()) return;
  materialize_dependents(it->second);
  drop_lazy(it->second);
  FEL_CHECK(hipStreamSynchronize(r.s));
  FEL_CHECK(hipFree(it->second.dev));
  r.reg.erase(it);
}

void felgpu_unregister_all_() {
  Runtime& r = R();
  if (!r.init) return;
  deferred_chain_flush();
  deferred_coef_flush();
  FEL_CHECK(hipStreamSynchronize(r.s));
  weno_release();
  rn_release();
  dense_src_release();
  memo_release();
  for (auto& kv : r.reg) FEL_CHECK(hipFree(kv.second.dev));
  r.reg.clear();
  for (auto& kv : r.transient) FEL_CHECK(hipFree(kv.second.dev));
  r.transient.clear();
}

// Device address of a registered host address (for callers that launch their
// own work on the library stream, e.g. the test harness).
void* felgpu_device_ptr_(const void* host) {
  deferred_coef_check(host, true);
  Range& g = must_find(host, "felgpu_device_ptr_");
  settle(g, true);  // the caller may read or write it directly
  note_write(g);    // so its contents count as changed
  return static_cast<char*>(g.dev) +
         (reinterpret_cast<uintptr_t>(host) - g.host) / sizeof(double) * (size_t)g.esize;
}

void felgpu_sync_() { sync(); }

void* felgpu_stream_() { return (void*)stream(); }

// Event timers: enable/disable, read (total ms, count) of timer id, reset.
void felgpu_prof_enable_(const int* on) {
  ensure_init();
  R().prof = (*on != 0);
}

void felgpu_prof_read_(const int* id, double* ms_total, int64_t* count) {
  Runtime& r = R();
  prof_collect();
  if (*id >= 0 && *id < (int)r.timers.size()) {
    *ms_total = r.timers[*id].ms;
    *count = r.timers[*id].n;
  } else {
    *ms_total = 0.0;
    *count = 0;
  }
}

void felgpu_prof_reset_() {
  Runtime& r = R();
  prof_collect();
  for (auto& t : r.timers) {
    t.ms = 0.0;
    t.n = 0;
  }
}

}  // extern "C"
