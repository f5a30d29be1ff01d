// k_out.hip -- device parts of the output cadences: the j-slice of
// srcs/mk_slice_j.f, the vertex values of the VTK writers srcs/mkvtk_phi.f
// and srcs/mkvtk_p.f, and the wall-row interface crossings of
// srcs/find_interface_positions*.f90.  Only the reduced data leaves the
// device (a slice, the vertex values, a few crossings); the files are
// formatted by the Fortran host (host/fel_driver.f90).
#include "felgpu.hpp"
#include "felgpu_visc.hpp"

namespace fel {

constexpr int kZMinus = 2;  // nID(Z_MINUS) as 0-based (srcs/param.h)

// srcs/mk_slice_j.f:25-53: the 8-point average of q at plane j with the
// offsets (ip, jp, kp), on i = 0..ni, k = kst..nk (kst = 0 without a
// Z_MINUS neighbour); the work slice is cleared first.  Values of q of any
// element type, averaged in FP64 in the reference's order.
template <class T>
void slice_j(const Dims d, int j, int ip, int jp, int kp, int kst, const T* q, double* wrk,
             double* out) {
  const int ni = d.ni, nk = d.nk;
  launch_flat((ni + 1) * (nk + 1), [=] __device__(int m) {
    const int i = m % (ni + 1), k = m / (ni + 1);
    wrk[m] = 0.0;
    if (k < kst) return;
    const int X = ip, Y = jp * d.sx, Z = kp * d.sx * d.sy;
    const int c = d.ix(i, j, k);
    const double s = (val(q[c]) + val(q[c + X]) + val(q[c + Y]) + val(q[c + X + Y]) +
                      val(q[c + Z]) + val(q[c + X + Z]) + val(q[c + Y + Z]) +
                      val(q[c + X + Y + Z])) *
                     0.125;
    wrk[m] = s;
    out[m] = s;
  });
}

// srcs/mkvtk_phi.f:15-80 and srcs/mkvtk_p.f:15-84: the writers first fill
// the ghost layer 0 / n+1 of their (gathered) copy -- x periodic, y Neumann,
// z Neumann (phi) or periodic (p), in the order x, y, z -- then write the
// 8-point average at every vertex (i,j,k) in [0,n]^3.  Each fill copies
// along one axis from a cell the earlier fills left final, so the filled
// value at (i,j,k) is q(mx(i), my(j), mz(k)) with the per-axis maps below;
// the kernel reads through those maps and leaves q untouched.
__device__ __forceinline__ int map_periodic(int i, int n) { return i == 0 ? n : (i == n + 1 ? 1 : i); }
__device__ __forceinline__ int map_neumann(int i, int n) { return i == 0 ? 1 : (i == n + 1 ? n : i); }

template <class T>
void vtk_points(const Dims d, int kind, const T* q, double* out) {
  const int ni = d.ni, nj = d.nj, nk = d.nk;
  const int nx = ni + 1, ny = nj + 1;
  launch_flat(nx * ny * (nk + 1), [=] __device__(int m) {
    const int i = m % nx, r = m / nx, j = r % ny, k = r / ny;
    auto at = [&](int a, int b, int c) {
      const int mz = kind == 0 ? map_neumann(c, nk) : map_periodic(c, nk);
      return val(q[d.ix(map_periodic(a, ni), map_neumann(b, nj), mz)]);
    };
    out[m] = (+at(i, j, k) + at(i + 1, j, k) + at(i, j + 1, k) + at(i + 1, j + 1, k) +
              at(i, j, k + 1) + at(i + 1, j, k + 1) + at(i, j + 1, k + 1) +
              at(i + 1, j + 1, k + 1)) *
             0.125;
  });
}

// srcs/find_interface_positions_upper.f90:13-24 / _lower.f90:17-30 on row
// jrow: one thread per k scans i = 1..ni for phi(i) > 0.5 > phi(i+1) and
// stores x = (i-1/2) dx + (1/2 - phi1) dx/(phi2 - phi1) - 3/4 xl of every
// crossing (first_only: of the first) at x[k*ni + n], n = 0..cnt[k]-1.
template <class T>
void crossings(const Dims d, int jrow, double dx, double xl, int first_only, const T* q, double* x,
               int* cnt) {
  const int ni = d.ni;
  launch_flat(d.nk, [=] __device__(int kk) {
    const int k = kk + 1;
    int n = 0;
    for (int i = 1; i <= ni; i++) {
      const double p1 = val(q[d.ix(i, jrow, k)]), p2 = val(q[d.ix(i + 1, jrow, k)]);
      if (p1 > 0.5 && p2 < 0.5) {
        x[kk * ni + n] = ((double)i - 0.5) * dx + (0.5 - p1) * dx / (p2 - p1) - xl * 0.75;
        n++;
        if (first_only) break;
      }
    }
    cnt[kk] = n;
  });
}

}  // namespace fel

using namespace fel;

extern "C" {

// srcs/mk_slice_j.f:1-2 (ipara must be 0).  wrk_slice and qk_slice are
// (0:ni, 0:nk) host arrays.
void mk_slice_j_(const int* ipara, const int* ID, const int* nID, const int* ndiv, const int* ni,
                 const int* nj, const int* nk, const int* j, const int* ip, const int* jp,
                 const int* kp, double* wrk_slice, const double* qk, double* qk_slice) {
  (void)ID;
  if (*ipara != 0 || *ndiv != 1) {
    std::fprintf(stderr, "felgpu: mk_slice_j_: only the single-domain case is supported\n");
    std::abort();
  }
  const Dims d = make_dims(*ni, *nj, *nk);
  if (*j < -2 || *j + *jp > *nj + 3) {
    std::fprintf(stderr, "felgpu: mk_slice_j_: plane j=%d outside the array\n", *j);
    std::abort();
  }
  const size_t ns = (size_t)(*ni + 1) * (*nk + 1);
  const int kst = nID[kZMinus] < 0 ? 0 : 1;
  with_elem(elem_size(qk), [&](auto t) {
    using T = decltype(t);
    Call c("mk_slice_j_");
    const T* q = c.in_t<T>(qk, d.n3, "qk");
    double* w = c.out_t<double>(wrk_slice, ns, "wrk_slice");
    double* o = c.io_t<double>(qk_slice, ns, "qk_slice");
    slice_j<T>(d, *j, *ip, *jp, *kp, kst, q, w, o);
    c.finish();
  });
}

// Vertex values of the reference's VTK writers: kind 0 = srcs/mkvtk_phi.f
// (x periodic, y and z Neumann), kind 1 = srcs/mkvtk_p.f (x and z periodic,
// y Neumann); out[(ni+1)(nj+1)(nk+1)], i fastest.  q is not modified.
void felgpu_vtk_points_(const int* ni, const int* nj, const int* nk, const double* q,
                        const int* kind, double* out) {
  const Dims d = make_dims(*ni, *nj, *nk);
  const size_t n = (size_t)(*ni + 1) * (*nj + 1) * (*nk + 1);
  with_elem(elem_size(q), [&](auto t) {
    using T = decltype(t);
    Call c("felgpu_vtk_points_");
    const T* dq = c.in_t<T>(q, d.n3, "q");
    double* o = c.out_t<double>(out, n, "out");
    vtk_points<T>(d, *kind, dq, o);
    c.finish();
  });
}

// Interface crossings on row jrow (srcs/find_interface_positions*.f90):
// x[k*ni + m], m < cnt[k], for k = 1..nk (x: ni*nk doubles, cnt: nk ints);
// *first_only != 0 keeps the first crossing of each row (the _lower form).
void felgpu_interface_crossings_(const int* ni, const int* nj, const int* nk, const double* phi,
                                 const int* jrow, const double* dx, const double* xl,
                                 const int* first_only, double* x, int* cnt) {
  const Dims d = make_dims(*ni, *nj, *nk);
  if (*jrow < -2 || *jrow > *nj + 3) {
    std::fprintf(stderr, "felgpu: felgpu_interface_crossings_: row j=%d outside the array\n", *jrow);
    std::abort();
  }
  with_elem(elem_size(phi), [&](auto t) {
    using T = decltype(t);
    Call c("felgpu_interface_crossings_");
    const T* q = c.in_t<T>(phi, d.n3, "phi");
    double* dx_ = c.out_t<double>(x, (size_t)d.ni * d.nk, "x");
    // the counts are ints: carried in a double-sized scratch and copied back
    int* dc = static_cast<int*>(scratch(16, sizeof(int) * (size_t)d.nk));
    crossings<T>(d, *jrow, *dx, *xl, *first_only, q, dx_, dc);
    FEL_CHECK(hipMemcpyAsync(cnt, dc, sizeof(int) * (size_t)d.nk, hipMemcpyDeviceToHost, stream()));
    c.finish();
    sync();
  });
}

}  // extern "C"
