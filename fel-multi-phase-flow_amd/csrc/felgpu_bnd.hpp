// felgpu_bnd.hpp -- ghost-fill operations shared by the boundary entry
// points and the viscous Gauss-Seidel solve (which refreshes ghosts between
// colour pairs, srcs/solu_sor4.f:176-180).
#pragma once

#include "felgpu.hpp"

namespace fel {

// Navier-slip ghost coefficients of srcs/bndu.f:39-47 (73-81) at wall cell
// (i, jw, k): ghost = c1 * (-+uwall) + c2 * (first interior value), with the
// slip length ls = 1/(phi_av/l1 + (1-phi_av)/l2) of the face-averaged phi.
template <class P>
struct WallCoef {
  P c1, c2;
};

template <class P>
__device__ __forceinline__ WallCoef<P> wall_coef(const Dims& d, int i, int jw, int k, double dy,
                                                 const P* l1, const P* l2, const P* phi) {
  const int ip = (i + 1) % (d.ni + 3);
  const int cw = d.ix(i, jw, k);
  const P phi_av = (phi[cw] + phi[d.ix(ip, jw, k)]) / 2.0;
  const P inv1 = phi_av / l1[cw];
  const P inv2 = (1.0 - phi_av) / l2[cw];
  const P ls = 1.0 / (inv1 + inv2);
  return WallCoef<P>{(2.0 * dy) / (2.0 * ls + dy), (2.0 * ls - dy) / (2.0 * ls + dy)};
}

// bndu (srcs/bndu.f) followed by bnd_periodic of component `swept` (0,1,2),
// one composed launch; swept = -1 is bndu alone.  T: velocity element type,
// P: element type of the slip lengths and the volume fraction.
// `comps` (bit q for component q) restricts the writes to those components;
// a component left out must already hold exactly the values the call would
// write (its interior unchanged since a refresh that wrote them).  `lean`
// leaves out the deep wall rows (j = -2, -1, nj+2, nj+3), which no stencil
// of the viscous solve reads; a caller using it rewrites them afterwards.
template <class T, class P>
void bc_fused(const Dims d, T* u, T* v, T* w, double uwall, double dy, const P* l1, const P* l2,
              const P* phi, bool lo, bool hi, int swept, int comps, bool lean = false);
// srcs/bnd_periodic.f:4-46 in one in-place pass.
template <class T>
void periodic(const Dims d, T* q);
template <class T>
void neumann(const Dims d, T* q, bool lo, bool hi);

}  // namespace fel
