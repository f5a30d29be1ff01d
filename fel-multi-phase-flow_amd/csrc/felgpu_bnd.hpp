// felgpu_bnd.hpp -- ghost-fill operations shared by the boundary entry
// points and the viscous Gauss-Seidel solve (which refreshes ghosts between
// colour pairs, srcs/solu_sor4.f:176-180).
#pragma once

#include "felgpu.hpp"

namespace fel {

// bndu (srcs/bndu.f) followed by bnd_periodic of component `swept` (0,1,2),
// one composed launch; swept = -1 is bndu alone.  T: velocity element type,
// P: element type of the slip lengths and the volume fraction.
// `comps` (bit q for component q) restricts the writes to those components;
// a component left out must already hold exactly the values the call would
// write (its interior unchanged since a refresh that wrote them).
template <class T, class P>
void bc_fused(const Dims d, T* u, T* v, T* w, double uwall, double dy, const P* l1, const P* l2,
              const P* phi, bool lo, bool hi, int swept, int comps);
// srcs/bnd_periodic.f:4-46 in one in-place pass.
template <class T>
void periodic(const Dims d, T* q);
template <class T>
void neumann(const Dims d, T* q, bool lo, bool hi);

}  // namespace fel
