// felgpu_bnd.hpp -- ghost-fill operations shared by the boundary entry
// points and the viscous Gauss-Seidel solve (which refreshes ghosts between
// colour pairs, srcs/solu_sor4.f:176-180).
#pragma once

#include "felgpu.hpp"

namespace fel {

// srcs/bndu.f in one launch (composed passes, see k_bnd.hip).
void bndu(const Dims d, double* u, double* v, double* w, double uwall, double dy, const double* l1,
          const double* l2, const double* phi, bool lo, bool hi);
// srcs/bndu.f as the reference's three ordered passes (cross-check of bc_fused).
void bndu_phased(const Dims d, double* u, double* v, double* w, double uwall, double dy,
                 const double* l1, const double* l2, const double* phi, bool lo, bool hi);
// bndu followed by bnd_periodic of component `swept` (0,1,2), one launch;
// swept = -1 is bndu alone.
void bc_fused(const Dims d, double* u, double* v, double* w, double uwall, double dy, const double* l1,
              const double* l2, const double* phi, bool lo, bool hi, int swept);
// srcs/bnd_periodic.f:4-46 in one in-place pass.
void periodic(const Dims d, double* q);
void neumann(const Dims d, double* q, bool lo, bool hi);

}  // namespace fel
