/*
 * fel_oracle.h -- CPU restatement of the per-timestep solver of
 * toruneko1129/fel-multi-phase-flow (Fortran, srcs/ Fortran sources).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * host driver, the Python package) includes, links or calls this code.  It is
 * the checker that the GPU path is compared against in tests/, in
 * __graft_entry__.smoke() and in the cpu_baseline leg of bench.py.
 *
 * Every routine keeps the argument list of the Fortran subroutine it
 * restates (all arguments by reference, column-major arrays with bounds
 * (-2:ni+3,-2:nj+3,-2:nk+3)), under the name orc_<subroutine>, so one test
 * harness can call the oracle, the compiled reference and the GPU library
 * with identical actual arguments.
 */
#ifndef FEL_ORACLE_H
#define FEL_ORACLE_H

#include <math.h>
#include <stddef.h>

/* Flat offset of (i,j,k) in an array declared (-2:ni+3,-2:nj+3,-2:nk+3). */
#define OIX(i, j, k) \
  ((ptrdiff_t)((i) + 2) + (ptrdiff_t)sx * ((ptrdiff_t)((j) + 2) + (ptrdiff_t)sy * (ptrdiff_t)((k) + 2)))

#define ODIMS(ni, nj, nk)              \
  const ptrdiff_t sx = (ni) + 6;       \
  const ptrdiff_t sy = (nj) + 6;       \
  const ptrdiff_t sz = (nk) + 6;       \
  const ptrdiff_t n3 = sx * sy * sz;   \
  (void)n3

/* Wall/periodic neighbour table of the reference (srcs/param.h:1-8). */
enum { ORC_X_MINUS = 1, ORC_Y_MINUS = 2, ORC_Z_MINUS = 3,
       ORC_X_PLUS = 4, ORC_Y_PLUS = 5, ORC_Z_PLUS = 6 };

/* Cell normal = mean of the 8 surrounding vertex gradients
 * (srcs/solphi_mthinc1.f:81-107, same sums in srcs/solphi_mthinc3.f:47-73). */
static inline void cell_normal(ptrdiff_t sx, ptrdiff_t sy, const double *phix, const double *phiy,
                               const double *phiz, int i, int j, int k, double *nx, double *ny,
                               double *nz) {
  const double *a[3] = {phix, phiy, phiz};
  double r[3];
  for (int c = 0; c < 3; c++) {
    const double *p = a[c];
    r[c] = (+p[OIX(i - 1, j - 1, k - 1)] + p[OIX(i, j - 1, k - 1)] + p[OIX(i - 1, j, k - 1)] +
            p[OIX(i, j, k - 1)] + p[OIX(i - 1, j - 1, k)] + p[OIX(i, j - 1, k)] +
            p[OIX(i - 1, j, k)] + p[OIX(i, j, k)]) *
           0.125;
  }
  *nx = r[0];
  *ny = r[1];
  *nz = r[2];
}

/* srcs/mthinc_x.h:1-14 followed by srcs/mthinc_tanh.h:1-9: 8-point Gauss sum
 * of wght/(1+exp(-2 bet (n.x+d))). */
static inline double mthinc_tanh_sum(double rnx, double rny, double rnz, double xg1, double xg2,
                                     double yg1, double yg2, double zg1, double zg2, double wght,
                                     double bet, double ddd, double xs[8]) {
  const double x1 = rnx * xg1, x2 = rnx * xg2, y1 = rny * yg1, y2 = rny * yg2;
  const double z1 = rnz * zg1, z2 = rnz * zg2;
  xs[0] = x1 + y1 + z1;
  xs[1] = x2 + y1 + z1;
  xs[2] = x1 + y2 + z1;
  xs[3] = x2 + y2 + z1;
  xs[4] = x1 + y1 + z2;
  xs[5] = x2 + y1 + z2;
  xs[6] = x1 + y2 + z2;
  xs[7] = x2 + y2 + z2;
  double s = 0.0;
  for (int m = 0; m < 8; m++) {
    const double t = wght / (1.0 + exp(-(2.0 * bet * (xs[m] + ddd))));
    s = (m == 0) ? t : s + t;
  }
  return s;
}

/* The 45 coefficient arrays of srcs/cal_arith_coef_vis.f (and of
 * srcs/solu_sor4.f's argument list), in argument order. */
enum {
  CU_W, CU_E, CU_S, CU_N, CU_B, CU_T, CU_P, CU_VWS, CU_VES, CU_VWN, CU_VEN, CU_WWB, CU_WEB, CU_WWT, CU_WET,
  CV_W, CV_E, CV_S, CV_N, CV_B, CV_T, CV_P, CV_USW, CV_UNW, CV_USE, CV_UNE, CV_WSB, CV_WNB, CV_WST, CV_WNT,
  CW_W, CW_E, CW_S, CW_N, CW_B, CW_T, CW_P, CW_UBW, CW_UTW, CW_UBE, CW_UTE, CW_VBS, CW_VTS, CW_VBN, CW_VTN,
  NCOEF
};

#endif
