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

/* The 45 coefficient arrays of srcs/cal_arith_coef_vis.f (and of
 * srcs/solu_sor4.f's argument list), in argument order. */
enum {
  CU_W, CU_E, CU_S, CU_N, CU_B, CU_T, CU_P, CU_VWS, CU_VES, CU_VWN, CU_VEN, CU_WWB, CU_WEB, CU_WWT, CU_WET,
  CV_W, CV_E, CV_S, CV_N, CV_B, CV_T, CV_P, CV_USW, CV_UNW, CV_USE, CV_UNE, CV_WSB, CV_WNB, CV_WST, CV_WNT,
  CW_W, CW_E, CW_S, CW_N, CW_B, CW_T, CW_P, CW_UBW, CW_UTW, CW_UBE, CW_UTE, CW_VBS, CW_VTS, CW_VBN, CW_VTN,
  NCOEF
};

#endif
