// lucore_cpu.hpp -- the reference's C++ host API for the CPU determinant
// path, with the same names and argument lists on complex128 (double2):
//
//   LUCoreCPU::CALL         LU-fqhe/inc/LUCoreCPU.h:26,
//                           LU-fqhe/src/LUCoreCPU.cu:39-107
//   LUCoreCPU::determinant  LU-fqhe/inc/LUCoreCPU.h:27,
//                           LU-fqhe/src/LUCoreCPU.cu:112-267
//
// The reference's Monte Carlo loop gives the CPU share of each iteration's
// matrices to CALL (LU-fqhe/src/MonteCarloLoop.cu:343).  Both functions run
// the reference's Crout algorithm (row scaling on |Re|, the `>=` pivot scan,
// full row swaps, complex-divide column scaling, sequential diagonal product)
// in FP64 with the reference's rounding steps, implemented over the C ABI of
// lufqhe.h (lufqhe_cpu_crout_det, lufqhe_cpu_pair_sum).  Where the reference
// exit()s (a row with max |Re| < 1e-24, LU-fqhe/src/LUCoreCPU.cu:158) these
// throw LUCoreCPU::Error.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "lufqhe.h"
#include "run_info.hpp"

namespace LUCoreCPU {

class Error : public std::runtime_error {
 public:
  Error(const std::string& what, int status) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

// *result = sum over i < num/2 of det(M_2i) * det(M_2i+1) for the num
// matrices at `matrices` (row stride loadSize, matrix stride
// matSide*loadSize), on cpuCnt threads.  The pair products are added in
// matrix order, so the sum does not depend on cpuCnt (the reference adds
// per-thread partial sums in completion order).  The matrices are left
// unchanged (the reference leaves LU factors in them; nothing reads them).
// rank is only used in error messages, as in the reference.
void CALL(double2* matrices, double2* result, int num, RunInfo* rInfo, int rank, int cpuCnt);

// *det = det of the matSide x matSide matrix at inputMatrix (row stride
// loadSize).  As in the reference, the matrix is factored in place: its LU
// factors (unit-diagonal L below, U on and above the diagonal, rows in pivot
// order) are left there.
void determinant(double2* inputMatrix, double2* det, int matSide, int loadSize);

}  // namespace LUCoreCPU
