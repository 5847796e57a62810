// lucore_gpu.hpp -- the reference's C++ host API for the GPU determinant
// path, with the same names and argument lists, on complex128 (double2) and
// HIP streams, implemented over the C ABI of lufqhe.h.
//
//   LUCoreGPU::CALL             LU-fqhe/inc/LUCoreGPU.h:47,
//                               LU-fqhe/src/LUCoreGPU.cu:38-106
//   LUCoreGPU::globalReduction  LU-fqhe/inc/LUCoreGPU.h:50,
//                               LU-fqhe/src/LUCoreGPU.cu:113-156
//   RunInfo                     run_info.hpp
//
// Semantics follow the reference build (WRITE_DET, PERFORM_PROD,
// LU-fqhe/Makefile:3): CALL leaves det(M_m) in M_m[0][0] and, for even m,
// det(M_m) * det(M_m+1); globalReduction sums the even [0][0] entries into
// d_Reduction[th_num][0].  Errors throw LUCoreGPU::Error (the reference
// exit()s, LU-fqhe/src/LUCoreGPU.cu:102-105).
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "lufqhe.h"
#include "run_info.hpp"

namespace LUCoreGPU {

class Error : public std::runtime_error {
 public:
  Error(const std::string& what, int status) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

// Pivot policy used by CALL (default LUFQHE_POLICY_CROUT_REF, the rule of
// the reference CPU Crout, LU-fqhe/src/LUCoreCPU.cu:145-213).
void setPivotPolicy(int policy);

// Determinants of the nMat matrices at d_Matrices[th_num] (row stride
// loadSize, matrix stride matSide*loadSize), written in place as above, on
// *stream.  rank is only used in error messages, as in the reference.
void CALL(double2** d_Matrices, int nMat, RunInfo* rInfo, int rank, int th_num,
          hipStream_t* stream);

// d_Reduction[th_num][0] = sum over i < nMat/2 of d_Matrices[th_num] M_2i[0][0].
// The matrix stride is the one the last CALL on this thread used.
void globalReduction(double2** d_Matrices, double2** d_Reduction, int nMat, int rank,
                     int th_num, hipStream_t stream);

}  // namespace LUCoreGPU
