// run_info.hpp -- the sizes of one run that the determinant path reads: the
// part of the reference's RunInfo (LU-fqhe/inc/RunInfo.h,
// LU-fqhe/src/RunInfo.cpp:46-129) used by LUCoreGPU::CALL and
// LUCoreCPU::CALL, with the same getters.
#pragma once

// The reference derives loadSize from the __PADDED__ build macro
// (LU-fqhe/src/RunInfo.cpp:72-76); here it is a constructor flag.
class RunInfo {
 public:
  RunInfo(int nMperBl, int matSide, int nElem, bool padded = false, int nodes = 1);

  int getNMperBl() const { return nMperBl_; }
  int getMatSide() const { return matSide_; }
  int getLoadSize() const { return loadSize_; }
  int getNElem() const { return nElem_; }
  int getNElemPnode() const { return nElemPNode_; }
  int getScrSpace() const { return scrSpace_; }
  int getScrSpace2() const { return scrSpace2_; }

  // Smallest power of two >= n (LU-fqhe/src/RunInfo.cpp:119-129).
  static int p2Above(int n);

 private:
  int nMperBl_, matSide_, nElem_, nElemPNode_, scrSpace_, scrSpace2_, loadSize_;
};
