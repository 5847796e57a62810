// run_info.cpp -- RunInfo (include/run_info.hpp).
#include "../../include/run_info.hpp"

RunInfo::RunInfo(int nMperBl, int matSide, int nElem, bool padded, int nodes)
    : nMperBl_(nMperBl),
      matSide_(matSide),
      nElem_(nElem),
      nElemPNode_(nodes > 0 ? nElem / nodes : nElem),
      scrSpace_(p2Above(matSide)),
      scrSpace2_(16),
      loadSize_(padded ? p2Above(matSide) : matSide) {}

int RunInfo::p2Above(int n) {
  if (n < 0) return 0;
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}
