// WT.h -- drop-in for the reference's wavelet-tree build function
//   pair<WTnode*,long*> WT(symbol* s, uintT n, uintT sigma)
// (wavelet/levelWT/WT.h:22-43; the same header in serialWT/ddWT/recWT).
// Same types and macros: POINTERS, INT (symbol = uintT) and LONG (64-bit
// intT/uintT, as in common/parallel.h:71-81).  The build runs on the GPU
// (wtg_wt_host); both arrays are malloc'ed and the caller frees them, as in
// wavelet/levelWT/WTTime.C:99-100,135.  Unreached node slots, which the
// reference leaves uninitialised, are {0, 0, UINT_T_MAX, UINT_T_MAX, UINT_T_MAX}.
#ifndef WTG_DROPIN_WT_H
#define WTG_DROPIN_WT_H
#include <climits>
#include <cstdlib>
#include <utility>
#include <vector>
#include "wtg.h"
#include "wtg_runtime.hpp"

#ifndef _PARALLEL_H  // the reference's parallel.h defines these
#if defined(LONG)
typedef long intT;
typedef unsigned long uintT;
#define INT_T_MAX LONG_MAX
#define UINT_T_MAX ULONG_MAX
#else
typedef int intT;
typedef unsigned int uintT;
#define INT_T_MAX INT_MAX
#define UINT_T_MAX UINT_MAX
#endif
#endif

using namespace std;

#define POINTERS 1

typedef uintT intOffset;

struct WTnode {
  intOffset bitmapPtr;
  uintT length;
#ifdef POINTERS
  uintT parent, leftChild, rightChild;
#endif
};

#ifdef INT
typedef uintT symbol;
#else
typedef unsigned char symbol;
#endif

inline pair<WTnode*, long*> WT(symbol* s, uintT n, uintT sigma) {
  int lg = 0;
  for (uint64_t b = (uint64_t)sigma - 1; sigma > 1 && b > 0; b >>= 1) lg++;
  const uint32_t L = lg > 1 ? (uint32_t)lg : 1u;
  const uint64_t nw = wtg_levels_words(n, L);
  long* wt = (long*)malloc((nw ? nw : 1) * sizeof(long));
  WTnode* nodes = (WTnode*)malloc(sizeof(WTnode) * 2 * (uint64_t)sigma);
  std::vector<wtg_node> tmp(2 * (uint64_t)sigma);
  wtg::check(wtg_wt_host(wtg::default_context(), s, (int)sizeof(symbol), n, sigma,
                         reinterpret_cast<uint64_t*>(wt), tmp.data()));
  auto narrow = [](uint64_t v) -> uintT { return v == WTG_NONE ? (uintT)UINT_T_MAX : (uintT)v; };
  for (uint64_t v = 0; v < 2 * (uint64_t)sigma; v++) {
    nodes[v].bitmapPtr = (intOffset)tmp[v].bitmapPtr;
    nodes[v].length = (uintT)tmp[v].length;
#ifdef POINTERS
    nodes[v].parent = narrow(tmp[v].parent);
    nodes[v].leftChild = narrow(tmp[v].leftChild);
    nodes[v].rightChild = narrow(tmp[v].rightChild);
#endif
  }
  return make_pair(nodes, wt);
}

#endif
