// Harness around the reference's own suffix array, the PBBS DC3 (skew)
// algorithm of suffixArray/parallelKS/pks.C (compiled unmodified next to this
// file by oracle/Makefile into oracle/_ref/libref_pks.so).  TEST INFRASTRUCTURE
// ONLY: it gives the CPU suffix array that the GPU suffix array and the BWT of
// the FM-index workload are checked against.  pks.C sorts the text shifted by
// one (every byte + 1, zero padding past the end, pks.C:242-244), so a proper
// prefix sorts before its extensions, as divsufsort does.
#include <stdint.h>
#include <stdlib.h>
#include <utility>
#include "parallel.h"

std::pair<uintT*, uintT*> suffixArray(unsigned char* s, long n, bool findLCPs);

extern "C" int ref_suffix_array(const uint8_t* s, uint64_t n, int64_t* sa) {
  if (n < 3 || n >= (1ull << 31)) return -1;
  uintT* r = suffixArray(const_cast<unsigned char*>(s), (long)n, false).first;
  if (!r) return -2;
  for (uint64_t i = 0; i < n; i++) sa[i] = (int64_t)r[i];
  free(r);
  return 0;
}

extern "C" int ref_sa_workers(void) {
#ifdef OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
