// ref_harness.cpp -- test-only shim that exposes the reference's own
// `pair<WTnode*,long*> WT(symbol*, uintT, uintT)` (wavelet/*/WT.h:43) through a
// C symbol so that tests and the CPU baseline can call the compiled reference
// builders.  It is compiled together with an unmodified reference WT.C from
// /root/reference by oracle/Makefile; nothing here is product code.
#include <stdint.h>
#include <string.h>
#include <time.h>
#include <utility>
#include <cstdlib>
#include "parallel.h"
#include "utils.h"
#include "WT.h"

extern "C" {

// width of the reference `symbol` type in bytes (1 = uchar, 4 = uint, 8 = ulong)
int ref_symbol_bytes() { return (int)sizeof(symbol); }
int ref_uintT_bytes() { return (int)sizeof(uintT); }

static double now_s() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// Builds the tree once (like timeWT's untimed call) and copies the outputs.
// words_out: ceil(n*L/64) uint64; nodes_out: 2*sigma*5 uint64 (bitmapPtr,
// length, parent, leftChild, rightChild; UINT_T_MAX widened to UINT64_MAX) or NULL.
int ref_wt(const void* s_in, uint64_t n, uint64_t sigma, uint64_t* words_out,
           uint64_t* nodes_out) {
  symbol* s = (symbol*)malloc((n ? n : 1) * sizeof(symbol));
  memcpy(s, s_in, n * sizeof(symbol));
  std::pair<WTnode*, long*> R = WT(s, (uintT)n, (uintT)sigma);
  int levels = std::max(1, utils::log2Up((uintT)sigma));
  uint64_t nw = (n * (uint64_t)levels + 63) / 64;
  memcpy(words_out, R.second, nw * 8);
  if (nodes_out) {
    for (uint64_t v = 0; v < 2 * sigma; v++) {
      const WTnode& x = R.first[v];
      uint64_t f[5] = {(uint64_t)x.bitmapPtr, (uint64_t)x.length, (uint64_t)x.parent,
                       (uint64_t)x.leftChild, (uint64_t)x.rightChild};
      for (int k = 0; k < 5; k++) {
        if (k != 0 && k != 1 && f[k] == (uint64_t)UINT_T_MAX) f[k] = UINT64_MAX;
        nodes_out[v * 5 + k] = f[k];
      }
    }
  }
  free(R.first);
  free(R.second);
  free(s);
  return levels;
}

// Timing like wavelet/levelWT/WTTime.C:96-103: one untimed build, then `rounds`
// timed builds; writes each round's wall seconds into times[].
int ref_wt_time(const void* s_in, uint64_t n, uint64_t sigma, int rounds, double* times) {
  symbol* s = (symbol*)malloc((n ? n : 1) * sizeof(symbol));
  memcpy(s, s_in, n * sizeof(symbol));
  std::pair<WTnode*, long*> R = WT(s, (uintT)n, (uintT)sigma);
  for (int r = 0; r < rounds; r++) {
    free(R.first);
    free(R.second);
    double t0 = now_s();
    R = WT(s, (uintT)n, (uintT)sigma);
    times[r] = now_s() - t0;
  }
  free(R.first);
  free(R.second);
  free(s);
  return 0;
}

int ref_workers() { return getWorkers(); }
}
