// The bridging of a band brick's nodes left without a value (partcache.rs:872-909),
// as one sequential pass over the brick's 729 nodes. The device runs it on one lane of
// the brick's block (finish_brick); the same code builds for the host, where the
// tests check it against the rule restated.
//
// Node states: 0 read exactly, 1 past the clamp (far) with no sign found, 2
// undecided, 3 signed at +-clamp by the flood fill, the centre's ball or the centre's
// sign, 4 bridged in an earlier pass, 5 bridged in this pass.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define ICAM_HD __host__ __device__
#else
#define ICAM_HD
#endif

namespace icam {

// Each pass gives every open node (states 1, 2) with a known neighbour (26 around it)
// the mean of the neighbours known when the pass began (states 0, 3, 4, summed with x
// fastest, then y, then z), held within the bracket each exactly known neighbour
// (states 0, 3) sets, |f(n) - f(m)| <= |n - m| h, on top of [-clamp, clamp] (the middle
// of an empty bracket); nodes bridged in a pass count as known from the next. Passes
// run until no node is open; when a pass moves none, the rest take +clamp (air).
// Returns the number of nodes bridged.
ICAM_HD inline long bridge_nodes(double* val, uint8_t* st, double h, double clamp) {
  long bridged = 0;
  for (int pass = 0; pass < 729; ++pass) {
    bool left = false, moved = false;
    for (int n = 0; n < 729; ++n) {
      if (st[n] != 1 && st[n] != 2) continue;
      const int i = n % 9, j = (n / 9) % 9, k = n / 81;
      double sum = 0.0, lo = -clamp, hi = clamp;
      int cnt = 0;
      for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
          for (int di = -1; di <= 1; ++di) {
            if (!di && !dj && !dk) continue;
            const int a = i + di, b = j + dj, c = k + dk;
            if (a < 0 || a > 8 || b < 0 || b > 8 || c < 0 || c > 8) continue;
            const int m = a + 9 * b + 81 * c;
            const uint8_t s = st[m];
            if (s == 0 || s == 3 || s == 4) sum += val[m], ++cnt;
            if (s == 0 || s == 3) {
              const double reach = sqrt(double(di * di + dj * dj + dk * dk)) * h;
              lo = fmax(lo, val[m] - reach);
              hi = fmin(hi, val[m] + reach);
            }
          }
      if (!cnt) {
        left = true;
        continue;
      }
      const double v = sum / cnt;
      val[n] = lo <= hi ? fmin(fmax(v, lo), hi) : 0.5 * (lo + hi);
      st[n] = 5;
      moved = true;
      ++bridged;
    }
    for (int n = 0; n < 729; ++n)
      if (st[n] == 5) st[n] = 4;
    if (!left) break;
    if (!moved) {
      for (int n = 0; n < 729; ++n)
        if (st[n] == 1 || st[n] == 2) val[n] = clamp, st[n] = 4, ++bridged;
      break;
    }
  }
  return bridged;
}

}  // namespace icam
