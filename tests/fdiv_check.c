/* Exhaustive check of swe_div500 (simplewinderosion_amd/csrc/swe_fdiv.h):
 * every one of the 2^32 float32 bit patterns x must give the same bits as
 * the IEEE quotient x / 500.0f.  Prints the number of mismatches and the
 * first one found.  Built and run by tests/test_fdiv.py. */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../simplewinderosion_amd/csrc/swe_fdiv.h"

int main(void) {
  long long bad = 0;
  uint32_t first = 0;
  int have_first = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (long long hi = 0; hi < (1LL << 16); hi++) {
    for (uint32_t lo = 0; lo < 65536u; lo++) {
      const uint32_t u = ((uint32_t)hi << 16) | lo;
      float x, a, r;
      uint32_t ua, ur;
      memcpy(&x, &u, 4);
      a = swe_div500(x);
      r = x / 500.0f;
      memcpy(&ua, &a, 4);
      memcpy(&ur, &r, 4);
      if (ua != ur) {
        bad++;
#pragma omp critical
        if (!have_first) {
          have_first = 1;
          first = u;
        }
      }
    }
  }
  printf("mismatches %lld first 0x%08x\n", bad, first);
  return bad != 0;
}
