// x / 12 through div12 (gpu-comcot_amd/csrc/fastdiv.h) against IEEE division
// for every float; prints the mismatch count where |x| >= 2^-124 or x = 0,
// and the largest |x| that mismatches at all.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "fastdiv.h"

int main() {
  unsigned long long bad = 0, bad_any = 0;
  float worst = 0.0f;
#pragma omp parallel for reduction(+ : bad, bad_any) reduction(max : worst)
  for (long long k = 0; k < (1ll << 32); ++k) {
    const uint32_t u = (uint32_t)k;
    float x;
    memcpy(&x, &u, 4);
    if (!(fabsf(x) <= 3.4028235e38f)) continue;
    const float want = x / 12.0f, got = gc::div12(x);
    uint32_t a, b;
    memcpy(&a, &want, 4);
    memcpy(&b, &got, 4);
    if (a == b) continue;
    ++bad_any;
    if (fabsf(x) > worst) worst = fabsf(x);
    if (fabsf(x) >= 0x1p-124f || x == 0.0f) ++bad;
  }
  printf("%llu %llu %a\n", bad, bad_any, (double)worst);
  return 0;
}
