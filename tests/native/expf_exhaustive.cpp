// Host build of the device expf (rgsplat_amd/csrc/rgs_math.h) checked against
// glibc expf on all 2^32 inputs. Prints the mismatch count.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdint>
#include "rgs_math.h"

int main() {
  long long mism = 0;
#pragma omp parallel for reduction(+ : mism) schedule(static)
  for (long long b = 0; b < (1LL << 32); ++b) {
    uint32_t u = (uint32_t)b;
    float x;
    std::memcpy(&x, &u, 4);
    if (std::isnan(x)) continue;
    float a = rgs::rgs_expf(x), r = expf(x);
    uint32_t ua, ur;
    std::memcpy(&ua, &a, 4);
    std::memcpy(&ur, &r, 4);
    if (ua != ur) ++mism;
  }
  std::printf("%lld\n", mism);
  return 0;
}
