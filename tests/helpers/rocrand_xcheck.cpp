// Prints rocRAND philox4x32_10 outputs for (seed, subsequence, offset) triples
// read from stdin (hex); used only to cross-check the oracle's Philox.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

int main() {
  unsigned long long seed, sub, off;
  while (std::scanf("%llx %llx %llx", &seed, &sub, &off) == 3) {
    rocrand_device::philox4x32_10_engine eng(seed, sub, off);
    unsigned a = eng(), b = eng(), c = eng(), d = eng();
    std::printf("%08x %08x %08x %08x\n", a, b, c, d);
  }
  return 0;
}
