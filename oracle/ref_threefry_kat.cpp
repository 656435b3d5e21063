// Known-answer dump for Threefry2x64-20 from the reference's own vendored
// Random123 header, compiled in place from the reference checkout (see
// oracle/Makefile, target _ref/threefry_kat).  TEST INFRASTRUCTURE ONLY: the
// output pins the oracle's and the device kernels' Threefry restatements.
//
// Input lines on stdin: "ctr0 ctr1 key0 key1" in hex; output lines:
// "ctr0 ctr1 key0 key1 out0 out1" in hex.
#include <cinttypes>
#include <cstdio>

#include "random123/threefry.h"

int main() {
  unsigned long long c0, c1, k0, k1;
  r123::Threefry2x64 gen;
  while (std::scanf("%llx %llx %llx %llx", &c0, &c1, &k0, &k1) == 4) {
    r123::Threefry2x64::ctr_type ctr = {{c0, c1}};
    r123::Threefry2x64::key_type key = {{k0, k1}};
    const r123::Threefry2x64::ctr_type r = gen(ctr, key);
    std::printf("%016llx %016llx %016llx %016llx %016llx %016llx\n", c0, c1,
                k0, k1, (unsigned long long)r[0], (unsigned long long)r[1]);
  }
  return 0;
}
