// CPU check (test infrastructure) of ac::sqrt_rn_fix, the correction step of the distance kernels' square root:
// for every tested float x >= 2^-104 and every approximation s within 1 ulp of sqrt(x) (the correctly rounded root
// and its two neighbours, standing in for the hardware root), the corrected value must be the correctly rounded
// root.  Built with hipcc and run on the host by tests/test_capi_host.py; argv[1] is the stride over float bit
// patterns (1 = every float).  Prints the count checked and exits 1 on any mismatch.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../application_classification_amd/csrc/ac_sqrt_rn.hpp"

static float bits_to_float(unsigned u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static unsigned float_to_bits(float f) {
  unsigned u;
  std::memcpy(&u, &f, 4);
  return u;
}

int main(int argc, char** argv) {
  const unsigned stride = argc > 1 ? (unsigned)std::atoi(argv[1]) : 1u;
  const unsigned lo = float_to_bits(0x1p-104f), hi = 0x7F800000u;  // [2^-104, +inf)
  unsigned long long checked = 0, wrong = 0;
  auto check = [&](float x) {
    const float t = std::sqrt(x);
    for (int d = -1; d <= 1; ++d) {
      if (t == 0.f && d < 0) continue;
      const float s = bits_to_float(float_to_bits(t) + d);
      const float r = ac::sqrt_rn_fix(x, s);
      ++checked;
      if (float_to_bits(r) != float_to_bits(t)) {
        if (wrong < 5) std::printf("x = %a: root %a, from %a got %a\n", x, t, s, r);
        ++wrong;
      }
    }
  };
  check(0.f);
  for (unsigned u = lo; u < hi; u += stride) check(bits_to_float(u));
  check(bits_to_float(hi - 1));
  std::printf("sqrt_rn_fix: %llu cases, %llu wrong\n", checked, wrong);
  return wrong ? 1 : 0;
}
