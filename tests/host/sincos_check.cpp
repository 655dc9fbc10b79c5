// Host build of the FP64 path's sincos (csrc/sincos64.hpp) for
// tests/test_sincos64.py: reads arguments (one per line) on stdin and prints
// "sin cos" for each with 17 significant digits.
#include "sincos64.hpp"

#include <cstdio>

int main() {
  double x;
  while (std::scanf("%lf", &x) == 1) {
    double s, c;
    glst::sincos_cw(x, &s, &c);
    std::printf("%.17e %.17e\n", s, c);
  }
  return 0;
}
