// Harness around the reference's core/drivers.hpp (included from the
// reference tree at build time, never copied).  Reads lines
//   kind constant amplitude p1 p2 p3 mul add t
// and prints getCurrentScalarValue(t) as a hex float for the driver built by
// the reference factory of `kind`, after `* mul` and `+ add` when those are
// not 1 and 0.  Kinds: const pulse sine posine halfsine step trap gimp gstep.
// TEST INFRASTRUCTURE ONLY.
#include <cstdio>
#include <cstring>
#include <string>

#include "drivers.hpp"

using SD = ScalarDriver<double>;

static SD make(const std::string& k, double c, double a, double p1, double p2, double p3) {
  if (k == "const") return SD::getConstantDriver(c);
  if (k == "pulse") return SD::getPulseDriver(c, a, p1, p2);
  if (k == "sine") return SD::getSineDriver(c, a, p1, p2);
  if (k == "posine") return SD::getPosSineDriver(c, a, p1, p2);
  if (k == "halfsine") return SD::getHalfSineDriver(c, a, p1, p2);
  if (k == "step") return SD::getStepDriver(c, a, p1, p2);
  if (k == "trap") return SD::getTrapezoidDriver(c, a, p1, p2, p3);
  if (k == "gimp") return SD::getGaussianImpulseDriver(c, a, p1, p2);
  if (k == "gstep") return SD::getGaussianStepDriver(c, a, p1, p2);
  return SD::getConstantDriver(0);
}

int main() {
  char kind[32];
  double c, a, p1, p2, p3, mul, add, t;
  while (std::scanf("%31s %lf %lf %lf %lf %lf %lf %lf %lf", kind, &c, &a, &p1, &p2, &p3, &mul, &add, &t) == 9) {
    SD d = make(kind, c, a, p1, p2, p3);
    if (mul != 1.0) d = d * mul;
    if (add != 0.0) d = d + add;
    std::printf("%a\n", d.getCurrentScalarValue(t));
  }
  return 0;
}
