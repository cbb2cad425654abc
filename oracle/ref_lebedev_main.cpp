// Driver for the reference's Lebedev rules (compiled only where the reference
// sources are present). Prints "x y z w" for the requested rule orders.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "sphere_lebedev_rule.hpp"

int main(int argc, char** argv) {
  for (int i = 1; i < argc; ++i) {
    const int order = std::atoi(argv[i]);
    std::vector<double> x(order), y(order), z(order), w(order);
    ld_by_order(order, x.data(), y.data(), z.data(), w.data());
    std::printf("rule %d\n", order);
    for (int k = 0; k < order; ++k) std::printf("%.17g %.17g %.17g %.17g\n", x[k], y[k], z[k], w[k]);
  }
  return 0;
}
