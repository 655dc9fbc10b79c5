// Prints the host Chebyshev plan of the mixed structure factors for the
// Python test (tests/test_host_cheb.py): Bessel values J_m(z) of bessel_j,
// the term counts of cheb_terms, and, for the 288,000-charge and
// 1,000-water workloads, every alpha group's plan (plan_chebyshev: terms,
// radial magnitudes, rho and the coefficient matrix).  Host code only.
#include "runtime.hpp"

#include <cstdio>

// the engine's definitions (engine.hip): no design files, no global error
std::string glst::tdesign_dir() { return ""; }
void glst::set_global_error(const std::string &) {}

int main() {
  const double zs[] = {0.3, 2.0, 10.21, 14.44, 25.0};
  for (double z : zs) {
    double J[60];
    glst::bessel_j(z, 60, J);
    std::printf("bessel %.17g", z);
    for (double v : J) std::printf(" %.17g", v);
    std::printf("\nterms %.17g %u\n", z, glst::cheb_terms(z));
  }
  struct Case {
    const char *name;
    double L, rcut, tol;
  };
  const Case cases[] = {{"288k", 142.0, 10.0, 1e-4}, {"1k", 31.04, 9.0, 1e-5}};
  for (const Case &c : cases) {
    glst::SystemSetup su = glst::setup_system(GLST_CT_FLOAT, c.L, c.L, c.L,
                                              c.rcut, c.tol);
    glst::SplitPlan P = glst::make_splits(su.nodes, su.K, 1024);
    const std::vector<float> tab =
        glst::plan_chebyshev(su.nodes, su.grid.h, P.sf_groups);
    std::printf("case %s h %.17g rho %.17g groups %zu table %zu\n", c.name,
                su.grid.h, glst::cheb_rho(su.grid.h), P.sf_groups.size(),
                tab.size());
    for (size_t g = 0; g < P.sf_groups.size(); ++g) {
      const glst::dev::SfGroup &sg = P.sf_groups[g];
      std::printf("group %zu nr %u m %u coff %u kappa", g, sg.nr, sg.m,
                  sg.coff);
      for (double k : su.nodes.kappa[g]) std::printf(" %.17g", k);
      std::printf("\n");
      if (sg.m == 0) continue;
      for (uint32_t i = 0; i < sg.nr; ++i) {
        std::printf("row %zu %u", g, i);
        for (uint32_t m = 0; m < sg.m; ++m)
          std::printf(" %.9g", tab[sg.coff + i * sg.m + m]);
        std::printf("\n");
      }
    }
  }
  return 0;
}
