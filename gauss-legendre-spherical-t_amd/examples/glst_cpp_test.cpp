// C++ facade check: glst_force<double> and glst_force<float> on a 1,000-water
// box against the direct Coulomb sum, the accumulate/zero semantics of the
// fx()/en() views, and the exceptions the reference throws for host errors.
// Exit status 0 on success.
#include <glst_force.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

namespace {

int failures = 0;

void expect(bool ok, const char *what) {
  std::printf("%s: %s\n", ok ? "ok  " : "FAIL", what);
  if (!ok) ++failures;
}

template <typename CT>
void check_path(const char *name, const hip_container<double> *in,
                const std::vector<double> *ref, size_t n) {
  glst_force<CT> f(static_cast<unsigned>(n), 1e-5, 31.04, 31.04, 31.04, 9.0);
  f.calc_ener_force(in[0].d_array().data(), in[1].d_array().data(),
                    in[2].d_array().data(), in[3].d_array().data());
  hip_container<double> fx, fy, fz, en;
  f.get_ef(fx, fy, fz, en);
  double num = 0, den = 0;
  const std::vector<double> *got[3] = {&fx.h_array(), &fy.h_array(),
                                       &fz.h_array()};
  for (int c = 0; c < 3; ++c)
    for (size_t i = 0; i < n; ++i) {
      const double d = (*got[c])[i] - ref[c][i];
      num += d * d;
      den += ref[c][i] * ref[c][i];
    }
  const double rel = std::sqrt(num / den);
  std::printf("%s relative force error vs direct sum: %.3e\n", name, rel);
  expect(rel < 1e-5, name);

  // outputs accumulate; the driver zeroes the views between steps
  f.calc_ener_force(in[0].d_array().data(), in[1].d_array().data(),
                    in[2].d_array().data(), in[3].d_array().data());
  hip_container<double> e2x, e2y, e2z, e2e;
  f.get_ef(e2x, e2y, e2z, e2e);
  // x + y + x + y and 2 (x + y) differ only by rounding
  double scale = 0.0, worst = 0.0;
  for (size_t i = 0; i < n; ++i) scale = std::max(scale, std::abs(en[i]));
  for (size_t i = 0; i < n; ++i)
    worst = std::max(worst, std::abs(e2e[i] - 2 * en[i]));
  expect(worst <= 1e-12 * scale, "second step accumulates");
  f.en()[0].set(0.0);
  f.fx()[0].set(0.0);
  f.fy()[0].set(0.0);
  f.fz()[0].set(0.0);
  f.calc_ener_force(in[0].d_array().data(), in[1].d_array().data(),
                    in[2].d_array().data(), in[3].d_array().data());
  f.get_ef(e2x, e2y, e2z, e2e);
  expect(e2e[5] == en[5] && e2x[7] == fx[7], "set(0.0) on the views clears");
}

}  // namespace

int main() {
  const unsigned nw = 1000;
  const size_t n = 3 * nw;
  std::vector<double> h[4];
  for (auto &v : h) v.resize(n);
  if (glst_make_water_box(nw, 0, 31.04, 1, h[0].data(), h[1].data(),
                          h[2].data(), h[3].data()) != GLST_OK)
    return 2;
  hip_container<double> in[4] = {hip_container<double>(h[0]),
                                 hip_container<double>(h[1]),
                                 hip_container<double>(h[2]),
                                 hip_container<double>(h[3])};
  hip_container<double> c[4] = {hip_container<double>(n),
                                hip_container<double>(n),
                                hip_container<double>(n),
                                hip_container<double>(n)};
  if (glst_coulomb_direct(in[0].d_array().data(), in[1].d_array().data(),
                          in[2].d_array().data(), in[3].d_array().data(),
                          static_cast<uint32_t>(n), c[0].d_array().data(),
                          c[1].d_array().data(), c[2].d_array().data(),
                          c[3].d_array().data(), nullptr) != GLST_OK)
    return 2;
  std::vector<double> ref[4];
  for (int i = 0; i < 4; ++i) {
    c[i].transfer_to_host();
    ref[i] = c[i].h_array();
  }
  check_path<double>("glst_force<double>", in, ref, n);
  check_path<float>("glst_force<float>", in, ref, n);

  bool threw = false;
  try {
    glst_force<double> bad(10, -1.0, 10, 10, 10, 5.0);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  expect(threw, "invalid tol throws std::runtime_error");
  threw = false;
  try {
    glst_force<double> g;
    g.calc_sf();
  } catch (const std::runtime_error &) {
    threw = true;
  }
  expect(threw, "uninitialised object throws");
  std::printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
