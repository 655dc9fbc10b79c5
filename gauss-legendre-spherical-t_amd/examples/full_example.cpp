// full_example: the reference driver (test/full_example.cu) on the MI355X
// engine.  Same CLI, environment variables, stage timings and report lines,
// so log parsers written for the reference read this output unchanged.
//
//   full_example [sys] [tol] [box_dim] [rcut]
//   full_example [sys] [tol] [box_dim] [ncell.x] [ncell.y] [ncell.z]
//
// [sys] is a CHARMM file prefix (sys.psf / sys.cor) or a synthetic water box
// "water:NWATER[:SEED[:NION]]".  GLST_CT=float selects the mixed path
// (default double, as the reference's typedef); GLST_BENCH_ITERS sets the
// iteration count (default 100).
#include <glst_force.hpp>
#include <glst_io.h>

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <string>
#include <vector>

namespace {

bool env_true(const char *name) {
  const char *v = std::getenv(name);
  if (!v) return false;
  std::string s(v);
  return !(s.empty() || s == "0" || s == "false" || s == "FALSE" ||
           s == "off" || s == "OFF" || s == "no" || s == "NO");
}

std::string env_or(const char *name, const std::string &fallback) {
  const char *v = std::getenv(name);
  return v ? std::string(v) : fallback;
}

double avg(const std::vector<double> &x) {
  double s = 0.0;
  for (double v : x) s += v;
  return s / static_cast<double>(x.size());
}

double stdev(const std::vector<double> &x) {
  const double m = avg(x);
  double s = 0.0;
  for (double v : x) s += (v - m) * (v - m);
  return std::sqrt(s / static_cast<double>(x.size()));
}

void load_system(const std::string &sys, std::vector<double> &rx,
                 std::vector<double> &ry, std::vector<double> &rz,
                 std::vector<double> &qc) {
  if (sys.rfind("water:", 0) == 0) {
    unsigned long nw = 0, seed = 1, nion = 0;
    std::string rest = sys.substr(6);
    nw = std::stoul(rest);
    size_t p = rest.find(':');
    if (p != std::string::npos) {
      rest = rest.substr(p + 1);
      seed = std::stoul(rest);
      p = rest.find(':');
      if (p != std::string::npos) nion = std::stoul(rest.substr(p + 1));
    }
    const size_t n = 3 * nw + 2 * nion;
    rx.resize(n);
    ry.resize(n);
    rz.resize(n);
    qc.resize(n);
    // the box is generated at the density of liquid water
    const double L = std::cbrt(static_cast<double>(nw) / 0.0334);
    if (glst_make_water_box(nw, nion, L, seed, rx.data(), ry.data(),
                            rz.data(), qc.data()) != GLST_OK)
      throw std::runtime_error("glst_make_water_box failed");
    return;
  }
  uint64_t natom = 0;
  if (glst_psf_natom((sys + ".psf").c_str(), &natom) != GLST_OK)
    throw std::runtime_error("Failed to open file \"" + sys + ".psf\"");
  rx.resize(natom);
  ry.resize(natom);
  rz.resize(natom);
  qc.resize(natom);
  if (glst_read_cor((sys + ".cor").c_str(), natom, rx.data(), ry.data(),
                    rz.data()) != GLST_OK)
    throw std::runtime_error("Number of atoms differs from size of array");
  if (glst_read_psf_charges((sys + ".psf").c_str(), natom, qc.data()) !=
      GLST_OK)
    throw std::runtime_error("Number of atoms differs from size of array");
  glst_recenter(natom, rx.data(), ry.data(), rz.data());
}

// utils/utils.cpp:60-168: min/max/avg/RMS of |error|, each also normalised by
// the RMS of the GLST value.
void print_error_report(const std::vector<double> *glst,
                        const std::vector<double> *coul, size_t natom,
                        double tol) {
  const char *names[4] = {"  X Force Error: ", "  Y Force Error: ",
                          "  Z Force Error: ", "   Energy Error: "};
  std::ios fmt(nullptr);
  fmt.copyfmt(std::cout);
  std::cout << std::endl;
  std::cout << "Target error: " << std::scientific << tol << std::endl;
  std::cout << std::endl;
  for (int c = 0; c < 4; ++c) {
    double mn = std::numeric_limits<double>::max();
    double mx = std::numeric_limits<double>::lowest();
    double av = 0.0, rms = 0.0, ref = 0.0;
    for (size_t i = 0; i < natom; ++i) {
      const double e = coul[c][i] - glst[c][i];
      const double a = std::abs(e);
      mn = std::min(mn, a);
      mx = std::max(mx, a);
      av += a;
      rms += e * e;
      ref += glst[c][i] * glst[c][i];
    }
    av /= static_cast<double>(natom);
    rms = std::sqrt(rms / static_cast<double>(natom));
    ref = std::sqrt(ref / static_cast<double>(natom));
    std::cout << names[c] << std::endl;
    std::cout << "         Absolute Min: " << std::showpos << mn
              << " (Norm: " << mn / ref << ")" << std::endl;
    std::cout << "         Absolute Max: " << std::showpos << mx
              << " (Norm: " << mx / ref << ")" << std::endl;
    std::cout << "              Average: " << std::showpos << av
              << " (Norm: " << av / ref << ")" << std::endl;
    std::cout << "    Root Mean Squared: " << rms << " (Norm: " << rms / ref
              << ")" << std::endl;
    std::cout << std::endl;
  }
  std::cout.copyfmt(fmt);
}

template <typename CT>
int run(int argc, char **argv, const std::string &sys, double tol,
        double box, double rcut, int ncx, int ncy, int ncz) {
  std::vector<double> hx, hy, hz, hq;
  load_system(sys, hx, hy, hz, hq);
  const size_t natom = hx.size();
  hip_container<double> rx(hx), ry(hy), rz(hz), qc(hq);

  std::unique_ptr<glst_force<CT>> glst;
  if (argc == 5)
    glst = std::make_unique<glst_force<CT>>(natom, tol, box, box, box, rcut);
  else
    glst = std::make_unique<glst_force<CT>>(natom, tol, box, box, box, ncx,
                                            ncy, ncz);

  // setup summary (printed by the reference's init(), src/glst_force.cu:245-293)
  const glst_info_t inf = glst->info();
  std::cout << std::endl;
  std::cout << "          Number of atoms: " << inf.natom << std::endl;
  std::cout << "    System dimensions [A]: " << box << " x " << box << " x "
            << box << std::endl;
  std::cout << "          Number of cells: " << inf.ncell[0] << ", "
            << inf.ncell[1] << ", " << inf.ncell[2] << std::endl;
  std::cout << "      Cell dimensions [A]: " << inf.cell_dim[0] << " x "
            << inf.cell_dim[1] << " x " << inf.cell_dim[2] << std::endl;
  std::cout << "  Total space covered [A]: "
            << inf.ncell[0] * inf.cell_dim[0] << " x "
            << inf.ncell[1] * inf.cell_dim[1] << " x "
            << inf.ncell[2] * inf.cell_dim[2] << std::endl;
  std::cout << "           Number of GPUs: " << 1 << std::endl;
  std::cout << std::endl;
  std::cout << "  Number of alpha groups: " << inf.ngroup << std::endl;
  std::cout << "       Total number of cubature nodes: " << inf.num_nodes
            << std::endl;
  for (unsigned g = 0; g < inf.ngroup; ++g)
    std::cout << "  Number of cubature nodes in group " << g << ": "
              << glst->group_info(g).num_nodes << std::endl;
  std::cout << "  Number of cubature nodes per device: ~" << inf.num_nodes
            << std::endl;
  std::cout << std::endl;

  const size_t iters = std::stoul(env_or("GLST_BENCH_ITERS", "100"));
  std::vector<std::vector<double>> t(8, std::vector<double>(iters));
  using clk = std::chrono::high_resolution_clock;
  auto ms = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  for (size_t it = 0; it < iters; ++it) {
    std::cerr << "\rIteration " << it << std::flush;
    const auto t0 = clk::now();
    glst->assign_atoms(rx.d_array().data(), ry.d_array().data(),
                       rz.d_array().data(), qc.d_array().data());
    (void)hipDeviceSynchronize();
    const auto t1 = clk::now();
    glst->calc_sf();
    (void)hipDeviceSynchronize();
    const auto t2 = clk::now();
    glst->sum_rmt_sf();
    (void)hipDeviceSynchronize();
    const auto t3 = clk::now();
    glst->calc_lr_ef();
    (void)hipDeviceSynchronize();
    const auto t4 = clk::now();
    glst->calc_sr_ef();
    (void)hipDeviceSynchronize();
    const auto t5 = clk::now();
    glst->comm_ef();
    (void)hipDeviceSynchronize();
    const auto t6 = clk::now();
    t[0][it] = ms(t0, t1);
    t[1][it] = ms(t1, t2);
    t[2][it] = ms(t2, t3);
    t[3][it] = ms(t3, t4);
    t[4][it] = ms(t4, t5);
    t[5][it] = ms(t5, t6);
    t[6][it] = ms(t0, t4);
    t[7][it] = ms(t0, t6);
    if (it + 1 < iters) {
      glst->fx()[0].set(0.0);
      glst->fy()[0].set(0.0);
      glst->fz()[0].set(0.0);
      glst->en()[0].set(0.0);
    }
  }
  std::cerr << "\rFinished " << iters << " calculations" << std::endl;
  std::cout << "Finished " << iters << " calculations" << std::endl;
  std::cout << std::endl;
  const char *labels[8] = {"                   Assign atoms to cells: ",
                           "             Calculate structure factors: ",
                           "            Sum remote structure factors: ",
                           "  Calculate long-range energy and forces: ",
                           " Calculate short-range energy and forces: ",
                           "         Communicate energies and forces: ",
                           "                 Long-Range GLST Runtime: ",
                           "                      Total GLST Runtime: "};
  for (int s = 0; s < 8; ++s) {
    if (s == 6)
      std::cout << "----------------------------------------------------------"
                   "----------------"
                << std::endl;
    std::cout << labels[s] << avg(t[s]) << " ms (+/- " << stdev(t[s])
              << " ms)" << std::endl;
  }

  hip_container<double> fx, fy, fz, en;
  glst->get_ef(fx, fy, fz, en);

  hip_container<double> cx(natom), cy(natom), cz(natom), ce(natom);
  const auto c0 = clk::now();
  if (glst_coulomb_direct(rx.d_array().data(), ry.d_array().data(),
                          rz.d_array().data(), qc.d_array().data(),
                          static_cast<uint32_t>(natom), cx.d_array().data(),
                          cy.d_array().data(), cz.d_array().data(),
                          ce.d_array().data(), nullptr) != GLST_OK)
    throw std::runtime_error(glst_global_last_error());
  const auto c1 = clk::now();
  std::cout << std::endl;
  std::cout << "Coulomb Runtime: " << ms(c0, c1) << " ms" << std::endl;
  cx.transfer_to_host();
  cy.transfer_to_host();
  cz.transfer_to_host();
  ce.transfer_to_host();
  const std::vector<double> g4[4] = {fx.h_array(), fy.h_array(), fz.h_array(),
                                     en.h_array()};
  const std::vector<double> c4[4] = {cx.h_array(), cy.h_array(), cz.h_array(),
                                     ce.h_array()};
  print_error_report(g4, c4, natom, tol);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 5 && argc != 7) {
    std::cout << "Usage: " << argv[0] << " [sys] [tol] [box_dim] [rcut]"
              << std::endl;
    std::cout << "    OR" << std::endl;
    std::cout << "       " << argv[0]
              << " [sys] [tol] [box_dim] [ncell.x] [ncell.y] [ncell.z]"
              << std::endl;
    return EXIT_FAILURE;
  }
  std::ofstream bench_out;
  std::streambuf *old = nullptr;
  if (!env_true("GLST_BENCH_TERMINAL")) {
    if (const char *name = std::getenv("GLST_BENCH_OUT")) {
      bench_out.open(name);
      if (!bench_out) {
        std::cerr << "Could not open GLST_BENCH_OUT file: \"" << name << "\""
                  << std::endl;
        return EXIT_FAILURE;
      }
      old = std::cout.rdbuf(bench_out.rdbuf());
    }
  }
  const std::string sys = argv[1];
  const double tol = std::stod(argv[2]);
  const double box = std::stod(argv[3]);
  double rcut = 0.0;
  int ncx = 0, ncy = 0, ncz = 0;
  if (argc == 5) {
    rcut = std::stod(argv[4]);
  } else {
    ncx = std::stoi(argv[4]);
    ncy = std::stoi(argv[5]);
    ncz = std::stoi(argv[6]);
  }
  std::cout << "BENCH_INPUT";
  std::cout << " sweep=" << env_or("GLST_BENCH_SWEEP", "manual");
  std::cout << " sys=" << env_or("GLST_BENCH_SYS", sys);
  std::cout << " tol=" << tol;
  std::cout << " box_dim=" << box << "," << box << "," << box;
  std::cout << " cut=" << rcut;
  std::cout << " requested_gpus=" << env_or("GLST_BENCH_GPUS", "");
  std::cout << std::endl;
  int rc = 0;
  try {
    if (env_or("GLST_CT", "double") == "float")
      rc = run<float>(argc, argv, sys, tol, box, rcut, ncx, ncy, ncz);
    else
      rc = run<double>(argc, argv, sys, tol, box, rcut, ncx, ncy, ncz);
  } catch (const std::exception &e) {
    std::cerr << e.what() << std::endl;
    rc = EXIT_FAILURE;
  }
  if (old) std::cout.rdbuf(old);
  return rc;
}
