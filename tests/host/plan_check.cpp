// Host-only check of the engine's per-handle planning (no GPU): for the
// three single-system workloads (3,000, 288,000 and 1,000,796 charges), both
// calculation types, every back-projection split target the autotuner tries
// and both shell-sum tile limits, it runs setup_system, layout_nodes,
// make_splits and plan_pipeline in the order a handle and its autotuner do,
// and checks every table index against the sizes the kernels are launched
// with.  Built with AddressSanitizer and UBSan by tests/test_host_plan.py,
// so an out-of-bounds write in this host code aborts the run.
#include "runtime.hpp"

#include <cstdio>

std::string glst::tdesign_dir() { return ""; }
void glst::set_global_error(const std::string &) {}

namespace {

int failures = 0;

#define CHECK(cond, ...)                                       \
  do {                                                         \
    if (!(cond)) {                                             \
      std::fprintf(stderr, "FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
      std::fprintf(stderr, __VA_ARGS__);                       \
      std::fprintf(stderr, "\n");                              \
      ++failures;                                              \
    }                                                          \
  } while (0)

struct Workload {
  const char *name;
  uint32_t natom;
  double box, rcut, tol;
};

void check_layout(const glst::SystemSetup &su, const glst::NodeLayout &L) {
  const uint32_t K = su.K, Kp = su.Kp;
  CHECK(L.k64.size() == Kp && L.group.size() == Kp, "node tables %zu", L.k64.size());
  std::vector<uint8_t> seen(K, 0);
  for (uint32_t r = 0; r < K; ++r) {
    const uint32_t d = L.dev_of_ref[r];
    CHECK(d < K, "dev_of_ref[%u] = %u", r, d);
    if (d < K) {
      CHECK(!seen[d], "device node %u twice", d);
      seen[d] = 1;
      CHECK(L.group[d] == su.nodes.group[r], "group of node %u", r);
    }
  }
  const size_t ntask = L.node0.size();
  CHECK(L.nr.size() == ntask && L.koff.size() == ntask &&
            L.dir64.size() == ntask && L.dir32.size() == ntask,
        "task tables");
  for (size_t t = 0; t < ntask; ++t) {
    CHECK(static_cast<uint64_t>(L.node0[t]) + L.nr[t] <= K, "task %zu nodes", t);
    CHECK(static_cast<uint64_t>(L.koff[t]) + L.nr[t] <= L.kappa64.size(),
          "task %zu kappa", t);
  }
}

void check_splits(const glst::SystemSetup &su, const glst::NodeLayout &L,
                  const glst::SplitPlan &P, uint32_t natom, uint32_t target) {
  const size_t G = su.nodes.dirs.size();
  const uint32_t ntask = static_cast<uint32_t>(L.node0.size());
  CHECK(P.split.size() == P.nsplit + 1u, "split table %zu vs %u",
        P.split.size(), P.nsplit);
  CHECK(P.split.front() == 0 && P.split.back() == ntask,
        "split ends %u %u (tasks %u)", P.split.front(), P.split.back(), ntask);
  for (size_t q = 1; q < P.split.size(); ++q)
    CHECK(P.split[q - 1] <= P.split[q], "split %zu decreases", q);
  CHECK(P.lr_groups.size() == G && P.sf_groups.size() == G, "groups");
  uint32_t base = 0;
  for (size_t g = 0; g < G && g < P.lr_groups.size(); ++g) {
    const auto &lg = P.lr_groups[g];
    const auto &sg = P.sf_groups[g];
    const uint32_t M = static_cast<uint32_t>(su.nodes.dirs[g].size());
    CHECK(sg.task0 == base && sg.ntask == M, "sf group %zu", g);
    CHECK(lg.nsplit >= 1 && lg.split0 + lg.nsplit <= P.nsplit,
          "lr group %zu splits %u+%u of %u", g, lg.split0, lg.nsplit, P.nsplit);
    CHECK(lg.nsplit <= 65535u, "lr group %zu grid.y %u", g, lg.nsplit);
    for (uint32_t q = lg.split0; q <= lg.split0 + lg.nsplit && q < P.split.size(); ++q) {
      const uint32_t b = P.split[q];
      CHECK(b >= base && b <= base + M, "group %zu split boundary %u", g, b);
      CHECK(b == base + M || (b - base) % glst::kSplitAlign == 0,
            "group %zu boundary %u not on %u directions", g, b,
            glst::kSplitAlign);
    }
    base += M;
  }
  // the partial-sum buffer the handle sizes from this plan
  const uint64_t bytes = static_cast<uint64_t>(P.nsplit) * natom * 32u;
  CHECK(bytes / 32u / natom == P.nsplit, "partial bytes overflow");
  (void)target;
}

void check_pipeline(const glst::SystemSetup &su, const glst::PipelinePlan &P,
                    uint32_t T) {
  const size_t G = su.nodes.dirs.size();
  const uint32_t ntile = su.Kp / T;
  CHECK(P.tile_end.size() == G && P.rmt_need.size() == G, "pipeline sizes");
  for (size_t g = 0; g < P.tile_end.size(); ++g) {
    CHECK(P.tile_end[g] <= ntile, "tile_end[%zu] = %u > %u", g, P.tile_end[g], ntile);
    if (g > 0) CHECK(P.tile_end[g - 1] <= P.tile_end[g], "tile_end decreases");
    const uint64_t end = static_cast<uint64_t>(su.nodes.first[g]) + su.nodes.count[g];
    // chunk g holds only nodes of groups <= g
    if (g + 1 < G)
      CHECK(static_cast<uint64_t>(P.tile_end[g]) * T <= end, "chunk %zu overruns", g);
    const uint32_t j = P.rmt_need[g];
    CHECK(j >= g && j < G, "rmt_need[%zu] = %u", g, j);
    if (j < G)
      CHECK(j + 1 == G || static_cast<uint64_t>(P.tile_end[j]) * T >= end,
            "chunk %u does not cover group %zu", j, g);
  }
  if (G) CHECK(P.tile_end[G - 1] == ntile, "last chunk ends at %u", P.tile_end[G - 1]);
}

// The Chebyshev matrices of the mixed structure factors: every group that
// takes the Chebyshev form reads NR x M entries from its offset.
void check_chebyshev(const glst::SystemSetup &su, const glst::SplitPlan &P) {
  std::vector<glst::dev::SfGroup> groups = P.sf_groups;
  const std::vector<float> tab =
      glst::plan_chebyshev(su.nodes, su.grid.h, groups);
  for (size_t g = 0; g < groups.size(); ++g) {
    const glst::dev::SfGroup &sg = groups[g];
    if (sg.m == 0) continue;
    CHECK(sg.m <= glst::dev::kChebMaxTerms, "group %zu: %u terms", g, sg.m);
    CHECK(g < su.nodes.kappa.size(), "group %zu has no radial nodes", g);
    if (g >= su.nodes.kappa.size()) continue;
    const uint64_t need = static_cast<uint64_t>(sg.coff) +
                          su.nodes.kappa[g].size() * sg.m;
    CHECK(need <= tab.size(), "group %zu reads %llu of %zu entries", g,
          static_cast<unsigned long long>(need), tab.size());
    CHECK(sg.nr == su.nodes.kappa[g].size(), "group %zu: nr %u vs %zu", g,
          sg.nr, su.nodes.kappa[g].size());
  }
}

}  // namespace

int main() {
  const Workload work[] = {
      {"1,000 waters", 3000, 31.04, 9.0, 1e-5},
      {"1,000 waters, L 44.7", 3000, 44.7, 9.0, 1e-4},
      {"96,000 waters", 288000, 142.0, 10.0, 1e-4},
      {"333,000 waters + 898 ion pairs", 1000796, 215.0, 12.0, 1e-5},
  };
  const size_t lds_limit = 160 * 1024;
  const uint32_t targets[] = {16384, 1024, 2048, 4096, 8192, 16384, 32768, 65536};
  for (const Workload &w : work) {
    for (int ct : {GLST_CT_FLOAT, GLST_CT_DOUBLE}) {
      const glst::SystemSetup su =
          glst::setup_system(ct, w.box, w.box, w.box, w.rcut, w.tol);
      const glst::NodeLayout L = glst::layout_nodes(su.nodes, su.K, su.Kp);
      check_layout(su, L);
      const uint32_t ntile = (w.natom + 255) / 256;
      for (uint32_t lr_blocks : targets) {
        const uint32_t target = (lr_blocks + ntile - 1) / ntile;
        const glst::SplitPlan P = glst::make_splits(su.nodes, su.K, target);
        check_splits(su, L, P, w.natom, target);
        for (uint32_t tmax : {1u, glst::dev::kRmtMaxNodes}) {
          const uint32_t T = glst::dev::rmt_nodes_per_group(
              su.grid.nx, su.grid.ny, su.grid.nz, lds_limit, tmax);
          if (T == 0) continue;
          check_pipeline(su, glst::plan_pipeline(su.nodes, P.sf_groups.size(),
                                                 su.Kp, T), T);
        }
        // the matrix-core shell sums take 1, 2 or 4 nodes per workgroup
        for (uint32_t T : {1u, 2u, 4u})
          check_pipeline(su, glst::plan_pipeline(su.nodes, P.sf_groups.size(),
                                                 su.Kp, T), T);
        if (ct == GLST_CT_FLOAT) check_chebyshev(su, P);
      }
      std::printf("%s (ct %d): grid %u^3, K %u, %zu groups: checked\n", w.name,
                  ct, su.grid.nx, su.K, su.nodes.dirs.size());
    }
  }
  std::printf("%d failures\n", failures);
  return failures ? 1 : 0;
}
