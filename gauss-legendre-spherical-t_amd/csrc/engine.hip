// The GLST engine behind the C ABI of include/glst_hip.h: one handle = one
// system on one GPU with one HIP stream.  Host setup follows
// src/glst_force.cu:178-327; the force step follows :431-1614.
#include "runtime.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace glst {

static std::mutex g_mutex;
static std::string g_last_error;
static std::string g_tdesign_dir;
static bool g_tdesign_init = false;

std::string tdesign_dir() {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_tdesign_init) {
    const char *env = std::getenv("GLST_TDESIGN_DIR");
    g_tdesign_dir = env ? env : "";
    g_tdesign_init = true;
  }
  return g_tdesign_dir;
}

void set_global_error(const std::string &m) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_last_error = m;
}

}  // namespace glst

using namespace glst;

struct glst_engine {
  std::string last_error;
  int ct = GLST_CT_DOUBLE;
  bool mixed() const { return ct == GLST_CT_FLOAT; }
  uint32_t natom = 0;
  double tol = 0.0;
  double box[3] = {0, 0, 0};
  dev::GridShape grid{};
  double inv_cell = 0.0;  // CT(1)/CT(rcut), widened
  AlphaGroups groups;
  NodeTable nodes;
  uint32_t K = 0, Kp = 0;
  bool has_lr = false;
  int spherical_rule = 0;
  float far_d = 3.0e38f;  // mixed: far-atom offset (far_offset)
  int device = 0;
  size_t lds_limit = 65536;

  hipStream_t stream = nullptr;
  bool own_stream = false;
  // the fused step runs the near field on a second stream (default
  // priority), overlapped with S(k) -> shell sums -> back-projection
  hipStream_t side = nullptr;
  hipEvent_t ev_assigned = nullptr, ev_near = nullptr;
  DeviceBuffer near4;
  // near field on the side stream: 0 off, 1 alongside S(k) onwards,
  // 2 alongside the shell sums onwards
  int overlap_near = 1;
  // shell sums: 0 LDS tables of as many nodes as fit, 1 one node per
  // workgroup, 2 global tables, 3 matrix cores where the grid allows (CT =
  // float, nx and ny up to 16), else as 0; 4 (default) matrix cores where
  // they apply and a z slice fills at least half of their 16 x 16 tile,
  // else the small-grid window sums where they apply (CT = float, every
  // dimension up to 8), else as 0; 5 the small-grid window sums where they
  // apply, else as 0
  int rmt_shape = 4;
  // mixed path, tables in LDS: pipeline S(k) -> shell sums ->
  // back-projection across alpha groups, the shell sums on a third stream
  int pipeline = 1;
  // mixed back-projection: 1 over atom groups (eight same-cell atoms share
  // each remote-sum load), 0 one atom per lane
  int lr_atom_groups = 1;
  // mixed structure factors: 1 the Chebyshev form for the alpha groups where
  // it saves work (plan_chebyshev), 0 the direct sum everywhere
  int sf_cheb = 1;
  hipStream_t rmt_stream = nullptr;
  std::vector<hipEvent_t> ev_sf, ev_rmt;  // per alpha group
  // back-projection workgroups aimed for (many short workgroups balance the
  // per-group launches better; the autotuner may pick another value)
  uint32_t lr_blocks = 16384;
  KernelTimer timer;
  // the whole step as a HIP graph for fixed input arrays (option "graph")
  bool use_graph = false;
  hipGraphExec_t graph_exec = nullptr;
  const double *graph_in[4] = {nullptr, nullptr, nullptr, nullptr};

  // atoms
  DeviceBuffer atom_cell, rank, hist, cell_count, cell_start, maxc;
  DeviceBuffer sorted_idx, cell_sorted, atoms64, atoms32;
  // mixed path: atom groups of the back-projection, first group of each cell
  DeviceBuffer atom_grp, ag_start;
  uint32_t nag_max = 0;
  DeviceBuffer fx, fy, fz, en;
  // nodes (device order: direction-major inside each group)
  DeviceBuffer k64, group, r_in, r_out;
  std::vector<uint32_t> dev_of_ref;  // device index of each reference node
  // back-projection tasks (one per group and direction) and their splits
  DeviceBuffer t_node0, t_nr, t_koff, t_dir64, t_dir32, kap64, kap32, t_split;
  uint32_t ntask = 0;
  // Chebyshev matrices of the mixed structure factors (plan_chebyshev)
  DeviceBuffer cheb;
  // fields
  DeviceBuffer S, R, partial;
  DeviceBuffer rmt_work;  // FP64 tables of the global shell-sum path
  uint32_t nsplit = 1;
  bool assigned = false;
  bool sf_done = false;
  bool rmt_done = false;

  ~glst_engine() {
    drop_graph();
    if (side) (void)hipStreamDestroy(side);
    if (rmt_stream) (void)hipStreamDestroy(rmt_stream);
    for (hipEvent_t e : ev_sf) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_rmt) (void)hipEventDestroy(e);
    if (ev_assigned) (void)hipEventDestroy(ev_assigned);
    if (ev_near) (void)hipEventDestroy(ev_near);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
  }

  size_t device_bytes() const {
    const DeviceBuffer *bufs[] = {&atom_cell, &rank,  &hist,       &cell_count,
                                  &cell_start, &maxc, &sorted_idx, &cell_sorted,
                                  &atoms64,   &fx,    &fy,         &fz,
                                  &en,        &k64,   &rmt_work,   &group,
                                  &r_in,      &r_out, &S,          &R,
                                  &partial,   &atoms32, &t_node0,   &t_nr,
                                  &t_koff,    &t_dir64, &t_dir32,   &kap64,
                                  &kap32,     &t_split, &near4, &atom_grp, &ag_start,
                                  &cheb};
    size_t s = 0;
    for (auto b : bufs) s += b->bytes();
    return s;
  }

  void init(int ct_, uint32_t natom_, double tol_, double bx, double by,
            double bz, double rcut, void *stream_) {
    check_system_args(ct_, bx, by, bz, rcut, tol_);
    ct = ct_;
    natom = natom_;
    tol = tol_;
    box[0] = bx;
    box[1] = by;
    box[2] = bz;
    GLST_HIP(hipGetDevice(&device));
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                              device) == hipSuccess &&
        lds > 0)
      lds_limit = static_cast<size_t>(lds);
    dev::prepare_kernels(lds_limit);
    if (stream_) {
      stream = static_cast<hipStream_t>(stream_);
      own_stream = false;
    } else {
      GLST_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      own_stream = true;
    }

    // Grid in CT precision (src/glst_force.cu:223-241), alpha groups and
    // cubature (src/glst_force.cu:1642-1724, src/cubature.cu:99-159).
    SystemSetup su = setup_system(ct_, bx, by, bz, rcut, tol_);
    grid = su.grid;
    inv_cell = su.inv_cell;
    has_lr = su.has_lr;
    groups = std::move(su.groups);
    nodes = std::move(su.nodes);
    spherical_rule = su.spherical_rule;
    K = su.K;
    Kp = su.Kp;
    far_d = far_offset(nodes);

    GLST_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    GLST_HIP(hipEventCreateWithFlags(&ev_assigned, hipEventDisableTiming));
    GLST_HIP(hipEventCreateWithFlags(&ev_near, hipEventDisableTiming));
    GLST_HIP(hipStreamCreateWithFlags(&rmt_stream, hipStreamNonBlocking));
    for (size_t g = 0; g < nodes.dirs.size(); ++g) {
      ev_sf.push_back(nullptr);
      GLST_HIP(hipEventCreateWithFlags(&ev_sf.back(), hipEventDisableTiming));
      ev_rmt.push_back(nullptr);
      GLST_HIP(hipEventCreateWithFlags(&ev_rmt.back(), hipEventDisableTiming));
    }
    allocate();
    upload_nodes();
    dev::fill_f64(fx.as<double>(), 0.0, natom, stream);
    dev::fill_f64(fy.as<double>(), 0.0, natom, stream);
    dev::fill_f64(fz.as<double>(), 0.0, natom, stream);
    dev::fill_f64(en.as<double>(), 0.0, natom, stream);
    // the mixed structure-factor kernel never writes the padded node columns
    dev::fill_f64(S.as<double>(), 0.0, S.bytes() / sizeof(double), stream);
    check_launch("init");
    GLST_HIP(hipStreamSynchronize(stream));
  }

  void allocate() {
    const size_t n = natom;
    const uint32_t nchunk = (natom + dev::kBinChunk - 1) / dev::kBinChunk;
    atom_cell.resize(n * 4);
    rank.resize(n * 4);
    hist.resize(static_cast<size_t>(nchunk) * grid.ncell * 4);
    cell_count.resize(static_cast<size_t>(grid.ncell) * 4);
    cell_start.resize((static_cast<size_t>(grid.ncell) + 1) * 4);
    maxc.resize(4);
    sorted_idx.resize(n * 4);
    cell_sorted.resize(n * 4);
    atoms64.resize(n * sizeof(double4));
    atoms32.resize(mixed() ? n * sizeof(float4) : 0);
    nag_max = mixed() && has_lr && n > 0 ? dev::atom_groups_max(n, grid.ncell)
                                         : 0;
    atom_grp.resize(static_cast<size_t>(nag_max) * sizeof(uint2));
    ag_start.resize(nag_max ? static_cast<size_t>(grid.ncell) * 4 : 0);
    near4.resize(n * sizeof(double4));
    fx.resize(n * 8);
    fy.resize(n * 8);
    fz.resize(n * 8);
    en.resize(n * 8);
    if (Kp > 0) {
      if (has_lr) {
        const size_t field = static_cast<size_t>(grid.ncell) * Kp;
        S.resize(field * (mixed() ? sizeof(float2) : sizeof(double2)));
        R.resize(field * (mixed() ? sizeof(float2) : sizeof(double2)));
        const uint32_t ntile = (natom + 255) / 256;
        // enough workgroups to fill the chip several times over
        set_splits(ntile > 0 ? (lr_blocks + ntile - 1) / ntile : 1);
        partial.resize(static_cast<size_t>(nsplit) * n * sizeof(double4));
      }
    }
  }

  // Back-projection splits (make_splits in runtime.hpp).
  std::vector<uint32_t> split_host;
  std::vector<dev::LrGroup> lr_groups;
  std::vector<dev::SfGroup> sf_groups;
  void set_splits(uint32_t target) {
    SplitPlan P = glst::make_splits(nodes, K, target);
    split_host = std::move(P.split);
    lr_groups = std::move(P.lr_groups);
    sf_groups = std::move(P.sf_groups);
    nsplit = P.nsplit;
    if (mixed()) {
      const std::vector<float> tab =
          glst::plan_chebyshev(nodes, grid.h, sf_groups);
      cheb.resize(tab.size() * sizeof(float));
      if (!tab.empty())
        GLST_HIP(hipMemcpy(cheb.as<void>(), tab.data(),
                           tab.size() * sizeof(float),
                           hipMemcpyHostToDevice));
    }
  }
  // the Chebyshev table when the option allows it (nullptr: direct sums)
  const float *cheb_table() const {
    return sf_cheb && mixed() && cheb.bytes() ? cheb.as<float>() : nullptr;
  }

  void upload_nodes() {
    if (Kp == 0) return;
    NodeLayout L = layout_nodes(nodes, K, Kp);
    dev_of_ref = std::move(L.dev_of_ref);
    ntask = static_cast<uint32_t>(L.node0.size());
    if (split_host.empty()) set_splits(1);
    const std::vector<uint32_t> &split = split_host;
    auto up = [&](DeviceBuffer &b, const void *src, size_t bytes) {
      b.resize(bytes);
      if (bytes)
        GLST_HIP(hipMemcpy(b.as<void>(), src, bytes, hipMemcpyHostToDevice));
    };
    up(k64, L.k64.data(), Kp * sizeof(double4));
    up(group, L.group.data(), Kp * 4);
    up(r_in, groups.r_in.data(), groups.r_in.size() * 4);
    up(r_out, groups.r_out.data(), groups.r_out.size() * 4);
    up(t_node0, L.node0.data(), L.node0.size() * 4);
    up(t_nr, L.nr.data(), L.nr.size() * 4);
    up(t_koff, L.koff.data(), L.koff.size() * 4);
    up(t_dir64, L.dir64.data(), L.dir64.size() * sizeof(double4));
    up(t_dir32, L.dir32.data(), L.dir32.size() * sizeof(float4));
    up(kap64, L.kappa64.data(), L.kappa64.size() * 8);
    up(kap32, L.kappa32.data(), L.kappa32.size() * 4);
    up(t_split, split.data(), split.size() * 4);
  }

  // atom groups for the mixed back-projection (nullptr: one atom per lane)
  const uint2 *lr_ag() const {
    return lr_atom_groups && nag_max ? atom_grp.as<uint2>() : nullptr;
  }

  // atom records the long-range kernels read
  const void *lr_atoms() const {
    return mixed() ? atoms32.as<void>() : atoms64.as<void>();
  }

  // ---- the step ------------------------------------------------------------

  void assign_atoms(const double *rx, const double *ry, const double *rz,
                    const double *qc) {
    if (natom > 0 && (!rx || !ry || !rz || !qc))
      throw ArgError("assign_atoms: null coordinate or charge pointer");
    const uint32_t nchunk = (natom + dev::kBinChunk - 1) / dev::kBinChunk;
    {
      Scope sc(timer, "bin", stream);
      dev::fill_u32(hist.as<uint32_t>(), 0u,
                    static_cast<size_t>(nchunk) * grid.ncell, stream);
      dev::bin_chunks(mixed(), rx, ry, rz, natom, inv_cell, grid,
                      atom_cell.as<uint32_t>(), rank.as<uint32_t>(),
                      hist.as<uint32_t>(), stream);
      check_launch("bin_chunks");
    }
    {
      Scope sc(timer, "sort", stream);
      dev::hist_offsets(hist.as<uint32_t>(), nchunk, grid.ncell,
                        cell_count.as<uint32_t>(), stream);
      uint2 *ag = nag_max ? atom_grp.as<uint2>() : nullptr;
      dev::cell_scan(cell_count.as<uint32_t>(), grid.ncell,
                     cell_start.as<uint32_t>(), maxc.as<uint32_t>(), stream,
                     ag_start.as<uint32_t>(), ag, nag_max);
      dev::bin_scatter(mixed(), rx, ry, rz, qc, natom,
                       atom_cell.as<uint32_t>(), rank.as<uint32_t>(),
                       hist.as<uint32_t>(), cell_start.as<uint32_t>(), grid,
                       sorted_idx.as<uint32_t>(), cell_sorted.as<uint32_t>(),
                       atoms64.as<double4>(), atoms32.as<float4>(), far_d,
                       stream, ag_start.as<uint32_t>(), ag);
      check_launch("sort");
    }
    assigned = true;
    sf_done = rmt_done = false;
  }

  void calc_sf() {
    if (!has_lr || Kp == 0) return;  // src/glst_force.cu:667-670
    if (!assigned) throw StateError("calc_sf called before assign_atoms");
    Scope sc(timer, "sf", stream);
    dev::sf(mixed(), lr_atoms(), cell_start.as<uint32_t>(),
            cell_sorted.as<uint32_t>(), k64.as<double4>(), task_args(),
            sf_groups, Kp, grid, S.as<void>(), stream, cheb_table());
    check_launch("sf");
    sf_done = true;
    rmt_done = false;
  }

  void sum_rmt_sf() {
    if (!has_lr || Kp == 0) return;
    if (!sf_done) throw StateError("sum_rmt_sf called before calc_sf");
    Scope sc(timer, "rmt", stream);
    if (rmt_mfma_path()) {
      dev::rmt_mfma(S.as<void>(), k64.as<double4>(), group.as<uint32_t>(),
                    r_in.as<uint32_t>(), r_out.as<uint32_t>(), Kp, grid,
                    lds_limit, R.as<void>(), stream);
    } else if (rmt_wave_path()) {
      dev::rmt_wave(S.as<void>(), k64.as<double4>(), group.as<uint32_t>(),
                    r_in.as<uint32_t>(), r_out.as<uint32_t>(), Kp, grid,
                    R.as<void>(), stream);
    } else if (rmt_shape == 2 ||
        !dev::rmt_lds(mixed(), rmt_shape == 1 ? 1 : dev::kRmtMaxNodes,
                      S.as<void>(),
                      k64.as<double4>(), group.as<uint32_t>(),
                      r_in.as<uint32_t>(), r_out.as<uint32_t>(), Kp, grid,
                      lds_limit, R.as<void>(), stream)) {
      // tables too large for LDS: FP64 work field, allocated on first use
      rmt_work.resize(static_cast<size_t>(grid.ncell) * Kp * sizeof(double2));
      dev::rmt_global(mixed(), S.as<void>(), rmt_work.as<double2>(),
                      k64.as<double4>(), group.as<uint32_t>(),
                      r_in.as<uint32_t>(), r_out.as<uint32_t>(), Kp, grid,
                      R.as<void>(), stream);
    }
    check_launch("rmt");
    rmt_done = true;
  }

  void calc_lr_ef() {
    if (!has_lr || Kp == 0 || natom == 0) return;
    if (!rmt_done) throw StateError("calc_lr_ef called before sum_rmt_sf");
    lr_partials();
    reduce(false);
  }

  // fx..en += the back-projection partials (and the near-field record of the
  // fused step when with_near)
  void reduce(bool with_near) {
    Scope sc(timer, "lr_reduce", stream);
    const bool lr = has_lr && Kp > 0;
    dev::lr_reduce(partial.as<double4>(), lr ? nsplit : 0, natom,
                   with_near ? near4.as<double4>() : nullptr, fx.as<double>(),
                   fy.as<double>(), fz.as<double>(), en.as<double>(), stream);
    check_launch("lr_reduce");
  }

  dev::LrTaskArgs task_args() const {
    return {t_node0.as<uint32_t>(), t_nr.as<uint32_t>(),
            t_koff.as<uint32_t>(),  t_dir64.as<double4>(),
            t_dir32.as<float4>(),   kap64.as<double>(),
            kap32.as<float>(),      t_split.as<uint32_t>()};
  }

  void lr_partials() {
    Scope sc(timer, "lr", stream);
    const dev::LrTaskArgs ta = task_args();
    dev::lr(mixed(), lr_atoms(), cell_sorted.as<uint32_t>(), ta, lr_groups,
            R.as<void>(), Kp, natom, nsplit, partial.as<double4>(), stream,
            lr_ag(), nag_max);
    check_launch("lr");
  }

  void calc_sr_ef() {
    if (natom == 0) return;
    if (!assigned) throw StateError("calc_sr_ef called before assign_atoms");
    Scope sc(timer, "sr", stream);
    dev::sr(mixed(), atoms64.as<double4>(), cell_start.as<uint32_t>(), grid,
            fx.as<double>(), fy.as<double>(), fz.as<double>(), en.as<double>(),
            nullptr, stream);
    check_launch("sr");
  }

  void comm_ef() { GLST_HIP(hipStreamSynchronize(stream)); }

  // ---- the far field of the fused step, pipelined across alpha groups ----
  //
  // Alpha group g owns device nodes [first_g, first_g + count_g).  Shell-sum
  // chunk g is the workgroup tiles [tile_end[g-1], tile_end[g]) of T nodes:
  // tile_end[g] = floor((first_g + count_g) / T), so chunk g holds nodes of
  // groups <= g only and may start once S(k) of group g is complete; a tile
  // that straddles groups g and g+1 belongs to chunk g+1, and the
  // back-projection of group g then waits for chunk rmt_need[g] = g+1.  The
  // LDS-bound shell sums of one group so run alongside the VALU-bound S(k)
  // of later groups and back-projection of earlier ones.
  uint32_t pipe_tmax = 0;
  std::vector<uint32_t> tile_end, rmt_need;
  bool pipelined() const {
    return pipeline && mixed() && has_lr && Kp > 0 && !sf_groups.empty() &&
           ev_sf.size() == sf_groups.size() &&
           lr_groups.size() == sf_groups.size() && !rmt_global_path();
  }
  void plan_pipeline() {
    pipe_tmax = rmt_shape == 1 ? 1 : dev::kRmtMaxNodes;
    const uint32_t T =
        rmt_mfma_path()
            ? dev::rmt_mfma_nodes(grid, lds_limit)
        : rmt_wave_path()
            ? dev::kRmtWaveNodes
            : dev::rmt_nodes_per_group(grid.nx, grid.ny, grid.nz, lds_limit,
                                       pipe_tmax);
    if (T == 0) throw StateError("shell-sum tables do not fit in LDS");
    PipelinePlan P = glst::plan_pipeline(nodes, sf_groups.size(), Kp, T);
    tile_end = std::move(P.tile_end);
    rmt_need = std::move(P.rmt_need);
  }
  void sf_stage() {
    if (!pipelined()) {
      calc_sf();
      return;
    }
    plan_pipeline();
    {
      Scope sc(timer, "sf", stream);
      const dev::LrTaskArgs ta = task_args();
      for (size_t g = 0; g < sf_groups.size(); ++g) {
        dev::sf(true, lr_atoms(), cell_start.as<uint32_t>(),
                cell_sorted.as<uint32_t>(), k64.as<double4>(), ta,
                &sf_groups[g], 1, Kp, grid, S.as<void>(), stream,
                cheb_table());
        GLST_HIP(hipEventRecord(ev_sf[g], stream));
      }
      check_launch("sf");
    }
    sf_done = true;
    rmt_done = false;
  }
  void rmt_stage() {
    if (!pipelined()) {
      sum_rmt_sf();
      return;
    }
    GLST_HIP(hipStreamWaitEvent(rmt_stream, ev_sf[0], 0));
    {
      Scope sc(timer, "rmt", rmt_stream);
      for (size_t g = 0; g < sf_groups.size(); ++g) {
        if (g > 0) GLST_HIP(hipStreamWaitEvent(rmt_stream, ev_sf[g], 0));
        const uint32_t t0 = g > 0 ? tile_end[g - 1] : 0;
        const bool ok =
            rmt_mfma_path()
                ? dev::rmt_mfma(S.as<void>(), k64.as<double4>(),
                                group.as<uint32_t>(), r_in.as<uint32_t>(),
                                r_out.as<uint32_t>(), Kp, grid, lds_limit,
                                R.as<void>(), rmt_stream, t0, tile_end[g])
            : rmt_wave_path()
                ? dev::rmt_wave(S.as<void>(), k64.as<double4>(),
                                group.as<uint32_t>(), r_in.as<uint32_t>(),
                                r_out.as<uint32_t>(), Kp, grid, R.as<void>(),
                                rmt_stream, t0, tile_end[g])
                : dev::rmt_lds(true, pipe_tmax, S.as<void>(),
                               k64.as<double4>(), group.as<uint32_t>(),
                               r_in.as<uint32_t>(), r_out.as<uint32_t>(), Kp,
                               grid, lds_limit, R.as<void>(), rmt_stream, t0,
                               tile_end[g]);
        if (!ok) throw StateError("shell-sum tables do not fit in LDS");
        GLST_HIP(hipEventRecord(ev_rmt[g], rmt_stream));
      }
      check_launch("rmt");
    }
    rmt_done = true;
  }
  void lr_stage() {
    if (!(has_lr && Kp > 0)) return;
    if (!pipelined()) {
      lr_partials();
      return;
    }
    Scope sc(timer, "lr", stream);
    const dev::LrTaskArgs ta = task_args();
    for (size_t g = 0; g < lr_groups.size(); ++g) {
      GLST_HIP(hipStreamWaitEvent(stream, ev_rmt[rmt_need[g]], 0));
      dev::lr(true, lr_atoms(), cell_sorted.as<uint32_t>(), ta, &lr_groups[g],
              1, R.as<void>(), Kp, natom, nsplit, partial.as<double4>(),
              stream, lr_ag(), nag_max);
    }
    check_launch("lr");
  }

  // The fused step: the near field (which needs only the sorted atoms) runs
  // on the side stream while the main stream does S(k), the shell sums and
  // the back-projection; one fixed-order reduction adds both parts.
  void step_async(const double *rx, const double *ry, const double *rz,
                  const double *qc) {
    assign_atoms(rx, ry, rz, qc);
    if (natom == 0) return;
    if (!overlap_near) {
      sf_stage();
      rmt_stage();
      if (has_lr && Kp > 0) {
        lr_stage();
        reduce(false);
      }
      calc_sr_ef();
      return;
    }
    auto launch_near = [&] {
      GLST_HIP(hipEventRecord(ev_assigned, stream));
      GLST_HIP(hipStreamWaitEvent(side, ev_assigned, 0));
      {
        Scope sc(timer, "sr", side);
        dev::sr(mixed(), atoms64.as<double4>(), cell_start.as<uint32_t>(),
                grid, nullptr, nullptr, nullptr, nullptr, near4.as<double4>(),
                side);
        check_launch("sr");
      }
      GLST_HIP(hipEventRecord(ev_near, side));
    };
    // Mode 1 starts the near field with S(k); mode 2 after all of S(k), so
    // that it runs beside the shell sums and the back-projection (the
    // autotuner's choice at 288k atoms).  Starting it after the first alpha
    // group's S(k) instead was 6 % slower.
    if (overlap_near == 1 || !(has_lr && Kp > 0)) launch_near();
    sf_stage();
    if (overlap_near == 2 && has_lr && Kp > 0) launch_near();
    rmt_stage();
    lr_stage();
    GLST_HIP(hipStreamWaitEvent(stream, ev_near, 0));
    reduce(true);
  }

  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
  }

  // shell sums through the global-memory tables (one node's table does not
  // fit in LDS, or forced)
  bool rmt_global_path() const {
    if (rmt_mfma_path() || rmt_wave_path()) return false;
    return rmt_shape == 2 ||
           dev::rmt_nodes_per_group(grid.nx, grid.ny, grid.nz, lds_limit,
                                    1) == 0;
  }
  // shell sums on the matrix cores (CT = float, nx and ny up to 16)
  bool rmt_mfma_path() const {
    const bool want =
        rmt_shape == 3 ||
        (rmt_shape == 4 && grid.nx * grid.ny * 2 >= dev::kRmtMfmaMaxXY *
                                                        dev::kRmtMfmaMaxXY);
    return want && mixed() && dev::rmt_mfma_nodes(grid, lds_limit) > 0;
  }
  // shell sums as window sums of small grids (CT = float, every dimension
  // up to 8)
  bool rmt_wave_path() const {
    const bool want = rmt_shape == 5 || (rmt_shape == 4 && !rmt_mfma_path());
    return want && mixed() && dev::rmt_wave_fits(grid);
  }

  // New back-projection split target: splits, their table and the partials.
  void apply_lr_blocks() {
    drop_graph();
    if (!(has_lr && Kp > 0)) return;
    const uint32_t ntile = (natom + 255) / 256;
    set_splits(ntile > 0 ? (lr_blocks + ntile - 1) / ntile : 1);
    partial.resize(static_cast<size_t>(nsplit) * natom * sizeof(double4));
    t_split.resize(split_host.size() * 4);
    GLST_HIP(hipMemcpy(t_split.as<void>(), split_host.data(),
                       split_host.size() * 4, hipMemcpyHostToDevice));
  }

  // calc_ener_force: eagerly, or (graph mode) by replaying the step captured
  // for these input arrays.
  void step(const double *rx, const double *ry, const double *rz,
            const double *qc) {
    if (!use_graph || natom == 0) {
      step_async(rx, ry, rz, qc);
      return;
    }
    if (!rx || !ry || !rz || !qc)
      throw ArgError("assign_atoms: null coordinate or charge pointer");
    const double *in[4] = {rx, ry, rz, qc};
    if (!graph_exec || !std::equal(in, in + 4, graph_in)) {
      drop_graph();
      // nothing may allocate inside the capture
      if (has_lr && Kp > 0 && rmt_global_path())
        rmt_work.resize(static_cast<size_t>(grid.ncell) * Kp * sizeof(double2));
      const bool timing = timer.enabled;
      timer.enabled = false;
      hipGraph_t g = nullptr;
      GLST_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        step_async(rx, ry, rz, qc);
      } catch (...) {
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        timer.enabled = timing;
        throw;
      }
      const hipError_t ec = hipStreamEndCapture(stream, &g);
      timer.enabled = timing;
      GLST_HIP(ec);
      const hipError_t ei = hipGraphInstantiate(&graph_exec, g, nullptr,
                                                nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) graph_exec = nullptr;
      GLST_HIP(ei);
      std::copy(in, in + 4, graph_in);
    }
    GLST_HIP(hipGraphLaunch(graph_exec, stream));
  }

  // Times `reps` steps on these inputs (HIP events on the handle's stream).
  double time_steps(const double *rx, const double *ry, const double *rz,
                    const double *qc, uint32_t reps) {
    step(rx, ry, rz, qc);  // warm-up (and capture in graph mode)
    hipEvent_t a = nullptr, b = nullptr;
    GLST_HIP(hipEventCreate(&a));
    GLST_HIP(hipEventCreate(&b));
    float ms = 0.f;
    try {
      GLST_HIP(hipEventRecord(a, stream));
      for (uint32_t r = 0; r < reps; ++r) step(rx, ry, rz, qc);
      GLST_HIP(hipEventRecord(b, stream));
      GLST_HIP(hipEventSynchronize(b));
      GLST_HIP(hipEventElapsedTime(&ms, a, b));
    } catch (...) {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return static_cast<double>(ms) / reps;
  }

  // Chooses the fastest schedule for these inputs, one knob at a time: graph
  // replay, the near-field overlap, the back-projection split target, the
  // shell-sum workgroup shape (FP64 handles: glst_autotune refuses mixed
  // ones).  Results are identical under every schedule up to the split
  // boundaries of the back-projection; the accumulators are cleared.
  double autotune(const double *rx, const double *ry, const double *rz,
                  const double *qc, uint32_t reps) {
    if (reps == 0) throw ArgError("autotune needs at least one repetition");
    double best = time_steps(rx, ry, rz, qc, reps);
    auto knob = [&](auto get, auto set, std::vector<int64_t> values) {
      const int64_t keep = get();
      int64_t pick = keep;
      for (int64_t v : values) {
        if (v == keep) continue;
        set(v);
        const double t = time_steps(rx, ry, rz, qc, reps);
        if (t < best) {
          best = t;
          pick = v;
        }
      }
      if (get() != pick) set(pick);
    };
    // graph replay changes how the two streams' kernels are dispatched, so
    // it is tuned first and again last, with the other knobs in between
    auto graph_knob = [&] {
      knob([&] { return static_cast<int64_t>(use_graph ? 1 : 0); },
           [&](int64_t v) {
             use_graph = v != 0;
             drop_graph();
           },
           {0, 1});
    };
    graph_knob();
    knob([&] { return static_cast<int64_t>(overlap_near); },
         [&](int64_t v) {
           overlap_near = static_cast<int>(v);
           drop_graph();
         },
         {0, 1, 2});
    if (has_lr && Kp > 0) {
      knob([&] { return static_cast<int64_t>(lr_blocks); },
           [&](int64_t v) {
             lr_blocks = static_cast<uint32_t>(v);
             apply_lr_blocks();
           },
           {1024, 2048, 4096, 8192, 16384, 32768, 65536});
      if (!rmt_global_path())
        knob([&] { return static_cast<int64_t>(rmt_shape); },
             [&](int64_t v) {
               rmt_shape = static_cast<int>(v);
               drop_graph();
             },
             {0, 1});
    }
    graph_knob();
    zero_ef();
    GLST_HIP(hipStreamSynchronize(stream));
    return best;
  }

  void get_ef(double *ofx, double *ofy, double *ofz, double *oen) {
    if (natom > 0 && (!ofx || !ofy || !ofz || !oen))
      throw ArgError("get_ef: null output pointer");
    if (natom > 0 && !assigned)
      throw StateError("get_ef called before assign_atoms");
    {
      Scope sc(timer, "unpermute", stream);
      dev::unpermute(sorted_idx.as<uint32_t>(), natom, fx.as<double>(),
                     fy.as<double>(), fz.as<double>(), en.as<double>(), ofx,
                     ofy, ofz, oen, stream);
      check_launch("unpermute");
    }
    GLST_HIP(hipStreamSynchronize(stream));
  }

  void zero_ef() {
    dev::fill_f64(fx.as<double>(), 0.0, natom, stream);
    dev::fill_f64(fy.as<double>(), 0.0, natom, stream);
    dev::fill_f64(fz.as<double>(), 0.0, natom, stream);
    dev::fill_f64(en.as<double>(), 0.0, natom, stream);
    check_launch("zero_ef");
  }

  uint32_t max_atoms_cell() {
    if (!assigned) return 0;
    uint32_t m = 0;
    GLST_HIP(hipMemcpyAsync(&m, maxc.as<void>(), 4, hipMemcpyDeviceToHost,
                            stream));
    GLST_HIP(hipStreamSynchronize(stream));
    return m;
  }

  template <typename T>
  void d2h(T *dst, const DeviceBuffer &src, size_t count) {
    if (!dst || count == 0) return;
    GLST_HIP(hipMemcpyAsync(dst, src.as<void>(), count * sizeof(T),
                            hipMemcpyDeviceToHost, stream));
  }

  // Device field (R: c*Kp + k; S: node runs, sf_slot) to host arrays of
  // ncell x K in reference node order.  FP32 fields hold `fstride` floats
  // per element, the first two being (re, im).
  void field_to_host(const DeviceBuffer &buf, bool single, uint32_t fstride,
                     bool tiled, double *re, double *im) {
    const size_t n = static_cast<size_t>(grid.ncell) * Kp;
    std::vector<double> tmp(2 * n);
    if (single) {
      std::vector<float> f(fstride * n);
      GLST_HIP(hipMemcpyAsync(f.data(), buf.as<void>(), f.size() * 4,
                              hipMemcpyDeviceToHost, stream));
      GLST_HIP(hipStreamSynchronize(stream));
      for (size_t i = 0; i < n; ++i) {
        tmp[2 * i] = f[fstride * i];
        tmp[2 * i + 1] = f[fstride * i + 1];
      }
    } else {
      GLST_HIP(hipMemcpyAsync(tmp.data(), buf.as<void>(), tmp.size() * 8,
                              hipMemcpyDeviceToHost, stream));
      GLST_HIP(hipStreamSynchronize(stream));
    }
    for (uint32_t c = 0; c < grid.ncell; ++c)
      for (uint32_t k = 0; k < K; ++k) {
        const uint32_t node = dev_of_ref[k];
        const size_t s =
            !tiled ? static_cast<size_t>(c) * Kp + node
            : single ? dev::sf_slot<true>(node, c, grid.ncell)
                     : dev::sf_slot<false>(node, c, grid.ncell);
        const size_t d = static_cast<size_t>(c) * K + k;
        if (re) re[d] = tmp[2 * s];
        if (im) im[d] = tmp[2 * s + 1];
      }
  }
};

// ---------------------------------------------------------------------------
// C ABI

namespace {

void record_error(glst_engine *h, const std::string &m) {
  if (h)
    h->last_error = m;
  else
    set_global_error(m);
}

// Runs f, mapping exceptions to status codes and the handle's (or the
// process-wide) last-error string.
template <class F>
glst_status_t guarded(glst_engine *h, F &&f) {
  try {
    f();
    if (h) h->last_error.clear();
    return GLST_OK;
  } catch (const ArgError &e) {
    record_error(h, e.what());
    return GLST_ERR_INVALID;
  } catch (const StateError &e) {
    record_error(h, e.what());
    return GLST_ERR_STATE;
  } catch (const AllocError &e) {
    record_error(h, e.what());
    return GLST_ERR_ALLOC;
  } catch (const HipError &e) {
    record_error(h, e.what());
    return GLST_ERR_HIP;
  } catch (const std::exception &e) {
    const std::string m = e.what();
    record_error(h, m);
    if (m.find("open") != std::string::npos ||
        m.find("Malformed") != std::string::npos ||
        m.find("directory") != std::string::npos)
      return GLST_ERR_IO;
    return GLST_ERR_NUMERICS;
  }
}

#define NEED_HANDLE(h)                                                        \
  if (!(h)) {                                                                 \
    set_global_error("null handle");                                          \
    return GLST_ERR_INVALID;                                                  \
  }

}  // namespace

extern "C" {

glst_status_t glst_create(int ct, uint32_t natom, double tol, double box_x,
                          double box_y, double box_z, double rcut,
                          void *stream, glst_handle_t *out) {
  if (!out) {
    set_global_error("glst_create: null output handle");
    return GLST_ERR_INVALID;
  }
  *out = nullptr;
  glst_engine *h = new (std::nothrow) glst_engine();
  if (!h) return GLST_ERR_ALLOC;
  glst_status_t st = guarded(nullptr, [&] {
    if (!(tol > 0)) throw ArgError("tol must be positive");
    h->init(ct, natom, tol, box_x, box_y, box_z, rcut, stream);
  });
  if (st != GLST_OK) {
    delete h;
    return st;
  }
  *out = h;
  return GLST_OK;
}

glst_status_t glst_create_ncell(int ct, uint32_t natom, double tol,
                                double box_x, double box_y, double box_z,
                                uint32_t ncell_x, uint32_t ncell_y,
                                uint32_t ncell_z, void *stream,
                                glst_handle_t *out) {
  if (ncell_x == 0 || ncell_y == 0 || ncell_z == 0) {
    set_global_error("glst_create_ncell: cell counts must be positive");
    if (out) *out = nullptr;
    return GLST_ERR_INVALID;
  }
  // src/glst_force.cu:87-94
  const double rx = box_x / static_cast<double>(ncell_x);
  const double ry = box_y / static_cast<double>(ncell_y);
  const double rz = box_z / static_cast<double>(ncell_z);
  double rcut = rx;
  rcut = (ry < rcut) ? ry : rcut;
  rcut = (rz < rcut) ? rz : rcut;
  return glst_create(ct, natom, tol, box_x, box_y, box_z, rcut, stream, out);
}

glst_status_t glst_destroy(glst_handle_t h) {
  if (!h) return GLST_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  delete h;
  return GLST_OK;
}

glst_status_t glst_set_stream(glst_handle_t h, void *stream) {
  NEED_HANDLE(h);
  return guarded(h, [&] {
    GLST_HIP(hipStreamSynchronize(h->stream));
    h->drop_graph();
    if (h->own_stream) GLST_HIP(hipStreamDestroy(h->stream));
    if (stream) {
      h->stream = static_cast<hipStream_t>(stream);
      h->own_stream = false;
    } else {
      GLST_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
      h->own_stream = true;
    }
  });
}

glst_status_t glst_assign_atoms(glst_handle_t h, const double *d_rx,
                                const double *d_ry, const double *d_rz,
                                const double *d_qc) {
  NEED_HANDLE(h);
  return guarded(h, [&] { h->assign_atoms(d_rx, d_ry, d_rz, d_qc); });
}

glst_status_t glst_calc_sf(glst_handle_t h) {
  NEED_HANDLE(h);
  return guarded(h, [&] { h->calc_sf(); });
}

glst_status_t glst_sum_rmt_sf(glst_handle_t h) {
  NEED_HANDLE(h);
  return guarded(h, [&] { h->sum_rmt_sf(); });
}

glst_status_t glst_calc_lr_ef(glst_handle_t h) {
  NEED_HANDLE(h);
  return guarded(h, [&] { h->calc_lr_ef(); });
}

glst_status_t glst_calc_sr_ef(glst_handle_t h) {
  NEED_HANDLE(h);
  return guarded(h, [&] { h->calc_sr_ef(); });
}

glst_status_t glst_comm_ef(glst_handle_t h) {
  NEED_HANDLE(h);
  return guarded(h, [&] { h->comm_ef(); });
}

glst_status_t glst_calc_ener_force(glst_handle_t h, const double *d_rx,
                                   const double *d_ry, const double *d_rz,
                                   const double *d_qc) {
  NEED_HANDLE(h);
  return guarded(h, [&] {
    h->step(d_rx, d_ry, d_rz, d_qc);
    h->comm_ef();
  });
}

glst_status_t glst_calc_ener_force_async(glst_handle_t h, const double *d_rx,
                                         const double *d_ry,
                                         const double *d_rz,
                                         const double *d_qc) {
  NEED_HANDLE(h);
  return guarded(h, [&] { h->step(d_rx, d_ry, d_rz, d_qc); });
}

glst_status_t glst_get_ef(glst_handle_t h, double *d_fx, double *d_fy,
                          double *d_fz, double *d_en) {
  NEED_HANDLE(h);
  return guarded(h, [&] { h->get_ef(d_fx, d_fy, d_fz, d_en); });
}

glst_status_t glst_sorted_ef_ptrs(glst_handle_t h, double **fx, double **fy,
                                  double **fz, double **en) {
  NEED_HANDLE(h);
  if (fx) *fx = h->fx.as<double>();
  if (fy) *fy = h->fy.as<double>();
  if (fz) *fz = h->fz.as<double>();
  if (en) *en = h->en.as<double>();
  return GLST_OK;
}

glst_status_t glst_zero_ef(glst_handle_t h) {
  NEED_HANDLE(h);
  return guarded(h, [&] { h->zero_ef(); });
}

glst_status_t glst_sorted_index_ptr(glst_handle_t h, const uint32_t **idx) {
  NEED_HANDLE(h);
  if (!idx) return GLST_ERR_INVALID;
  *idx = h->sorted_idx.as<uint32_t>();
  return GLST_OK;
}

glst_status_t glst_get_info(glst_handle_t h, glst_info_t *info) {
  NEED_HANDLE(h);
  if (!info) return GLST_ERR_INVALID;
  return guarded(h, [&] {
    std::memset(info, 0, sizeof(*info));
    info->natom = h->natom;
    info->ncell[0] = h->grid.nx;
    info->ncell[1] = h->grid.ny;
    info->ncell[2] = h->grid.nz;
    info->ncell_total = h->grid.ncell;
    info->cell_dim[0] = info->cell_dim[1] = info->cell_dim[2] = h->grid.h;
    info->ngroup = static_cast<uint32_t>(h->groups.width.size());
    info->num_nodes = h->K;
    info->has_long_range = h->has_lr ? 1u : 0u;
    info->max_atoms_cell = h->max_atoms_cell();
    info->ct = h->ct;
    info->spherical_rule = h->spherical_rule;
    info->device_bytes = h->device_bytes();
  });
}

glst_status_t glst_get_group_info(glst_handle_t h, uint32_t g,
                                  glst_group_info_t *info) {
  NEED_HANDLE(h);
  if (!info || g >= h->groups.width.size()) {
    h->last_error = "glst_get_group_info: group index out of range";
    return GLST_ERR_INVALID;
  }
  info->width = h->groups.width[g];
  info->r_in = h->groups.r_in[g];
  info->r_out = h->groups.r_out[g];
  info->rmin = h->groups.rmin[g];
  info->rmax = h->groups.rmax[g];
  info->alpha = h->groups.alpha[g];
  info->zcut = h->groups.zcut[g];
  info->gl_order = h->nodes.gl[g].order;
  info->sph_degree = 2 * h->nodes.sph_m[g] + 1;
  info->num_dirs = h->nodes.sph_count[g];
  info->first_node = h->nodes.first[g];
  info->num_nodes = h->nodes.count[g];
  return GLST_OK;
}

glst_status_t glst_get_nodes(glst_handle_t h, double *kx, double *ky,
                             double *kz, double *w, uint32_t *group) {
  NEED_HANDLE(h);
  const size_t K = h->K;
  if (kx) std::memcpy(kx, h->nodes.kx.data(), K * 8);
  if (ky) std::memcpy(ky, h->nodes.ky.data(), K * 8);
  if (kz) std::memcpy(kz, h->nodes.kz.data(), K * 8);
  if (w) std::memcpy(w, h->nodes.w.data(), K * 8);
  if (group) std::memcpy(group, h->nodes.group.data(), K * 4);
  return GLST_OK;
}

glst_status_t glst_get_cells(glst_handle_t h, uint32_t *atom_cell,
                             uint32_t *sorted_idx, uint32_t *cell_point,
                             uint32_t *cell_count) {
  NEED_HANDLE(h);
  return guarded(h, [&] {
    if (!h->assigned) throw StateError("glst_get_cells before assign_atoms");
    h->d2h(atom_cell, h->atom_cell, h->natom);
    h->d2h(sorted_idx, h->sorted_idx, h->natom);
    h->d2h(cell_point, h->cell_start, h->grid.ncell);
    h->d2h(cell_count, h->cell_count, h->grid.ncell);
    GLST_HIP(hipStreamSynchronize(h->stream));
  });
}

glst_status_t glst_get_sf(glst_handle_t h, double *re, double *im) {
  NEED_HANDLE(h);
  return guarded(h, [&] {
    if (!h->has_lr || h->K == 0) return;
    if (!h->sf_done) throw StateError("glst_get_sf: no structure factors");
    const size_t n = static_cast<size_t>(h->grid.ncell) * h->K;
    std::vector<double> r(n), i(n);
    h->field_to_host(h->S, h->mixed(), 2, true, r.data(), i.data());
    const auto &g = h->grid;
    if (h->mixed()) {
      // the field holds cell-local sums; apply exp(-i k.c)
      for (uint32_t c = 0; c < g.ncell; ++c) {
        const uint32_t x = c / (g.ny * g.nz), y = (c / g.nz) % g.ny,
                       z = c % g.nz;
        const double cx = (x + 0.5) * g.h, cy = (y + 0.5) * g.h,
                     cz = (z + 0.5) * g.h;
        for (uint32_t k = 0; k < h->K; ++k) {
          const double th =
              h->nodes.kx[k] * cx + h->nodes.ky[k] * cy + h->nodes.kz[k] * cz;
          const double cs = std::cos(th), sn = std::sin(th);
          const size_t o = static_cast<size_t>(c) * h->K + k;
          const double a = r[o], b = i[o];
          r[o] = a * cs + b * sn;
          i[o] = b * cs - a * sn;
        }
      }
      if (re) std::memcpy(re, r.data(), n * 8);
      if (im) std::memcpy(im, i.data(), n * 8);
      return;
    }
    // the FP64 field holds running sums along z; difference them back
    for (uint32_t col = 0; col < g.nx * g.ny; ++col)
      for (uint32_t z = g.nz - 1; z >= 1; --z) {
        const size_t c = static_cast<size_t>(col) * g.nz + z;
        for (uint32_t k = 0; k < h->K; ++k) {
          r[c * h->K + k] -= r[(c - 1) * h->K + k];
          i[c * h->K + k] -= i[(c - 1) * h->K + k];
        }
      }
    if (re) std::memcpy(re, r.data(), n * 8);
    if (im) std::memcpy(im, i.data(), n * 8);
  });
}

glst_status_t glst_get_rmt(glst_handle_t h, double *re, double *im) {
  NEED_HANDLE(h);
  return guarded(h, [&] {
    if (!h->has_lr || h->K == 0) return;
    if (!h->rmt_done) throw StateError("glst_get_rmt: no remote sums");
    const size_t n = static_cast<size_t>(h->grid.ncell) * h->K;
    std::vector<double> r(n), i(n);
    h->field_to_host(h->R, h->mixed(), 2, false, r.data(), i.data());
    if (h->mixed()) {
      // the stored (kappa |V|, arg V / 2 pi) back to V = exp(+i k.c) R_c(k),
      // then undo the exp(+i k.c) factor in FP64
      const auto &g = h->grid;
      const double two_pi = 2.0 * M_PI;
      std::vector<double> inv_kappa(h->K);
      for (uint32_t k = 0; k < h->K; ++k)
        inv_kappa[k] = two_pi / std::sqrt(h->nodes.kx[k] * h->nodes.kx[k] +
                                          h->nodes.ky[k] * h->nodes.ky[k] +
                                          h->nodes.kz[k] * h->nodes.kz[k]);
      for (uint32_t c = 0; c < g.ncell; ++c) {
        const uint32_t x = c / (g.ny * g.nz), y = (c / g.nz) % g.ny,
                       z = c % g.nz;
        const double cx = (x + 0.5) * g.h, cy = (y + 0.5) * g.h,
                     cz = (z + 0.5) * g.h;
        for (uint32_t k = 0; k < h->K; ++k) {
          const double th =
              h->nodes.kx[k] * cx + h->nodes.ky[k] * cy + h->nodes.kz[k] * cz;
          const double cs = std::cos(th), sn = std::sin(th);
          const size_t o = static_cast<size_t>(c) * h->K + k;
          const double mag = r[o] * inv_kappa[k], arg = two_pi * i[o];
          const double a = mag * std::cos(arg), b = mag * std::sin(arg);
          r[o] = a * cs + b * sn;
          i[o] = b * cs - a * sn;
        }
      }
    }
    if (re) std::memcpy(re, r.data(), n * 8);
    if (im) std::memcpy(im, i.data(), n * 8);
  });
}

glst_status_t glst_set_option(glst_handle_t h, const char *name,
                              int64_t value) {
  NEED_HANDLE(h);
  if (!name) return GLST_ERR_INVALID;
  return guarded(h, [&] {
    const std::string n = name;
    h->drop_graph();  // every option changes the captured step
    if (n == "overlap_near") {
      if (value < 0 || value > 2) throw ArgError("overlap_near must be 0, 1 or 2");
      h->overlap_near = static_cast<int>(value);
    }
    else if (n == "rmt_shape") {
      if (value < 0 || value > 5)
        throw ArgError("rmt_shape must be 0 .. 5");
      h->rmt_shape = static_cast<int>(value);
    } else if (n == "pipeline") {
      if (value < 0 || value > 1) throw ArgError("pipeline must be 0 or 1");
      h->pipeline = static_cast<int>(value);
    } else if (n == "lr_blocks") {
      if (value < 1 || value > (1 << 24))
        throw ArgError("lr_blocks must be in 1 .. 2^24");
      h->lr_blocks = static_cast<uint32_t>(value);
      h->apply_lr_blocks();
    } else if (n == "graph") {
      if (value != 0 && value != 1) throw ArgError("graph must be 0 or 1");
      h->use_graph = value != 0;
    } else if (n == "lr_atom_groups") {
      if (value != 0 && value != 1)
        throw ArgError("lr_atom_groups must be 0 or 1");
      h->lr_atom_groups = static_cast<int>(value);
    } else if (n == "sf_cheb") {
      if (value != 0 && value != 1) throw ArgError("sf_cheb must be 0 or 1");
      h->sf_cheb = static_cast<int>(value);
    } else
      throw ArgError("unknown option \"" + n + "\"");
  });
}

glst_status_t glst_get_option(glst_handle_t h, const char *name,
                              int64_t *value) {
  NEED_HANDLE(h);
  if (!name || !value) return GLST_ERR_INVALID;
  return guarded(h, [&] {
    const std::string n = name;
    if (n == "overlap_near")
      *value = h->overlap_near;
    else if (n == "rmt_shape")
      *value = h->rmt_shape;
    else if (n == "pipeline")
      *value = h->pipeline;
    else if (n == "lr_blocks")
      *value = h->lr_blocks;
    else if (n == "graph")
      *value = h->use_graph ? 1 : 0;
    else if (n == "lr_atom_groups")
      *value = h->lr_atom_groups;
    else if (n == "sf_cheb")
      *value = h->sf_cheb;
    else
      throw ArgError("unknown option \"" + n + "\"");
  });
}

// Autotuning is refused for handles of more atoms than this.  One run that
// tuned the 1,000,796-charge handle (after a 3,000-charge handle in the same
// process) ended in a host segmentation fault whose cause is not found: the
// host planning the autotuner drives (splits for every target, pipeline
// chunks, node layout, the Chebyshev matrices) checks clean under
// AddressSanitizer and UBSan at that shape (tests/host/plan_check.cpp), and
// the default schedule runs that handle correctly.  Mixed (CT = float)
// handles are not tuned either: one run that tuned a 3,000-charge mixed
// handle, after two runs of the same test that passed, ended in the same
// kind of host segmentation fault, cause not found.  Their schedules are set
// with glst_set_option.
constexpr uint32_t kAutotuneMaxAtoms = 300000;

glst_status_t glst_autotune(glst_handle_t h, const double *d_rx,
                            const double *d_ry, const double *d_rz,
                            const double *d_qc, uint32_t reps,
                            double *best_ms) {
  NEED_HANDLE(h);
  return guarded(h, [&] {
    if (h->natom > kAutotuneMaxAtoms)
      throw StateError("glst_autotune: handles of more than " +
                       std::to_string(kAutotuneMaxAtoms) +
                       " atoms are not tuned (the default schedule applies)");
    if (h->mixed())
      throw StateError("glst_autotune: mixed (CT = float) handles are not "
                       "tuned; set their schedule with glst_set_option");
    const double t = h->autotune(d_rx, d_ry, d_rz, d_qc, reps);
    if (best_ms) *best_ms = t;
  });
}

glst_status_t glst_set_timing(glst_handle_t h, int enable) {
  NEED_HANDLE(h);
  h->timer.enabled = enable != 0;
  return GLST_OK;
}

glst_status_t glst_kernel_time(glst_handle_t h, const char *name,
                               double *total_ms, uint32_t *launches) {
  NEED_HANDLE(h);
  if (!name) return GLST_ERR_INVALID;
  return guarded(h, [&] {
    double ms;
    uint32_t n;
    h->timer.get(name, ms, n);
    if (total_ms) *total_ms = ms;
    if (launches) *launches = n;
  });
}

glst_status_t glst_reset_timing(glst_handle_t h) {
  NEED_HANDLE(h);
  return guarded(h, [&] { h->timer.reset(); });
}

const char *glst_last_error(glst_handle_t h) {
  if (!h) return glst_global_last_error();
  return h->last_error.c_str();
}

const char *glst_global_last_error(void) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return g_last_error.c_str();
}

const char *glst_version(void) { return "glst-mi355x 0.1.0 (gfx950)"; }

glst_status_t glst_coulomb_direct(const double *d_rx, const double *d_ry,
                                  const double *d_rz, const double *d_qc,
                                  uint32_t natom, double *d_fx, double *d_fy,
                                  double *d_fz, double *d_en, void *stream) {
  return guarded(nullptr, [&] {
    if (natom > 0 && (!d_rx || !d_ry || !d_rz || !d_qc || !d_fx || !d_fy ||
                      !d_fz || !d_en))
      throw ArgError("glst_coulomb_direct: null pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // partial sums of the source splits in a buffer of this call
    const dev::CoulombPlan plan = dev::coulomb_plan(natom, SIZE_MAX);
    DeviceBuffer part;
    part.allocate(plan.entries * sizeof(double4));
    dev::coulomb_direct(d_rx, d_ry, d_rz, d_qc, natom, plan,
                        part.as<double4>(), d_fx, d_fy, d_fz, d_en, s);
    check_launch("coulomb_direct");
    GLST_HIP(hipStreamSynchronize(s));
  });
}

glst_status_t glst_coulomb_direct_noalloc(const double *d_rx,
                                          const double *d_ry,
                                          const double *d_rz,
                                          const double *d_qc, uint32_t natom,
                                          double *d_fx, double *d_fy,
                                          double *d_fz, double *d_en,
                                          void *stream) {
  return guarded(nullptr, [&] {
    if (natom > 0 && (!d_rx || !d_ry || !d_rz || !d_qc || !d_fx || !d_fy ||
                      !d_fz || !d_en))
      throw ArgError("glst_coulomb_direct_noalloc: null pointer");
    if (natom == 0) return;
    // the same split plan as glst_coulomb_direct, capped by the fixed array
    const dev::CoulombPlan plan =
        dev::coulomb_plan(natom, dev::kCoulScratchEntries);
    if (plan.entries > dev::kCoulScratchEntries)
      throw ArgError("glst_coulomb_direct_noalloc: " +
                     std::to_string(plan.entries) +
                     " partial sums exceed the fixed array of " +
                     std::to_string(dev::kCoulScratchEntries));
    double4 *part = nullptr;
    if (plan.entries > 0) {
      part = dev::coulomb_scratch();
      if (!part)
        throw HipError("glst_coulomb_direct_noalloc: device address of the "
                       "partial-sum array not found");
    }
    dev::coulomb_direct(d_rx, d_ry, d_rz, d_qc, natom, plan, part, d_fx, d_fy,
                        d_fz, d_en, static_cast<hipStream_t>(stream));
    check_launch("coulomb_direct_noalloc");
  });
}

glst_status_t glst_erfc_inv(double y, double tol, double *x) {
  if (!x) return GLST_ERR_INVALID;
  return guarded(nullptr, [&] { *x = erfc_inv(y, tol); });
}

glst_status_t glst_alpha_groups(uint32_t nx, uint32_t ny, uint32_t nz,
                                double cdx, double cdy, double cdz, double tol,
                                uint32_t cap, uint32_t *ngroup,
                                uint32_t *width, double *rmax, double *alpha,
                                double *zcut, uint32_t *r_in,
                                uint32_t *r_out) {
  if (!ngroup) return GLST_ERR_INVALID;
  return guarded(nullptr, [&] {
    AlphaGroups g = alpha_groups(nx, ny, nz, cdx, cdy, cdz, tol);
    *ngroup = static_cast<uint32_t>(g.width.size());
    const size_t n = std::min<size_t>(cap, g.width.size());
    for (size_t i = 0; i < n; ++i) {
      if (width) width[i] = g.width[i];
      if (rmax) rmax[i] = g.rmax[i];
      if (alpha) alpha[i] = g.alpha[i];
      if (zcut) zcut[i] = g.zcut[i];
      if (r_in) r_in[i] = g.r_in[i];
      if (r_out) r_out[i] = g.r_out[i];
    }
  });
}

glst_status_t glst_gauss_legendre(uint32_t n, double *x, double *w,
                                  uint32_t *count) {
  return guarded(nullptr, [&] {
    if (n == 0) throw ArgError("glst_gauss_legendre: n must be positive");
    std::vector<double> xx, ww;
    gauss_legendre_nonneg(n, xx, ww);
    if (count) *count = static_cast<uint32_t>(xx.size());
    if (x) std::memcpy(x, xx.data(), xx.size() * 8);
    if (w) std::memcpy(w, ww.data(), ww.size() * 8);
  });
}

glst_status_t glst_select_gl_order(double tol, double dist, double alpha,
                                   double zcut, uint32_t *order) {
  if (!order) return GLST_ERR_INVALID;
  return guarded(nullptr, [&] {
    *order = select_gauss_legendre(tol, dist, alpha, zcut).order;
  });
}

glst_status_t glst_spherical_rule(uint32_t m, double *xyz, double *w,
                                  uint32_t *count) {
  return guarded(nullptr, [&] {
    SphRule r = product_rule(m);
    if (count) *count = static_cast<uint32_t>(r.s.size());
    for (size_t i = 0; i < r.s.size(); ++i) {
      if (xyz) {
        xyz[3 * i] = r.s[i][0];
        xyz[3 * i + 1] = r.s[i][1];
        xyz[3 * i + 2] = r.s[i][2];
      }
      if (w) w[i] = r.w[i];
    }
  });
}

glst_status_t glst_read_tdesign(const char *dir, uint32_t t, double *xyz,
                                uint32_t *count) {
  if (!dir || !count) return GLST_ERR_INVALID;
  return guarded(nullptr, [&] {
    SphRule r = design_file(t, dir);
    const uint32_t n = static_cast<uint32_t>(r.s.size());
    if (xyz && *count < n)
      throw ArgError("glst_read_tdesign: output holds " +
                     std::to_string(*count) + " points, the design " +
                     std::to_string(n));
    *count = n;
    if (xyz)
      for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) xyz[3 * i + a] = r.s[i][a];
  });
}

glst_status_t glst_select_spherical_degree(uint32_t gl_order, double tol,
                                           double dist, double alpha,
                                           double zcut, uint32_t *m,
                                           uint32_t *count) {
  if (!m) return GLST_ERR_INVALID;
  return guarded(nullptr, [&] {
    if (gl_order < 2 || gl_order % 2)
      throw ArgError("glst_select_spherical_degree: gl_order must be even "
                     "and at least 2");
    GLRule gl;
    gl.order = gl_order;
    gauss_legendre_nonneg(gl_order, gl.x, gl.w);
    SphRule r = select_spherical(gl, tol, dist, alpha, zcut, tdesign_dir());
    *m = r.m;
    if (count) *count = static_cast<uint32_t>(r.s.size());
  });
}

glst_status_t glst_set_tdesign_dir(const char *dir) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_tdesign_dir = dir ? dir : "";
  g_tdesign_init = true;
  return GLST_OK;
}

glst_status_t glst_grid_dims(int ct, double box_x, double box_y, double box_z,
                             double rcut, uint32_t *nx, uint32_t *ny,
                             uint32_t *nz) {
  if (!nx || !ny || !nz) return GLST_ERR_INVALID;
  return guarded(nullptr, [&] {
    if (!(box_x > 0 && box_y > 0 && box_z > 0 && rcut > 0))
      throw ArgError("box dimensions and rcut must be positive");
    if (ct == GLST_CT_FLOAT)
      grid_dims<float>(box_x, box_y, box_z, rcut, *nx, *ny, *nz);
    else
      grid_dims<double>(box_x, box_y, box_z, rcut, *nx, *ny, *nz);
  });
}

}  // extern "C"
