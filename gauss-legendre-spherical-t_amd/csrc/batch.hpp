// The batched multi-system engine (include/glst_batch.h), shared by its C
// ABI (batch.hip) and the serving loop (serve.hip).
//
// Every system is set up exactly as a single glst_force (runtime.hpp: grid,
// alpha groups, cubature, device node order) and cached as a BatchSystem; a
// batch packs any list of such systems into shared arenas (atoms, cells,
// splits, S/R fields, histogram, partial rows), refers to the node and task
// tables of their grid shapes in device-resident ShapeTables, and runs each
// stage of the force step as one launch over all of them (S(k) and the
// back-projection: one launch per radial node count).  A workgroup finds its
// system through a run table of (first workgroup, system, group data)
// records.  The arenas only grow, so re-packing a batch for a new list of
// systems allocates nothing once they are large enough, and copies no table
// of a shape seen before.
#pragma once

#include "runtime.hpp"

#include <map>
#include <memory>

namespace glst {

// The host setup of a cell grid: it depends only on the grid dimensions,
// the cell edge, CT and tol (not on the atoms or the exact box), so systems
// of one grid share it.
struct ShapeSetup {
  SystemSetup su;
  NodeLayout nl;  // device node/task order (empty without a far field)
  // CT = float: Chebyshev form of S(k) per alpha group (plan_chebyshev; m 0
  // keeps the direct sum), the matrices at cheb_coff of `cheb`
  std::vector<uint32_t> cheb_m, cheb_coff;
  std::vector<float> cheb;
};

std::shared_ptr<const ShapeSetup> make_shape_setup(int ct, const double box[3],
                                                   double rcut, double tol);

// One system of a batch: its atom count and grid setup.
struct BatchSystem {
  uint32_t natom = 0;
  std::shared_ptr<const ShapeSetup> shape;
  const SystemSetup &su() const { return shape->su; }
  const NodeLayout &nl() const { return shape->nl; }
};

std::shared_ptr<const BatchSystem> make_batch_system(
    uint32_t natom, std::shared_ptr<const ShapeSetup> shape);

// Node, task, kappa and shell-radius tables of grid shapes, resident on the
// device.  A shape's tables are appended (host) the first time a batch packs
// a system of that shape and uploaded once; every system of the shape refers
// to them by offset, so packing a batch copies no table.  A server's batches
// share one set.  Device capacity grows with headroom; a growth waits for
// the device, reallocates and uploads everything again, and clear() starts
// over (both bump `gen`: a batch configured under an older generation must
// be configured again, which also drops its captured graph).
struct ShapeTables {
  struct Offsets {
    uint32_t node0 = 0, task0 = 0, kap0 = 0, group0 = 0, cheb0 = 0;
  };
  std::vector<double4> k64;  // per node (kx, ky, kz, W)
  std::vector<uint32_t> group;
  std::vector<uint32_t> r_in, r_out;  // per alpha group
  std::vector<uint32_t> t_node0, t_nr, t_koff;  // per task; koff global
  std::vector<double4> t_dir64;
  std::vector<float4> t_dir32;
  std::vector<double> kap64;
  std::vector<float> kap32;
  // Chebyshev matrices of all shapes, and per task the offset of its
  // group's matrix
  std::vector<float> cheb;
  std::vector<uint32_t> t_coff;
  DeviceBuffer d_k64, d_group, d_r_in, d_r_out, d_node0, d_nr, d_koff,
      d_dir64, d_dir32, d_kap64, d_kap32, d_cheb, d_coff;
  std::map<const ShapeSetup *,
           std::pair<std::shared_ptr<const ShapeSetup>, Offsets>>
      reg;
  size_t on_dev = 0;  // shapes (in order of addition) already uploaded
  std::vector<const ShapeSetup *> order;
  uint64_t gen = 1;

  // offsets of the shape's tables, appending them (host) if new
  const Offsets &add(const std::shared_ptr<const ShapeSetup> &shape);
  // uploads the shapes added since the last call (on stream s; returns when
  // the host copies may change again)
  void upload(hipStream_t s);
  void clear();
  size_t host_bytes() const;
  size_t device_bytes() const;
  // table bytes of one shape
  static size_t shape_bytes(const ShapeSetup &shape);
};

// Shapes a table set holds before a configure starts it over.
constexpr size_t kMaxShapeTables = 1024;

// A launch of a batched stage under construction: run table and workgroups.
struct Plan {
  std::vector<dev::Seg> seg;
  uint32_t nblk = 0;
  uint32_t nr = 0;
  void add(uint32_t sys, uint32_t blocks, uint32_t a = 0, uint32_t b = 0) {
    if (blocks == 0) return;
    seg.push_back({nblk, sys, a, b});
    nblk += blocks;
  }
};

// Host layout of a whole batch: SysDev records (offsets into the shape
// tables and the batch's arenas), arena sizes, the split table and the run
// tables of every launch.
struct BatchLayout {
  int ct = GLST_CT_DOUBLE;
  std::vector<std::shared_ptr<const BatchSystem>> cs;
  std::vector<dev::SysDev> sys;
  uint64_t natom = 0, ncell = 0, hist = 0, field = 0, partial = 0;
  uint64_t nag = 0;  // mixed: atom-group entries of the back-projection
  size_t lds = 0;          // dynamic LDS of the shell-sum launch
  size_t table_bytes = 0;  // shape tables of the batch's distinct shapes
  std::vector<uint32_t> t_split;
  Plan bin, hist_p, scatter, sr, reduce, rmt, sf_f64, lr_f64;
  // the shell sums split by kernel (per-stage launches): the mixed systems
  // whose grids fit dev::rmt_wave (rmt_small, when asked for) and the rest
  // (rmt_rest); rmt holds every system (the fused step's phase)
  Plan rmt_small, rmt_rest;
  std::map<uint32_t, Plan> sf_nr, lr_nr;  // radial count (0: any) -> plan
  // mixed: the S(k) and back-projection units of the fused step (all radial
  // counts, largest first; see dev::FusedStep)
  Plan fsf, flr;

  bool mixed() const { return ct == GLST_CT_FLOAT; }
  size_t device_bytes() const;  // arena and table bytes this layout needs
};

// Packs systems (all of precision ct) into one batch layout, adding their
// shapes to `tables` (host side).  lds_limit: the device's LDS per workgroup
// (a system whose one-node cell table does not fit is rejected); lr_blocks:
// back-projection workgroups aimed for over the whole batch; cheb: CT =
// float S(k) in Chebyshev form where the shapes planned it; wave: CT = float
// shell sums of the grids up to 8 cells a side as window sums (rmt_small).
// back-projection workgroups a batch aims for unless set (option lr_blocks)
#ifndef GLST_BATCH_LR_BLOCKS
#define GLST_BATCH_LR_BLOCKS 4096
#endif

BatchLayout pack_batch(int ct,
                       const std::vector<std::shared_ptr<const BatchSystem>> &cs,
                       ShapeTables &tables, size_t lds_limit,
                       uint32_t lr_blocks = 4096, bool cheb = true,
                       bool wave = true);

}  // namespace glst

struct glst_batch {
  std::string last_error;
  glst::BatchLayout B;
  int device = 0;
  size_t lds_limit = 65536;
  hipStream_t stream = nullptr, side = nullptr;
  bool own_stream = false;
  hipEvent_t ev_assigned = nullptr, ev_near = nullptr;
  int overlap_near = 1;
  uint32_t lr_blocks = GLST_BATCH_LR_BLOCKS;
  // mixed S(k) in Chebyshev form where the shapes planned it (option
  // sf_cheb; the persistent and one-launch-per-stage modes sum directly)
  int sf_cheb = 1;
  // mixed shell sums of grids up to 8 cells a side as window sums (option
  // rmt_wave; the persistent fused step keeps the LDS tables)
  int rmt_wave = 1;
  bool assigned = false;
  // HIP graph of the step (graph_mode 1: captured on the first step with
  // given inputs; 2: on the second consecutive step with the same inputs,
  // so batches used once are never captured)
  int graph_mode = 0;
  hipGraphExec_t graph_exec = nullptr;
  const double *graph_in[4] = {nullptr, nullptr, nullptr, nullptr};
  const double *last_in[4] = {nullptr, nullptr, nullptr, nullptr};
  // mixed path: 0 runs every stage as its own launch(es), S(k) and the
  // back-projection one per radial count; 3 one launch per stage; 1 and 2
  // the whole step as one persistent launch (dev::step_fused variants).
  // fused_share: the launch takes 1/fused_share of the device's resident
  // workgroups (a server's two slots may run theirs at once).
  int fused = 0;
  unsigned fused_share = 1;
  uint32_t fused_nwg = 0;   // workgroups of the configured batch's launch
  size_t fused_lds = 0;     // its dynamic LDS
  long long fused_timeout = 0;
  glst::DeviceBuffer fsync;  // barrier counter and failure word
  uint32_t *fail_host = nullptr, *fail_dev = nullptr;  // mapped host flag
  // per cent added to an arena when it grows (0: exact sizes)
  unsigned arena_headroom = 25;
  // shape tables (shared with the other batches of a server) and the
  // generation they were packed under
  std::shared_ptr<glst::ShapeTables> tables;
  uint64_t tables_gen = 0;
  // arenas (grow-only)
  glst::DeviceBuffer d_sys, d_seg;
  glst::DeviceBuffer atom_cell, rank, hist, cell_count, cell_start, maxc;
  glst::DeviceBuffer sorted_idx, cell_sorted, atoms64, atoms32, near4;
  glst::DeviceBuffer atom_grp, ag_start;  // mixed back-projection groups
  glst::DeviceBuffer fx, fy, fz, en;
  glst::DeviceBuffer t_split;
  glst::DeviceBuffer S, R, partial;
  std::vector<glst::dev::Seg> host_seg;  // run tables (kept for the copy)
  // launches (run tables inside d_seg)
  glst::dev::BatchLaunch L_bin{}, L_hist{}, L_scatter{}, L_sr{}, L_reduce{},
      L_rmt{}, L_fsf{}, L_flr{}, L_rmt_small{}, L_rmt_rest{};
  std::vector<glst::dev::BatchLaunch> L_sf, L_lr;

  bool mixed() const { return B.mixed(); }
  uint32_t nsys() const { return static_cast<uint32_t>(B.sys.size()); }
  // packed under shape tables that have since been reallocated or cleared
  bool stale() const { return tables_gen != tables->gen; }

  ~glst_batch();
  // device, streams and events (once)
  void open(void *stream_);
  // (re)pack for these systems: uploads tables and run tables, grows arenas,
  // clears S and the accumulators
  void configure(int ct,
                 std::vector<std::shared_ptr<const glst::BatchSystem>> cs);
  void zero_ef();
  // frees every arena (the next configure allocates them for its batch)
  void release_arenas();
  void step_async(const double *rx, const double *ry, const double *rz,
                  const double *qc);
  // step_async, or a replay of the step captured for these inputs
  void step(const double *rx, const double *ry, const double *rz,
            const double *qc);
  void drop_graph();
  // fastest of the candidate schedules (overlap_near, lr_blocks) for these
  // inputs, `reps` steps each; returns its ms per step
  double autotune(const double *rx, const double *ry, const double *rz,
                  const double *qc, uint32_t reps);
  // unpermuted outputs in packed original order (stream-ordered, no sync)
  void get_ef_async(double *ofx, double *ofy, double *ofz, double *oen);
  // after the stream's work is complete: throws if a fused step failed
  // (a barrier timed out; its outputs are not valid); the flag stays set
  // until the batch is configured again
  void check_failed() const;
  // the next step runs as one persistent launch
  bool use_fused() const {
    return (fused == 1 || fused == 2) && mixed() && fused_nwg > 0;
  }
  // workgroups and LDS of the fused launch for the configured batch (after
  // configure and whenever fused or fused_share change)
  void plan_fused();
  size_t device_bytes() const;

 private:
  glst::dev::LrTaskArgs task_args() const;
  void near_field(hipStream_t s);
  void step_fused(const double *rx, const double *ry, const double *rz,
                  const double *qc);
};
