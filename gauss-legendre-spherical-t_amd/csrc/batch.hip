// The batched multi-system engine behind include/glst_batch.h (see
// batch.hpp for the design).
#include "glst_batch.h"
#include "batch.hpp"

#include <cstring>
#include <new>

using namespace glst;

namespace {

uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

namespace glst {

std::shared_ptr<const ShapeSetup> make_shape_setup(int ct, const double box[3],
                                                   double rcut, double tol) {
  auto p = std::make_shared<ShapeSetup>();
  p->su = setup_system(ct, box[0], box[1], box[2], rcut, tol);
  if (p->su.has_lr && p->su.K > 0) {
    p->nl = layout_nodes(p->su.nodes, p->su.K, p->su.Kp);
    if (ct == GLST_CT_FLOAT) {
      std::vector<dev::SfGroup> groups =
          make_splits(p->su.nodes, p->su.K, 1).sf_groups;
      p->cheb = plan_chebyshev(p->su.nodes, p->su.grid.h, groups);
      for (const dev::SfGroup &g : groups) {
        p->cheb_m.push_back(g.m);
        p->cheb_coff.push_back(g.coff);
      }
    }
  }
  return p;
}

std::shared_ptr<const BatchSystem> make_batch_system(
    uint32_t natom, std::shared_ptr<const ShapeSetup> shape) {
  if (natom == 0) throw ArgError("a system without atoms");
  auto b = std::make_shared<BatchSystem>();
  b->natom = natom;
  b->shape = std::move(shape);
  return b;
}

size_t BatchLayout::device_bytes() const {
  const size_t n = natom, c = ncell;
  size_t b = 0;
  b += n * 4 * 4;                         // atom_cell, rank, sorted, cell
  b += n * sizeof(double4) * 2;           // atoms64, near4
  b += mixed() ? n * sizeof(float4) : 0;  // atoms32
  b += nag * sizeof(uint2) + (nag ? c * 4 : 0);  // atom groups, ag_start
  b += n * 8 * 4;                         // fx..en
  b += hist * 4 + (c + 1) * 4 + c * 4 + sys.size() * 4;
  b += table_bytes + t_split.size() * 4;
  b += field * (mixed() ? 8 : 16) * 2;  // S and R
  b += partial * sizeof(double4);
  b += sys.size() * sizeof(dev::SysDev);
  size_t nseg = bin.seg.size() + hist_p.seg.size() + scatter.seg.size() +
                sr.seg.size() + reduce.seg.size() + rmt.seg.size() +
                rmt_small.seg.size() + rmt_rest.seg.size() +
                sf_f64.seg.size() + lr_f64.seg.size();
  for (auto &kv : sf_nr) nseg += kv.second.seg.size();
  for (auto &kv : lr_nr) nseg += kv.second.seg.size();
  b += nseg * sizeof(dev::Seg);
  return b;
}

size_t ShapeTables::shape_bytes(const ShapeSetup &shape) {
  const NodeLayout &L = shape.nl;
  return L.k64.size() * (sizeof(double4) + 4) +
         L.node0.size() * (16 + sizeof(double4) + sizeof(float4)) +
         L.kappa64.size() * 12 + shape.su.groups.r_in.size() * 8 +
         shape.cheb.size() * 4;
}

const ShapeTables::Offsets &ShapeTables::add(
    const std::shared_ptr<const ShapeSetup> &shape) {
  auto it = reg.find(shape.get());
  if (it != reg.end()) return it->second.second;
  const NodeLayout &L = shape->nl;
  Offsets o;
  o.node0 = static_cast<uint32_t>(k64.size());
  o.task0 = static_cast<uint32_t>(t_node0.size());
  o.kap0 = static_cast<uint32_t>(kap64.size());
  o.group0 = static_cast<uint32_t>(r_in.size());
  o.cheb0 = static_cast<uint32_t>(cheb.size());
  if (k64.size() + L.k64.size() > 0xFFFFFFFFull ||
      t_node0.size() + L.node0.size() > 0xFFFFFFFFull)
    throw ArgError("shape tables exceed 2^32 entries");
  k64.insert(k64.end(), L.k64.begin(), L.k64.end());
  group.insert(group.end(), L.group.begin(), L.group.end());
  for (size_t t = 0; t < L.node0.size(); ++t) {
    t_node0.push_back(L.node0[t]);  // local to the system
    t_nr.push_back(L.nr[t]);
    t_koff.push_back(o.kap0 + L.koff[t]);
    t_dir64.push_back(L.dir64[t]);
    t_dir32.push_back(L.dir32[t]);
  }
  // the Chebyshev matrix of each task's group (one task per direction,
  // group-major)
  {
    const auto &dirs = shape->su.nodes.dirs;
    size_t t = 0;
    for (size_t g = 0; g < dirs.size() && t < L.node0.size(); ++g)
      for (size_t j = 0; j < dirs[g].size() && t < L.node0.size(); ++j, ++t)
        t_coff.push_back(o.cheb0 + (g < shape->cheb_coff.size()
                                        ? shape->cheb_coff[g]
                                        : 0u));
    for (; t < L.node0.size(); ++t) t_coff.push_back(o.cheb0);
  }
  cheb.insert(cheb.end(), shape->cheb.begin(), shape->cheb.end());
  kap64.insert(kap64.end(), L.kappa64.begin(), L.kappa64.end());
  kap32.insert(kap32.end(), L.kappa32.begin(), L.kappa32.end());
  r_in.insert(r_in.end(), shape->su.groups.r_in.begin(),
              shape->su.groups.r_in.end());
  r_out.insert(r_out.end(), shape->su.groups.r_out.begin(),
               shape->su.groups.r_out.end());
  order.push_back(shape.get());
  return reg.emplace(shape.get(), std::make_pair(shape, o)).first->second.second;
}

void ShapeTables::upload(hipStream_t s) {
  if (on_dev == order.size()) return;
  struct Arr {
    DeviceBuffer *d;
    const void *h;
    size_t n, size;
  };
  const Arr arrs[] = {
      {&d_k64, k64.data(), k64.size(), sizeof(double4)},
      {&d_group, group.data(), group.size(), 4},
      {&d_r_in, r_in.data(), r_in.size(), 4},
      {&d_r_out, r_out.data(), r_out.size(), 4},
      {&d_node0, t_node0.data(), t_node0.size(), 4},
      {&d_nr, t_nr.data(), t_nr.size(), 4},
      {&d_koff, t_koff.data(), t_koff.size(), 4},
      {&d_dir64, t_dir64.data(), t_dir64.size(), sizeof(double4)},
      {&d_dir32, t_dir32.data(), t_dir32.size(), sizeof(float4)},
      {&d_kap64, kap64.data(), kap64.size(), 8},
      {&d_kap32, kap32.data(), kap32.size(), 4},
      {&d_coff, t_coff.data(), t_coff.size(), 4},
      {&d_cheb, cheb.data(), cheb.size(), 4}};
  constexpr size_t kArr = sizeof(arrs) / sizeof(arrs[0]);
  // entries of the shapes already on the device
  size_t done[kArr] = {};
  if (on_dev > 0) {
    const Offsets &o = reg.at(order[on_dev]).second;
    const size_t groups = o.group0;
    const size_t tasks = o.task0, nodes = o.node0, kaps = o.kap0;
    const size_t v[kArr] = {nodes, nodes, groups, groups, tasks, tasks, tasks,
                            tasks, tasks, kaps,   kaps,  tasks, o.cheb0};
    std::copy(v, v + kArr, done);
  }
  bool grow = false;
  for (const Arr &a : arrs) grow = grow || a.n * a.size > a.d->bytes();
  if (grow) {
    // batches in flight read the old copies: wait for them, then
    // reallocate with headroom and upload every shape again
    GLST_HIP(hipDeviceSynchronize());
    for (const Arr &a : arrs)
      if (a.n * a.size > a.d->bytes()) {
        const size_t want = std::max(a.n * a.size, 2 * a.d->bytes());
        a.d->release();
        a.d->allocate(want);
      }
    std::fill(done, done + kArr, size_t{0});
    ++gen;
  }
  for (size_t i = 0; i < kArr; ++i) {
    const Arr &a = arrs[i];
    if (a.n > done[i])
      GLST_HIP(hipMemcpyAsync(a.d->as<char>() + done[i] * a.size,
                              static_cast<const char *>(a.h) + done[i] * a.size,
                              (a.n - done[i]) * a.size, hipMemcpyHostToDevice,
                              s));
  }
  on_dev = order.size();
  // the host vectors may grow (and move) with the next shape
  GLST_HIP(hipStreamSynchronize(s));
}

void ShapeTables::clear() {
  for (auto *v : {&group, &r_in, &r_out, &t_node0, &t_nr, &t_koff, &t_coff})
    v->clear();
  cheb.clear();
  k64.clear();
  t_dir64.clear();
  t_dir32.clear();
  kap64.clear();
  kap32.clear();
  reg.clear();
  order.clear();
  on_dev = 0;
  ++gen;
}

size_t ShapeTables::host_bytes() const {
  return k64.size() * (sizeof(double4) + 4) +
         t_node0.size() * (16 + sizeof(double4) + sizeof(float4)) +
         kap64.size() * 12 + r_in.size() * 8 + cheb.size() * 4;
}

size_t ShapeTables::device_bytes() const {
  size_t b = 0;
  for (const DeviceBuffer *d : {&d_k64, &d_group, &d_r_in, &d_r_out, &d_node0,
                                &d_nr, &d_koff, &d_dir64, &d_dir32, &d_kap64,
                                &d_kap32, &d_coff, &d_cheb})
    b += d->bytes();
  return b;
}

// 4 with 256-thread workgroups: the 512-system mix ran at 32.5 ms against
// 34.9 ms for one 1024-thread workgroup per CU (2: 36.4, 8: 34.0)
#ifndef GLST_BATCH_RMT_LDS_DIV
#define GLST_BATCH_RMT_LDS_DIV 4
#endif
constexpr size_t kBatchRmtLdsDiv = GLST_BATCH_RMT_LDS_DIV;

BatchLayout pack_batch(int ct,
                       const std::vector<std::shared_ptr<const BatchSystem>> &cs,
                       ShapeTables &tables, size_t lds_limit,
                       uint32_t lr_blocks, bool cheb, bool wave) {
  if (cs.empty()) throw ArgError("a batch needs at least one system");
  const uint32_t nsys = static_cast<uint32_t>(cs.size());
  BatchLayout B;
  B.ct = ct;
  B.cs = cs;
  // back-projection splits: enough workgroups over the whole batch
  uint64_t tiles = 0;
  for (const auto &c : cs)
    if (c->su().has_lr && c->su().K > 0) tiles += div_up(c->natom, 256);
  const uint32_t target =
      tiles ? static_cast<uint32_t>(
                  std::max<uint64_t>(1, (lr_blocks + tiles - 1) / tiles))
            : 1;
  std::vector<SplitPlan> splits(nsys);
  std::vector<uint32_t> task_first(nsys, 0);
  uint32_t split_base = 0;
  std::map<const ShapeSetup *, bool> seen;
  for (uint32_t s = 0; s < nsys; ++s) {
    const SystemSetup &u = cs[s]->su();
    const uint32_t na = cs[s]->natom;
    const bool lr = u.has_lr && u.K > 0;
    const ShapeTables::Offsets off = tables.add(cs[s]->shape);
    if (!seen[cs[s]->shape.get()]) {
      seen[cs[s]->shape.get()] = true;
      B.table_bytes += ShapeTables::shape_bytes(*cs[s]->shape);
    }
    task_first[s] = off.task0;
    dev::SysDev d{};
    d.natom = na;
    d.atom0 = static_cast<uint32_t>(B.natom);
    d.nx = u.grid.nx;
    d.ny = u.grid.ny;
    d.nz = u.grid.nz;
    d.ncell = u.grid.ncell;
    d.cell0 = static_cast<uint32_t>(B.ncell);
    d.nchunk = div_up(na, dev::kBinChunk);
    d.hist0 = B.hist;
    d.h = u.grid.h;
    d.inv_cell = u.inv_cell;
    d.far_d = far_offset(u.nodes);
    d.ntile = div_up(na, 256);
    d.group0 = off.group0;
    d.Kp = lr ? u.Kp : 0;
    d.node0 = off.node0;
    d.field0 = B.field;
    d.T = 1;
    d.split0 = split_base;
    d.partial0 = B.partial;
    d.ag0 = static_cast<uint32_t>(B.nag);
    d.agtile = 0;
    B.natom += na;
    B.ncell += u.grid.ncell;
    B.hist += static_cast<uint64_t>(d.nchunk) * u.grid.ncell;
    if (B.natom > 0xFFFFFFFFull || B.ncell > 0x7FFFFFFFull)
      throw ArgError("batch exceeds 2^32 atoms or 2^31 cells");
    if (lr) {
      // One launch covers every system with one dynamic LDS size, the
      // largest tile's: tiles sized for 1/kBatchRmtLdsDiv of the LDS let
      // that many workgroups share a CU (a grid whose one-node table needs
      // more takes the whole LDS).
      d.T = dev::rmt_nodes_per_group(u.grid.nx, u.grid.ny, u.grid.nz,
                                     lds_limit / kBatchRmtLdsDiv,
                                     dev::kRmtMaxNodes);
      if (d.T == 0)
        d.T = dev::rmt_nodes_per_group(u.grid.nx, u.grid.ny, u.grid.nz,
                                       lds_limit, dev::kRmtMaxNodes);
      if (d.T == 0)
        throw ArgError("system " + std::to_string(s) + ": its " +
                       std::to_string(u.grid.ncell) +
                       "-cell grid is too large for the batched engine; "
                       "use a single-system handle");
      B.lds = std::max(B.lds, dev::rmt_lds_bytes(u.grid.nx, u.grid.ny,
                                                 u.grid.nz, d.T));
      SplitPlan &P = splits[s];
      P = make_splits(u.nodes, u.K, target);
      d.nsplit = P.nsplit;
      // split boundaries as task indices of the shape tables; a split ends
      // where the next one begins, so every system's splits are followed by
      // the end of its task range
      for (uint32_t q = 0; q <= P.nsplit; ++q)
        B.t_split.push_back(off.task0 + P.split[q]);
      B.field += static_cast<uint64_t>(u.grid.ncell) * u.Kp;
      B.partial += static_cast<uint64_t>(P.nsplit) * na;
      split_base += P.nsplit + 1;
      if (ct == GLST_CT_FLOAT) {  // back-projection over atom groups
        const uint32_t nag = dev::atom_groups_max(na, u.grid.ncell);
        d.agtile = nag / dev::kAgGroupsPerBlock;
        B.nag += nag;
        if (B.nag > 0xFFFFFFFFull)
          throw ArgError("batch exceeds 2^32 atom groups");
      }
    }
    B.sys.push_back(d);
  }
  if (B.t_split.empty()) B.t_split.push_back(0);
  // units of the fused step's S(k) and back-projection phases, largest
  // radial count first
  struct FusedRun {
    uint32_t nr, sys, blocks, a, b;
  };
  std::vector<FusedRun> fsf, flr;
  // run tables
  for (uint32_t s = 0; s < nsys; ++s) {
    const dev::SysDev &d = B.sys[s];
    B.bin.add(s, d.nchunk);
    B.hist_p.add(s, div_up(d.ncell, 256));
    B.scatter.add(s, div_up(d.natom, 256));
    B.reduce.add(s, div_up(d.natom, 256));
    B.sr.add(s, d.ncell);
    if (d.Kp == 0) continue;
    B.rmt.add(s, d.Kp / d.T);
    const dev::GridShape gs{d.nx, d.ny, d.nz, d.ncell, d.h};
    if (wave && ct == GLST_CT_FLOAT && dev::rmt_wave_fits(gs))
      B.rmt_small.add(s, d.Kp / dev::kRmtWaveNodes);
    else
      B.rmt_rest.add(s, d.Kp / d.T);
    B.sf_f64.add(s, d.nx * d.ny * (d.Kp / 256));
    B.lr_f64.add(s, d.ntile * d.nsplit);
    const SplitPlan &P = splits[s];
    const uint32_t task0 = task_first[s];  // first task of the system
    for (size_t g = 0; g < P.sf_groups.size(); ++g) {
      const dev::SfGroup &sg = P.sf_groups[g];
      const dev::LrGroup &lg = P.lr_groups[g];
      if (sg.ntask == 0) continue;
      const uint32_t key = sg.nr <= 40 ? sg.nr : 0;
      // S(k) launches: direct sums per radial count, Chebyshev forms per
      // term count (keys kSfChebKey + M)
      const std::vector<uint32_t> &cm = cs[s]->shape->cheb_m;
      const uint32_t m = cheb && g < cm.size() ? cm[g] : 0u;
      const uint32_t skey = m ? dev::kSfChebKey + m : key;
      Plan &pf = B.sf_nr[skey];
      pf.nr = skey;
      pf.add(s, d.ncell * div_up(sg.ntask, 64),
             task0 + sg.task0, sg.ntask);
      Plan &pl = B.lr_nr[key];
      pl.nr = key;
      // atom-group tiles for the unrolled radial counts (see lr_batch)
      const uint32_t tiles = key != 0 && d.agtile ? d.agtile : d.ntile;
      pl.add(s, tiles * lg.nsplit, d.split0 + lg.split0, lg.nsplit);
      if (ct == GLST_CT_FLOAT) {
        fsf.push_back({sg.nr, s,
                       div_up(d.ncell, dev::kFusedSfCells) *
                           div_up(sg.ntask, 64),
                       task0 + sg.task0, sg.ntask});
        const bool unrolled =
            sg.nr <= static_cast<uint32_t>(dev::kFusedMaxNr) && d.agtile;
        flr.push_back({sg.nr, s, (unrolled ? d.agtile : d.ntile) * lg.nsplit,
                       d.split0 + lg.split0, lg.nsplit});
      }
    }
  }
  auto by_nr = [](const FusedRun &x, const FusedRun &y) { return x.nr > y.nr; };
  std::stable_sort(fsf.begin(), fsf.end(), by_nr);
  std::stable_sort(flr.begin(), flr.end(), by_nr);
  for (const FusedRun &r : fsf) B.fsf.add(r.sys, r.blocks, r.a, r.b);
  for (const FusedRun &r : flr) B.flr.add(r.sys, r.blocks, r.a, r.b);
  return B;
}

}  // namespace glst

glst_batch::~glst_batch() {
  // a step still running may write the failure flag or the arenas
  if (stream) (void)hipStreamSynchronize(stream);
  drop_graph();
  if (fail_host) (void)hipHostFree(fail_host);
  if (side) (void)hipStreamDestroy(side);
  if (ev_assigned) (void)hipEventDestroy(ev_assigned);
  if (ev_near) (void)hipEventDestroy(ev_near);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

void glst_batch::open(void *stream_) {
  GLST_HIP(hipGetDevice(&device));
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                            device) == hipSuccess &&
      lds > 0)
    lds_limit = static_cast<size_t>(lds);
  dev::prepare_kernels(lds_limit);
  if (stream_) {
    stream = static_cast<hipStream_t>(stream_);
  } else {
    GLST_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    own_stream = true;
  }
  GLST_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  GLST_HIP(hipEventCreateWithFlags(&ev_assigned, hipEventDisableTiming));
  GLST_HIP(hipEventCreateWithFlags(&ev_near, hipEventDisableTiming));
  if (!tables) tables = std::make_shared<ShapeTables>();
}

namespace {

template <typename T>
void upload(DeviceBuffer &b, const std::vector<T> &v, hipStream_t s,
            unsigned headroom) {
  b.grow(v.size() * sizeof(T), headroom);
  if (!v.empty())
    GLST_HIP(hipMemcpyAsync(b.as<void>(), v.data(), v.size() * sizeof(T),
                            hipMemcpyHostToDevice, s));
}

}  // namespace

void glst_batch::configure(
    int ct, std::vector<std::shared_ptr<const glst::BatchSystem>> cs) {
  // the previous batch's work must be done before its tables change
  GLST_HIP(hipStreamSynchronize(stream));
  drop_graph();
  std::fill(last_in, last_in + 4, nullptr);
  if (tables->order.size() > kMaxShapeTables) {
    // bounded: start over (other batches sharing the tables see the new
    // generation and configure again)
    GLST_HIP(hipDeviceSynchronize());
    tables->clear();
  }
  B = pack_batch(ct, cs, *tables, lds_limit, lr_blocks, sf_cheb != 0,
                 rmt_wave != 0);
  tables->upload(stream);
  tables_gen = tables->gen;
  const size_t N = B.natom, C = B.ncell;
  const unsigned hr = arena_headroom;
  atom_cell.grow(N * 4, hr);
  rank.grow(N * 4, hr);
  sorted_idx.grow(N * 4, hr);
  cell_sorted.grow(N * 4, hr);
  atoms64.grow(N * sizeof(double4), hr);
  if (mixed()) atoms32.grow(N * sizeof(float4), hr);
  atom_grp.grow(B.nag * sizeof(uint2), hr);
  ag_start.grow(B.nag ? C * 4 : 0, hr);
  near4.grow(N * sizeof(double4), hr);
  for (DeviceBuffer *b : {&fx, &fy, &fz, &en}) b->grow(N * 8, hr);
  hist.grow(std::max<uint64_t>(B.hist, 1) * 4, hr);
  cell_count.grow(C * 4, hr);
  cell_start.grow((C + 1) * 4, hr);
  maxc.grow(nsys() * 4, hr);
  const size_t fbytes = mixed() ? sizeof(float2) : sizeof(double2);
  S.grow(B.field * fbytes, hr);
  R.grow(B.field * fbytes, hr);
  partial.grow(B.partial * sizeof(double4), hr);
  // B and host_seg stay unchanged until the next configure, which first
  // waits for this batch's stream: the copies may still be in flight
  upload(d_sys, B.sys, stream, hr);
  upload(t_split, B.t_split, stream, hr);
  // all run tables in one buffer; launches hold offsets until it exists
  std::vector<dev::Seg> &all = host_seg;
  all.clear();
  std::vector<std::pair<dev::BatchLaunch *, size_t>> where;
  auto stage = [&](const Plan &p, dev::BatchLaunch &L) {
    L.nseg = static_cast<uint32_t>(p.seg.size());
    L.nblk = p.nblk;
    L.nr = p.nr;
    where.emplace_back(&L, all.size());
    all.insert(all.end(), p.seg.begin(), p.seg.end());
  };
  L_sf.clear();
  L_lr.clear();
  // sized first so the pointers taken by `stage` stay valid
  L_sf.resize(mixed() ? B.sf_nr.size() : 1);
  L_lr.resize(mixed() ? B.lr_nr.size() : 1);
  stage(B.bin, L_bin);
  stage(B.hist_p, L_hist);
  stage(B.scatter, L_scatter);
  stage(B.sr, L_sr);
  stage(B.reduce, L_reduce);
  stage(B.rmt, L_rmt);
  stage(B.rmt_small, L_rmt_small);
  stage(B.rmt_rest, L_rmt_rest);
  stage(B.fsf, L_fsf);
  stage(B.flr, L_flr);
  if (mixed()) {
    size_t i = 0;
    for (auto &kv : B.sf_nr) stage(kv.second, L_sf[i++]);
    i = 0;
    for (auto &kv : B.lr_nr) stage(kv.second, L_lr[i++]);
  } else {
    stage(B.sf_f64, L_sf[0]);
    stage(B.lr_f64, L_lr[0]);
  }
  upload(d_seg, all, stream, hr);
  for (auto &w : where) w.first->seg = d_seg.as<dev::Seg>() + w.second;
  // the S(k) kernels never write padded node slots
  if (B.field)
    dev::fill_f64(S.as<double>(), 0.0, B.field * fbytes / sizeof(double),
                  stream);
  zero_ef();
  assigned = false;
  // a failed fused step left nothing behind that a new configuration uses
  if (fail_host) *fail_host = 0u;
  plan_fused();
  check_launch("batch configure");
}

void glst_batch::plan_fused() {
  fused_nwg = 0;
  if (!mixed() || (fused != 1 && fused != 2)) return;
  fused_lds = dev::fused_lds_bytes(B.lds);
  const uint32_t cap = dev::fused_workgroups(fused_lds, fused);
  if (cap == 0) throw StateError("the fused step does not fit on this device");
  if (!fail_host) {
    GLST_HIP(hipHostMalloc(reinterpret_cast<void **>(&fail_host), 4,
                           hipHostMallocMapped));
    *fail_host = 0u;
    GLST_HIP(hipHostGetDevicePointer(reinterpret_cast<void **>(&fail_dev),
                                     fail_host, 0));
    // a barrier gives up after 10 s (a legitimate phase takes milliseconds)
    fused_timeout = 10 * dev::wall_clock_rate();
  }
  fsync.grow(2 * sizeof(uint32_t), 0);
  // no more workgroups than the largest phase has units
  uint64_t units = std::max<uint64_t>(
      {B.bin.nblk, B.hist_p.nblk, B.sys.size(), B.scatter.nblk,
       static_cast<uint64_t>(B.fsf.nblk) + B.sr.nblk, B.rmt.nblk, B.flr.nblk,
       B.reduce.nblk, 1});
  const uint32_t share = std::max(1u, cap / std::max(1u, fused_share));
  fused_nwg = static_cast<uint32_t>(std::min<uint64_t>(units, share));
}

void glst_batch::check_failed() const {
  if (fail_host && *static_cast<volatile uint32_t *>(fail_host) != 0u)
    throw StateError(
        "a fused batched step failed: a grid barrier timed out because not "
        "all of its workgroups were resident; its outputs are not valid");
}

void glst_batch::step_fused(const double *rx, const double *ry,
                            const double *rz, const double *qc) {
  dev::FusedStep f{};
  f.rx = rx;
  f.ry = ry;
  f.rz = rz;
  f.qc = qc;
  f.sys = d_sys.as<dev::SysDev>();
  f.nsys = nsys();
  f.bin = L_bin;
  f.hist_p = L_hist;
  f.scatter = L_scatter;
  f.sf = L_fsf;
  f.sr = L_sr;
  f.rmt = L_rmt;
  f.lr = L_flr;
  f.reduce = L_reduce;
  f.atom_cell = atom_cell.as<uint32_t>();
  f.rank = rank.as<uint32_t>();
  f.hist = hist.as<uint32_t>();
  f.count = cell_count.as<uint32_t>();
  f.start = cell_start.as<uint32_t>();
  f.maxc = maxc.as<uint32_t>();
  f.sorted_idx = sorted_idx.as<uint32_t>();
  f.cell_sorted = cell_sorted.as<uint32_t>();
  f.atoms64 = atoms64.as<double4>();
  f.atoms32 = atoms32.as<float4>();
  f.ag_start = B.nag ? ag_start.as<uint32_t>() : nullptr;
  f.ag = B.nag ? atom_grp.as<uint2>() : nullptr;
  f.S = S.as<float2>();
  f.R = R.as<float2>();
  f.partial = partial.as<double4>();
  f.near4 = near4.as<double4>();
  f.fx = fx.as<double>();
  f.fy = fy.as<double>();
  f.fz = fz.as<double>();
  f.en = en.as<double>();
  const ShapeTables &tb = *tables;
  f.k64 = tb.d_k64.as<double4>();
  f.group = tb.d_group.as<uint32_t>();
  f.r_in = tb.d_r_in.as<uint32_t>();
  f.r_out = tb.d_r_out.as<uint32_t>();
  f.tk = task_args();
  f.sync = fsync.as<uint32_t>();
  f.fail = fail_dev;
  f.timeout = fused_timeout;
  dev::step_fused(f, fused_nwg, fused_lds, stream, fused);
  check_launch("fused step");
}

dev::LrTaskArgs glst_batch::task_args() const {
  const ShapeTables &t = *tables;
  return {t.d_node0.as<uint32_t>(), t.d_nr.as<uint32_t>(),
          t.d_koff.as<uint32_t>(),  t.d_dir64.as<double4>(),
          t.d_dir32.as<float4>(),   t.d_kap64.as<double>(),
          t.d_kap32.as<float>(),    t_split.as<uint32_t>()};
}

void glst_batch::release_arenas() {
  GLST_HIP(hipStreamSynchronize(stream));
  for (DeviceBuffer *b :
       {&d_sys, &d_seg, &atom_cell, &rank, &hist, &cell_count, &cell_start,
        &maxc, &sorted_idx, &cell_sorted, &atoms64, &atoms32, &near4, &fx, &fy,
        &fz, &en, &t_split, &S, &R, &partial, &atom_grp, &ag_start})
    b->release();
}

void glst_batch::zero_ef() {
  for (DeviceBuffer *b : {&fx, &fy, &fz, &en})
    dev::fill_f64(b->as<double>(), 0.0, B.natom, stream);
  check_launch("zero_ef");
}

void glst_batch::near_field(hipStream_t s) {
  dev::sr_batch(mixed(), atoms64.as<double4>(), cell_start.as<uint32_t>(),
                d_sys.as<dev::SysDev>(), L_sr, nullptr, nullptr, nullptr,
                nullptr, near4.as<double4>(), s);
  check_launch("near field");
}

void glst_batch::step_async(const double *rx, const double *ry,
                            const double *rz, const double *qc) {
  if (!rx || !ry || !rz || !qc)
    throw ArgError("calc_ener_force: null coordinate or charge pointer");
  if (use_fused()) {
    step_fused(rx, ry, rz, qc);
    assigned = true;
    return;
  }
  const dev::SysDev *sys = d_sys.as<dev::SysDev>();
  // assign_atoms of every system
  dev::fill_u32(hist.as<uint32_t>(), 0u, std::max<uint64_t>(B.hist, 1),
                stream);
  dev::bin_chunks_batch(mixed(), rx, ry, rz, sys, L_bin,
                        atom_cell.as<uint32_t>(), rank.as<uint32_t>(),
                        hist.as<uint32_t>(), stream);
  dev::hist_offsets_batch(hist.as<uint32_t>(), sys, L_hist,
                          cell_count.as<uint32_t>(), stream);
  uint2 *ag = B.nag ? atom_grp.as<uint2>() : nullptr;
  dev::cell_scan_batch(cell_count.as<uint32_t>(), sys, nsys(),
                       cell_start.as<uint32_t>(), maxc.as<uint32_t>(), stream,
                       ag_start.as<uint32_t>(), ag);
  dev::bin_scatter_batch(mixed(), rx, ry, rz, qc, sys, L_scatter,
                         atom_cell.as<uint32_t>(), rank.as<uint32_t>(),
                         hist.as<uint32_t>(), cell_start.as<uint32_t>(),
                         sorted_idx.as<uint32_t>(), cell_sorted.as<uint32_t>(),
                         atoms64.as<double4>(), atoms32.as<float4>(), stream,
                         ag_start.as<uint32_t>(), ag);
  check_launch("assign_atoms");
  assigned = true;
  if (overlap_near) {
    GLST_HIP(hipEventRecord(ev_assigned, stream));
    GLST_HIP(hipStreamWaitEvent(side, ev_assigned, 0));
    near_field(side);
    GLST_HIP(hipEventRecord(ev_near, side));
  }
  // far field
  const void *atoms = mixed() ? atoms32.as<void>()
                              : static_cast<const void *>(atoms64.as<void>());
  const dev::LrTaskArgs ta = task_args();
  const ShapeTables &tb = *tables;
  if (fused == 3 && mixed())
    dev::sf_all_batch(atoms, cell_start.as<uint32_t>(),
                      cell_sorted.as<uint32_t>(), ta, sys, L_fsf,
                      S.as<void>(), stream);
  else
    dev::sf_batch(mixed(), atoms, cell_start.as<uint32_t>(),
                  cell_sorted.as<uint32_t>(), tb.d_k64.as<double4>(), ta, sys,
                  L_sf, S.as<void>(), stream, tb.d_cheb.as<float>(),
                  tb.d_coff.as<uint32_t>());
  check_launch("calc_sf");
  dev::rmt_batch(mixed(), S.as<void>(), tb.d_k64.as<double4>(),
                 tb.d_group.as<uint32_t>(), tb.d_r_in.as<uint32_t>(),
                 tb.d_r_out.as<uint32_t>(), sys, L_rmt_rest, B.lds,
                 R.as<void>(), stream);
  dev::rmt_wave_batch(S.as<void>(), tb.d_k64.as<double4>(),
                      tb.d_group.as<uint32_t>(), tb.d_r_in.as<uint32_t>(),
                      tb.d_r_out.as<uint32_t>(), sys, L_rmt_small,
                      R.as<void>(), stream);
  check_launch("sum_rmt_sf");
  if (fused == 3 && mixed())
    dev::lr_all_batch(atoms, cell_sorted.as<uint32_t>(), ta, sys, L_flr,
                      R.as<void>(), partial.as<double4>(), stream, ag);
  else
    dev::lr_batch(mixed(), atoms, cell_sorted.as<uint32_t>(), ta, sys, L_lr,
                  R.as<void>(), partial.as<double4>(), stream, ag);
  check_launch("calc_lr_ef");
  if (overlap_near)
    GLST_HIP(hipStreamWaitEvent(stream, ev_near, 0));
  else
    near_field(stream);
  dev::lr_reduce_batch(partial.as<double4>(), sys, L_reduce,
                       near4.as<double4>(), fx.as<double>(), fy.as<double>(),
                       fz.as<double>(), en.as<double>(), stream);
  check_launch("reduce");
}

void glst_batch::drop_graph() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  graph_exec = nullptr;
}

void glst_batch::step(const double *rx, const double *ry, const double *rz,
                      const double *qc) {
  const double *in[4] = {rx, ry, rz, qc};
  if (graph_mode == 0 || !rx || !ry || !rz || !qc) {
    step_async(rx, ry, rz, qc);
    return;
  }
  if (graph_exec && std::equal(in, in + 4, graph_in)) {
    GLST_HIP(hipGraphLaunch(graph_exec, stream));
    assigned = true;
    return;
  }
  if (graph_mode == 2 && !std::equal(in, in + 4, last_in)) {
    std::copy(in, in + 4, last_in);
    step_async(rx, ry, rz, qc);
    return;
  }
  drop_graph();
  hipGraph_t g = nullptr;
  GLST_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  try {
    step_async(rx, ry, rz, qc);
  } catch (...) {
    (void)hipStreamEndCapture(stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  GLST_HIP(hipStreamEndCapture(stream, &g));
  const hipError_t ei = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) graph_exec = nullptr;
  GLST_HIP(ei);
  std::copy(in, in + 4, graph_in);
  GLST_HIP(hipGraphLaunch(graph_exec, stream));
}

double glst_batch::autotune(const double *rx, const double *ry,
                            const double *rz, const double *qc,
                            uint32_t reps) {
  if (reps == 0) throw ArgError("autotune needs at least one repetition");
  hipEvent_t a = nullptr, b = nullptr;
  GLST_HIP(hipEventCreate(&a));
  GLST_HIP(hipEventCreate(&b));
  auto time = [&] {
    step(rx, ry, rz, qc);  // warm-up (and capture in graph mode)
    step(rx, ry, rz, qc);
    GLST_HIP(hipEventRecord(a, stream));
    for (uint32_t r = 0; r < reps; ++r) step(rx, ry, rz, qc);
    GLST_HIP(hipEventRecord(b, stream));
    GLST_HIP(hipEventSynchronize(b));
    check_failed();
    float ms = 0.f;
    GLST_HIP(hipEventElapsedTime(&ms, a, b));
    return static_cast<double>(ms) / reps;
  };
  double best;
  try {
    best = time();
    if (mixed()) {  // one persistent launch or a launch per stage
      const int keep = fused;
      int pick = keep;
      for (int v : {0, 1, 2, 3}) {
        if (v == keep) continue;
        fused = v;
        try {
          plan_fused();
        } catch (const StateError &) {  // this device cannot run it
          fused = keep;
          plan_fused();
          continue;
        }
        drop_graph();
        const double t = time();
        if (t < best) {
          best = t;
          pick = v;
        }
      }
      fused = pick;
      plan_fused();
      drop_graph();
    }
    const int keep_overlap = overlap_near;
    int pick_overlap = keep_overlap;
    for (int v : {0, 1}) {
      if (v == keep_overlap) continue;
      overlap_near = v;
      drop_graph();
      const double t = time();
      if (t < best) {
        best = t;
        pick_overlap = v;
      }
    }
    overlap_near = pick_overlap;
    drop_graph();
    if (B.partial > 0) {
      const uint32_t keep = lr_blocks;
      uint32_t pick = keep;
      for (uint32_t v : {1024u, 2048u, 4096u, 8192u, 16384u, 32768u, 65536u}) {
        if (v == keep) continue;
        lr_blocks = v;
        configure(B.ct, B.cs);
        const double t = time();
        if (t < best) {
          best = t;
          pick = v;
        }
      }
      lr_blocks = pick;
      configure(B.ct, B.cs);
    }
  } catch (...) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  zero_ef();
  GLST_HIP(hipStreamSynchronize(stream));
  return best;
}

void glst_batch::get_ef_async(double *ofx, double *ofy, double *ofz,
                              double *oen) {
  if (!ofx || !ofy || !ofz || !oen)
    throw ArgError("get_ef: null output pointer");
  if (!assigned) throw StateError("get_ef called before calc_ener_force");
  dev::unpermute(sorted_idx.as<uint32_t>(), static_cast<uint32_t>(B.natom),
                 fx.as<double>(), fy.as<double>(), fz.as<double>(),
                 en.as<double>(), ofx, ofy, ofz, oen, stream);
  check_launch("get_ef");
}

size_t glst_batch::device_bytes() const {
  const DeviceBuffer *bufs[] = {
      &d_sys,   &d_seg,   &atom_cell, &rank,    &hist,    &cell_count,
      &cell_start, &maxc, &sorted_idx, &cell_sorted, &atoms64, &atoms32,
      &near4,   &fx,      &fy,        &fz,      &en,      &t_split,
      &S,       &R,       &partial, &atom_grp, &ag_start};
  size_t b = tables ? tables->device_bytes() : 0;
  for (auto p : bufs) b += p->bytes();
  return b;
}

namespace {

std::vector<std::shared_ptr<const BatchSystem>> make_systems(
    int ct, uint32_t nsys, const uint32_t *natom, const double *box,
    const double *rcut, double tol) {
  if (nsys == 0) throw ArgError("a batch needs at least one system");
  if (!natom || !box || !rcut) throw ArgError("null system description");
  for (uint32_t s = 0; s < nsys; ++s) {
    check_system_args(ct, box[3 * s], box[3 * s + 1], box[3 * s + 2], rcut[s],
                      tol);
    if (natom[s] == 0)
      throw ArgError("system " + std::to_string(s) + " has no atoms");
  }
  std::vector<std::shared_ptr<const BatchSystem>> cs;
  cs.reserve(nsys);
  for (uint32_t s = 0; s < nsys; ++s)
    cs.push_back(make_batch_system(
        natom[s], make_shape_setup(ct, box + 3 * s, rcut[s], tol)));
  return cs;
}

}  // namespace

#define NEED_BATCH(b)                                                         \
  if (!(b)) {                                                                 \
    set_global_error("null batch");                                           \
    return GLST_ERR_INVALID;                                                  \
  }

extern "C" {

glst_status_t glst_batch_create(int ct, uint32_t nsys, const uint32_t *natom,
                                const double *box, const double *rcut,
                                double tol, void *stream, glst_batch_t *out) {
  if (!out) {
    set_global_error("glst_batch_create: null output handle");
    return GLST_ERR_INVALID;
  }
  *out = nullptr;
  glst_batch *b = new (std::nothrow) glst_batch();
  if (!b) return GLST_ERR_ALLOC;
  const glst_status_t st = run_guarded(nullptr, [&] {
    auto cs = make_systems(ct, nsys, natom, box, rcut, tol);
    b->open(stream);
    b->configure(ct, std::move(cs));
  });
  if (st != GLST_OK) {
    delete b;
    return st;
  }
  *out = b;
  return GLST_OK;
}

glst_status_t glst_batch_destroy(glst_batch_t b) {
  NEED_BATCH(b);
  delete b;
  return GLST_OK;
}

glst_status_t glst_batch_device_bytes(int ct, uint32_t nsys,
                                      const uint32_t *natom, const double *box,
                                      const double *rcut, double tol,
                                      size_t *bytes) {
  if (!bytes) {
    set_global_error("glst_batch_device_bytes: null output");
    return GLST_ERR_INVALID;
  }
  return run_guarded(nullptr, [&] {
    // LDS capacity of gfx950 (160 KiB per workgroup); the tables are only
    // sized here, on the host
    ShapeTables sizing;
    *bytes = pack_batch(ct, make_systems(ct, nsys, natom, box, rcut, tol),
                        sizing, 160 * 1024)
                 .device_bytes();
  });
}

glst_status_t glst_batch_num_systems(glst_batch_t b, uint32_t *nsys) {
  NEED_BATCH(b);
  if (!nsys) return GLST_ERR_INVALID;
  *nsys = b->nsys();
  return GLST_OK;
}

glst_status_t glst_batch_atom_offset(glst_batch_t b, uint32_t sys,
                                     uint32_t *offset) {
  NEED_BATCH(b);
  return run_guarded(&b->last_error, [&] {
    if (!offset) throw ArgError("null output");
    if (sys >= b->nsys()) throw ArgError("system index out of range");
    *offset = b->B.sys[sys].atom0;
  });
}

glst_status_t glst_batch_get_info(glst_batch_t b, uint32_t sys,
                                  glst_info_t *info) {
  NEED_BATCH(b);
  return run_guarded(&b->last_error, [&] {
    if (!info) throw ArgError("null output");
    if (sys >= b->nsys()) throw ArgError("system index out of range");
    const SystemSetup &u = b->B.cs[sys]->su();
    std::memset(info, 0, sizeof(*info));
    info->natom = b->B.sys[sys].natom;
    info->ncell[0] = u.grid.nx;
    info->ncell[1] = u.grid.ny;
    info->ncell[2] = u.grid.nz;
    info->ncell_total = u.grid.ncell;
    info->cell_dim[0] = info->cell_dim[1] = info->cell_dim[2] = u.grid.h;
    info->ngroup = static_cast<uint32_t>(u.groups.width.size());
    info->num_nodes = u.K;
    info->has_long_range = u.has_lr ? 1u : 0u;
    if (b->assigned) {
      uint32_t m = 0;
      GLST_HIP(hipMemcpyAsync(&m, b->maxc.as<uint32_t>() + sys, 4,
                              hipMemcpyDeviceToHost, b->stream));
      GLST_HIP(hipStreamSynchronize(b->stream));
      info->max_atoms_cell = m;
    }
    info->ct = b->B.ct;
    info->spherical_rule = u.spherical_rule;
    info->device_bytes = b->device_bytes();
  });
}

glst_status_t glst_batch_calc_ener_force_async(glst_batch_t b,
                                               const double *d_rx,
                                               const double *d_ry,
                                               const double *d_rz,
                                               const double *d_qc) {
  NEED_BATCH(b);
  return run_guarded(&b->last_error,
                     [&] { b->step(d_rx, d_ry, d_rz, d_qc); });
}

glst_status_t glst_batch_calc_ener_force(glst_batch_t b, const double *d_rx,
                                         const double *d_ry,
                                         const double *d_rz,
                                         const double *d_qc) {
  NEED_BATCH(b);
  return run_guarded(&b->last_error, [&] {
    b->step(d_rx, d_ry, d_rz, d_qc);
    GLST_HIP(hipStreamSynchronize(b->stream));
    b->check_failed();
  });
}

glst_status_t glst_batch_get_ef(glst_batch_t b, double *d_fx, double *d_fy,
                                double *d_fz, double *d_en) {
  NEED_BATCH(b);
  return run_guarded(&b->last_error, [&] {
    b->get_ef_async(d_fx, d_fy, d_fz, d_en);
    GLST_HIP(hipStreamSynchronize(b->stream));
    b->check_failed();
  });
}

glst_status_t glst_batch_zero_ef(glst_batch_t b) {
  NEED_BATCH(b);
  return run_guarded(&b->last_error, [&] { b->zero_ef(); });
}

glst_status_t glst_batch_set_option(glst_batch_t b, const char *name,
                                    int64_t value) {
  NEED_BATCH(b);
  return run_guarded(&b->last_error, [&] {
    if (!name) throw ArgError("null option name");
    const std::string n = name;
    b->drop_graph();  // every option changes the captured step
    if (n == "overlap_near") {
      if (value != 0 && value != 1)
        throw ArgError("overlap_near must be 0 or 1 for a batch");
      b->overlap_near = static_cast<int>(value);
    } else if (n == "lr_blocks") {
      if (value < 1 || value > (1 << 24))
        throw ArgError("lr_blocks must be in 1 .. 2^24");
      b->lr_blocks = static_cast<uint32_t>(value);
      b->configure(b->B.ct, b->B.cs);
    } else if (n == "graph") {
      if (value < 0 || value > 2) throw ArgError("graph must be 0, 1 or 2");
      b->graph_mode = static_cast<int>(value);
    } else if (n == "sf_cheb") {
      if (value != 0 && value != 1) throw ArgError("sf_cheb must be 0 or 1");
      b->sf_cheb = static_cast<int>(value);
      b->configure(b->B.ct, b->B.cs);
    } else if (n == "rmt_wave") {
      if (value != 0 && value != 1) throw ArgError("rmt_wave must be 0 or 1");
      b->rmt_wave = static_cast<int>(value);
      b->configure(b->B.ct, b->B.cs);
    } else if (n == "fused") {
      if (value < 0 || value > 3) throw ArgError("fused must be 0 .. 3");
      if (value != 0 && !b->mixed())
        throw ArgError("fused launch modes need the mixed path (CT = float)");
      GLST_HIP(hipStreamSynchronize(b->stream));
      b->fused = static_cast<int>(value);
      b->plan_fused();
    } else {
      throw ArgError("unknown option \"" + n + "\"");
    }
  });
}

glst_status_t glst_batch_autotune(glst_batch_t b, const double *d_rx,
                                  const double *d_ry, const double *d_rz,
                                  const double *d_qc, uint32_t reps,
                                  double *best_ms) {
  NEED_BATCH(b);
  return run_guarded(&b->last_error, [&] {
    const double t = b->autotune(d_rx, d_ry, d_rz, d_qc, reps);
    if (best_ms) *best_ms = t;
  });
}

glst_status_t glst_batch_set_stream(glst_batch_t b, void *stream) {
  NEED_BATCH(b);
  return run_guarded(&b->last_error, [&] {
    GLST_HIP(hipStreamSynchronize(b->stream));
    b->drop_graph();
    if (b->own_stream) GLST_HIP(hipStreamDestroy(b->stream));
    b->own_stream = false;
    b->stream = static_cast<hipStream_t>(stream);
    if (!b->stream) {
      GLST_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
      b->own_stream = true;
    }
  });
}

const char *glst_batch_last_error(glst_batch_t b) {
  if (!b) return "null batch";
  return b->last_error.c_str();
}

}  // extern "C"
