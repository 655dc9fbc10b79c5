// Serving loop behind include/glst_serve.h: continuous batching of force
// requests over the batched engine (batch.hpp) with two batches in flight,
// admission by device budget, grid-setup cache, and one set of device-
// resident shape tables shared by both slots.
#include "glst_serve.h"
#include "batch.hpp"

#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <new>
#include <tuple>

using namespace glst;

namespace {

// Grid setups cached by (grid dimensions, cell edge) under the server's CT
// and tol: systems of one grid share them whatever their atoms or exact box.
// Batch systems (a setup plus an atom count) are cached by (grid, cell edge,
// natom), so a resident system's repeat requests reuse the same object and a
// batch of the same composition is not re-packed.  Both caches are bounded;
// least recently used entries go first (queued requests and the configured
// batch hold their own references).
using GridKey = std::tuple<uint32_t, uint32_t, uint32_t, double>;
using SysKey = std::tuple<uint32_t, uint32_t, uint32_t, double, uint32_t>;

template <class Key, class Value>
struct Lru {
  std::map<Key, std::pair<Value, typename std::list<Key>::iterator>> map;
  std::list<Key> order;  // front = most recent
  size_t cap;
  explicit Lru(size_t c) : cap(c) {}
  Value *find(const Key &k) {
    auto it = map.find(k);
    if (it == map.end()) return nullptr;
    order.splice(order.begin(), order, it->second.second);
    return &it->second.first;
  }
  Value &insert(const Key &k, Value v) {
    order.push_front(k);
    auto &slot = map.emplace(k, std::make_pair(std::move(v), order.begin()))
                     .first->second.first;
    while (map.size() > cap) {
      map.erase(order.back());
      order.pop_back();
    }
    return slot;
  }
};

struct SysEntry {
  std::shared_ptr<const BatchSystem> sys;
  size_t bytes = 0;  // device footprint of a batch of this system alone
};

constexpr size_t kMaxCachedGrids = 1024;
constexpr size_t kMaxCachedSystems = 8192;

struct Request {
  uint64_t ticket;
  std::shared_ptr<const BatchSystem> sys;
  size_t bytes;
  const double *in[4];
  double *out[4];
};

// Batches in flight at once.  Each slot owns a batch (arenas and its own
// stream), the packed request inputs/outputs and the copy-run table; while
// one slot's batch runs on the device, the next is admitted, packed and
// launched on the other.
constexpr size_t kSlots = 2;

struct Slot {
  glst_batch batch;
  DeviceBuffer in[4], out[4], runs;
  std::vector<dev::CopyRun> host_runs;
  std::vector<const BatchSystem *> composition;  // of the configured batch
  std::vector<uint64_t> tickets;                 // of the batch in flight
  uint64_t atoms = 0;
  hipEvent_t done = nullptr;
  bool busy = false;
  ~Slot() {
    if (done) (void)hipEventDestroy(done);
  }
};

}  // namespace

struct glst_server {
  std::string last_error;
  int ct = GLST_CT_DOUBLE;
  double tol = 0.0;
  size_t budget = 0;  // whole server; each slot's batch gets budget / kSlots
  uint32_t max_batch = 0;
  size_t lds_limit = 65536;
  std::vector<std::unique_ptr<Slot>> slots;
  std::deque<size_t> in_flight;  // slot indices, oldest first
  Lru<GridKey, std::shared_ptr<const ShapeSetup>> grids{kMaxCachedGrids};
  Lru<SysKey, SysEntry> systems{kMaxCachedSystems};
  std::deque<Request> queue;
  std::vector<uint64_t> done;
  uint64_t next_ticket = 1;
  glst_server_stats_t stats{};

  size_t share() const { return budget / kSlots; }

  ~glst_server() {
    // batches still in flight finish before their arenas go away
    for (auto &sl : slots) {
      (void)hipStreamSynchronize(sl->batch.stream);
      (void)hipStreamSynchronize(sl->batch.side);
    }
  }

  void open(void *stream) {
    // the slots' batches share one set of shape tables
    auto tables = std::make_shared<ShapeTables>();
    for (size_t i = 0; i < kSlots; ++i) {
      auto sl = std::make_unique<Slot>();
      sl->batch.tables = tables;
      // the caller's stream serves the first slot; the others get their own
      sl->batch.open(i == 0 ? stream : nullptr);
      // a composition run twice in a row (resident systems) is captured as
      // a HIP graph and replayed from then on
      sl->batch.graph_mode = 2;
      // with the fused step, both slots' persistent launches must be
      // resident at once: each takes half of the device
      sl->batch.fused_share = kSlots;
      GLST_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
      slots.push_back(std::move(sl));
    }
    lds_limit = slots[0]->batch.lds_limit;
  }

  SysEntry &lookup(uint32_t natom, const double *box, double rcut) {
    uint32_t n[3];
    if (ct == GLST_CT_FLOAT)
      grid_dims<float>(box[0], box[1], box[2], rcut, n[0], n[1], n[2]);
    else
      grid_dims<double>(box[0], box[1], box[2], rcut, n[0], n[1], n[2]);
    const SysKey skey{n[0], n[1], n[2], rcut, natom};
    if (SysEntry *e = systems.find(skey)) {
      ++stats.setup_hits;
      return *e;
    }
    const GridKey gkey{n[0], n[1], n[2], rcut};
    std::shared_ptr<const ShapeSetup> shape;
    if (auto *g = grids.find(gkey)) {
      shape = *g;
      ++stats.setup_hits;
    } else {
      shape = grids.insert(gkey, make_shape_setup(ct, box, rcut, tol));
      ++stats.setups;
    }
    SysEntry e;
    e.sys = make_batch_system(natom, shape);
    ShapeTables sizing;  // host-side sizing only
    e.bytes = pack_batch(ct, {e.sys}, sizing, lds_limit).device_bytes() +
              8 * sizeof(double) * static_cast<size_t>(natom);
    return systems.insert(skey, std::move(e));
  }

  void submit(uint32_t natom, const double *box, double rcut,
              const double *const *inp, double *const *outp,
              uint64_t *ticket) {
    if (!box || !ticket) throw ArgError("null box or ticket pointer");
    if (natom == 0) throw ArgError("a request without atoms");
    for (int k = 0; k < 4; ++k)
      if (!inp[k] || !outp[k]) throw ArgError("null input or output array");
    check_system_args(ct, box[0], box[1], box[2], rcut, tol);
    SysEntry &e = lookup(natom, box, rcut);
    if (e.bytes > share()) {
      ++stats.rejected;
      throw AllocError("request needs " + std::to_string(e.bytes) +
                       " device bytes; a batch of this server may hold " +
                       std::to_string(share()) + " (budget " +
                       std::to_string(budget) + " over " +
                       std::to_string(kSlots) + " batches in flight)");
    }
    Request r;
    r.ticket = next_ticket++;
    r.sys = e.sys;
    r.bytes = e.bytes;
    for (int k = 0; k < 4; ++k) {
      r.in[k] = inp[k];
      r.out[k] = outp[k];
    }
    queue.push_back(r);
    ++stats.submitted;
    *ticket = r.ticket;
  }

  // Admit the oldest queued requests into free slot `sl` and launch them.
  void launch(Slot &sl) {
    size_t take = 0, bytes = 0;
    while (take < queue.size() && (max_batch == 0 || take < max_batch)) {
      const size_t b = queue[take].bytes;
      if (take > 0 && bytes + b > share()) break;
      bytes += b;
      ++take;
    }
    std::vector<std::shared_ptr<const BatchSystem>> cs;
    std::vector<const BatchSystem *> comp;
    uint64_t natom = 0;
    for (size_t i = 0; i < take; ++i) {
      cs.push_back(queue[i].sys);
      comp.push_back(queue[i].sys.get());
      natom += queue[i].sys->natom;
    }
    glst_batch &b = sl.batch;
    if (comp != sl.composition || b.stale()) {
      b.configure(ct, cs);
      if (b.device_bytes() > share()) {
        // arenas grown by earlier batches: size them for this one alone
        b.release_arenas();
        b.arena_headroom = 0;
        b.configure(ct, cs);
        b.arena_headroom = 25;
      }
      sl.composition = comp;
      ++stats.reconfigures;
    } else {
      b.zero_ef();
    }
    hipStream_t s = b.stream;
    for (int k = 0; k < 4; ++k) {
      sl.in[k].grow(natom * sizeof(double), 25);
      sl.out[k].grow(natom * sizeof(double), 25);
    }
    // gather the requests' inputs into the packed arrays, scatter back after
    sl.host_runs.clear();
    uint32_t max_n = 0;
    for (size_t i = 0; i < take; ++i) {
      const uint32_t off = b.B.sys[i].atom0, n = b.B.sys[i].natom;
      max_n = std::max(max_n, n);
      for (int k = 0; k < 4; ++k)
        sl.host_runs.push_back({queue[i].in[k], sl.in[k].as<double>() + off, n});
    }
    for (size_t i = 0; i < take; ++i) {
      const uint32_t off = b.B.sys[i].atom0, n = b.B.sys[i].natom;
      for (int k = 0; k < 4; ++k)
        sl.host_runs.push_back(
            {sl.out[k].as<double>() + off, queue[i].out[k], n});
    }
    sl.runs.grow(sl.host_runs.size() * sizeof(dev::CopyRun), 25);
    GLST_HIP(hipMemcpyAsync(sl.runs.as<void>(), sl.host_runs.data(),
                            sl.host_runs.size() * sizeof(dev::CopyRun),
                            hipMemcpyHostToDevice, s));
    const uint32_t nrun = static_cast<uint32_t>(4 * take);
    dev::copy_runs(sl.runs.as<dev::CopyRun>(), nrun, max_n, s);
    check_launch("gather requests");
    // step() replays the slot's captured graph when its graph mode asks for
    // one and the composition and packed inputs are unchanged
    b.step(sl.in[0].as<double>(), sl.in[1].as<double>(),
           sl.in[2].as<double>(), sl.in[3].as<double>());
    b.get_ef_async(sl.out[0].as<double>(), sl.out[1].as<double>(),
                   sl.out[2].as<double>(), sl.out[3].as<double>());
    dev::copy_runs(sl.runs.as<dev::CopyRun>() + nrun, nrun, max_n, s);
    check_launch("scatter results");
    GLST_HIP(hipEventRecord(sl.done, s));
    sl.tickets.clear();
    for (size_t i = 0; i < take; ++i) sl.tickets.push_back(queue[i].ticket);
    sl.atoms = natom;
    sl.busy = true;
    queue.erase(queue.begin(), queue.begin() + take);
    ++stats.batches;
  }

  uint32_t retire(Slot &sl) {
    sl.batch.check_failed();
    for (uint64_t t : sl.tickets) done.push_back(t);
    const uint32_t n = static_cast<uint32_t>(sl.tickets.size());
    stats.completed += n;
    stats.atoms += sl.atoms;
    sl.tickets.clear();
    sl.busy = false;
    return n;
  }

  // Launch while a slot is free and requests are queued, retire finished
  // batches; with `block`, wait for the oldest batch in flight when nothing
  // else could be done.
  uint32_t advance(bool block) {
    uint32_t completed = 0;
    bool launched = false;
    for (;;) {
      // retire finished batches, oldest first
      while (!in_flight.empty()) {
        Slot &sl = *slots[in_flight.front()];
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) break;
        GLST_HIP(q);
        completed += retire(sl);
        in_flight.pop_front();
      }
      if (queue.empty()) break;
      size_t free_slot = kSlots;
      for (size_t i = 0; i < kSlots; ++i)
        if (!slots[i]->busy) {
          free_slot = i;
          break;
        }
      if (free_slot == kSlots) break;
      launch(*slots[free_slot]);
      in_flight.push_back(free_slot);
      stats.max_in_flight =
          std::max<uint64_t>(stats.max_in_flight, in_flight.size());
      launched = true;
    }
    if (block && completed == 0 && !in_flight.empty()) {
      Slot &sl = *slots[in_flight.front()];
      GLST_HIP(hipEventSynchronize(sl.done));
      completed += retire(sl);
      in_flight.pop_front();
      (void)launched;
    }
    return completed;
  }

  uint32_t unfinished() const {
    size_t n = queue.size();
    for (size_t i : in_flight) n += slots[i]->tickets.size();
    return static_cast<uint32_t>(n);
  }
};

#define NEED_SERVER(v)                                                        \
  if (!(v)) {                                                                 \
    set_global_error("null server");                                          \
    return GLST_ERR_INVALID;                                                  \
  }

extern "C" {

glst_status_t glst_server_create(int ct, double tol, size_t budget_bytes,
                                 uint32_t max_batch, void *stream,
                                 glst_server_t *out) {
  if (!out) {
    set_global_error("glst_server_create: null output handle");
    return GLST_ERR_INVALID;
  }
  *out = nullptr;
  glst_server *v = new (std::nothrow) glst_server();
  if (!v) return GLST_ERR_ALLOC;
  const glst_status_t st = run_guarded(nullptr, [&] {
    if (ct != GLST_CT_DOUBLE && ct != GLST_CT_FLOAT)
      throw ArgError("ct must be GLST_CT_DOUBLE or GLST_CT_FLOAT");
    if (!(tol > 0)) throw ArgError("tol must be positive");
    if (budget_bytes == 0) throw ArgError("the device budget must be positive");
    v->ct = ct;
    v->tol = tol;
    v->budget = budget_bytes;
    v->max_batch = max_batch;
    v->open(stream);
  });
  if (st != GLST_OK) {
    delete v;
    return st;
  }
  *out = v;
  return GLST_OK;
}

glst_status_t glst_server_destroy(glst_server_t srv) {
  NEED_SERVER(srv);
  delete srv;
  return GLST_OK;
}

glst_status_t glst_server_submit(glst_server_t srv, uint32_t natom,
                                 const double *box, double rcut,
                                 const double *d_rx, const double *d_ry,
                                 const double *d_rz, const double *d_qc,
                                 double *d_fx, double *d_fy, double *d_fz,
                                 double *d_en, uint64_t *ticket) {
  NEED_SERVER(srv);
  const double *in[4] = {d_rx, d_ry, d_rz, d_qc};
  double *out[4] = {d_fx, d_fy, d_fz, d_en};
  return run_guarded(&srv->last_error,
                     [&] { srv->submit(natom, box, rcut, in, out, ticket); });
}

glst_status_t glst_server_step(glst_server_t srv, uint32_t *ncompleted) {
  NEED_SERVER(srv);
  return run_guarded(&srv->last_error, [&] {
    const uint32_t n = srv->advance(true);
    if (ncompleted) *ncompleted = n;
  });
}

glst_status_t glst_server_poll(glst_server_t srv, uint32_t *ncompleted) {
  NEED_SERVER(srv);
  return run_guarded(&srv->last_error, [&] {
    const uint32_t n = srv->advance(false);
    if (ncompleted) *ncompleted = n;
  });
}

glst_status_t glst_server_completed(glst_server_t srv, uint64_t *tickets,
                                    uint32_t cap, uint32_t *n) {
  NEED_SERVER(srv);
  return run_guarded(&srv->last_error, [&] {
    if (!n) throw ArgError("null count pointer");
    const size_t m = std::min<size_t>(cap, srv->done.size());
    if (m && !tickets) throw ArgError("null ticket array");
    std::copy(srv->done.begin(), srv->done.begin() + m, tickets);
    srv->done.erase(srv->done.begin(), srv->done.begin() + m);
    *n = static_cast<uint32_t>(m);
  });
}

glst_status_t glst_server_pending(glst_server_t srv, uint32_t *n) {
  NEED_SERVER(srv);
  if (!n) return GLST_ERR_INVALID;
  *n = srv->unfinished();
  return GLST_OK;
}

glst_status_t glst_server_get_stats(glst_server_t srv,
                                    glst_server_stats_t *stats) {
  NEED_SERVER(srv);
  if (!stats) return GLST_ERR_INVALID;
  *stats = srv->stats;
  return GLST_OK;
}

glst_status_t glst_server_set_option(glst_server_t srv, const char *name,
                                     int64_t value) {
  NEED_SERVER(srv);
  return run_guarded(&srv->last_error, [&] {
    if (!name) throw ArgError("null option name");
    const std::string n = name;
    if (n == "fused") {
      if (value < 0 || value > 3) throw ArgError("fused must be 0 .. 3");
      if (value != 0 && srv->ct != GLST_CT_FLOAT)
        throw ArgError("fused launch modes need the mixed path (CT = float)");
      for (auto &sl : srv->slots) {
        GLST_HIP(hipStreamSynchronize(sl->batch.stream));
        sl->batch.drop_graph();
        sl->batch.fused = static_cast<int>(value);
        // a slot is planned when it is next configured
        if (!sl->composition.empty()) sl->batch.plan_fused();
      }
    } else if (n == "graph") {
      if (value < 0 || value > 2) throw ArgError("graph must be 0, 1 or 2");
      for (auto &sl : srv->slots) {
        GLST_HIP(hipStreamSynchronize(sl->batch.stream));
        sl->batch.drop_graph();
        sl->batch.graph_mode = static_cast<int>(value);
      }
    } else if (n == "max_batch") {
      if (value < 0 || value > (int64_t(1) << 31))
        throw ArgError("max_batch must be in 0 .. 2^31");
      srv->max_batch = static_cast<uint32_t>(value);
    } else {
      throw ArgError("unknown option \"" + n + "\"");
    }
  });
}

const char *glst_server_last_error(glst_server_t srv) {
  if (!srv) return "null server";
  return srv->last_error.c_str();
}

}  // extern "C"
