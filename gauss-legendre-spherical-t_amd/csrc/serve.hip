// Serving loop behind include/glst_serve.h: continuous batching of force
// requests over the batched engine (batch.hpp), admission by device budget,
// per-shape setup cache.
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

}  // namespace

struct glst_server {
  std::string last_error;
  int ct = GLST_CT_DOUBLE;
  double tol = 0.0;
  size_t budget = 0;
  uint32_t max_batch = 0;
  glst_batch batch;
  bool batch_open = false;
  std::vector<const BatchSystem *> composition;  // of the configured batch
  Lru<GridKey, std::shared_ptr<const ShapeSetup>> grids{kMaxCachedGrids};
  Lru<SysKey, SysEntry> systems{kMaxCachedSystems};
  std::deque<Request> queue;
  std::vector<uint64_t> done;
  uint64_t next_ticket = 1;
  glst_server_stats_t stats{};
  // packed inputs/outputs of the current batch and the copy-run table
  DeviceBuffer in[4], out[4], runs;
  std::vector<dev::CopyRun> host_runs;

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
    e.bytes = pack_batch(ct, {e.sys}, batch.lds_limit).device_bytes() +
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
    if (e.bytes > budget) {
      ++stats.rejected;
      throw AllocError("request needs " + std::to_string(e.bytes) +
                       " device bytes; the server's budget is " +
                       std::to_string(budget));
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

  uint32_t step() {
    if (queue.empty()) return 0;
    // admission: oldest first while the budget and max_batch allow
    size_t take = 0, bytes = 0;
    while (take < queue.size() && (max_batch == 0 || take < max_batch)) {
      const size_t b = queue[take].bytes;
      if (take > 0 && bytes + b > budget) break;
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
    if (!batch_open) {
      batch.open(nullptr);
      batch_open = true;
    }
    if (comp != composition) {
      batch.configure(ct, std::move(cs));
      composition = comp;
      ++stats.reconfigures;
    } else {
      batch.zero_ef();
    }
    hipStream_t s = batch.stream;
    for (int k = 0; k < 4; ++k) {
      in[k].reserve(natom * sizeof(double));
      out[k].reserve(natom * sizeof(double));
    }
    // gather the requests' inputs into the packed arrays
    host_runs.clear();
    for (size_t i = 0; i < take; ++i) {
      const uint32_t off = batch.B.sys[i].atom0, n = batch.B.sys[i].natom;
      for (int k = 0; k < 4; ++k)
        host_runs.push_back({queue[i].in[k], in[k].as<double>() + off, n});
    }
    for (size_t i = 0; i < take; ++i) {
      const uint32_t off = batch.B.sys[i].atom0, n = batch.B.sys[i].natom;
      for (int k = 0; k < 4; ++k)
        host_runs.push_back({out[k].as<double>() + off, queue[i].out[k], n});
    }
    runs.reserve(host_runs.size() * sizeof(dev::CopyRun));
    GLST_HIP(hipMemcpyAsync(runs.as<void>(), host_runs.data(),
                            host_runs.size() * sizeof(dev::CopyRun),
                            hipMemcpyHostToDevice, s));
    const uint32_t nrun = static_cast<uint32_t>(4 * take);
    dev::copy_runs(runs.as<dev::CopyRun>(), nrun, s);
    check_launch("gather requests");
    batch.step_async(in[0].as<double>(), in[1].as<double>(),
                     in[2].as<double>(), in[3].as<double>());
    batch.get_ef_async(out[0].as<double>(), out[1].as<double>(),
                       out[2].as<double>(), out[3].as<double>());
    dev::copy_runs(runs.as<dev::CopyRun>() + nrun, nrun, s);
    check_launch("scatter results");
    GLST_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < take; ++i) done.push_back(queue[i].ticket);
    queue.erase(queue.begin(), queue.begin() + take);
    ++stats.batches;
    stats.completed += take;
    stats.atoms += natom;
    return static_cast<uint32_t>(take);
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
    v->batch.open(stream);
    v->batch_open = true;
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
    const uint32_t n = srv->step();
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
  *n = static_cast<uint32_t>(srv->queue.size());
  return GLST_OK;
}

glst_status_t glst_server_get_stats(glst_server_t srv,
                                    glst_server_stats_t *stats) {
  NEED_SERVER(srv);
  if (!stats) return GLST_ERR_INVALID;
  *stats = srv->stats;
  return GLST_OK;
}

const char *glst_server_last_error(glst_server_t srv) {
  if (!srv) return "null server";
  return srv->last_error.c_str();
}

}  // extern "C"
