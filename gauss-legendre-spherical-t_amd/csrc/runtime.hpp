// Host runtime shared by the single-system engine (engine.hip) and the
// batched engine (batch.hip): error types, device buffers, kernel timing,
// and the host setup of one system (grid, alpha groups, cubature nodes in
// device order, back-projection splits).
#pragma once

#include "glst_hip.h"
#include "kernels.hpp"
#include "numerics.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace glst {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct AllocError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define GLST_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw HipError(std::string("HIP error: ") + hipGetErrorString(e_) +    \
                     " at " + __FILE__ + ":" + std::to_string(__LINE__));     \
  } while (0)

inline void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw HipError(std::string("HIP launch error in ") + what + ": " +
                   hipGetErrorString(e));
}

class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  ~DeviceBuffer() { release(); }
  // exactly `bytes` (reallocates on any change)
  void resize(size_t bytes) {
    if (bytes == bytes_) return;
    release();
    allocate(bytes);
  }
  // at least `bytes`; never shrinks (buffers reused across batches)
  void reserve(size_t bytes) {
    if (bytes <= bytes_) return;
    release();
    allocate(bytes);
  }
  // reserve with `headroom` percent more on a growth, so that sizes that
  // vary from batch to batch stop reallocating
  void grow(size_t bytes, unsigned headroom) {
    if (bytes <= bytes_) return;
    release();
    allocate(bytes + bytes / 100 * headroom);
  }
  void allocate(size_t bytes) {
    if (bytes == 0) return;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw AllocError("hipMalloc of " + std::to_string(bytes) +
                       " bytes failed: " + hipGetErrorString(e));
    }
    ptr_ = p;
    bytes_ = bytes;
  }
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(ptr_);
  }
  size_t bytes() const { return bytes_; }

 private:
  void *ptr_ = nullptr;
  size_t bytes_ = 0;
};

class KernelTimer {
 public:
  ~KernelTimer() {
    for (auto &kv : stats_)
      for (auto &pr : kv.second.pending) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
      }
    for (auto e : pool_) (void)hipEventDestroy(e);
  }
  bool enabled = false;
  hipEvent_t take() {
    if (!pool_.empty()) {
      hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e;
    GLST_HIP(hipEventCreate(&e));
    return e;
  }
  void begin(const char *name, hipStream_t s) {
    if (!enabled) return;
    cur_name_ = name;
    cur_start_ = take();
    GLST_HIP(hipEventRecord(cur_start_, s));
  }
  // Called from a destructor: never throws; a failed record drops the
  // sample.
  void end(hipStream_t s) noexcept {
    if (!enabled) return;
    try {
      hipEvent_t stop = take();
      if (hipEventRecord(stop, s) == hipSuccess)
        stats_[cur_name_].pending.emplace_back(cur_start_, stop);
    } catch (...) {
    }
  }
  void collect() {
    for (auto &kv : stats_) {
      for (auto &pr : kv.second.pending) {
        GLST_HIP(hipEventSynchronize(pr.second));
        float ms = 0.f;
        GLST_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
        kv.second.ms += ms;
        kv.second.n += 1;
        pool_.push_back(pr.first);
        pool_.push_back(pr.second);
      }
      kv.second.pending.clear();
    }
  }
  void get(const std::string &name, double &ms, uint32_t &n) {
    collect();
    auto it = stats_.find(name);
    ms = it == stats_.end() ? 0.0 : it->second.ms;
    n = it == stats_.end() ? 0 : it->second.n;
  }
  void reset() {
    collect();
    stats_.clear();
  }

 private:
  struct Stat {
    double ms = 0.0;
    uint32_t n = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  };
  std::map<std::string, Stat> stats_;
  std::vector<hipEvent_t> pool_;
  std::string cur_name_;
  hipEvent_t cur_start_{};
};

struct Scope {
  KernelTimer &t;
  hipStream_t s;
  Scope(KernelTimer &timer, const char *name, hipStream_t stream)
      : t(timer), s(stream) {
    t.begin(name, s);
  }
  ~Scope() { t.end(s); }
};

// Process-wide last error (functions without a handle) and the t-design
// directory (GLST_TDESIGN_DIR or glst_set_tdesign_dir); engine.hip.
void set_global_error(const std::string &m);
std::string tdesign_dir();

// Runs f, mapping exceptions to status codes; the message goes to *err (a
// handle's last error) or, with err == nullptr, to the process-wide one.
template <class F>
glst_status_t run_guarded(std::string *err, F &&f) {
  auto record = [&](const std::string &m) {
    if (err)
      *err = m;
    else
      set_global_error(m);
  };
  try {
    f();
    if (err) err->clear();
    return GLST_OK;
  } catch (const ArgError &e) {
    record(e.what());
    return GLST_ERR_INVALID;
  } catch (const StateError &e) {
    record(e.what());
    return GLST_ERR_STATE;
  } catch (const AllocError &e) {
    record(e.what());
    return GLST_ERR_ALLOC;
  } catch (const HipError &e) {
    record(e.what());
    return GLST_ERR_HIP;
  } catch (const std::exception &e) {
    const std::string m = e.what();
    record(m);
    if (m.find("open") != std::string::npos ||
        m.find("Malformed") != std::string::npos ||
        m.find("directory") != std::string::npos)
      return GLST_ERR_IO;
    return GLST_ERR_NUMERICS;
  }
}

// ---- host setup of one system ------------------------------------------------

// Grid, alpha groups and cubature of one system (src/glst_force.cu:178-327,
// src/cubature.cu:99-159).  The grid uses CT arithmetic (ct = GLST_CT_*).
struct SystemSetup {
  dev::GridShape grid{};
  double inv_cell = 0.0;  // CT(1)/CT(rcut), widened
  bool has_lr = false;    // all n_d > 2 (src/glst_force.cu:667-670)
  AlphaGroups groups;
  NodeTable nodes;
  uint32_t K = 0, Kp = 0;  // nodes, padded to a multiple of 256
  int spherical_rule = 0;  // 0 product rule, 1 design files
};

inline void check_system_args(int ct, double bx, double by, double bz,
                              double rcut, double tol) {
  if (ct != GLST_CT_DOUBLE && ct != GLST_CT_FLOAT)
    throw ArgError("ct must be GLST_CT_DOUBLE or GLST_CT_FLOAT");
  if (!(bx > 0 && by > 0 && bz > 0 && rcut > 0))
    throw ArgError("box dimensions and rcut must be positive");
  if (!(tol > 0)) throw ArgError("tol must be positive");
}

inline SystemSetup setup_system(int ct, double bx, double by, double bz,
                                double rcut, double tol) {
  check_system_args(ct, bx, by, bz, rcut, tol);
  SystemSetup S;
  double cell;
  if (ct == GLST_CT_FLOAT) {
    grid_dims<float>(bx, by, bz, rcut, S.grid.nx, S.grid.ny, S.grid.nz);
    const float cf = static_cast<float>(rcut);
    cell = static_cast<double>(cf);
    S.inv_cell = static_cast<double>(1.0f / cf);
  } else {
    grid_dims<double>(bx, by, bz, rcut, S.grid.nx, S.grid.ny, S.grid.nz);
    cell = rcut;
    S.inv_cell = 1.0 / rcut;
  }
  S.grid.h = cell;
  const uint64_t nc = static_cast<uint64_t>(S.grid.nx) * S.grid.ny * S.grid.nz;
  if (nc == 0 || nc > 0x7FFFFFFFull)
    throw ArgError("cell grid of " + std::to_string(nc) +
                   " cells is out of range");
  S.grid.ncell = static_cast<uint32_t>(nc);
  S.has_lr = S.grid.nx > 2 && S.grid.ny > 2 && S.grid.nz > 2;
  S.groups = alpha_groups(S.grid.nx, S.grid.ny, S.grid.nz, cell, cell, cell,
                          tol);
  const std::string dir = tdesign_dir();
  S.spherical_rule = dir.empty() ? 0 : 1;
  S.nodes = build_nodes(tol, S.groups, dir);
  S.K = static_cast<uint32_t>(S.nodes.kx.size());
  S.Kp = (S.K + 255u) / 256u * 256u;
  return S;
}

// Offset from the cell centre (|d|_1, Angstrom) beyond which the mixed
// path's FP32 phases of a system's nodes could pass 63 revolutions (kFarAtom
// in kernels.hpp): 63 / max_k |k| / 2 pi.  No nodes: never.
inline float far_offset(const NodeTable &nodes) {
  double kmax = 0.0;
  for (const auto &kap : nodes.kappa)
    for (double k : kap) kmax = std::max(kmax, std::fabs(k));
  if (!(kmax > 0.0)) return 3.0e38f;
  constexpr double kTwoPi = 6.283185307179586476925286766559;
  return static_cast<float>(63.0 / (kmax / kTwoPi));
}

// Node and task tables in device order.  Inside each alpha group the nodes
// are direction-major: task (group, direction m) covers nodes first + m*nr
// .. + nr - 1 with k = kappa_i s_m.  Node indices are local to the system;
// koff indexes kappa64/kappa32 of the system.
struct NodeLayout {
  std::vector<double4> k64;     // (kx, ky, kz, W) per device node, Kp
  std::vector<uint32_t> group;  // group per device node, Kp
  std::vector<uint32_t> node0, nr, koff;  // per task
  std::vector<double4> dir64;
  std::vector<float4> dir32;
  std::vector<double> kappa64;  // kappa_i per group
  std::vector<float> kappa32;   // kappa_i / 2 pi
  std::vector<uint32_t> dev_of_ref;  // device index of each reference node
};

inline NodeLayout layout_nodes(const NodeTable &nodes, uint32_t K,
                               uint32_t Kp) {
  NodeLayout L;
  const double inv2pi = 1.0 / 6.283185307179586476925286766559;
  L.k64.assign(Kp, make_double4(0, 0, 0, 0));
  L.group.assign(Kp, 0);
  L.dev_of_ref.assign(K, 0);
  for (size_t g = 0; g < nodes.dirs.size(); ++g) {
    const uint32_t nr = static_cast<uint32_t>(nodes.kappa[g].size());
    const uint32_t M = static_cast<uint32_t>(nodes.dirs[g].size());
    const uint32_t first = nodes.first[g];
    const uint32_t koff = static_cast<uint32_t>(L.kappa64.size());
    for (double kap : nodes.kappa[g]) {
      L.kappa64.push_back(kap);
      L.kappa32.push_back(static_cast<float>(kap * inv2pi));
    }
    for (uint32_t m = 0; m < M; ++m) {
      const auto &sd = nodes.dirs[g][m];
      L.node0.push_back(first + m * nr);
      L.nr.push_back(nr);
      L.koff.push_back(koff);
      L.dir64.push_back(make_double4(sd[0], sd[1], sd[2], 0.0));
      L.dir32.push_back(make_float4(static_cast<float>(sd[0]),
                                    static_cast<float>(sd[1]),
                                    static_cast<float>(sd[2]), 0.f));
      for (uint32_t i = 0; i < nr; ++i) {
        const uint32_t ref = first + i * M + m;  // reference order
        const uint32_t dev = first + m * nr + i;  // direction-major
        L.dev_of_ref[ref] = dev;
        L.k64[dev] = make_double4(nodes.kx[ref], nodes.ky[ref], nodes.kz[ref],
                                  nodes.w[ref]);
        L.group[dev] = static_cast<uint32_t>(g);
      }
    }
  }
  return L;
}

// Back-projection splits: contiguous task ranges that never cross an alpha
// group (all tasks of a split then share the radial count and kappa_i);
// each group gets about target * K_g / K splits of equal size.  Task and
// split indices are local to the system.
struct SplitPlan {
  std::vector<uint32_t> split;         // nsplit + 1 task boundaries
  std::vector<dev::LrGroup> lr_groups;  // per group: nr, first split, splits
  std::vector<dev::SfGroup> sf_groups;  // per group: nr, first task, tasks
  uint32_t nsplit = 0;
};

constexpr uint32_t kSplitAlign = 8;

inline SplitPlan make_splits(const NodeTable &nodes, uint32_t K,
                             uint32_t target) {
  SplitPlan P;
  P.split.assign(1, 0);
  uint32_t base = 0;
  for (size_t g = 0; g < nodes.dirs.size(); ++g) {
    const uint32_t M = static_cast<uint32_t>(nodes.dirs[g].size());
    const double share =
        static_cast<double>(target) * nodes.count[g] / std::max<uint32_t>(K, 1);
    // split boundaries fall on multiples of kSplitAlign directions (the
    // mixed back-projection serves directions in passes of that many lanes)
    const uint32_t units = (M + kSplitAlign - 1) / kSplitAlign;
    uint32_t ns = static_cast<uint32_t>(share + 0.5);
    ns = std::max<uint32_t>(1, std::min<uint32_t>(ns, units));
    P.lr_groups.push_back({static_cast<uint32_t>(nodes.kappa[g].size()),
                           static_cast<uint32_t>(P.split.size() - 1), ns});
    P.sf_groups.push_back(
        {static_cast<uint32_t>(nodes.kappa[g].size()), base, M});
    for (uint32_t q = 1; q <= ns; ++q) {
      const uint64_t u = (static_cast<uint64_t>(units) * q) / ns;
      P.split.push_back(base + static_cast<uint32_t>(
                                   std::min<uint64_t>(u * kSplitAlign, M)));
    }
    base += M;
  }
  if (P.split.size() == 1) P.split.push_back(0);  // no tasks at all
  P.nsplit = static_cast<uint32_t>(P.split.size() - 1);
  return P;
}

// ---- Chebyshev form of the mixed structure factors -------------------------
//
// The atoms of a cell lie within half a cell edge of its centre in every
// axis, so p = s.d stays in [-rho, rho] (cheb_rho: sqrt(3) h / 2 with a
// margin for rounding), and (Jacobi-Anger, x = p / rho, eps_0 = 1, eps_m = 2)
//   exp(-i kappa p) = sum_m (-i)^m eps_m J_m(kappa rho) T_m(x).
// S at the NR radial nodes of a direction is then a fixed NR x M matrix
// applied to the cell's M Chebyshev moments sum_j q_j T_m(x_j), which cost two
// FMAs per (atom, term) against a multiply, a sine, a cosine and two FMAs
// per (atom, radial node).  M is the smallest count whose tail bound
// 2 sum_{m >= M} |J_m(kappa_max rho)| is below kChebTol.
constexpr double kChebTol = 3e-8;
using dev::kChebMaxTerms;
using dev::cheb_rho;

// J_0(z) .. J_{n-1}(z) for z >= 0: Miller's backward recurrence
// J_{m-1} = (2m / z) J_m - J_{m+1} from well above max(n, z), normalised by
// J_0 + 2 (J_2 + J_4 + ...) = 1.
inline void bessel_j(double z, uint32_t n, double *out) {
  if (n == 0) return;
  if (!(z > 0.0)) {
    out[0] = 1.0;
    for (uint32_t m = 1; m < n; ++m) out[m] = 0.0;
    return;
  }
  const double big = std::max(static_cast<double>(n), z);
  const uint32_t top =
      2 * ((static_cast<uint32_t>(big) + 20 +
            static_cast<uint32_t>(std::sqrt(40.0 * big))) / 2);
  std::vector<double> b(top + 2, 0.0);
  b[top] = 1e-30;
  for (uint32_t m = top; m >= 1; --m) {
    b[m - 1] = (2.0 * m / z) * b[m] - b[m + 1];
    if (std::fabs(b[m - 1]) > 1e250)
      for (uint32_t k = m - 1; k <= top; ++k) b[k] *= 1e-250;
  }
  double norm = b[0];
  for (uint32_t m = 2; m <= top; m += 2) norm += 2.0 * b[m];
  for (uint32_t m = 0; m < n; ++m) out[m] = m <= top ? b[m] / norm : 0.0;
}

// Terms needed for phases up to z radians (kChebMaxTerms + 1: more than the
// kernels hold).
inline uint32_t cheb_terms(double z) {
  constexpr uint32_t kTail = 40;
  std::vector<double> J(kChebMaxTerms + kTail + 1);
  bessel_j(z, static_cast<uint32_t>(J.size()), J.data());
  for (uint32_t M = 1; M <= kChebMaxTerms; ++M) {
    double tail = 0.0;
    for (uint32_t m = M; m < J.size(); ++m) tail += std::fabs(J[m]);
    if (2.0 * tail < kChebTol) return M;
  }
  return kChebMaxTerms + 1;
}

// Chooses the Chebyshev form for the groups where it takes fewer issue
// slots per (atom, direction) than the direct sum (2 M + 4 plus the matrix,
// NR M over at least 32 atoms, against about 6 NR: the sine and cosine cost
// about 1.6 FMA slots each) and fills their matrices: for group g with
// sf_groups[g].m = M > 0, table[coff + i M + m] = eps_m J_m(kappa_i rho) times
// the real (m even: (-1)^(m/2)) or imaginary (m odd: (-1)^((m+1)/2)) part of
// (-i)^m.  Groups left at m = 0 keep the direct sum.
inline std::vector<float> plan_chebyshev(const NodeTable &nodes, double h,
                                         std::vector<dev::SfGroup> &groups) {
  std::vector<float> table;
  const double rho = cheb_rho(h);
  for (size_t g = 0; g < groups.size() && g < nodes.kappa.size(); ++g) {
    dev::SfGroup &sg = groups[g];
    sg.m = 0;
    sg.coff = 0;
    const std::vector<double> &kap = nodes.kappa[g];
    if (kap.empty() || sg.ntask == 0) continue;
    const double zmax = *std::max_element(kap.begin(), kap.end()) * rho;
    const uint32_t M = cheb_terms(zmax);
    const double nr = static_cast<double>(kap.size());
    if (M > kChebMaxTerms || 2.0 * M + 4.0 + nr * M / 32.0 >= 6.0 * nr)
      continue;
    sg.m = M;
    sg.coff = static_cast<uint32_t>(table.size());
    std::vector<double> J(M);
    for (double k : kap) {
      bessel_j(k * rho, M, J.data());
      for (uint32_t m = 0; m < M; ++m) {
        const double eps = m == 0 ? 1.0 : 2.0;
        const double sign = (m % 2 == 0) ? ((m / 2) % 2 ? -1.0 : 1.0)
                                         : (((m + 1) / 2) % 2 ? -1.0 : 1.0);
        table.push_back(static_cast<float>(eps * sign * J[m]));
      }
    }
  }
  return table;
}

// Shell-sum chunks of the far field pipelined across alpha groups.  Group g
// owns device nodes [first_g, first_g + count_g); the shell sums run in
// workgroup tiles of T consecutive nodes, Kp / T tiles in all.  Chunk g is
// the tiles [tile_end[g-1], tile_end[g]) with tile_end[g] =
// floor((first_g + count_g) / T) (the last chunk ends at Kp / T), so chunk g
// holds nodes of groups <= g only and may start once S(k) of group g is
// complete.  A tile that straddles groups g and g+1 belongs to chunk g+1, so
// the back-projection of group g waits for chunk rmt_need[g], the first
// chunk whose end covers the group's last node.
struct PipelinePlan {
  std::vector<uint32_t> tile_end;  // per group: end of its chunk (tiles)
  std::vector<uint32_t> rmt_need;  // per group: chunk its back-projection needs
};

inline PipelinePlan plan_pipeline(const NodeTable &nodes, size_t G,
                                  uint32_t Kp, uint32_t T) {
  PipelinePlan P;
  if (G == 0 || T == 0) return P;
  if (nodes.first.size() < G || nodes.count.size() < G)
    throw StateError("pipeline plan: fewer node groups than alpha groups");
  const uint32_t ntile = Kp / T;
  P.tile_end.assign(G, ntile);
  P.rmt_need.assign(G, static_cast<uint32_t>(G - 1));
  for (size_t g = 0; g + 1 < G; ++g) {
    const uint64_t end = static_cast<uint64_t>(nodes.first[g]) + nodes.count[g];
    P.tile_end[g] = static_cast<uint32_t>(std::min<uint64_t>(end / T, ntile));
    if (g > 0) P.tile_end[g] = std::max(P.tile_end[g], P.tile_end[g - 1]);
  }
  for (size_t g = 0; g < G; ++g) {
    const uint64_t end = static_cast<uint64_t>(nodes.first[g]) + nodes.count[g];
    size_t j = g;
    while (j + 1 < G && static_cast<uint64_t>(P.tile_end[j]) * T < end) ++j;
    P.rmt_need[g] = static_cast<uint32_t>(j);
  }
  return P;
}

}  // namespace glst
