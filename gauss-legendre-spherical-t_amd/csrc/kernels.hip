// HIP kernels of the GLST force step, written for gfx950 (CDNA4, wave64).
//
// Mapping choices:
//  * assign_atoms: a deterministic, atomics-free stable counting sort.  Each
//    1024-atom chunk is bitonic-sorted by (cell, index) in LDS, chunk
//    histograms are scanned per cell, and atoms scatter straight to their
//    final slot.  The order inside a cell is by original index, which is what
//    a stable radix sort keyed by cell produces.
//  * calc_sf (mixed): one wave per (cell, 64 directions of an alpha group),
//    lane = direction, the group's radial magnitudes unrolled as registers;
//    atom records are wave-uniform scalar loads, so the inner loop is VALU
//    and transcendental work only.  S is stored in node runs (sf_slot), so
//    each wave writes whole runs of its cell.
//  * sum_rmt_sf: a workgroup holds the whole cell grid of one or more nodes
//    in LDS (read as z-lines of the node runs) and does the 3-D summed-volume
//    table and the shell gather there.
//  * calc_lr_ef: lane = atom, 64 consecutive cell-sorted atoms per wave (no
//    lanes idle at cell boundaries), one launch per alpha group with its
//    radial magnitudes unrolled; remote sums are per-lane loads that
//    coalesce to the 1-2 cells a wave spans.
//  * calc_sr_ef: one 256-thread workgroup per cell, a cell of n atoms
//    covered floor(256/n) times by interleaved neighbour phases.
//  * Every kernel body is a __device__ function shared by the single-system
//    kernel and its batched twin, which looks its system up in a run table.
#include "kernels.hpp"
#include "sincos64.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <array>
#include <cmath>
#include <type_traits>
#include <mutex>
#include <stdexcept>
#include <utility>

namespace glst {
namespace dev {
namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

// Atoms per atom group of the mixed back-projection: consecutive sorted atoms
// of one cell (a cell of n atoms has ceil(n / kAgAtoms) groups).
#ifndef GLST_AG_ATOMS
#define GLST_AG_ATOMS 4
#endif
constexpr uint32_t kAgAtoms = GLST_AG_ATOMS;
static_assert(kAgAtoms == 4 || kAgAtoms == 8, "atoms per group: 4 or 8");

// Blocks are dealt to the 8 XCDs round-robin (linear block b runs on XCD
// b % 8).  Renumber so that each XCD processes a contiguous range of logical
// tiles, which keeps neighbouring tiles in the same L2.
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t q = nb >> 3, r = nb & 7u;
  const uint32_t x = b & 7u, i = b >> 3;
  return x * q + (x < r ? x : r) + i;
}

// Tuning knobs (compile-time): FP32 partial-sum lengths before the FP64 flush
// (atoms for the node-per-lane S(k), directions for the back-projection).
#ifndef GLST_SF_FLUSH
#define GLST_SF_FLUSH 32
#endif
#ifndef GLST_LR_FLUSH_TASKS
#define GLST_LR_FLUSH_TASKS 4
#endif
#ifndef GLST_LR_EN_FLUSH_TASKS
#define GLST_LR_EN_FLUSH_TASKS 16
#endif
#ifndef GLST_ACCURATE_SINCOS
#define GLST_ACCURATE_SINCOS 0
#endif

// sin(2 pi t), cos(2 pi t) for t in revolutions.  The hardware V_SIN_F32 /
// V_COS_F32 take revolutions directly and reduce the argument themselves:
// on gfx950 their absolute error stays below 1.3e-7 for |t| up to 64
// revolutions (measured against FP64 sin/cos over 4M uniform samples); the
// ISA's input domain is |t| <= 256.  The phases here are k.d with d the
// offset from the cell centre, a few revolutions at most for atoms inside
// the box (|t| > 256 needs an atom hundreds of Angstrom outside it), so no
// V_FRACT_F32 is spent (GLST_NO_FRACT=0 restores it).
#if GLST_ACCURATE_SINCOS
__device__ inline float sin_rev(float t) { return sinpif(2.0f * t); }
__device__ inline float cos_rev(float t) { return cospif(2.0f * t); }
#else
__device__ inline float sin_rev(float t) { return __builtin_amdgcn_sinf(t); }
__device__ inline float cos_rev(float t) { return __builtin_amdgcn_cosf(t); }
#endif
#ifndef GLST_NO_FRACT
#define GLST_NO_FRACT 1
#endif
#if GLST_NO_FRACT
__device__ inline float fract_rev(float t) { return t; }
#else
__device__ inline float fract_rev(float t) { return __builtin_amdgcn_fractf(t); }
#endif
// Phases of atoms marked kFarAtom (offsets far outside their cell) are
// reduced to [0, 1) revolutions whatever GLST_NO_FRACT says.
template <bool FAR>
__device__ inline float reduce_rev(float t) {
  if constexpr (FAR)
    return __builtin_amdgcn_fractf(t);
  else
    return fract_rev(t);
}

// ---------------------------------------------------------------------------
// fills

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_f64_kernel(double *p, double v, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// binning

// c_d = (int)(CT(r_d) * inv), clamped to [0, n-1] (src/glst_force.cu:369-377).
// Written as a compare chain so no out-of-range float->int conversion is ever
// executed; the result equals truncate-then-clamp for every input (NaN -> 0).
template <typename CT>
__device__ inline uint32_t cell_coord(double r, CT inv, uint32_t n) {
  const CT v = static_cast<CT>(r) * inv;
  if (!(v >= static_cast<CT>(0))) return 0u;
  if (v >= static_cast<CT>(n)) return n - 1u;
  const int c = static_cast<int>(v);
  return static_cast<uint32_t>(c) < n ? static_cast<uint32_t>(c) : n - 1u;
}

// One chunk of kBinChunk atoms of one system (pointers at the system's first
// atom and histogram), sorted by NT threads (kBinChunk / NT keys each) in
// `keys` (kBinChunk entries of LDS).  ZERO: the chunk first clears its own
// histogram row (a launch that has not cleared the histogram before).
template <typename CT, uint32_t NT = kBinChunk, bool ZERO = false>
__device__ __forceinline__ void bin_chunk_body(
    const double *__restrict__ rx, const double *__restrict__ ry,
    const double *__restrict__ rz, uint32_t natom, CT inv, uint32_t nx,
    uint32_t ny, uint32_t nz, uint32_t ncell, uint32_t *__restrict__ atom_cell,
    uint32_t *__restrict__ rank, uint32_t *__restrict__ hist, uint32_t chunk,
    unsigned long long *__restrict__ keys) {
  static_assert(kBinChunk % NT == 0, "threads must divide the chunk");
  constexpr uint32_t E = kBinChunk / NT;
  if (ZERO)
    for (uint32_t c = threadIdx.x; c < ncell; c += NT)
      hist[static_cast<size_t>(chunk) * ncell + c] = 0u;
#pragma unroll
  for (uint32_t e = 0; e < E; ++e) {
    const uint32_t t = threadIdx.x + e * NT;
    const uint32_t i = chunk * kBinChunk + t;
    uint32_t cell = 0xFFFFFFFFu;  // sorts after every real cell
    if (i < natom) {
      const uint32_t cx = cell_coord<CT>(rx[i], inv, nx);
      const uint32_t cy = cell_coord<CT>(ry[i], inv, ny);
      const uint32_t cz = cell_coord<CT>(rz[i], inv, nz);
      cell = (cx * ny + cy) * nz + cz;
      atom_cell[i] = cell;
    }
    keys[t] = (static_cast<unsigned long long>(cell) << 32) | t;
  }
  __syncthreads();
  for (uint32_t k = 2; k <= kBinChunk; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (uint32_t e = 0; e < E; ++e) {
        const uint32_t t = threadIdx.x + e * NT;
        const uint32_t p = t ^ j;
        if (p > t) {
          const unsigned long long a = keys[t], b = keys[p];
          const bool ascending = (t & k) == 0;
          if ((a > b) == ascending) {
            keys[t] = b;
            keys[p] = a;
          }
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (uint32_t e = 0; e < E; ++e) {
    const uint32_t t = threadIdx.x + e * NT;
    const unsigned long long key = keys[t];
    const uint32_t c = static_cast<uint32_t>(key >> 32);
    if (c != 0xFFFFFFFFu) {
      // first sorted position holding this cell
      uint32_t lo = 0, hi = t;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (static_cast<uint32_t>(keys[mid] >> 32) < c)
          lo = mid + 1;
        else
          hi = mid;
      }
      const uint32_t local = static_cast<uint32_t>(key & 0xFFFFFFFFull);
      rank[chunk * kBinChunk + local] = t - lo;
      const bool last = (t == kBinChunk - 1) ||
                        (static_cast<uint32_t>(keys[t + 1] >> 32) != c);
      if (last) hist[static_cast<size_t>(chunk) * ncell + c] = t - lo + 1;
    }
  }
}

template <typename CT>
__global__ __launch_bounds__(1024) void bin_chunk_kernel(
    const double *__restrict__ rx, const double *__restrict__ ry,
    const double *__restrict__ rz, uint32_t natom, CT inv, uint32_t nx,
    uint32_t ny, uint32_t nz, uint32_t ncell, uint32_t *__restrict__ atom_cell,
    uint32_t *__restrict__ rank, uint32_t *__restrict__ hist) {
  __shared__ unsigned long long keys[kBinChunk];
  bin_chunk_body<CT>(rx, ry, rz, natom, inv, nx, ny, nz, ncell, atom_cell,
                     rank, hist, blockIdx.x, keys);
}

__device__ __forceinline__ void hist_offsets_body(uint32_t *__restrict__ hist,
                                                  uint32_t nchunk,
                                                  uint32_t ncell,
                                                  uint32_t *__restrict__ count,
                                                  uint32_t c) {
  if (c >= ncell) return;
  // the counts of kBatch chunks are loaded together, then overwritten with
  // their offsets: one memory round trip per batch instead of per chunk
  constexpr uint32_t kBatch = 32;
  uint32_t run = 0, ch = 0;
  for (; ch + kBatch <= nchunk; ch += kBatch) {
    uint32_t v[kBatch];
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u)
      v[u] = hist[static_cast<size_t>(ch + u) * ncell + c];
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u) {
      hist[static_cast<size_t>(ch + u) * ncell + c] = run;
      run += v[u];
    }
  }
  for (; ch < nchunk; ++ch) {
    const size_t o = static_cast<size_t>(ch) * ncell + c;
    const uint32_t v = hist[o];
    hist[o] = run;
    run += v;
  }
  count[c] = run;
}

__global__ void hist_offsets_kernel(uint32_t *__restrict__ hist,
                                    uint32_t nchunk, uint32_t ncell,
                                    uint32_t *__restrict__ count) {
  hist_offsets_body(hist, nchunk, ncell, count,
                    blockIdx.x * blockDim.x + threadIdx.x);
}

// start[c] = base + exclusive scan of count (start[ncell] too when
// write_end); *maxc = max count.  With ag_start: ag_start[c] = exclusive scan
// of the cells' atom-group counts ceil(count / kAgAtoms), and the entries of
// ag from the total up to nag_max are set to (0, 0) (bin_scatter fills the
// others).  NT threads, `lds` 3 * NT words.
template <uint32_t NT = 1024>
__device__ __forceinline__ void cell_scan_body(
    const uint32_t *__restrict__ count, uint32_t ncell,
    uint32_t *__restrict__ start, uint32_t *__restrict__ maxc, uint32_t base,
    bool write_end, uint32_t *__restrict__ lds,
    uint32_t *__restrict__ ag_start = nullptr, uint2 *__restrict__ ag = nullptr,
    uint32_t nag_max = 0) {
  uint32_t *const sums = lds, *const maxs = lds + NT, *const gsum = lds + 2 * NT;
  const uint32_t tid = threadIdx.x;
  const uint32_t seg = (ncell + NT - 1) / NT;
  const uint32_t b = tid * seg;
  const uint32_t e = (b + seg < ncell) ? b + seg : ncell;
  uint32_t s = 0, m = 0, gs = 0;
  for (uint32_t c = b; c < e; ++c) {
    const uint32_t n = count[c];
    s += n;
    m = n > m ? n : m;
    gs += (n + kAgAtoms - 1) / kAgAtoms;
  }
  sums[tid] = s;
  maxs[tid] = m;
  gsum[tid] = gs;
  __syncthreads();
  for (uint32_t off = 1; off < NT; off <<= 1) {
    const uint32_t add = tid >= off ? sums[tid - off] : 0u;
    const uint32_t mx = tid >= off ? maxs[tid - off] : 0u;
    const uint32_t gadd = tid >= off ? gsum[tid - off] : 0u;
    __syncthreads();
    sums[tid] += add;
    maxs[tid] = mx > maxs[tid] ? mx : maxs[tid];
    gsum[tid] += gadd;
    __syncthreads();
  }
  uint32_t run = base + sums[tid] - s;  // exclusive prefix of this segment
  uint32_t grun = gsum[tid] - gs;
  for (uint32_t c = b; c < e; ++c) {
    start[c] = run;
    run += count[c];
    if (ag_start) {
      ag_start[c] = grun;
      grun += (count[c] + kAgAtoms - 1) / kAgAtoms;
    }
  }
  if (tid == NT - 1) {
    if (write_end) start[ncell] = base + sums[NT - 1];
    *maxc = maxs[NT - 1];
  }
  if (ag)
    for (uint32_t k = gsum[NT - 1] + tid; k < nag_max; k += NT)
      ag[k] = make_uint2(0u, 0u);
}

__global__ __launch_bounds__(1024) void cell_scan_kernel(
    const uint32_t *__restrict__ count, uint32_t ncell,
    uint32_t *__restrict__ start, uint32_t *__restrict__ maxc,
    uint32_t *__restrict__ ag_start, uint2 *__restrict__ ag,
    uint32_t nag_max) {
  __shared__ uint32_t lds[3 * 1024];
  cell_scan_body(count, ncell, start, maxc, 0u, true, lds, ag_start, ag,
                 nag_max);
}

// Atom i of one system (inputs, atom_cell, rank, hist and start at the
// system's first atom / histogram / cell; start holds batch-wide slots);
// sorted_idx receives atom0 + i, cell_sorted the system-local cell.
template <bool MIXED>
__device__ __forceinline__ void bin_scatter_body(
    const double *__restrict__ rx, const double *__restrict__ ry,
    const double *__restrict__ rz, const double *__restrict__ qc,
    uint32_t natom, const uint32_t *__restrict__ atom_cell,
    const uint32_t *__restrict__ rank, const uint32_t *__restrict__ hist,
    const uint32_t *__restrict__ start, uint32_t ncell, uint32_t ny,
    uint32_t nz, double h, uint32_t *__restrict__ sorted_idx,
    uint32_t *__restrict__ cell_sorted, double4 *__restrict__ atoms64,
    float4 *__restrict__ atoms32, float far_d, uint32_t i, uint32_t atom0,
    const uint32_t *__restrict__ ag_start = nullptr,
    uint2 *__restrict__ ag = nullptr) {
  if (i >= natom) return;
  const uint32_t c = atom_cell[i];
  const uint32_t ch = i / kBinChunk;
  const uint32_t r = hist[static_cast<size_t>(ch) * ncell + c] + rank[i];
  const uint32_t dst = start[c] + r;
  sorted_idx[dst] = atom0 + i;
  if (ag && r % kAgAtoms == 0) {  // the first atom of its group
    const uint32_t n = start[c + 1] - start[c] - r;
    ag[ag_start[c] + r / kAgAtoms] =
        make_uint2(dst, n < kAgAtoms ? n : kAgAtoms);
  }
  const double x = rx[i], y = ry[i], z = rz[i], q = qc[i];
  atoms64[dst] = make_double4(x, y, z, q);
  uint32_t far = 0u;
  if (MIXED) {
    // Coordinates relative to the cell centre, rounded once from FP64: the
    // large part of every phase k.r is carried by the FP64 cell centre.
    const uint32_t ix = c / (ny * nz), iy = (c / nz) % ny, iz = c % nz;
    const double cx = (static_cast<double>(ix) + 0.5) * h;
    const double cy = (static_cast<double>(iy) + 0.5) * h;
    const double cz = (static_cast<double>(iz) + 0.5) * h;
    const double dx = x - cx, dy = y - cy, dz = z - cz;
    atoms32[dst] = make_float4(static_cast<float>(dx), static_cast<float>(dy),
                               static_cast<float>(dz), static_cast<float>(q));
    // (a NaN offset is not marked: its phases are NaN either way)
    far = fabs(dx) + fabs(dy) + fabs(dz) > static_cast<double>(far_d)
              ? kFarAtom
              : 0u;
  }
  cell_sorted[dst] = c | far;
}

template <bool MIXED>
__global__ void bin_scatter_kernel(
    const double *__restrict__ rx, const double *__restrict__ ry,
    const double *__restrict__ rz, const double *__restrict__ qc,
    uint32_t natom, const uint32_t *__restrict__ atom_cell,
    const uint32_t *__restrict__ rank, const uint32_t *__restrict__ hist,
    const uint32_t *__restrict__ start, uint32_t ncell, uint32_t ny,
    uint32_t nz, double h, uint32_t *__restrict__ sorted_idx,
    uint32_t *__restrict__ cell_sorted, double4 *__restrict__ atoms64,
    float4 *__restrict__ atoms32, float far_d,
    const uint32_t *__restrict__ ag_start, uint2 *__restrict__ ag) {
  bin_scatter_body<MIXED>(rx, ry, rz, qc, natom, atom_cell, rank, hist, start,
                          ncell, ny, nz, h, sorted_idx, cell_sorted, atoms64,
                          atoms32, far_d,
                          blockIdx.x * blockDim.x + threadIdx.x, 0u, ag_start,
                          ag);
}

// ---------------------------------------------------------------------------
// structure factors

constexpr uint32_t kSfBlock = 256;

// The FP64 structure-factor kernel walks one z-column of cells per lane and
// stores the running sum along z, i.e. the first pass of the summed-volume
// table of src/glst_force.cu:702-724, so the shell-sum kernel starts at the y
// pass.  S is stored in node runs (sf_slot).

// FP64: S_c(k) = sum_j q_j (cos(k.r_j) - i sin(k.r_j)), absolute coordinates
// (src/glst_force.cu:601-664).  blockIdx.x = z-column, blockIdx.y = node tile.
__device__ __forceinline__ void sf_f64_body(
    const double4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const double4 *__restrict__ k64, uint32_t nz, uint32_t ncell,
    double2 *__restrict__ S, uint32_t col, uint32_t node) {
  const double4 k = k64[node];
  double pre = 0.0, pim = 0.0;  // running z-prefix
  for (uint32_t z = 0; z < nz; ++z) {
    const uint32_t cell = col * nz + z;
    const uint32_t a0 = start[cell], a1 = start[cell + 1];
    double re0 = 0.0, im0 = 0.0, re1 = 0.0, im1 = 0.0;
    uint32_t j = a0;
    for (; j + 2 <= a1; j += 2) {
      const double4 a = atoms[j];
      const double4 b = atoms[j + 1];
      double sa, ca, sb, cb;
      sincos_cw(k.x * a.x + k.y * a.y + k.z * a.z, &sa, &ca);
      sincos_cw(k.x * b.x + k.y * b.y + k.z * b.z, &sb, &cb);
      re0 += a.w * ca;
      im0 -= a.w * sa;
      re1 += b.w * cb;
      im1 -= b.w * sb;
    }
    if (j < a1) {
      const double4 a = atoms[j];
      double sa, ca;
      sincos_cw(k.x * a.x + k.y * a.y + k.z * a.z, &sa, &ca);
      re0 += a.w * ca;
      im0 -= a.w * sa;
    }
    pre += re0 + re1;
    pim += im0 + im1;
    S[sf_slot<false>(node, cell, ncell)] = make_double2(pre, pim);
  }
}

__global__ __launch_bounds__(kSfBlock) void sf_f64_kernel(
    const double4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const double4 *__restrict__ k64, uint32_t nz, uint32_t ncell,
    double2 *__restrict__ S) {
  sf_f64_body(atoms, start, k64, nz, ncell, S, blockIdx.x,
              blockIdx.y * kSfBlock + threadIdx.x);
}


// ---------------------------------------------------------------------------
// shell sums

struct GlobalGrid {
  const double2 *p;
  uint32_t ld, node;
  __device__ double2 at(uint32_t c) const {
    return p[static_cast<size_t>(c) * ld + node];
  }
};

// Inclusion-exclusion on the summed-volume table P over the inclusive box
// [x0,x1]x[y0,y1]x[z0,z1] (src/glst_force.cu:774-834, same term order).
template <class G>
__device__ inline double2 box_sum(const G &P, uint32_t x0, uint32_t y0,
                                  uint32_t z0, uint32_t x1, uint32_t y1,
                                  uint32_t z1, uint32_t ny, uint32_t nz) {
  const bool xb = x0 == 0, yb = y0 == 0, zb = z0 == 0;
  const uint32_t xm = xb ? 0 : x0 - 1, ym = yb ? 0 : y0 - 1,
                 zm = zb ? 0 : z0 - 1;
  auto id = [&](uint32_t x, uint32_t y, uint32_t z) {
    return (x * ny + y) * nz + z;
  };
  const double2 zero = make_double2(0.0, 0.0);
  const double2 g0 = P.at(id(x1, y1, z1));
  const double2 g1 = xb ? zero : P.at(id(xm, y1, z1));
  const double2 g2 = yb ? zero : P.at(id(x1, ym, z1));
  const double2 g3 = zb ? zero : P.at(id(x1, y1, zm));
  const double2 g4 = (xb || yb) ? zero : P.at(id(xm, ym, z1));
  const double2 g5 = (xb || zb) ? zero : P.at(id(xm, y1, zm));
  const double2 g6 = (yb || zb) ? zero : P.at(id(x1, ym, zm));
  const double2 g7 = (xb || yb || zb) ? zero : P.at(id(xm, ym, zm));
  return make_double2(g0.x - g1.x - g2.x - g3.x + g4.x + g5.x + g6.x - g7.x,
                      g0.y - g1.y - g2.y - g3.y + g4.y + g5.y + g6.y - g7.y);
}

// Cube of Chebyshev radius r around (x,y,z), clamped to the grid
// (src/glst_force.cu:836-859).
template <class G>
__device__ inline double2 cube_sum(const G &P, uint32_t x, uint32_t y,
                                   uint32_t z, uint32_t r, uint32_t nx,
                                   uint32_t ny, uint32_t nz) {
  const uint32_t x0 = x < r ? 0 : x - r, y0 = y < r ? 0 : y - r,
                 z0 = z < r ? 0 : z - r;
  uint32_t x1 = x + r, y1 = y + r, z1 = z + r;
  x1 = x1 >= nx ? nx - 1 : x1;
  y1 = y1 >= ny ? ny - 1 : y1;
  z1 = z1 >= nz ? nz - 1 : z1;
  return box_sum(P, x0, y0, z0, x1, y1, z1, ny, nz);
}

// arg(x + i y) / 2 pi in revolutions, in [-1/2, 1/2]; (0, 0) gives 0.  With
// t = min(|x|, |y|) / max(|x|, |y|) in [0, 1], atan(t) / 2 pi = t P(t^2) for a
// degree-7 polynomial P fitted here (least squares on Chebyshev points with
// Lawson reweighting toward the minimax error); the octant then follows from
// |y| > |x| and the signs.  Emulated in FP32 with an exact reciprocal, the
// error stays below 4.3e-8 revolutions (2.7e-7 rad) against atan2 in FP64
// over 2e6 random (x, y); V_RCP_F32 adds at most one ulp of t.
__device__ inline float arg_rev(float x, float y) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  const float t = mx > 0.f ? mn * __builtin_amdgcn_rcpf(mx) : 0.f;
  const float s = t * t;
  float p = -0.0006453040987253189f;
  p = fmaf(p, s, 0.0034795973915606737f);
  p = fmaf(p, s, -0.008898723870515823f);
  p = fmaf(p, s, 0.015346036292612553f);
  p = fmaf(p, s, -0.022136272862553596f);
  p = fmaf(p, s, 0.031745944172143936f);
  p = fmaf(p, s, -0.05304612219333649f);
  p = fmaf(p, s, 0.15915483236312866f);
  float r = p * t;  // [0, 1/8]
  r = ay > ax ? 0.25f - r : r;
  r = x < 0.f ? 0.5f - r : r;
  return y < 0.f ? -r : r;
}

// Mixed-path remote sum of one (cell, node), stored in polar form: for
// V = exp(+i k.c) R_c(k) (FP64 in) the pair (kappa |V|, arg V / 2 pi) in
// FP32, kappa = |k| / 2 pi in revolutions per Angstrom.  The back-projection
// then forms Re and kappa Im of exp(i k.d) V as |V| cos(t) and kappa |V|
// sin(t) with t = kappa p + arg V / 2 pi: one FMA per pair for the phase and
// none for a complex product.
__device__ inline void store_rmt_mixed(void *R, size_t off, double re,
                                       double im, float kappa_rev) {
  const float fr = static_cast<float>(re), fi = static_cast<float>(im);
  const float mag = __builtin_amdgcn_sqrtf(fmaf(fr, fr, fi * fi));
  static_cast<float2 *>(R)[off] =
      make_float2(kappa_rev * mag, arg_rev(fr, fi));
}

// |k| / 2 pi of a node (revolutions per Angstrom).
__device__ inline float kappa_rev(const double4 &k) {
  constexpr double kInv2Pi = 0.15915494309189533577;
  return static_cast<float>(sqrt(k.x * k.x + k.y * k.y + k.z * k.z) * kInv2Pi);
}

// R_c(k) = W_k (cube_out - cube_in); the mixed path multiplies by
// exp(+i k.c) = a b c (the cell's three FP64 axis factors) and stores FP32,
// the FP64 path stores R_c(k).
template <bool MIXED, class G>
__device__ inline void shell_store(const G &P, uint32_t x, uint32_t y,
                                   uint32_t z, uint32_t node, const double4 &k,
                                   uint32_t rin, uint32_t rout, uint32_t nx,
                                   uint32_t ny, uint32_t nz, double2 a,
                                   double2 b, double2 c, uint32_t Kp,
                                   void *R) {
  const double2 o = cube_sum(P, x, y, z, rout, nx, ny, nz);
  const double2 in = cube_sum(P, x, y, z, rin, nx, ny, nz);
  const double re = (o.x - in.x) * k.w, im = (o.y - in.y) * k.w;
  const size_t off = static_cast<size_t>((x * ny + y) * nz + z) * Kp + node;
  if (MIXED) {
    const double pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
    const double cr = pr * c.x - pi * c.y, ci = pr * c.y + pi * c.x;
    store_rmt_mixed(R, off, re * cr - im * ci, re * ci + im * cr,
                    kappa_rev(k));
  } else {
    static_cast<double2 *>(R)[off] = make_double2(re, im);
  }
}

// exp(+i k_d (i + 1/2) h) for i < n (one axis of the cell-centre phase).
__device__ inline void axis_phase(double2 *dst, double kd, double h,
                                  uint32_t i) {
  double s, c;
  sincos(kd * ((static_cast<double>(i) + 0.5) * h), &s, &c);
  *dst = make_double2(c, s);
}

// ---- LDS-resident tables (grids whose table of one node fits in LDS) ----

// threads of a per-tile shell-sum workgroup (the batched engine): 256, with
// tiles sized for a quarter of the LDS, lets four workgroups share a CU
#ifndef GLST_RMT_BLOCK
#define GLST_RMT_BLOCK 256
#endif
constexpr uint32_t kRmtBlock = GLST_RMT_BLOCK;
// Values held in registers per thread while a line of the table is scanned
// (8: 175 VGPRs for the persistent kernel against 204 with 16, and 2 %
// faster alone)
#ifndef GLST_RMT_ZBATCH
#define GLST_RMT_ZBATCH 8
#endif
constexpr uint32_t kRmtBatch = GLST_RMT_ZBATCH;
// Threads of the persistent single-system workgroups.
#ifndef GLST_RMT_PBLOCK
#define GLST_RMT_PBLOCK 512
#endif
// waves per SIMD the persistent kernel's register budget must allow (2 is
// what one 512-thread workgroup per CU needs; more leaves registers for the
// kernels that run beside it)
#ifndef GLST_RMT_MIN_WAVES
#define GLST_RMT_MIN_WAVES 1
#endif
constexpr uint32_t kRmtPBlock = GLST_RMT_PBLOCK;

// n / d for n < 2^22 and 0 < d, with rd = 1.0f / d: the float quotient is
// within one of the exact one, and one correction step fixes it.
__device__ inline uint32_t fast_div(uint32_t n, uint32_t d, float rd) {
  uint32_t q = static_cast<uint32_t>(static_cast<float>(n) * rd);
  const int32_t r = static_cast<int32_t>(n - q * d);
  q = r < 0 ? q - 1 : (r >= static_cast<int32_t>(d) ? q + 1 : q);
  return q;
}

// Inclusive running sum along one line of the LDS table (n elements, stride
// elements apart), in batches of kRmtBatch.
__device__ inline void lds_scan(double2 *p, uint32_t n, uint32_t stride) {
  double2 run = make_double2(0.0, 0.0);
  for (uint32_t i0 = 0; i0 < n; i0 += kRmtBatch) {
    const uint32_t nb = (n - i0 < kRmtBatch) ? n - i0 : kRmtBatch;
    double2 v[kRmtBatch];
#pragma unroll
    for (uint32_t u = 0; u < kRmtBatch; ++u)  // clamped: always in the line
      v[u] = p[(i0 + (u < nb ? u : nb - 1)) * stride];
#pragma unroll
    for (uint32_t u = 0; u < kRmtBatch; ++u)
      if (u < nb) {
        run.x += v[u].x;
        run.y += v[u].y;
        p[(i0 + u) * stride] = run;
      }
  }
}

// Bordered tables.  One node's summed-volume table in LDS carries a zero
// plane before the first cell of every axis: cell (x, y, z) sits at
// ((x+1) (ny+1) + (y+1)) Zp + (z+1), Zp = rmt_zpitch(nz) (odd, so that
// neighbouring z-lines fall on different banks).  The inclusion-exclusion
// corner before x0 of a box starting at x0 = 0 then reads the zero plane
// instead of being selected away: every corner is one LDS read, unmasked,
// and the sum keeps box_sum's term order (subtracting an exact zero).

// Table offsets of a cube of Chebyshev radius r around (x, ., z), clamped to
// the grid: x planes and z indices of its upper corner (bx1, z1) and of the
// planes just before it (bxm, zm).
struct CubeXZ {
  uint32_t bx1, bxm, z1, zm;
};

__device__ inline CubeXZ cube_xz(uint32_t x, uint32_t z, uint32_t r,
                                 uint32_t nx, uint32_t nz, uint32_t plane) {
  CubeXZ c;
  c.bx1 = ((x + r >= nx) ? nx : x + r + 1) * plane;
  c.bxm = (x < r ? 0 : x - r) * plane;
  c.z1 = (z + r >= nz) ? nz : z + r + 1;
  c.zm = z < r ? 0 : z - r;
  return c;
}

// The 8 corners of a cube at one row from the bordered table P: rows o1
// (last row of the cube + 1, times zp) and om (first row, times zp).
struct CubeCorners {
  double2 v[8];
};
__device__ __forceinline__ CubeCorners cube_corners(const double2 *P,
                                                    const CubeXZ &c,
                                                    uint32_t o1, uint32_t om) {
  CubeCorners k;
  k.v[0] = P[c.bx1 + o1 + c.z1];  // (x1, y1, z1)
  k.v[1] = P[c.bxm + o1 + c.z1];  // (x0-1, y1, z1)
  k.v[2] = P[c.bx1 + om + c.z1];  // (x1, y0-1, z1)
  k.v[3] = P[c.bx1 + o1 + c.zm];  // (x1, y1, z0-1)
  k.v[4] = P[c.bxm + om + c.z1];  // (x0-1, y0-1, z1)
  k.v[5] = P[c.bxm + o1 + c.zm];  // (x0-1, y1, z0-1)
  k.v[6] = P[c.bx1 + om + c.zm];  // (x1, y0-1, z0-1)
  k.v[7] = P[c.bxm + om + c.zm];  // (x0-1, y0-1, z0-1)
  return k;
}
// inclusion-exclusion in box_sum's term order
__device__ __forceinline__ double2 cube_sum(const CubeCorners &k) {
  const double2 *r = k.v;
  return make_double2(
      r[0].x - r[1].x - r[2].x - r[3].x + r[4].x + r[5].x + r[6].x - r[7].x,
      r[0].y - r[1].y - r[2].y - r[3].y + r[4].y + r[5].y + r[6].y - r[7].y);
}

// Geometry of a workgroup's bordered tables (rmt_table_elems): T tables of
// ts elements, then the phase factors (T * nax per buffer).  Lanes of the
// passes run over the nodes first, then over a line coordinate padded to
// ypad / zpad (multiples of rmt_line_group(T)), which with rmt_table_stride
// keeps every LDS cycle free of bank conflicts.
struct RmtGeom {
  uint32_t T, lgT, nx, ny, nz, ncell, nax, zp, plane, ts, ypad, zpad;
  float rny, rypad, rzpad;
  __device__ RmtGeom(uint32_t T_, uint32_t nx_, uint32_t ny_, uint32_t nz_)
      : T(T_), lgT(31u - __builtin_clz(T_)), nx(nx_), ny(ny_), nz(nz_),
        ncell(nx_ * ny_ * nz_), nax(nx_ + ny_ + nz_), zp(rmt_zpitch(nz_)),
        plane((ny_ + 1) * rmt_zpitch(nz_)),
        ts(rmt_table_stride(rmt_table_elems(nx_, ny_, nz_), T_)) {
    const uint32_t lg = rmt_line_group(T);
    ypad = (ny + lg - 1) / lg * lg;
    zpad = (nz + lg - 1) / lg * lg;
    rny = 1.0f / static_cast<float>(ny);
    rypad = 1.0f / static_cast<float>(ypad);
    rzpad = 1.0f / static_cast<float>(zpad);
  }
};

// Axis factor e (< T nax) of nodes node0..: exp(+i k_d (i + 1/2) h).
__device__ inline double2 rmt_phase(const double4 *__restrict__ k64,
                                    const RmtGeom &g, uint32_t node0, double h,
                                    uint32_t e) {
  const uint32_t t = e / g.nax, a = e - t * g.nax;
  const double4 k = k64[node0 + t];
  double2 r;
  if (a < g.nx)
    axis_phase(&r, k.x, h, a);
  else if (a < g.nx + g.ny)
    axis_phase(&r, k.y, h, a - g.nx);
  else
    axis_phase(&r, k.z, h, a - g.nx - g.ny);
  return r;
}

// Zero planes before each axis of every table (the passes touch interior
// cells only, so they stay zero across tiles).
template <uint32_t BS>
__device__ inline void rmt_border(double2 *P, const RmtGeom &g) {
  const uint32_t nborder = g.plane + g.nx * g.zp + g.nx * g.ny;
  for (uint32_t e = threadIdx.x; e < g.T * nborder; e += BS) {
    const uint32_t t = e / nborder;
    uint32_t b = e - t * nborder, idx;
    if (b < g.plane) {
      idx = b;  // the x plane before the grid
    } else if ((b -= g.plane) < g.nx * g.zp) {
      idx = (1 + b / g.zp) * g.plane + b % g.zp;  // the y row before the grid
    } else {
      b -= g.nx * g.zp;
      idx = (1 + b / g.ny) * g.plane + (1 + b % g.ny) * g.zp;  // z slot
    }
    P[t * g.ts + idx] = make_double2(0.0, 0.0);
  }
}

// Tables and phase factors of nodes node0.. are in LDS (after a barrier).
// (2) Mixed path, unless the tables already hold them (ZDONE): turns the
// cell-local sums into S_c = exp(-i k.c) local and takes the running sum
// along z (the FP64 slots hold z-running sums already).  (3) Running sums along y, then x.  `between` runs next (the
// persistent kernel fetches its next tile there).  (4) Shells: lanes on
// consecutive (node, z), every thread walking y; the mixed path applies
// w exp(+i k.c) with the node weight folded into the thread's x and z
// factors.
template <uint32_t BS, bool MIXED, bool ZDONE, class F>
__device__ __forceinline__ void rmt_compute(
    double2 *P, const double2 *phase, const RmtGeom &g,
    const double4 *__restrict__ k64, const uint32_t *__restrict__ group,
    const uint32_t *__restrict__ r_in, const uint32_t *__restrict__ r_out,
    uint32_t Kp, uint32_t node0, void *__restrict__ R, F between) {
  const uint32_t T = g.T, nx = g.nx, ny = g.ny, nz = g.nz, nax = g.nax;
  const uint32_t zp = g.zp, plane = g.plane, ts = g.ts, lgT = g.lgT;
  const uint32_t tid = threadIdx.x;
  if (MIXED && !ZDONE) {
    // (2) z-lines
    for (uint32_t l = tid; l < T * nx * g.ypad; l += BS) {
      const uint32_t t = l & (T - 1), q = l >> lgT;
      const uint32_t x = fast_div(q, g.ypad, g.rypad), y = q - x * g.ypad;
      if (y >= ny) continue;
      const double2 *ph = phase + t * nax;
      const double2 a = ph[x], b = ph[nx + y];
      const double pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
      const double2 *pz = ph + nx + ny;
      double2 *p = P + t * ts + (x + 1) * plane + (y + 1) * zp + 1;
      double2 run = make_double2(0.0, 0.0);
      for (uint32_t z0 = 0; z0 < nz; z0 += kRmtBatch) {
        const uint32_t nb = (nz - z0 < kRmtBatch) ? nz - z0 : kRmtBatch;
        double2 v[kRmtBatch];
#pragma unroll
        for (uint32_t u = 0; u < kRmtBatch; ++u)  // clamped: always in the line
          v[u] = p[z0 + (u < nb ? u : nb - 1)];
#pragma unroll
        for (uint32_t u = 0; u < kRmtBatch; ++u)
          if (u < nb) {
            const double2 c = pz[z0 + u];
            const double er = pr * c.x - pi * c.y, ei = pr * c.y + pi * c.x;
            // v * conj(e), accumulated along z
            run.x += v[u].x * er + v[u].y * ei;
            run.y += v[u].y * er - v[u].x * ei;
            p[z0 + u] = run;
          }
      }
    }
    __syncthreads();
  }
  // (3) inclusive prefix along y, then x (src/glst_force.cu:726-772)
  for (uint32_t l = tid; l < T * nx * g.zpad; l += BS) {
    const uint32_t t = l & (T - 1), q = l >> lgT;
    const uint32_t x = fast_div(q, g.zpad, g.rzpad), z = q - x * g.zpad;
    if (z < nz) lds_scan(P + t * ts + (x + 1) * plane + zp + (z + 1), ny, zp);
  }
  __syncthreads();
  for (uint32_t l = tid; l < T * ny * g.zpad; l += BS) {
    const uint32_t t = l & (T - 1), q = l >> lgT;
    const uint32_t y = fast_div(q, g.zpad, g.rzpad), z = q - y * g.zpad;
    if (z < nz) lds_scan(P + t * ts + plane + (y + 1) * zp + (z + 1), nx, plane);
  }
  __syncthreads();
  between();
  // (4) shells; y is cut into segments so that the workgroup stays busy
  const uint32_t glines = T * nx * g.zpad;
  uint32_t nseg = BS / glines;
  nseg = nseg < 1 ? 1 : (nseg > ny ? ny : nseg);
  const uint32_t ylen = (ny + nseg - 1) / nseg;
  for (uint32_t task = tid; task < glines * nseg; task += BS) {
    const uint32_t t = task & (T - 1);
    const uint32_t q = task >> lgT;
    const uint32_t xs = fast_div(q, g.zpad, g.rzpad), z = q - xs * g.zpad;
    if (z >= nz) continue;
    const uint32_t x = xs % nx, seg = xs / nx;
    const uint32_t node = node0 + t;
    const uint32_t gr = group[node];
    const uint32_t rin = r_in[gr], rout = r_out[gr];
    const double4 kn = k64[node];
    const double w = kn.w;
    const float kr = MIXED ? kappa_rev(kn) : 0.f;
    const double2 *Pt = P + t * ts;
    const CubeXZ co = cube_xz(x, z, rout, nx, nz, plane);
    const CubeXZ ci = cube_xz(x, z, rin, nx, nz, plane);
    const double2 *ph = phase + t * nax;
    double qr = 0.0, qi = 0.0;
    if (MIXED) {  // w exp(+i k_x c_x) exp(+i k_z c_z)
      const double2 a = ph[x], c = ph[nx + ny + z];
      qr = (a.x * c.x - a.y * c.y) * w;
      qi = (a.x * c.y + a.y * c.x) * w;
    }
    const uint32_t yb = seg * ylen;
    const uint32_t ye = (yb + ylen < ny) ? yb + ylen : ny;
    // row offsets of both cubes, advanced with y: the upper row is
    // min(y + r + 1, ny), the lower one max(y - r, 0)
    uint32_t o1o = ((yb + rout >= ny) ? ny : yb + rout + 1) * zp;
    uint32_t omo = (yb < rout ? 0 : yb - rout) * zp;
    uint32_t o1i = ((yb + rin >= ny) ? ny : yb + rin + 1) * zp;
    uint32_t omi = (yb < rin ? 0 : yb - rin) * zp;
    for (uint32_t y = yb; y < ye; ++y) {
      // all 16 corners are read before any is used
      const CubeCorners ko = cube_corners(Pt, co, o1o, omo);
      const CubeCorners ki = cube_corners(Pt, ci, o1i, omi);
      o1o += (y + rout + 1 < ny) ? zp : 0u;
      omo += (y >= rout) ? zp : 0u;
      o1i += (y + rin + 1 < ny) ? zp : 0u;
      omi += (y >= rin) ? zp : 0u;
      const double2 o = cube_sum(ko), in = cube_sum(ki);
      const double dr = o.x - in.x, di = o.y - in.y;
      const size_t off = static_cast<size_t>((x * ny + y) * nz + z) * Kp + node;
      if (MIXED) {
        const double2 b = ph[nx + y];
        const double er = qr * b.x - qi * b.y, ei = qr * b.y + qi * b.x;
        store_rmt_mixed(R, off, dr * er - di * ei, dr * ei + di * er, kr);
      } else {
        static_cast<double2 *>(R)[off] = make_double2(dr * w, di * w);
      }
    }
  }
}

// z-lines loaded straight into the tables.  Line l of a tile is node
// t = l & (T-1) and column q = l >> lg T (x = q / ny, y = q - x ny); its
// slots sf_slot(node0 + t, q nz + z) are sf_run apart.  The first LEN of them
// travel through registers (the persistent kernel fetches them while the
// previous tile's shells are gathered), the rest of a longer line is read
// when the line is stored.  Storing writes the running sums along z into
// the node's table; the mixed path first turns the cell-local sums into
// S_c = exp(-i k.c) local in FP64 (the FP64 slots hold z-running sums).
template <bool MIXED, uint32_t LEN>
struct RmtLine;
template <uint32_t LEN>
struct RmtLine<true, LEN> {
  static constexpr uint32_t kLen = LEN;
  float2 v[LEN];
};
template <uint32_t LEN>
struct RmtLine<false, LEN> {
  static constexpr uint32_t kLen = LEN;
  double re[LEN], im[LEN];  // (a double2 array stays in scratch memory)
};
// line lengths: the persistent kernel's prefetch, the batched tiles' loads
template <bool MIXED>
using RmtLineP = RmtLine<MIXED, MIXED ? 24u : 16u>;
template <bool MIXED>
using RmtLineB = RmtLine<MIXED, 8u>;

template <bool MIXED>
__device__ __forceinline__ double2 rmt_slot(const void *__restrict__ Sv,
                                            size_t i) {
  if constexpr (MIXED) {
    const float2 f = static_cast<const float2 *>(Sv)[i];
    return make_double2(f.x, f.y);
  } else {
    return static_cast<const double2 *>(Sv)[i];
  }
}

template <bool MIXED, uint32_t LEN>
__device__ __forceinline__ void rmt_line_fetch(const void *__restrict__ Sv,
                                               const RmtGeom &g,
                                               uint32_t node0, uint32_t l,
                                               RmtLine<MIXED, LEN> &ln) {
  constexpr uint32_t L = LEN, G = sf_run<MIXED>();
  const uint32_t t = l & (g.T - 1), q = l >> g.lgT;
  const size_t base = sf_slot<MIXED>(node0 + t, q * g.nz, g.ncell);
#pragma unroll
  for (uint32_t z = 0; z < L; ++z) {
    const uint32_t zz = z < g.nz ? z : g.nz - 1;  // clamped: in the line
    if constexpr (MIXED) {
      ln.v[z] = static_cast<const float2 *>(Sv)[base + G * zz];
    } else {
      const double2 v = static_cast<const double2 *>(Sv)[base + G * zz];
      ln.re[z] = v.x;
      ln.im[z] = v.y;
    }
  }
}

template <bool MIXED, uint32_t LEN>
__device__ __forceinline__ void rmt_line_put(double2 *P,
                                             const double2 *phase,
                                             const RmtGeom &g,
                                             const void *__restrict__ Sv,
                                             uint32_t node0, uint32_t l,
                                             const RmtLine<MIXED, LEN> &ln) {
  constexpr uint32_t L = LEN, G = sf_run<MIXED>();
  const uint32_t t = l & (g.T - 1), q = l >> g.lgT;
  const uint32_t x = fast_div(q, g.ny, g.rny), y = q - x * g.ny;
  double2 *p = P + t * g.ts + (x + 1) * g.plane + (y + 1) * g.zp + 1;
  const size_t base = sf_slot<MIXED>(node0 + t, q * g.nz, g.ncell);
  if constexpr (MIXED) {
    const double2 *ph = phase + t * g.nax;
    const double2 a = ph[x], b = ph[g.nx + y];
    const double pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
    const double2 *pz = ph + g.nx + g.ny;
    double2 run = make_double2(0.0, 0.0);
    auto step = [&](uint32_t z, double2 v) {
      const double2 c = pz[z];
      const double er = pr * c.x - pi * c.y, ei = pr * c.y + pi * c.x;
      // v * conj(e), accumulated along z
      run.x += v.x * er + v.y * ei;
      run.y += v.y * er - v.x * ei;
      p[z] = run;
    };
#pragma unroll
    for (uint32_t z = 0; z < L; ++z)
      if (z < g.nz) step(z, make_double2(ln.v[z].x, ln.v[z].y));
    for (uint32_t z = L; z < g.nz; ++z)
      step(z, rmt_slot<true>(Sv, base + G * z));
  } else {
    // past the line's end the fetch repeated its last slot, so the clamped
    // store rewrites that slot's value (a loop bounded by nz instead would
    // keep the line in scratch memory)
#pragma unroll
    for (uint32_t z = 0; z < L; ++z)
      p[z < g.nz ? z : g.nz - 1] = make_double2(ln.re[z], ln.im[z]);
    for (uint32_t z = L; z < g.nz; ++z) p[z] = rmt_slot<false>(Sv, base + G * z);
  }
}

// Shell sums of one tile of T consecutive nodes by one workgroup (the
// batched engine: every workgroup's tile may belong to another grid).
template <bool MIXED>
__device__ __forceinline__ void rmt_lds_body(
    const void *__restrict__ Sv, const double4 *__restrict__ k64,
    const uint32_t *__restrict__ group, const uint32_t *__restrict__ r_in,
    const uint32_t *__restrict__ r_out, uint32_t Kp, uint32_t T,
    uint32_t nx, uint32_t ny, uint32_t nz, double h, void *__restrict__ R,
    uint32_t tile) {
  extern __shared__ double2 P[];  // [T][ts], then [T][nx+ny+nz] phases
  const RmtGeom g(T, nx, ny, nz);
  double2 *phase = P + T * g.ts;
  const uint32_t node0 = tile * T;
  if (MIXED)
    for (uint32_t e = threadIdx.x; e < T * g.nax; e += kRmtBlock)
      phase[e] = rmt_phase(k64, g, node0, h, e);
  rmt_border<kRmtBlock>(P, g);
  __syncthreads();  // the phase factors
  const uint32_t nline = T * nx * ny;
  for (uint32_t l = threadIdx.x; l < nline; l += kRmtBlock) {
    RmtLineB<MIXED> ln;
    rmt_line_fetch(Sv, g, node0, l, ln);
    rmt_line_put(P, phase, g, Sv, node0, l, ln);
  }
  __syncthreads();
  rmt_compute<kRmtBlock, MIXED, true>(P, phase, g, k64, group, r_in, r_out,
                                      Kp, node0, R, [] {});
}

// Shell sums of tiles tile0 .. tile0+ntile-1 (T nodes each) by persistent
// workgroups, one or two per CU.  With at least 8 workgroups, XCD x (the
// workgroups b with b % 8 == x) owns a contiguous eighth of the tiles and
// its workgroups take every (count on x)-th tile of it, so the tiles in
// flight on one XCD are consecutive and their R rows meet in its L2.  While
// a workgroup gathers one tile's shells, the next tile's z-lines are already
// on their way into registers and its phase factors are written to the
// second of two phase buffers.
template <bool MIXED>
__global__ __launch_bounds__(kRmtPBlock, GLST_RMT_MIN_WAVES) void rmt_lds_kernel(
    const void *__restrict__ Sv, const double4 *__restrict__ k64,
    const uint32_t *__restrict__ group, const uint32_t *__restrict__ r_in,
    const uint32_t *__restrict__ r_out, uint32_t Kp, uint32_t T,
    uint32_t tile0, uint32_t ntile, uint32_t nx, uint32_t ny, uint32_t nz,
    double h, void *__restrict__ R) {
  // [T][ts] tables, then two buffers of [T][nx+ny+nz] phase factors
  extern __shared__ double2 P[];
  const RmtGeom g(T, nx, ny, nz);
  const uint32_t tid = threadIdx.x;
  const uint32_t nwg = gridDim.x;
  uint32_t tile, end, step;
  if (nwg >= 8) {
    const uint32_t x = blockIdx.x & 7u, q = ntile >> 3, r = ntile & 7u;
    const uint32_t begin = x * q + (x < r ? x : r);
    tile = begin + (blockIdx.x >> 3);
    end = begin + q + (x < r ? 1u : 0u);
    step = (nwg - x + 7) >> 3;
  } else {
    tile = blockIdx.x;
    end = ntile;
    step = nwg;
  }
  const uint32_t nph = MIXED ? T * g.nax : 0u;
  const uint32_t nline = T * nx * ny;
  double2 *phase0 = P + T * g.ts;
  rmt_border<kRmtPBlock>(P, g);
  RmtLineP<MIXED> ln;
  if (tile < end) {
    const uint32_t node0 = (tile0 + tile) * T;
    for (uint32_t e = tid; e < nph; e += kRmtPBlock)
      phase0[e] = rmt_phase(k64, g, node0, h, e);
    if (tid < nline) rmt_line_fetch(Sv, g, node0, tid, ln);
  }
  __syncthreads();  // the first tile's phase factors
  uint32_t buf = 0;
  while (tile < end) {  // uniform over the workgroup
    const uint32_t node0 = (tile0 + tile) * T;
    const double2 *phase = phase0 + buf * nph;
    if (tid < nline) rmt_line_put(P, phase, g, Sv, node0, tid, ln);
    for (uint32_t l = tid + kRmtPBlock; l < nline; l += kRmtPBlock) {
      RmtLineP<MIXED> more;
      rmt_line_fetch(Sv, g, node0, l, more);
      rmt_line_put(P, phase, g, Sv, node0, l, more);
    }
    __syncthreads();
    const uint32_t next = tile + step;
    rmt_compute<kRmtPBlock, MIXED, true>(
        P, phase, g, k64, group, r_in, r_out, Kp, node0, R, [&] {
          if (next < end) {
            const uint32_t n1 = (tile0 + next) * T;
            double2 *ph1 = phase0 + (buf ^ 1u) * nph;
            for (uint32_t e = tid; e < nph; e += kRmtPBlock)
              ph1[e] = rmt_phase(k64, g, n1, h, e);
            if (tid < nline) rmt_line_fetch(Sv, g, n1, tid, ln);
          }
        });
    __syncthreads();  // the tables are free, the next phases written
    tile = next;
    buf ^= 1u;
  }
}

// ---- shell sums on the matrix cores (mixed path, nx and ny up to 16) ----
//
// A cube sum is separable: B_r(c) sums Y = exp(-i k.c') S_c' over the cells
// c' with |x'-x|, |y'-y|, |z'-z| <= r.  Per node, the running sums of Y
// along z sit in LDS in FP32 with a zero before each line, so the z window
// of a cell is the difference of two of them.  For every z slice the x and
// y windows of both cubes are products with 0/1 band matrices on the
// matrix cores (v_mfma_f32_16x16x4f32: exact products, FP32 sums).  Pass 1
// forms Q^T = T^T Kx (rows y, columns x) from the z-windowed slice T; its
// accumulator registers are, one for one, the B operands of pass 2,
// W^T = Ky Q^T, whose k index runs over y in the order the registers hold
// it.  Every sum is a direct sum of at most nz + 32 terms, so FP32 suffices
// where the summed-volume tables need FP64 for their differences of large
// running sums.
//
// Operand layouts of the 16x16x4 FP32 instruction for lane l (c = l % 16,
// g = l / 16): A[c][g], B[g][c], accumulator register v = D[4g + v][c].
typedef float rmt_f4 __attribute__((ext_vector_type(4)));

__device__ inline float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// NPW consecutive S slots of one cell (the nodes n0 .. n0+NPW-1 of one node
// run, n0 a multiple of NPW)
template <uint32_t NPW>
__device__ __forceinline__ void rmt_mfma_load(const float2 *__restrict__ p,
                                              float2 (&v)[NPW]) {
  if constexpr (NPW == 1) {
    v[0] = p[0];
  } else {
#pragma unroll
    for (uint32_t h = 0; h < NPW / 2; ++h) {
      const float4 a = reinterpret_cast<const float4 *>(p)[h];
      v[2 * h] = make_float2(a.x, a.y);
      v[2 * h + 1] = make_float2(a.z, a.w);
    }
  }
}

// One tile of the matrix-core shell sums: nodes n0 .. n0+NPW-1 (see
// rmt_mfma_kernel).
template <uint32_t NPW>
__device__ __forceinline__ void rmt_mfma_tile(
    const float2 *__restrict__ S, const double4 *__restrict__ k64,
    const uint32_t *__restrict__ group, const uint32_t *__restrict__ r_in,
    const uint32_t *__restrict__ r_out, uint32_t Kp, uint32_t n0,
    uint32_t nx, uint32_t ny, uint32_t nz, double h, float2 *__restrict__ R,
    float2 *__restrict__ rmt_mfma_lds) {
  const uint32_t tid = threadIdx.x;
  const uint32_t nax = nx + ny + nz, zp = rmt_zpitch(nz);
  const uint32_t nline = nx * ny, ts = nline * zp, ncell = nline * nz;
  float2 *tab = rmt_mfma_lds;
  float2 *ph = rmt_mfma_lds + NPW * ts;
  float2 *stage = ph + NPW * nax;  // line factors
  // axis factors exp(+i k_d (i + 1/2) h), FP64 arguments, FP32 values
  for (uint32_t e = tid; e < NPW * nax; e += kRmtMfmaBlock) {
    const uint32_t t = e / nax, a = e - t * nax;
    const double4 k = k64[n0 + t];
    double2 r;
    if (a < nx)
      axis_phase(&r, k.x, h, a);
    else if (a < nx + ny)
      axis_phase(&r, k.y, h, a - nx);
    else
      axis_phase(&r, k.z, h, a - nx - ny);
    ph[e] = make_float2(static_cast<float>(r.x), static_cast<float>(r.y));
  }
  __syncthreads();
  // line factors exp(+i (k_x c_x + k_y c_y)) and the zero before each line
  // (nline <= 256: one line per thread)
  if (tid < nline) {
    const uint32_t x = tid / ny, y = tid - x * ny;
#pragma unroll
    for (uint32_t t = 0; t < NPW; ++t) {
      stage[t * nline + tid] = cmulf(ph[t * nax + x], ph[t * nax + nx + y]);
      tab[t * ts + tid * zp] = make_float2(0.f, 0.f);
    }
  }
  __syncthreads();
  // (1a) Y = S conj(exp(i k.c)): sixteen lanes per line, consecutive lanes
  // on consecutive z, so a wave reads consecutive cells
  {
    // batches of 8 lines per thread, all 8 loads in flight at once
    constexpr uint32_t kB = 8, kLs = kRmtMfmaBlock / 16;
    const uint32_t zl = tid & 15u;
    for (uint32_t z = zl; z < nz; z += 16) {
      for (uint32_t lb = tid >> 4; lb < nline; lb += kB * kLs) {
        float2 v[kB][NPW];
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u) {
          const uint32_t l = lb + u * kLs < nline ? lb + u * kLs : lb;
          rmt_mfma_load<NPW>(S + sf_slot<true>(n0, l * nz + z, ncell), v[u]);
        }
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u) {
          const uint32_t l = lb + u * kLs;
          if (l < nline) {
#pragma unroll
            for (uint32_t t = 0; t < NPW; ++t) {
              const float2 e = cmulf(stage[t * nline + l],
                                     ph[t * nax + nx + ny + z]);
              tab[t * ts + l * zp + z + 1] =
                  make_float2(v[u][t].x * e.x + v[u][t].y * e.y,
                              v[u][t].y * e.x - v[u][t].x * e.y);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  // (1b) running sums along z, one thread per line and node
  if (tid < nline) {
#pragma unroll
    for (uint32_t t = 0; t < NPW; ++t) {
      float2 *p = tab + t * ts + tid * zp + 1;
      float2 run = make_float2(0.f, 0.f);
      for (uint32_t z0 = 0; z0 < nz; z0 += 8) {
        float2 a[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u)  // clamped: always in the line
          a[u] = p[z0 + u < nz ? z0 + u : nz - 1];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u)
          if (z0 + u < nz) {
            run.x += a[u].x;
            run.y += a[u].y;
            p[z0 + u] = run;
          }
      }
    }
  }
  __syncthreads();
  // (2) each wave takes every fourth z slice, for all NPW nodes
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t c = tid & 15u, g4 = (tid >> 4) & 3u;
  uint32_t rin[NPW], rout[NPW];
  float wk[NPW], kr[NPW];
  // band matrices: pass 1's B (row x = 4m + g4, column x' = c) and pass 2's
  // A (row y' = c, k slot g4 of instruction v: y = 4 g4 + v)
  float bxo[NPW][4], bxi[NPW][4], ayo[NPW][4], ayi[NPW][4];
#pragma unroll
  for (uint32_t t = 0; t < NPW; ++t) {
    const uint32_t gr = group[n0 + t];
    rin[t] = r_in[gr];
    rout[t] = r_out[gr];
    const double4 kn = k64[n0 + t];
    wk[t] = static_cast<float>(kn.w);
    kr[t] = kappa_rev(kn);
#pragma unroll
    for (uint32_t m = 0; m < 4; ++m) {
      const int dx = static_cast<int>(c) - static_cast<int>(4 * m + g4);
      const uint32_t ax = static_cast<uint32_t>(dx < 0 ? -dx : dx);
      bxo[t][m] = ax <= rout[t] ? 1.f : 0.f;
      bxi[t][m] = ax <= rin[t] ? 1.f : 0.f;
      const int dy = static_cast<int>(c) - static_cast<int>(4 * g4 + m);
      const uint32_t ay = static_cast<uint32_t>(dy < 0 ? -dy : dy);
      ayo[t][m] = ay <= rout[t] ? 1.f : 0.f;
      ayi[t][m] = ay <= rin[t] ? 1.f : 0.f;
    }
  }
  for (uint32_t z = wave; z < nz; z += kRmtMfmaBlock / 64) {
    float2 val[NPW][4];
#pragma unroll
    for (uint32_t t = 0; t < NPW; ++t) {
      const uint32_t hio = z + rout[t] + 1 < nz ? z + rout[t] + 1 : nz;
      const uint32_t loo = z > rout[t] ? z - rout[t] : 0u;
      const uint32_t hii = z + rin[t] + 1 < nz ? z + rin[t] + 1 : nz;
      const uint32_t loi = z > rin[t] ? z - rin[t] : 0u;
      // the z windows of this lane's cells (x = 4m + g4, y = c)
      float aor[4], aoi[4], air[4], aii[4];
#pragma unroll
      for (uint32_t m = 0; m < 4; ++m) {
        const uint32_t x = 4 * m + g4;
        float2 o = make_float2(0.f, 0.f), in = o;
        if (x < nx && c < ny) {
          const float2 *p = tab + t * ts + (x * ny + c) * zp;
          const float2 a = p[hio], b = p[loo], d = p[hii], e = p[loi];
          o = make_float2(a.x - b.x, a.y - b.y);
          in = make_float2(d.x - e.x, d.y - e.y);
        }
        aor[m] = o.x;
        aoi[m] = o.y;
        air[m] = in.x;
        aii[m] = in.y;
      }
      rmt_f4 qor = {0.f, 0.f, 0.f, 0.f}, qoi = qor, qir = qor, qii = qor;
#pragma unroll
      for (uint32_t m = 0; m < 4; ++m) {
        qor = __builtin_amdgcn_mfma_f32_16x16x4f32(aor[m], bxo[t][m], qor, 0,
                                                   0, 0);
        qoi = __builtin_amdgcn_mfma_f32_16x16x4f32(aoi[m], bxo[t][m], qoi, 0,
                                                   0, 0);
        qir = __builtin_amdgcn_mfma_f32_16x16x4f32(air[m], bxi[t][m], qir, 0,
                                                   0, 0);
        qii = __builtin_amdgcn_mfma_f32_16x16x4f32(aii[m], bxi[t][m], qii, 0,
                                                   0, 0);
      }
      rmt_f4 wor = {0.f, 0.f, 0.f, 0.f}, woi = wor, wir = wor, wii = wor;
#pragma unroll
      for (uint32_t v = 0; v < 4; ++v) {
        wor = __builtin_amdgcn_mfma_f32_16x16x4f32(ayo[t][v], qor[v], wor, 0,
                                                   0, 0);
        woi = __builtin_amdgcn_mfma_f32_16x16x4f32(ayo[t][v], qoi[v], woi, 0,
                                                   0, 0);
        wir = __builtin_amdgcn_mfma_f32_16x16x4f32(ayi[t][v], qir[v], wir, 0,
                                                   0, 0);
        wii = __builtin_amdgcn_mfma_f32_16x16x4f32(ayi[t][v], qii[v], wii, 0,
                                                   0, 0);
      }
      // register v of lane (c, g4): the cell (x' = c, y' = 4 g4 + v, z);
      // V = w exp(+i k.c) (outer - inner) in polar form
      const float2 *pt = ph + t * nax;
      const uint32_t cx = c < nx ? c : 0u;
      float2 e0 = cmulf(pt[cx], pt[nx + ny + z]);
      e0.x *= wk[t];
      e0.y *= wk[t];
#pragma unroll
      for (uint32_t v = 0; v < 4; ++v) {
        const uint32_t yo = 4 * g4 + v < ny ? 4 * g4 + v : 0u;
        const float dr = wor[v] - wir[v], di = woi[v] - wii[v];
        const float2 e = cmulf(e0, pt[nx + yo]);
        const float vr = dr * e.x - di * e.y, vi = dr * e.y + di * e.x;
        const float mag = __builtin_amdgcn_sqrtf(fmaf(vr, vr, vi * vi));
        val[t][v] = make_float2(kr[t] * mag, arg_rev(vr, vi));
      }
    }
    // each cell's NPW consecutive nodes with one store
    if (c < nx) {
#pragma unroll
      for (uint32_t v = 0; v < 4; ++v) {
        const uint32_t yo = 4 * g4 + v;
        if (yo < ny) {
          float2 *dst = R + static_cast<size_t>((c * ny + yo) * nz + z) * Kp + n0;
          if constexpr (NPW == 1) {
            dst[0] = val[0][v];
          } else {
#pragma unroll
            for (uint32_t hh = 0; hh < NPW / 2; ++hh)
              reinterpret_cast<float4 *>(dst)[hh] =
                  make_float4(val[2 * hh][v].x, val[2 * hh][v].y,
                              val[2 * hh + 1][v].x, val[2 * hh + 1][v].y);
          }
        }
      }
    }
  }
}

// Shell sums of tiles tile0 .. tile0+ntile-1 of NPW nodes, one workgroup of
// four waves per tile.  (1) Consecutive threads read consecutive cells' S
// slots of the tile's nodes and write Y into the tables, which then take
// their running sums along z, one thread per (node, line).  (2) Each wave
// takes every fourth z slice and sums it for all NPW nodes (independent
// chains of matrix instructions), so each lane holds a cell's NPW
// consecutive nodes and stores them at once (R is cell-major: a lane
// storing one node alone writes 8 bytes into a line of its own).  The tiles
// of XCD x (workgroups b with b % 8 == x) are a contiguous eighth of the
// range, so the tiles that share S node runs and R lines run on one XCD at
// once.  (Persistent workgroups, two per CU taking their XCD's tiles in
// rounds, ran 0.68 ms alone at 288k atoms against 0.72 ms, but the tuned
// step, where S(k) and the back-projection share the CUs, took 4.77 ms
// against 4.72.)
template <uint32_t NPW>
__global__ __launch_bounds__(kRmtMfmaBlock) void rmt_mfma_kernel(
    const float2 *__restrict__ S, const double4 *__restrict__ k64,
    const uint32_t *__restrict__ group, const uint32_t *__restrict__ r_in,
    const uint32_t *__restrict__ r_out, uint32_t Kp, uint32_t tile0,
    uint32_t ntile, uint32_t nx, uint32_t ny, uint32_t nz, double h,
    float2 *__restrict__ R) {
  static_assert(NPW == 1 || NPW == 2 || NPW == 4, "nodes per workgroup");
  // [NPW][nx ny][zp] tables, [NPW][nax] axis factors, [NPW][nx ny] line
  // factors
  extern __shared__ float2 rmt_mfma_lds[];
  uint32_t tile = blockIdx.x;
  if (gridDim.x >= 8) {
    const uint32_t x = blockIdx.x & 7u, q = ntile >> 3, r = ntile & 7u;
    tile = x * q + (x < r ? x : r) + (blockIdx.x >> 3);
  }
  rmt_mfma_tile<NPW>(S, k64, group, r_in, r_out, Kp, (tile0 + tile) * NPW, nx,
                     ny, nz, h, R, rmt_mfma_lds);
}

// ---- shell sums of small grids (mixed path, every dimension up to 8) ----
//
// The cube sums are separable, and on these grids a line has at most 8
// cells, so each axis is one pass over lines held in registers: a lane
// loads a line, takes its running sums (the cells past the grid hold zero,
// so the last running sum repeats to the 8th slot) and forms the windows of
// both cubes, |i' - i| <= r, as differences of two running sums.  The
// half-widths select compile-time windows in a switch, so every register
// index is a constant.  z: Y = S exp(-i k.c) from S; x, then y: the windows
// of the previous pass, the lines read across through LDS.  All sums are
// FP32 over at most 8 terms.
template <int W>
__device__ __forceinline__ void rmt_win8(const float2 (&P)[8],
                                         float2 (&out)[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int hi = i + W < 7 ? i + W : 7;
    const int lo = i - W - 1;
    out[i] = lo >= 0 ? make_float2(P[hi].x - P[lo].x, P[hi].y - P[lo].y)
                     : P[hi];
  }
}

__device__ __forceinline__ void rmt_window8(const float2 (&P)[8],
                                            float2 (&out)[8], uint32_t r) {
  switch (r < 7u ? r : 7u) {
    case 0: rmt_win8<0>(P, out); break;
    case 1: rmt_win8<1>(P, out); break;
    case 2: rmt_win8<2>(P, out); break;
    case 3: rmt_win8<3>(P, out); break;
    case 4: rmt_win8<4>(P, out); break;
    case 5: rmt_win8<5>(P, out); break;
    case 6: rmt_win8<6>(P, out); break;
    default: rmt_win8<7>(P, out); break;
  }
}

// Workgroup barrier that waits for this wave's LDS operations only, so that
// global loads issued before it stay in flight (__syncthreads also drains
// them).  The memory clobber keeps the compiler from moving memory accesses
// across it.
__device__ __forceinline__ void rmt_lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

__device__ __forceinline__ void rmt_prefix8(float2 (&v)[8]) {
#pragma unroll
  for (int i = 1; i < 8; ++i) {
    v[i].x += v[i - 1].x;
    v[i].y += v[i - 1].y;
  }
}

// One tile: nodes n0 .. n0 + kRmtWaveNodes - 1, M of them at a time, M <=
// min(16, 256 / L) for passes of at most L lines (6^3: 7 nodes of 36 lines,
// the last of the three groups two nodes; the threads of node slots past
// the tile compute on the tile's last node and store nothing).
// Thread t serves node n0 + step + t % M and line t / M: consecutive
// threads take consecutive nodes of one cell, so that each S load of a node
// run (four adjacent slots of a cell) and each R store (cell-major, the
// nodes of a cell adjacent) is one contiguous segment, and the window
// tables, [cell][M] in LDS, are read and written without bank conflicts.
__device__ __forceinline__ void rmt_wave_tile(
    const float2 *__restrict__ S, const double4 *__restrict__ k64,
    const uint32_t *__restrict__ group, const uint32_t *__restrict__ r_in,
    const uint32_t *__restrict__ r_out, uint32_t Kp, uint32_t n0,
    uint32_t nx, uint32_t ny, uint32_t nz, double h,
    float2 *__restrict__ R) {
  constexpr uint32_t kAx = 3 * kRmtWaveMaxDim;
  // M ncell <= 2048: ncell <= L^(3/2) and M <= 256 / L, so M ncell <=
  // 256 sqrt(L) <= 256 * 8
  constexpr uint32_t kWin = 2048;
  static_assert(kRmtWaveBlock == 256 && kRmtWaveNodes == 16,
                "thread map of the small-grid shell sums");
  __shared__ float2 win_o[kWin], win_i[kWin];
  __shared__ float2 axis[kRmtWaveNodes][kAx];
  // per node: r_out, r_in, weight, |k| / 2 pi (read from LDS in the loop, so
  // that no vector load there waits behind the S prefetch: the vector
  // memory counter completes in order)
  __shared__ uint4 npar[kRmtWaveNodes];
  const uint32_t tid = threadIdx.x;
  const uint32_t nax = nx + ny + nz;
  if (tid < kRmtWaveNodes) {
    const uint32_t gr = group[n0 + tid];
    const double4 kn = k64[n0 + tid];
    npar[tid] = make_uint4(r_out[gr], r_in[gr],
                           __float_as_uint(static_cast<float>(kn.w)),
                           __float_as_uint(kappa_rev(kn)));
  }
  // axis factors exp(+i k_d (i + 1/2) h): FP64 arguments, FP32 values
  for (uint32_t e = tid; e < kRmtWaveNodes * nax; e += kRmtWaveBlock) {
    const uint32_t t = e / nax, a = e - t * nax;
    const double4 k = k64[n0 + t];
    double2 r;
    if (a < nx)
      axis_phase(&r, k.x, h, a);
    else if (a < nx + ny)
      axis_phase(&r, k.y, h, a - nx);
    else
      axis_phase(&r, k.z, h, a - nx - ny);
    axis[t][a] = make_float2(static_cast<float>(r.x), static_cast<float>(r.y));
  }
  const uint32_t lxy = nx * ny, lyz = ny * nz, lxz = nx * nz;
  const uint32_t lmax = lxy > lyz ? (lxy > lxz ? lxy : lxz)
                                  : (lyz > lxz ? lyz : lxz);
  // the largest power of two that fits, or the largest count if that
  // takes fewer groups (6^3: 7, three groups against four): a power of two
  // keeps a group's S loads on whole node runs
  const uint32_t mmax = 256u / lmax < kRmtWaveNodes ? 256u / lmax
                                                    : kRmtWaveNodes;
  uint32_t mp = kRmtWaveNodes;
  while (mp > mmax) mp >>= 1;
  const uint32_t M = (kRmtWaveNodes + mmax - 1) / mmax <
                             (kRmtWaveNodes + mp - 1) / mp
                         ? mmax
                         : mp;
  const uint32_t q = tid / M, j = tid - q * M;
  const uint32_t ncell = nx * ny * nz;
  // this thread's z-line of S for the nodes of step s, in flight while the
  // previous step's x and y passes run; the indexes are clamped into the
  // grid, so all eight loads are unconditional and in flight together (the
  // extra values are not used)
  float2 sv[8];
  const uint32_t qf = q < lxy ? q : 0u;
  auto fetch = [&](uint32_t s) {
    const uint32_t nf =
        n0 + (s + j < kRmtWaveNodes ? s + j : kRmtWaveNodes - 1);
#pragma unroll
    for (uint32_t z = 0; z < 8; ++z)
      sv[z] = S[sf_slot<true>(nf, qf * nz + (z < nz ? z : nz - 1), ncell)];
  };
  fetch(0);
  __syncthreads();  // the axis factors and node parameters
  for (uint32_t step = 0; step < kRmtWaveNodes; step += M) {
    const bool live = step + j < kRmtWaveNodes;
    const uint32_t t = live ? step + j : kRmtWaveNodes - 1;
    const uint32_t nd = n0 + t;
    const uint4 par = npar[t];
    const uint32_t ro = par.x, ri = par.y;
    const float2 *pt = axis[t];
    // (z) line (x, y): Y = S conj(exp(i k.c)), windows along z
    if (q < lxy) {
      const uint32_t x = q / ny, y = q - x * ny;
      const float2 exy = cmulf(pt[x], pt[nx + y]);
      float2 v[8];
#pragma unroll
      for (uint32_t z = 0; z < 8; ++z) {
        v[z] = make_float2(0.f, 0.f);
        if (z < nz) {
          const float2 e = cmulf(exy, pt[nx + ny + z]);
          v[z] = make_float2(sv[z].x * e.x + sv[z].y * e.y,
                             sv[z].y * e.x - sv[z].x * e.y);
        }
      }
      if (step + M < kRmtWaveNodes) fetch(step + M);
      rmt_prefix8(v);
      float2 wo[8], wi[8];
      rmt_window8(v, wo, ro);
      rmt_window8(v, wi, ri);
#pragma unroll
      for (uint32_t z = 0; z < 8; ++z)
        if (z < nz) {
          win_o[(q * nz + z) * M + j] = wo[z];
          win_i[(q * nz + z) * M + j] = wi[z];
        }
    }
    rmt_lds_barrier();
    // (x) line (y, z): cells (x ny + y) nz + z
    if (q < lyz) {
      const uint32_t y = q / nz, z = q - y * nz;
      float2 vo[8], vi[8];
#pragma unroll
      for (uint32_t x = 0; x < 8; ++x) {
        vo[x] = vi[x] = make_float2(0.f, 0.f);
        if (x < nx) {
          vo[x] = win_o[((x * ny + y) * nz + z) * M + j];
          vi[x] = win_i[((x * ny + y) * nz + z) * M + j];
        }
      }
      rmt_prefix8(vo);
      rmt_prefix8(vi);
      float2 wo[8], wi[8];
      rmt_window8(vo, wo, ro);
      rmt_window8(vi, wi, ri);
#pragma unroll
      for (uint32_t x = 0; x < 8; ++x)
        if (x < nx) {
          win_o[((x * ny + y) * nz + z) * M + j] = wo[x];
          win_i[((x * ny + y) * nz + z) * M + j] = wi[x];
        }
    }
    rmt_lds_barrier();
    // (y) line (x, z): the two cubes, then V = w exp(+i k.c) (outer -
    // inner) in polar form
    if (q < lxz && live) {
      const uint32_t x = q / nz, z = q - x * nz;
      float2 vo[8], vi[8];
#pragma unroll
      for (uint32_t y = 0; y < 8; ++y) {
        vo[y] = vi[y] = make_float2(0.f, 0.f);
        if (y < ny) {
          vo[y] = win_o[((x * ny + y) * nz + z) * M + j];
          vi[y] = win_i[((x * ny + y) * nz + z) * M + j];
        }
      }
      rmt_prefix8(vo);
      rmt_prefix8(vi);
      float2 co[8], ci[8];
      rmt_window8(vo, co, ro);
      rmt_window8(vi, ci, ri);
      const float wk = __uint_as_float(par.z), kr = __uint_as_float(par.w);
      float2 e0 = cmulf(pt[x], pt[nx + ny + z]);
      e0.x *= wk;
      e0.y *= wk;
#pragma unroll
      for (uint32_t y = 0; y < 8; ++y)
        if (y < ny) {
          const float dr = co[y].x - ci[y].x, di = co[y].y - ci[y].y;
          const float2 e = cmulf(e0, pt[nx + y]);
          const float vr = dr * e.x - di * e.y, vi2 = dr * e.y + di * e.x;
          const float mag = __builtin_amdgcn_sqrtf(fmaf(vr, vr, vi2 * vi2));
          R[static_cast<size_t>((x * ny + y) * nz + z) * Kp + nd] =
              make_float2(kr * mag, arg_rev(vr, vi2));
        }
    }
    rmt_lds_barrier();  // the windows are free for the next nodes
  }
}

// Shell sums of tiles tile0 .. tile0 + ntile - 1 on a small grid, one
// workgroup per tile; the tiles of XCD x (workgroups b with b % 8 == x) are
// a contiguous eighth of the range.
// Workgroups per CU the register budget is held to.  1: the kernel takes
// the 150 VGPRs it needs (3 waves per SIMD, no scratch).  Held to 4 (128
// VGPRs, 12 bytes of scratch per lane) the 512-system batch ran 4 % faster,
// but a served stream of two concurrent slots then returned one answer of
// NaN forces, with no cause found; that build is not used.
#ifndef GLST_RMT_WAVE_MIN_WG
#define GLST_RMT_WAVE_MIN_WG 1
#endif
__global__ __launch_bounds__(kRmtWaveBlock, GLST_RMT_WAVE_MIN_WG) void rmt_wave_kernel(
    const float2 *__restrict__ S, const double4 *__restrict__ k64,
    const uint32_t *__restrict__ group, const uint32_t *__restrict__ r_in,
    const uint32_t *__restrict__ r_out, uint32_t Kp, uint32_t tile0,
    uint32_t ntile, uint32_t nx, uint32_t ny, uint32_t nz, double h,
    float2 *__restrict__ R) {
  uint32_t tile = blockIdx.x;
  if (gridDim.x >= 8) {
    const uint32_t x = blockIdx.x & 7u, q = ntile >> 3, r = ntile & 7u;
    tile = x * q + (x < r ? x : r) + (blockIdx.x >> 3);
  }
  rmt_wave_tile(S, k64, group, r_in, r_out, Kp,
                (tile0 + tile) * kRmtWaveNodes, nx, ny, nz, h, R);
}

// ---- global-memory tables (grids too large for LDS) ----

// G[c*Kp + k] = S (node runs) converted to FP64 cell-major order;
// the mixed path applies exp(-i k.c) to the cell-local sums and takes the
// running sum along z (the FP64 slots hold z-running sums already).  One
// thread per (node, z-column).
template <bool MIXED>
__global__ void rmt_global_load_kernel(const void *__restrict__ Sv,
                                       const double4 *__restrict__ k64,
                                       double2 *__restrict__ G, uint32_t Kp,
                                       uint32_t col0, uint32_t nx,
                                       uint32_t ny, uint32_t nz, double h) {
  const uint32_t node = blockIdx.x * 256 + threadIdx.x;
  const uint32_t col = col0 + blockIdx.y;
  if (node >= Kp || col >= nx * ny) return;
  const uint32_t ncell = nx * ny * nz;
  if (!MIXED) {
    for (uint32_t z = 0; z < nz; ++z) {
      const uint32_t cell = col * nz + z;
      G[static_cast<size_t>(cell) * Kp + node] =
          static_cast<const double2 *>(Sv)[sf_slot<false>(node, cell, ncell)];
    }
    return;
  }
  const uint32_t x = col / ny, y = col % ny;
  const double4 k = k64[node];
  double2 a, b;
  axis_phase(&a, k.x, h, x);
  axis_phase(&b, k.y, h, y);
  const double pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
  double2 run = make_double2(0.0, 0.0);
  for (uint32_t z = 0; z < nz; ++z) {
    const uint32_t cell = col * nz + z;
    double2 c;
    axis_phase(&c, k.z, h, z);
    const double er = pr * c.x - pi * c.y, ei = pr * c.y + pi * c.x;
    const float2 v = static_cast<const float2 *>(Sv)[sf_slot<true>(node, cell, ncell)];
    run.x += v.x * er + v.y * ei;
    run.y += v.y * er - v.x * ei;
    G[static_cast<size_t>(cell) * Kp + node] = run;
  }
}

__global__ void prefix_axis_kernel(double2 *__restrict__ S, uint32_t Kp,
                                   uint32_t line0, uint32_t nlines, uint32_t n,
                                   uint32_t stride, uint32_t inner) {
  // line l covers cells base(l) + i*stride, i < n; base(l) enumerates the
  // other two axes with `inner` consecutive cells per outer step.
  const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t l = line0 + blockIdx.y;
  if (node >= Kp || l >= nlines) return;
  const uint32_t base = (l / inner) * inner * n + (l % inner);
  double2 run = make_double2(0.0, 0.0);
  for (uint32_t i = 0; i < n; ++i) {
    const size_t o = static_cast<size_t>(base + i * stride) * Kp + node;
    run.x += S[o].x;
    run.y += S[o].y;
    S[o] = run;
  }
}

template <bool MIXED>
__global__ void rmt_gather_kernel(const double2 *__restrict__ P,
                                  const double4 *__restrict__ k64,
                                  const uint32_t *__restrict__ group,
                                  const uint32_t *__restrict__ r_in,
                                  const uint32_t *__restrict__ r_out,
                                  uint32_t Kp, uint32_t cell0, uint32_t nx,
                                  uint32_t ny, uint32_t nz, double h,
                                  void *__restrict__ R) {
  const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t cell = cell0 + blockIdx.y;
  if (node >= Kp || cell >= nx * ny * nz) return;
  const uint32_t g = group[node];
  const double4 k = k64[node];
  const uint32_t x = cell / (ny * nz), y = (cell / nz) % ny, z = cell % nz;
  double2 a = make_double2(1.0, 0.0), b = a, c = a;
  if (MIXED) {
    axis_phase(&a, k.x, h, x);
    axis_phase(&b, k.y, h, y);
    axis_phase(&c, k.z, h, z);
  }
  GlobalGrid grid{P, Kp, node};
  shell_store<MIXED>(grid, x, y, z, node, k, r_in[g], r_out[g], nx, ny, nz,
                     a, b, c, Kp, R);
}

// ---------------------------------------------------------------------------
// long-range back-projection

constexpr uint32_t kLrBlock = 256;

// Nodes are stored direction-major inside each group: a task is one
// (group, direction s) with its n_r radial nodes k_i = kappa_i s at
// consecutive indices, so every lane projects its atom once, p = s.r, and
// each radial node costs one multiply for the phase kappa_i p.
struct LrTasks {
  const uint32_t *node0;   // first node of the task
  const uint32_t *nr;      // radial nodes in the task
  const uint32_t *koff;    // offset of the task's kappa values
  const double4 *dir64;    // s (FP64)
  const float4 *dir32;     // s (FP32)
  const double *kappa64;   // kappa (1/A)
  const float *kappa32;    // kappa / (2 pi) (revolutions per A)
  const uint32_t *split;   // task range of each split: [split[s], split[s+1])
};

// FP64 (src/glst_force.cu:1060-1135): en_a = q_a sum_k Re(e^{ik.r_a} R_c(k)),
// F_a = q_a sum_k k Im(e^{ik.r_a} R_c(k)), absolute coordinates.
// Atom tile `tile` of one system (atoms, cell_sorted at its first atom, R at
// its field) over the tasks of split `split`; the result goes to partial row
// `prow` (partial at the system's partial block).
__device__ __forceinline__ void lr_f64_body(
    const double4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    const LrTasks &tk, const double2 *__restrict__ R, uint32_t Kp,
    uint32_t natom, uint32_t tile, uint32_t split, uint32_t prow,
    double4 *__restrict__ partial) {
  const uint32_t i = tile * kLrBlock + threadIdx.x;
  const bool valid = i < natom;
  const uint32_t ii = valid ? i : natom - 1;
  const double4 a = atoms[ii];
  const double2 *Rc = R + static_cast<size_t>(cell_sorted[ii] & ~kFarAtom) * Kp;
  double fx = 0.0, fy = 0.0, fz = 0.0, en = 0.0;
  const uint32_t t1 = tk.split[split + 1];
  for (uint32_t t = tk.split[split]; t < t1; ++t) {
    const uint32_t n0 = tk.node0[t], nr = tk.nr[t];
    const double *kap = tk.kappa64 + tk.koff[t];
    const double4 d = tk.dir64[t];
    const double p = d.x * a.x + d.y * a.y + d.z * a.z;
    // two radial nodes per iteration: independent chains, and the sine and
    // cosine coefficients stay in registers across the loop
    double e = 0.0, b = 0.0, e1 = 0.0, b1 = 0.0;
    uint32_t j = 0;
    for (; j + 2 <= nr; j += 2) {
      const double k0 = kap[j], k1 = kap[j + 1];
      const double2 r0 = Rc[n0 + j], r1 = Rc[n0 + j + 1];
      double s0, c0, s1, c1;
      sincos_cw(k0 * p, &s0, &c0);
      sincos_cw(k1 * p, &s1, &c1);
      e += c0 * r0.x - s0 * r0.y;
      b += k0 * (c0 * r0.y + s0 * r0.x);
      e1 += c1 * r1.x - s1 * r1.y;
      b1 += k1 * (c1 * r1.y + s1 * r1.x);
    }
    if (j < nr) {
      const double kj = kap[j];
      const double2 r = Rc[n0 + j];
      double sn, cs;
      sincos_cw(kj * p, &sn, &cs);
      e += cs * r.x - sn * r.y;
      b += kj * (cs * r.y + sn * r.x);
    }
    e += e1;
    b += b1;
    en += e;
    fx += d.x * b;
    fy += d.y * b;
    fz += d.z * b;
  }
  if (valid)
    partial[static_cast<size_t>(prow) * natom + i] =
        make_double4(a.w * fx, a.w * fy, a.w * fz, a.w * en);
}

__global__ __launch_bounds__(kLrBlock) void lr_f64_kernel(
    const double4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    LrTasks tk, const double2 *__restrict__ R, uint32_t Kp, uint32_t natom,
    uint32_t ntile, double4 *__restrict__ partial) {
  lr_f64_body(atoms, cell_sorted, tk, R, Kp, natom,
              xcd_remap(blockIdx.x, ntile), blockIdx.y, blockIdx.y, partial);
}

// Mixed, fixed radial count: every split lies inside one alpha group, so all
// its tasks have the same number NR of radial nodes and share the same
// kappa_i.  With NR a template parameter the kappa_i sit in registers for the
// whole split, the NR remote sums of a direction are fetched together with
// immediate offsets, and the radial loop unrolls completely.  R holds
// (A, phi) = (kappa |V|, arg V / 2 pi) per (cell, node), so per node:
// t = kappa_i p + phi (one FMA), sin, cos, b += A sin t, e_i += A cos t.
// The energy sums e_i are kept per radial node and divided by kappa_i when
// they are flushed (en = sum_i |V_i| cos t_i).
template <int NR, bool FAR>
__device__ __forceinline__ void lr_mixed_dirs(
    const float4 &a, const float2 *__restrict__ Rc, const LrTasks &tk,
    uint32_t tb, uint32_t te, double &FX, double &FY, double &FZ,
    double &EN) {
  float kap[NR];
  const float *kp = tk.kappa32 + tk.koff[tb];
#pragma unroll
  for (int i = 0; i < NR; ++i) kap[i] = kp[i];
  float fx = 0.f, fy = 0.f, fz = 0.f, e[NR];
#pragma unroll
  for (int i = 0; i < NR; ++i) e[i] = 0.f;
  auto flush_en = [&] {
    float en = 0.f;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      en = fmaf(e[i], 1.0f / kap[i], en);
      e[i] = 0.f;
    }
    EN += en;
  };
  // Force sums go to FP64 every GLST_LR_FLUSH_TASKS directions, energy
  // sums every GLST_LR_EN_FLUSH_TASKS; the inner loop has no branch.
  constexpr uint32_t kEnEvery = GLST_LR_EN_FLUSH_TASKS / GLST_LR_FLUSH_TASKS;
  uint32_t rounds = 0;
  for (uint32_t t0 = tb; t0 < te; t0 += GLST_LR_FLUSH_TASKS) {
    const uint32_t t1 =
        (te - t0 < GLST_LR_FLUSH_TASKS) ? te : t0 + GLST_LR_FLUSH_TASKS;
    for (uint32_t t = t0; t < t1; ++t) {
      const float2 *rp = Rc + tk.node0[t];
      const float4 d = tk.dir32[t];
      float2 r[NR];
#pragma unroll
      for (int i = 0; i < NR; ++i) r[i] = rp[i];
      const float p = d.x * a.x + d.y * a.y + d.z * a.z;
      float b0 = 0.f, b1 = 0.f;
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        const float tt = reduce_rev<FAR>(fmaf(kap[i], p, r[i].y));
        const float sn = sin_rev(tt), cs = cos_rev(tt);
        e[i] = fmaf(r[i].x, cs, e[i]);
        if (i & 1)
          b1 = fmaf(r[i].x, sn, b1);
        else
          b0 = fmaf(r[i].x, sn, b0);
      }
      const float b = b0 + b1;
      fx = fmaf(d.x, b, fx);
      fy = fmaf(d.y, b, fy);
      fz = fmaf(d.z, b, fz);
    }
    FX += fx;
    FY += fy;
    FZ += fz;
    fx = fy = fz = 0.f;
    if (++rounds == kEnEvery) {
      flush_en();
      rounds = 0;
    }
  }
  flush_en();
}

// Any radial count (larger than the unrolled cases).
template <bool FAR>
__device__ __forceinline__ void lr_mixed_dirs_any(
    const float4 &a, const float2 *__restrict__ Rc, const LrTasks &tk,
    uint32_t tb, uint32_t te, double &FX, double &FY, double &FZ,
    double &EN) {
  float fx = 0.f, fy = 0.f, fz = 0.f, en = 0.f;
  uint32_t pending = 0;
  for (uint32_t t = tb; t < te; ++t) {
    const uint32_t n0 = tk.node0[t], nr = tk.nr[t];
    const float *kap = tk.kappa32 + tk.koff[t];
    const float4 d = tk.dir32[t];
    const float p = d.x * a.x + d.y * a.y + d.z * a.z;
    float e = 0.f, b = 0.f;
    for (uint32_t j = 0; j < nr; ++j) {
      const float kj = kap[j];
      const float2 r = Rc[n0 + j];
      const float tt = reduce_rev<FAR>(fmaf(kj, p, r.y));
      const float sn = sin_rev(tt), cs = cos_rev(tt);
      e = fmaf(r.x * __builtin_amdgcn_rcpf(kj), cs, e);
      b = fmaf(r.x, sn, b);
    }
    en += e;
    fx = fmaf(d.x, b, fx);
    fy = fmaf(d.y, b, fy);
    fz = fmaf(d.z, b, fz);
    if (++pending == GLST_LR_FLUSH_TASKS) {
      FX += fx;
      FY += fy;
      FZ += fz;
      EN += en;
      fx = fy = fz = en = 0.f;
      pending = 0;
    }
  }
  FX += fx;
  FY += fy;
  FZ += fz;
  EN += en;
}

// One instantiation per radial count (NR = 0: any count); the host launches
// one grid per alpha group, blockIdx.y + split0 being the split.
template <int NR>
__device__ __forceinline__ void lr_mixed_nr_body(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    const LrTasks &tk, const float2 *__restrict__ R, uint32_t Kp,
    uint32_t natom, uint32_t tile, uint32_t split, uint32_t prow,
    double4 *__restrict__ partial) {
  const uint32_t i = tile * kLrBlock + threadIdx.x;
  const bool valid = i < natom;
  const uint32_t ii = valid ? i : natom - 1;
  const float4 a = atoms[ii];
  const uint32_t cs = cell_sorted[ii];
  const float2 *Rc = R + static_cast<size_t>(cs & ~kFarAtom) * Kp;
  double FX = 0.0, FY = 0.0, FZ = 0.0, EN = 0.0;
  const uint32_t tb = tk.split[split], te = tk.split[split + 1];
  // uniform: a wave with a far atom reduces all its phases
  const bool far = __ballot(valid && (cs & kFarAtom)) != 0;
  if (tb < te) {
    if constexpr (NR > 0) {
      if (far)
        lr_mixed_dirs<NR, true>(a, Rc, tk, tb, te, FX, FY, FZ, EN);
      else
        lr_mixed_dirs<NR, false>(a, Rc, tk, tb, te, FX, FY, FZ, EN);
    } else {
      if (far)
        lr_mixed_dirs_any<true>(a, Rc, tk, tb, te, FX, FY, FZ, EN);
      else
        lr_mixed_dirs_any<false>(a, Rc, tk, tb, te, FX, FY, FZ, EN);
    }
  }
  if (valid) {
    const double qf = static_cast<double>(a.w) * kTwoPi;  // k = 2 pi k_rev
    partial[static_cast<size_t>(prow) * natom + i] =
        make_double4(qf * FX, qf * FY, qf * FZ, static_cast<double>(a.w) * EN);
  }
}

template <int NR>
__global__ __launch_bounds__(kLrBlock) void lr_mixed_nr_kernel(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    LrTasks tk, const float2 *__restrict__ R, uint32_t Kp, uint32_t natom,
    uint32_t ntile, uint32_t split0, double4 *__restrict__ partial) {
  const uint32_t split = split0 + blockIdx.y;
  lr_mixed_nr_body<NR>(atoms, cell_sorted, tk, R, Kp, natom,
                       xcd_remap(blockIdx.x, ntile), split, split, partial);
}

typedef void (*LrNrKernel)(const float4 *, const uint32_t *, LrTasks,
                           const float2 *, uint32_t, uint32_t, uint32_t,
                           uint32_t, double4 *);

template <int... N>
constexpr auto lr_nr_table(std::integer_sequence<int, N...>) {
  return std::array<LrNrKernel, sizeof...(N)>{{&lr_mixed_nr_kernel<N>...}};
}

// index NR (1..kLrMaxUnrolled) -> unrolled kernel; 0 -> any radial count
constexpr int kLrMaxUnrolled = 40;
const auto kLrNrKernels =
    lr_nr_table(std::make_integer_sequence<int, kLrMaxUnrolled + 1>{});

// ---------------------------------------------------------------------------
// Mixed back-projection over atom groups.  An atom group is up to kAgAtoms
// consecutive sorted atoms of one cell (4 by default: a cell of n atoms has
// ceil(n / kAgAtoms) groups); ag[g] = (first atom, count), count 0 past the
// last group.  Eight lanes serve one group: lane m takes directions tb + m,
// tb + m + 8, ... of the split, loads that direction's NR remote sums
// (A, phi) of the group's cell once per pass and applies them to all the
// group's atoms, so a pair costs one FMA for the phase, sin, cos and two
// FMAs, and the vector memory path carries one R value per kAgAtoms pairs.
// A reduce-scatter over the eight lanes then gives lane m the sums of atom
// m mod kAgAtoms (lanes m >= kAgAtoms hold copies and write nothing).
constexpr uint32_t kAgLanes = 8;
constexpr uint32_t kAgPerBlock = kLrBlock / kAgLanes;
static_assert(kAgPerBlock == kAgGroupsPerBlock, "groups per workgroup");

// v[0 .. n) -> v[0 .. n/2): the lane keeps the half of the atoms selected by
// its `mask` bit and adds the partner lane's values of that half.
template <class T, int N>
__device__ __forceinline__ void ag_halve(T (&v)[N], int n, uint32_t lane,
                                         uint32_t mask) {
  const bool hi = (lane & mask) != 0;
#pragma unroll
  for (int j = 0; j < N / 2; ++j) {
    if (j < n / 2) {
      const T keep = hi ? v[j + n / 2] : v[j];
      const T send = hi ? v[j] : v[j + n / 2];
      v[j] = keep + __shfl_xor(send, mask, kAgLanes);
    }
  }
}

// Sum over the 8 lanes of a group of v[u] (u = atom of the group); lane m
// returns the sum of atom m mod kAgAtoms.
template <class T>
__device__ __forceinline__ T ag_reduce_scatter(T (&v)[kAgAtoms],
                                               uint32_t lane) {
  if constexpr (kAgAtoms == 8) {
    ag_halve(v, 8, lane, 4);
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kAgAtoms; ++j)
      v[j] += __shfl_xor(v[j], 4, kAgLanes);
  }
  ag_halve(v, 4, lane, 2);
  ag_halve(v, 2, lane, 1);
  return v[0];
}

// Atom tile `tile` of one system (atoms, cell_sorted at its first atom, whose
// index in the group table's numbering is atom_base; R at its field).
template <int NR>
__device__ __forceinline__ void lr_ag_body(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    const uint2 *__restrict__ ag, uint32_t atom_base, const LrTasks &tk,
    const float2 *__restrict__ R, uint32_t Kp, uint32_t natom, uint32_t tile,
    uint32_t split, uint32_t prow, double4 *__restrict__ partial) {
  const uint32_t lane = threadIdx.x & (kAgLanes - 1);
  const uint2 g = ag[tile * kAgPerBlock + threadIdx.x / kAgLanes];
  const uint32_t cnt = g.y, first = cnt ? g.x - atom_base : 0u;
  const uint32_t tb = tk.split[split], te = tk.split[split + 1];
  // uniform per wave: every group of the wave is padding (nothing to write)
  if (__ballot(cnt != 0) == 0) return;
  if (tb >= te) {  // a split without directions adds zero
    if (lane < cnt)
      partial[static_cast<size_t>(prow) * natom + first + lane] =
          make_double4(0.0, 0.0, 0.0, 0.0);
    return;
  }
  const float2 *Rc =
      R + static_cast<size_t>(cell_sorted[first] & ~kFarAtom) * Kp;
  float ax[kAgAtoms], ay[kAgAtoms], az[kAgAtoms];
  uint32_t farg = 0u;
#pragma unroll
  for (uint32_t u = 0; u < kAgAtoms; ++u) {
    const uint32_t j = first + (u < cnt ? u : 0u);
    const float4 a = atoms[j];
    ax[u] = a.x;
    ay[u] = a.y;
    az[u] = a.z;
    farg |= cnt ? cell_sorted[j] & kFarAtom : 0u;
  }
  // uniform: a wave with a far atom reduces all its phases
  const bool far = __ballot(farg != 0u) != 0;
  float kap[NR], ikap[NR];
  const float *kp = tk.kappa32 + tk.koff[tb];
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    kap[i] = kp[i];
    // uniform: kept in a scalar register (readfirstlane moves 32-bit words)
    ikap[i] = __int_as_float(
        __builtin_amdgcn_readfirstlane(__float_as_int(1.0f / kap[i])));
  }
  // Forces in FP32 per lane; the potential's terms cancel far more, so each
  // pass's energy sum goes into FP64.
  float fx[kAgAtoms], fy[kAgAtoms], fz[kAgAtoms];
  double en[kAgAtoms];
#pragma unroll
  for (uint32_t u = 0; u < kAgAtoms; ++u) {
    fx[u] = fy[u] = fz[u] = 0.f;
    en[u] = 0.0;
  }
  // One pass: lane direction t, all atoms of the group.  FULL: every lane
  // has a direction (the last pass of a split may leave lanes idle).
  auto pass = [&](uint32_t t0, auto full_tag, auto far_tag) {
    constexpr bool FULL = decltype(full_tag)::value;
    constexpr bool FAR = decltype(far_tag)::value;
    const uint32_t t = t0 + lane;
    const bool on = FULL || t < te;
    const uint32_t tt = on ? t : tb;
    const float4 d = tk.dir32[tt];
    const float2 *rp = Rc + tk.node0[tt];
    float A[NR], V[NR], ph[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const float2 r = rp[i];
      A[i] = on ? r.x : 0.f;  // idle lanes add nothing
      V[i] = A[i] * ikap[i];  // |V|
      ph[i] = r.y;
    }
#pragma unroll
    for (uint32_t u = 0; u < kAgAtoms; ++u) {
      const float p = d.x * ax[u] + d.y * ay[u] + d.z * az[u];
      float e = 0.f, b = 0.f;
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        const float th = reduce_rev<FAR>(fmaf(kap[i], p, ph[i]));
        e = fmaf(V[i], cos_rev(th), e);
        b = fmaf(A[i], sin_rev(th), b);
      }
      en[u] += static_cast<double>(e);
      fx[u] = fmaf(d.x, b, fx[u]);
      fy[u] = fmaf(d.y, b, fy[u]);
      fz[u] = fmaf(d.z, b, fz[u]);
    }
  };
  uint32_t t0 = tb;
  if (far) {
    for (; t0 + kAgLanes <= te; t0 += kAgLanes)
      pass(t0, std::true_type{}, std::true_type{});
    if (t0 < te) pass(t0, std::false_type{}, std::true_type{});
  } else {
    for (; t0 + kAgLanes <= te; t0 += kAgLanes)
      pass(t0, std::true_type{}, std::false_type{});
    if (t0 < te) pass(t0, std::false_type{}, std::false_type{});
  }
  const float FX = ag_reduce_scatter(fx, lane);
  const float FY = ag_reduce_scatter(fy, lane);
  const float FZ = ag_reduce_scatter(fz, lane);
  const double EN = ag_reduce_scatter(en, lane);
  if (lane < cnt) {
    const uint32_t i = first + lane;
    const double q = static_cast<double>(atoms[i].w);
    const double qf = q * kTwoPi;  // k = 2 pi k_rev
    partial[static_cast<size_t>(prow) * natom + i] =
        make_double4(qf * FX, qf * FY, qf * FZ, q * EN);
  }
}

// Waves per SIMD the atom-group kernel's register budget allows: 4 (128
// VGPRs) up to GLST_LR_AG_4WAVE_NR radial nodes, where that costs at most a
// few bytes of scratch (one 8-byte spill and one reload, 12 B per lane, at
// 16 and 19 nodes) and made the kernel 3.5 % faster at 288k atoms; else
// whatever the kernel needs, except at 23 and 24 nodes.  Those compile to
// 169 and 172 VGPRs uncapped, just past the 168 that three waves allow;
// capped at three they spill 20-44 B per lane and the batched step of 512
// small systems (most of whose groups have 13-26 nodes) runs 1.6 % faster.
#ifndef GLST_LR_AG_4WAVE_NR
#define GLST_LR_AG_4WAVE_NR 19
#endif
constexpr bool lr_ag_3wave(int nr) { return nr == 23 || nr == 24; }
template <int NR>
constexpr int lr_ag_min_waves() {
  return NR <= GLST_LR_AG_4WAVE_NR ? 4 : lr_ag_3wave(NR) ? 3 : 1;
}
template <int NR>
__global__ __launch_bounds__(kLrBlock, lr_ag_min_waves<NR>()) void lr_ag_kernel(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    const uint2 *__restrict__ ag, LrTasks tk, const float2 *__restrict__ R,
    uint32_t Kp, uint32_t natom, uint32_t ntile, uint32_t split0,
    double4 *__restrict__ partial) {
  const uint32_t split = split0 + blockIdx.y;
  lr_ag_body<NR>(atoms, cell_sorted, ag, 0u, tk, R, Kp, natom,
                 xcd_remap(blockIdx.x, ntile), split, split, partial);
}

typedef void (*LrAgKernel)(const float4 *, const uint32_t *, const uint2 *,
                           LrTasks, const float2 *, uint32_t, uint32_t,
                           uint32_t, uint32_t, double4 *);

template <int... N>
constexpr auto lr_ag_table(std::integer_sequence<int, N...>) {
  return std::array<LrAgKernel, sizeof...(N)>{{&lr_ag_kernel<N + 1>...}};
}

// index NR - 1 (NR = 1 .. kLrMaxUnrolled)
const auto kLrAgKernels =
    lr_ag_table(std::make_integer_sequence<int, kLrMaxUnrolled>{});

// Mixed structure factors, one instantiation per radial count NR: lane = one
// direction s of an alpha group, one wave per (cell, 64 directions), atom
// records are wave-uniform.  Per atom the lane projects p = s.d once; each of
// the NR radial nodes then costs one multiply for its phase kappa_i p, fract,
// sin and cos and two FMAs.  The kappa_i are wave-uniform registers.  The
// lane stores the cell-local sums sum_j q_j exp(-i k.d_j) as FP32 pairs in
// the node runs of sf_slot (FP32 running sums over at most 128 atoms,
// denser cells add their further chunks to the stored value); the
// shell-sum kernel applies the FP64 cell-centre phase exp(-i k.c).
constexpr uint32_t kSfDirs = 64;
constexpr uint32_t kSfChunk = 128;

// Cells cell0 + w of one system (start at its first cell, S at its field),
// wave w of the workgroup taking one cell, for direction chunk `chunk` of
// the group whose tasks are task0 .. +ntask (nodes of task t: node0[t] + i,
// i < NR, the tasks' nodes consecutive).  The lanes of a wave hold 64 NR
// consecutive nodes, so their stores fill whole node runs of the cell (all
// but the chunk's first and last run, which it shares with the neighbouring
// chunks).
template <int NR>
__device__ __forceinline__ void sf_mixed_nr_body(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ cell_sorted, const LrTasks &tk,
    uint32_t task0, uint32_t ntask, uint32_t ncell, float2 *__restrict__ S,
    uint32_t cell0, uint32_t chunk) {
  const uint32_t w = threadIdx.x / kSfDirs, lane = threadIdx.x % kSfDirs;
  const uint32_t cell = cell0 + w;
  const uint32_t t = task0 + chunk * kSfDirs + lane;
  const bool active = t < task0 + ntask;
  const uint32_t tt = active ? t : task0;
  float kap[NR];
  const float *kp = tk.kappa32 + tk.koff[task0];
#pragma unroll
  for (int i = 0; i < NR; ++i) kap[i] = kp[i];
  const float4 d = tk.dir32[tt];
  const uint32_t n0 = tk.node0[tt];
  auto slot = [&](int i) -> float2 & {
    return S[sf_slot<true>(n0 + i, cell, ncell)];
  };
  // uniform over the wave (its cell): does the cell hold a far atom?  One
  // vector read of the cell's cell_sorted entries, then a ballot.
  bool far = false;
  if (cell < ncell) {
    uint32_t f = 0u;
    for (uint32_t j = start[cell] + lane; j < start[cell + 1]; j += kSfDirs)
      f |= cell_sorted[j] & kFarAtom;
    far = __ballot(f != 0u) != 0;
  }
  if (cell < ncell && active) {
    const uint32_t a0 = start[cell], a1 = start[cell + 1];
    if (a0 == a1) {
#pragma unroll
      for (int i = 0; i < NR; ++i) slot(i) = make_float2(0.f, 0.f);
    }
    for (uint32_t jb = a0; jb < a1; jb += kSfChunk) {
      const uint32_t je = (jb + kSfChunk < a1) ? jb + kSfChunk : a1;
      float re[NR], im[NR];
#pragma unroll
      for (int i = 0; i < NR; ++i) re[i] = im[i] = 0.f;
      auto atoms_loop = [&](auto far_tag) {
        constexpr bool FAR = decltype(far_tag)::value;
        for (uint32_t j = jb; j < je; ++j) {
          const float4 a = atoms[j];
          const float p = d.x * a.x + d.y * a.y + d.z * a.z;
#pragma unroll
          for (int i = 0; i < NR; ++i) {
            const float ph = reduce_rev<FAR>(kap[i] * p);
            re[i] = fmaf(a.w, cos_rev(ph), re[i]);
            im[i] = fmaf(-a.w, sin_rev(ph), im[i]);
          }
        }
      };
      if (far)
        atoms_loop(std::true_type{});
      else
        atoms_loop(std::false_type{});
      if (jb == a0) {
#pragma unroll
        for (int i = 0; i < NR; ++i) slot(i) = make_float2(re[i], im[i]);
      } else {  // cells denser than one chunk
#pragma unroll
        for (int i = 0; i < NR; ++i) {
          const float2 o = slot(i);
          slot(i) = make_float2(o.x + re[i], o.y + im[i]);
        }
      }
    }
  }
}

// Any radial count (beyond the unrolled cases).
__device__ __forceinline__ void sf_mixed_any_body(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ cell_sorted, const LrTasks &tk,
    uint32_t task0, uint32_t ntask, uint32_t ncell, float2 *__restrict__ S,
    uint32_t cell0, uint32_t chunk) {
  // wave w of the workgroup takes cell cell0 + w
  const uint32_t lane = threadIdx.x % kSfDirs;
  const uint32_t cell = cell0 + threadIdx.x / kSfDirs;
  const uint32_t t = task0 + chunk * kSfDirs + lane;
  const bool active = t < task0 + ntask;
  const uint32_t tt = active ? t : task0;
  const float *kap = tk.kappa32 + tk.koff[task0];
  const uint32_t nr = tk.nr[task0];
  const float4 d = tk.dir32[tt];
  const uint32_t n0 = tk.node0[tt];
  if (cell < ncell) {  // uniform over the wave
    const uint32_t a0 = start[cell], a1 = start[cell + 1];
    // uniform: does the cell hold a far atom (one vector read, a ballot)?
    uint32_t f = 0u;
    for (uint32_t j = a0 + lane; j < a1; j += kSfDirs)
      f |= cell_sorted[j] & kFarAtom;
    const bool far = __ballot(f != 0u) != 0;
    for (uint32_t i = 0; i < nr; ++i) {
      double RE = 0.0, IM = 0.0;
      for (uint32_t jb = a0; jb < a1; jb += kSfChunk) {
        const uint32_t je = (jb + kSfChunk < a1) ? jb + kSfChunk : a1;
        float re = 0.f, im = 0.f;
        for (uint32_t j = jb; j < je; ++j) {
          const float4 a = atoms[j];
          const float t = kap[i] * (d.x * a.x + d.y * a.y + d.z * a.z);
          const float ph = far ? reduce_rev<true>(t) : reduce_rev<false>(t);
          re = fmaf(a.w, cos_rev(ph), re);
          im = fmaf(-a.w, sin_rev(ph), im);
        }
        RE += re;
        IM += im;
      }
      if (active)
        S[sf_slot<true>(n0 + i, cell, ncell)] =
            make_float2(static_cast<float>(RE), static_cast<float>(IM));
    }
  }
}

// Chebyshev form (plan_chebyshev): per atom the lane projects x = s.d / rho
// and adds q T_m(x), m < M, to its M moments (the recurrence T_{m+1} =
// 2 x T_m - T_{m-1} carried on q T_m: two FMAs per term); then S_i =
// sum_m C[i M + m] moment_m over the group's matrix C (wave-uniform, scalar
// loads), even m into the real part, odd m into the imaginary part.  A cell
// with a far atom, or with an atom farther than dlim from its centre in some
// axis (binned into an edge cell from outside the grid), takes the direct
// sum instead (uniform over the wave: one vector read of the cell's atoms
// and a ballot).
template <int M>
__device__ __forceinline__ void sf_cheb_body(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ cell_sorted, const LrTasks &tk,
    uint32_t task0, uint32_t ntask, uint32_t ncell, float2 *__restrict__ S,
    uint32_t cell0, uint32_t chunk, const float *__restrict__ C,
    float inv_rho, float dlim) {
  const uint32_t lane = threadIdx.x % kSfDirs;
  const uint32_t cell = cell0 + threadIdx.x / kSfDirs;
  if (cell >= ncell) return;  // uniform over the wave
  const uint32_t a0 = start[cell], a1 = start[cell + 1];
  uint32_t bad = 0u;
  for (uint32_t j = a0 + lane; j < a1; j += kSfDirs) {
    const float4 a = atoms[j];
    const float m = fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z)));
    bad |= (cell_sorted[j] & kFarAtom) | (m > dlim ? 1u : 0u);
  }
  if (__ballot(bad != 0u) != 0) {
    sf_mixed_any_body(atoms, start, cell_sorted, tk, task0, ntask, ncell, S,
                      cell0, chunk);
    return;
  }
  const uint32_t t = task0 + chunk * kSfDirs + lane;
  if (t >= task0 + ntask) return;
  const float4 d = tk.dir32[t];
  const float sx = d.x * inv_rho, sy = d.y * inv_rho, sz = d.z * inv_rho;
  float mom[M];
#pragma unroll
  for (int m = 0; m < M; ++m) mom[m] = 0.f;
  for (uint32_t j = a0; j < a1; ++j) {
    const float4 a = atoms[j];
    const float x = fmaf(sx, a.x, fmaf(sy, a.y, sz * a.z));
    const float q = a.w;
    mom[0] += q;
    if constexpr (M > 1) {
      float u0 = q, u1 = q * x;
      mom[1] += u1;
      const float x2 = x + x;
#pragma unroll
      for (int m = 2; m < M; ++m) {
        const float u = fmaf(x2, u1, -u0);
        mom[m] += u;
        u0 = u1;
        u1 = u;
      }
    }
  }
  const uint32_t nr = tk.nr[task0];
  const uint32_t n0 = tk.node0[t];
  for (uint32_t i = 0; i < nr; ++i) {
    const float *c = C + i * M;
    float re = 0.f, im = 0.f;
#pragma unroll
    for (int m = 0; m < M; m += 2) re = fmaf(c[m], mom[m], re);
#pragma unroll
    for (int m = 1; m < M; m += 2) im = fmaf(c[m], mom[m], im);
    S[sf_slot<true>(n0 + i, cell, ncell)] = make_float2(re, im);
  }
}

template <int M>
__global__ __launch_bounds__(kSfDirs) void sf_cheb_kernel(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ cell_sorted, LrTasks tk, uint32_t task0,
    uint32_t ntask, uint32_t ncell, float2 *__restrict__ S,
    const float *__restrict__ C, float inv_rho, float dlim) {
  // gridDim.x = cells
  const uint32_t wg = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x,
                                gridDim.x * gridDim.y);
  const uint32_t ncg = gridDim.x;
  sf_cheb_body<M>(atoms, start, cell_sorted, tk, task0, ntask, ncell, S,
                  wg % ncg, wg / ncg, C, inv_rho, dlim);
}

typedef void (*SfChebKernel)(const float4 *, const uint32_t *,
                             const uint32_t *, LrTasks, uint32_t, uint32_t,
                             uint32_t, float2 *, const float *, float, float);

// index M - 1 (M = 1 .. kChebMaxTerms)
template <int... N>
constexpr auto sf_cheb_table(std::integer_sequence<int, N...>) {
  return std::array<SfChebKernel, sizeof...(N)>{{&sf_cheb_kernel<N + 1>...}};
}

const auto kSfChebKernels =
    sf_cheb_table(std::make_integer_sequence<int, kChebMaxTerms>{});

// (cell, direction chunk) of a workgroup, numbered so that each XCD runs
// consecutive cells.  amdgpu_waves_per_eu(1, 10) lets the compiler size the
// registers of the unrolled radial loop for as few as one wave per SIMD.
template <int NR>
__global__ __launch_bounds__(kSfDirs)
__attribute__((amdgpu_waves_per_eu(1, 10)))
void sf_mixed_nr_kernel(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ cell_sorted, LrTasks tk, uint32_t task0,
    uint32_t ntask, uint32_t ncell, float2 *__restrict__ S) {
  // gridDim.x = cells
  const uint32_t wg = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x,
                                gridDim.x * gridDim.y);
  const uint32_t ncg = gridDim.x;
  if constexpr (NR > 0)
    sf_mixed_nr_body<NR>(atoms, start, cell_sorted, tk, task0, ntask, ncell,
                         S, wg % ncg, wg / ncg);
  else
    sf_mixed_any_body(atoms, start, cell_sorted, tk, task0, ntask, ncell, S,
                      wg % ncg, wg / ncg);
}

typedef void (*SfNrKernel)(const float4 *, const uint32_t *,
                           const uint32_t *, LrTasks, uint32_t, uint32_t,
                           uint32_t, float2 *);

// index NR (1..kSfMaxUnrolled) -> unrolled kernel; 0 -> any radial count
template <int... N>
constexpr auto sf_nr_table(std::integer_sequence<int, N...>) {
  return std::array<SfNrKernel, sizeof...(N)>{{&sf_mixed_nr_kernel<N>...}};
}

const auto kSfNrKernels =
    sf_nr_table(std::make_integer_sequence<int, kSfMaxUnrolled + 1>{});

// f[i] += sum of the node-range partials (fixed order), plus the near-field
// record when one is given.
__device__ __forceinline__ void lr_reduce_body(
    const double4 *__restrict__ partial, uint32_t nsplit, uint32_t natom,
    const double4 *__restrict__ near, double *__restrict__ fx,
    double *__restrict__ fy, double *__restrict__ fz, double *__restrict__ en,
    uint32_t i) {
  if (i >= natom) return;
  double4 s = nsplit ? partial[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  for (uint32_t p = 1; p < nsplit; ++p) {
    const double4 v = partial[static_cast<size_t>(p) * natom + i];
    s.x += v.x;
    s.y += v.y;
    s.z += v.z;
    s.w += v.w;
  }
  if (near) {
    const double4 v = near[i];
    s.x += v.x;
    s.y += v.y;
    s.z += v.z;
    s.w += v.w;
  }
  fx[i] += s.x;
  fy[i] += s.y;
  fz[i] += s.z;
  en[i] += s.w;
}

__global__ void lr_reduce_kernel(const double4 *__restrict__ partial,
                                 uint32_t nsplit, uint32_t natom,
                                 const double4 *__restrict__ near,
                                 double *__restrict__ fx,
                                 double *__restrict__ fy,
                                 double *__restrict__ fz,
                                 double *__restrict__ en) {
  lr_reduce_body(partial, nsplit, natom, near, fx, fy, fz, en,
                 blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// near field: bare Coulomb over the cell and its 26 neighbours, clamped at the
// grid edge (src/glst_force.cu:1170-1544).
//
// One 256-thread workgroup per cell.  A cell of n atoms is covered S =
// floor(256/n) times: lane (i, s) takes atom i against every S-th atom
// (phase s) of each neighbour range, so a cell of 85 atoms keeps 255 lanes
// busy instead of 128 lanes in two 64-atom passes (denser cells run in
// 256-atom chunks with S = 1).  The neighbour cells z-1 .. z+1 of one
// (x, y) column are consecutive in the sorted order, so the 27 neighbour
// cells are 9 atom ranges.  Each range is staged through LDS tiles of up to
// 256 atom records, loaded once by the workgroup and read by every lane
// (the lanes of one phase read the same record); the S partial sums of an
// atom are then added in LDS in a fixed order, in the same buffer.
constexpr uint32_t kSrBlock = 256;
constexpr uint32_t kSrMaxPhase = 16;

// (F, e) / q_i of one pair; r = r_i - r_j, e = q_i q_j / |r|,
// F = e r / |r|^2.  q_i multiplies the sums over a neighbour range once,
// when they are flushed.  Mixed: the separation in FP64, the rest in FP32.
// SELF: the pair may be the atom with itself (only in its own column), which
// then adds nothing.
template <bool MIXED>
struct PairAcc;

// A near-field LDS tile record.  Mixed tiles carry q_j already rounded to
// FP32 in the low word of .w (rounded once per staged atom, not per pair).
template <bool MIXED>
__device__ __forceinline__ double4 sr_tile_record(const double4 &a) {
  if constexpr (MIXED) {
    const float q = static_cast<float>(a.w);
    return make_double4(a.x, a.y, a.z,
                        __hiloint2double(0, __float_as_int(q)));
  } else {
    return a;
  }
}

template <>
struct PairAcc<true> {
  float fx = 0.f, fy = 0.f, fz = 0.f, en = 0.f;
  template <bool SELF>
  __device__ void add(const double4 &ai, const double4 &aj, bool self) {
    const float rx = static_cast<float>(ai.x - aj.x);
    const float ry = static_cast<float>(ai.y - aj.y);
    const float rz = static_cast<float>(ai.z - aj.z);
    float r2 = rx * rx + ry * ry + rz * rz;
    if (SELF) r2 = self ? INFINITY : r2;  // no self interaction
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float e = __int_as_float(__double2loint(aj.w)) * ir;
    const float f = e * ir * ir;
    fx += f * rx;
    fy += f * ry;
    fz += f * rz;
    en += e;
  }
};

template <>
struct PairAcc<false> {
  double fx = 0.0, fy = 0.0, fz = 0.0, en = 0.0;
  template <bool SELF>
  __device__ void add(const double4 &ai, const double4 &aj, bool self) {
    const double rx = ai.x - aj.x, ry = ai.y - aj.y, rz = ai.z - aj.z;
    const double r2 = rx * rx + ry * ry + rz * rz;
    // 1/|r|: the FP32 hardware estimate (relative error about 1e-7) refined
    // by two Newton steps in FP64, each of which squares the relative error
    // (1e-7 -> 1e-14 -> FP64 rounding); cheaper than an FP64 square root and
    // division.  Near-field separations lie far inside FP32's range.
    double ir = static_cast<double>(__builtin_amdgcn_rsqf(static_cast<float>(r2)));
    const double h = 0.5 * r2;
    ir = ir * fma(-h, ir * ir, 1.5);
    ir = ir * fma(-h, ir * ir, 1.5);
    if (SELF) ir = self ? 0.0 : ir;  // no self interaction
    const double e = aj.w * ir;
    const double f = e * ir * ir;
    fx += f * rx;
    fy += f * ry;
    fz += f * rz;
    en += e;
  }
};

// STORE: write (fx, fy, fz, en) to out4[i] (the caller adds it later);
// otherwise accumulate into fx..en.
// Cell `cell` of one system (start at the system's first cell, holding
// batch-wide atom slots; atoms and outputs batch-wide).
template <bool MIXED, bool STORE>
__device__ __forceinline__ void sr_body(
    const double4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    uint32_t nx, uint32_t ny, uint32_t nz, double *__restrict__ fx,
    double *__restrict__ fy, double *__restrict__ fz,
    double *__restrict__ en, double4 *__restrict__ out4, uint32_t cell,
    double4 *__restrict__ tile) {
  // `tile` (kSrBlock records of LDS): the neighbour tiles, then the per-lane
  // partial sums (8 KB, so that two workgroups fit beside the shell sums'
  // tables)
  double4 *const part = tile;
  const uint32_t ix = cell / (ny * nz), iy = (cell / nz) % ny, iz = cell % nz;
  const uint32_t a0 = start[cell], a1 = start[cell + 1];
  const uint32_t n = a1 - a0;
  if (n == 0) return;  // uniform over the workgroup
  uint32_t S = n >= kSrBlock ? 1 : kSrBlock / n;
  S = S > kSrMaxPhase ? kSrMaxPhase : S;
  const uint32_t chunk = kSrBlock / S;  // atoms per pass
  const uint32_t tid = threadIdx.x;
  const uint32_t li = tid / S, s = tid - li * S;
  // waves whose lanes all lie past chunk * S only help load the tiles
  const bool wave_busy = (tid & ~63u) < chunk * S;
  for (uint32_t base = a0; base < a1; base += chunk) {
    const uint32_t i = base + li;
    const bool valid = li < chunk && i < a1;
    const uint32_t ii = valid ? i : a0;
    double FX = 0.0, FY = 0.0, FZ = 0.0, EN = 0.0;
    const double4 ai = atoms[ii];
    using Q = typename std::conditional<MIXED, float, double>::type;
    const Q qi = static_cast<Q>(ai.w);
    for (int dx = -1; dx <= 1; ++dx) {
      if ((ix == 0 && dx < 0) || (ix == nx - 1 && dx > 0)) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        if ((iy == 0 && dy < 0) || (iy == ny - 1 && dy > 0)) continue;
        {
          // The neighbour cells iz-1 .. iz+1 of one (x, y) column are
          // consecutive in the sorted order: one contiguous atom range,
          // staged in tiles of up to kSrBlock atoms.
          const uint32_t col = ((ix + dx) * ny + (iy + dy)) * nz;
          const uint32_t z0 = iz > 0 ? iz - 1 : iz;
          const uint32_t z1 = iz + 1 < nz ? iz + 1 : iz;
          const uint32_t b0 = start[col + z0], b1 = start[col + z1 + 1];
          // Mixed: FP32 sums over one neighbour column, flushed to FP64.
          PairAcc<MIXED> acc;
          // The self pair exists only in the atom's own column: the other
          // columns run the pair loop without the test (uniform branch).
          auto sweep = [&](auto self_tag) {
            constexpr bool SELF = decltype(self_tag)::value;
            for (uint32_t t0 = b0; t0 < b1; t0 += kSrBlock) {
              const uint32_t tn = (b1 - t0 < kSrBlock) ? b1 - t0 : kSrBlock;
              __syncthreads();  // the previous tile has been consumed
              if (tid < tn) tile[tid] = sr_tile_record<MIXED>(atoms[t0 + tid]);
              __syncthreads();
              if (wave_busy) {
                uint32_t j = s;
                for (; j + 3 * S < tn; j += 4 * S) {
                  const double4 q0 = tile[j], q1 = tile[j + S],
                                q2 = tile[j + 2 * S], q3 = tile[j + 3 * S];
                  acc.template add<SELF>(ai, q0, t0 + j == ii);
                  acc.template add<SELF>(ai, q1, t0 + j + S == ii);
                  acc.template add<SELF>(ai, q2, t0 + j + 2 * S == ii);
                  acc.template add<SELF>(ai, q3, t0 + j + 3 * S == ii);
                }
                for (; j < tn; j += S)
                  acc.template add<SELF>(ai, tile[j], t0 + j == ii);
              }
            }
          };
          if (dx == 0 && dy == 0)
            sweep(std::true_type{});
          else
            sweep(std::false_type{});
          // exact products in FP64 for the mixed path (24 x 24 bits)
          FX += static_cast<double>(qi) * acc.fx;
          FY += static_cast<double>(qi) * acc.fy;
          FZ += static_cast<double>(qi) * acc.fz;
          EN += static_cast<double>(qi) * acc.en;
        }
      }
    }
    __syncthreads();  // the last tile has been consumed
    part[tid] = valid ? make_double4(FX, FY, FZ, EN)
                      : make_double4(0.0, 0.0, 0.0, 0.0);
    __syncthreads();
    if (tid < chunk && base + tid < a1) {
      double4 t = part[tid * S];
      for (uint32_t u = 1; u < S; ++u) {
        const double4 v = part[tid * S + u];
        t.x += v.x;
        t.y += v.y;
        t.z += v.z;
        t.w += v.w;
      }
      const uint32_t o = base + tid;
      if (STORE) {
        out4[o] = t;
      } else {
        fx[o] += t.x;
        fy[o] += t.y;
        fz[o] += t.z;
        en[o] += t.w;
      }
    }
    __syncthreads();
  }
}

template <bool MIXED, bool STORE>
__global__ __launch_bounds__(kSrBlock) void sr_kernel(
    const double4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    uint32_t nx, uint32_t ny, uint32_t nz, double *__restrict__ fx,
    double *__restrict__ fy, double *__restrict__ fz,
    double *__restrict__ en, double4 *__restrict__ out4) {
  __shared__ double4 tile[kSrBlock];
  sr_body<MIXED, STORE>(atoms, start, nx, ny, nz, fx, fy, fz, en, out4,
                        xcd_remap(blockIdx.x, nx * ny * nz), tile);
}

// ---------------------------------------------------------------------------

__global__ void unpermute_kernel(const uint32_t *__restrict__ sorted_idx,
                                 uint32_t natom, const double *__restrict__ fx,
                                 const double *__restrict__ fy,
                                 const double *__restrict__ fz,
                                 const double *__restrict__ en,
                                 double *__restrict__ ofx,
                                 double *__restrict__ ofy,
                                 double *__restrict__ ofz,
                                 double *__restrict__ oen) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natom) return;
  const uint32_t o = sorted_idx[i];
  ofx[o] = fx[i];
  ofy[o] = fy[i];
  ofz[o] = fz[i];
  oen[o] = en[i];
}

constexpr uint32_t kCoulBlock = 256;

// Direct all-pairs sum of utils/coulomb.cu:14-81 over the source range
// [j_begin, j_end) of split blockIdx.y, j_begin = blockIdx.y * jchunk (jchunk
// a multiple of kCoulBlock).  With one split the sums go straight to the
// outputs; otherwise split s writes (fx, fy, fz, en) of atom i to
// part[s * natom + i] and coulomb_reduce_kernel adds the splits in order.
__global__ __launch_bounds__(kCoulBlock) void coulomb_kernel(
    const double *__restrict__ rx, const double *__restrict__ ry,
    const double *__restrict__ rz, const double *__restrict__ qc,
    uint32_t natom, uint32_t jchunk, double *__restrict__ fx,
    double *__restrict__ fy, double *__restrict__ fz, double *__restrict__ en,
    double4 *__restrict__ part) {
  __shared__ double4 tile[kCoulBlock];
  const uint32_t i = blockIdx.x * kCoulBlock + threadIdx.x;
  const uint32_t ii = i < natom ? i : 0;
  const double xi = rx[ii], yi = ry[ii], zi = rz[ii], qi = qc[ii];
  const uint32_t j_begin = blockIdx.y * jchunk;
  const uint32_t j_end =
      (natom - j_begin < jchunk) ? natom : j_begin + jchunk;
  double FX = 0.0, FY = 0.0, FZ = 0.0, EN = 0.0;
  for (uint32_t j0 = j_begin; j0 < j_end; j0 += kCoulBlock) {
    __syncthreads();
    const uint32_t j = j0 + threadIdx.x;
    if (j < j_end) tile[threadIdx.x] = make_double4(rx[j], ry[j], rz[j], qc[j]);
    __syncthreads();
    const uint32_t n = (j_end - j0 < kCoulBlock) ? j_end - j0 : kCoulBlock;
    for (uint32_t t = 0; t < n; ++t) {
      const double4 b = tile[t];
      const double dx = xi - b.x, dy = yi - b.y, dz = zi - b.z;
      double r2 = dx * dx + dy * dy + dz * dz;
      r2 = (j0 + t == i) ? INFINITY : r2;
      const double ir = 1.0 / sqrt(r2);
      const double e = qi * b.w * ir;
      const double f = e * ir * ir;
      FX += f * dx;
      FY += f * dy;
      FZ += f * dz;
      EN += e;
    }
  }
  if (i >= natom) return;
  if (part) {
    part[(size_t)blockIdx.y * natom + i] = make_double4(FX, FY, FZ, EN);
  } else {
    fx[i] = FX;
    fy[i] = FY;
    fz[i] = FZ;
    en[i] = EN;
  }
}

// outputs of atom i = sum over splits s = 0, 1, ... of part[s * natom + i]
__global__ __launch_bounds__(256) void coulomb_reduce_kernel(
    const double4 *__restrict__ part, uint32_t nsplit, uint32_t natom,
    double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz,
    double *__restrict__ en) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= natom) return;
  double4 a = part[i];
  for (uint32_t s = 1; s < nsplit; ++s) {
    const double4 b = part[(size_t)s * natom + i];
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    a.w += b.w;
  }
  fx[i] = a.x;
  fy[i] = a.y;
  fz[i] = a.z;
  en[i] = a.w;
}

inline uint32_t blocks_for(size_t n, uint32_t b) {
  return static_cast<uint32_t>((n + b - 1) / b);
}

// ---------------------------------------------------------------------------
// batched systems: every workgroup finds its run (system, and alpha group
// where the launch is per radial count) by binary search over the runs'
// first workgroups, then runs the single-system body on its system's slice.
// The *_unit functions take the workgroup's number within its launch (blk);
// the per-stage kernels and the fused step share them.

__device__ __forceinline__ uint32_t run_of(const Seg *__restrict__ seg,
                                           uint32_t nseg, uint32_t blk) {
  uint32_t lo = 0, hi = nseg;  // the last run with blk0 <= blk
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (seg[mid].blk0 <= blk)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// threads of the batched histogram, scatter and reduce workgroups
constexpr uint32_t kBatchBlock = 256;

template <typename CT, uint32_t NT, bool ZERO>
__device__ __forceinline__ void bin_chunk_unit(
    const double *__restrict__ rx, const double *__restrict__ ry,
    const double *__restrict__ rz, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg, uint32_t *__restrict__ atom_cell,
    uint32_t *__restrict__ rank, uint32_t *__restrict__ hist, uint32_t blk,
    unsigned long long *__restrict__ keys) {
  const Seg g = seg[run_of(seg, nseg, blk)];
  const SysDev &d = sys[g.sys];
  bin_chunk_body<CT, NT, ZERO>(rx + d.atom0, ry + d.atom0, rz + d.atom0,
                               d.natom, static_cast<CT>(d.inv_cell), d.nx,
                               d.ny, d.nz, d.ncell, atom_cell + d.atom0,
                               rank + d.atom0, hist + d.hist0, blk - g.blk0,
                               keys);
}

__device__ __forceinline__ void hist_offsets_unit(
    uint32_t *__restrict__ hist, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg, uint32_t *__restrict__ count,
    uint32_t blk) {
  const Seg g = seg[run_of(seg, nseg, blk)];
  const SysDev &d = sys[g.sys];
  hist_offsets_body(hist + d.hist0, d.nchunk, d.ncell, count + d.cell0,
                    (blk - g.blk0) * kBatchBlock + threadIdx.x);
}

// system s; start[cell0 + ncell] is the next system's first slot, the last
// system writes the batch's end
template <uint32_t NT>
__device__ __forceinline__ void cell_scan_unit(
    const uint32_t *__restrict__ count, const SysDev *__restrict__ sys,
    uint32_t nsys, uint32_t *__restrict__ start, uint32_t *__restrict__ maxc,
    uint32_t *__restrict__ ag_start, uint2 *__restrict__ ag, uint32_t s,
    uint32_t *__restrict__ lds) {
  const SysDev &d = sys[s];
  const bool groups = ag && d.agtile;
  cell_scan_body<NT>(count + d.cell0, d.ncell, start + d.cell0, maxc + s,
                     d.atom0, s + 1 == nsys, lds,
                     groups ? ag_start + d.cell0 : nullptr,
                     groups ? ag + d.ag0 : nullptr, d.agtile * kAgPerBlock);
}

template <bool MIXED>
__device__ __forceinline__ void bin_scatter_unit(
    const double *__restrict__ rx, const double *__restrict__ ry,
    const double *__restrict__ rz, const double *__restrict__ qc,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg, uint32_t nseg,
    const uint32_t *__restrict__ atom_cell, const uint32_t *__restrict__ rank,
    const uint32_t *__restrict__ hist, const uint32_t *__restrict__ start,
    uint32_t *__restrict__ sorted_idx, uint32_t *__restrict__ cell_sorted,
    double4 *__restrict__ atoms64, float4 *__restrict__ atoms32,
    const uint32_t *__restrict__ ag_start, uint2 *__restrict__ ag,
    uint32_t blk) {
  const Seg g = seg[run_of(seg, nseg, blk)];
  const SysDev &d = sys[g.sys];
  const bool groups = ag && d.agtile;
  bin_scatter_body<MIXED>(
      rx + d.atom0, ry + d.atom0, rz + d.atom0, qc + d.atom0, d.natom,
      atom_cell + d.atom0, rank + d.atom0, hist + d.hist0, start + d.cell0,
      d.ncell, d.ny, d.nz, d.h, sorted_idx, cell_sorted, atoms64, atoms32,
      d.far_d, (blk - g.blk0) * kBatchBlock + threadIdx.x, d.atom0,
      groups ? ag_start + d.cell0 : nullptr, groups ? ag + d.ag0 : nullptr);
}

template <bool MIXED>
__device__ __forceinline__ void rmt_unit(
    const void *__restrict__ Sv, const double4 *__restrict__ k64,
    const uint32_t *__restrict__ group, const uint32_t *__restrict__ r_in,
    const uint32_t *__restrict__ r_out, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg, void *__restrict__ R,
    uint32_t blk) {
  const Seg g = seg[run_of(seg, nseg, blk)];
  const SysDev &d = sys[g.sys];
  const void *Ss;
  void *Rs;
  if (MIXED) {
    Ss = static_cast<const float2 *>(Sv) + d.field0;
    Rs = static_cast<float2 *>(R) + d.field0;
  } else {
    Ss = static_cast<const double2 *>(Sv) + d.field0;
    Rs = static_cast<double2 *>(R) + d.field0;
  }
  rmt_lds_body<MIXED>(Ss, k64 + d.node0, group + d.node0, r_in + d.group0,
                      r_out + d.group0, d.Kp, d.T, d.nx, d.ny, d.nz, d.h, Rs,
                      blk - g.blk0);
}

// a run's workgroups are (split, atom tile) of one system and alpha group
template <int NR>
__device__ __forceinline__ void lr_mixed_unit(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    const LrTasks &tk, const float2 *__restrict__ R,
    const SysDev *__restrict__ sys, const Seg &g, uint32_t blk,
    double4 *__restrict__ partial) {
  const SysDev &d = sys[g.sys];
  const uint32_t local = blk - g.blk0;
  const uint32_t split = g.a + local / d.ntile;
  lr_mixed_nr_body<NR>(atoms + d.atom0, cell_sorted + d.atom0, tk,
                       R + d.field0, d.Kp, d.natom, local % d.ntile, split,
                       split - d.split0, partial + d.partial0);
}

// a run's workgroups are (split, atom-group tile) of one system and group
template <int NR>
__device__ __forceinline__ void lr_ag_unit(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    const uint2 *__restrict__ ag, const LrTasks &tk,
    const float2 *__restrict__ R, const SysDev *__restrict__ sys, const Seg &g,
    uint32_t blk, double4 *__restrict__ partial) {
  const SysDev &d = sys[g.sys];
  const uint32_t local = blk - g.blk0;
  const uint32_t split = g.a + local / d.agtile;
  lr_ag_body<NR>(atoms + d.atom0, cell_sorted + d.atom0, ag + d.ag0, d.atom0,
                 tk, R + d.field0, d.Kp, d.natom, local % d.agtile, split,
                 split - d.split0, partial + d.partial0);
}

__device__ __forceinline__ void lr_reduce_unit(
    const double4 *__restrict__ partial, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg,
    const double4 *__restrict__ near, double *__restrict__ fx,
    double *__restrict__ fy, double *__restrict__ fz, double *__restrict__ en,
    uint32_t blk) {
  const Seg g = seg[run_of(seg, nseg, blk)];
  const SysDev &d = sys[g.sys];
  lr_reduce_body(partial + d.partial0, d.nsplit, d.natom,
                 near ? near + d.atom0 : nullptr, fx + d.atom0, fy + d.atom0,
                 fz + d.atom0, en + d.atom0,
                 (blk - g.blk0) * kBatchBlock + threadIdx.x);
}

template <bool MIXED, bool STORE>
__device__ __forceinline__ void sr_unit(
    const double4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg, uint32_t nseg,
    double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz,
    double *__restrict__ en, double4 *__restrict__ out4, uint32_t blk,
    double4 *__restrict__ tile) {
  const Seg g = seg[run_of(seg, nseg, blk)];
  const SysDev &d = sys[g.sys];
  sr_body<MIXED, STORE>(atoms, start + d.cell0, d.nx, d.ny, d.nz, fx, fy, fz,
                        en, out4, blk - g.blk0, tile);
}

template <typename CT>
__global__ __launch_bounds__(kBinChunk) void bin_chunk_batch_kernel(
    const double *__restrict__ rx, const double *__restrict__ ry,
    const double *__restrict__ rz, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg, uint32_t *__restrict__ atom_cell,
    uint32_t *__restrict__ rank, uint32_t *__restrict__ hist) {
  __shared__ unsigned long long keys[kBinChunk];
  bin_chunk_unit<CT, kBinChunk, false>(rx, ry, rz, sys, seg, nseg, atom_cell,
                                       rank, hist, blockIdx.x, keys);
}

__global__ __launch_bounds__(kBatchBlock) void hist_offsets_batch_kernel(
    uint32_t *__restrict__ hist, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg, uint32_t *__restrict__ count) {
  hist_offsets_unit(hist, sys, seg, nseg, count, blockIdx.x);
}

__global__ __launch_bounds__(1024) void cell_scan_batch_kernel(
    const uint32_t *__restrict__ count, const SysDev *__restrict__ sys,
    uint32_t nsys, uint32_t *__restrict__ start, uint32_t *__restrict__ maxc,
    uint32_t *__restrict__ ag_start, uint2 *__restrict__ ag) {
  __shared__ uint32_t lds[3 * 1024];
  cell_scan_unit<1024>(count, sys, nsys, start, maxc, ag_start, ag,
                       blockIdx.x, lds);
}

template <bool MIXED>
__global__ __launch_bounds__(kBatchBlock) void bin_scatter_batch_kernel(
    const double *__restrict__ rx, const double *__restrict__ ry,
    const double *__restrict__ rz, const double *__restrict__ qc,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg, uint32_t nseg,
    const uint32_t *__restrict__ atom_cell, const uint32_t *__restrict__ rank,
    const uint32_t *__restrict__ hist, const uint32_t *__restrict__ start,
    uint32_t *__restrict__ sorted_idx, uint32_t *__restrict__ cell_sorted,
    double4 *__restrict__ atoms64, float4 *__restrict__ atoms32,
    const uint32_t *__restrict__ ag_start, uint2 *__restrict__ ag) {
  bin_scatter_unit<MIXED>(rx, ry, rz, qc, sys, seg, nseg, atom_cell, rank,
                          hist, start, sorted_idx, cell_sorted, atoms64,
                          atoms32, ag_start, ag, blockIdx.x);
}

__global__ __launch_bounds__(kSfBlock) void sf_f64_batch_kernel(
    const double4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const double4 *__restrict__ k64, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg, double2 *__restrict__ S) {
  const Seg g = seg[run_of(seg, nseg, blockIdx.x)];
  const SysDev &d = sys[g.sys];
  const uint32_t local = blockIdx.x - g.blk0, ncol = d.nx * d.ny;
  sf_f64_body(atoms, start + d.cell0, k64 + d.node0, d.nz, d.ncell,
              S + d.field0, local % ncol,
              (local / ncol) * kSfBlock + threadIdx.x);
}

template <int NR>
__global__ __launch_bounds__(kSfDirs)
__attribute__((amdgpu_waves_per_eu(1, 10)))
void sf_mixed_batch_kernel(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ cell_sorted, LrTasks tk,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg,
    uint32_t nseg, float2 *__restrict__ S) {
  const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
  const Seg g = seg[run_of(seg, nseg, wg)];
  const SysDev &d = sys[g.sys];
  const uint32_t local = wg - g.blk0;
  const uint32_t ncg = d.ncell;
  if constexpr (NR > 0)
    sf_mixed_nr_body<NR>(atoms, start + d.cell0, cell_sorted, tk, g.a, g.b,
                         d.ncell, S + d.field0, local % ncg, local / ncg);
  else
    sf_mixed_any_body(atoms, start + d.cell0, cell_sorted, tk, g.a, g.b,
                      d.ncell, S + d.field0, local % ncg, local / ncg);
}

// The Chebyshev form over a batch's runs (system, alpha group): the group's
// matrix at coff[first task], rho and dlim from the system's cell edge.
template <int M>
__global__ __launch_bounds__(kSfDirs) void sf_cheb_batch_kernel(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ cell_sorted, LrTasks tk,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg,
    uint32_t nseg, float2 *__restrict__ S, const float *__restrict__ cheb,
    const uint32_t *__restrict__ coff) {
  const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
  const Seg g = seg[run_of(seg, nseg, wg)];
  const SysDev &d = sys[g.sys];
  const uint32_t local = wg - g.blk0;
  const uint32_t ncg = d.ncell;
  sf_cheb_body<M>(atoms, start + d.cell0, cell_sorted, tk, g.a, g.b, d.ncell,
                  S + d.field0, local % ncg, local / ncg, cheb + coff[g.a],
                  static_cast<float>(1.0 / cheb_rho(d.h)),
                  static_cast<float>(cheb_dlim(d.h)));
}

typedef void (*SfChebBatchKernel)(const float4 *, const uint32_t *,
                                  const uint32_t *, LrTasks, const SysDev *,
                                  const Seg *, uint32_t, float2 *,
                                  const float *, const uint32_t *);

// index M - 1 (M = 1 .. kChebMaxTerms)
template <int... N>
constexpr auto sf_cheb_batch_table(std::integer_sequence<int, N...>) {
  return std::array<SfChebBatchKernel, sizeof...(N)>{
      {&sf_cheb_batch_kernel<N + 1>...}};
}

const auto kSfChebBatchKernels =
    sf_cheb_batch_table(std::make_integer_sequence<int, kChebMaxTerms>{});

typedef void (*SfBatchKernel)(const float4 *, const uint32_t *,
                              const uint32_t *, LrTasks, const SysDev *,
                              const Seg *, uint32_t, float2 *);

template <int... N>
constexpr auto sf_batch_table(std::integer_sequence<int, N...>) {
  return std::array<SfBatchKernel, sizeof...(N)>{
      {&sf_mixed_batch_kernel<N>...}};
}

const auto kSfBatchKernels =
    sf_batch_table(std::make_integer_sequence<int, kSfMaxUnrolled + 1>{});

template <bool MIXED>
__global__ __launch_bounds__(kRmtBlock) void rmt_lds_batch_kernel(
    const void *__restrict__ Sv, const double4 *__restrict__ k64,
    const uint32_t *__restrict__ group, const uint32_t *__restrict__ r_in,
    const uint32_t *__restrict__ r_out, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg, void *__restrict__ R) {
  rmt_unit<MIXED>(Sv, k64, group, r_in, r_out, sys, seg, nseg, R,
                  xcd_remap(blockIdx.x, gridDim.x));
}

// a run's workgroups are the kRmtWaveNodes-node tiles of one small system
__global__ __launch_bounds__(kRmtWaveBlock, GLST_RMT_WAVE_MIN_WG) void rmt_wave_batch_kernel(
    const float2 *__restrict__ S, const double4 *__restrict__ k64,
    const uint32_t *__restrict__ group, const uint32_t *__restrict__ r_in,
    const uint32_t *__restrict__ r_out, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg, float2 *__restrict__ R) {
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const Seg g = seg[run_of(seg, nseg, blk)];
  const SysDev &d = sys[g.sys];
  rmt_wave_tile(S + d.field0, k64 + d.node0, group + d.node0,
                r_in + d.group0, r_out + d.group0, d.Kp,
                (blk - g.blk0) * kRmtWaveNodes, d.nx, d.ny, d.nz, d.h,
                R + d.field0);
}

__global__ __launch_bounds__(kLrBlock) void lr_f64_batch_kernel(
    const double4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    LrTasks tk, const double2 *__restrict__ R, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg,
    double4 *__restrict__ partial) {
  const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
  const Seg g = seg[run_of(seg, nseg, wg)];
  const SysDev &d = sys[g.sys];
  const uint32_t local = wg - g.blk0;
  const uint32_t prow = local / d.ntile;
  lr_f64_body(atoms + d.atom0, cell_sorted + d.atom0, tk, R + d.field0, d.Kp,
              d.natom, local % d.ntile, d.split0 + prow, prow,
              partial + d.partial0);
}

template <int NR>
__global__ __launch_bounds__(kLrBlock) void lr_mixed_batch_kernel(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    LrTasks tk, const float2 *__restrict__ R, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg,
    double4 *__restrict__ partial) {
  const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
  lr_mixed_unit<NR>(atoms, cell_sorted, tk, R, sys, seg[run_of(seg, nseg, wg)],
                    wg, partial);
}

typedef void (*LrBatchKernel)(const float4 *, const uint32_t *, LrTasks,
                              const float2 *, const SysDev *, const Seg *,
                              uint32_t, double4 *);

// The batched back-projection over atom groups.  No four-wave register cap
// (on the batch's mix of small systems it cost 2-3 %); the three-wave cap at
// 23 and 24 radial nodes as above.
template <int NR>
__global__ __launch_bounds__(kLrBlock, lr_ag_3wave(NR) ? 3 : 1)
void lr_ag_batch_kernel(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    const uint2 *__restrict__ ag, LrTasks tk, const float2 *__restrict__ R,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg, uint32_t nseg,
    double4 *__restrict__ partial) {
  const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
  lr_ag_unit<NR>(atoms, cell_sorted, ag, tk, R, sys,
                 seg[run_of(seg, nseg, wg)], wg, partial);
}

typedef void (*LrAgBatchKernel)(const float4 *, const uint32_t *,
                                const uint2 *, LrTasks, const float2 *,
                                const SysDev *, const Seg *, uint32_t,
                                double4 *);

template <int... N>
constexpr auto lr_ag_batch_table(std::integer_sequence<int, N...>) {
  return std::array<LrAgBatchKernel, sizeof...(N)>{
      {&lr_ag_batch_kernel<N + 1>...}};
}

// index NR - 1 (NR = 1 .. kLrMaxUnrolled)
const auto kLrAgBatchKernels =
    lr_ag_batch_table(std::make_integer_sequence<int, kLrMaxUnrolled>{});

template <int... N>
constexpr auto lr_batch_table(std::integer_sequence<int, N...>) {
  return std::array<LrBatchKernel, sizeof...(N)>{
      {&lr_mixed_batch_kernel<N>...}};
}

const auto kLrBatchKernels =
    lr_batch_table(std::make_integer_sequence<int, kLrMaxUnrolled + 1>{});

__global__ __launch_bounds__(kBatchBlock) void lr_reduce_batch_kernel(
    const double4 *__restrict__ partial, const SysDev *__restrict__ sys,
    const Seg *__restrict__ seg, uint32_t nseg,
    const double4 *__restrict__ near, double *__restrict__ fx,
    double *__restrict__ fy, double *__restrict__ fz,
    double *__restrict__ en) {
  lr_reduce_unit(partial, sys, seg, nseg, near, fx, fy, fz, en, blockIdx.x);
}

template <bool MIXED, bool STORE>
__global__ __launch_bounds__(kSrBlock) void sr_batch_kernel(
    const double4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg, uint32_t nseg,
    double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz,
    double *__restrict__ en, double4 *__restrict__ out4) {
  __shared__ double4 tile[kSrBlock];
  sr_unit<MIXED, STORE>(atoms, start, sys, seg, nseg, fx, fy, fz, en, out4,
                        xcd_remap(blockIdx.x, gridDim.x), tile);
}

// ---------------------------------------------------------------------------
// The mixed batched step as one persistent launch (FusedStep): every
// workgroup walks the units of each phase (binning chunks; histogram
// offsets; cell scans; scatter; S(k) and near field; shell sums;
// back-projection; reduction) with stride gridDim.x, and a grid-wide barrier
// separates the phases.  All gridDim.x workgroups must be resident at once
// (fused_workgroups); a barrier that waits longer than f.timeout wall-clock
// ticks marks the step failed (f.sync[1] and the host flag f.fail) and every
// workgroup leaves, so the launch always ends.

static_assert(kFusedBlock == kSrBlock && kFusedBlock == kRmtBlock &&
                  kFusedBlock == kLrBlock && kFusedBlock == kBatchBlock,
              "the fused step runs every body with one workgroup size");
static_assert(kFusedBlock % kSfDirs == 0 &&
                  kFusedSfCells == kFusedBlock / kSfDirs,
              "one S(k) wave per cell");

// Arrive at barrier n (1, 2, ...).  Every wave waits until its stores have
// reached the L2; then one thread makes the workgroup's writes visible at
// agent scope (a release fence: the XCD's L2 writes its dirty lines back),
// counts the arrival, waits for n * gridDim.x of them, and drops stale lines
// of the CU's L1 and the XCD's L2 (an acquire fence) before the workgroup
// goes on.  False: the step has failed.
__device__ __forceinline__ bool fused_barrier(uint32_t *__restrict__ sync,
                                              uint32_t *__restrict__ fail,
                                              long long timeout, uint32_t n,
                                              uint32_t *__restrict__ ok) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    atomicAdd(sync, 1u);
    const uint32_t target = n * gridDim.x;
    const long long t0 = wall_clock64();
    uint32_t good = 1u;
    while (__hip_atomic_load(sync, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(sync + 1, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT) != 0u ||
          wall_clock64() - t0 > timeout) {
        __hip_atomic_store(sync + 1, 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(fail, 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
        good = 0u;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *ok = good;
  }
  __syncthreads();
  return *ok != 0u;
}

// S(k) of unit (cell group of kFusedSfCells cells, direction chunk): the
// unrolled body for radial counts up to kFusedMaxNr, else the any-count body
template <int N>
__device__ __forceinline__ void sf_fused_nr(
    uint32_t nr, const float4 *__restrict__ atoms,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ cell_sorted,
    const LrTasks &tk, uint32_t task0, uint32_t ntask, uint32_t ncell,
    float2 *__restrict__ S, uint32_t cell0, uint32_t chunk) {
  if constexpr (N == 0) {
    sf_mixed_any_body(atoms, start, cell_sorted, tk, task0, ntask, ncell, S,
                      cell0, chunk);
  } else {
    if (nr == static_cast<uint32_t>(N))
      sf_mixed_nr_body<N>(atoms, start, cell_sorted, tk, task0, ntask, ncell,
                          S, cell0, chunk);
    else
      sf_fused_nr<N - 1>(nr, atoms, start, cell_sorted, tk, task0, ntask,
                         ncell, S, cell0, chunk);
  }
}

template <int N>
__device__ __forceinline__ void lr_fused_nr(
    uint32_t nr, const float4 *__restrict__ atoms,
    const uint32_t *__restrict__ cell_sorted, const uint2 *__restrict__ ag,
    const LrTasks &tk, const float2 *__restrict__ R,
    const SysDev *__restrict__ sys, const Seg &g, uint32_t blk,
    double4 *__restrict__ partial) {
  if constexpr (N == 0) {
    lr_mixed_unit<0>(atoms, cell_sorted, tk, R, sys, g, blk, partial);
  } else {
    if (nr == static_cast<uint32_t>(N))
      lr_ag_unit<N>(atoms, cell_sorted, ag, tk, R, sys, g, blk, partial);
    else
      lr_fused_nr<N - 1>(nr, atoms, cell_sorted, ag, tk, R, sys, g, blk,
                         partial);
  }
}

// S(k) unit u of the fused plan: kFusedSfCells cells of one direction
// chunk of one system's alpha group
__device__ __forceinline__ void sf_fused_unit(
    const float4 *__restrict__ atoms32, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ cell_sorted, const LrTasks &tk,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg,
    uint32_t nseg, float2 *__restrict__ S, uint32_t u) {
  const Seg g = seg[run_of(seg, nseg, u)];
  const SysDev &d = sys[g.sys];
  const uint32_t local = u - g.blk0;
  const uint32_t ncg = (d.ncell + kFusedSfCells - 1) / kFusedSfCells;
  const uint32_t nr = tk.nr[g.a];
  sf_fused_nr<kFusedMaxNr>(nr <= kFusedMaxNr ? nr : 0u, atoms32,
                           start + d.cell0, cell_sorted, tk, g.a, g.b, d.ncell,
                           S + d.field0, (local % ncg) * kFusedSfCells,
                           local / ncg);
}

// back-projection unit u of the fused plan
__device__ __forceinline__ void lr_fused_unit(
    const float4 *__restrict__ atoms32,
    const uint32_t *__restrict__ cell_sorted, const uint2 *__restrict__ ag,
    const LrTasks &tk, const float2 *__restrict__ R,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg,
    uint32_t nseg, double4 *__restrict__ partial, uint32_t u) {
  const Seg g = seg[run_of(seg, nseg, u)];
  const uint32_t nr = tk.nr[tk.split[g.a]];
  // the plan gives atom-group tiles exactly to these runs
  const bool unrolled = nr <= kFusedMaxNr && sys[g.sys].agtile != 0u;
  lr_fused_nr<kFusedMaxNr>(unrolled ? nr : 0u, atoms32, cell_sorted, ag, tk, R,
                           sys, g, u, partial);
}

// One launch per stage for every radial count (option "fused" 3): the fused
// plan's S(k) and back-projection units as ordinary launches.
__global__ __launch_bounds__(kFusedBlock) void sf_all_batch_kernel(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ cell_sorted, LrTasks tk,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg,
    uint32_t nseg, float2 *__restrict__ S) {
  sf_fused_unit(atoms, start, cell_sorted, tk, sys, seg, nseg, S,
                xcd_remap(blockIdx.x, gridDim.x));
}

__global__ __launch_bounds__(kFusedBlock)
__attribute__((amdgpu_waves_per_eu(3)))
void lr_all_batch_kernel(
    const float4 *__restrict__ atoms, const uint32_t *__restrict__ cell_sorted,
    const uint2 *__restrict__ ag, LrTasks tk, const float2 *__restrict__ R,
    const SysDev *__restrict__ sys, const Seg *__restrict__ seg, uint32_t nseg,
    double4 *__restrict__ partial) {
  lr_fused_unit(atoms, cell_sorted, ag, tk, R, sys, seg, nseg, partial,
                xcd_remap(blockIdx.x, gridDim.x));
}

// WAVES: waves per SIMD the registers are held to (2: no spills at
// kFusedMaxNr 28; 3: spills in the largest unrolled bodies)
template <int WAVES>
__global__ __launch_bounds__(kFusedBlock)
__attribute__((amdgpu_waves_per_eu(WAVES)))
void step_fused_kernel(FusedStep f) {
  extern __shared__ double2 fused_lds[];
  __shared__ uint32_t ok;
  const uint32_t W = gridDim.x;
  // the workgroups of one XCD take consecutive units of every phase
  const uint32_t v = xcd_remap(blockIdx.x, W);
  const LrTasks tk{f.tk.node0, f.tk.nr,      f.tk.koff,    f.tk.dir64,
                   f.tk.dir32, f.tk.kappa64, f.tk.kappa32, f.tk.split};
  uint32_t n = 0;
  auto barrier = [&] {
    return fused_barrier(f.sync, f.fail, f.timeout, ++n, &ok);
  };
  // binning chunks (each clears its own histogram row first)
  for (uint32_t u = v; u < f.bin.nblk; u += W) {
    __syncthreads();
    bin_chunk_unit<float, kFusedBlock, true>(
        f.rx, f.ry, f.rz, f.sys, f.bin.seg, f.bin.nseg, f.atom_cell, f.rank,
        f.hist, u, reinterpret_cast<unsigned long long *>(fused_lds));
  }
  if (!barrier()) return;
  for (uint32_t u = v; u < f.hist_p.nblk; u += W)
    hist_offsets_unit(f.hist, f.sys, f.hist_p.seg, f.hist_p.nseg, f.count, u);
  if (!barrier()) return;
  for (uint32_t u = v; u < f.nsys; u += W) {
    __syncthreads();
    cell_scan_unit<kFusedBlock>(f.count, f.sys, f.nsys, f.start, f.maxc,
                                f.ag_start, f.ag, u,
                                reinterpret_cast<uint32_t *>(fused_lds));
  }
  if (!barrier()) return;
  for (uint32_t u = v; u < f.scatter.nblk; u += W)
    bin_scatter_unit<true>(f.rx, f.ry, f.rz, f.qc, f.sys, f.scatter.seg,
                           f.scatter.nseg, f.atom_cell, f.rank, f.hist,
                           f.start, f.sorted_idx, f.cell_sorted, f.atoms64,
                           f.atoms32, f.ag_start, f.ag, u);
  if (!barrier()) return;
  // S(k) units first (the longest), then the near field's cells
  for (uint32_t u = v; u < f.sf.nblk + f.sr.nblk; u += W) {
    if (u < f.sf.nblk) {
      sf_fused_unit(f.atoms32, f.start, f.cell_sorted, tk, f.sys, f.sf.seg,
                    f.sf.nseg, f.S, u);
    } else {
      __syncthreads();
      sr_unit<true, true>(f.atoms64, f.start, f.sys, f.sr.seg, f.sr.nseg,
                          nullptr, nullptr, nullptr, nullptr, f.near4,
                          u - f.sf.nblk,
                          reinterpret_cast<double4 *>(fused_lds));
    }
  }
  if (!barrier()) return;
  for (uint32_t u = v; u < f.rmt.nblk; u += W) {
    __syncthreads();
    rmt_unit<true>(f.S, f.k64, f.group, f.r_in, f.r_out, f.sys, f.rmt.seg,
                   f.rmt.nseg, f.R, u);
  }
  if (!barrier()) return;
  for (uint32_t u = v; u < f.lr.nblk; u += W)
    lr_fused_unit(f.atoms32, f.cell_sorted, f.ag, tk, f.R, f.sys, f.lr.seg,
                  f.lr.nseg, f.partial, u);
  if (!barrier()) return;
  for (uint32_t u = v; u < f.reduce.nblk; u += W)
    lr_reduce_unit(f.partial, f.sys, f.reduce.seg, f.reduce.nseg, f.near4,
                   f.fx, f.fy, f.fz, f.en, u);
}

// run blockIdx.x, copied by gridDim.y workgroups
__global__ void copy_runs_kernel(const CopyRun *__restrict__ runs) {
  const CopyRun r = runs[blockIdx.x];
  const uint32_t step = gridDim.y * blockDim.x;
  for (uint32_t i = blockIdx.y * blockDim.x + threadIdx.x; i < r.n; i += step)
    r.dst[i] = r.src[i];
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers

// Compute units and LDS bytes per CU of the device (prepare_kernels); the
// defaults are MI355X's.
static uint32_t g_cu_count = 256;
static size_t g_lds_per_cu = 160 * 1024;

void prepare_kernels(size_t lds_limit) {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount,
                              dev) == hipSuccess &&
        v > 0)
      g_cu_count = static_cast<uint32_t>(v);
    if (hipDeviceGetAttribute(
            &v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) ==
            hipSuccess &&
        v > 0)
      g_lds_per_cu = static_cast<size_t>(v);
  }
  (void)hipGetLastError();
  const int lds = static_cast<int>(lds_limit);
  const void *fns[] = {reinterpret_cast<const void *>(&rmt_lds_kernel<true>),
                       reinterpret_cast<const void *>(&rmt_lds_kernel<false>),
                       reinterpret_cast<const void *>(&rmt_lds_batch_kernel<true>),
                       reinterpret_cast<const void *>(&rmt_lds_batch_kernel<false>),
                       reinterpret_cast<const void *>(&step_fused_kernel<2>),
                       reinterpret_cast<const void *>(&step_fused_kernel<3>),
                       reinterpret_cast<const void *>(&rmt_mfma_kernel<1>),
                       reinterpret_cast<const void *>(&rmt_mfma_kernel<2>),
                       reinterpret_cast<const void *>(&rmt_mfma_kernel<4>)};
  for (const void *f : fns) {
    const hipError_t e = hipFuncSetAttribute(
        f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) {
      (void)hipGetLastError();
    }
  }
}

void fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s) {
  if (n == 0) return;
  fill_u32_kernel<<<blocks_for(n, 256), 256, 0, s>>>(p, value, n);
}

void fill_f64(double *p, double value, size_t n, hipStream_t s) {
  if (n == 0) return;
  fill_f64_kernel<<<blocks_for(n, 256), 256, 0, s>>>(p, value, n);
}

void bin_chunks(bool ct_float, const double *rx, const double *ry,
                const double *rz, uint32_t natom, double inv_cell,
                const GridShape &g, uint32_t *atom_cell, uint32_t *rank,
                uint32_t *hist, hipStream_t s) {
  const uint32_t nchunk = blocks_for(natom, kBinChunk);
  if (nchunk == 0) return;
  if (ct_float)
    bin_chunk_kernel<float><<<nchunk, kBinChunk, 0, s>>>(
        rx, ry, rz, natom, static_cast<float>(inv_cell), g.nx, g.ny, g.nz,
        g.ncell, atom_cell, rank, hist);
  else
    bin_chunk_kernel<double><<<nchunk, kBinChunk, 0, s>>>(
        rx, ry, rz, natom, inv_cell, g.nx, g.ny, g.nz, g.ncell, atom_cell,
        rank, hist);
}

void hist_offsets(uint32_t *hist, uint32_t nchunk, uint32_t ncell,
                  uint32_t *count, hipStream_t s) {
  hist_offsets_kernel<<<blocks_for(ncell, 256), 256, 0, s>>>(hist, nchunk,
                                                             ncell, count);
}

void cell_scan(const uint32_t *count, uint32_t ncell, uint32_t *start,
               uint32_t *maxc, hipStream_t s, uint32_t *ag_start, uint2 *ag,
               uint32_t nag_max) {
  cell_scan_kernel<<<1, 1024, 0, s>>>(count, ncell, start, maxc,
                                      ag ? ag_start : nullptr, ag, nag_max);
}

void bin_scatter(bool mixed, const double *rx, const double *ry,
                 const double *rz, const double *qc, uint32_t natom,
                 const uint32_t *atom_cell, const uint32_t *rank,
                 const uint32_t *hist, const uint32_t *start,
                 const GridShape &g, uint32_t *sorted_idx,
                 uint32_t *cell_sorted, double4 *atoms64, float4 *atoms32,
                 float far_d, hipStream_t s, const uint32_t *ag_start,
                 uint2 *ag) {
  if (natom == 0) return;
  if (mixed)
    bin_scatter_kernel<true><<<blocks_for(natom, 256), 256, 0, s>>>(
        rx, ry, rz, qc, natom, atom_cell, rank, hist, start, g.ncell, g.ny,
        g.nz, g.h, sorted_idx, cell_sorted, atoms64, atoms32, far_d, ag_start,
        ag);
  else
    bin_scatter_kernel<false><<<blocks_for(natom, 256), 256, 0, s>>>(
        rx, ry, rz, qc, natom, atom_cell, rank, hist, start, g.ncell, g.ny,
        g.nz, g.h, sorted_idx, cell_sorted, atoms64, atoms32, far_d, ag_start,
        ag);
}

void sf(bool mixed, const void *atoms, const uint32_t *start,
        const uint32_t *cell_sorted, const double4 *k64,
        const LrTaskArgs &ta, const SfGroup *groups, size_t ngroup,
        uint32_t Kp, const GridShape &g, void *S, hipStream_t s,
        const float *cheb) {
  if (!mixed) {
    const dim3 grid(g.nx * g.ny, Kp / kSfBlock);
    sf_f64_kernel<<<grid, kSfBlock, 0, s>>>(static_cast<const double4 *>(atoms),
                                            start, k64, g.nz, g.ncell,
                                            static_cast<double2 *>(S));
    return;
  }
  const LrTasks tk{ta.node0, ta.nr,      ta.koff,   ta.dir64,
                   ta.dir32, ta.kappa64, ta.kappa32, ta.split};
  const float inv_rho = static_cast<float>(1.0 / cheb_rho(g.h));
  const float dlim = static_cast<float>(cheb_dlim(g.h));
  for (size_t q = 0; q < ngroup; ++q) {
    const SfGroup &gr = groups[q];
    if (gr.ntask == 0) continue;
    const dim3 grid(g.ncell, (gr.ntask + kSfDirs - 1) / kSfDirs);
    if (cheb && gr.m >= 1 && gr.m <= kChebMaxTerms) {
      kSfChebKernels[gr.m - 1]<<<grid, kSfDirs, 0, s>>>(
          static_cast<const float4 *>(atoms), start, cell_sorted, tk,
          gr.task0, gr.ntask, g.ncell, static_cast<float2 *>(S),
          cheb + gr.coff, inv_rho, dlim);
      continue;
    }
    const int nr = gr.nr <= static_cast<uint32_t>(kSfMaxUnrolled)
                       ? static_cast<int>(gr.nr)
                       : 0;
    kSfNrKernels[nr]<<<grid, kSfDirs, 0, s>>>(
        static_cast<const float4 *>(atoms), start, cell_sorted, tk, gr.task0,
        gr.ntask, g.ncell, static_cast<float2 *>(S));
  }
}

bool rmt_lds(bool mixed, uint32_t tmax, const void *S, const double4 *k64,
             const uint32_t *group, const uint32_t *r_in,
             const uint32_t *r_out, uint32_t Kp, const GridShape &g,
             size_t lds_limit, void *R, hipStream_t s, uint32_t tile0,
             uint32_t tile1) {
  const uint32_t T = rmt_nodes_per_group(g.nx, g.ny, g.nz, lds_limit, tmax);
  if (T == 0) return false;
  const uint32_t all = Kp / T;  // Kp is a multiple of 256
  if (tile1 > all) tile1 = all;
  if (tile0 >= tile1) return true;
  const uint32_t ntile = tile1 - tile0;
  const size_t lds = rmt_lds_bytes(g.nx, g.ny, g.nz, T);
  // persistent workgroups: as many as are resident at once (LDS allows one
  // or two kRmtPBlock-thread workgroups per CU)
  size_t per_cu = g_lds_per_cu / lds;
  const size_t wave_cap = 2048 / kRmtPBlock;  // 32 waves per CU
  per_cu = per_cu < 1 ? 1 : (per_cu > wave_cap ? wave_cap : per_cu);
  const uint32_t nwg = static_cast<uint32_t>(
      std::min<size_t>(ntile, static_cast<size_t>(g_cu_count) * per_cu));
  if (mixed) {
    rmt_lds_kernel<true><<<nwg, kRmtPBlock, lds, s>>>(
        S, k64, group, r_in, r_out, Kp, T, tile0, ntile, g.nx, g.ny, g.nz,
        g.h, R);
  } else {
    rmt_lds_kernel<false><<<nwg, kRmtPBlock, lds, s>>>(
        S, k64, group, r_in, r_out, Kp, T, tile0, ntile, g.nx, g.ny, g.nz,
        g.h, R);
  }
  return true;
}

// Nodes per tile of the matrix-core shell sums: two (two workgroups of 63 KB
// share a CU at 15^3) where they fit, else one.
#ifndef GLST_RMT_MFMA_NPW
#define GLST_RMT_MFMA_NPW 2
#endif
uint32_t rmt_mfma_nodes(const GridShape &g, size_t lds_limit) {
  if (g.nx > kRmtMfmaMaxXY || g.ny > kRmtMfmaMaxXY || g.ncell == 0) return 0;
  for (uint32_t npw = GLST_RMT_MFMA_NPW; npw >= 1; npw >>= 1)
    if (rmt_mfma_lds_bytes(g.nx, g.ny, g.nz, npw) <= lds_limit) return npw;
  return 0;
}

bool rmt_mfma(const void *S, const double4 *k64, const uint32_t *group,
              const uint32_t *r_in, const uint32_t *r_out, uint32_t Kp,
              const GridShape &g, size_t lds_limit, void *R, hipStream_t s,
              uint32_t tile0, uint32_t tile1) {
  const uint32_t npw = rmt_mfma_nodes(g, lds_limit);
  if (npw == 0) return false;
  const uint32_t all = Kp / npw;  // Kp is a multiple of 256
  if (tile1 > all) tile1 = all;
  if (tile0 >= tile1) return true;
  const uint32_t ntile = tile1 - tile0;
  const size_t lds = rmt_mfma_lds_bytes(g.nx, g.ny, g.nz, npw);
  const float2 *Sf = static_cast<const float2 *>(S);
  float2 *Rf = static_cast<float2 *>(R);
  if (npw == 4)
    rmt_mfma_kernel<4><<<ntile, kRmtMfmaBlock, lds, s>>>(
        Sf, k64, group, r_in, r_out, Kp, tile0, ntile, g.nx, g.ny, g.nz, g.h,
        Rf);
  else if (npw == 2)
    rmt_mfma_kernel<2><<<ntile, kRmtMfmaBlock, lds, s>>>(
        Sf, k64, group, r_in, r_out, Kp, tile0, ntile, g.nx, g.ny, g.nz, g.h,
        Rf);
  else
    rmt_mfma_kernel<1><<<ntile, kRmtMfmaBlock, lds, s>>>(
        Sf, k64, group, r_in, r_out, Kp, tile0, ntile, g.nx, g.ny, g.nz, g.h,
        Rf);
  return true;
}

bool rmt_wave(const void *S, const double4 *k64, const uint32_t *group,
              const uint32_t *r_in, const uint32_t *r_out, uint32_t Kp,
              const GridShape &g, void *R, hipStream_t s, uint32_t tile0,
              uint32_t tile1) {
  if (!rmt_wave_fits(g)) return false;
  const uint32_t all = Kp / kRmtWaveNodes;  // Kp is a multiple of 256
  if (tile1 > all) tile1 = all;
  if (tile0 >= tile1) return true;
  const uint32_t ntile = tile1 - tile0;
  rmt_wave_kernel<<<ntile, kRmtWaveBlock, 0, s>>>(
      static_cast<const float2 *>(S), k64, group, r_in, r_out, Kp, tile0,
      ntile, g.nx, g.ny, g.nz, g.h, static_cast<float2 *>(R));
  return true;
}

void rmt_global(bool mixed, const void *S, double2 *G, const double4 *k64,
                const uint32_t *group, const uint32_t *r_in,
                const uint32_t *r_out, uint32_t Kp, const GridShape &g,
                void *R, hipStream_t s) {
  const uint32_t bx = blocks_for(Kp, 256);
  constexpr uint32_t kMaxY = 65535;  // grid.y limit
  const uint32_t ncol = g.nx * g.ny;
  for (uint32_t c0 = 0; c0 < ncol; c0 += kMaxY) {
    const uint32_t cnt = (ncol - c0 < kMaxY) ? ncol - c0 : kMaxY;
    if (mixed)
      rmt_global_load_kernel<true><<<dim3(bx, cnt), 256, 0, s>>>(
          S, k64, G, Kp, c0, g.nx, g.ny, g.nz, g.h);
    else
      rmt_global_load_kernel<false><<<dim3(bx, cnt), 256, 0, s>>>(
          S, k64, G, Kp, c0, g.nx, g.ny, g.nz, g.h);
  }
  auto prefix = [&](uint32_t nlines, uint32_t n, uint32_t stride,
                    uint32_t inner) {
    for (uint32_t l0 = 0; l0 < nlines; l0 += kMaxY) {
      const uint32_t cnt = (nlines - l0 < kMaxY) ? nlines - l0 : kMaxY;
      prefix_axis_kernel<<<dim3(bx, cnt), 256, 0, s>>>(G, Kp, l0, nlines, n,
                                                       stride, inner);
    }
  };
  prefix(g.nx * g.nz, g.ny, g.nz, g.nz);                // y lines
  prefix(g.ny * g.nz, g.nx, g.ny * g.nz, g.ny * g.nz);  // x lines
  for (uint32_t c0 = 0; c0 < g.ncell; c0 += kMaxY) {
    const uint32_t cnt = (g.ncell - c0 < kMaxY) ? g.ncell - c0 : kMaxY;
    if (mixed)
      rmt_gather_kernel<true><<<dim3(bx, cnt), 256, 0, s>>>(
          G, k64, group, r_in, r_out, Kp, c0, g.nx, g.ny, g.nz, g.h, R);
    else
      rmt_gather_kernel<false><<<dim3(bx, cnt), 256, 0, s>>>(
          G, k64, group, r_in, r_out, Kp, c0, g.nx, g.ny, g.nz, g.h, R);
  }
}

uint32_t atom_groups_max(uint32_t natom, uint32_t ncell) {
  const uint64_t n =
      (static_cast<uint64_t>(natom) + (kAgAtoms - 1ull) * ncell) / kAgAtoms;
  // whole workgroups of groups
  return static_cast<uint32_t>((n + kAgPerBlock) / kAgPerBlock * kAgPerBlock);
}

void lr(bool mixed, const void *atoms, const uint32_t *cell_sorted,
        const LrTaskArgs &ta, const LrGroup *groups, size_t ngroup,
        const void *R, uint32_t Kp, uint32_t natom, uint32_t nsplit,
        double4 *partial, hipStream_t s, const uint2 *ag, uint32_t nag_max) {
  if (natom == 0 || nsplit == 0) return;
  const uint32_t ntile = blocks_for(natom, kLrBlock);
  const dim3 grid(ntile, nsplit);
  const LrTasks tk{ta.node0, ta.nr, ta.koff, ta.dir64, ta.dir32,
                   ta.kappa64, ta.kappa32, ta.split};
  if (mixed) {
    for (size_t q = 0; q < ngroup; ++q) {
      const LrGroup &gr = groups[q];
      if (gr.nsplit == 0) continue;
      if (ag && gr.nr >= 1 && gr.nr <= static_cast<uint32_t>(kLrMaxUnrolled)) {
        const uint32_t ngt = nag_max / kAgPerBlock;
        kLrAgKernels[gr.nr - 1]<<<dim3(ngt, gr.nsplit), kLrBlock, 0, s>>>(
            static_cast<const float4 *>(atoms), cell_sorted, ag, tk,
            static_cast<const float2 *>(R), Kp, natom, ngt, gr.split0,
            partial);
        continue;
      }
      const LrNrKernel k =
          kLrNrKernels[gr.nr <= static_cast<uint32_t>(kLrMaxUnrolled) ? gr.nr : 0];
      k<<<dim3(ntile, gr.nsplit), kLrBlock, 0, s>>>(
          static_cast<const float4 *>(atoms), cell_sorted, tk,
          static_cast<const float2 *>(R), Kp, natom, ntile, gr.split0,
          partial);
    }
  } else
    lr_f64_kernel<<<grid, kLrBlock, 0, s>>>(
        static_cast<const double4 *>(atoms), cell_sorted, tk,
        static_cast<const double2 *>(R), Kp, natom, ntile, partial);
}

void lr_reduce(const double4 *partial, uint32_t nsplit, uint32_t natom,
               const double4 *near, double *fx, double *fy, double *fz,
               double *en, hipStream_t s) {
  if (natom == 0) return;
  lr_reduce_kernel<<<blocks_for(natom, 256), 256, 0, s>>>(
      partial, nsplit, natom, near, fx, fy, fz, en);
}

void sr(bool mixed, const double4 *atoms, const uint32_t *start,
        const GridShape &g, double *fx, double *fy, double *fz, double *en,
        double4 *out4, hipStream_t s) {
  const uint32_t nb = g.ncell;
  if (out4) {
    if (mixed)
      sr_kernel<true, true><<<nb, kSrBlock, 0, s>>>(atoms, start, g.nx, g.ny,
                                                    g.nz, fx, fy, fz, en, out4);
    else
      sr_kernel<false, true><<<nb, kSrBlock, 0, s>>>(
          atoms, start, g.nx, g.ny, g.nz, fx, fy, fz, en, out4);
  } else {
    if (mixed)
      sr_kernel<true, false><<<nb, kSrBlock, 0, s>>>(
          atoms, start, g.nx, g.ny, g.nz, fx, fy, fz, en, out4);
    else
      sr_kernel<false, false><<<nb, kSrBlock, 0, s>>>(
          atoms, start, g.nx, g.ny, g.nz, fx, fy, fz, en, out4);
  }
}

void unpermute(const uint32_t *sorted_idx, uint32_t natom, const double *fx,
               const double *fy, const double *fz, const double *en,
               double *ofx, double *ofy, double *ofz, double *oen,
               hipStream_t s) {
  if (natom == 0) return;
  unpermute_kernel<<<blocks_for(natom, 256), 256, 0, s>>>(
      sorted_idx, natom, fx, fy, fz, en, ofx, ofy, ofz, oen);
}

CoulombPlan coulomb_plan(uint32_t natom, size_t max_entries) {
  CoulombPlan p{1, natom, 0};
  if (natom == 0) return p;
  const uint32_t blocks = blocks_for(natom, kCoulBlock);
  uint32_t nsplit = (kCoulTargetBlocks + blocks - 1) / blocks;
  nsplit = std::min<uint32_t>(nsplit, kCoulMaxSplits);
  nsplit = static_cast<uint32_t>(
      std::min<size_t>(nsplit, std::max<size_t>(1, max_entries / natom)));
  if (nsplit <= 1) return p;
  // source ranges of whole tiles; recount so that no split is empty
  const uint32_t per = (natom + nsplit - 1) / nsplit;
  p.jchunk = (per + kCoulBlock - 1) / kCoulBlock * kCoulBlock;
  p.nsplit = (natom + p.jchunk - 1) / p.jchunk;
  if (p.nsplit <= 1) return CoulombPlan{1, natom, 0};
  p.entries = static_cast<size_t>(p.nsplit) * natom;
  return p;
}

// Partial sums of the allocation-free direct sum: one array per GPU in the
// library's code object, held until the process ends.  Host code never
// passes `scratch` itself to a kernel (in host code the name is a stand-in,
// not a device address); it passes the address hipGetSymbolAddress returns
// for the current device.
static __device__ __align__(256) double4 scratch[kCoulScratchEntries];

double4 *coulomb_scratch() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double4 *addr[kMaxDevices] = {};
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 ||
      device >= kMaxDevices)
    return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess)
      return nullptr;
    addr[device] = static_cast<double4 *>(p);
  }
  return addr[device];
}

void coulomb_direct(const double *rx, const double *ry, const double *rz,
                    const double *qc, uint32_t natom, const CoulombPlan &p,
                    double4 *part, double *fx, double *fy, double *fz,
                    double *en, hipStream_t s) {
  if (natom == 0) return;
  const dim3 grid(blocks_for(natom, kCoulBlock), p.nsplit);
  coulomb_kernel<<<grid, kCoulBlock, 0, s>>>(rx, ry, rz, qc, natom, p.jchunk,
                                             fx, fy, fz, en,
                                             p.nsplit > 1 ? part : nullptr);
  if (p.nsplit > 1)
    coulomb_reduce_kernel<<<blocks_for(natom, 256), 256, 0, s>>>(
        part, p.nsplit, natom, fx, fy, fz, en);
}

// ---- batched launchers -----------------------------------------------------

void bin_chunks_batch(bool ct_float, const double *rx, const double *ry,
                      const double *rz, const SysDev *sys,
                      const BatchLaunch &L, uint32_t *atom_cell,
                      uint32_t *rank, uint32_t *hist, hipStream_t s) {
  if (L.nblk == 0) return;
  if (ct_float)
    bin_chunk_batch_kernel<float><<<L.nblk, kBinChunk, 0, s>>>(
        rx, ry, rz, sys, L.seg, L.nseg, atom_cell, rank, hist);
  else
    bin_chunk_batch_kernel<double><<<L.nblk, kBinChunk, 0, s>>>(
        rx, ry, rz, sys, L.seg, L.nseg, atom_cell, rank, hist);
}

void hist_offsets_batch(uint32_t *hist, const SysDev *sys,
                        const BatchLaunch &L, uint32_t *count, hipStream_t s) {
  if (L.nblk == 0) return;
  hist_offsets_batch_kernel<<<L.nblk, 256, 0, s>>>(hist, sys, L.seg, L.nseg,
                                                   count);
}

void cell_scan_batch(const uint32_t *count, const SysDev *sys, uint32_t nsys,
                     uint32_t *start, uint32_t *maxc, hipStream_t s,
                     uint32_t *ag_start, uint2 *ag) {
  if (nsys == 0) return;
  cell_scan_batch_kernel<<<nsys, 1024, 0, s>>>(count, sys, nsys, start, maxc,
                                               ag ? ag_start : nullptr, ag);
}

void bin_scatter_batch(bool mixed, const double *rx, const double *ry,
                       const double *rz, const double *qc, const SysDev *sys,
                       const BatchLaunch &L, const uint32_t *atom_cell,
                       const uint32_t *rank, const uint32_t *hist,
                       const uint32_t *start, uint32_t *sorted_idx,
                       uint32_t *cell_sorted, double4 *atoms64,
                       float4 *atoms32, hipStream_t s,
                       const uint32_t *ag_start, uint2 *ag) {
  if (L.nblk == 0) return;
  if (!ag) ag_start = nullptr;
  if (mixed)
    bin_scatter_batch_kernel<true><<<L.nblk, 256, 0, s>>>(
        rx, ry, rz, qc, sys, L.seg, L.nseg, atom_cell, rank, hist, start,
        sorted_idx, cell_sorted, atoms64, atoms32, ag_start, ag);
  else
    bin_scatter_batch_kernel<false><<<L.nblk, 256, 0, s>>>(
        rx, ry, rz, qc, sys, L.seg, L.nseg, atom_cell, rank, hist, start,
        sorted_idx, cell_sorted, atoms64, atoms32, ag_start, ag);
}

void sf_batch(bool mixed, const void *atoms, const uint32_t *start,
              const uint32_t *cell_sorted, const double4 *k64,
              const LrTaskArgs &ta, const SysDev *sys,
              const std::vector<BatchLaunch> &launches, void *S,
              hipStream_t s, const float *cheb, const uint32_t *coff) {
  const LrTasks tk{ta.node0, ta.nr,      ta.koff,   ta.dir64,
                   ta.dir32, ta.kappa64, ta.kappa32, ta.split};
  for (const BatchLaunch &L : launches) {
    if (L.nblk == 0) continue;
    if (!mixed) {
      sf_f64_batch_kernel<<<L.nblk, kSfBlock, 0, s>>>(
          static_cast<const double4 *>(atoms), start, k64, sys, L.seg, L.nseg,
          static_cast<double2 *>(S));
      continue;
    }
    if (L.nr > kSfChebKey) {
      const uint32_t M = L.nr - kSfChebKey;
      if (M > kChebMaxTerms || !cheb || !coff)
        throw std::runtime_error("sf_batch: no Chebyshev kernel for this launch");
      kSfChebBatchKernels[M - 1]<<<L.nblk, kSfDirs, 0, s>>>(
          static_cast<const float4 *>(atoms), start, cell_sorted, tk, sys,
          L.seg, L.nseg, static_cast<float2 *>(S), cheb, coff);
      continue;
    }
    const int nr = L.nr <= static_cast<uint32_t>(kSfMaxUnrolled)
                       ? static_cast<int>(L.nr)
                       : 0;
    kSfBatchKernels[nr]<<<L.nblk, kSfDirs, 0, s>>>(static_cast<const float4 *>(atoms), start,
                               cell_sorted, tk, sys, L.seg, L.nseg,
                               static_cast<float2 *>(S));
  }
}

void rmt_batch(bool mixed, const void *S, const double4 *k64,
               const uint32_t *group, const uint32_t *r_in,
               const uint32_t *r_out, const SysDev *sys, const BatchLaunch &L,
               size_t lds_bytes, void *R, hipStream_t s) {
  if (L.nblk == 0) return;
  if (mixed) {
    rmt_lds_batch_kernel<true><<<L.nblk, kRmtBlock, lds_bytes, s>>>(
        S, k64, group, r_in, r_out, sys, L.seg, L.nseg, R);
  } else {
    rmt_lds_batch_kernel<false><<<L.nblk, kRmtBlock, lds_bytes, s>>>(
        S, k64, group, r_in, r_out, sys, L.seg, L.nseg, R);
  }
}

void rmt_wave_batch(const void *S, const double4 *k64, const uint32_t *group,
                    const uint32_t *r_in, const uint32_t *r_out,
                    const SysDev *sys, const BatchLaunch &L, void *R,
                    hipStream_t s) {
  if (L.nblk == 0) return;
  rmt_wave_batch_kernel<<<L.nblk, kRmtWaveBlock, 0, s>>>(
      static_cast<const float2 *>(S), k64, group, r_in, r_out, sys, L.seg,
      L.nseg, static_cast<float2 *>(R));
}

void lr_batch(bool mixed, const void *atoms, const uint32_t *cell_sorted,
              const LrTaskArgs &ta, const SysDev *sys,
              const std::vector<BatchLaunch> &launches, const void *R,
              double4 *partial, hipStream_t s, const uint2 *ag) {
  const LrTasks tk{ta.node0, ta.nr,      ta.koff,   ta.dir64,
                   ta.dir32, ta.kappa64, ta.kappa32, ta.split};
  for (const BatchLaunch &L : launches) {
    if (L.nblk == 0) continue;
    if (!mixed) {
      lr_f64_batch_kernel<<<L.nblk, kLrBlock, 0, s>>>(
          static_cast<const double4 *>(atoms), cell_sorted, tk,
          static_cast<const double2 *>(R), sys, L.seg, L.nseg, partial);
      continue;
    }
    if (ag && L.nr >= 1 && L.nr <= static_cast<uint32_t>(kLrMaxUnrolled)) {
      kLrAgBatchKernels[L.nr - 1]<<<L.nblk, kLrBlock, 0, s>>>(
          static_cast<const float4 *>(atoms), cell_sorted, ag, tk,
          static_cast<const float2 *>(R), sys, L.seg, L.nseg, partial);
      continue;
    }
    kLrBatchKernels[L.nr <= static_cast<uint32_t>(kLrMaxUnrolled) ? L.nr : 0]
        <<<L.nblk, kLrBlock, 0, s>>>(static_cast<const float4 *>(atoms),
                                     cell_sorted, tk,
                                     static_cast<const float2 *>(R), sys,
                                     L.seg, L.nseg, partial);
  }
}

void lr_reduce_batch(const double4 *partial, const SysDev *sys,
                     const BatchLaunch &L, const double4 *near, double *fx,
                     double *fy, double *fz, double *en, hipStream_t s) {
  if (L.nblk == 0) return;
  lr_reduce_batch_kernel<<<L.nblk, 256, 0, s>>>(partial, sys, L.seg, L.nseg,
                                                near, fx, fy, fz, en);
}

void sr_batch(bool mixed, const double4 *atoms, const uint32_t *start,
              const SysDev *sys, const BatchLaunch &L, double *fx, double *fy,
              double *fz, double *en, double4 *out4, hipStream_t s) {
  if (L.nblk == 0) return;
  if (out4) {
    if (mixed)
      sr_batch_kernel<true, true><<<L.nblk, kSrBlock, 0, s>>>(
          atoms, start, sys, L.seg, L.nseg, fx, fy, fz, en, out4);
    else
      sr_batch_kernel<false, true><<<L.nblk, kSrBlock, 0, s>>>(
          atoms, start, sys, L.seg, L.nseg, fx, fy, fz, en, out4);
  } else {
    if (mixed)
      sr_batch_kernel<true, false><<<L.nblk, kSrBlock, 0, s>>>(
          atoms, start, sys, L.seg, L.nseg, fx, fy, fz, en, out4);
    else
      sr_batch_kernel<false, false><<<L.nblk, kSrBlock, 0, s>>>(
          atoms, start, sys, L.seg, L.nseg, fx, fy, fz, en, out4);
  }
}

size_t fused_lds_bytes(size_t rmt_lds) {
  // binning keys, cell-scan words and near-field tiles share the space
  const size_t other = std::max<size_t>(
      {kBinChunk * sizeof(unsigned long long), 3 * kFusedBlock * 4,
       kSrBlock * sizeof(double4)});
  return std::max(rmt_lds, other);
}

static const void *fused_kernel(int variant) {
  return variant == 2 ? reinterpret_cast<const void *>(&step_fused_kernel<3>)
                      : reinterpret_cast<const void *>(&step_fused_kernel<2>);
}

uint32_t fused_workgroups(size_t lds, int variant) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, fused_kernel(variant), static_cast<int>(kFusedBlock),
          lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return per_cu > 0 ? static_cast<uint32_t>(per_cu) * g_cu_count : 0u;
}

long long wall_clock_rate() {
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) !=
          hipSuccess ||
      khz <= 0) {
    (void)hipGetLastError();
    return 100000000LL;  // 100 MHz
  }
  return static_cast<long long>(khz) * 1000;
}

void sf_all_batch(const void *atoms, const uint32_t *start,
                  const uint32_t *cell_sorted, const LrTaskArgs &ta,
                  const SysDev *sys, const BatchLaunch &L, void *S,
                  hipStream_t s) {
  if (L.nblk == 0) return;
  const LrTasks tk{ta.node0, ta.nr,      ta.koff,   ta.dir64,
                   ta.dir32, ta.kappa64, ta.kappa32, ta.split};
  sf_all_batch_kernel<<<L.nblk, kFusedBlock, 0, s>>>(
      static_cast<const float4 *>(atoms), start, cell_sorted, tk, sys, L.seg,
      L.nseg, static_cast<float2 *>(S));
}

void lr_all_batch(const void *atoms, const uint32_t *cell_sorted,
                  const LrTaskArgs &ta, const SysDev *sys,
                  const BatchLaunch &L, const void *R, double4 *partial,
                  hipStream_t s, const uint2 *ag) {
  if (L.nblk == 0) return;
  const LrTasks tk{ta.node0, ta.nr,      ta.koff,   ta.dir64,
                   ta.dir32, ta.kappa64, ta.kappa32, ta.split};
  lr_all_batch_kernel<<<L.nblk, kFusedBlock, 0, s>>>(
      static_cast<const float4 *>(atoms), cell_sorted, ag, tk,
      static_cast<const float2 *>(R), sys, L.seg, L.nseg, partial);
}

void step_fused(const FusedStep &f, uint32_t nwg, size_t lds, hipStream_t s,
                int variant) {
  // a fill kernel, not a memset: the step may be captured in a graph
  fill_u32(f.sync, 0u, 2, s);
  if (variant == 2)
    step_fused_kernel<3><<<nwg, kFusedBlock, lds, s>>>(f);
  else
    step_fused_kernel<2><<<nwg, kFusedBlock, lds, s>>>(f);
}

void copy_runs(const CopyRun *runs, uint32_t nruns, uint32_t max_n,
               hipStream_t s) {
  if (nruns == 0) return;
  // about four elements per thread, at most 64 workgroups per run
  const uint32_t per = blocks_for(max_n, 256 * 4);
  const uint32_t wgs = per < 1 ? 1 : (per > 64 ? 64 : per);
  copy_runs_kernel<<<dim3(nruns, wgs), 256, 0, s>>>(runs);
}

}  // namespace dev
}  // namespace glst
