// Launchers for the GLST HIP kernels (gfx950).  Every launcher enqueues work
// on `stream` and returns immediately; shapes are validated by the caller
// (engine.hip) before any launch.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace glst {
namespace dev {

// Atoms per workgroup of the chunked stable counting sort.
constexpr uint32_t kBinChunk = 1024;

struct GridShape {
  uint32_t nx, ny, nz, ncell;
  double h;  // cell edge (double of the CT value)
};

// Once per process and device, before any launch: lets the LDS-resident
// shell-sum kernels use up to lds_limit bytes of dynamic LDS (set outside
// the launches so that a step can be captured in a HIP graph).
void prepare_kernels(size_t lds_limit);

// ---- binning and stable counting sort (assign_atoms) ----------------------
void fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s);
void fill_f64(double *p, double value, size_t n, hipStream_t s);

// Cell of every atom (CT arithmetic of src/glst_force.cu:356-386), the
// stable rank of each atom among the atoms of its chunk with the same cell,
// and the per-chunk cell histogram hist[chunk][cell] (must be zeroed).
void bin_chunks(bool ct_float, const double *rx, const double *ry,
                const double *rz, uint32_t natom, double inv_cell,
                const GridShape &g, uint32_t *atom_cell, uint32_t *rank,
                uint32_t *hist, hipStream_t s);
// hist[chunk][cell] -> exclusive offsets over chunks; count[cell] = total.
void hist_offsets(uint32_t *hist, uint32_t nchunk, uint32_t ncell,
                  uint32_t *count, hipStream_t s);
// start[0..ncell] = exclusive scan of count; *maxc = max count.  With ag
// (mixed path): ag_start[c] = first atom group of cell c, and ag entries past
// the last group up to nag_max = (0, 0).
void cell_scan(const uint32_t *count, uint32_t ncell, uint32_t *start,
               uint32_t *maxc, hipStream_t s, uint32_t *ag_start = nullptr,
               uint2 *ag = nullptr, uint32_t nag_max = 0);
// Scatter atoms to cell-sorted order; writes sorted_idx, the sorted atoms'
// cells and the sorted atom records: double4 absolute (x,y,z,q) always, and
// for the mixed path float4 cell-centre-relative (x,y,z,q).  With ag: the
// (first atom, count) entry of every atom group, from cell_scan's ag_start.
// The mixed path's FP32 offsets from the cell centre carry phases k.d of at
// most kappa_max |d|_1 revolutions; the hardware sine and cosine take
// |t| <= 256 (measured accurate to 64).  An atom whose |d|_1 exceeds far_d
// (63 / kappa_max, set by the caller; only atoms far outside the grid, which
// the binning clamps into its edge cells) gets kFarAtom set in its
// cell_sorted entry, and S(k) and the back-projection reduce its phases
// with V_FRACT first.  Readers of cell_sorted mask the bit.
constexpr uint32_t kFarAtom = 1u << 31;
void bin_scatter(bool mixed, const double *rx, const double *ry,
                 const double *rz, const double *qc, uint32_t natom,
                 const uint32_t *atom_cell, const uint32_t *rank,
                 const uint32_t *hist, const uint32_t *start,
                 const GridShape &g, uint32_t *sorted_idx,
                 uint32_t *cell_sorted, double4 *atoms64, float4 *atoms32,
                 float far_d, hipStream_t s,
                 const uint32_t *ag_start = nullptr, uint2 *ag = nullptr);

// ---- structure factors (calc_sf) -------------------------------------------
// S(k) is stored in node runs: the slots of nodes G m .. G m + G - 1 of
// every cell form one contiguous run, cell-major inside it (mixed: float2
// slots, G = 4; FP64: double2, G = 2; 32-byte runs, one HBM sector).  The
// wave that computes a cell's sums writes whole runs, so fewer lines leave
// L2 partly written (mixed S(k) at 288k atoms writes 1.39 times its bytes
// with G = 4, 2.4 with 2, 1.13 with 8).  A shell-sum tile of T < G nodes
// reads part of each run and its neighbours the rest while the lines are in
// L2; G = 8 made that cost 5 % on the 18^3 grid (T = 1), G = 4 nothing.
#ifndef GLST_SF_RUN_MIXED
#define GLST_SF_RUN_MIXED 4
#endif
template <bool MIXED>
__host__ __device__ constexpr uint32_t sf_run() {
  return MIXED ? GLST_SF_RUN_MIXED : 2u;
}
template <bool MIXED>
__host__ __device__ inline size_t sf_slot(uint32_t node, uint32_t cell,
                                          uint32_t ncell) {
  constexpr uint32_t G = sf_run<MIXED>();
  return (static_cast<size_t>(node / G) * ncell + cell) * G + node % G;
}

// Mixed structure factors: one wave per (cell, 64 directions of a group), the
// NR radial nodes of a direction unrolled for NR <= kSfMaxUnrolled (larger
// counts run a generic kernel); each lane stores its direction's NR slots.
constexpr int kSfMaxUnrolled = 40;

struct LrTaskArgs;
// Chebyshev form of the mixed structure factors (plan_chebyshev in
// runtime.hpp): a cell's atoms within dlim = (1 + kChebMargin) h / 2 of its
// centre in every axis have |s.d| <= rho = sqrt(3) dlim.
constexpr double kChebMargin = 1e-3;
constexpr uint32_t kChebMaxTerms = 40;
__host__ __device__ inline double cheb_dlim(double h) {
  return (1.0 + kChebMargin) * 0.5 * h;
}
__host__ __device__ inline double cheb_rho(double h) {
  return 1.7320508075688772 * cheb_dlim(h);
}
// Batched S(k) launches in Chebyshev form carry kSfChebKey + M in
// BatchLaunch::nr (direct sums: the radial count, at most 40).
constexpr uint32_t kSfChebKey = 64;

// Tasks (directions) of one alpha group for the mixed structure factors; m
// Chebyshev terms (0: the direct sum) with their matrix at coff of the
// Chebyshev table.
struct SfGroup {
  uint32_t nr, task0, ntask;
  uint32_t m = 0, coff = 0;
};
// FP64: S[sf_slot(k, c)] (double2) = running sum along z of
// sum_{j in cell} q_j exp(-i k.r_j).  Mixed: S[sf_slot(k, c)] (float2) =
// sum_{j in c} q_j exp(-i k.d_j) with d_j the position relative to the cell
// centre; one lane per direction of an alpha group, the group's radii
// unrolled (padded node columns are not written).
// The mixed path launches once per listed group (groups[0..ngroup)); the
// FP64 path ignores the list and covers every node in one launch.  Mixed
// groups with m > 0 take the Chebyshev form (cheb: the device table of
// plan_chebyshev for this grid's cell edge); a cell holding an atom more
// than half a cell edge (with the margin) from its centre, or a far atom,
// takes the direct sum.
void sf(bool mixed, const void *atoms, const uint32_t *start,
        const uint32_t *cell_sorted, const double4 *k64,
        const LrTaskArgs &tasks, const SfGroup *groups, size_t ngroup,
        uint32_t Kp, const GridShape &g, void *S, hipStream_t s,
        const float *cheb = nullptr);
inline void sf(bool mixed, const void *atoms, const uint32_t *start,
               const uint32_t *cell_sorted, const double4 *k64,
               const LrTaskArgs &tasks, const std::vector<SfGroup> &groups,
               uint32_t Kp, const GridShape &g, void *S, hipStream_t s,
               const float *cheb = nullptr) {
  sf(mixed, atoms, start, cell_sorted, k64, tasks, groups.data(),
     groups.size(), Kp, g, S, s, cheb);
}

// ---- remote (shell) sums (sum_rmt_sf) --------------------------------------
// LDS layout of the shell-sum workgroups.  A workgroup holds the summed-
// volume tables of T nodes (T = 1 or a power of two), T tables of
// rmt_table_stride elements (double2) plus two buffers of T * (nx+ny+nz)
// phase factors (the persistent kernel fills one while it reads the other).
// A table has a zero plane before the first cell of each axis and z-lines
// of odd pitch: (nx+1) x (ny+1) x rmt_zpitch(nz) elements.  An LDS cycle
// serves 16 consecutive lanes of a 16-byte access, conflict-free when their
// elements differ mod 16 (or coincide).  Lanes run over (node, then a line
// coordinate padded to a multiple of 16/T), so the stride is that size
// rounded up to 16/T mod 16 (1 from T = 16 on): the T nodes then take
// disjoint slots and the 16/T line positions fill the rest.
__host__ __device__ inline uint32_t rmt_zpitch(uint32_t nz) {
  return (nz + 1) | 1u;
}
__host__ __device__ inline uint32_t rmt_table_elems(uint32_t nx, uint32_t ny,
                                                    uint32_t nz) {
  return (nx + 1) * (ny + 1) * rmt_zpitch(nz);
}
__host__ __device__ inline uint32_t rmt_table_stride(uint32_t elems,
                                                     uint32_t T) {
  if (T <= 1) return elems;
  const uint32_t r = T >= 16 ? 1u : 16u / T;
  return elems + ((r - (elems & 15u)) & 15u);
}
// Lanes per node group of an LDS cycle: line coordinates are padded to a
// multiple of this.
__host__ __device__ inline uint32_t rmt_line_group(uint32_t T) {
  return T >= 16 ? 1u : 16u / T;
}
inline size_t rmt_lds_bytes(uint32_t nx, uint32_t ny, uint32_t nz,
                            uint32_t T) {
  return static_cast<size_t>(T) *
         (rmt_table_stride(rmt_table_elems(nx, ny, nz), T) +
          2 * (nx + ny + nz)) *
         16;
}
// Largest T <= tmax (a power of two) whose tables fit lds_limit bytes; 0 when
// not even one node's table fits.
inline uint32_t rmt_nodes_per_group(uint32_t nx, uint32_t ny, uint32_t nz,
                                    size_t lds_limit, uint32_t tmax) {
  for (uint32_t T = tmax; T >= 2; T >>= 1)
    if (rmt_lds_bytes(nx, ny, nz, T) <= lds_limit) return T;
  return rmt_lds_bytes(nx, ny, nz, 1) <= lds_limit ? 1u : 0u;
}
constexpr uint32_t kRmtMaxNodes = 64;

// Workgroup tiles [tile0, tile1) of T nodes each (T from
// rmt_nodes_per_group with this tmax; tile1 is clamped to Kp / T), so that
// the shell sums of a node range can start as soon as its S is complete.
// False: one node's tables do not fit in LDS (nothing launched).
bool rmt_lds(bool mixed, uint32_t tmax, const void *S, const double4 *k64,
             const uint32_t *group, const uint32_t *r_in,
             const uint32_t *r_out, uint32_t Kp, const GridShape &g,
             size_t lds_limit, void *R, hipStream_t s, uint32_t tile0 = 0,
             uint32_t tile1 = 0xffffffffu);
// Shell sums on the matrix cores (CT = float, nx and ny up to 16):
// workgroups of kRmtMfmaBlock threads take tiles of NPW consecutive nodes
// and hold NPW z-running-sum tables of nx * ny * rmt_zpitch(nz) float2, the
// nodes' nx + ny + nz axis factors and nx ny line factors, in LDS.
constexpr uint32_t kRmtMfmaBlock = 256;
constexpr uint32_t kRmtMfmaMaxXY = 16;
inline size_t rmt_mfma_lds_bytes(uint32_t nx, uint32_t ny, uint32_t nz,
                                 uint32_t npw) {
  const size_t nline = static_cast<size_t>(nx) * ny;
  return npw * (nline * (rmt_zpitch(nz) + 1) + nx + ny + nz) * 8;
}
// Nodes per workgroup tile of the matrix-core shell sums on this grid (0:
// the grid is too wide for them, or one node's table does not fit).
uint32_t rmt_mfma_nodes(const GridShape &g, size_t lds_limit);
// Tiles [tile0, tile1) of rmt_mfma_nodes() nodes (tile1 clamped to
// Kp / nodes).  False: the grid is not supported (nothing launched).
bool rmt_mfma(const void *S, const double4 *k64, const uint32_t *group,
              const uint32_t *r_in, const uint32_t *r_out, uint32_t Kp,
              const GridShape &g, size_t lds_limit, void *R, hipStream_t s,
              uint32_t tile0 = 0, uint32_t tile1 = 0xffffffffu);
// Shell sums of small grids (CT = float, every dimension up to
// kRmtWaveMaxDim): workgroups of kRmtWaveBlock threads take tiles of
// kRmtWaveNodes consecutive nodes, up to 16 at a time, and sum them as
// three passes of FP32 window sums along lines of at most 8 cells, with the
// lines exchanged through LDS (static, 35 KB per workgroup).
constexpr uint32_t kRmtWaveMaxDim = 8;
constexpr uint32_t kRmtWaveNodes = 16;
constexpr uint32_t kRmtWaveBlock = 256;
inline bool rmt_wave_fits(const GridShape &g) {
  return g.ncell > 0 && g.nx <= kRmtWaveMaxDim && g.ny <= kRmtWaveMaxDim &&
         g.nz <= kRmtWaveMaxDim;
}
// Tiles [tile0, tile1) of kRmtWaveNodes nodes (tile1 clamped to
// Kp / kRmtWaveNodes).  False: the grid does not fit (nothing launched).
bool rmt_wave(const void *S, const double4 *k64, const uint32_t *group,
              const uint32_t *r_in, const uint32_t *r_out, uint32_t Kp,
              const GridShape &g, void *R, hipStream_t s, uint32_t tile0 = 0,
              uint32_t tile1 = 0xffffffffu);
// Global-memory tables: G (ncell*Kp double2, scratch) receives S in FP64
// cell-major order, is prefix-summed in place, and the shells are gathered
// from it.  S is left unchanged.
void rmt_global(bool mixed, const void *S, double2 *G, const double4 *k64,
                const uint32_t *group, const uint32_t *r_in,
                const uint32_t *r_out, uint32_t Kp, const GridShape &g,
                void *R, hipStream_t s);

// ---- long-range back-projection (calc_lr_ef) -------------------------------
// Task table of the direction-major node order: task t covers nodes
// node0[t] .. node0[t]+nr[t]-1, all along direction dir[t], with radial
// magnitudes kappa[koff[t] ..]; split s owns tasks split[s] .. split[s+1]-1.
struct LrTaskArgs {
  const uint32_t *node0, *nr, *koff;
  const double4 *dir64;
  const float4 *dir32;
  const double *kappa64;
  const float *kappa32;
  const uint32_t *split;
};
// The splits of one alpha group: splits split0 .. split0+nsplit-1, all of
// whose tasks have nr radial nodes.
struct LrGroup {
  uint32_t nr, split0, nsplit;
};
// partial[split*natom + i] = (fx, fy, fz, en) of sorted atom i over the
// split's tasks.
// The mixed path launches once per listed group; the FP64 path covers
// every split in one launch.
// ag (mixed path, optional): the atom groups of atom_groups(), nag_max of
// them; groups with an unrolled radial count then use the atom-group kernel.
void lr(bool mixed, const void *atoms, const uint32_t *cell_sorted,
        const LrTaskArgs &tasks, const LrGroup *groups, size_t ngroup,
        const void *R, uint32_t Kp, uint32_t natom, uint32_t nsplit,
        double4 *partial, hipStream_t s, const uint2 *ag = nullptr,
        uint32_t nag_max = 0);
// Atom groups of the mixed back-projection: up to 4 consecutive sorted atoms
// of one cell, written by cell_scan and bin_scatter.  atom_groups_max bounds
// their number (a multiple of the kAgGroupsPerBlock groups of one
// workgroup).
constexpr uint32_t kAgGroupsPerBlock = 32;
uint32_t atom_groups_max(uint32_t natom, uint32_t ncell);
inline void lr(bool mixed, const void *atoms, const uint32_t *cell_sorted,
               const LrTaskArgs &tasks, const std::vector<LrGroup> &groups,
               const void *R, uint32_t Kp, uint32_t natom, uint32_t nsplit,
               double4 *partial, hipStream_t s, const uint2 *ag = nullptr,
               uint32_t nag_max = 0) {
  lr(mixed, atoms, cell_sorted, tasks, groups.data(), groups.size(), R, Kp,
     natom, nsplit, partial, s, ag, nag_max);
}
// f[i] += sum over splits (fixed order) + near[i] (when near != nullptr).
void lr_reduce(const double4 *partial, uint32_t nsplit, uint32_t natom,
               const double4 *near, double *fx, double *fy, double *fz,
               double *en, hipStream_t s);

// ---- near field (calc_sr_ef) -----------------------------------------------
// Accumulates into fx..en, or (out4 != nullptr) stores (fx, fy, fz, en) per
// sorted atom into out4.
void sr(bool mixed, const double4 *atoms, const uint32_t *start,
        const GridShape &g, double *fx, double *fy, double *fz, double *en,
        double4 *out4, hipStream_t s);

// ---- result order (get_ef) ---------------------------------------------------
void unpermute(const uint32_t *sorted_idx, uint32_t natom, const double *fx,
               const double *fy, const double *fz, const double *en,
               double *ofx, double *ofy, double *ofz, double *oen,
               hipStream_t s);

// ---- batched systems ----------------------------------------------------------
// Many independent systems in one set of launches.  Every system owns a
// contiguous range of the batch's atoms, cells, nodes, tasks, splits, S/R
// slots, histogram and partial rows; a batched launch covers all systems,
// each workgroup finding its system through a run table (Seg).

// One system of a batch as the batched kernels see it.
struct SysDev {
  uint32_t natom, atom0;              // atoms [atom0, atom0 + natom)
  uint32_t nx, ny, nz, ncell, cell0;  // cells [cell0, cell0 + ncell)
  uint32_t nchunk;                    // binning chunks of kBinChunk atoms
  uint32_t Kp, node0;                 // nodes [node0, node0 + Kp)
  uint32_t group0;                    // alpha groups at r_in/r_out[group0..]
  uint32_t ntile;                     // back-projection atom tiles (256)
  uint32_t nsplit, split0;            // splits [split0, split0 + nsplit)
  uint32_t T;                         // shell-sum nodes per workgroup
  uint32_t ag0, agtile;               // mixed: atom groups [ag0, + agtile
                                      // * kAgGroupsPerBlock), agtile WGs
  float far_d;                        // mixed: far-atom offset (kFarAtom)
  uint64_t hist0;                     // histogram [nchunk][ncell]
  uint64_t field0;                    // S and R slots [field0, + ncell*Kp)
  uint64_t partial0;                  // partial rows [nsplit][natom]
  double h, inv_cell;                 // cell edge, CT(1)/CT(rcut) widened
};

// A run of a batched launch's workgroups that belongs to one system (and,
// for launches per radial count, one alpha group): workgroups blk0 up to the
// next run's blk0.  a, b: launch-specific (S(k): first task and task count
// of the group; back-projection: first split and split count of the group).
struct Seg {
  uint32_t blk0, sys, a, b;
};

// One launch of a batched stage: device run table, runs, workgroups, and
// for S(k) / back-projection the radial count of its groups.
struct BatchLaunch {
  const Seg *seg;
  uint32_t nseg, nblk, nr;
};

void bin_chunks_batch(bool ct_float, const double *rx, const double *ry,
                      const double *rz, const SysDev *sys,
                      const BatchLaunch &L, uint32_t *atom_cell,
                      uint32_t *rank, uint32_t *hist, hipStream_t s);
void hist_offsets_batch(uint32_t *hist, const SysDev *sys,
                        const BatchLaunch &L, uint32_t *count, hipStream_t s);
// one workgroup per system; maxc[s] = largest cell of system s; with ag:
// the atom groups of every system with agtile > 0 (ag_start per cell)
void cell_scan_batch(const uint32_t *count, const SysDev *sys, uint32_t nsys,
                     uint32_t *start, uint32_t *maxc, hipStream_t s,
                     uint32_t *ag_start = nullptr, uint2 *ag = nullptr);
void bin_scatter_batch(bool mixed, const double *rx, const double *ry,
                       const double *rz, const double *qc, const SysDev *sys,
                       const BatchLaunch &L, const uint32_t *atom_cell,
                       const uint32_t *rank, const uint32_t *hist,
                       const uint32_t *start, uint32_t *sorted_idx,
                       uint32_t *cell_sorted, double4 *atoms64,
                       float4 *atoms32, hipStream_t s,
                       const uint32_t *ag_start = nullptr,
                       uint2 *ag = nullptr);
// mixed: one launch per radial count (L.nr), or per Chebyshev term count
// (L.nr = kSfChebKey + M: cheb holds the shapes' matrices, coff per task the
// offset of its group's); FP64: one launch (L.nr unused)
void sf_batch(bool mixed, const void *atoms, const uint32_t *start,
              const uint32_t *cell_sorted, const double4 *k64,
              const LrTaskArgs &tasks, const SysDev *sys,
              const std::vector<BatchLaunch> &launches, void *S,
              hipStream_t s, const float *cheb = nullptr,
              const uint32_t *coff = nullptr);
void rmt_batch(bool mixed, const void *S, const double4 *k64,
               const uint32_t *group, const uint32_t *r_in,
               const uint32_t *r_out, const SysDev *sys, const BatchLaunch &L,
               size_t lds_bytes, void *R, hipStream_t s);
// mixed systems whose grids fit rmt_wave: kRmtWaveNodes-node tiles
void rmt_wave_batch(const void *S, const double4 *k64, const uint32_t *group,
                    const uint32_t *r_in, const uint32_t *r_out,
                    const SysDev *sys, const BatchLaunch &L, void *R,
                    hipStream_t s);
// mixed with ag: launches with 1 <= nr <= 40 run over atom groups (their
// plans count agtile workgroups per split, the others ntile)
void lr_batch(bool mixed, const void *atoms, const uint32_t *cell_sorted,
              const LrTaskArgs &tasks, const SysDev *sys,
              const std::vector<BatchLaunch> &launches, const void *R,
              double4 *partial, hipStream_t s, const uint2 *ag = nullptr);
void lr_reduce_batch(const double4 *partial, const SysDev *sys,
                     const BatchLaunch &L, const double4 *near, double *fx,
                     double *fy, double *fz, double *en, hipStream_t s);
void sr_batch(bool mixed, const double4 *atoms, const uint32_t *start,
              const SysDev *sys, const BatchLaunch &L, double *fx, double *fy,
              double *fz, double *en, double4 *out4, hipStream_t s);

// ---- the mixed batched step as one persistent launch ---------------------------
// Every phase of the step (binning, histogram offsets, cell scans, scatter,
// S(k) with the near field, shell sums, back-projection, reduction) runs in
// one kernel of gridDim.x resident workgroups, each walking the phase's
// units with stride gridDim.x; grid-wide barriers separate the phases.
// Units are the workgroups of the per-stage launches except for S(k), whose
// units hold kFusedSfCells cells (one wave each) of one direction chunk.
// S(k) and the back-projection of radial counts up to kFusedMaxNr run the
// unrolled bodies (over atom groups), others the any-count bodies; the plan
// of `lr` counts agtile workgroups per split exactly for the unrolled runs.
constexpr uint32_t kFusedBlock = 256;
constexpr uint32_t kFusedSfCells = 4;
#ifndef GLST_FUSED_MAX_NR
#define GLST_FUSED_MAX_NR 28
#endif
constexpr int kFusedMaxNr = GLST_FUSED_MAX_NR;
struct FusedStep {
  const double *rx, *ry, *rz, *qc;
  const SysDev *sys;
  uint32_t nsys;
  BatchLaunch bin, hist_p, scatter, sf, sr, rmt, lr, reduce;
  uint32_t *atom_cell, *rank, *hist, *count, *start, *maxc;
  uint32_t *sorted_idx, *cell_sorted;
  double4 *atoms64;
  float4 *atoms32;
  uint32_t *ag_start;
  uint2 *ag;
  float2 *S, *R;
  double4 *partial, *near4;
  double *fx, *fy, *fz, *en;
  const double4 *k64;
  const uint32_t *group, *r_in, *r_out;
  LrTaskArgs tk;
  uint32_t *sync;     // [0] barrier arrivals, [1] failure (cleared at launch)
  uint32_t *fail;     // host-mapped failure flag (set on a barrier time-out)
  long long timeout;  // wall-clock ticks a barrier may wait
};
// The fused plan's S(k) and back-projection units (FusedStep::sf, lr) as
// one ordinary launch per stage for every radial count (mixed path).
void sf_all_batch(const void *atoms, const uint32_t *start,
                  const uint32_t *cell_sorted, const LrTaskArgs &tasks,
                  const SysDev *sys, const BatchLaunch &L, void *S,
                  hipStream_t s);
void lr_all_batch(const void *atoms, const uint32_t *cell_sorted,
                  const LrTaskArgs &tasks, const SysDev *sys,
                  const BatchLaunch &L, const void *R, double4 *partial,
                  hipStream_t s, const uint2 *ag);
// Dynamic LDS of the fused step for a batch whose shell sums need
// rmt_lds bytes.
size_t fused_lds_bytes(size_t rmt_lds);
// Variant 1 holds the kernel to 2 waves per SIMD (252 VGPRs, no spills),
// variant 2 to 3 waves (168 VGPRs; the largest unrolled bodies spill).
// Workgroups of the fused step that are resident at once on the current
// device with this dynamic LDS (0 if the query fails).
uint32_t fused_workgroups(size_t lds, int variant);
// Wall-clock ticks per second of the current device.
long long wall_clock_rate();
// Clears f.sync and launches the step with nwg workgroups (at most
// fused_workgroups(lds, variant), less what other persistent launches that
// may run beside it hold).
void step_fused(const FusedStep &f, uint32_t nwg, size_t lds, hipStream_t s,
                int variant);

// ---- copies of many arrays in one launch (serving loop) -----------------------
struct CopyRun {
  const double *src;
  double *dst;
  uint32_t n;
};
// dst[i] = src[i], i < n, for every run (one workgroup per run)
void copy_runs(const CopyRun *runs, uint32_t nruns, uint32_t max_n,
               hipStream_t s);

// ---- direct all-pairs Coulomb (utils/coulomb.cu) -----------------------------
// The source atoms are split into `nsplit` ranges of `jchunk` (a multiple of
// the 256-atom tile) when the atoms alone give too few workgroups; each
// split's partial sums take one double4 per atom (`entries` in all), added
// in split order.
constexpr uint32_t kCoulTargetBlocks = 2048;  // workgroups aimed for
constexpr uint32_t kCoulMaxSplits = 64;
struct CoulombPlan {
  uint32_t nsplit;  // source ranges (1: no partial sums)
  uint32_t jchunk;  // atoms per source range
  size_t entries;   // double4 partial sums needed (0 when nsplit == 1)
};
// at most `max_entries` partial sums
CoulombPlan coulomb_plan(uint32_t natom, size_t max_entries);
// The library's fixed partial-sum array of the allocation-free direct sum
// (32 MiB per GPU): its device address on the current device, looked up
// once per device; nullptr if the lookup fails.
constexpr size_t kCoulScratchEntries = size_t(1) << 20;
double4 *coulomb_scratch();
void coulomb_direct(const double *rx, const double *ry, const double *rz,
                    const double *qc, uint32_t natom, const CoulombPlan &p,
                    double4 *part, double *fx, double *fy, double *fz,
                    double *en, hipStream_t s);

}  // namespace dev
}  // namespace glst
