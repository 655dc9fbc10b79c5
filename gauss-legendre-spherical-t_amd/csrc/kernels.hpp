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
// start[0..ncell] = exclusive scan of count; *maxc = max count.
void cell_scan(const uint32_t *count, uint32_t ncell, uint32_t *start,
               uint32_t *maxc, hipStream_t s);
// Scatter atoms to cell-sorted order; writes sorted_idx, the sorted atoms'
// cells and the sorted atom records: double4 absolute (x,y,z,q) always, and
// for the mixed path float4 cell-centre-relative (x,y,z,q).
void bin_scatter(bool mixed, const double *rx, const double *ry,
                 const double *rz, const double *qc, uint32_t natom,
                 const uint32_t *atom_cell, const uint32_t *rank,
                 const uint32_t *hist, const uint32_t *start,
                 const GridShape &g, uint32_t *sorted_idx,
                 uint32_t *cell_sorted, double4 *atoms64, float4 *atoms32,
                 hipStream_t s);

// ---- structure factors (calc_sf) -------------------------------------------
// S(k) is stored in node-pair tiles: the slots of nodes 2m and 2m+1 of all
// cells form one contiguous run, cell-major inside it, so the shell-sum
// workgroup of a node pair reads one contiguous block.
__host__ __device__ inline size_t sf_slot(uint32_t node, uint32_t cell,
                                          uint32_t ncell) {
  return ((static_cast<size_t>(node >> 1) * ncell + cell) << 1) + (node & 1);
}

struct LrTaskArgs;
// Tasks (directions) of one alpha group for the mixed structure factors.
struct SfGroup {
  uint32_t nr, task0, ntask;
};
// FP64: S[sf_slot(k, c)] (double2) = running sum along z of
// sum_{j in cell} q_j exp(-i k.r_j).  Mixed: S[sf_slot(k, c)] (float2) =
// sum_{j in c} q_j exp(-i k.d_j) with d_j the position relative to the cell
// centre; one lane per direction of an alpha group, the group's radii
// unrolled (padded node columns are not written).
void sf(bool mixed, const void *atoms, const uint32_t *start,
        const double4 *k64, const LrTaskArgs &tasks,
        const std::vector<SfGroup> &groups, uint32_t Kp, const GridShape &g,
        void *S, hipStream_t s);

// ---- remote (shell) sums (sum_rmt_sf) --------------------------------------
// R_c(k) = W_k * sum_{c' : r_in < cheb(c,c') <= r_out} S_c'(k); the mixed
// path stores exp(+i k.center_c) R_c(k) as float2, the FP64 path R_c(k) as
// double2, both at R[c*Kp + k].  LDS-resident tables of tmax (1 or 2) nodes
// per workgroup; returns false when the table of one node does not fit in
// lds_limit bytes (the caller then uses rmt_global).
bool rmt_lds(bool mixed, uint32_t tmax, const void *S, const double4 *k64,
             const uint32_t *group, const uint32_t *r_in,
             const uint32_t *r_out, uint32_t Kp, const GridShape &g,
             size_t lds_limit, void *R, hipStream_t s);
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
void lr(bool mixed, const void *atoms, const uint32_t *cell_sorted,
        const LrTaskArgs &tasks, const std::vector<LrGroup> &groups,
        const void *R, uint32_t Kp, uint32_t natom,
        uint32_t nsplit, double4 *partial, hipStream_t s);
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

// ---- direct all-pairs Coulomb (utils/coulomb.cu) -----------------------------
void coulomb_direct(const double *rx, const double *ry, const double *rz,
                    const double *qc, uint32_t natom, double *fx, double *fy,
                    double *fz, double *en, hipStream_t s);

}  // namespace dev
}  // namespace glst
