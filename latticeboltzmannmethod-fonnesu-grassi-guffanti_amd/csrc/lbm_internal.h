// Internal interface between the C ABI (lbm_capi.hip) and the gfx950
// kernels (lbm_kernels.hip).  Not installed; the public interface is
// include/lbm_amd.h.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/lbm_amd.h"

namespace lbm {

enum Storage : int32_t { kF64 = 0, kF32 = 1, kF16 = 2 };
enum Policy : int32_t { kVelocity = 0, kPressure = 1, kNoop = 2 };
enum EdgeKind : uint8_t { kEdgePlain = 0, kEdgeInlet = 1, kEdgeOutlet = 2 };

// Rows handled by one work-group: four wave64s stacked along the
// contiguous (row) axis of one column.
constexpr int kBlockRows = 256;

// Device-side description of the lattice.
//  * Populations are stored for wet nodes only, in column-major order (rows
//    ascending within a column), 9 * `plane` elements per buffer (`plane` a
//    multiple of 64), in one of two layouts:
//      blocked: the 9 directions of compact nodes 64k .. 64k+63 together,
//               (c, d) at 64*(9k + d) + c%64 -- every lattice whose buffer
//               stays below 4 GiB, unless the context asks for planes;
//      planes:  (c, d) at d*plane + c -- larger lattices (up to 3 planes
//               below 4 GiB).
//  * Wet mask: one bit per node, 64-bit words along rows, one guard word
//    above and below every column and one guard column on each side, all
//    zero: word (j, k) at wet[(j+1)*wl + (k+1)], bit i%64 of k = i/64.
//  * sb (same layout as wet): compact index of the first wet node of
//    (column j, word k); node (i, j) is at sb + popcount of the wet bits of
//    its word below bit i%64.
//  * ldr = NR rounded up to 64 (the row tiling of the kernels).
//  * Edge slots: every perimeter node owns slot e (top row e = j, bottom
//    row NC + j, left column 2NC + i, right column 2NC + NR + i); they hold
//    the inlet/outlet kind, the prescribed values and the frozen
//    populations whose source lies outside the domain.
struct Geometry {
    int32_t nr, nc;
    int64_t ldr;
    int64_t plane;
    const uint64_t *wet;
    const int32_t *sb;
    int64_t wl;
    int64_t E;
    const uint8_t *edge_kind;
};

struct StepArgs {
    Geometry g;
    const void *fin;      // state read by the step (F* of the previous step, or P)
    void *fout;           // F* written by the step
    // Direction base pointers of fin / fout (direction d at fin + d * 64
    // elements blocked, d * plane in planes), passed explicitly so the
    // kernels address every population as a scalar base plus a 32-bit
    // vector offset: compact node c of any direction at byte offset
    // (c + 8 (c & block_mask)) * esize from the direction's base.
    const void *fin_plane[9];
    void *fout_plane[9];
    uint32_t block_mask;    // ~63 blocked, 0 in planes
    uint32_t two_planes;    // bytes from the base of direction d to that of d + 2
    uint32_t launch_lds;    // host side: dynamic LDS per work-group of the step launches (occupancy_lds)
    const void *frozen;   // [9][E] populations with out-of-domain source, storage type
    const double *bc_prev;  // [3][E]: ux, uy, rho prescribed at the previous step
    const double *bc_cur;   // [3][E]: values prescribed at this step
    int32_t inlet_policy, outlet_policy;
    int32_t collision;      // 0 BGK, 1 TRT
    int32_t segs;           // row segments (of kBlockRows) per column
    int32_t esize;          // bytes per stored population
    int32_t wpc;            // step kernel: waves per column
    int64_t nblocks;
    double bgk_t, bgk_omt;              // 1/tau, 1 - 1/tau
    double trt_inv_te, trt_a, trt_b;    // 1/tau+, 0.5/tau+, 0.5/tau-
    float sh_a, sh_b;                   // shifted paths: 0.5/tau+, 0.5/tau- (BGK: both 0.5/tau)
    float scale, inv_scale;             // FP16 storage scale
    int32_t f16_sr;                     // FP16 storage rounds stochastically
    double *macro;          // optional [3][NC][NR]: rho, ux, uy
    int32_t *flag;          // bit 0: NaN density seen
    // Wet-run table of the run kernel (kRunLaunch): per run, 4 int32 --
    // column, first row, compact index of the first row, wet rows (<= 64,
    // all within 128 rows of the first and within one 64-node block of the
    // compact index); only rows of segments the perimeter kernel does not
    // take.
    const int32_t *runs;
    int64_t nruns;
    // Perimeter items of the perimeter kernel: per 64-row segment holding a
    // wet perimeter node (the segments the step kernel skips), 2 int32 --
    // column and segment index.
    const int32_t *perim;
    int64_t nperim;
    // Tabled wet runs (kRunTableLaunch): the same runs, per run 8 int32 --
    // column, first row, compact index of the first row, wet rows, the
    // compact indices of the first row's left and right neighbour rows
    // (rank in columns j-1 and j+1), 2 unused -- and per run 64 lane
    // records (uint32): row - first row (bits 0-6), wet rows of column j-1
    // from the first row up to the lane's row (7-13), the same for column
    // j+1 (14-20), and the lane's source-is-wet bits of directions 1..8
    // (21-28).
    const int32_t *trun;
    const uint32_t *tlane;
    // Partially saturated bounce-back of the wet obstacles (FP64 storage
    // only; every node then runs through the general path).  For a wet
    // node n and direction d whose source n - c_d is a wet obstacle, the
    // population entering n at the next step is computed by n itself at
    // the end of this step (PSBounceBackPolicy.hpp:177-297) and kept in
    // ps_val instead of taking n's own opposite slot.
    int32_t psbb;
    double ps_t;             // 1 / tau of the policy
    const uint16_t *ps_mask; // per wet node (compact index): bit d set when n - c_d is a wet obstacle
    const double *ps_b;      // [9][plane]: weight b of that obstacle
    double *ps_val;          // [9][plane]: populations entering n through those slots (read, then rewritten)
};

// Launch shape value selecting the run kernel: one wave per run of up to 64
// consecutive wet rows of a column instead of R 64-row segments.
constexpr int32_t kRunLaunch = LBM_LAUNCH_WET_RUNS;
// The same runs with each lane's addressing read from a table.
constexpr int32_t kRunTableLaunch = LBM_LAUNCH_RUN_TABLE;

// Fills fin_plane / fout_plane and two_planes from fin / fout, the layout
// (block_mask), the plane and the element size.
void set_planes(StepArgs &a);

// Dynamic LDS per work-group that leaves room for exactly `groups_per_cu`
// work-groups (2..7) in a compute unit's 160 KiB, or 0 (no limit).  The
// step kernels' own LDS (at most 1 KiB) does not change the count.
constexpr uint32_t occupancy_lds(int32_t groups_per_cu)
{
    return groups_per_cu >= 2 && groups_per_cu <= 7 ? 163840u / (uint32_t)(groups_per_cu + 1) + 64u : 0u;
}

// A batch step: per sample two complete argument sets, args[2 s + p] for
// buffer parity p (p = 0: input in the sample's current buffer).
struct BatchArgs {
    const StepArgs *args;
    const int32_t *tiles;  // per tile: sample, column, first row (R segments of 64 rows)
    int64_t ntiles;
    const int32_t *ptiles;  // perimeter table: segments holding a wet perimeter node (64 rows each)
    int64_t nptiles;
    int32_t parity;
    int32_t direct;  // set by launch_batch_step: one wave per tile instead of a persistent grid
    // wet-run table of the ensemble (launch shape kRunLaunch): per run, 5
    // int32 -- sample, column, first row, compact index of the first row,
    // wet rows (each sample's own run table, as run_kernel reads it)
    const int32_t *runs;
    int64_t nruns;
};

// Per-sample FP64 sums (readout's five) of a batch's current state:
// out[5 s + k].  partials needs 5 * ntiles doubles; tile_begin has
// nsamples + 1 entries (sample s owns tiles [tile_begin[s], tile_begin[s+1])).
hipError_t launch_batch_reduce(int32_t storage, bool gather, int32_t R, const BatchArgs &b,
                               const int64_t *tile_begin, int64_t nsamples, double *partials, double *out,
                               hipStream_t s);

// Segments of 64 rows per wave the step kernels use by default.
int32_t default_segments(int32_t storage);
// Occupancy of the step kernels until one is set or tuned (work-groups per CU).
int32_t default_occupancy(int32_t storage);

// Kernel launchers (lbm_kernels.hip).  All are asynchronous on `s`.
// R: 64-row segments per wave (1, 2, 4 or 8; <= 0 = default_segments).
hipError_t launch_step(int32_t storage, bool gather, const StepArgs &a, hipStream_t s, int32_t R = 0);
// One time step of every lattice of a batch: on a persistent grid of
// `blocks` work-groups (rounded up to a multiple of 8), or with one wave per
// tile when blocks <= 0; R = kRunLaunch: one wave per entry of the batch's
// wet-run table (the collide-only first step then takes the tiles, built
// for R = 1).
hipError_t launch_batch_step(int32_t storage, bool gather, int32_t R, const BatchArgs &b, int64_t blocks,
                             hipStream_t s);
// P (or F* when gather is false) of every node in the reference layout
// (NR*NC*9 doubles, zeros on non-wet nodes); optional FP64 partial sums
// per block (5 each) for lbm_reduce.
hipError_t launch_readout(int32_t storage, bool gather, const StepArgs &a, double *out, double *partials,
                          hipStream_t s);
hipError_t launch_upload(int32_t storage, const StepArgs &a, const double *src, void *dst, void *frozen,
                         hipStream_t s);
hipError_t launch_init_eq(int32_t storage, const StepArgs &a, void *dst, void *frozen, double *macro,
                          hipStream_t s);
hipError_t launch_lift_drag(int32_t storage, const StepArgs &a, const void *fstar, const int64_t *ij,
                            const int32_t *masks, int64_t n, double *per_point, double *partials,
                            int64_t *nblocks_out, hipStream_t s);
hipError_t launch_reduce_partials(const double *partials, int64_t nblocks, int32_t ncomp, double *out,
                                  hipStream_t s);
int64_t readout_blocks(const StepArgs &a);
// rho = 1, u = 0 over a field of n nodes ([3][n] doubles).
hipError_t launch_fill_macro(double *macro, int64_t n, hipStream_t s);

}  // namespace lbm
