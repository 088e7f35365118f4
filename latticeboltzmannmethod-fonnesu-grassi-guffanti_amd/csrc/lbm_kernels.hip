// gfx950 kernels of the D2Q9 engine.
//
// One fused kernel per time step, pull scheme, A-B buffers holding the
// post-collision populations F* of the previous step.  For a wet node x
// and direction d with source s = x - c_d (c_d = the push offset of d in
// the reference's stream, BGKCollision.hpp:179-187):
//   s wet            -> P_d(x) = F*_d(s)                 (stream)
//   s in the domain,
//   not wet          -> P_d(x) = F*_opp(d)(x)            (half-way bounce-back,
//                                                         BounceBackPolicy.hpp:68-187)
//   s outside        -> P_d(x) = frozen value, or the Zou-He closure on a
//                       velocity/pressure boundary node (ZouHePolicy.hpp:58-123)
// then the moments, the equilibrium and the BGK/TRT collision of the same
// node, written to F*(x).  Each wet slot has exactly one writer in the
// reference's push order (its source node), so the gather reproduces the
// push + bounce-back + Zou-He sequence of Lattice::perform_lbm
// (Lattice.hpp:351-410) slot for slot.
//
// Storage is compact: only wet nodes are stored, in column-major order
// (rows ascending within a column).  A node's index is the base of its
// (column, 64-row segment) plus the number of wet rows above it in the
// segment -- a popcount of the bit-packed wet mask below the lane
// (v_mbcnt).  Neighbours one row up/down are the previous/next compact
// entries, so every source index follows from three bases and three
// popcounts per wave column.  A wave then touches only wet bytes.
//
// The FP64 path evaluates every expression in the reference's order with
// no contraction (the library is built with -ffp-contract=off), so it
// matches the Serial policy bit for bit.  The shifted paths store
// g = f - w (FP32) or (f - w) * scale (FP16) and compute in FP32.
#include "lbm_internal.h"

namespace lbm {

namespace {

// Push offsets of direction d (rows, columns), BGKCollision.hpp:179-187.
__device__ constexpr int kOffI[9] = {0, 0, -1, 0, 1, -1, -1, 1, 1};
__device__ constexpr int kOffJ[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
__device__ constexpr int kOpp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
// Lattice weights (DefaultEquilibrium.hpp:68-76).
__device__ constexpr double kW[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

// FP32 weights as an unevaluated sum hi + lo (48 significant bits).  The
// rounded weights alone all carry the same relative error, +7.45e-9 (4/9,
// 1/9 and 1/36 share one significand), which would give every FP32
// equilibrium 1 + 7.45e-9 times the node's mass and momentum: a mass
// source of the same sign at every node and step.
constexpr float w_hi(double w) { return (float)w; }
constexpr float w_lo(double w) { return (float)(w - (double)(float)w); }
__device__ constexpr float kWh[9] = {w_hi(4.0 / 9.0), w_hi(1.0 / 9.0), w_hi(1.0 / 9.0), w_hi(1.0 / 9.0),
                                     w_hi(1.0 / 9.0), w_hi(1.0 / 36.0), w_hi(1.0 / 36.0), w_hi(1.0 / 36.0),
                                     w_hi(1.0 / 36.0)};
__device__ constexpr float kWl[9] = {w_lo(4.0 / 9.0), w_lo(1.0 / 9.0), w_lo(1.0 / 9.0), w_lo(1.0 / 9.0),
                                     w_lo(1.0 / 9.0), w_lo(1.0 / 36.0), w_lo(1.0 / 36.0), w_lo(1.0 / 36.0),
                                     w_lo(1.0 / 36.0)};
// 6 w in the same form (6 * kWh is not exact in FP32: 6 * float(1/9)
// rounds, which would bias the equilibrium momentum by 2e-8); 2 w is
// exactly 2 kWh + 2 kWl.
__device__ constexpr float kW6h[9] = {w_hi(24.0 / 9.0), w_hi(6.0 / 9.0), w_hi(6.0 / 9.0), w_hi(6.0 / 9.0),
                                      w_hi(6.0 / 9.0), w_hi(6.0 / 36.0), w_hi(6.0 / 36.0), w_hi(6.0 / 36.0),
                                      w_hi(6.0 / 36.0)};
__device__ constexpr float kW6l[9] = {w_lo(24.0 / 9.0), w_lo(6.0 / 9.0), w_lo(6.0 / 9.0), w_lo(6.0 / 9.0),
                                      w_lo(6.0 / 9.0), w_lo(6.0 / 36.0), w_lo(6.0 / 36.0), w_lo(6.0 / 36.0),
                                      w_lo(6.0 / 36.0)};
// 2/3 and 1/6 of the Zou-He closures in the same form.
constexpr float k23h = w_hi(2.0 / 3.0), k23l = w_lo(2.0 / 3.0);
constexpr float k16h = w_hi(1.0 / 6.0), k16l = w_lo(1.0 / 6.0);
__device__ __forceinline__ float mul23(float x) { return __builtin_fmaf(k23h, x, k23l * x); }
__device__ __forceinline__ float mul16(float x) { return __builtin_fmaf(k16h, x, k16l * x); }

// Boundary types of the reference's btypes enum (aliases.hpp:28-38).
enum : int { BT_NONE = 0, BT_TOP, BT_RIGHT, BT_BOTTOM, BT_LEFT, BT_CORNER };

// Global (1) and constant (4) address spaces: accesses through the first
// are vector global loads/stores with a scalar base; reads through the
// second, of data no kernel writes while it runs, are scalar loads where
// the address is wave-uniform.
#define LBM_GLOBAL __attribute__((address_space(1)))
#define LBM_CONST __attribute__((address_space(4)))

// Maps a launch's blocks (dealt round-robin to the 8 XCDs by the hardware)
// onto the column-major tile order in chunks: XCD x takes chunks x, x + 8,
// x + 16, ... of `chunk` consecutive blocks, so each XCD works through a few
// neighbouring columns at a time (their row segments share its L2) and all
// eight move through memory together.  Blocks past the last whole group of
// 8 chunks keep their order.  Chunk sizes (step_chunk, kXcdChunkRuns,
// kXcdChunk): the row-segment step kernels take one column per chunk when
// a column holds at least 32 work-groups, else two; the wet-run and tabled
// kernels 16 blocks (FP64) or 32 (FP32, FP16; the two-run FP16 tabled
// waves 16, the same 128 runs); the others (ensembles, set-up) 1024
// blocks.
// Measured against one contiguous eighth of the tile order per XCD, one
// process per variant, best launch shape per case: porous 8192^2 FP32 +5 %,
// FP64 +5 %, FP16 +3 %; open 8192^2 FP32 +2 %, FP16 +5 %; 4096^2 cavity
// FP32 +5 %, FP16 +6 %, FP64 +5 %; 512-lattice ensemble +3 %.  With the
// blocked population layout and the tuned occupancy, these chunks then
// beat the earlier 8 columns / 64 blocks / 256 blocks by 1-7 % (porous
// 8192^2 FP32 +1 %, FP64 +6 %, FP16 +3 %; open 8192^2 FP32 +7 %; 4096^2
// cavity FP64 +3 %, FP16 +1.6 %; ensemble +1 %).  Four or eight columns
// per chunk were 2-4 % slower for the one-segment shapes; one column for
// the cavity's 16 work-groups per column 1.5 % slower than two.
// Block counts stay below 2^31 (the geometry caps the 64-row segments at
// 2^30), so the arithmetic is 32-bit.
constexpr uint32_t kXcdChunk = 1024;
template <typename T>
constexpr uint32_t kXcdChunkRuns = sizeof(T) == 8 ? 16 : 32;

__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb, uint32_t chunk = kXcdChunk)
{
    const uint32_t whole = (nb / (8 * chunk)) * (8 * chunk);
    if (b >= whole) return b;
    const uint32_t x = b & 7, k = b >> 3;
    return ((k / chunk) * 8 + x) * chunk + (k % chunk);
}

// Chunk of the row-segment step kernels (wpc waves per column, four per
// work-group): one column when it holds at least 32 work-groups, else two.
__device__ __forceinline__ uint32_t step_chunk(int32_t wpc)
{
    const uint32_t per_column = (uint32_t)max(1, wpc / (kBlockRows / 64));
    return per_column >= 32 ? per_column : 2 * per_column;
}

__device__ __forceinline__ int64_t edge_slot(int i, int j, int nr, int nc)
{
    if (i == 0) return j;
    if (i == nr - 1) return (int64_t)nc + j;
    if (j == 0) return 2 * (int64_t)nc + i;
    return 2 * (int64_t)nc + nr + i;
}

// NodeIdentification.hpp:45-110 for a perimeter node.
__device__ __forceinline__ int edge_btype(int i, int j, int nr, int nc)
{
    bool top = i == 0, bottom = i == nr - 1, left = j == 0, right = j == nc - 1;
    if ((top || bottom) && (left || right)) return BT_CORNER;
    if (top) return BT_TOP;
    if (bottom) return BT_BOTTOM;
    if (left) return BT_LEFT;
    if (right) return BT_RIGHT;
    return BT_NONE;
}

// Node position of this thread: column j, row i, wave-uniform row base i0.
struct Tile {
    int i, j;
    int64_t i0;
    int lane;
    bool valid;
};

// Dense tiling used by the set-up/readout kernels: a work-group covers
// kBlockRows consecutive rows of one column.
__device__ __forceinline__ Tile tile_of(const StepArgs &a)
{
    Tile t;
    int64_t b = xcd_remap(blockIdx.x, (uint32_t)a.nblocks);
    t.j = (int)(b / a.segs);
    int64_t seg = b - (int64_t)t.j * a.segs;
    int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    t.lane = threadIdx.x & 63;
    t.i0 = seg * kBlockRows + wave * 64;
    t.valid = t.i0 < a.g.ldr;
    t.i = (int)t.i0 + t.lane;
    return t;
}

// The 3x3 wet neighbourhood of a wave and the compact-index bases of its
// three columns, all wave-uniform (scalar loads): for the columns j-1, j,
// j+1 the words whose bit r is the wet bit of row i0+r-1 (up), i0+r (mid)
// and i0+r+1 (dn), and the compact index of the first wet node of the
// segment (base).
struct Nbhd {
    uint64_t up[3], mid[3], dn[3];
    int32_t base[3];
};

// EAGER: every word and base is needed before the wave's first population
// load; naming the loaded values in an empty asm statement keeps the
// compiler from sinking the neighbour columns' loads below a test on the
// middle column, which would cost the wave a second round trip to L2.
template <bool EAGER = false>
__device__ __forceinline__ Nbhd load_nbhd(const Geometry &g, int j, int64_t i0)
{
    Nbhd h;
    // The mask and the segment bases never change while a kernel runs
    // (constant address space: scalar loads from any kernel).  Word indices
    // are 32-bit: the geometry caps the mask at 2^31 words.
    const uint32_t wl = (uint32_t)g.wl;
    const uint32_t k0 = (uint32_t)j * wl + (uint32_t)(i0 >> 6);  // column j-1, word k-1
    const LBM_CONST uint64_t *wet = (const LBM_CONST uint64_t *)g.wet;
    const LBM_CONST int32_t *sb = (const LBM_CONST int32_t *)g.sb;
    uint64_t w[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const uint32_t kc = k0 + (uint32_t)c * wl;
        const LBM_CONST uint64_t *col = wet + kc;  // constant offsets below: merged loads
        w[c][0] = col[0];
        w[c][1] = col[1];
        w[c][2] = col[2];
        h.base[c] = (sb + kc)[1];
    }
    if (EAGER)
        asm volatile("" ::"s"(w[0][0]), "s"(w[0][1]), "s"(w[0][2]), "s"(w[1][0]), "s"(w[1][1]), "s"(w[1][2]),
                     "s"(w[2][0]), "s"(w[2][1]), "s"(w[2][2]), "s"(h.base[0]), "s"(h.base[1]), "s"(h.base[2]));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        h.up[c] = (w[c][1] << 1) | (w[c][0] >> 63);
        h.mid[c] = w[c][1];
        h.dn[c] = (w[c][1] >> 1) | (w[c][2] << 63);
    }
    return h;
}

// True when every node of the wave and every source of every node is wet:
// no bounce-back, no perimeter node (guard bits are zero), plain streaming.
__device__ __forceinline__ bool wave_is_bulk(const Nbhd &h)
{
    uint64_t all = ~0ull;
#pragma unroll
    for (int c = 0; c < 3; ++c) all &= h.up[c] & h.mid[c] & h.dn[c];
    return all == ~0ull;
}

// Per-lane mask: bit 0 = the node itself is wet, bit d = the source of
// direction d is wet.
__device__ __forceinline__ uint32_t lane_mask(const Nbhd &h, int l)
{
    uint32_t m = (uint32_t)((h.mid[1] >> l) & 1);
    m |= (uint32_t)((h.mid[0] >> l) & 1) << 1;  // d1: source (i, j-1)
    m |= (uint32_t)((h.dn[1] >> l) & 1) << 2;   // d2: source (i+1, j)
    m |= (uint32_t)((h.mid[2] >> l) & 1) << 3;  // d3: source (i, j+1)
    m |= (uint32_t)((h.up[1] >> l) & 1) << 4;   // d4: source (i-1, j)
    m |= (uint32_t)((h.dn[0] >> l) & 1) << 5;   // d5: source (i+1, j-1)
    m |= (uint32_t)((h.dn[2] >> l) & 1) << 6;   // d6: source (i+1, j+1)
    m |= (uint32_t)((h.up[2] >> l) & 1) << 7;   // d7: source (i-1, j+1)
    m |= (uint32_t)((h.up[0] >> l) & 1) << 8;   // d8: source (i-1, j-1)
    return m;
}

// Number of set bits of w below this lane.
__device__ __forceinline__ uint32_t below(uint64_t w)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(w >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)w, 0u));
}

// Compact indices of this lane's node and of the source of every
// direction (meaningful where the node/source is wet).  Row i of column c
// sits at base[c] + below(mid[c]); row i-1 is the entry before it, row
// i+1 the entry after it when row i is wet.
struct Idx {
    int32_t src[9];
};

__device__ __forceinline__ Idx lane_idx(const Nbhd &h, int lane)
{
    const int32_t l = h.base[0] + (int32_t)below(h.mid[0]);
    const int32_t c = h.base[1] + (int32_t)below(h.mid[1]);
    const int32_t r = h.base[2] + (int32_t)below(h.mid[2]);
    const int32_t wl_ = __builtin_amdgcn_inverse_ballot_w64(h.mid[0]) ? 1 : 0;
    const int32_t wr_ = __builtin_amdgcn_inverse_ballot_w64(h.mid[2]) ? 1 : 0;
    Idx x;
    x.src[0] = c;           // own node
    x.src[1] = l;           // (i, j-1)
    x.src[2] = c + 1;       // (i+1, j)
    x.src[3] = r;           // (i, j+1)
    x.src[4] = c - 1;       // (i-1, j)
    x.src[5] = l + wl_;     // (i+1, j-1)
    x.src[6] = r + wr_;     // (i+1, j+1)
    x.src[7] = r - 1;       // (i-1, j+1)
    x.src[8] = l - 1;       // (i-1, j-1)
    return x;
}

// Element access through a wave-uniform base pointer and a 32-bit byte
// offset, the addressing form global loads/stores take natively (scalar
// base + vector offset).  The context only accepts lattices whose
// population offsets stay below 2^32 bytes (the whole buffer blocked, up to
// three planes in planes).  The accesses go through global-address-space
// pointers: a direction's base pointer read from memory (the batch
// kernels' per-sample arguments) is otherwise a generic pointer, and
// generic accesses become flat instructions with 64-bit vector addresses.
template <typename T>
__device__ __forceinline__ const LBM_GLOBAL T *gaddr(const void *base, uint32_t byte_off)
{
    return (const LBM_GLOBAL T *)((const LBM_GLOBAL char *)base + byte_off);
}

template <typename T>
__device__ __forceinline__ LBM_GLOBAL T *gaddr(void *base, uint32_t byte_off)
{
    return (LBM_GLOBAL T *)((LBM_GLOBAL char *)base + byte_off);
}

template <typename T>
__device__ __forceinline__ T ld(const T *__restrict__ base, uint32_t byte_off)
{
    return *gaddr<T>((const void *)base, byte_off);
}

template <typename T>
__device__ __forceinline__ void st(T *__restrict__ base, uint32_t byte_off, T v)
{
    *gaddr<T>((void *)base, byte_off) = v;
}

// Byte offset of compact node n from its direction's base pointer
// (fin_plane / fout_plane): n in planes, n + 8 (n - n mod 64) blocked (the
// 9 directions of 64 consecutive nodes together).  The layout is a scalar
// mask, so both cost the same two instructions.  Both keep 64 consecutive
// nodes from a 64-aligned index in one stretch per direction, so
// whole-line stores stay whole lines.  Blocked, a wave's nine loads of a
// 64-node range fall in one 9 * 64-element stretch instead of nine places
// a plane apart; on MI355X that made the step kernels 4-7 % faster in FP64
// and FP32 and 3-15 % in FP16 (8192^2 porous and open media, 4096^2
// cavity), and the ensemble 8 %.
template <typename T>
__device__ __forceinline__ uint32_t slot_bytes(const StepArgs &a, uint32_t n)
{
    return (n + 8u * (n & a.block_mask)) * (uint32_t)sizeof(T);
}

// Streams the 9 populations of an interior wet node: one load per
// direction, its address chosen between the neighbour (stream) and the
// node's own opposite slot (bounce-back).  Opposite directions are two
// direction strides apart (two_planes bytes), so both candidates are
// reached from the base of the pair's lower direction with a non-negative
// 32-bit offset.
template <typename T, typename WetFn>
__device__ __forceinline__ void gather_interior_by(const StepArgs &a, const Idx &x, WetFn wet_src, T *p)
{
    const uint32_t twoP = a.two_planes;
    const uint32_t own = slot_bytes<T>(a, (uint32_t)x.src[0]);
    p[0] = ld((const T *)a.fin_plane[0], own);
#pragma unroll
    for (int d = 1; d < 9; ++d) {
        const bool wet = wet_src(d);
        const int o = kOpp[d];
        const bool low = d < o;  // d is the lower direction of its pair
        const T *base = (const T *)a.fin_plane[low ? d : o];
        uint32_t nb = slot_bytes<T>(a, (uint32_t)x.src[d]) + (low ? 0u : twoP);
        uint32_t ob = own + (low ? twoP : 0u);
        p[d] = ld(base, wet ? nb : ob);
    }
}

// (s & x) | (~s & y) in one instruction.  Written out because the compiler
// turns the plain expression back into a test and a select when s is a
// sign-extended bit.
__device__ __forceinline__ uint32_t bit_insert(uint32_t s, uint32_t x, uint32_t y)
{
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(s), "v"(x), "v"(y));
    return r;
}

// Per-lane mask form (bit d of m = the source of direction d is wet): bit
// d, sign-extended, selects between the two offsets by a bit-field insert
// (two instructions per direction where a test and a select take three).
template <typename T>
__device__ __forceinline__ void gather_interior(const StepArgs &a, const Idx &x, uint32_t m, T *p)
{
    const uint32_t twoP = a.two_planes;
    const uint32_t own = slot_bytes<T>(a, (uint32_t)x.src[0]);
    p[0] = ld((const T *)a.fin_plane[0], own);
#pragma unroll
    for (int d = 1; d < 9; ++d) {
        const uint32_t sel = (uint32_t)((int32_t)(m << (31 - d)) >> 31);  // all ones: the source is wet
        const int o = kOpp[d];
        const bool low = d < o;
        const T *base = (const T *)a.fin_plane[low ? d : o];
        const uint32_t nb = slot_bytes<T>(a, (uint32_t)x.src[d]) + (low ? 0u : twoP);
        const uint32_t ob = own + (low ? twoP : 0u);
        p[d] = ld(base, bit_insert(sel, nb, ob));
    }
}

// Wave-uniform source masks (bit l of src[d] = the source of direction d of
// lane l is wet): each select is one v_cndmask on a scalar lane mask.
struct SrcMasks {
    uint64_t src[9];
};

__device__ __forceinline__ SrcMasks src_masks(const Nbhd &h)
{
    SrcMasks s;
    s.src[0] = h.mid[1];
    s.src[1] = h.mid[0];  // (i, j-1)
    s.src[2] = h.dn[1];   // (i+1, j)
    s.src[3] = h.mid[2];  // (i, j+1)
    s.src[4] = h.up[1];   // (i-1, j)
    s.src[5] = h.dn[0];   // (i+1, j-1)
    s.src[6] = h.dn[2];   // (i+1, j+1)
    s.src[7] = h.up[2];   // (i-1, j+1)
    s.src[8] = h.up[0];   // (i-1, j-1)
    return s;
}

// Same for a perimeter node: sources outside the domain read the frozen
// slot of the node's edge entry.
template <typename T>
__device__ __forceinline__ void gather_edge(const StepArgs &a, const T *__restrict__ frozen, int i, int j,
                                            const Idx &x, int64_t e, uint32_t m, T *p)
{
    const uint32_t own = slot_bytes<T>(a, (uint32_t)x.src[0]);
    p[0] = ld((const T *)a.fin_plane[0], own);
#pragma unroll
    for (int d = 1; d < 9; ++d) {
        int si = i - kOffI[d], sj = j - kOffJ[d];
        bool out = si < 0 || si >= a.g.nr || sj < 0 || sj >= a.g.nc;
        if (out) {
            p[d] = frozen[d * a.g.E + e];
        } else {
            bool wet = (m >> d) & 1;
            p[d] = wet ? ld((const T *)a.fin_plane[d], slot_bytes<T>(a, (uint32_t)x.src[d]))
                       : ld((const T *)a.fin_plane[kOpp[d]], own);
        }
    }
}

// Writes the 9 post-collision populations of the node with compact index n.
// Each store helper takes its own opaque copy of the byte offset (empty
// asm): the compiler then cannot hoist the nine direction addresses shared by
// the plain and the streaming branch into 64-bit vector pointers, and keeps
// every store in the scalar-base + 32-bit-offset form.
template <typename T>
__device__ __forceinline__ void store9(const StepArgs &a, int64_t n, const T *o)
{
    uint32_t b = slot_bytes<T>(a, (uint32_t)n);
    asm volatile("" : "+v"(b));
#pragma unroll
    for (int d = 0; d < 9; ++d) st((T *)a.fout_plane[d], b, o[d]);
}

// The 9 stores of one node of a segment whose wave writes whole, aligned
// cache lines (64 contiguous wet nodes from a 64-aligned slot): stream
// them past the caches.  A partial line written that way would cost HBM a
// read-modify-write, so every other store goes through L2, which merges
// the pieces.  The empty asm statements keep the compiler from merging
// these stores with plain ones, which would drop the hint.
template <typename T>
__device__ __forceinline__ void store9_stream(const StepArgs &a, int64_t n, const T *o)
{
    uint32_t b = slot_bytes<T>(a, (uint32_t)n);
    asm volatile("" : "+v"(b)::"memory");
#pragma unroll
    for (int d = 0; d < 9; ++d) __builtin_nontemporal_store(o[d], gaddr<T>(a.fout_plane[d], b));
    asm volatile("" ::: "memory");
}

// ---------------------------------------------------------------- FP64 --

// DefaultEquilibrium.hpp:62-76, same expression order.
__device__ __forceinline__ void feq64(double rho, double ux, double uy, double *e)
{
    const double two9 = (2.0 / 9.0);
    const double one18 = (1.0 / 18.0);
    const double one36 = (1.0 / 36.0);
    double u2 = ux * ux + uy * uy;
    e[0] = two9 * rho * (2 - 3 * u2);
    e[1] = one18 * rho * (2 + 6 * ux + 9 * ux * ux - 3 * u2);
    e[2] = one18 * rho * (2 + 6 * uy + 9 * uy * uy - 3 * u2);
    e[3] = one18 * rho * (2 - 6 * ux + 9 * ux * ux - 3 * u2);
    e[4] = one18 * rho * (2 - 6 * uy + 9 * uy * uy - 3 * u2);
    e[5] = one36 * rho * (1 + 3 * (ux + uy) + 9 * ux * uy + 3 * u2);
    e[6] = one36 * rho * (1 - 3 * (ux - uy) - 9 * ux * uy + 3 * u2);
    e[7] = one36 * rho * (1 - 3 * (ux + uy) + 9 * ux * uy + 3 * u2);
    e[8] = one36 * rho * (1 + 3 * (ux - uy) - 9 * ux * uy + 3 * u2);
}

// DefaultEquilibrium.hpp:145-153
__device__ __forceinline__ void moments64(const double *p, double &rho, double &ux, double &uy)
{
    rho = p[0] + p[1] + p[2] + p[3] + p[4] + p[5] + p[6] + p[7] + p[8];
    double rhoinv = 1.0 / rho;
    ux = rhoinv * (p[1] + p[5] + p[8] - (p[3] + p[6] + p[7]));
    uy = rhoinv * (p[2] + p[5] + p[6] - (p[4] + p[7] + p[8]));
}

// ZouHePolicy.hpp:70-118 (velocity prescribed).  The populations are
// copied to scalars, the branch of the node's side rewrites its three
// unknowns, and all nine are stored back unconditionally: stores into
// different array elements on different branches would otherwise be merged
// through a pointer phi and push the array into scratch memory.
__device__ __forceinline__ void zouhe_velocity64(int bt, double *p, double ux, double uy)
{
    const double two_thirds = 2.0 / 3.0;
    const double one_sixth = 1.0 / 6.0;
    const double one_half = 0.5;
    double p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3], p4 = p[4], p5 = p[5], p6 = p[6], p7 = p[7], p8 = p[8];
    double rho, ru, rv;
    if (bt == BT_TOP) {
        rho = (p0 + p1 + p3 + 2.0 * (p2 + p5 + p6)) / (1.0 + uy);
        ru = rho * ux; rv = rho * uy;
        p4 = p2 - two_thirds * rv;
        p7 = p5 - one_sixth * rv - one_half * ru + one_half * (p1 - p3);
        p8 = p6 - one_sixth * rv + one_half * ru - one_half * (p1 - p3);
    } else if (bt == BT_BOTTOM) {
        rho = (p0 + p1 + p3 + 2.0 * (p4 + p7 + p8)) / (1.0 - uy);
        ru = rho * ux; rv = rho * uy;
        p2 = p4 + two_thirds * rv;
        p5 = p7 + one_sixth * rv + one_half * ru - one_half * (p1 - p3);
        p6 = p8 + one_sixth * rv - one_half * ru + one_half * (p1 - p3);
    } else if (bt == BT_LEFT) {
        rho = (p0 + p2 + p4 + 2.0 * (p3 + p7 + p6)) / (1.0 - ux);
        ru = rho * ux; rv = rho * uy;
        p1 = p3 + two_thirds * ru;
        p5 = p7 + one_sixth * ru + one_half * rv - one_half * (p2 - p4);
        p8 = p6 + one_sixth * ru - one_half * rv + one_half * (p2 - p4);
    } else if (bt == BT_RIGHT) {
        rho = (p0 + p2 + p4 + 2.0 * (p1 + p5 + p8)) / (1.0 + ux);
        ru = rho * ux; rv = rho * uy;
        p3 = p1 - two_thirds * ru;
        p7 = p5 - one_sixth * ru - one_half * rv + one_half * (p2 - p4);
        p6 = p8 - one_sixth * ru + one_half * rv - one_half * (p2 - p4);
    }
    // corners and NONE are left alone (ZouHePolicy.hpp:119-120)
    p[1] = p1; p[2] = p2; p[3] = p3; p[4] = p4; p[5] = p5; p[6] = p6; p[7] = p7; p[8] = p8;
}

// Pressure closure: density prescribed, normal velocity from the known
// populations, tangential velocity zero, unknowns from the velocity closure.
__device__ __forceinline__ void zouhe_pressure64(int bt, double *p, double rho)
{
    double ux = 0.0, uy = 0.0;
    if (bt == BT_TOP) uy = (p[0] + p[1] + p[3] + 2.0 * (p[2] + p[5] + p[6])) / rho - 1.0;
    else if (bt == BT_BOTTOM) uy = 1.0 - (p[0] + p[1] + p[3] + 2.0 * (p[4] + p[7] + p[8])) / rho;
    else if (bt == BT_LEFT) ux = 1.0 - (p[0] + p[2] + p[4] + 2.0 * (p[3] + p[7] + p[6])) / rho;
    else if (bt == BT_RIGHT) ux = (p[0] + p[2] + p[4] + 2.0 * (p[1] + p[5] + p[8])) / rho - 1.0;
    else return;
    const double two_thirds = 2.0 / 3.0;
    const double one_sixth = 1.0 / 6.0;
    const double one_half = 0.5;
    double p1 = p[1], p2 = p[2], p3 = p[3], p4 = p[4], p5 = p[5], p6 = p[6], p7 = p[7], p8 = p[8];
    double ru = rho * ux, rv = rho * uy;
    if (bt == BT_TOP) {
        p4 = p2 - two_thirds * rv;
        p7 = p5 - one_sixth * rv - one_half * ru + one_half * (p1 - p3);
        p8 = p6 - one_sixth * rv + one_half * ru - one_half * (p1 - p3);
    } else if (bt == BT_BOTTOM) {
        p2 = p4 + two_thirds * rv;
        p5 = p7 + one_sixth * rv + one_half * ru - one_half * (p1 - p3);
        p6 = p8 + one_sixth * rv - one_half * ru + one_half * (p1 - p3);
    } else if (bt == BT_LEFT) {
        p1 = p3 + two_thirds * ru;
        p5 = p7 + one_sixth * ru + one_half * rv - one_half * (p2 - p4);
        p8 = p6 + one_sixth * ru - one_half * rv + one_half * (p2 - p4);
    } else {
        p3 = p1 - two_thirds * ru;
        p7 = p5 - one_sixth * ru - one_half * rv + one_half * (p2 - p4);
        p6 = p8 - one_sixth * ru + one_half * rv - one_half * (p2 - p4);
    }
    p[1] = p1; p[2] = p2; p[3] = p3; p[4] = p4; p[5] = p5; p[6] = p6; p[7] = p7; p[8] = p8;
}

// Pre-collision populations of wet node (i, j) in FP64, with boundary
// handling.  Returns the open-boundary policy of the node (-1 = fluid).
// Partially saturated bounce-back, gather side: slots whose source is a
// wet obstacle take the value the node computed for them at the end of the
// previous step (psbb_update) instead of its own opposite slot.
__device__ __forceinline__ void psbb_gather(const StepArgs &a, int64_t n, double *p)
{
    const uint32_t mask = a.ps_mask[n];
    if (!mask) return;
#pragma unroll
    for (int d = 1; d < 9; ++d)
        if ((mask >> d) & 1u) p[d] = a.ps_val[(int64_t)d * a.g.plane + n];
}

// PSBounceBackPolicy<2>::update_boundaries (PSBounceBackPolicy.hpp:177-297)
// for the slots of node n whose source o = n - c_d is a wet obstacle: with
// q = opp(d), p_q = F*_q(n) (what n streamed into o), P_d the population
// n held in slot d before this step, e the equilibrium of n at this step
// (DefaultEquilibrium forms, as the policy restates them) and rho_o = 1
// (an obstacle's global_rho keeps its initial value, Lattice.hpp:295):
//   P'_d = p_q - (1 - b) (p_q - e_q) t + b ((P_d - e_d) - (p_q - 2 w_d rho_o))
// with t = 1 / tau of the policy and 2 w_d written 2 * (1/18) on the axes
// and 1/36 on the diagonals, in the reference's operation order.
__device__ __forceinline__ void psbb_update(const StepArgs &a, int64_t n, const double *p, const double *eq,
                                            const double *o)
{
    const uint32_t mask = a.ps_mask[n];
    if (!mask) return;
    constexpr double one18 = (1.0 / 18.0);
    constexpr double one36 = (1.0 / 36.0);
    const double rho_o = 1.0;
    const double t = a.ps_t;
#pragma unroll
    for (int d = 1; d < 9; ++d) {
        if (!((mask >> d) & 1u)) continue;
        const int q = kOpp[d];
        const int64_t k = (int64_t)d * a.g.plane + n;
        const double b = a.ps_b[k];
        const double wr = d < 5 ? 2 * one18 * rho_o : one36 * rho_o;
        const double pq = o[q];
        a.ps_val[k] = pq - (1.0 - b) * (pq - eq[q]) * t + b * ((p[d] - eq[d]) - (pq - wr));
    }
}

template <bool GATHER>
__device__ __forceinline__ int load_p64(const StepArgs &a, int i, int j, const Idx &x, uint32_t m, bool edge,
                                        int64_t &e, double *p)
{
    int policy = -1;
    if (!edge) {
        if (GATHER) {
            gather_interior(a, x, m, p);
            if (a.psbb) psbb_gather(a, x.src[0], p);
        } else {
#pragma unroll
            for (int d = 0; d < 9; ++d) p[d] = ld((const double *)a.fin_plane[d], slot_bytes<double>(a, x.src[0]));
        }
        return -1;
    }
    e = edge_slot(i, j, a.g.nr, a.g.nc);
    if (GATHER) {
        gather_edge(a, (const double *)a.frozen, i, j, x, e, m, p);
        if (a.psbb) psbb_gather(a, x.src[0], p);
    } else {
#pragma unroll
        for (int d = 0; d < 9; ++d) p[d] = ld((const double *)a.fin_plane[d], slot_bytes<double>(a, x.src[0]));
    }
    int kind = a.g.edge_kind[e];
    if (kind != kEdgePlain) {
        policy = kind == kEdgeInlet ? a.inlet_policy : a.outlet_policy;
        if (GATHER) {
            int bt = edge_btype(i, j, a.g.nr, a.g.nc);
            if (policy == kVelocity) zouhe_velocity64(bt, p, a.bc_prev[e], a.bc_prev[a.g.E + e]);
            else if (policy == kPressure) zouhe_pressure64(bt, p, a.bc_prev[2 * a.g.E + e]);
        }
    }
    return policy;
}

// Moments, equilibrium, collision and store of one wet node in FP64, in
// the reference's expression order (update_macro, calc_equilibrium,
// collide: Lattice.hpp:351-382).  Fluid and pressure nodes take their
// moments; velocity-prescribed and no-op nodes keep rho = 1
// (Lattice.hpp:295; the Zou-He density write is lost on a by-value copy,
// ZouHePolicy.hpp:58,71) and the prescribed u.
__device__ __forceinline__ void collide_f64(const StepArgs &a, const double *p, int policy, int64_t e, int i, int j,
                                            double *o, double *eq_out = nullptr)
{
    double rho, ux, uy;
    if (policy == kVelocity || policy == kNoop) {
        rho = 1.0;
        ux = a.bc_cur[e];
        uy = a.bc_cur[a.g.E + e];
        double s = p[0] + p[1] + p[2] + p[3] + p[4] + p[5] + p[6] + p[7] + p[8];
        if (s != s) atomicOr(a.flag, 1);
    } else {
        moments64(p, rho, ux, uy);
        if (rho != rho) atomicOr(a.flag, 1);
    }
    if (a.macro) {
        int64_t q = (int64_t)j * a.g.nr + i, N = (int64_t)a.g.nr * a.g.nc;
        LBM_GLOBAL double *mq = (LBM_GLOBAL double *)a.macro + q;
        mq[0] = rho;
        mq[N] = ux;
        mq[2 * N] = uy;
    }
    double eq[9];
    feq64(rho, ux, uy, eq);
    if (eq_out)
#pragma unroll
        for (int d = 0; d < 9; ++d) eq_out[d] = eq[d];
    if (a.collision == 0) {
        // BGKCollision.hpp:140-153
        const double t_ = a.bgk_t, o_m_t = a.bgk_omt;
#pragma unroll
        for (int d = 0; d < 9; ++d) o[d] = o_m_t * p[d] + t_ * eq[d];
    } else {
        // TRTCollision.hpp:155-188
        const double ca = a.trt_a, cb = a.trt_b;
        o[0] = p[0] - a.trt_inv_te * (p[0] - eq[0]);
#pragma unroll
        for (int d = 1; d < 9; ++d) {
            const int q = kOpp[d];
            o[d] = p[d] - ca * (p[d] + p[q] - eq[d] - eq[q]) - cb * (p[d] - p[q] - eq[d] + eq[q]);
        }
    }
}

__device__ __forceinline__ void collide_store_f64(const StepArgs &a, const double *p, int policy, int64_t e, int i,
                                                  int j, int64_t n)
{
    double o[9];
    collide_f64(a, p, policy, e, i, j, o);
    store9(a, n, o);
}

// One node of any kind (perimeter, open boundary, bounce-back).
template <bool GATHER>
__device__ __forceinline__ void node_general_f64(const StepArgs &a, int i, int j, int lane, const Nbhd &h)
{
    uint32_t m = lane_mask(h, lane);
    if (!(m & 1)) return;
    const bool edge = (i == 0) | (i == a.g.nr - 1) | (j == 0) | (j == a.g.nc - 1);
    const Idx x = lane_idx(h, lane);
    double p[9];
    int64_t e = 0;
    int policy = load_p64<GATHER>(a, i, j, x, m, edge, e, p);
    if (a.psbb) {
        double o[9], eq[9];
        collide_f64(a, p, policy, e, i, j, o, eq);
        psbb_update(a, x.src[0], p, eq, o);
        store9(a, x.src[0], o);
        return;
    }
    collide_store_f64(a, p, policy, e, i, j, x.src[0]);
}

// ------------------------------------------------------- shifted FP32 --

template <typename T>
__device__ __forceinline__ float to_g(T v, float inv_scale);
template <>
__device__ __forceinline__ float to_g<float>(float v, float) { return v; }
template <>
__device__ __forceinline__ float to_g<_Float16>(_Float16 v, float inv_scale) { return (float)v * inv_scale; }

template <typename T>
__device__ __forceinline__ T from_g(float g, float scale);
template <>
__device__ __forceinline__ float from_g<float>(float g, float) { return g; }
template <>
__device__ __forceinline__ _Float16 from_g<_Float16>(float g, float scale) { return (_Float16)(g * scale); }

// Zou-He closures on shifted populations g = f - w.  The weights of the
// known sets sum to 1 and opposite weights are equal, so the closures keep
// their form with f -> g and rho -> 1 + (sum) as below.  Same scalar
// structure as the FP64 closures.
__device__ __forceinline__ void zouhe_velocity32(int bt, float *g, float ux, float uy)
{
    // rho - 1 is formed directly, e.g. TOP: rho = (1 + G) / (1 + uy) gives
    // rho - 1 = (G - uy) / (1 + uy); rho * u = u + (rho - 1) * u.  Rounding
    // 1 + G first would cost the low bits of G.
    float g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3], g4 = g[4], g5 = g[5], g6 = g[6], g7 = g[7], g8 = g[8];
    float dr, ru, rv;
    if (bt == BT_TOP) {
        dr = ((g0 + g1 + g3 + 2.f * (g2 + g5 + g6)) - uy) / (1.f + uy);
        ru = __builtin_fmaf(dr, ux, ux); rv = __builtin_fmaf(dr, uy, uy);
        g4 = g2 - mul23(rv);
        g7 = g5 - mul16(rv) - 0.5f * ru + 0.5f * (g1 - g3);
        g8 = g6 - mul16(rv) + 0.5f * ru - 0.5f * (g1 - g3);
    } else if (bt == BT_BOTTOM) {
        dr = ((g0 + g1 + g3 + 2.f * (g4 + g7 + g8)) + uy) / (1.f - uy);
        ru = __builtin_fmaf(dr, ux, ux); rv = __builtin_fmaf(dr, uy, uy);
        g2 = g4 + mul23(rv);
        g5 = g7 + mul16(rv) + 0.5f * ru - 0.5f * (g1 - g3);
        g6 = g8 + mul16(rv) - 0.5f * ru + 0.5f * (g1 - g3);
    } else if (bt == BT_LEFT) {
        dr = ((g0 + g2 + g4 + 2.f * (g3 + g7 + g6)) + ux) / (1.f - ux);
        ru = __builtin_fmaf(dr, ux, ux); rv = __builtin_fmaf(dr, uy, uy);
        g1 = g3 + mul23(ru);
        g5 = g7 + mul16(ru) + 0.5f * rv - 0.5f * (g2 - g4);
        g8 = g6 + mul16(ru) - 0.5f * rv + 0.5f * (g2 - g4);
    } else if (bt == BT_RIGHT) {
        dr = ((g0 + g2 + g4 + 2.f * (g1 + g5 + g8)) - ux) / (1.f + ux);
        ru = __builtin_fmaf(dr, ux, ux); rv = __builtin_fmaf(dr, uy, uy);
        g3 = g1 - mul23(ru);
        g7 = g5 - mul16(ru) - 0.5f * rv + 0.5f * (g2 - g4);
        g6 = g8 - mul16(ru) + 0.5f * rv - 0.5f * (g2 - g4);
    }
    g[1] = g1; g[2] = g2; g[3] = g3; g[4] = g4; g[5] = g5; g[6] = g6; g[7] = g7; g[8] = g8;
}

// drho0 = rho0 - 1 is formed in FP64 by the caller.
__device__ __forceinline__ void zouhe_pressure32(int bt, float *g, float rho0, float drho0)
{
    float g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3], g4 = g[4], g5 = g[5], g6 = g[6], g7 = g[7], g8 = g[8];
    float ux = 0.f, uy = 0.f;
    if (bt == BT_TOP) uy = ((g0 + g1 + g3 + 2.f * (g2 + g5 + g6)) - drho0) / rho0;
    else if (bt == BT_BOTTOM) uy = (drho0 - (g0 + g1 + g3 + 2.f * (g4 + g7 + g8))) / rho0;
    else if (bt == BT_LEFT) ux = (drho0 - (g0 + g2 + g4 + 2.f * (g3 + g7 + g6))) / rho0;
    else if (bt == BT_RIGHT) ux = ((g0 + g2 + g4 + 2.f * (g1 + g5 + g8)) - drho0) / rho0;
    else return;
    float ru = rho0 * ux, rv = rho0 * uy;
    if (bt == BT_TOP) {
        g4 = g2 - mul23(rv);
        g7 = g5 - mul16(rv) - 0.5f * ru + 0.5f * (g1 - g3);
        g8 = g6 - mul16(rv) + 0.5f * ru - 0.5f * (g1 - g3);
    } else if (bt == BT_BOTTOM) {
        g2 = g4 + mul23(rv);
        g5 = g7 + mul16(rv) + 0.5f * ru - 0.5f * (g1 - g3);
        g6 = g8 + mul16(rv) - 0.5f * ru + 0.5f * (g1 - g3);
    } else if (bt == BT_LEFT) {
        g1 = g3 + mul23(ru);
        g5 = g7 + mul16(ru) + 0.5f * rv - 0.5f * (g2 - g4);
        g8 = g6 + mul16(ru) - 0.5f * rv + 0.5f * (g2 - g4);
    } else {
        g3 = g1 - mul23(ru);
        g7 = g5 - mul16(ru) - 0.5f * rv + 0.5f * (g2 - g4);
        g6 = g8 - mul16(ru) + 0.5f * rv - 0.5f * (g2 - g4);
    }
    g[1] = g1; g[2] = g2; g[3] = g3; g[4] = g4; g[5] = g5; g[6] = g6; g[7] = g7; g[8] = g8;
}

template <typename T, bool GATHER>
__device__ __forceinline__ int load_g(const StepArgs &a, int i, int j, const Idx &x, uint32_t m, bool edge,
                                      int64_t &e, float *g)
{
    T raw[9];
    int policy = -1;
    const uint32_t own = slot_bytes<T>(a, (uint32_t)x.src[0]);
    if (!edge) {
        if (GATHER) gather_interior(a, x, m, raw);
        else
#pragma unroll
            for (int d = 0; d < 9; ++d) raw[d] = ld((const T *)a.fin_plane[d], own);
#pragma unroll
        for (int d = 0; d < 9; ++d) g[d] = to_g<T>(raw[d], a.inv_scale);
        return -1;
    }
    e = edge_slot(i, j, a.g.nr, a.g.nc);
    if (GATHER) gather_edge(a, (const T *)a.frozen, i, j, x, e, m, raw);
    else
#pragma unroll
        for (int d = 0; d < 9; ++d) raw[d] = ld((const T *)a.fin_plane[d], own);
#pragma unroll
    for (int d = 0; d < 9; ++d) g[d] = to_g<T>(raw[d], a.inv_scale);
    int kind = a.g.edge_kind[e];
    if (kind != kEdgePlain) {
        policy = kind == kEdgeInlet ? a.inlet_policy : a.outlet_policy;
        if (GATHER) {
            int bt = edge_btype(i, j, a.g.nr, a.g.nc);
            if (policy == kVelocity) {
                zouhe_velocity32(bt, g, (float)a.bc_prev[e], (float)a.bc_prev[a.g.E + e]);
            } else if (policy == kPressure) {
                double r0 = a.bc_prev[2 * a.g.E + e];
                zouhe_pressure32(bt, g, (float)r0, (float)(r0 - 1.0));
            }
        }
    }
    return policy;
}

// Moments of shifted populations, summed in FP64.  FP32 values whose
// exponents lie within 29 binades of each other add up exactly in FP64, so
// the sums below are exact for shifted populations (and accurate to 2^-53
// of the largest term otherwise); each result is rounded once.  Opposite
// pairs are formed first: their sums and differences are reused by the
// collision.
struct Moments {
    float s[4], a[4];  // pair sums / differences, pairs (1,3) (2,4) (5,7) (6,8)
    double drho, jx, jy;
};

__device__ __forceinline__ void moments_shifted(const float *g, Moments &m)
{
    const double g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3], g4 = g[4], g5 = g[5], g6 = g[6], g7 = g[7], g8 = g[8];
    const double s0 = g1 + g3, a0 = g1 - g3;
    const double s1 = g2 + g4, a1 = g2 - g4;
    const double s2 = g5 + g7, a2 = g5 - g7;
    const double s3 = g6 + g8, a3 = g6 - g8;
    m.drho = g0 + ((s0 + s1) + (s2 + s3));
    // x along +column, y along -row (DefaultEquilibrium.hpp:148-149)
    m.jx = a0 + (a2 - a3);
    m.jy = a1 + (a2 + a3);
    m.s[0] = (float)s0; m.s[1] = (float)s1; m.s[2] = (float)s2; m.s[3] = (float)s3;
    m.a[0] = (float)a0; m.a[1] = (float)a1; m.a[2] = (float)a2; m.a[3] = (float)a3;
}

// Moments, equilibrium and collision of one wet node on shifted
// populations, FP32 arithmetic.  F16 = false (FP32 storage): g = f - w and
// the moments are summed exactly in FP64 (moments_shifted).  F16 = true:
// g = (f - w) S with S the FP16 storage scale, a power of two, so every
// FP32 operation below gives exactly S (or S^2) times what it gives on
// f - w and the scale never needs applying; the momentum is summed in
// FP32, whose rounding is far below the storage's 11-bit resolution, and
// the density exactly.
//
// Per opposite pair (d, q = opp d), with c = c_d and j = rho u:
//   geq_d + geq_q = 2 w (drho + (4.5 (c.j)^2 - 1.5 j.j) / rho)
//   geq_d - geq_q = 6 w (c.j)
// (the D2Q9 equilibrium written in j: rho (3 c.u) = 3 c.j), and the rest
// equilibrium is the remainder drho - sum of the pairs, so the nine carry
// the node's mass up to one rounding.  The weights are applied in two
// parts (hi + lo) so the equilibrium momentum is j to FP32 accuracy, not
// (1 + 7.45e-9) j.  1/rho only scales the second-order term, where the
// hardware reciprocal's error is far below FP32 resolution of the result
// (scaled form: 1 / (S rho) = 1 / (S + S drho)).
// TRT on pairs of n = geq - g:  o_d = g_d + ca (n_d + n_q) + cb (n_d - n_q)
// (BGK: ca = cb = omega / 2).
template <bool F16>
__device__ __forceinline__ void collide_shifted(const StepArgs &a, const float *g, int policy, int64_t e, int i, int j,
                                                float *o)
{
    float ps[4], pa[4];  // pair sums / differences, pairs (1,3) (2,4) (5,7) (6,8)
    float drho, jx, jy;
    double md, mx, my;   // moments of f - w for the macro output
    if constexpr (F16) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int d = k < 2 ? k + 1 : k + 3;
            ps[k] = g[d] + g[kOpp[d]];
            pa[k] = g[d] - g[kOpp[d]];
        }
        // the density exactly (FP64, rounded once): the rest equilibrium is
        // the remainder drho - sum of the pairs, so this sum is what keeps
        // the collision's mass balance to one rounding
        const double dsum = (double)g[0] + ((((double)g[1] + g[3]) + ((double)g[2] + g[4])) +
                                            (((double)g[5] + g[7]) + ((double)g[6] + g[8])));
        drho = (float)dsum;
        jx = pa[0] + (pa[2] - pa[3]);
        jy = pa[1] + (pa[2] + pa[3]);
        md = dsum * a.inv_scale;
        mx = (double)jx * a.inv_scale;
        my = (double)jy * a.inv_scale;
    } else {
        Moments m;
        moments_shifted(g, m);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ps[k] = m.s[k];
            pa[k] = m.a[k];
        }
        md = m.drho;
        mx = m.jx;
        my = m.jy;
        drho = (float)md;
        jx = (float)mx;
        jy = (float)my;
    }
    if (md != md) atomicOr(a.flag, 1);
    double rho_out, ux_out, uy_out;
    if (policy == kVelocity || policy == kNoop) {
        // rho = 1 and the prescribed velocity (see collide_store_f64)
        const float ux = (float)a.bc_cur[e], uy = (float)a.bc_cur[a.g.E + e];
        drho = 0.f;
        jx = F16 ? ux * a.scale : ux;
        jy = F16 ? uy * a.scale : uy;
        rho_out = 1.0;
        ux_out = ux;
        uy_out = uy;
    } else {
        rho_out = 1.0 + md;
        ux_out = mx / rho_out;
        uy_out = my / rho_out;
    }
    if (a.macro) {
        int64_t q = (int64_t)j * a.g.nr + i, N = (int64_t)a.g.nr * a.g.nc;
        LBM_GLOBAL double *mq = (LBM_GLOBAL double *)a.macro + q;
        mq[0] = rho_out;
        mq[N] = ux_out;
        mq[2 * N] = uy_out;
    }
    const float rinv = __builtin_amdgcn_rcpf((F16 ? a.scale : 1.f) + drho);
    const float j15 = 1.5f * __builtin_fmaf(jx, jx, jy * jy);
    const float ca = a.sh_a, cb = a.sh_b;
    // c.j of the first direction of each pair: 1 (+x), 2 (+y), 5 (+x+y), 6 (-x+y)
    const float cj[4] = {jx, jy, jx + jy, jy - jx};
    constexpr int kD[4] = {1, 2, 5, 6};
    float seq_sum;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int d = kD[k], q = kOpp[d];
        const float c = cj[k];
        const float sw = drho + __builtin_fmaf(4.5f * c, c, -j15) * rinv;
        const float seq = __builtin_fmaf(2.f * kWh[d], sw, (2.f * kWl[d]) * sw);
        const float aeq = __builtin_fmaf(kW6h[d], c, kW6l[d] * c);
        seq_sum = k == 0 ? seq : seq_sum + seq;
        const float ds = seq - ps[k];
        const float da = aeq - pa[k];
        const float t = cb * da;
        o[d] = g[d] + __builtin_fmaf(ca, ds, t);
        o[q] = g[q] + __builtin_fmaf(ca, ds, -t);
    }
    const float geq0 = drho - seq_sum;
    o[0] = __builtin_fmaf(2.f * ca, geq0 - g[0], g[0]);
}

// Converts the post-collision shifted populations to the storage type and
// stores them (whole_lines: non-temporal, see store9_stream).  FP16 values
// arrive scaled (collide_shifted<true>).  FP16 rounds stochastically unless
// the context asked for round-to-nearest: 13 random bits are added below
// the FP16 cut of the FP32 value, which is then converted toward zero
// (sign-magnitude, so the same for both signs; two values per packed
// conversion).  The bits come from one hash per node of its index and its
// new rest population, a different window per direction.  Each rounding
// mode stores in its own branch so the uniform flag stays a branch instead
// of both conversions being computed and selected.
template <typename T>
__device__ __forceinline__ void encode_store(const StepArgs &a, int64_t n, const float *o, bool whole_lines)
{
    T raw[9];
    if constexpr (sizeof(T) == 2) {
        // Range guard: a scaled value at or beyond the largest FP16 number
        // (65504) would be clamped by the conversion toward zero or become
        // infinite to nearest; bit 1 of the flag reports it (the storage
        // scale is too large for the flow).  NaNs pass fmaxf unnoticed:
        // the density check above reports them.
        const float mx = fmaxf(fmaxf(fmaxf(fabsf(o[0]), fabsf(o[1])), fmaxf(fabsf(o[2]), fabsf(o[3]))),
                               fmaxf(fmaxf(fabsf(o[4]), fabsf(o[5])),
                                     fmaxf(fmaxf(fabsf(o[6]), fabsf(o[7])), fabsf(o[8]))));
        if (mx >= 65504.f) atomicOr(a.flag, 2);
        if (a.f16_sr) {
            uint32_t h = (uint32_t)n * 0x9E3779B9u ^ __float_as_uint(o[0]);
            h ^= h >> 16;
            h *= 0x85EBCA6Bu;
            h ^= h >> 13;
            h *= 0xC2B2AE35u;
            h ^= h >> 16;
            float u[10];
#pragma unroll
            for (int d = 0; d < 9; ++d) {
                const uint32_t r = __builtin_amdgcn_alignbit(h, h, 3 * d) & 0x1FFFu;
                u[d] = __uint_as_float(__float_as_uint(o[d]) + r);
            }
            u[9] = 0.f;
#pragma unroll
            for (int d = 0; d < 9; d += 2) {
                const auto p = __builtin_amdgcn_cvt_pkrtz(u[d], u[d + 1]);
                raw[d] = (T)p[0];
                if (d + 1 < 9) raw[d + 1] = (T)p[1];
            }
            if (whole_lines) store9_stream(a, n, raw);
            else store9(a, n, raw);
            return;
        }
    }
#pragma unroll
    for (int d = 0; d < 9; ++d) raw[d] = (T)o[d];
    if (whole_lines) store9_stream(a, n, raw);
    else store9(a, n, raw);
}

template <typename T>
__device__ __forceinline__ void collide_store_shifted(const StepArgs &a, const float *g, int policy, int64_t e, int i,
                                                      int j, int64_t n)
{
    float o[9];
    collide_shifted<sizeof(T) == 2>(a, g, policy, e, i, j, o);
    encode_store<T>(a, n, o, false);
}

template <typename T, bool GATHER>
__device__ __forceinline__ void node_general_shifted(const StepArgs &a, int i, int j, int lane, const Nbhd &h)
{
    uint32_t m = lane_mask(h, lane);
    if (!(m & 1)) return;
    const bool edge = (i == 0) | (i == a.g.nr - 1) | (j == 0) | (j == a.g.nc - 1);
    const Idx x = lane_idx(h, lane);
    float g[9];
    int64_t e = 0;
    int policy = load_g<T, GATHER>(a, i, j, x, m, edge, e, g);
    if constexpr (sizeof(T) == 2) {
        // the closures above work on f - w; the FP16 collision on (f - w) S
#pragma unroll
        for (int d = 0; d < 9; ++d) g[d] *= a.scale;
    }
    collide_store_shifted<T>(a, g, policy, e, i, j, x.src[0]);
}

// ------------------------------------------------------------ step kernel --

template <typename T> struct ComputeOf { using type = float; };
template <> struct ComputeOf<double> { using type = double; };

// All wet nodes of one 64-row segment through the general path (perimeter
// nodes: frozen slots, Zou-He, open nodes; every node of the collide-only
// first step).
template <typename T, bool GATHER>
__device__ __forceinline__ void general_segment(const StepArgs &a, int j, int64_t i0, int lane)
{
    const Nbhd h = load_nbhd(a.g, j, i0);
    if (h.mid[1] == 0) return;
    const int i = (int)i0 + lane;
    if constexpr (sizeof(T) == 8) node_general_f64<GATHER>(a, i, j, lane, h);
    else node_general_shifted<T, GATHER>(a, i, j, lane, h);
}

// True when the segment of column j starting at row i0 holds a wet node on
// the domain perimeter (column 0 or NC-1, row 0 or NR-1).  Those segments
// belong to the perimeter kernel; the step kernel skips them.
__device__ __forceinline__ bool perimeter_segment(const Geometry &g, int j, int64_t i0, uint64_t mid)
{
    if (mid == 0) return false;
    if (j == 0 || j == g.nc - 1) return true;
    const int32_t last = g.nr - 1, r0 = (int32_t)i0;
    return (r0 == 0 && (mid & 1)) || ((uint32_t)(last - r0) < 64u && ((mid >> (last - r0)) & 1));
}

// One wave's share of a time step: R consecutive 64-row segments of
// column j starting at row row0 (callers pass row0 < ldr).  The
// collide-only first step (!GATHER) runs every node through the general
// path.  Otherwise segments holding a wet perimeter node are left to the
// perimeter kernel, and all others -- including those whose perimeter rows
// are dry walls, which the zero guard bits turn into bounce-back -- stream
// every direction when their 3x3 neighbourhood is all wet (a scalar,
// wave-uniform test) and otherwise select each source between the
// neighbour and the node's own opposite slot with the wave-uniform source
// masks.
//
// Loads first, for every segment and every lane, with no branch between
// them: a load under a branch leaves its register to be merged at the
// join, and the merge waits for the data, which serialised the segments
// (one round trip to memory per segment instead of one per wave).  Dry
// lanes therefore read their own slot (bounce-back address, always in
// range: the compact range holds one slot more than the lattice uses), and
// segments the wave does not compute (perimeter, all dry, past the
// column's end) read slot 0.  Collision and stores then run per segment
// under the wet lanes' mask.  Keeping the general path out of this kernel
// keeps its register allocation to what the fast path needs.
template <typename T, bool GATHER, int R>
__device__ __forceinline__ void step_wave(const StepArgs &a, int j, int64_t row0, int lane)
{
    using V = typename ComputeOf<T>::type;
    if (!GATHER) {
#pragma unroll 1
        for (int r = 0; r < R; ++r) {
            const int64_t i0 = row0 + 64 * r;
            if (i0 >= a.g.ldr) break;
            general_segment<T, GATHER>(a, j, i0, lane);
        }
        return;
    }
    T raw[R][9];
    int32_t own[R];
    uint64_t wet[R];   // the segment's wet lanes; 0 = segment not computed here
    uint32_t lines = 0;  // segments whose stores are whole, aligned lines
    // FP64: the wave issues its loads at raised priority, so waves about
    // to load go ahead of waves colliding (measured +1.7 % on the 8192^2
    // porous and open media; FP32, FP16 and the wet-run waves lose with it)
    if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(3);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t i0 = row0 + 64 * r;
        const bool in = r == 0 || i0 < a.g.ldr;
        Nbhd h = load_nbhd<true>(a.g, j, in ? i0 : row0);
        const bool go = in && !perimeter_segment(a.g, j, i0, h.mid[1]) && h.mid[1] != 0;
        const bool bulk = go && wave_is_bulk(h);
        if (bulk && (h.base[1] & 63) == 0) lines |= 1u << r;
        SrcMasks sm = src_masks(h);
#pragma unroll
        for (int d = 1; d < 9; ++d) sm.src[d] = bulk ? ~0ull : (go ? sm.src[d] & h.mid[1] : 0ull);
        if (!go) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                h.up[c] = h.mid[c] = h.dn[c] = 0;
                h.base[c] = 0;
            }
        }
        wet[r] = go ? h.mid[1] : 0ull;
        const Idx x = lane_idx(h, lane);
        own[r] = x.src[0];
        gather_interior_by(a, x, [&sm](int d) { return __builtin_amdgcn_inverse_ballot_w64(sm.src[d]); }, raw[r]);
    }
    // FP16, one segment: the scheduler packed loaded halves between the
    // loads and waited for them there; this keeps all nine loads ahead of
    // any use of their data (see run_wave).
    if constexpr (sizeof(T) == 2 && R == 1)
        asm volatile("" : "+v"(raw[0][0]), "+v"(raw[0][1]), "+v"(raw[0][2]), "+v"(raw[0][3]), "+v"(raw[0][4]),
                     "+v"(raw[0][5]), "+v"(raw[0][6]), "+v"(raw[0][7]), "+v"(raw[0][8]));
    if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(0);
    if constexpr (sizeof(T) == 4 && R == 2) {
        // FP32, two segments: collide both, then store direction by direction
        // across them.  The cache line shared by the two segments' ranges
        // then gets both of its parts back to back instead of nine stores
        // apart, by which time the wave's share of L2 may have evicted the
        // first part (measured: +6 % HBM writes in segment order).  With
        // more segments, or 2- and 8-byte storage, the longer live ranges
        // cost more in register spills than this saves.
        const bool act[2] = {__builtin_amdgcn_inverse_ballot_w64(wet[0]), __builtin_amdgcn_inverse_ballot_w64(wet[1])};
        T out[R][9];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!act[r]) continue;
            const int i = (int)row0 + 64 * r + lane;
            collide_shifted<false>(a, raw[r], -1, 0, i, j, out[r]);
        }
        uint32_t b0 = slot_bytes<T>(a, (uint32_t)own[0]), b1 = slot_bytes<T>(a, (uint32_t)own[1]);
        if ((lines & 3u) == 3u) {
            asm volatile("" : "+v"(b0), "+v"(b1)::"memory");
#pragma unroll
            for (int d = 0; d < 9; ++d) {
                __builtin_nontemporal_store(out[0][d], gaddr<T>(a.fout_plane[d], b0));
                __builtin_nontemporal_store(out[1][d], gaddr<T>(a.fout_plane[d], b1));
            }
            asm volatile("" ::: "memory");
        } else {
            asm volatile("" : "+v"(b0), "+v"(b1));
#pragma unroll
            for (int d = 0; d < 9; ++d) {
                if (act[0]) st((T *)a.fout_plane[d], b0, out[0][d]);
                if (act[1]) st((T *)a.fout_plane[d], b1, out[1][d]);
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (wet[r] == 0 || !__builtin_amdgcn_inverse_ballot_w64(wet[r])) continue;
            const int i = (int)row0 + 64 * r + lane;
            V p[9];
#pragma unroll
            for (int d = 0; d < 9; ++d) p[d] = (V)raw[r][d];  // FP16: scaled, see collide_shifted
            if constexpr (sizeof(T) == 8) {
                double out[9];
                collide_f64(a, p, -1, 0, i, j, out);
                if ((lines >> r) & 1) store9_stream(a, own[r], out);
                else store9(a, own[r], out);
            } else {
                float out[9];
                collide_shifted<sizeof(T) == 2>(a, p, -1, 0, i, j, out);
                encode_store<T>(a, own[r], out, (lines >> r) & 1);
            }
        }
    }
}

// One wave's share of a gathering step in the run layout: run t of the
// wet-run table, up to 64 consecutive wet rows of column j from row i0,
// all within 128 rows and within one 64-node block of the compact index.
// Lane l takes the l-th of them, so in porous media
// no lane idles on a solid node (a 64-row segment leaves those lanes
// empty).  A lane's own compact index is the run's first index plus l.
// Its row comes from the column's mask words: each wet row of the window
// writes itself to this wave's LDS slots at its rank.  The neighbour
// columns' words and segment bases are scalars (5 words and 3 bases per
// column around the window); every lane selects its own from them, and
// the neighbour indices and source masks follow as in lane_idx.  Lanes
// past the run read their first row's slot (bounce-back addresses, in
// range) and do not store.  Loads come before any branch (see step_wave).
template <typename T>
__device__ __forceinline__ void run_lane(const StepArgs &a, int i, int j, int s0, int n, int lane, bool act,
                                         int32_t rank0, int32_t rank2, uint32_t wet);

template <typename T>
__device__ __forceinline__ void run_wave(const StepArgs &a, int j, int i0, int s0, int n, int lane, int32_t *rows)
{
    const int w0 = i0 >> 6;
    // words w0-1 .. w0+3 of columns j-1, j, j+1 and the segment bases of
    // words w0 .. w0+2.  A column keeps three zero guard words past its last
    // segment, so these stay inside it with no clamping: constant offsets
    // from one pointer per column, merged loads, and one wait for all of
    // them (named together in the asm statement below).
    uint64_t W[3][5];
    int32_t B[3][3];
    const uint32_t wl = (uint32_t)a.g.wl;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const uint32_t kc = (uint32_t)(j + c) * wl + (uint32_t)w0;  // word w0-1 of the column
        const LBM_CONST uint64_t *col = (const LBM_CONST uint64_t *)a.g.wet + kc;
        const LBM_CONST int32_t *scol = (const LBM_CONST int32_t *)a.g.sb + kc + 1;
#pragma unroll
        for (int q = 0; q < 5; ++q) W[c][q] = col[q];
#pragma unroll
        for (int q = 0; q < 3; ++q) B[c][q] = scol[q];
    }
    asm volatile("" ::"s"(W[0][0]), "s"(W[0][1]), "s"(W[0][2]), "s"(W[0][3]), "s"(W[0][4]), "s"(W[1][0]),
                 "s"(W[1][1]), "s"(W[1][2]), "s"(W[1][3]), "s"(W[1][4]), "s"(W[2][0]), "s"(W[2][1]), "s"(W[2][2]),
                 "s"(W[2][3]), "s"(W[2][4]), "s"(B[0][0]), "s"(B[0][1]), "s"(B[0][2]), "s"(B[1][0]), "s"(B[1][1]),
                 "s"(B[1][2]), "s"(B[2][0]), "s"(B[2][1]), "s"(B[2][2]));
    // rows of the run: wet row r of the window goes to rows[rank]
    int prefix = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        uint64_t m = W[1][q + 1];
        if (q == 0) m &= ~0ull << (i0 & 63);
        if ((m >> lane) & 1) {
            const int r = prefix + (int)below(m);
            if (r < n) rows[r] = 64 * (w0 + q) + lane;
        }
        prefix += __builtin_popcountll(m);
    }
    __builtin_amdgcn_wave_barrier();
    const bool act = lane < n;
    const int i = act ? rows[lane] : i0;
    const int kk = (i >> 6) - w0;  // 0, 1 or 2
    const int bp = i & 63;
    auto pick = [kk](uint64_t x0, uint64_t x1, uint64_t x2) { return kk == 0 ? x0 : (kk == 1 ? x1 : x2); };
    uint32_t wu[3], wm[3], wd[3];
    int32_t rank[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const uint64_t cur = pick(W[c][1], W[c][2], W[c][3]);
        const uint64_t up = (cur << 1) | (pick(W[c][0], W[c][1], W[c][2]) >> 63);
        const uint64_t dn = (cur >> 1) | (pick(W[c][2], W[c][3], W[c][4]) << 63);
        wu[c] = (uint32_t)(up >> bp) & 1u;
        wm[c] = (uint32_t)(cur >> bp) & 1u;
        wd[c] = (uint32_t)(dn >> bp) & 1u;
        const int32_t base = kk == 0 ? B[c][0] : (kk == 1 ? B[c][1] : B[c][2]);
        rank[c] = base + (int32_t)__builtin_popcountll(cur & ((1ull << bp) - 1ull));
    }
    const uint32_t wet = act ? (wm[0] << 1) | (wd[1] << 2) | (wm[2] << 3) | (wu[1] << 4) | (wd[0] << 5) |
                                   (wd[2] << 6) | (wu[2] << 7) | (wu[0] << 8)
                             : 0u;
    run_lane<T>(a, i, j, s0, n, lane, act, rank[0], rank[2], wet);
}

// The rest of a wet-run wave once a lane knows its row i, whether it holds a
// node (act), the compact indices of its row in columns j-1 and j+1 (rank0,
// rank2: the first wet row at or below row i there) and its source-is-wet
// bits (bit d for direction d): gather (run_gather), then collide and store
// (run_finish).
template <typename T>
__device__ __forceinline__ void run_gather(const StepArgs &a, int s0, int lane, bool act, int32_t rank0, int32_t rank2,
                                           uint32_t wet, T *raw)
{
    const int32_t own = act ? s0 + lane : s0;
    Idx x;
    x.src[0] = own;
    x.src[1] = rank0;
    x.src[2] = own + 1;
    x.src[3] = rank2;
    x.src[4] = own - 1;
    x.src[5] = rank0 + (int32_t)((wet >> 1) & 1u);
    x.src[6] = rank2 + (int32_t)((wet >> 3) & 1u);
    x.src[7] = rank2 - 1;
    x.src[8] = rank0 - 1;
    gather_interior(a, x, wet, raw);
}

template <typename T>
__device__ __forceinline__ void run_finish(const StepArgs &a, int i, int j, int s0, int n, int lane, bool act,
                                           const T *raw)
{
    using V = typename ComputeOf<T>::type;
    if (!act) return;
    const int32_t own = s0 + lane;
    V p[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) p[d] = (V)raw[d];  // FP16: scaled, see collide_shifted
    if constexpr (sizeof(T) == 8) {
        // An FP64 wave whose data has arrived pauses about 128 cycles
        // before it collides and stores: spreading the memory requests
        // this way measured 1.6 % faster on the 8192^2 porous medium (wet
        // runs at 4 work-groups per CU), 0.3 % on the 4096^2 cavity.
        asm volatile("" ::"v"(p[0]));
        __builtin_amdgcn_s_sleep(2);
    }
    const bool lines = n == 64 && (s0 & 63) == 0;
    if constexpr (sizeof(T) == 8) {
        double out[9];
        collide_f64(a, p, -1, 0, i, j, out);
        if (lines) store9_stream(a, own, out);
        else store9(a, own, out);
    } else {
        float out[9];
        collide_shifted<sizeof(T) == 2>(a, p, -1, 0, i, j, out);
        encode_store<T>(a, own, out, lines);
    }
}

template <typename T>
__device__ __forceinline__ void run_lane(const StepArgs &a, int i, int j, int s0, int n, int lane, bool act,
                                         int32_t rank0, int32_t rank2, uint32_t wet)
{
    T raw[9];
    run_gather<T>(a, s0, lane, act, rank0, rank2, wet, raw);
    // All nine loads issue before any of their values is used: without
    // this the scheduler interleaves the first moments with the later
    // loads and waits for the early data in between, so a wave spends
    // several memory round trips where one suffices.
    asm volatile("" : "+v"(raw[0]), "+v"(raw[1]), "+v"(raw[2]), "+v"(raw[3]), "+v"(raw[4]), "+v"(raw[5]),
                 "+v"(raw[6]), "+v"(raw[7]), "+v"(raw[8]));
    run_finish<T>(a, i, j, s0, n, lane, act, raw);
}

// The time-step kernel of one lattice: wave w of the (XCD-remapped) grid
// owns segments [w mod wpc] * R .. of column w / wpc.
// (A multiply-high by host-computed magic numbers in place of the division
// by wpc measured 0.4 % slower for FP64 on the 8192^2 porous medium and no
// faster for FP32, so the division stays.)
template <typename T, bool GATHER, int R>
__global__ __launch_bounds__(kBlockRows) void step_kernel(StepArgs a)
{
    const int64_t b = xcd_remap(blockIdx.x, (uint32_t)a.nblocks, step_chunk(a.wpc));
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t gw = b * (kBlockRows / 64) + wave;
    const int j = (int)(gw / a.wpc);
    if (j >= a.g.nc) return;
    const int64_t row0 = (gw - (int64_t)j * a.wpc) * (64 * R);
    step_wave<T, GATHER, R>(a, j, row0, lane);
}

// The time-step kernel of one lattice in the run layout: one wave per run
// of the wet-run table (XCD-remapped, four waves per work-group).
template <typename T>
__global__ __launch_bounds__(kBlockRows) void run_kernel(StepArgs a)
{
    __shared__ int32_t rows[kBlockRows / 64][64];
    // every kernel argument the prologue needs, in one round trip
    asm volatile("" ::"s"(a.nblocks), "s"(a.nruns), "s"(a.runs), "s"(a.g.wet), "s"(a.g.sb), "s"(a.g.wl));
    const int64_t b = xcd_remap(blockIdx.x, (uint32_t)a.nblocks, kXcdChunkRuns<T>);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t = b * (kBlockRows / 64) + wave;
    if (t >= a.nruns) return;
    const LBM_CONST int32_t *tp = (const LBM_CONST int32_t *)a.runs + 4 * t;
    run_wave<T>(a, __builtin_amdgcn_readfirstlane(tp[0]), __builtin_amdgcn_readfirstlane(tp[1]),
                __builtin_amdgcn_readfirstlane(tp[2]), __builtin_amdgcn_readfirstlane(tp[3]), threadIdx.x & 63,
                rows[wave]);
}

// The wet runs with every lane's addressing read from the tables built with
// the geometry (StepArgs::trun, tlane): the run entry (scalar) and the
// lane's record (one 4-byte vector load) are the only loads before the
// population loads, and they issue together.  An FP16 wave takes two
// consecutive runs and issues the loads of both before either collides
// (kTabRunsPerWave); its XCD chunks are halved so that a chunk still holds
// the same runs.  With the runs cut at the layout's blocks this measured
// +7 % for FP16 on the 8192^2 porous medium (same bits); before, when runs
// straddled the blocks, it was no faster, and for FP32 it was slower.
template <typename T>
constexpr int kTabRunsPerWave = sizeof(T) == 2 ? 2 : 1;

// One tabled run's lane: row, neighbour ranks and source-is-wet bits from
// the lane's record (see StepArgs::tlane), then the gather.
struct TabLane {
    int i, j, s0, n;
    bool act;
    int32_t rank0, rank2;
    uint32_t wet;
};

__device__ __forceinline__ TabLane tab_lane(const StepArgs &a, int64_t t, int lane)
{
    const LBM_CONST int32_t *tp = (const LBM_CONST int32_t *)a.trun + 8 * t;
    const uint32_t rec = ((const LBM_GLOBAL uint32_t *)a.tlane)[64 * t + lane];
    TabLane r;
    r.j = __builtin_amdgcn_readfirstlane(tp[0]);
    const int i0 = __builtin_amdgcn_readfirstlane(tp[1]);
    r.s0 = __builtin_amdgcn_readfirstlane(tp[2]);
    r.n = __builtin_amdgcn_readfirstlane(tp[3]);
    const int32_t bl = __builtin_amdgcn_readfirstlane(tp[4]);
    const int32_t br = __builtin_amdgcn_readfirstlane(tp[5]);
    r.act = lane < r.n;
    r.i = i0 + (int)(rec & 127u);
    r.rank0 = bl + (int32_t)((rec >> 7) & 127u);
    r.rank2 = br + (int32_t)((rec >> 14) & 127u);
    r.wet = r.act ? ((rec >> 21) & 255u) << 1 : 0u;
    return r;
}

template <typename T>
__global__ __launch_bounds__(kBlockRows) void tab_kernel(StepArgs a)
{
    asm volatile("" ::"s"(a.nblocks), "s"(a.nruns), "s"(a.trun), "s"(a.tlane));  // one round trip
    const int64_t b = xcd_remap(blockIdx.x, (uint32_t)a.nblocks, kXcdChunkRuns<T> / kTabRunsPerWave<T>);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if constexpr (kTabRunsPerWave<T> == 2) {
        // runs t0 and t0 + 1; a last odd run reads its own run twice and
        // stores once
        const int64_t t0 = 2 * (b * (kBlockRows / 64) + wave);
        if (t0 >= a.nruns) return;
        const bool two = t0 + 1 < a.nruns;
        const int lane = threadIdx.x & 63;
        const TabLane x = tab_lane(a, t0, lane);
        TabLane y = tab_lane(a, two ? t0 + 1 : t0, lane);
        y.act = y.act && two;
        y.wet = two ? y.wet : 0u;
        T rx[9], ry[9];
        run_gather<T>(a, x.s0, lane, x.act, x.rank0, x.rank2, x.wet, rx);
        run_gather<T>(a, y.s0, lane, y.act, y.rank0, y.rank2, y.wet, ry);
        // all eighteen loads issue before any value is used (see run_lane)
        asm volatile("" : "+v"(rx[0]), "+v"(rx[1]), "+v"(rx[2]), "+v"(rx[3]), "+v"(rx[4]), "+v"(rx[5]),
                     "+v"(rx[6]), "+v"(rx[7]), "+v"(rx[8]), "+v"(ry[0]), "+v"(ry[1]), "+v"(ry[2]), "+v"(ry[3]),
                     "+v"(ry[4]), "+v"(ry[5]), "+v"(ry[6]), "+v"(ry[7]), "+v"(ry[8]));
        run_finish<T>(a, x.i, x.j, x.s0, x.n, lane, x.act, rx);
        run_finish<T>(a, y.i, y.j, y.s0, y.n, lane, y.act, ry);
        return;
    }
    const int64_t t = b * (kBlockRows / 64) + wave;
    if (t >= a.nruns) return;
    const int lane = threadIdx.x & 63;
    const TabLane x = tab_lane(a, t, lane);
    run_lane<T>(a, x.i, x.j, x.s0, x.n, lane, x.act, x.rank0, x.rank2, x.wet);
}

// The perimeter share of a gathering step of one lattice: one wave per
// entry (column, segment) of the perimeter table, which lists exactly the
// segments the step kernels skip (perimeter_segment), so every wave works.
template <typename T>
__global__ __launch_bounds__(kBlockRows) void perimeter_kernel(StepArgs a)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t w = (int64_t)blockIdx.x * (kBlockRows / 64) + wave;
    if (w >= a.nperim) return;
    const LBM_CONST int32_t *item = (const LBM_CONST int32_t *)a.perim + 2 * w;
    general_segment<T, true>(a, item[0], 64 * (int64_t)item[1], threadIdx.x & 63);
}

// Entry t of a batch tile table (sample, column, first row) and the step
// arguments of sample s, read through the constant address space: neither
// changes while a kernel runs, so the wave-uniform reads become scalar
// loads and the arguments live in scalar registers.
struct TileRef {
    int s, j, row0;
};

__device__ __forceinline__ TileRef tile_at(const int32_t *table, int64_t t)
{
    const LBM_CONST int32_t *p = (const LBM_CONST int32_t *)table + 3 * t;
    return {__builtin_amdgcn_readfirstlane(p[0]), __builtin_amdgcn_readfirstlane(p[1]),
            __builtin_amdgcn_readfirstlane(p[2])};
}

__device__ __forceinline__ StepArgs args_at(const BatchArgs &bt, int s)
{
    static_assert(sizeof(StepArgs) % 4 == 0, "word copy");
    const LBM_CONST uint32_t *src = (const LBM_CONST uint32_t *)(bt.args + 2 * (int64_t)s + bt.parity);
    StepArgs a;
    uint32_t *dst = (uint32_t *)&a;
#pragma unroll
    for (unsigned k = 0; k < sizeof(StepArgs) / 4; ++k) dst[k] = src[k];
    return a;
}

// The perimeter share of a gathering step of a batch: one wave per entry
// (sample, column, first row) of the perimeter table, which lists every
// 64-row segment holding a wet perimeter node.
// A gathering time step with every wet node through the general path, one
// wave per 64-row segment: the step of lattices with partially saturated
// bounce-back obstacles (psbb_gather / psbb_update), FP64 storage.
__global__ __launch_bounds__(kBlockRows) void general_kernel(StepArgs a)
{
    const Tile t = tile_of(a);
    if (!t.valid) return;
    general_segment<double, true>(a, t.j, t.i0, t.lane);
}

// One time step of every lattice of an ensemble in wet-run waves: one wave
// per entry of the ensemble's run table (sample, then the sample's own run
// entry), the XCDs taking the table in chunks (xcd_remap).  The samples'
// perimeter segments run in batch_perimeter_kernel, as for the tiles.
template <typename T>
__global__ __launch_bounds__(kBlockRows) void batch_run_kernel(BatchArgs bt)
{
    __shared__ int32_t rows[kBlockRows / 64][64];
    asm volatile("" ::"s"(bt.nruns), "s"(bt.runs), "s"(bt.args), "s"(bt.parity));  // one round trip
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t nb = (uint32_t)((bt.nruns + kBlockRows / 64 - 1) / (kBlockRows / 64));
    const int64_t t = (int64_t)xcd_remap(blockIdx.x, nb) * (kBlockRows / 64) + wave;
    if (t >= bt.nruns) return;
    const LBM_CONST int32_t *e = (const LBM_CONST int32_t *)bt.runs + 5 * t;
    const int s = __builtin_amdgcn_readfirstlane(e[0]);
    const int j = __builtin_amdgcn_readfirstlane(e[1]);
    const int i0 = __builtin_amdgcn_readfirstlane(e[2]);
    const int s0 = __builtin_amdgcn_readfirstlane(e[3]);
    const int n = __builtin_amdgcn_readfirstlane(e[4]);
    const StepArgs a = args_at(bt, s);
    run_wave<T>(a, j, i0, s0, n, threadIdx.x & 63, rows[wave]);
}

template <typename T>
__global__ __launch_bounds__(kBlockRows) void batch_perimeter_kernel(BatchArgs bt)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t = (int64_t)blockIdx.x * (kBlockRows / 64) + wave;
    if (t >= bt.nptiles) return;
    const TileRef tr = tile_at(bt.ptiles, t);
    const StepArgs a = args_at(bt, tr.s);
    general_segment<T, true>(a, tr.j, tr.row0, threadIdx.x & 63);
}

// The time-step kernel of a batch of independent lattices: a persistent
// grid walks the tile table (sample, column, first row; R segments per
// tile).  The table is split into 8 contiguous chunks, one per XCD (blocks
// are dealt round-robin to the XCDs), so each XCD's L2 sees consecutive
// tiles of the same lattices.  Each sample's step arguments are prepared on
// the host (one set per buffer parity) and read in place through a
// wave-uniform pointer, so they live in scalar registers or are reloaded
// from the scalar cache instead of occupying vector registers.
template <typename T, bool GATHER, int R, bool DIRECT>
__global__ __launch_bounds__(kBlockRows) void batch_kernel(BatchArgs bt)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if constexpr (DIRECT) {
        // one wave per tile; the XCDs take the table in chunks (xcd_remap).
        // The grid has ceil(ntiles / 4) work-groups (one when the table is
        // empty), counted here from ntiles rather than read from the
        // dispatch packet: one scalar round trip less before the tile.
        asm volatile("" ::"s"(bt.ntiles), "s"(bt.tiles), "s"(bt.args), "s"(bt.parity));  // one round trip
        const uint32_t nb = (uint32_t)((bt.ntiles + kBlockRows / 64 - 1) / (kBlockRows / 64));
        const int64_t t = (int64_t)xcd_remap(blockIdx.x, nb) * (kBlockRows / 64) + wave;
        if (t >= bt.ntiles) return;
        const TileRef tr = tile_at(bt.tiles, t);
        const StepArgs a = args_at(bt, tr.s);
        // the direction pointers are loaded with the sample's other arguments,
        // not after its mask words arrive
        asm volatile("" ::"s"(a.fin_plane[0]), "s"(a.fin_plane[1]), "s"(a.fin_plane[2]), "s"(a.fin_plane[5]),
                     "s"(a.fin_plane[6]), "s"(a.block_mask), "s"(a.two_planes));
        step_wave<T, GATHER, R>(a, tr.j, tr.row0, lane);
        return;
    }
    const int64_t xcd = blockIdx.x & 7;
    const int64_t per_xcd_blocks = (int64_t)gridDim.x >> 3;
    const int64_t local_wave = (int64_t)(blockIdx.x >> 3) * (kBlockRows / 64) + wave;
    const int64_t stride = per_xcd_blocks * (kBlockRows / 64);
    const int64_t chunk = (bt.ntiles + 7) >> 3;
    const int64_t first = xcd * chunk;
    const int64_t last = first + chunk < bt.ntiles ? first + chunk : bt.ntiles;
#pragma unroll 1
    for (int64_t t = first + local_wave; t < last; t += stride) {
        const TileRef tr = tile_at(bt.tiles, t);
        const StepArgs a = args_at(bt, tr.s);
        step_wave<T, GATHER, R>(a, tr.j, tr.row0, lane);
    }
}

// ------------------------------------------------- readout / reductions --

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sums NC values of every thread over the block; thread 0 writes them.
template <int NC>
__device__ __forceinline__ void block_sum(double *v, double *dst)
{
    __shared__ double red[kBlockRows / 64][NC];
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        double s = wave_sum(v[c]);
        if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            double s = 0.0;
            for (int w = 0; w < kBlockRows / 64; ++w) s += red[w][c];
            dst[c] = s;
        }
    }
}

// Pre-collision populations P of wet node (i, j) in FP64: gathered from F*
// (GATHER) or read from the stored state as is (!GATHER).  Returns the
// open-boundary policy of the node (-1 for all other nodes).
template <typename T, bool GATHER>
__device__ __forceinline__ int node_state(const StepArgs &a, int i, int j, int lane, const Nbhd &h, uint32_t m,
                                          double *f)
{
    const bool edge = (i == 0) | (i == a.g.nr - 1) | (j == 0) | (j == a.g.nc - 1);
    const Idx x = lane_idx(h, lane);
    int64_t e = 0;
    int policy;
    if constexpr (sizeof(T) == 8) {
        policy = load_p64<GATHER>(a, i, j, x, m, edge, e, f);
    } else {
        float g[9];
        policy = load_g<T, GATHER>(a, i, j, x, m, edge, e, g);
#pragma unroll
        for (int d = 0; d < 9; ++d) f[d] = (double)g[d] + kW[d];
    }
    return policy;
}

// Adds a node's contribution to the five sums: wet mass, fluid momentum
// x / y, fluid mass, fluid count.
__device__ __forceinline__ void add_sums(const double *f, int policy, double *acc)
{
    double rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    acc[0] += rho;
    if (policy < 0) {
        acc[1] += f[1] + f[5] + f[8] - (f[3] + f[6] + f[7]);
        acc[2] += f[2] + f[5] + f[6] - (f[4] + f[7] + f[8]);
        acc[3] += rho;
        acc[4] += 1.0;
    }
}

// Writes the pre-collision populations P in FP64 in the reference layout
// (zeros on the other nodes) and/or per-block partial sums (add_sums).
template <typename T, bool GATHER>
__global__ __launch_bounds__(kBlockRows) void readout(StepArgs a, double *out, double *partials)
{
    Tile t = tile_of(a);
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (t.valid) {
        const Nbhd h = load_nbhd(a.g, t.j, t.i0);
        uint32_t m = lane_mask(h, t.lane);
        const int i = t.i, j = t.j;
        const int64_t N = (int64_t)a.g.nr * a.g.nc;
        const int64_t q = (int64_t)j * a.g.nr + i;
        if (m & 1) {
            double f[9];
            const int policy = node_state<T, GATHER>(a, i, j, t.lane, h, m, f);
            if (out) {
#pragma unroll
                for (int d = 0; d < 9; ++d) out[d * N + q] = f[d];
            }
            add_sums(f, policy, acc);
        } else if (out && i < a.g.nr) {
#pragma unroll
            for (int d = 0; d < 9; ++d) out[d * N + q] = 0.0;
        }
    }
    if (partials) block_sum<5>(acc, partials + (int64_t)blockIdx.x * 5);
}

// Batch reductions, pass 1: one wave per tile of the batch's tile table
// sums its nodes (add_sums, FP64) and lane 0 writes partials[5 t + k].
template <typename T, bool GATHER, int R>
__global__ __launch_bounds__(kBlockRows) void batch_partials(BatchArgs bt, double *partials)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
#pragma unroll 1
    for (int64_t t = (int64_t)blockIdx.x * (kBlockRows / 64) + wave; t < bt.ntiles;
         t += (int64_t)gridDim.x * (kBlockRows / 64)) {
        const TileRef tr = tile_at(bt.tiles, t);
        const int j = tr.j, row0 = tr.row0;
        const StepArgs a = args_at(bt, tr.s);
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int r = 0; r < R; ++r) {
            const int64_t i0 = row0 + 64 * r;
            if (i0 >= a.g.ldr) break;
            const Nbhd h = load_nbhd(a.g, j, i0);
            const uint32_t m = lane_mask(h, lane);
            if (m & 1) {
                double f[9];
                const int policy = node_state<T, GATHER>(a, (int)i0 + lane, j, lane, h, m, f);
                add_sums(f, policy, acc);
            }
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double v = wave_sum(acc[k]);
            if (lane == 0) partials[5 * t + k] = v;
        }
    }
}

// Batch reductions, pass 2: block s folds the partials of sample s's
// tiles [tile_begin[s], tile_begin[s + 1]) in a fixed order.
__global__ __launch_bounds__(kBlockRows) void batch_fold(const double *partials, const int64_t *tile_begin,
                                                         double *out)
{
    const int64_t s = blockIdx.x;
    const int64_t t0 = tile_begin[s], t1 = tile_begin[s + 1];
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t t = t0 + threadIdx.x; t < t1; t += kBlockRows)
#pragma unroll
        for (int k = 0; k < 5; ++k) v[k] += partials[5 * t + k];
    block_sum<5>(v, out + 5 * s);
}

// Converts a reference-layout FP64 state to the storage format on the wet
// nodes and fills the frozen slots of the perimeter.
template <typename T>
__global__ __launch_bounds__(kBlockRows) void upload(StepArgs a, const double *src, T *frozen)
{
    Tile t = tile_of(a);
    if (!t.valid) return;
    const Nbhd h = load_nbhd(a.g, t.j, t.i0);
    if (!((h.mid[1] >> t.lane) & 1)) return;
    const int i = t.i, j = t.j;
    const int64_t N = (int64_t)a.g.nr * a.g.nc, q = (int64_t)j * a.g.nr + i;
    const uint32_t off = slot_bytes<T>(a, (uint32_t)lane_idx(h, t.lane).src[0]);
    const bool edge = (i == 0) | (i == a.g.nr - 1) | (j == 0) | (j == a.g.nc - 1);
    const int64_t e = edge ? edge_slot(i, j, a.g.nr, a.g.nc) : 0;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        double f = src[d * N + q];
        T v;
        if constexpr (sizeof(T) == 8) v = f;
        else v = from_g<T>((float)(f - kW[d]), a.scale);
        st((T *)a.fout_plane[d], off, v);
        if (edge) frozen[d * a.g.E + e] = v;
    }
}

// Lattice.hpp:294-328: feq(1, 0) = w_d on the wet nodes (0 in shifted
// storage), rho = 1 and u = 0 everywhere.
template <typename T>
__global__ __launch_bounds__(kBlockRows) void init_eq(StepArgs a, T *frozen, double *macro)
{
    Tile t = tile_of(a);
    if (!t.valid) return;
    const int i = t.i, j = t.j;
    if (macro && i < a.g.nr) {
        const int64_t N = (int64_t)a.g.nr * a.g.nc, q = (int64_t)j * a.g.nr + i;
        macro[q] = 1.0;
        macro[N + q] = 0.0;
        macro[2 * N + q] = 0.0;
    }
    const Nbhd h = load_nbhd(a.g, t.j, t.i0);
    if (!((h.mid[1] >> t.lane) & 1)) return;
    const uint32_t off = slot_bytes<T>(a, (uint32_t)lane_idx(h, t.lane).src[0]);
    const bool edge = (i == 0) | (i == a.g.nr - 1) | (j == 0) | (j == a.g.nc - 1);
    const int64_t e = edge ? edge_slot(i, j, a.g.nr, a.g.nc) : 0;
    double e9[9];
    feq64(1.0, 0.0, 0.0, e9);
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        T v;
        if constexpr (sizeof(T) == 8) v = e9[d];
        else v = from_g<T>((float)(e9[d] - kW[d]), a.scale);
        st((T *)a.fout_plane[d], off, v);
        if (edge) frozen[d * a.g.E + e] = v;
    }
}

// rho = 1, u = 0 on every node of a [3][n] field.
__global__ __launch_bounds__(kBlockRows) void fill_macro(double *macro, int64_t n)
{
    const int64_t q = (int64_t)blockIdx.x * kBlockRows + threadIdx.x;
    if (q >= n) return;
    macro[q] = 1.0;
    macro[n + q] = 0.0;
    macro[2 * n + q] = 0.0;
}

// Compact index of wet node (i, j), any thread.
__device__ __forceinline__ int64_t cidx_of(const Geometry &g, int64_t i, int64_t j)
{
    const int64_t w = (j + 1) * g.wl + (i >> 6) + 1;
    const uint64_t word = g.wet[w];
    const uint64_t lowbits = (i & 63) ? (word & ((1ull << (i & 63)) - 1)) : 0ull;
    return (int64_t)g.sb[w] + __popcll(lowbits);
}

template <typename T>
__device__ __forceinline__ double fstar_at(const StepArgs &a, int64_t i, int64_t j, int d)
{
    const T v = ld((const T *)a.fin_plane[d], slot_bytes<T>(a, (uint32_t)cidx_of(a.g, i, j)));
    if constexpr (sizeof(T) == 8) return (double)v;
    else return (double)to_g<T>(v, a.inv_scale) + kW[d];
}

// FlowAnalyzer.hpp:386-430 for one point per thread, then a block sum.
template <typename T>
__global__ __launch_bounds__(kBlockRows) void lift_drag(StepArgs a, const int64_t *ij, const int32_t *masks,
                                                        int64_t n, double *per_point, double *partials)
{
    int64_t q = (int64_t)blockIdx.x * kBlockRows + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (q < n) {
        const int64_t i = ij[2 * q], j = ij[2 * q + 1];
        const int32_t msk = masks[q];
        const double h = sqrt(2.0);
        double local_lift = 0., local_drag = 0.;
        if (msk & (1 << 1)) local_drag -= fstar_at<T>(a, i, j + 1, 3);
        if (msk & (1 << 2)) local_lift -= fstar_at<T>(a, i - 1, j, 4);
        if (msk & (1 << 3)) local_drag += fstar_at<T>(a, i, j - 1, 1);
        if (msk & (1 << 4)) local_lift += fstar_at<T>(a, i + 1, j, 2);
        if (msk & (1 << 5)) {
            double v = fstar_at<T>(a, i - 1, j + 1, 7);
            local_drag -= (v) * h / 2.0;
            local_lift -= (v) * h / 2.0;
        }
        if (msk & (1 << 6)) {
            double v = fstar_at<T>(a, i - 1, j - 1, 8);
            local_drag += (v) * h / 2.0;
            local_lift -= (v) * h / 2.0;
        }
        if (msk & (1 << 7)) {
            double v = fstar_at<T>(a, i + 1, j - 1, 5);
            local_drag += (v) * h / 2.0;
            local_lift += (v) * h / 2.0;
        }
        if (msk & (1 << 8)) {
            double v = fstar_at<T>(a, i + 1, j + 1, 6);
            local_drag -= (v) * h / 2.0;
            local_lift += (v) * h / 2.0;
        }
        if (per_point) {
            per_point[2 * q] = local_lift;
            per_point[2 * q + 1] = local_drag;
        }
        acc[0] = local_lift;
        acc[1] = local_drag;
    }
    block_sum<2>(acc, partials + (int64_t)blockIdx.x * 2);
}

// Second pass: one block folds the per-block partials in a fixed order.
__global__ __launch_bounds__(kBlockRows) void reduce_partials(const double *partials, int64_t nb, int32_t ncomp,
                                                              double *out)
{
    for (int c = 0; c < ncomp; ++c) {
        double v = 0.0;
        for (int64_t b = threadIdx.x; b < nb; b += kBlockRows) v += partials[b * ncomp + c];
        double s[1] = {v};
        double r[1];
        block_sum<1>(s, r);
        if (threadIdx.x == 0) out[c] = r[0];
        __syncthreads();
    }
}

int64_t blocks_of(const StepArgs &a) { return (int64_t)a.segs * a.g.nc; }

template <typename K>
hipError_t launch(K kernel, int64_t nb, hipStream_t s, StepArgs a)
{
    a.nblocks = nb;
    set_planes(a);
    hipLaunchKernelGGL(kernel, dim3((unsigned)nb), dim3(kBlockRows), a.launch_lds, s, a);
    return hipGetLastError();
}

int64_t step_blocks(const StepArgs &a, int32_t R)
{
    const int64_t wpc = (a.g.ldr / 64 + R - 1) / R;
    const int64_t waves = wpc * a.g.nc;
    return (waves + (kBlockRows / 64) - 1) / (kBlockRows / 64);
}

template <typename T>
hipError_t launch_perimeter(const StepArgs &a, hipStream_t s)
{
    if (a.nperim == 0) return hipSuccess;
    return launch(perimeter_kernel<T>, (a.nperim + kBlockRows / 64 - 1) / (kBlockRows / 64), s, a);
}

template <typename T, int R>
hipError_t launch_step_t(bool gather, StepArgs a, hipStream_t s)
{
    a.wpc = (int32_t)((a.g.ldr / 64 + R - 1) / R);
    const int64_t nb = step_blocks(a, R);
    if (!gather) return launch(step_kernel<T, false, R>, nb, s, a);
    hipError_t e = launch(step_kernel<T, true, R>, nb, s, a);
    if (e != hipSuccess) return e;
    return launch_perimeter<T>(a, s);
}

template <typename T>
hipError_t launch_step_tab(StepArgs a, hipStream_t s)
{
    hipError_t e = hipSuccess;
    if (a.nruns > 0 && (!a.trun || !a.tlane)) return hipErrorInvalidValue;  // tables not built
    const int64_t per_block = (kBlockRows / 64) * kTabRunsPerWave<T>;
    if (a.nruns > 0) e = launch(tab_kernel<T>, (a.nruns + per_block - 1) / per_block, s, a);
    if (e != hipSuccess) return e;
    return launch_perimeter<T>(a, s);
}

template <typename T>
hipError_t launch_step_runs(StepArgs a, hipStream_t s)
{
    hipError_t e = hipSuccess;
    if (a.nruns > 0) e = launch(run_kernel<T>, (a.nruns + kBlockRows / 64 - 1) / (kBlockRows / 64), s, a);
    if (e != hipSuccess) return e;
    return launch_perimeter<T>(a, s);
}

template <typename T>
hipError_t launch_step_r(bool gather, int32_t R, const StepArgs &a, hipStream_t s)
{
    if (R == kRunLaunch || R == kRunTableLaunch) {
        if (gather) return R == kRunLaunch ? launch_step_runs<T>(a, s) : launch_step_tab<T>(a, s);
        R = default_segments(sizeof(T) == 8 ? kF64 : kF32);
    }
    switch (R) {
    case 1: return launch_step_t<T, 1>(gather, a, s);
    case 2: return launch_step_t<T, 2>(gather, a, s);
    case 8: return launch_step_t<T, 8>(gather, a, s);
    default: return launch_step_t<T, 4>(gather, a, s);
    }
}

template <typename T, int R>
hipError_t launch_batch_t(bool gather, const BatchArgs &b, int64_t blocks, hipStream_t s)
{
    dim3 grid((unsigned)blocks), block(kBlockRows);
    if (!gather) {
        if (b.direct) hipLaunchKernelGGL((batch_kernel<T, false, R, true>), grid, block, 0, s, b);
        else hipLaunchKernelGGL((batch_kernel<T, false, R, false>), grid, block, 0, s, b);
        return hipGetLastError();
    }
    if (b.direct) hipLaunchKernelGGL((batch_kernel<T, true, R, true>), grid, block, 0, s, b);
    else hipLaunchKernelGGL((batch_kernel<T, true, R, false>), grid, block, 0, s, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || b.nptiles == 0) return e;
    const int64_t pb = (b.nptiles + kBlockRows / 64 - 1) / (kBlockRows / 64);
    hipLaunchKernelGGL((batch_perimeter_kernel<T>), dim3((unsigned)pb), block, 0, s, b);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_r(bool gather, int32_t R, const BatchArgs &b, int64_t blocks, hipStream_t s)
{
    if (R == kRunLaunch) {
        if (!gather) return launch_batch_t<T, 1>(false, b, blocks, s);
        if (b.nruns > 0) {
            const int64_t nb = (b.nruns + kBlockRows / 64 - 1) / (kBlockRows / 64);
            hipLaunchKernelGGL((batch_run_kernel<T>), dim3((unsigned)nb), dim3(kBlockRows), 0, s, b);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        if (b.nptiles == 0) return hipSuccess;
        const int64_t pb = (b.nptiles + kBlockRows / 64 - 1) / (kBlockRows / 64);
        hipLaunchKernelGGL((batch_perimeter_kernel<T>), dim3((unsigned)pb), dim3(kBlockRows), 0, s, b);
        return hipGetLastError();
    }
    switch (R) {
    case 1: return launch_batch_t<T, 1>(gather, b, blocks, s);
    case 2: return launch_batch_t<T, 2>(gather, b, blocks, s);
    case 8: return launch_batch_t<T, 8>(gather, b, blocks, s);
    default: return launch_batch_t<T, 4>(gather, b, blocks, s);
    }
}

template <typename T, int R>
void launch_batch_partials_t(bool gather, const BatchArgs &b, double *partials, hipStream_t s)
{
    const int64_t nb = std::min<int64_t>((b.ntiles + 3) / 4, 8192);
    if (gather) hipLaunchKernelGGL((batch_partials<T, true, R>), dim3((unsigned)nb), dim3(kBlockRows), 0, s, b, partials);
    else hipLaunchKernelGGL((batch_partials<T, false, R>), dim3((unsigned)nb), dim3(kBlockRows), 0, s, b, partials);
}

template <typename T>
void launch_batch_partials_r(bool gather, int32_t R, const BatchArgs &b, double *partials, hipStream_t s)
{
    switch (R) {
    case 1: launch_batch_partials_t<T, 1>(gather, b, partials, s); break;
    case 2: launch_batch_partials_t<T, 2>(gather, b, partials, s); break;
    case 8: launch_batch_partials_t<T, 8>(gather, b, partials, s); break;
    default: launch_batch_partials_t<T, 4>(gather, b, partials, s); break;
    }
}

}  // namespace

void set_planes(StepArgs &a)
{
    const size_t stride = (size_t)(a.block_mask ? 64 : a.g.plane) * (size_t)a.esize;
    a.two_planes = (uint32_t)(2 * stride);
    for (int d = 0; d < 9; ++d) {
        a.fin_plane[d] = a.fin ? (const char *)a.fin + (size_t)d * stride : nullptr;
        a.fout_plane[d] = a.fout ? (char *)a.fout + (size_t)d * stride : nullptr;
    }
}

hipError_t launch_batch_reduce(int32_t storage, bool gather, int32_t R, const BatchArgs &b,
                               const int64_t *tile_begin, int64_t nsamples, double *partials, double *out,
                               hipStream_t s)
{
    if (R <= 0) R = default_segments(storage);
    if (b.ntiles > 0) {
        switch (storage) {
        case kF64: launch_batch_partials_r<double>(gather, R, b, partials, s); break;
        case kF32: launch_batch_partials_r<float>(gather, R, b, partials, s); break;
        default: launch_batch_partials_r<_Float16>(gather, R, b, partials, s); break;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(batch_fold, dim3((unsigned)nsamples), dim3(kBlockRows), 0, s, partials, tile_begin, out);
    return hipGetLastError();
}

// The shapes the autotuner picks on the 8192^2 porous and open media.
int32_t default_segments(int32_t storage) { return storage == kF16 ? 2 : 1; }

// For those default shapes, measured on the 8192^2 porous and open media and
// the 4096^2 cavity against no limit: FP32 R = 1 at 7 work-groups per CU
// 1-3 % faster in all three, FP64 R = 1 at 5 6-10 % faster, FP16 R = 2 at 7
// 9-17 % faster on the 8192^2 media and 4 % slower on the cavity.
int32_t default_occupancy(int32_t storage) { return storage == kF64 ? 5 : 7; }

hipError_t launch_step(int32_t storage, bool gather, const StepArgs &a, hipStream_t s, int32_t R)
{
    if (a.psbb && gather) return launch(general_kernel, blocks_of(a), s, a);
    if (R <= 0) R = default_segments(storage);
    switch (storage) {
    case kF64: return launch_step_r<double>(gather, R, a, s);
    case kF32: return launch_step_r<float>(gather, R, a, s);
    default: return launch_step_r<_Float16>(gather, R, a, s);
    }
}

hipError_t launch_batch_step(int32_t storage, bool gather, int32_t R, const BatchArgs &b0, int64_t blocks,
                             hipStream_t s)
{
    if (R <= 0) R = default_segments(storage);
    BatchArgs b = b0;
    if (R == kRunLaunch) blocks = 0;  // one wave per run; the first step takes one wave per tile
    b.direct = blocks <= 0;
    if (b.direct) blocks = std::max<int64_t>(1, (b.ntiles + kBlockRows / 64 - 1) / (kBlockRows / 64));
    else blocks = ((blocks + 7) / 8) * 8;  // whole rounds of the 8 XCDs
    switch (storage) {
    case kF64: return launch_batch_r<double>(gather, R, b, blocks, s);
    case kF32: return launch_batch_r<float>(gather, R, b, blocks, s);
    default: return launch_batch_r<_Float16>(gather, R, b, blocks, s);
    }
}

int64_t readout_blocks(const StepArgs &a) { return blocks_of(a); }

hipError_t launch_readout(int32_t storage, bool gather, const StepArgs &a0, double *out, double *partials,
                          hipStream_t s)
{
    StepArgs a = a0;
    a.nblocks = blocks_of(a);
    set_planes(a);
    dim3 grid((unsigned)a.nblocks), block(kBlockRows);
    switch (storage) {
    case kF64:
        if (gather) hipLaunchKernelGGL((readout<double, true>), grid, block, 0, s, a, out, partials);
        else hipLaunchKernelGGL((readout<double, false>), grid, block, 0, s, a, out, partials);
        break;
    case kF32:
        if (gather) hipLaunchKernelGGL((readout<float, true>), grid, block, 0, s, a, out, partials);
        else hipLaunchKernelGGL((readout<float, false>), grid, block, 0, s, a, out, partials);
        break;
    default:
        if (gather) hipLaunchKernelGGL((readout<_Float16, true>), grid, block, 0, s, a, out, partials);
        else hipLaunchKernelGGL((readout<_Float16, false>), grid, block, 0, s, a, out, partials);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_upload(int32_t storage, const StepArgs &a0, const double *src, void *dst, void *frozen,
                         hipStream_t s)
{
    StepArgs a = a0;
    a.nblocks = blocks_of(a);
    a.fin = nullptr;
    a.fout = dst;
    set_planes(a);
    dim3 grid((unsigned)a.nblocks), block(kBlockRows);
    switch (storage) {
    case kF64: hipLaunchKernelGGL(upload<double>, grid, block, 0, s, a, src, (double *)frozen); break;
    case kF32: hipLaunchKernelGGL(upload<float>, grid, block, 0, s, a, src, (float *)frozen); break;
    default: hipLaunchKernelGGL(upload<_Float16>, grid, block, 0, s, a, src, (_Float16 *)frozen); break;
    }
    return hipGetLastError();
}

hipError_t launch_init_eq(int32_t storage, const StepArgs &a0, void *dst, void *frozen, double *macro,
                          hipStream_t s)
{
    StepArgs a = a0;
    a.nblocks = blocks_of(a);
    a.fin = nullptr;
    a.fout = dst;
    set_planes(a);
    dim3 grid((unsigned)a.nblocks), block(kBlockRows);
    switch (storage) {
    case kF64: hipLaunchKernelGGL(init_eq<double>, grid, block, 0, s, a, (double *)frozen, macro); break;
    case kF32: hipLaunchKernelGGL(init_eq<float>, grid, block, 0, s, a, (float *)frozen, macro); break;
    default: hipLaunchKernelGGL(init_eq<_Float16>, grid, block, 0, s, a, (_Float16 *)frozen, macro); break;
    }
    return hipGetLastError();
}

hipError_t launch_lift_drag(int32_t storage, const StepArgs &a0, const void *fstar, const int64_t *ij,
                            const int32_t *masks, int64_t n, double *per_point, double *partials,
                            int64_t *nblocks_out, hipStream_t s)
{
    StepArgs a = a0;
    a.fin = fstar;
    a.fout = nullptr;
    set_planes(a);
    int64_t nb = (n + kBlockRows - 1) / kBlockRows;
    if (nb < 1) nb = 1;
    *nblocks_out = nb;
    dim3 grid((unsigned)nb), block(kBlockRows);
    switch (storage) {
    case kF64: hipLaunchKernelGGL(lift_drag<double>, grid, block, 0, s, a, ij, masks, n, per_point, partials); break;
    case kF32: hipLaunchKernelGGL(lift_drag<float>, grid, block, 0, s, a, ij, masks, n, per_point, partials); break;
    default: hipLaunchKernelGGL(lift_drag<_Float16>, grid, block, 0, s, a, ij, masks, n, per_point, partials); break;
    }
    return hipGetLastError();
}

hipError_t launch_fill_macro(double *macro, int64_t n, hipStream_t s)
{
    const int64_t nb = (n + kBlockRows - 1) / kBlockRows;
    hipLaunchKernelGGL(fill_macro, dim3((unsigned)nb), dim3(kBlockRows), 0, s, macro, n);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const double *partials, int64_t nblocks, int32_t ncomp, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(kBlockRows), 0, s, partials, nblocks, ncomp, out);
    return hipGetLastError();
}

}  // namespace lbm
