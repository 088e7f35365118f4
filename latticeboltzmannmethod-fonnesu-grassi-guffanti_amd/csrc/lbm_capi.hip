// C ABI of the D2Q9 engine (declared in include/lbm_amd.h).
//
// The context owns all device memory: two population buffers (A-B), the
// bit-packed wet mask, the perimeter edge table (inlet/outlet kind,
// prescribed values, frozen populations), the rho/u fields and reduction
// scratch.  State machine:
//   kNone  -- no populations yet
//   kP     -- buffer `cur` holds pre-collision populations P_t (after
//             init or upload); the next step starts with a collide-only pass
//   kFstar -- buffer `cur` holds F*_{t-1}; the next step gathers from it
// t counts completed reference time steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbm_amd.h"
#include "lbm_context.h"
#include "lbm_internal.h"

using namespace lbm;

namespace lbm {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

StepArgs make_args(const lbm_ctx *c)
{
    StepArgs a{};
    a.g.nr = (int32_t)c->nr;
    a.g.nc = (int32_t)c->nc;
    a.g.ldr = c->ldr;
    a.g.plane = c->plane;
    a.g.wet = c->d_wet;
    a.g.sb = c->d_sb;
    a.g.wl = c->wl;
    a.g.E = c->E;
    a.g.edge_kind = c->d_edge_kind;
    a.frozen = c->d_frozen;
    a.bc_prev = c->d_bc[c->bc_prev];
    a.bc_cur = c->d_bc[c->bc_cur];
    a.inlet_policy = c->desc.inlet_policy;
    a.outlet_policy = c->desc.outlet_policy;
    a.collision = c->desc.collision;
    a.segs = (int32_t)c->segs;
    a.esize = (int32_t)c->esize;
    a.block_mask = c->blocked ? ~63u : 0u;
    a.launch_lds = occupancy_lds(c->occupancy >= 0 ? c->occupancy : default_occupancy(c->storage));
    // The reference recomputes t = 1.0 / tau inside every call
    // (BGKCollision.hpp:140-141) and (1.0/tau_even) etc. inside every
    // expression (TRTCollision.hpp:155-188); IEEE division makes these the
    // same values when formed once here.
    a.bgk_t = 1.0 / c->desc.tau;
    a.bgk_omt = 1.0 - a.bgk_t;
    a.trt_inv_te = 1.0 / c->desc.tau_even;
    a.trt_a = 0.5 * (1.0 / c->desc.tau_even);
    a.trt_b = 0.5 * (1.0 / c->desc.tau_odd);
    if (c->desc.collision == LBM_BGK) {
        a.sh_a = a.sh_b = (float)(0.5 / c->desc.tau);
    } else {
        a.sh_a = (float)a.trt_a;
        a.sh_b = (float)a.trt_b;
    }
    a.scale = c->scale;
    a.inv_scale = 1.f / c->scale;
    a.f16_sr = (c->desc.options & LBM_OPT_F16_ROUND_TO_NEAREST) ? 0 : 1;
    a.flag = c->d_flag;
    a.runs = c->d_runs;
    a.nruns = c->nruns;
    a.perim = c->d_perim;
    a.nperim = c->nperim;
    a.trun = c->d_trun;
    a.tlane = c->d_tlane;
    a.psbb = c->psbb ? 1 : 0;
    a.ps_t = c->psbb ? 1.0 / c->ps_tau : 0.0;
    a.ps_mask = c->d_ps_mask;
    a.ps_b = c->d_ps_b;
    a.ps_val = c->d_ps_val;
    return a;
}

}  // namespace lbm

namespace {

// Steps per captured graph: even, so a replay leaves the buffer holding
// the state unchanged.
constexpr int64_t kGraphSteps = 32;

void drop_graphs(lbm_ctx *c)
{
    for (auto &row : c->graph)
        for (auto &g : row) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// Captures kGraphSteps gathering steps from the context's current buffer
// and prescribed values (both steps of a run after its first).
int capture_steps(lbm_ctx *c)
{
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    for (int64_t k = 0; k < kGraphSteps && err == hipSuccess; ++k) {
        StepArgs a = make_args(c);
        const int in = c->cur ^ (int)(k & 1);
        a.fin = c->d_F[in];
        a.fout = c->d_F[1 - in];
        a.macro = nullptr;
        err = launch_step(c->storage, true, a, c->stream, c->segments);
    }
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (err != hipSuccess) return fail(LBM_ERR_HIP, "step graph capture: %s", hipGetErrorString(err));
    if (e2 != hipSuccess) return fail(LBM_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
    err = hipGraphInstantiate(&c->graph[c->cur][c->bc_cur], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess) return fail(LBM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(err));
    return LBM_OK;
}

void free_psbb(lbm_ctx *c)
{
    (void)hipFree(c->d_ps_mask); c->d_ps_mask = nullptr;
    (void)hipFree(c->d_ps_b); c->d_ps_b = nullptr;
    (void)hipFree(c->d_ps_val); c->d_ps_val = nullptr;
    c->psbb = false;
    c->ps_tau = 0.0;
}

void free_geometry(lbm_ctx *c)
{
    drop_graphs(c);
    (void)hipFree(c->d_wet); c->d_wet = nullptr;
    (void)hipFree(c->d_sb); c->d_sb = nullptr;
    (void)hipFree(c->d_runs); c->d_runs = nullptr; c->nruns = 0;
    (void)hipFree(c->d_trun); c->d_trun = nullptr;
    (void)hipFree(c->d_tlane); c->d_tlane = nullptr;
    (void)hipFree(c->d_perim); c->d_perim = nullptr; c->nperim = 0;
    (void)hipFree(c->d_edge_kind); c->d_edge_kind = nullptr;
    (void)hipFree(c->d_frozen); c->d_frozen = nullptr;
    for (int k = 0; k < 2; ++k) {
        (void)hipFree(c->d_F[k]); c->d_F[k] = nullptr;
        (void)hipFree(c->d_bc[k]); c->d_bc[k] = nullptr;
        if (c->h_bc_pin[k]) (void)hipHostFree(c->h_bc_pin[k]);
        c->h_bc_pin[k] = nullptr;
    }
    for (int k = 0; k < 2; ++k) {
        (void)hipFree(c->d_macro[k]);
        c->d_macro[k] = nullptr;
    }
    c->macro_slot = 0;
    (void)hipFree(c->d_partials); c->d_partials = nullptr; c->partials_cap = 0;
    (void)hipFree(c->d_pts); c->d_pts = nullptr;
    (void)hipFree(c->d_masks); c->d_masks = nullptr;
    (void)hipFree(c->d_pp); c->d_pp = nullptr; c->pts_cap = 0;
    c->h_pts.clear();
    c->h_masks.clear();
    (void)hipFree(c->d_staging); c->d_staging = nullptr;
    free_psbb(c);
    c->has_geom = false;
    c->state = kNone;
}

int ensure_partials(lbm_ctx *c, int64_t n)
{
    if (n <= c->partials_cap) return LBM_OK;
    (void)hipFree(c->d_partials);
    c->d_partials = nullptr;
    HIP_TRY(hipMalloc(&c->d_partials, (size_t)n * sizeof(double)));
    c->partials_cap = n;
    return LBM_OK;
}

int ensure_staging(lbm_ctx *c)
{
    if (c->d_staging) return LBM_OK;
    HIP_TRY(hipMalloc(&c->d_staging, (size_t)(9 * c->nr * c->nc) * sizeof(double)));
    return LBM_OK;
}

// Pushes the host values h_bc into the device copy the next step reads as
// "current".  The copy that the last step read as "current" becomes
// "previous" and stays untouched.
int push_bc(lbm_ctx *c)
{
    if (c->bc_cur == c->bc_prev) c->bc_cur = 1 - c->bc_prev;
    int k = c->bc_cur;
    // The pinned staging copy k may still be in flight from an earlier push.
    HIP_TRY(hipEventSynchronize(c->bc_ev[k]));
    std::memcpy(c->h_bc_pin[k], c->h_bc.data(), c->h_bc.size() * sizeof(double));
    HIP_TRY(hipMemcpyAsync(c->d_bc[k], c->h_bc_pin[k], c->h_bc.size() * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipEventRecord(c->bc_ev[k], c->stream));
    return LBM_OK;
}

// Storage scale of the FP16 path: a power of two that puts the largest
// expected |f - w| (from the prescribed velocities and densities) near
// 2^11.  The estimate is at least 1/16, so the scale is at most 2^15 and
// values up to |f - w| = 1.99 stay below the FP16 limit (65504) whatever
// the boundary values: interior flows faster than the boundaries, and
// inlets that ramp up from rest, fit.  The scale does not change the
// relative resolution of normal FP16 numbers (2^-11); at 2^15 only
// |f - w| below 2e-9 falls into the subnormal range.
float choose_scale(const lbm_ctx *c)
{
    if (c->storage != kF16) return 1.f;
    if (c->desc.f16_scale > 0) return (float)c->desc.f16_scale;
    double umax = 0.0, dmax = 0.0;
    for (int w = 0; w < 2; ++w)
        for (int64_t e : c->open_e[w]) {
            umax = std::max(umax, std::hypot(c->h_bc[e], c->h_bc[c->E + e]));
            dmax = std::max(dmax, std::fabs(c->h_bc[2 * c->E + e] - 1.0));
        }
    double gmax = (4.0 / 9.0) * (dmax + 3.0 * umax + 6.0 * umax * umax);
    gmax = std::max(gmax, 1.0 / 16.0);
    double s = std::exp2(std::floor(std::log2(2048.0 / gmax)));
    s = std::min(std::max(s, 1.0), 32768.0);
    return (float)s;
}

bool on_perimeter(int64_t i, int64_t j, int64_t nr, int64_t nc)
{
    return i == 0 || j == 0 || i == nr - 1 || j == nc - 1;
}

int64_t edge_slot_host(int64_t i, int64_t j, int64_t nr, int64_t nc)
{
    if (i == 0) return j;
    if (i == nr - 1) return nc + j;
    if (j == 0) return 2 * nc + i;
    return 2 * nc + nr + i;
}

const int kOffI[9] = {0, 0, -1, 0, 1, -1, -1, 1, 1};
const int kOffJ[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};

// Scratch of the allocation-free reduction (lbm_reduce_device): the
// per-block partial sums of the readout kernel, 5 doubles per block of
// 256 rows x 1 column, for lattices of at most kScratchBlocks such blocks
// (ceil(NR / 256) * NC; e.g. 4096 x 8192).  Every device the library runs
// on holds its own copy (5 MiB) until the process ends.  Host code never
// passes the array's name to a kernel: it uses the device address that
// hipGetSymbolAddress returned for the current device (scratch_ptr).
constexpr int64_t kScratchBlocks = int64_t(1) << 17;
static __device__ __align__(256) double scratch[5 * kScratchBlocks];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// Looks the scratch array up on `device` (which must be current) once per
// device; called by lbm_create.
int init_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(LBM_ERR_UNSUPPORTED, "device ordinal %d >= %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return LBM_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail(LBM_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[device] = (double *)ptr;
    return LBM_OK;
}

double *scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}

bool wet_at(const lbm_ctx *c, int64_t i, int64_t j)
{
    return (c->h_wet[(size_t)((j + 1) * c->wl + 1 + i / 64)] >> (i % 64)) & 1;
}

// The copy of the rho/u fields the next writer (an LBM_STEP_MACRO step or
// init) fills: the one the last writer did not fill, once its last
// asynchronous download has read it.
int claim_macro(lbm_ctx *c, double **out)
{
    const int k = 1 - c->macro_slot;
    HIP_TRY(hipStreamWaitEvent(c->stream, c->macro_free[k], 0));
    c->macro_slot = k;
    *out = c->d_macro[k];
    return LBM_OK;
}

// Records the completion event of a new ticket on stream s.  The event
// slot of the ticket kTickets back is reused once that ticket completed.
int new_ticket(lbm_ctx *c, hipStream_t s, int64_t *ticket)
{
    const int64_t t = c->tick_seq + 1;
    hipEvent_t &ev = c->tick_ev[t % lbm_ctx::kTickets];
    if (!ev) HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    else HIP_TRY(hipEventSynchronize(ev));
    HIP_TRY(hipEventRecord(ev, s));
    c->tick_seq = t;
    if (ticket) *ticket = t;
    return LBM_OK;
}

// FP16 storage with a scale chosen from the prescribed values: after init
// and before the first step the state is exactly zero in every scale, so
// the scale follows values set in between.
void rescale_if_fresh(lbm_ctx *c)
{
    if (c->storage != kF16 || c->desc.f16_scale > 0 || !c->fresh_init) return;
    const float s = choose_scale(c);
    if (s != c->scale) {
        c->scale = s;
        drop_graphs(c);
    }
}

// Validates a lift/drag point set against the lattice (each direction bit
// must name a wet node inside the lattice) and makes it the one held on
// the device.  *np receives the point count.
int stage_points(lbm_ctx *c, const int64_t *ij, const int32_t *masks, int64_t n, int64_t *np)
{
    if ((ij == nullptr) != (masks == nullptr)) return fail(LBM_ERR_INVALID, "ij and masks go together");
    const int64_t *hij = ij ? ij : c->obst_ij.data();
    const int32_t *hm = ij ? masks : c->obst_mask.data();
    if (!ij) n = (int64_t)c->obst_mask.size();
    if (n < 0) return fail(LBM_ERR_INVALID, "negative point count");
    for (int64_t q = 0; q < n; ++q) {
        int64_t i = hij[2 * q], j = hij[2 * q + 1];
        int32_t m = hm[q];
        if (i < 0 || i >= c->nr || j < 0 || j >= c->nc) return fail(LBM_ERR_INVALID, "point %lld outside the lattice", (long long)q);
        if (m & ~0x1FE) return fail(LBM_ERR_INVALID, "point %lld: mask bits outside directions 1..8", (long long)q);
        for (int d = 1; d < 9; ++d)
            if ((m >> d) & 1) {
                int64_t ii = i + kOffI[d], jj = j + kOffJ[d];
                if (ii < 0 || ii >= c->nr || jj < 0 || jj >= c->nc)
                    return fail(LBM_ERR_INVALID, "point %lld: direction %d leaves the lattice", (long long)q, d);
                if (!wet_at(c, ii, jj))
                    return fail(LBM_ERR_INVALID, "point %lld: direction %d names a node that is not fluid, inlet or outlet",
                                (long long)q, d);
            }
    }
    *np = n;
    if (n == 0) return LBM_OK;
    const bool same = (int64_t)c->h_masks.size() == n && std::equal(hm, hm + n, c->h_masks.begin()) &&
                      std::equal(hij, hij + 2 * n, c->h_pts.begin());
    if (same) return LBM_OK;
    // kernels of earlier (asynchronous) calls may still read the old set
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (n > c->pts_cap) {
        (void)hipFree(c->d_pts); (void)hipFree(c->d_masks); (void)hipFree(c->d_pp);
        c->d_pts = nullptr; c->d_masks = nullptr; c->d_pp = nullptr; c->pts_cap = 0;
        HIP_TRY(hipMalloc(&c->d_pts, (size_t)(2 * n) * sizeof(int64_t)));
        HIP_TRY(hipMalloc(&c->d_masks, (size_t)n * sizeof(int32_t)));
        HIP_TRY(hipMalloc(&c->d_pp, (size_t)(2 * n) * sizeof(double)));
        c->pts_cap = n;
    }
    c->h_pts.clear();
    c->h_masks.clear();
    HIP_TRY(hipMemcpy(c->d_pts, hij, (size_t)(2 * n) * sizeof(int64_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_masks, hm, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
    c->h_pts.assign(hij, hij + 2 * n);
    c->h_masks.assign(hm, hm + n);
    return LBM_OK;
}

// Tables of the tabled wet runs (kRunTableLaunch), built from the run table
// on first use: per run the entry and the compact indices of its first row
// in the neighbouring columns, per lane one record (see StepArgs::tlane).
int ensure_run_tables(lbm_ctx *c)
{
    if (c->d_trun || c->nruns == 0) return LBM_OK;
    const int64_t nr = c->nr;
    auto word = [&](int64_t i, int64_t j) { return (size_t)((j + 1) * c->wl + 1 + i / 64); };
    auto wet = [&](int64_t i, int64_t j) {
        return i >= 0 && i < nr && j >= 0 && j < c->nc && ((c->h_wet[word(i, j)] >> (i % 64)) & 1);
    };
    // compact index of the first wet row at or below row i in column j
    auto rank = [&](int64_t i, int64_t j) {
        const size_t w = word(i, j);
        const uint64_t below = (i % 64) ? (c->h_wet[w] & ((1ull << (i % 64)) - 1ull)) : 0ull;
        return (int32_t)(c->h_sb[w] + __builtin_popcountll(below));
    };
    static const int di[9] = {0, 0, -1, 0, 1, -1, -1, 1, 1};
    static const int dj[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    std::vector<int32_t> trun((size_t)(8 * c->nruns), 0);
    std::vector<uint32_t> tlane((size_t)(64 * c->nruns), 0u);
    for (int64_t t = 0; t < c->nruns; ++t) {
        const int32_t *r = &c->h_runs[(size_t)(4 * t)];
        const int64_t j = r[0], i0 = r[1];
        const int64_t n = r[3];
        int32_t *e = &trun[(size_t)(8 * t)];
        e[0] = r[0]; e[1] = r[1]; e[2] = r[2]; e[3] = r[3];
        e[4] = rank(i0, j - 1);
        e[5] = rank(i0, j + 1);
        int64_t i = i0, lc = 0, rc = 0;
        for (int64_t l = 0; l < n; ++l, ++i) {
            while (!wet(i, j)) {  // the lane's row: the next wet row of column j
                lc += wet(i, j - 1);
                rc += wet(i, j + 1);
                ++i;
            }
            uint32_t bits = 0;
            for (int d = 1; d < 9; ++d)  // source of direction d: (i - di[d], j - dj[d])
                if (wet(i - di[d], j - dj[d])) bits |= 1u << (d - 1);
            tlane[(size_t)(64 * t + l)] = (uint32_t)(i - i0) | (uint32_t)lc << 7 | (uint32_t)rc << 14 | bits << 21;
            lc += wet(i, j - 1);
            rc += wet(i, j + 1);
        }
    }
    // both tables or neither: the context takes them only once both are on
    // the device
    int32_t *drun = nullptr;
    uint32_t *dlane = nullptr;
    hipError_t e = hipMalloc(&drun, trun.size() * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&dlane, tlane.size() * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpy(drun, trun.data(), trun.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dlane, tlane.data(), tlane.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(drun);
        (void)hipFree(dlane);
        return fail(e == hipErrorOutOfMemory ? LBM_ERR_NOMEM : LBM_ERR_HIP, "run tables: %s", hipGetErrorString(e));
    }
    c->d_trun = drun;
    c->d_tlane = dlane;
    return LBM_OK;
}

// Queues the lift/drag reduction of the staged points: per-point values in
// d_pp (when wanted), totals in d_out[0..1].
int queue_lift_drag(lbm_ctx *c, int64_t n, bool per_point)
{
    int64_t nb = (n + kBlockRows - 1) / kBlockRows;
    int r = ensure_partials(c, 2 * nb);
    if (r) return r;
    StepArgs a = make_args(c);
    HIP_TRY(launch_lift_drag(c->storage, a, c->d_F[c->cur], c->d_pts, c->d_masks, n, per_point ? c->d_pp : nullptr,
                             c->d_partials, &nb, c->stream));
    HIP_TRY(launch_reduce_partials(c->d_partials, nb, 2, c->d_out, c->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

const char *lbm_version(void) { return "lbm_amd 0.1 (gfx950)"; }

const char *lbm_last_error(void) { return g_err.c_str(); }

int lbm_create(const lbm_desc *desc, lbm_ctx **out)
{
    if (!desc || !out) return fail(LBM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (desc->n_rows < 3 || desc->n_cols < 3) return fail(LBM_ERR_INVALID, "lattice must be at least 3x3");
    if (desc->n_rows > (1 << 30) || desc->n_cols > (1 << 30)) return fail(LBM_ERR_INVALID, "lattice too large");
    if (desc->collision != LBM_BGK && desc->collision != LBM_TRT) return fail(LBM_ERR_INVALID, "unknown collision");
    if (desc->storage < LBM_STORE_F64 || desc->storage > LBM_STORE_F16_SHIFTED)
        return fail(LBM_ERR_INVALID, "unknown storage");
    for (int32_t p : {desc->inlet_policy, desc->outlet_policy})
        if (p < LBM_ZOUHE_VELOCITY || p > LBM_OPEN_NOOP) return fail(LBM_ERR_INVALID, "unknown open policy");
    if (desc->options & ~(int32_t)(LBM_OPT_F16_ROUND_TO_NEAREST | LBM_OPT_PLANAR_LAYOUT))
        return fail(LBM_ERR_INVALID, "unknown option bits");
    if (desc->collision == LBM_BGK && !(desc->tau > 0)) return fail(LBM_ERR_INVALID, "tau must be > 0");
    if (desc->collision == LBM_TRT && !(desc->tau_even > 0 && desc->tau_odd > 0))
        return fail(LBM_ERR_INVALID, "tau_even and tau_odd must be > 0");
    if (desc->f16_scale != 0.0) {
        // the FP16 collision relies on S being an exact power of two
        int ex = 0;
        const double m = std::isfinite(desc->f16_scale) ? std::frexp(desc->f16_scale, &ex) : 0.0;
        if (!(desc->f16_scale >= 1.0 && desc->f16_scale <= 16777216.0 && m == 0.5))
            return fail(LBM_ERR_INVALID, "f16_scale must be 0 or a power of two in [1, 2^24]");
    }
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (desc->device < 0 || desc->device >= ndev) return fail(LBM_ERR_INVALID, "no HIP device %d", desc->device);
    HIP_TRY(hipSetDevice(desc->device));
    {
        const int r = init_scratch(desc->device);
        if (r) return r;
    }
    lbm_ctx *c = new lbm_ctx();
    c->desc = *desc;
    c->storage = desc->storage;
    c->esize = desc->storage == kF64 ? 8 : desc->storage == kF32 ? 4 : 2;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(LBM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->copy_go, hipEventDisableTiming);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipEventCreateWithFlags(&c->bc_ev[k], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->macro_free[k], hipEventDisableTiming);
        // both field copies start free
        if (e == hipSuccess) e = hipEventRecord(c->macro_free[k], c->copy_stream);
    }
    if (e != hipSuccess) {
        lbm_destroy(c);
        return fail(LBM_ERR_HIP, "stream/event creation: %s", hipGetErrorString(e));
    }
    e = hipMalloc(&c->d_flag, 64);
    if (e == hipSuccess) e = hipMalloc(&c->d_out, 8 * sizeof(double));
    if (e == hipSuccess) e = hipMemsetAsync(c->d_flag, 0, 64, c->stream);
    if (e != hipSuccess) {
        lbm_destroy(c);
        return fail(LBM_ERR_HIP, "allocation: %s", hipGetErrorString(e));
    }
    *out = c;
    return LBM_OK;
}

int lbm_destroy(lbm_ctx *c)
{
    if (!c) return LBM_OK;
    if (c->batch) return fail(LBM_ERR_STATE, "context belongs to an ensemble (destroy the lbm_batch first)");
    (void)hipSetDevice(c->desc.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    free_geometry(c);
    (void)hipFree(c->d_flag);
    (void)hipFree(c->d_out);
    for (int k = 0; k < 2; ++k) {
        if (c->bc_ev[k]) (void)hipEventDestroy(c->bc_ev[k]);
        if (c->macro_free[k]) (void)hipEventDestroy(c->macro_free[k]);
    }
    for (hipEvent_t ev : c->tick_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->copy_go) (void)hipEventDestroy(c->copy_go);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LBM_OK;
}

int lbm_set_relaxation(lbm_ctx *c, double tau, double tau_even, double tau_odd)
{
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    c->version++;
    if (c->desc.collision == LBM_BGK && !(tau > 0)) return fail(LBM_ERR_INVALID, "tau must be > 0");
    if (c->desc.collision == LBM_TRT && !(tau_even > 0 && tau_odd > 0))
        return fail(LBM_ERR_INVALID, "tau_even and tau_odd must be > 0");
    c->desc.tau = tau;
    c->desc.tau_even = tau_even;
    c->desc.tau_odd = tau_odd;
    drop_graphs(c);
    return LBM_OK;
}

int lbm_set_partial_bounce_back(lbm_ctx *c, double tau, const double *b)
{
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    if (!c->has_geom) return fail(LBM_ERR_STATE, "no geometry");
    if (c->batch) return fail(LBM_ERR_STATE, "context belongs to an ensemble");
    if (c->state == kFstar)
        return fail(LBM_ERR_STATE, "partially saturated bounce-back is set before the run (after the geometry, "
                    "before or right after lbm_init_equilibrium / lbm_upload_populations)");
    if (c->storage != kF64 && tau > 0)
        return fail(LBM_ERR_UNSUPPORTED, "partially saturated bounce-back needs FP64 storage");
    const int64_t nob = (int64_t)c->obst_mask.size();
    if (tau > 0 && b)
        for (int64_t q = 0; q < nob; ++q)
            if (!std::isfinite(b[q])) return fail(LBM_ERR_INVALID, "weight of obstacle %lld is not finite", (long long)q);
    if (tau > 0 && !std::isfinite(tau)) return fail(LBM_ERR_INVALID, "tau must be finite");
    HIP_TRY(hipSetDevice(c->desc.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->version++;
    drop_graphs(c);
    free_psbb(c);
    if (!(tau > 0)) return LBM_OK;
    // per wet node n next to a wet obstacle o = n - c_d: bit d of the mask
    // and o's weight in plane d
    const size_t plane = (size_t)c->plane;
    std::vector<uint16_t> mask(plane, 0);
    std::vector<double> wb(9 * plane, 0.0);
    auto cidx = [&](int64_t i, int64_t j) {
        const size_t w = (size_t)((j + 1) * c->wl + 1 + i / 64);
        const uint64_t below = (i % 64) ? (c->h_wet[w] & ((1ull << (i % 64)) - 1ull)) : 0ull;
        return (int64_t)c->h_sb[w] + __builtin_popcountll(below);
    };
    for (int64_t q = 0; q < nob; ++q) {
        const int64_t i = c->obst_ij[2 * q], j = c->obst_ij[2 * q + 1];
        for (int d = 1; d < 9; ++d)
            if ((c->obst_mask[q] >> d) & 1) {
                // the obstacle's bit d names its wet neighbour at the push target o + c_d
                const int64_t n = cidx(i + kOffI[d], j + kOffJ[d]);
                mask[(size_t)n] |= (uint16_t)(1u << d);
                wb[(size_t)d * plane + (size_t)n] = b ? b[q] : 1.0;
            }
    }
    HIP_TRY(hipMalloc(&c->d_ps_mask, plane * sizeof(uint16_t)));
    HIP_TRY(hipMemcpy(c->d_ps_mask, mask.data(), plane * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_ps_b, 9 * plane * sizeof(double)));
    HIP_TRY(hipMemcpy(c->d_ps_b, wb.data(), 9 * plane * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_ps_val, 9 * plane * sizeof(double)));
    HIP_TRY(hipMemset(c->d_ps_val, 0, 9 * plane * sizeof(double)));
    c->psbb = true;
    c->ps_tau = tau;
    return LBM_OK;
}

int lbm_set_geometry(lbm_ctx *c, const uint8_t *cls)
{
    if (!c || !cls) return fail(LBM_ERR_INVALID, "null argument");
    if (c->batch) return fail(LBM_ERR_STATE, "context belongs to an ensemble (destroy the lbm_batch first)");
    c->version++;
    HIP_TRY(hipSetDevice(c->desc.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipStreamSynchronize(c->copy_stream));
    const int64_t nr = c->desc.n_rows, nc = c->desc.n_cols;
    // validation
    for (int64_t q = 0; q < nr * nc; ++q)
        if (cls[q] > LBM_OBSTACLE) return fail(LBM_ERR_INVALID, "node class %d at %lld out of range", cls[q], (long long)q);
    auto wet_cls = [](uint8_t k) { return k == LBM_FLUID || k == LBM_INLET || k == LBM_OUTLET; };
    for (int64_t j = 0; j < nc; ++j)
        for (int64_t i = 0; i < nr; ++i) {
            uint8_t k = cls[i + nr * j];
            if ((k == LBM_INLET || k == LBM_OUTLET) && !on_perimeter(i, j, nr, nc))
                return fail(LBM_ERR_INVALID, "inlet/outlet node (%lld, %lld) is not on the domain perimeter",
                            (long long)i, (long long)j);
            if (k == LBM_SOLID) {
                for (int d = 1; d < 9; ++d) {
                    int64_t ii = i + kOffI[d], jj = j + kOffJ[d];
                    if (ii >= 0 && ii < nr && jj >= 0 && jj < nc && wet_cls(cls[ii + nr * jj]))
                        return fail(LBM_ERR_UNSUPPORTED,
                                    "SOLID node (%lld, %lld) borders a wet node: unsupported", (long long)i,
                                    (long long)j);
                }
            }
        }
    // The step kernels index waves, blocks and mask words in 32 bits: at
    // most 2^30 64-row segments and 2^31 mask words (with the guard words
    // and columns); the wet mask alone would take 8 GiB.
    if ((nr + 63) / 64 * nc > (int64_t(1) << 30) || (nc + 2) * ((nr + 63) / 64 + 4) > (int64_t(1) << 31))
        return fail(LBM_ERR_UNSUPPORTED, "lattice of %lld x %lld nodes exceeds 2^30 64-row segments",
                    (long long)nr, (long long)nc);
    free_geometry(c);
    c->nr = nr;
    c->nc = nc;
    const int64_t nrw = (nr + 63) / 64;
    c->ldr = nrw * 64;
    // one zero guard word before each column's segments and three after
    // them (the wet-run waves read up to two words past a column's last
    // segment)
    c->wl = nrw + 4;
    c->E = 2 * nc + 2 * nr;
    c->segs = (c->ldr + kBlockRows - 1) / kBlockRows;
    // host tables
    std::vector<uint64_t> wet((size_t)((nc + 2) * c->wl), 0);
    std::vector<uint8_t> edge((size_t)c->E, kEdgePlain);
    for (int k = 0; k < 7; ++k) c->counts[k] = 0;
    c->n_wet = 0;
    for (int w = 0; w < 2; ++w) { c->open_ij[w].clear(); c->open_e[w].clear(); }
    c->obst_ij.clear();
    c->obst_mask.clear();
    for (int64_t j = 0; j < nc; ++j)
        for (int64_t i = 0; i < nr; ++i) {
            uint8_t k = cls[i + nr * j];
            c->counts[k]++;
            if (wet_cls(k)) {
                wet[(size_t)((j + 1) * c->wl + 1 + i / 64)] |= 1ull << (i % 64);
                c->n_wet++;
            }
            if (k == LBM_INLET || k == LBM_OUTLET) edge[(size_t)edge_slot_host(i, j, nr, nc)] = k == LBM_INLET ? kEdgeInlet : kEdgeOutlet;
        }
    // Compact storage: wet nodes only, column-major.  sb holds the compact
    // index of the first wet node of every (column, 64-row word).  A column
    // whose wet rows form one unbroken stretch of at least kAlignRows rows
    // starts at the first index congruent to the row of its first wet node
    // modulo 64 (at most 63 unused slots), so its 64-row segments fill
    // whole, aligned 64-element groups: wave loads and stores cover full
    // cache lines and stream.  Every other column follows the previous one
    // directly: behind the first obstacle its slots are misaligned anyway,
    // and packing measured faster there (a 512-lattice porous ensemble
    // +18 %, lattices of 128^2 +46 %, a porous 8192^2 +1 %), while aligning
    // the unbroken columns of an open 8192^2 channel is worth 4 %.  Nodes
    // are contiguous inside a column, which is all the neighbour arithmetic
    // (row i +- 1 = compact index +- 1) relies on.
    std::vector<int32_t> sb((size_t)((nc + 2) * c->wl), 0);
    int64_t slots = 0;
    {
        constexpr int64_t kAlignRows = 256;
        int64_t acc = 0;
        for (int64_t j = 0; j < nc; ++j) {
            const size_t w0 = (size_t)((j + 1) * c->wl + 1);
            // first wet row, number of wet rows and number of unbroken
            // stretches of the column
            int64_t r0 = -1, count = 0, stretches = 0;
            bool prev = false;  // row 64 k - 1 wet
            for (int64_t k = 0; k < nrw; ++k) {
                const uint64_t m = wet[w0 + k];
                if (m && r0 < 0) r0 = 64 * k + __builtin_ctzll(m);
                count += __builtin_popcountll(m);
                // stretch starts: wet rows whose row above is dry
                const uint64_t above = (m << 1) | (prev ? 1ull : 0ull);
                stretches += __builtin_popcountll(m & ~above);
                prev = (m >> 63) & 1;
            }
            if (stretches == 1 && count >= kAlignRows) acc += ((r0 - acc) % 64 + 64) % 64;
            for (int64_t k = 0; k < nrw; ++k) {
                sb[w0 + k] = (int32_t)acc;
                acc += __builtin_popcountll(wet[w0 + k]);
            }
        }
        slots = acc;
    }
    // At least one unused slot at the end of every plane: the step kernels'
    // dry lanes read their own slot, which past the last wet node of the
    // lattice is index `slots`.
    c->plane = ((slots + 1 + 63) / 64) * 64;
    // The kernels address a population as a direction base pointer plus a
    // 32-bit byte offset: spanning up to three planes, or the whole buffer
    // in the blocked layout, which lattices take when it fits.
    const uint64_t fbytes_all = 9ull * (uint64_t)c->plane * c->esize;
    if (fbytes_all / 3 >= (1ull << 32)) {
        c->nr = c->nc = 0;
        return fail(LBM_ERR_UNSUPPORTED, "lattice with %lld wet nodes exceeds the 32-bit population addressing "
                    "(3 * compact slots * element size must stay below 4 GiB)", (long long)c->n_wet);
    }
    c->blocked = !(c->desc.options & LBM_OPT_PLANAR_LAYOUT) && fbytes_all < (1ull << 32);
    // list order of the reference: row-major scan
    for (int64_t i = 0; i < nr; ++i)
        for (int64_t j = 0; j < nc; ++j) {
            uint8_t k = cls[i + nr * j];
            if (k == LBM_INLET || k == LBM_OUTLET) {
                int w = k == LBM_INLET ? 0 : 1;
                c->open_ij[w].push_back(i);
                c->open_ij[w].push_back(j);
                c->open_e[w].push_back(edge_slot_host(i, j, nr, nc));
            } else if (k == LBM_OBSTACLE) {
                int32_t m = 0;
                for (int d = 1; d < 9; ++d) {
                    int64_t ii = i + kOffI[d], jj = j + kOffJ[d];
                    if (ii >= 0 && ii < nr && jj >= 0 && jj < nc && wet_cls(cls[ii + nr * jj])) m |= 1 << d;
                }
                if (m) {
                    c->obst_ij.push_back(i);
                    c->obst_ij.push_back(j);
                    c->obst_mask.push_back(m);
                }
            }
        }
    c->counts[6] = (int64_t)c->obst_mask.size();
    c->h_bc.assign((size_t)(3 * c->E), 0.0);
    for (int64_t e = 0; e < c->E; ++e) c->h_bc[2 * c->E + e] = 1.0;
    // device allocations
    size_t fbytes = (size_t)(9 * c->plane) * c->esize;
    HIP_TRY(hipMalloc(&c->d_wet, wet.size() * sizeof(uint64_t)));
    HIP_TRY(hipMemcpy(c->d_wet, wet.data(), wet.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    c->h_wet = std::move(wet);
    HIP_TRY(hipMalloc(&c->d_sb, sb.size() * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(c->d_sb, sb.data(), sb.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    c->h_sb = sb;
    // Wet runs of the run kernel: in every interior column, the wet rows of
    // the segments the perimeter kernel does not take (the first segment
    // when row 0 is wet, the one holding row NR-1 when that row is wet),
    // cut into runs of at most 64 rows that span at most 128 rows and never
    // cross a multiple of 64 in the compact index (a block of the blocked
    // layout).  A wave's own-column loads and all its stores then fall in
    // one block per direction instead of two partial ones: on the 8192^2
    // porous medium this measured FP32 +7 %, FP64 +5 %, FP16 (tabled) +3 %
    // at each format's best launch, the same bits.  Entry: column, first
    // row, its compact index, number of rows.
    {
        std::vector<int32_t> runs;
        auto push = [&](int64_t j, int64_t i, int64_t s0, int64_t n) {
            const int32_t e[4] = {(int32_t)j, (int32_t)i, (int32_t)s0, (int32_t)n};
            runs.insert(runs.end(), e, e + 4);
        };
        const int64_t last = nr - 1;
        for (int64_t j = 1; j + 1 < nc; ++j) {
            const size_t w0 = (size_t)((j + 1) * c->wl + 1);
            const std::vector<uint64_t> &hw = c->h_wet;
            const int64_t klo = (hw[w0] & 1) ? 1 : 0;
            const int64_t khi = ((hw[w0 + (size_t)(last / 64)] >> (last % 64)) & 1) ? last / 64 : nrw;
            int64_t i_first = 0, s_first = 0, n = 0;
            for (int64_t k = klo; k < khi; ++k) {
                uint64_t m = hw[w0 + (size_t)k];
                int64_t slot = sb[w0 + (size_t)k];
                while (m) {
                    const int64_t i = 64 * k + __builtin_ctzll(m);
                    m &= m - 1;
                    if (n == 64 || (n > 0 && (i >= i_first + 128 || (slot & 63) == 0))) {
                        push(j, i_first, s_first, n);
                        n = 0;
                    }
                    if (n == 0) {
                        i_first = i;
                        s_first = slot;
                    }
                    ++n;
                    ++slot;
                }
            }
            if (n > 0) push(j, i_first, s_first, n);
        }
        c->nruns = (int64_t)(runs.size() / 4);
        HIP_TRY(hipMalloc(&c->d_runs, std::max<size_t>(runs.size(), 4) * sizeof(int32_t)));
        HIP_TRY(hipMemcpy(c->d_runs, runs.data(), runs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        c->h_runs = std::move(runs);
    }
    // Perimeter items: the 64-row segments holding a wet perimeter node,
    // which the step kernels skip (every wet segment of columns 0 and NC-1;
    // of the other columns the first segment when row 0 is wet and the one
    // holding row NR-1 when that row is wet), as (column, segment).
    {
        std::vector<int32_t> items;
        const std::vector<uint64_t> &hw = c->h_wet;
        const int64_t last = nr - 1, klast = last / 64;
        auto add = [&](int64_t j, int64_t k) {
            items.push_back((int32_t)j);
            items.push_back((int32_t)k);
        };
        for (int64_t j = 0; j < nc; ++j) {
            const size_t w0 = (size_t)((j + 1) * c->wl + 1);
            if (j == 0 || j == nc - 1) {
                for (int64_t k = 0; k < nrw; ++k)
                    if (hw[w0 + (size_t)k]) add(j, k);
                continue;
            }
            const bool top = hw[w0] & 1;
            const bool bottom = (hw[w0 + (size_t)klast] >> (last % 64)) & 1;
            if (top) add(j, 0);
            if (bottom && !(top && klast == 0)) add(j, klast);
        }
        c->nperim = (int64_t)(items.size() / 2);
        HIP_TRY(hipMalloc(&c->d_perim, std::max<size_t>(items.size(), 2) * sizeof(int32_t)));
        HIP_TRY(hipMemcpy(c->d_perim, items.data(), items.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMalloc(&c->d_edge_kind, (size_t)c->E));
    HIP_TRY(hipMemcpy(c->d_edge_kind, edge.data(), (size_t)c->E, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_frozen, (size_t)(9 * c->E) * c->esize));
    HIP_TRY(hipMemset(c->d_frozen, 0, (size_t)(9 * c->E) * c->esize));
    for (int k = 0; k < 2; ++k) {
        HIP_TRY(hipMalloc(&c->d_F[k], fbytes));
        HIP_TRY(hipMemset(c->d_F[k], 0, fbytes));
        HIP_TRY(hipMalloc(&c->d_bc[k], (size_t)(3 * c->E) * sizeof(double)));
        HIP_TRY(hipHostMalloc((void **)&c->h_bc_pin[k], (size_t)(3 * c->E) * sizeof(double), hipHostMallocDefault));
        HIP_TRY(hipMemcpy(c->d_bc[k], c->h_bc.data(), (size_t)(3 * c->E) * sizeof(double), hipMemcpyHostToDevice));
    }
    c->bc_cur = c->bc_prev = 0;
    // both copies of the fields start at rho = 1, u = 0 (Lattice.hpp:295-296):
    // the steps write the wet nodes only
    for (int k = 0; k < 2; ++k) {
        HIP_TRY(hipMalloc(&c->d_macro[k], (size_t)(3 * nr * nc) * sizeof(double)));
        HIP_TRY(launch_fill_macro(c->d_macro[k], nr * nc, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->segments == kRunTableLaunch) {  // the chosen shape reads tables of the new geometry
        const int r = ensure_run_tables(c);
        if (r) return r;
    }
    c->macro_slot = 0;
    c->has_geom = true;
    c->state = kNone;
    c->t = 0;
    c->cur = 0;
    c->macro_valid = false;
    c->fresh_init = false;
    return LBM_OK;
}

int lbm_node_counts(lbm_ctx *c, int64_t out[7])
{
    if (!c || !out) return fail(LBM_ERR_INVALID, "null argument");
    if (!c->has_geom) return fail(LBM_ERR_STATE, "no geometry");
    for (int k = 0; k < 7; ++k) out[k] = c->counts[k];
    return LBM_OK;
}

int lbm_open_nodes(lbm_ctx *c, int which, int64_t *ij, int64_t *count)
{
    if (!c || !count || (which != 0 && which != 1)) return fail(LBM_ERR_INVALID, "bad argument");
    if (!c->has_geom) return fail(LBM_ERR_STATE, "no geometry");
    *count = (int64_t)c->open_e[which].size();
    if (ij) std::copy(c->open_ij[which].begin(), c->open_ij[which].end(), ij);
    return LBM_OK;
}

int lbm_set_open_values(lbm_ctx *c, int which, const double *ux, const double *uy, const double *rho)
{
    if (!c || (which != 0 && which != 1)) return fail(LBM_ERR_INVALID, "bad argument");
    if (!c->has_geom) return fail(LBM_ERR_STATE, "no geometry");
    const auto &es = c->open_e[which];
    // all values are checked before any is taken: a failed call changes nothing
    for (size_t q = 0; q < es.size(); ++q) {
        if ((ux && !std::isfinite(ux[q])) || (uy && !std::isfinite(uy[q])))
            return fail(LBM_ERR_INVALID, "prescribed velocity of node %zu is not finite", q);
        if (rho && !(rho[q] > 0 && std::isfinite(rho[q])))
            return fail(LBM_ERR_INVALID, "prescribed density of node %zu must be finite and > 0", q);
    }
    HIP_TRY(hipSetDevice(c->desc.device));
    c->version++;
    for (size_t q = 0; q < es.size(); ++q) {
        if (ux) c->h_bc[es[q]] = ux[q];
        if (uy) c->h_bc[c->E + es[q]] = uy[q];
        if (rho) c->h_bc[2 * c->E + es[q]] = rho[q];
    }
    rescale_if_fresh(c);
    return push_bc(c);
}

int lbm_init_equilibrium(lbm_ctx *c)
{
    if (c) c->version++;
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    if (!c->has_geom) return fail(LBM_ERR_STATE, "no geometry");
    HIP_TRY(hipSetDevice(c->desc.device));
    c->scale = choose_scale(c);
    drop_graphs(c);
    double *macro = nullptr;
    int r = claim_macro(c, &macro);
    if (r) return r;
    StepArgs a = make_args(c);
    HIP_TRY(launch_init_eq(c->storage, a, c->d_F[c->cur], c->d_frozen, macro, c->stream));
    HIP_TRY(hipMemsetAsync(c->d_flag, 0, 64, c->stream));
    c->state = kP;
    c->t = 0;
    c->macro_valid = true;
    c->fresh_init = true;
    return LBM_OK;
}

int lbm_upload_populations(lbm_ctx *c, const double *P)
{
    if (c) c->version++;
    if (!c || !P) return fail(LBM_ERR_INVALID, "null argument");
    if (!c->has_geom) return fail(LBM_ERR_STATE, "no geometry");
    HIP_TRY(hipSetDevice(c->desc.device));
    int r = ensure_staging(c);
    if (r) return r;
    if (c->state == kNone) c->scale = choose_scale(c);
    drop_graphs(c);
    HIP_TRY(hipMemcpyAsync(c->d_staging, P, (size_t)(9 * c->nr * c->nc) * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    StepArgs a = make_args(c);
    HIP_TRY(launch_upload(c->storage, a, c->d_staging, c->d_F[c->cur], c->d_frozen, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->state = kP;
    c->fresh_init = false;
    return LBM_OK;
}

int lbm_download_populations(lbm_ctx *c, double *P)
{
    if (!c || !P) return fail(LBM_ERR_INVALID, "null argument");
    if (c->state == kNone) return fail(LBM_ERR_STATE, "no populations (call lbm_init_equilibrium)");
    HIP_TRY(hipSetDevice(c->desc.device));
    int r = ensure_staging(c);
    if (r) return r;
    StepArgs a = make_args(c);
    a.fin = c->d_F[c->cur];
    HIP_TRY(launch_readout(c->storage, c->state == kFstar, a, c->d_staging, nullptr, c->stream));
    HIP_TRY(hipMemcpyAsync(P, c->d_staging, (size_t)(9 * c->nr * c->nc) * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_download_after_collision(lbm_ctx *c, double *P)
{
    if (!c || !P) return fail(LBM_ERR_INVALID, "null argument");
    if (c->state != kFstar) return fail(LBM_ERR_STATE, "no step completed since the last init/upload");
    HIP_TRY(hipSetDevice(c->desc.device));
    int r = ensure_staging(c);
    if (r) return r;
    StepArgs a = make_args(c);
    a.fin = c->d_F[c->cur];
    HIP_TRY(launch_readout(c->storage, false, a, c->d_staging, nullptr, c->stream));
    HIP_TRY(hipMemcpyAsync(P, c->d_staging, (size_t)(9 * c->nr * c->nc) * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_step(lbm_ctx *c, int64_t n, uint32_t flags)
{
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    if (c->batch) return fail(LBM_ERR_STATE, "context belongs to an ensemble: step it with lbm_batch_step");
    c->version++;
    if (n < 0) return fail(LBM_ERR_INVALID, "negative step count");
    if (c->state == kNone) return fail(LBM_ERR_STATE, "no populations (call lbm_init_equilibrium)");
    HIP_TRY(hipSetDevice(c->desc.device));
    const bool macro = (flags & LBM_STEP_MACRO) != 0;
    for (int64_t s = 0; s < n;) {
        // After the first step of a run (collide-only pass, new prescribed
        // values) every step is the same launch on alternating buffers:
        // replay a captured loop of kGraphSteps of them.
        const bool steady = c->state == kFstar && c->bc_prev == c->bc_cur;
        if (!(flags & LBM_STEP_NO_GRAPH) && steady && n - s - (macro ? 1 : 0) >= kGraphSteps) {
            if (!c->graph[c->cur][c->bc_cur]) {
                int r = capture_steps(c);
                if (r) return r;
            }
            HIP_TRY(hipGraphLaunch(c->graph[c->cur][c->bc_cur], c->stream));
            c->t += kGraphSteps;
            s += kGraphSteps;
            continue;
        }
        StepArgs a = make_args(c);
        a.fin = c->d_F[c->cur];
        a.fout = c->d_F[1 - c->cur];
        a.macro = nullptr;
        if (macro && s == n - 1) {
            int r = claim_macro(c, &a.macro);
            if (r) return r;
        }
        HIP_TRY(launch_step(c->storage, c->state == kFstar, a, c->stream, c->segments));
        c->fresh_init = false;
        c->cur = 1 - c->cur;
        c->state = kFstar;
        c->bc_prev = c->bc_cur;
        c->t++;
        ++s;
    }
    if (n > 0) c->macro_valid = macro;
    if (flags & LBM_STEP_CHECK) {
        int32_t flag = 0;
        HIP_TRY(hipMemcpyAsync(&flag, c->d_flag, sizeof flag, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        // a population beyond the FP16 range is the cause of the NaNs that
        // usually follow it, so it is reported first
        if (flag & 2)
            return fail(LBM_ERR_RANGE, "FP16 storage: a scaled population reached 65504 at or before step %lld "
                        "(storage scale %g is too large for this flow)", (long long)c->t, (double)c->scale);
        if (flag & 1) return fail(LBM_ERR_NAN, "NaN density at or before step %lld", (long long)c->t);
    }
    return LBM_OK;
}

int lbm_set_launch(lbm_ctx *c, int32_t segments)
{
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    if (segments != 0 && segments != 1 && segments != 2 && segments != 4 && segments != 8 && segments != kRunLaunch &&
        segments != kRunTableLaunch)
        return fail(LBM_ERR_INVALID, "segments must be 0, 1, 2, 4, 8, %d (wet runs) or %d (tabled wet runs)", kRunLaunch,
                    kRunTableLaunch);
    if (segments == kRunTableLaunch) {
        if (!c->has_geom) return fail(LBM_ERR_STATE, "no geometry");
        HIP_TRY(hipSetDevice(c->desc.device));
        int r = ensure_run_tables(c);
        if (r) return r;
    }
    c->segments = segments;
    drop_graphs(c);
    return LBM_OK;
}

int lbm_get_layout(lbm_ctx *c, int32_t *layout)
{
    if (!c || !layout) return fail(LBM_ERR_INVALID, "null argument");
    if (!c->has_geom) return fail(LBM_ERR_STATE, "no geometry");
    *layout = c->blocked ? LBM_LAYOUT_BLOCKED : LBM_LAYOUT_PLANES;
    return LBM_OK;
}

int lbm_set_occupancy(lbm_ctx *c, int32_t groups_per_cu)
{
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    if (groups_per_cu != 0 && (groups_per_cu < 2 || groups_per_cu > 7))
        return fail(LBM_ERR_INVALID, "occupancy must be 0 (no limit) or 2..7 work-groups per compute unit");
    c->occupancy = groups_per_cu;
    drop_graphs(c);
    return LBM_OK;
}

int lbm_get_occupancy(lbm_ctx *c, int32_t *groups_per_cu)
{
    if (!c || !groups_per_cu) return fail(LBM_ERR_INVALID, "null argument");
    *groups_per_cu = c->occupancy >= 0 ? c->occupancy : default_occupancy(c->storage);
    return LBM_OK;
}

int lbm_get_launch(lbm_ctx *c, int32_t *segments)
{
    if (!c || !segments) return fail(LBM_ERR_INVALID, "null argument");
    *segments = c->segments > 0 ? c->segments : default_segments(c->storage);
    return LBM_OK;
}

int lbm_autotune(lbm_ctx *c, int32_t steps, int32_t *segments, double *ms_per_step)
{
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    if (steps < 1) return fail(LBM_ERR_INVALID, "steps must be >= 1");
    if (c->batch) return fail(LBM_ERR_STATE, "context belongs to an ensemble: tune it with lbm_batch_autotune");
    if (c->state == kNone) return fail(LBM_ERR_STATE, "no populations (call lbm_init_equilibrium)");
    // Every launch of a partially saturated bounce-back lattice rewrites the
    // carried obstacle values (ps_val) in place, so a timing launch would
    // change the state.
    if (c->psbb)
        return fail(LBM_ERR_STATE, "autotune: a lattice with partially saturated bounce-back runs one general "
                                   "kernel and would change its carried obstacle values; set the occupancy by hand");
    HIP_TRY(hipSetDevice(c->desc.device));
    {
        // the tabled-run candidate's tables, before the arguments that point
        // at them are formed
        const int r = ensure_run_tables(c);
        if (r) return r;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIP_TRY(hipEventCreate(&t0));
    hipError_t err = hipEventCreate(&t1);
    float best = -1.f;
    int32_t best_R = default_segments(c->storage), best_occ = 0;
    // Each launch recomputes the spare buffer from the current one, so the
    // state is the same after tuning.
    StepArgs a = make_args(c);
    a.bc_prev = a.bc_cur;
    a.fin = c->d_F[c->cur];
    a.fout = c->d_F[1 - c->cur];
    a.macro = nullptr;
    for (int32_t occ : {0, 7, 6, 5, 4, 3}) {
        a.launch_lds = occupancy_lds(occ);
        for (int32_t R : {1, 2, 4, 8, kRunLaunch, kRunTableLaunch}) {
            if (err == hipSuccess) err = launch_step(c->storage, true, a, c->stream, R);
            if (err == hipSuccess) err = hipEventRecord(t0, c->stream);
            for (int32_t k = 0; k < steps && err == hipSuccess; ++k)
                err = launch_step(c->storage, true, a, c->stream, R);
            if (err == hipSuccess) err = hipEventRecord(t1, c->stream);
            if (err == hipSuccess) err = hipEventSynchronize(t1);
            float ms = 0.f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, t0, t1);
            if (err == hipSuccess && (best < 0 || ms < best)) {
                best = ms;
                best_R = R;
                best_occ = occ;
            }
        }
    }
    (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (err != hipSuccess) return fail(LBM_ERR_HIP, "autotune: %s", hipGetErrorString(err));
    c->segments = best_R;
    c->occupancy = best_occ;
    drop_graphs(c);
    if (segments) *segments = best_R;
    if (ms_per_step) *ms_per_step = best / steps;
    return LBM_OK;
}

int lbm_get_time(lbm_ctx *c, int64_t *t)
{
    if (!c || !t) return fail(LBM_ERR_INVALID, "null argument");
    *t = c->t;
    return LBM_OK;
}

int lbm_download_macro(lbm_ctx *c, double *rho, double *u)
{
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    if (!c->macro_valid) return fail(LBM_ERR_STATE, "no macro fields (step with LBM_STEP_MACRO)");
    HIP_TRY(hipSetDevice(c->desc.device));
    size_t N = (size_t)(c->nr * c->nc);
    const double *m = c->d_macro[c->macro_slot];
    if (rho) HIP_TRY(hipMemcpyAsync(rho, m, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (u) HIP_TRY(hipMemcpyAsync(u, m + N, 2 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_download_macro_async(lbm_ctx *c, double *rho, double *u, int64_t *ticket)
{
    if (!c || !ticket) return fail(LBM_ERR_INVALID, "null argument");
    if (!c->macro_valid) return fail(LBM_ERR_STATE, "no macro fields (step with LBM_STEP_MACRO)");
    HIP_TRY(hipSetDevice(c->desc.device));
    const size_t N = (size_t)(c->nr * c->nc);
    const int k = c->macro_slot;
    const double *m = c->d_macro[k];
    HIP_TRY(hipEventRecord(c->copy_go, c->stream));
    HIP_TRY(hipStreamWaitEvent(c->copy_stream, c->copy_go, 0));
    if (rho) HIP_TRY(hipMemcpyAsync(rho, m, N * sizeof(double), hipMemcpyDeviceToHost, c->copy_stream));
    if (u) HIP_TRY(hipMemcpyAsync(u, m + N, 2 * N * sizeof(double), hipMemcpyDeviceToHost, c->copy_stream));
    HIP_TRY(hipEventRecord(c->macro_free[k], c->copy_stream));
    return new_ticket(c, c->copy_stream, ticket);
}

int lbm_wait(lbm_ctx *c, int64_t ticket)
{
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    if (ticket <= 0 || ticket > c->tick_seq) return fail(LBM_ERR_INVALID, "unknown ticket %lld", (long long)ticket);
    if (ticket <= c->tick_seq - lbm_ctx::kTickets) return LBM_OK;
    HIP_TRY(hipSetDevice(c->desc.device));
    HIP_TRY(hipEventSynchronize(c->tick_ev[ticket % lbm_ctx::kTickets]));
    return LBM_OK;
}

int lbm_host_alloc(int64_t bytes, void **ptr)
{
    if (!ptr || bytes < 0) return fail(LBM_ERR_INVALID, "bad argument");
    *ptr = nullptr;
    if (bytes == 0) return LBM_OK;
    HIP_TRY(hipHostMalloc(ptr, (size_t)bytes, hipHostMallocDefault));
    return LBM_OK;
}

int lbm_host_free(void *ptr)
{
    if (ptr) HIP_TRY(hipHostFree(ptr));
    return LBM_OK;
}

int lbm_lift_drag(lbm_ctx *c, const int64_t *ij, const int32_t *masks, int64_t n, double lift_drag[2],
                  double *per_point)
{
    if (!c || !lift_drag) return fail(LBM_ERR_INVALID, "null argument");
    if (c->state != kFstar) return fail(LBM_ERR_STATE, "no step completed since the last init/upload");
    HIP_TRY(hipSetDevice(c->desc.device));
    int r = stage_points(c, ij, masks, n, &n);
    if (r) return r;
    if (n == 0) {
        lift_drag[0] = lift_drag[1] = 0.0;
        return LBM_OK;
    }
    r = queue_lift_drag(c, n, per_point != nullptr);
    if (r) return r;
    HIP_TRY(hipMemcpyAsync(lift_drag, c->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (per_point)
        HIP_TRY(hipMemcpyAsync(per_point, c->d_pp, (size_t)(2 * n) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_lift_drag_async(lbm_ctx *c, const int64_t *ij, const int32_t *masks, int64_t n, double *lift_drag,
                        double *per_point, int64_t *ticket)
{
    if (!c || !lift_drag || !ticket) return fail(LBM_ERR_INVALID, "null argument");
    if (c->state != kFstar) return fail(LBM_ERR_STATE, "no step completed since the last init/upload");
    HIP_TRY(hipSetDevice(c->desc.device));
    int r = stage_points(c, ij, masks, n, &n);
    if (r) return r;
    if (n == 0) {
        lift_drag[0] = lift_drag[1] = 0.0;
        return new_ticket(c, c->stream, ticket);
    }
    r = queue_lift_drag(c, n, per_point != nullptr);
    if (r) return r;
    HIP_TRY(hipMemcpyAsync(lift_drag, c->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (per_point)
        HIP_TRY(hipMemcpyAsync(per_point, c->d_pp, (size_t)(2 * n) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return new_ticket(c, c->stream, ticket);
}

int lbm_reduce(lbm_ctx *c, double out[5])
{
    if (!c || !out) return fail(LBM_ERR_INVALID, "null argument");
    if (c->state == kNone) return fail(LBM_ERR_STATE, "no populations");
    HIP_TRY(hipSetDevice(c->desc.device));
    StepArgs a = make_args(c);
    a.fin = c->d_F[c->cur];
    int64_t nb = readout_blocks(a);
    int r = ensure_partials(c, 5 * nb);
    if (r) return r;
    HIP_TRY(launch_readout(c->storage, c->state == kFstar, a, nullptr, c->d_partials, c->stream));
    HIP_TRY(launch_reduce_partials(c->d_partials, nb, 5, c->d_out, c->stream));
    HIP_TRY(hipMemcpyAsync(out, c->d_out, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_reduce_device(lbm_ctx *c, double *out)
{
    if (!c || !out) return fail(LBM_ERR_INVALID, "null argument");
    if (c->state == kNone) return fail(LBM_ERR_STATE, "no populations");
    HIP_TRY(hipSetDevice(c->desc.device));
    StepArgs a = make_args(c);
    a.fin = c->d_F[c->cur];
    const int64_t nb = readout_blocks(a);
    if (nb > kScratchBlocks)
        return fail(LBM_ERR_UNSUPPORTED, "lattice needs %lld reduction blocks (ceil(NR / 256) * NC); the "
                    "allocation-free reduction holds %lld", (long long)nb, (long long)kScratchBlocks);
    double *partials = scratch_ptr(c->desc.device);
    if (!partials) return fail(LBM_ERR_STATE, "scratch of device %d not set up", c->desc.device);
    // the same kernels and partial layout as lbm_reduce: the same bits
    HIP_TRY(launch_readout(c->storage, c->state == kFstar, a, nullptr, partials, c->stream));
    HIP_TRY(launch_reduce_partials(partials, nb, 5, out, c->stream));
    return LBM_OK;
}

int lbm_synchronize(lbm_ctx *c)
{
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    HIP_TRY(hipSetDevice(c->desc.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_get_stream(lbm_ctx *c, void **stream)
{
    if (!c || !stream) return fail(LBM_ERR_INVALID, "null argument");
    *stream = (void *)c->stream;
    return LBM_OK;
}

int lbm_get_scale(lbm_ctx *c, double *scale)
{
    if (!c || !scale) return fail(LBM_ERR_INVALID, "null argument");
    *scale = c->scale;
    return LBM_OK;
}

}  // extern "C"
