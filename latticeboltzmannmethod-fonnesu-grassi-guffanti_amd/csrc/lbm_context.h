// The lattice context behind the C ABI's opaque lbm_ctx, shared by the
// single-lattice entry points (lbm_capi.hip) and the ensemble
// (lbm_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lbm_amd.h"
#include "lbm_internal.h"

namespace lbm {

// Sets the calling thread's last error message (lbm_last_error) and
// returns `code`.
int fail(int code, const char *fmt, ...);

#define HIP_TRY(expr)                                                                                        \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return ::lbm::fail(LBM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// kNone  -- no populations yet
// kP     -- buffer `cur` holds pre-collision populations P_t (after init or
//           upload); the next step starts with a collide-only pass
// kFstar -- buffer `cur` holds F*_{t-1}; the next step gathers from it
enum State { kNone = 0, kP = 1, kFstar = 2 };

}  // namespace lbm

struct lbm_ctx {
    lbm_desc desc{};
    int32_t storage = lbm::kF64;
    size_t esize = 8;
    hipStream_t stream = nullptr;
    // geometry
    bool has_geom = false;
    int64_t nr = 0, nc = 0, ldr = 0, plane = 0, wl = 0, E = 0, segs = 0;
    bool blocked = false;  // population layout: blocked (true) or 9 planes
    int32_t occupancy = -1;  // step kernels: work-groups per CU (2..7), 0 = no limit, -1 = default_occupancy
    uint64_t *d_wet = nullptr;
    std::vector<uint64_t> h_wet;  // host copy of the wet words (tile tables)
    std::vector<int32_t> h_sb;    // host copy of the segment bases
    int32_t *d_sb = nullptr;
    int32_t *d_runs = nullptr;  // wet-run table of the run kernel
    int64_t nruns = 0;
    std::vector<int32_t> h_runs;  // host copy (ensembles build theirs from it)
    int32_t *d_trun = nullptr;    // tabled wet runs (built on first use)
    uint32_t *d_tlane = nullptr;
    int32_t *d_perim = nullptr;  // perimeter items (column, segment) of the perimeter kernel
    int64_t nperim = 0;
    uint8_t *d_edge_kind = nullptr;
    void *d_frozen = nullptr;
    void *d_F[2] = {nullptr, nullptr};
    int cur = 0;
    // prescribed values: two device copies [3][E]; `bc_prev` is the one the
    // last step used as "current", `bc_cur` the one the next step uses.
    double *d_bc[2] = {nullptr, nullptr};
    double *h_bc_pin[2] = {nullptr, nullptr};
    hipEvent_t bc_ev[2] = {nullptr, nullptr};
    int bc_cur = 0, bc_prev = 0;
    std::vector<double> h_bc;  // [3][E] latest values
    // rho/u fields of LBM_STEP_MACRO steps: two device copies, written in
    // turn, so an asynchronous download of one (on copy_stream) overlaps the
    // steps that write the other; macro_free[k] is recorded on copy_stream
    // after the last download of copy k, and a step writing copy k waits
    // for it.
    double *d_macro[2] = {nullptr, nullptr};
    int macro_slot = 0;  // the copy the last LBM_STEP_MACRO step (or init) wrote
    hipStream_t copy_stream = nullptr;
    hipEvent_t macro_free[2] = {nullptr, nullptr};
    hipEvent_t copy_go = nullptr;  // fields complete on the main stream: the copy stream may start
    // completion events of asynchronous downloads, by ticket modulo kTickets
    static constexpr int kTickets = 64;
    hipEvent_t tick_ev[kTickets] = {};
    int64_t tick_seq = 0;
    int32_t *d_flag = nullptr;
    double *d_partials = nullptr;
    int64_t partials_cap = 0;
    double *d_out = nullptr;  // 8 doubles
    int64_t *d_pts = nullptr;
    int32_t *d_masks = nullptr;
    std::vector<int64_t> h_pts;   // the point set held in d_pts / d_masks
    std::vector<int32_t> h_masks;
    double *d_pp = nullptr;
    int64_t pts_cap = 0;
    double *d_staging = nullptr;
    // host-side geometry
    std::vector<int64_t> open_ij[2], open_e[2];
    std::vector<int64_t> obst_ij;
    std::vector<int32_t> obst_mask;
    int64_t counts[7] = {0, 0, 0, 0, 0, 0, 0};
    int64_t n_wet = 0;
    // dynamic state
    int state = lbm::kNone;
    int64_t t = 0;
    float scale = 1.f;
    // set by lbm_init_equilibrium, cleared by the first step or an upload:
    // the state is then exactly zero in shifted storage, so an FP16 scale
    // chosen from the prescribed values may still change
    bool fresh_init = false;
    bool macro_valid = false;
    // bumped by every call that changes what a step reads (geometry,
    // populations, prescribed values, relaxation) or steps the lattice
    uint64_t version = 0;
    // set while the context belongs to an ensemble (lbm_batch)
    struct lbm_batch *batch = nullptr;
    // 64-row segments per wave of the step kernel (0 = default_segments)
    int32_t segments = 0;
    // partially saturated bounce-back of the wet obstacles (lbm_set_partial_bounce_back)
    bool psbb = false;
    double ps_tau = 0.0;
    uint16_t *d_ps_mask = nullptr;
    double *d_ps_b = nullptr;
    double *d_ps_val = nullptr;
    // captured step loops, by [buffer holding the state][prescribed-value copy]
    hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
};

namespace lbm {

// Step arguments of a context (geometry, prescribed values, relaxation,
// storage scale); the caller sets fin/fout/macro.
StepArgs make_args(const lbm_ctx *c);

}  // namespace lbm
