// Ensembles of independent lattices (lbm_batch_* in include/lbm_amd.h).
//
// A batch keeps, per sample, two complete step-argument sets (one per
// buffer parity) in device memory and a tile table (sample, column, first
// row of R 64-row segments) covering every tile that holds a wet node.
// One time step of the whole ensemble is one launch of batch_kernel, a
// persistent grid that walks the table.  Runs of steps replay a captured
// HIP graph of kGraphSteps launches (an even number, so a replay leaves
// the buffer parity unchanged).
//
// Parity bookkeeping: sample s's state is in buffer base_cur[s] ^ phase,
// where phase counts the ensemble's steps mod 2 since the sample's
// arguments were last built; args[2 s + p] reads buffer base_cur[s] ^ p.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/lbm_amd.h"
#include "lbm_context.h"
#include "lbm_internal.h"

using namespace lbm;

namespace {

constexpr int64_t kGraphSteps = 32;

// Device copy of an argument array, uploaded from pinned staging memory on
// the batch stream (ordered after the launches that read the old copy).
struct ArgBuffer {
    StepArgs *dev = nullptr;
    StepArgs *pin = nullptr;
    hipEvent_t ev = nullptr;  // last upload out of `pin`
    size_t n = 0;
};

}  // namespace

struct lbm_batch {
    std::vector<lbm_ctx *> samples;
    int32_t storage = kF64;
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    int phase = 0;
    std::vector<int> base_cur;
    std::vector<uint64_t> seen;
    bool fresh = true;  // arguments never built
    // launch configuration and tile table
    int32_t R = 0;
    int64_t blocks = 0;
    int32_t tiles_R = -1;
    int64_t ntiles = 0;
    int32_t *d_tiles = nullptr;
    int64_t *d_tile_begin = nullptr;
    // segments holding a wet perimeter node (the general path's share)
    int32_t *d_ptiles = nullptr;
    int64_t nptiles = 0;
    // wet-run table (launch shape kRunLaunch): every sample's runs
    int32_t *d_runs = nullptr;
    int64_t nruns = 0;
    // "first" = each sample's arguments as lbm_step would build them for its
    // next step; "steady" = the same with the previous prescribed values set
    // to the current ones (every step after the first of a run)
    ArgBuffer first, steady;
    int32_t *d_flags = nullptr;
    double *d_partials = nullptr;
    int64_t partials_cap = 0;
    double *d_out = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int64_t wet = 0;
};

namespace {

void drop_graphs(lbm_batch *b)
{
    for (auto &g : b->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

int alloc_args(ArgBuffer &a, size_t n)
{
    a.n = n;
    HIP_TRY(hipMalloc(&a.dev, n * sizeof(StepArgs)));
    HIP_TRY(hipHostMalloc((void **)&a.pin, n * sizeof(StepArgs), hipHostMallocDefault));
    HIP_TRY(hipEventCreateWithFlags(&a.ev, hipEventDisableTiming));
    return LBM_OK;
}

void free_args(ArgBuffer &a)
{
    if (a.ev) (void)hipEventSynchronize(a.ev);
    (void)hipFree(a.dev);
    if (a.pin) (void)hipHostFree(a.pin);
    if (a.ev) (void)hipEventDestroy(a.ev);
    a = ArgBuffer{};
}

void fill_args(const lbm_batch *b, bool steady, StepArgs *out)
{
    for (size_t s = 0; s < b->samples.size(); ++s) {
        const lbm_ctx *c = b->samples[s];
        for (int p = 0; p < 2; ++p) {
            StepArgs a = make_args(c);
            if (steady) a.bc_prev = a.bc_cur;
            const int in = b->base_cur[s] ^ p;
            a.fin = c->d_F[in];
            a.fout = c->d_F[1 - in];
            a.macro = nullptr;
            a.flag = b->d_flags + s;
            set_planes(a);
            out[2 * s + p] = a;
        }
    }
}

int upload_args(lbm_batch *b, ArgBuffer &ab, bool steady)
{
    HIP_TRY(hipEventSynchronize(ab.ev));
    fill_args(b, steady, ab.pin);
    HIP_TRY(hipMemcpyAsync(ab.dev, ab.pin, ab.n * sizeof(StepArgs), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipEventRecord(ab.ev, b->stream));
    return LBM_OK;
}

// Tile table for R segments per tile: every (sample, column, 64R-row
// block) that contains a wet node, sample-major then column-major.
int build_tiles(lbm_batch *b, int32_t R)
{
    std::vector<int32_t> tiles;
    std::vector<int64_t> begin(b->samples.size() + 1, 0);
    for (size_t s = 0; s < b->samples.size(); ++s) {
        const lbm_ctx *c = b->samples[s];
        const int64_t nrw = c->ldr / 64;
        for (int64_t j = 0; j < c->nc; ++j) {
            const uint64_t *col = c->h_wet.data() + (size_t)((j + 1) * c->wl + 1);
            for (int64_t k0 = 0; k0 < nrw; k0 += R) {
                uint64_t any = 0;
                for (int64_t k = k0; k < std::min<int64_t>(k0 + R, nrw); ++k) any |= col[k];
                if (!any) continue;
                tiles.push_back((int32_t)s);
                tiles.push_back((int32_t)j);
                tiles.push_back((int32_t)(64 * k0));
            }
        }
        begin[s + 1] = (int64_t)(tiles.size() / 3);
    }
    // the previous table may still be read by queued launches
    HIP_TRY(hipStreamSynchronize(b->stream));
    (void)hipFree(b->d_tiles);
    b->d_tiles = nullptr;
    b->ntiles = (int64_t)(tiles.size() / 3);
    HIP_TRY(hipMalloc(&b->d_tiles, std::max<size_t>(tiles.size(), 3) * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_tile_begin, begin.data(), begin.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    b->tiles_R = R;
    drop_graphs(b);
    return LBM_OK;
}

// Perimeter table: every 64-row segment holding a wet perimeter node
// (column 0 or NC-1, row 0 or NR-1) -- the segments the step kernel leaves
// to the general path; the same test as perimeter_segment on the device.
int build_perimeter_tiles(lbm_batch *b)
{
    std::vector<int32_t> tiles;
    for (size_t s = 0; s < b->samples.size(); ++s) {
        const lbm_ctx *c = b->samples[s];
        const int64_t nrw = c->ldr / 64, last = c->nr - 1;
        for (int64_t j = 0; j < c->nc; ++j) {
            const uint64_t *col = c->h_wet.data() + (size_t)((j + 1) * c->wl + 1);
            for (int64_t k = 0; k < nrw; ++k) {
                const uint64_t mid = col[k];
                if (!mid) continue;
                const bool edge_col = j == 0 || j == c->nc - 1;
                const bool top = k == 0 && (mid & 1);
                const bool bottom = k == last / 64 && ((mid >> (last % 64)) & 1);
                if (!(edge_col || top || bottom)) continue;
                tiles.push_back((int32_t)s);
                tiles.push_back((int32_t)j);
                tiles.push_back((int32_t)(64 * k));
            }
        }
    }
    b->nptiles = (int64_t)(tiles.size() / 3);
    HIP_TRY(hipMalloc(&b->d_ptiles, std::max<size_t>(tiles.size(), 3) * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(b->d_ptiles, tiles.data(), tiles.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    return LBM_OK;
}

// Wet-run table of the ensemble: each sample's own run table (built with
// its geometry) behind the sample index, in sample order.
int build_runs(lbm_batch *b)
{
    std::vector<int32_t> runs;
    for (size_t s = 0; s < b->samples.size(); ++s) {
        const std::vector<int32_t> &r = b->samples[s]->h_runs;
        for (size_t k = 0; k + 3 < r.size(); k += 4) {
            runs.push_back((int32_t)s);
            runs.insert(runs.end(), r.begin() + (long)k, r.begin() + (long)k + 4);
        }
    }
    b->nruns = (int64_t)(runs.size() / 5);
    HIP_TRY(hipMalloc(&b->d_runs, std::max<size_t>(runs.size(), 5) * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(b->d_runs, runs.data(), runs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    return LBM_OK;
}

int64_t round8(int64_t v) { return std::max<int64_t>(8, (v + 7) / 8 * 8); }

// Brings the device arguments up to date with the samples: a sample whose
// context changed since the last build (new values, populations, ...) is
// synchronised with its own stream and its arguments are rebuilt.
int refresh(lbm_batch *b, bool &changed)
{
    changed = b->fresh;
    for (size_t s = 0; s < b->samples.size(); ++s) {
        lbm_ctx *c = b->samples[s];
        if (b->fresh || c->version != b->seen[s]) {
            HIP_TRY(hipStreamSynchronize(c->stream));
            b->base_cur[s] = c->cur ^ b->phase;
            b->seen[s] = c->version;
            changed = true;
        }
    }
    b->fresh = false;
    if (changed) {
        int r = upload_args(b, b->first, false);
        if (r) return r;
        r = upload_args(b, b->steady, true);
        if (r) return r;
    }
    return LBM_OK;
}

int common_state(const lbm_batch *b)
{
    const int st = b->samples[0]->state;
    for (const lbm_ctx *c : b->samples)
        if (c->state != st) return -1;
    return st;
}

BatchArgs batch_args(const lbm_batch *b, const ArgBuffer &ab, int parity)
{
    BatchArgs a;
    a.args = ab.dev;
    a.tiles = b->d_tiles;
    a.ntiles = b->ntiles;
    a.ptiles = b->d_ptiles;
    a.nptiles = b->nptiles;
    a.parity = parity;
    a.direct = 0;
    a.runs = b->d_runs;
    a.nruns = b->nruns;
    return a;
}

int capture_graph(lbm_batch *b, int phase)
{
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    for (int64_t k = 0; k < kGraphSteps && err == hipSuccess; ++k)
        err = launch_batch_step(b->storage, true, b->R, batch_args(b, b->steady, phase ^ (int)(k & 1)), b->blocks,
                                b->stream);
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (err != hipSuccess) return fail(LBM_ERR_HIP, "batch graph capture: %s", hipGetErrorString(err));
    if (e2 != hipSuccess) return fail(LBM_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
    err = hipGraphInstantiate(&b->graph[phase], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess) return fail(LBM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(err));
    return LBM_OK;
}

// Orders every sample's own stream after the ensemble's work so far.
int release_samples(lbm_batch *b)
{
    HIP_TRY(hipEventRecord(b->done, b->stream));
    for (lbm_ctx *c : b->samples) HIP_TRY(hipStreamWaitEvent(c->stream, b->done, 0));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_create(lbm_ctx *const *samples, int64_t n, lbm_batch **out)
{
    if (!samples || !out) return fail(LBM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n < 1) return fail(LBM_ERR_INVALID, "an ensemble needs at least one sample");
    if (n > (1 << 30)) return fail(LBM_ERR_INVALID, "too many samples");
    for (int64_t s = 0; s < n; ++s) {
        const lbm_ctx *c = samples[s];
        if (!c) return fail(LBM_ERR_INVALID, "sample %lld is null", (long long)s);
        if (c->batch) return fail(LBM_ERR_STATE, "sample %lld already belongs to an ensemble", (long long)s);
        if (!c->has_geom) return fail(LBM_ERR_STATE, "sample %lld has no geometry", (long long)s);
        if (c->storage != samples[0]->storage) return fail(LBM_ERR_INVALID, "samples differ in storage format");
        if (c->psbb)
            return fail(LBM_ERR_UNSUPPORTED, "sample %lld uses partially saturated bounce-back (single lattices only)",
                        (long long)s);
        if (c->desc.device != samples[0]->desc.device) return fail(LBM_ERR_INVALID, "samples are on different devices");
        for (int64_t q = 0; q < s; ++q)
            if (samples[q] == c) return fail(LBM_ERR_INVALID, "sample %lld appears twice", (long long)s);
    }
    lbm_batch *b = new lbm_batch();
    b->samples.assign(samples, samples + n);
    b->storage = samples[0]->storage;
    b->device = samples[0]->desc.device;
    b->base_cur.assign((size_t)n, 0);
    b->seen.assign((size_t)n, 0);
    for (const lbm_ctx *c : b->samples) b->wet += c->n_wet;
    auto bail = [&](int r) {
        lbm_batch_destroy(b);
        return r;
    };
    if (hipSetDevice(b->device) != hipSuccess) return bail(fail(LBM_ERR_HIP, "hipSetDevice"));
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess && cus > 0)
        b->cus = cus;
    b->R = default_segments(b->storage);
    b->blocks = 0;  // one wave per tile
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->t0);
    if (e == hipSuccess) e = hipEventCreate(&b->t1);
    if (e == hipSuccess) e = hipMalloc(&b->d_flags, (size_t)n * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemsetAsync(b->d_flags, 0, (size_t)n * sizeof(int32_t), b->stream);
    if (e == hipSuccess) e = hipMalloc(&b->d_tile_begin, (size_t)(n + 1) * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&b->d_out, (size_t)(5 * n) * sizeof(double));
    if (e != hipSuccess) return bail(fail(LBM_ERR_HIP, "ensemble set-up: %s", hipGetErrorString(e)));
    int r = alloc_args(b->first, (size_t)(2 * n));
    if (!r) r = alloc_args(b->steady, (size_t)(2 * n));
    if (!r) r = build_tiles(b, b->R);
    if (!r) r = build_perimeter_tiles(b);
    if (!r) r = build_runs(b);
    if (r) return bail(r);
    for (lbm_ctx *c : b->samples) c->batch = b;
    *out = b;
    return LBM_OK;
}

int lbm_batch_destroy(lbm_batch *b)
{
    if (!b) return LBM_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (lbm_ctx *c : b->samples)
        if (c->batch == b) c->batch = nullptr;
    free_args(b->first);
    free_args(b->steady);
    (void)hipFree(b->d_tiles);
    (void)hipFree(b->d_ptiles);
    (void)hipFree(b->d_runs);
    (void)hipFree(b->d_tile_begin);
    (void)hipFree(b->d_flags);
    (void)hipFree(b->d_partials);
    (void)hipFree(b->d_out);
    for (hipEvent_t ev : {b->done, b->t0, b->t1})
        if (ev) (void)hipEventDestroy(ev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return LBM_OK;
}

int lbm_batch_set_launch(lbm_batch *b, int32_t segments, int64_t blocks)
{
    if (!b) return fail(LBM_ERR_INVALID, "null batch");
    if (segments != 1 && segments != 2 && segments != 4 && segments != 8 && segments != kRunLaunch)
        return fail(LBM_ERR_INVALID, "segments must be 1, 2, 4, 8 or %d (wet runs)", kRunLaunch);
    if (blocks < 0 || blocks > (1 << 24)) return fail(LBM_ERR_INVALID, "grid size out of range");
    if (segments == kRunLaunch && blocks != 0)
        return fail(LBM_ERR_INVALID, "wet-run waves take one wave per run (blocks = 0)");
    HIP_TRY(hipSetDevice(b->device));
    const int64_t nb = blocks == 0 ? 0 : round8(blocks);
    // wet runs keep a one-segment tile table for the collide-only first
    // step and the reductions
    const int32_t tiles_R = segments == kRunLaunch ? 1 : segments;
    if (tiles_R != b->tiles_R) {
        int r = build_tiles(b, tiles_R);
        if (r) return r;
    }
    if (nb != b->blocks || segments != b->R) drop_graphs(b);
    b->R = segments;
    b->blocks = nb;
    return LBM_OK;
}

int lbm_batch_get_launch(lbm_batch *b, int32_t *segments, int64_t *blocks)
{
    if (!b) return fail(LBM_ERR_INVALID, "null batch");
    if (segments) *segments = b->R;
    if (blocks) *blocks = b->blocks;
    return LBM_OK;
}

int lbm_batch_autotune(lbm_batch *b, int32_t steps, int32_t *segments, int64_t *blocks, double *ms_per_step)
{
    if (!b) return fail(LBM_ERR_INVALID, "null batch");
    if (steps < 1) return fail(LBM_ERR_INVALID, "steps must be >= 1");
    if (common_state(b) <= kNone) return fail(LBM_ERR_STATE, "samples need populations, all in the same state");
    HIP_TRY(hipSetDevice(b->device));
    bool changed = false;
    int r = refresh(b, changed);
    if (r) return r;
    const int32_t Rs[] = {1, 2, 4, 8, kRunLaunch};
    const int64_t per_cu[] = {0, 2, 4, 8, 16};  // 0: one wave per tile
    float best = -1.f;
    int32_t best_R = b->R;
    int64_t best_nb = b->blocks;
    for (int32_t R : Rs) {
        r = build_tiles(b, R == kRunLaunch ? 1 : R);
        if (r) return r;
        for (int64_t k : per_cu) {
            if (R == kRunLaunch && k != 0) continue;  // wet runs: one wave per run only
            const int64_t nb = k == 0 ? 0 : round8(k * (int64_t)b->cus);
            // every launch recomputes the spare buffer from the current
            // state (fixed parity), so the state is unchanged
            const BatchArgs ba = batch_args(b, b->steady, b->phase);
            HIP_TRY(launch_batch_step(b->storage, true, R, ba, nb, b->stream));
            HIP_TRY(hipEventRecord(b->t0, b->stream));
            for (int32_t s = 0; s < steps; ++s) HIP_TRY(launch_batch_step(b->storage, true, R, ba, nb, b->stream));
            HIP_TRY(hipEventRecord(b->t1, b->stream));
            HIP_TRY(hipEventSynchronize(b->t1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, b->t0, b->t1));
            if (best < 0 || ms < best) {
                best = ms;
                best_R = R;
                best_nb = nb;
            }
        }
    }
    r = build_tiles(b, best_R == kRunLaunch ? 1 : best_R);
    if (r) return r;
    b->R = best_R;
    b->blocks = best_nb;
    drop_graphs(b);
    if (segments) *segments = best_R;
    if (blocks) *blocks = best_nb;
    if (ms_per_step) *ms_per_step = best / steps;
    return LBM_OK;
}

int lbm_batch_step(lbm_batch *b, int64_t n, uint32_t flags)
{
    if (!b) return fail(LBM_ERR_INVALID, "null batch");
    if (n < 0) return fail(LBM_ERR_INVALID, "negative step count");
    if (flags & LBM_STEP_MACRO) return fail(LBM_ERR_UNSUPPORTED, "LBM_STEP_MACRO is not available on an ensemble");
    const int st = common_state(b);
    if (st < 0) return fail(LBM_ERR_STATE, "samples are in different states (initialise or step them together)");
    if (st == kNone) return fail(LBM_ERR_STATE, "no populations (call lbm_init_equilibrium on every sample)");
    HIP_TRY(hipSetDevice(b->device));
    bool changed = false;
    int r = refresh(b, changed);
    if (r) return r;
    int64_t left = n;
    if (left > 0) {
        bool any_new_values = false;
        for (const lbm_ctx *c : b->samples) any_new_values |= c->bc_prev != c->bc_cur;
        if (st == kP || any_new_values) {
            // first step of a run: each sample's own previous values (and the
            // collide-only pass after an init or upload)
            HIP_TRY(launch_batch_step(b->storage, st == kFstar, b->R, batch_args(b, b->first, b->phase), b->blocks,
                                      b->stream));
            b->phase ^= 1;
            --left;
            for (lbm_ctx *c : b->samples) {
                c->bc_prev = c->bc_cur;
                c->state = kFstar;
                c->fresh_init = false;
            }
            // "first" now equals "steady"
            r = upload_args(b, b->first, false);
            if (r) return r;
        }
        if (!(flags & LBM_STEP_NO_GRAPH) && left >= kGraphSteps) {
            if (!b->graph[b->phase]) {
                r = capture_graph(b, b->phase);
                if (r) return r;
            }
            for (; left >= kGraphSteps; left -= kGraphSteps) HIP_TRY(hipGraphLaunch(b->graph[b->phase], b->stream));
        }
        for (; left > 0; --left) {
            HIP_TRY(launch_batch_step(b->storage, true, b->R, batch_args(b, b->steady, b->phase), b->blocks, b->stream));
            b->phase ^= 1;
        }
        for (size_t s = 0; s < b->samples.size(); ++s) {
            lbm_ctx *c = b->samples[s];
            c->cur = b->base_cur[s] ^ b->phase;
            c->state = kFstar;
            c->t += n;
            c->bc_prev = c->bc_cur;
            c->macro_valid = false;
            c->fresh_init = false;
        }
    }
    r = release_samples(b);
    if (r) return r;
    if (flags & LBM_STEP_CHECK) {
        std::vector<int32_t> fl(b->samples.size());
        HIP_TRY(hipMemcpyAsync(fl.data(), b->d_flags, fl.size() * sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        for (size_t s = 0; s < fl.size(); ++s) {
            if (fl[s] & 2)
                return fail(LBM_ERR_RANGE, "FP16 storage of sample %lld: a scaled population reached 65504 at or "
                            "before step %lld", (long long)s, (long long)b->samples[s]->t);
            if (fl[s] & 1)
                return fail(LBM_ERR_NAN, "NaN density in sample %lld at or before step %lld", (long long)s,
                            (long long)b->samples[s]->t);
        }
    }
    return LBM_OK;
}

int lbm_batch_reduce(lbm_batch *b, double *out)
{
    if (!b || !out) return fail(LBM_ERR_INVALID, "null argument");
    const int st = common_state(b);
    if (st <= kNone) return fail(LBM_ERR_STATE, "samples need populations, all in the same state");
    HIP_TRY(hipSetDevice(b->device));
    bool changed = false;
    int r = refresh(b, changed);
    if (r) return r;
    if (5 * b->ntiles > b->partials_cap) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        (void)hipFree(b->d_partials);
        b->d_partials = nullptr;
        b->partials_cap = 0;
        HIP_TRY(hipMalloc(&b->d_partials, (size_t)std::max<int64_t>(5 * b->ntiles, 5) * sizeof(double)));
        b->partials_cap = std::max<int64_t>(5 * b->ntiles, 5);
    }
    const int64_t n = (int64_t)b->samples.size();
    HIP_TRY(launch_batch_reduce(b->storage, st == kFstar, b->tiles_R, batch_args(b, b->first, b->phase),
                                b->d_tile_begin, n, b->d_partials, b->d_out, b->stream));
    HIP_TRY(hipMemcpyAsync(out, b->d_out, (size_t)(5 * n) * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_info(lbm_batch *b, int64_t *wet_nodes, int64_t *tiles)
{
    if (!b) return fail(LBM_ERR_INVALID, "null batch");
    if (wet_nodes) *wet_nodes = b->wet;
    if (tiles) *tiles = b->ntiles;
    return LBM_OK;
}

int lbm_batch_get_stream(lbm_batch *b, void **stream)
{
    if (!b || !stream) return fail(LBM_ERR_INVALID, "null argument");
    *stream = (void *)b->stream;
    return LBM_OK;
}

int lbm_batch_synchronize(lbm_batch *b)
{
    if (!b) return fail(LBM_ERR_INVALID, "null batch");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

}  // extern "C"
