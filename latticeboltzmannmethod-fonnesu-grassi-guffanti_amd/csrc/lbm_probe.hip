// Practical HBM ceilings on the current part, timed with events:
//  * a copy kernel with 16-byte lanes.  The step kernels read and write
//    every population once per step, the same traffic shape as a copy of
//    one population buffer, so this rate is the yardstick their achieved
//    bandwidth is set against (beside the nominal 8 TB/s);
//  * the access patterns a D2Q9 step can have, without its arithmetic: the
//    pull between two buffers of 9 planes (the layout the step kernels
//    use), the AA in-place pattern (one buffer: even steps read and write a
//    node's own 9 slots, odd steps its neighbours'), and the pull in a
//    blocked layout (the 9 populations of 64 consecutive nodes together).
//    They measure what a layout could gain before any kernel is written
//    for it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbm_context.h"

namespace {

using u32x4 = unsigned int __attribute__((ext_vector_type(4)));

constexpr int kCopyBlock = 256;

// One 16-byte element per lane.  STREAM: the stores bypass the caches
// (whole lines, so no read-modify-write).
template <bool STREAM>
__global__ __launch_bounds__(kCopyBlock) void copy_kernel(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                          int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kCopyBlock + threadIdx.x;
    if (i >= n) return;
    const u32x4 v = src[i];
    if (STREAM) __builtin_nontemporal_store(v, dst + i);
    else dst[i] = v;
}

struct Buffers {
    void *src = nullptr, *dst = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Buffers()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (stream) (void)hipStreamDestroy(stream);
        if (src) (void)hipFree(src);
        if (dst) (void)hipFree(dst);
    }
};

template <bool STREAM>
int time_copy(Buffers &b, int64_t n, int32_t reps, float *ms)
{
    const int64_t blocks = (n + kCopyBlock - 1) / kCopyBlock;
    auto launch = [&] {
        hipLaunchKernelGGL(copy_kernel<STREAM>, dim3((unsigned)blocks), dim3(kCopyBlock), 0, b.stream,
                           (const u32x4 *)b.src, (u32x4 *)b.dst, n);
        return hipGetLastError();
    };
    for (int w = 0; w < 2; ++w) HIP_TRY(launch());
    HIP_TRY(hipEventRecord(b.e0, b.stream));
    for (int r = 0; r < reps; ++r) HIP_TRY(launch());
    HIP_TRY(hipEventRecord(b.e1, b.stream));
    HIP_TRY(hipEventSynchronize(b.e1));
    HIP_TRY(hipEventElapsedTime(ms, b.e0, b.e1));
    return LBM_OK;
}

// ------------------------------------------------------- access patterns --
//
// A dense lattice of N = NR * NC nodes in column order; direction d's
// neighbour offset is off[d] = ci[d] + NR * cj[d] (the step's pull sources
// are n - off[d]).  Indices wrap modulo N, which keeps every access inside
// its plane and makes n -> n + off[d] a bijection, so the AA odd step's
// read set and write set per thread are the same nine slots and no two
// threads touch one slot.

constexpr int kPatternBlock = 256;

struct PatternOff {
    int64_t off[9];
};

__device__ __forceinline__ int64_t wrap_index(int64_t i, int64_t N) { return i < 0 ? i + N : (i >= N ? i - N : i); }

constexpr int kPatOpp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

// Every loaded value is named before the first store, so the nine loads of
// a thread are in flight together, as in the step kernels.
template <typename T>
__device__ __forceinline__ void pin9(T *v)
{
    asm volatile("" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]), "+v"(v[5]), "+v"(v[6]), "+v"(v[7]),
                 "+v"(v[8]));
}

// A-B pull: plane d of `out` at n <- plane d of `in` at n - off[d].
template <typename T>
__global__ __launch_bounds__(kPatternBlock) void pattern_ab(const T *__restrict__ in, T *__restrict__ out, int64_t N,
                                                            PatternOff o)
{
    const int64_t n = (int64_t)blockIdx.x * kPatternBlock + threadIdx.x;
    if (n >= N) return;
    T v[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) v[d] = in[(size_t)d * N + wrap_index(n - o.off[d], N)];
    pin9(v);
#pragma unroll
    for (int d = 0; d < 9; ++d) out[(size_t)d * N + n] = v[d];
}

// AA even step: read the node's own slots, write them back swapped
// (slot opp(d) <- slot d).
template <typename T>
__global__ __launch_bounds__(kPatternBlock) void pattern_aa_even(T *__restrict__ buf, int64_t N)
{
    const int64_t n = (int64_t)blockIdx.x * kPatternBlock + threadIdx.x;
    if (n >= N) return;
    T v[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) v[d] = buf[(size_t)d * N + n];
    pin9(v);
#pragma unroll
    for (int d = 0; d < 9; ++d) buf[(size_t)kPatOpp[d] * N + n] = v[d];
}

// AA odd step: read slot opp(d) at n - off[d] (what streams into n), write
// slot d at n + off[d] (what streams out of n): the same nine locations.
template <typename T>
__global__ __launch_bounds__(kPatternBlock) void pattern_aa_odd(T *__restrict__ buf, int64_t N, PatternOff o)
{
    const int64_t n = (int64_t)blockIdx.x * kPatternBlock + threadIdx.x;
    if (n >= N) return;
    T v[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) v[d] = buf[(size_t)kPatOpp[d] * N + wrap_index(n - o.off[d], N)];
    pin9(v);
#pragma unroll
    for (int d = 0; d < 9; ++d) buf[(size_t)d * N + wrap_index(n + o.off[d], N)] = v[kPatOpp[d]];
}

// Blocked pull: element (d, n) at ((n / 64) * 9 + d) * 64 + n % 64 (N a
// multiple of 64).
template <typename T>
__global__ __launch_bounds__(kPatternBlock) void pattern_blocked(const T *__restrict__ in, T *__restrict__ out,
                                                                 int64_t N, PatternOff o)
{
    const int64_t n = (int64_t)blockIdx.x * kPatternBlock + threadIdx.x;
    if (n >= N) return;
    auto at = [](int64_t m, int d) { return (size_t)(((m >> 6) * 9 + d) * 64 + (m & 63)); };
    T v[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) v[d] = in[at(wrap_index(n - o.off[d], N), d)];
    pin9(v);
#pragma unroll
    for (int d = 0; d < 9; ++d) out[at(n, d)] = v[d];
}

template <typename T>
int time_pattern(Buffers &b, int32_t pattern, int64_t N, const PatternOff &o, int32_t reps, float *ms)
{
    const unsigned blocks = (unsigned)((N + kPatternBlock - 1) / kPatternBlock);
    const dim3 grid(blocks), block(kPatternBlock);
    T *x = (T *)b.src, *y = (T *)b.dst;
    // launch k of a timed sequence; A-B and blocked alternate the buffers,
    // LBM_PATTERN_AA alternates even and odd steps
    auto launch = [&](int k) {
        switch (pattern) {
        case LBM_PATTERN_AB:
            hipLaunchKernelGGL(pattern_ab<T>, grid, block, 0, b.stream, (k & 1) ? y : x, (k & 1) ? x : y, N, o);
            break;
        case LBM_PATTERN_AA:
            if (k & 1) hipLaunchKernelGGL(pattern_aa_odd<T>, grid, block, 0, b.stream, x, N, o);
            else hipLaunchKernelGGL(pattern_aa_even<T>, grid, block, 0, b.stream, x, N);
            break;
        case LBM_PATTERN_AA_EVEN: hipLaunchKernelGGL(pattern_aa_even<T>, grid, block, 0, b.stream, x, N); break;
        case LBM_PATTERN_AA_ODD: hipLaunchKernelGGL(pattern_aa_odd<T>, grid, block, 0, b.stream, x, N, o); break;
        default:
            hipLaunchKernelGGL(pattern_blocked<T>, grid, block, 0, b.stream, (k & 1) ? y : x, (k & 1) ? x : y, N, o);
            break;
        }
        return hipGetLastError();
    };
    for (int w = 0; w < 2; ++w) HIP_TRY(launch(w));
    HIP_TRY(hipEventRecord(b.e0, b.stream));
    for (int r = 0; r < reps; ++r) HIP_TRY(launch(r));
    HIP_TRY(hipEventRecord(b.e1, b.stream));
    HIP_TRY(hipEventSynchronize(b.e1));
    HIP_TRY(hipEventElapsedTime(ms, b.e0, b.e1));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double *gbs)
{
    using lbm::fail;
    if (!gbs) return fail(LBM_ERR_INVALID, "null output");
    if (bytes < (1 << 20) || reps < 1 || reps > 1000)
        return fail(LBM_ERR_INVALID, "copy probe: bytes >= 1 MiB and 1 <= reps <= 1000 (got %lld, %d)",
                    (long long)bytes, reps);
    const int64_t n = bytes / 16;
    if ((n + kCopyBlock - 1) / kCopyBlock > 0x7fffffff) return fail(LBM_ERR_INVALID, "copy probe: size too large");
    HIP_TRY(hipSetDevice(device));
    Buffers b;
    if (hipMalloc(&b.src, n * 16) != hipSuccess || hipMalloc(&b.dst, n * 16) != hipSuccess) {
        (void)hipGetLastError();
        return fail(LBM_ERR_NOMEM, "copy probe: cannot allocate 2 x %lld bytes", (long long)(n * 16));
    }
    HIP_TRY(hipStreamCreate(&b.stream));
    HIP_TRY(hipEventCreate(&b.e0));
    HIP_TRY(hipEventCreate(&b.e1));
    HIP_TRY(hipMemsetAsync(b.src, 1, n * 16, b.stream));
    HIP_TRY(hipMemsetAsync(b.dst, 0, n * 16, b.stream));
    float ms_plain = 0.f, ms_stream = 0.f;
    int rc = time_copy<false>(b, n, reps, &ms_plain);
    if (rc != LBM_OK) return rc;
    rc = time_copy<true>(b, n, reps, &ms_stream);
    if (rc != LBM_OK) return rc;
    const double best_ms = ms_plain < ms_stream ? ms_plain : ms_stream;
    *gbs = 2.0 * (double)(n * 16) * reps / (best_ms * 1e-3) / 1e9;
    return LBM_OK;
}

int lbm_pattern_bandwidth(int32_t device, int32_t pattern, int32_t elem_bytes, int64_t n_rows, int64_t n_cols,
                          int32_t reps, double *gbs)
{
    using lbm::fail;
    if (!gbs) return fail(LBM_ERR_INVALID, "null output");
    if (pattern < LBM_PATTERN_AB || pattern > LBM_PATTERN_BLOCKED)
        return fail(LBM_ERR_INVALID, "pattern probe: unknown pattern %d", pattern);
    if (elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
        return fail(LBM_ERR_INVALID, "pattern probe: elem_bytes must be 2, 4 or 8 (got %d)", elem_bytes);
    if (n_rows < 2 || n_cols < 2 || n_rows > (1 << 20) || n_cols > (1 << 20) || reps < 2 || reps > 1000)
        return fail(LBM_ERR_INVALID, "pattern probe: 2 <= n_rows, n_cols <= 2^20 and 2 <= reps <= 1000");
    const int64_t N = n_rows * n_cols;
    if (N % 64 != 0) return fail(LBM_ERR_INVALID, "pattern probe: n_rows * n_cols must be a multiple of 64");
    const int64_t bytes = 9 * N * elem_bytes;
    if (bytes > (int64_t(1) << 36)) return fail(LBM_ERR_INVALID, "pattern probe: more than 64 GiB per buffer");
    if ((N + kPatternBlock - 1) / kPatternBlock > 0x7fffffff) return fail(LBM_ERR_INVALID, "pattern probe: too large");
    HIP_TRY(hipSetDevice(device));
    Buffers b;
    if (hipMalloc(&b.src, bytes) != hipSuccess || hipMalloc(&b.dst, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(LBM_ERR_NOMEM, "pattern probe: cannot allocate 2 x %lld bytes", (long long)bytes);
    }
    HIP_TRY(hipStreamCreate(&b.stream));
    HIP_TRY(hipEventCreate(&b.e0));
    HIP_TRY(hipEventCreate(&b.e1));
    HIP_TRY(hipMemsetAsync(b.src, 0, bytes, b.stream));
    HIP_TRY(hipMemsetAsync(b.dst, 0, bytes, b.stream));
    static const int ci[9] = {0, 0, -1, 0, 1, -1, -1, 1, 1};
    static const int cj[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    PatternOff o;
    for (int d = 0; d < 9; ++d) o.off[d] = ci[d] + n_rows * cj[d];
    float ms = 0.f;
    int rc = elem_bytes == 2   ? time_pattern<uint16_t>(b, pattern, N, o, reps, &ms)
             : elem_bytes == 4 ? time_pattern<uint32_t>(b, pattern, N, o, reps, &ms)
                               : time_pattern<uint64_t>(b, pattern, N, o, reps, &ms);
    if (rc != LBM_OK) return rc;
    *gbs = 18.0 * (double)N * elem_bytes * reps / (ms * 1e-3) / 1e9;
    return LBM_OK;
}

}  // extern "C"
