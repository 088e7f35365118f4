// Practical HBM ceiling of a plain stream on the current part: a copy
// kernel with 16-byte lanes, timed with events.  The step kernels read and
// write every population once per step, the same traffic shape as a copy
// of one population buffer, so this rate is the yardstick their achieved
// bandwidth is set against (beside the nominal 8 TB/s).
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

}  // extern "C"
