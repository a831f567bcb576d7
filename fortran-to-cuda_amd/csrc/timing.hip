// HIP events on a context's stream, for timing submitted work from the host without a framework:
// the elapsed time between two events recorded on the stream the kernels run on.
#include <hip/hip_runtime.h>

#include "../../include/fort_scope.h"
#include "xcd.hpp"

// Grid-stride copy of n doubles as 16-byte pairs, four pairs in flight per thread; the odd last double
// (if any) is copied by thread 0.
__global__ __launch_bounds__(256) void ftc_copy_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                                       size_t n)
{
    const double2 *__restrict__ s2 = reinterpret_cast<const double2 *>(src);
    double2 *__restrict__ d2 = reinterpret_cast<double2 *>(dst);
    const size_t n2 = n / 2, stride = (size_t)gridDim.x * blockDim.x;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; e + 3 * stride < n2; e += 4 * stride) {
        const double2 a0 = s2[e], a1 = s2[e + stride], a2 = s2[e + 2 * stride], a3 = s2[e + 3 * stride];
        d2[e] = a0;
        d2[e + stride] = a1;
        d2[e + 2 * stride] = a2;
        d2[e + 3 * stride] = a3;
    }
    for (; e < n2; e += stride) d2[e] = s2[e];
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) dst[n - 1] = src[n - 1];
}

typedef double ftc_d2v __attribute__((ext_vector_type(2)));

// The copy above with nontemporal loads and stores.
__global__ __launch_bounds__(256) void ftc_copy_nt_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                                          size_t n)
{
    const ftc_d2v *__restrict__ s2 = reinterpret_cast<const ftc_d2v *>(src);
    ftc_d2v *__restrict__ d2 = reinterpret_cast<ftc_d2v *>(dst);
    const size_t n2 = n / 2, stride = (size_t)gridDim.x * blockDim.x;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; e + 3 * stride < n2; e += 4 * stride) {
        const ftc_d2v a0 = __builtin_nontemporal_load(&s2[e]), a1 = __builtin_nontemporal_load(&s2[e + stride]);
        const ftc_d2v a2 = __builtin_nontemporal_load(&s2[e + 2 * stride]);
        const ftc_d2v a3 = __builtin_nontemporal_load(&s2[e + 3 * stride]);
        __builtin_nontemporal_store(a0, &d2[e]);
        __builtin_nontemporal_store(a1, &d2[e + stride]);
        __builtin_nontemporal_store(a2, &d2[e + 2 * stride]);
        __builtin_nontemporal_store(a3, &d2[e + 3 * stride]);
    }
    for (; e < n2; e += stride) d2[e] = s2[e];
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) dst[n - 1] = src[n - 1];
}

// Reads n doubles (16 bytes per lane, four in flight); the sum is stored only if it equals `never`, which
// the zero-filled input never gives, so the loads cannot be dropped.
__global__ __launch_bounds__(256) void ftc_read_kernel(const double *__restrict__ src, double *__restrict__ out,
                                                       size_t n, double never)
{
    const double2 *__restrict__ s2 = reinterpret_cast<const double2 *>(src);
    const size_t n2 = n / 2, stride = (size_t)gridDim.x * blockDim.x;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    for (; e + 3 * stride < n2; e += 4 * stride) {
        const double2 a0 = s2[e], a1 = s2[e + stride], a2 = s2[e + 2 * stride], a3 = s2[e + 3 * stride];
        acc += ((a0.x + a0.y) + (a1.x + a1.y)) + ((a2.x + a2.y) + (a3.x + a3.y));
    }
    for (; e < n2; e += stride) acc += s2[e].x + s2[e].y;
    if (acc == never) out[0] = acc;
}

// Writes n doubles (16 bytes per lane).
__global__ __launch_bounds__(256) void ftc_write_kernel(double *__restrict__ dst, size_t n)
{
    double2 *__restrict__ d2 = reinterpret_cast<double2 *>(dst);
    const size_t n2 = n / 2, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += stride) d2[e] = double2{1.0, 2.0};
}

__global__ void ftc_iota(double *__restrict__ x, long long n, double scale)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = scale * (double)i;
}

// Counts the visits of every tile under one of the tile orders of xcd.hpp, one thread per workgroup: scheme
// 0 launch order, 1 XCD-contiguous, 2 the banded walk of `total` tiles by gridDim.x workgroups, 3 the
// contraction order. A tile number past the last (which a correct order never gives) is counted in the
// extra slot counts[total], so every store stays in the array.
__global__ void ftc_tile_visit_kernel(int scheme, unsigned total, unsigned *counts)
{
    auto visit = [&](unsigned long long t) { atomicAdd(&counts[t < total ? t : total], 1u); };
    if (scheme == 2) {
        const ftc::TileWalk w = ftc::banded_walk(total, gridDim.x, blockIdx.x);
        for (unsigned t = w.first; t < w.end; t += w.step) visit(t);
    } else if (scheme == 3) {
        int mt, nt;
        long long b;
        ftc::gemm_block_tile(mt, nt, b);
        visit((unsigned long long)mt + gridDim.x * ((unsigned long long)nt + (unsigned long long)gridDim.y * b));
    } else {
        visit(ftc::launch_tile(scheme));
    }
}

// The FP64 MFMA issue rate: each wave runs `iters` steps of eight independent v_mfma_f64_16x16x4f64
// accumulations with no memory traffic in the loop; one double per lane is stored at the end (never read).
// Declaring four waves per SIMD holds a wave to 128 registers, so the accumulators stay where the MFMAs
// write them (with the larger budget the compiler keeps them in VGPRs and copies them through AGPRs around
// every step).
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(4, 4))) void
ftc_mfma_f64_kernel(double *__restrict__ out, int iters, double a0)
{
    typedef double f64x4 __attribute__((ext_vector_type(4)));
    f64x4 acc[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) acc[x] = f64x4{0.0, 0.0, 0.0, 0.0};
    const double a = a0 + threadIdx.x, b = 0.5 * a0 + threadIdx.x;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int x = 0; x < 8; ++x) acc[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[x], 0, 0, 0);
    double sum = 0.0;
#pragma unroll
    for (int x = 0; x < 8; ++x) sum += (acc[x][0] + acc[x][1]) + (acc[x][2] + acc[x][3]);
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

extern "C" {

const void *ftc_iota_kernel(void) { return reinterpret_cast<const void *>(&ftc_iota); }

// Measures a plain device-to-device copy of n doubles (reps launches of a grid-stride copy kernel on a
// private stream): the practical HBM ceiling, and the calibration kernel for the FETCH_SIZE/WRITE_SIZE
// counters (each launch reads and writes exactly 8*n bytes). Writes the average milliseconds per launch.

int ftc_copy_probe(size_t n, int reps, float *ms_per_launch)
{
    if (!ms_per_launch || n == 0 || reps <= 0) return FORT_SCOPE_ARGUMENT;
    double *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int status = FORT_SCOPE_RESOURCE;
    if (hipMalloc(&a, n * sizeof(double)) == hipSuccess && hipMalloc(&b, n * sizeof(double)) == hipSuccess &&
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess) {
        (void)hipMemsetAsync(a, 0, n * sizeof(double), s);
        const unsigned blocks = 256 * 32;
        hipLaunchKernelGGL(ftc_copy_kernel, dim3(blocks), dim3(256), 0, s, a, b, n);
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(ftc_copy_kernel, dim3(blocks), dim3(256), 0, s, a, b, n);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) == hipSuccess) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *ms_per_launch = ms / reps;
            status = FORT_SCOPE_OK;
        } else {
            status = FORT_SCOPE_EXECUTION;
        }
    }
    (void)hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return status;
}

// The HBM rates a kernel can reach on this part: kind 0 copy (the kernel above), 1 copy with nontemporal
// loads and stores, 2 read only, 3 write only, over n doubles (2n for the copies), reps timed launches after
// one warm-up launch on a private stream. n must be even.
int ftc_bandwidth_probe(int kind, size_t n, int reps, float *ms_per_launch)
{
    if (!ms_per_launch || n == 0 || (n & 1) || reps <= 0 || kind < 0 || kind > 3) return FORT_SCOPE_ARGUMENT;
    double *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int status = FORT_SCOPE_RESOURCE;
    if (hipMalloc(&a, n * sizeof(double)) == hipSuccess && hipMalloc(&b, n * sizeof(double)) == hipSuccess &&
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess) {
        (void)hipMemsetAsync(a, 0, n * sizeof(double), s);
        const unsigned blocks = 256 * 32;
        auto launch = [&]() {
            if (kind == 0) hipLaunchKernelGGL(ftc_copy_kernel, dim3(blocks), dim3(256), 0, s, a, b, n);
            else if (kind == 1) hipLaunchKernelGGL(ftc_copy_nt_kernel, dim3(blocks), dim3(256), 0, s, a, b, n);
            else if (kind == 2) hipLaunchKernelGGL(ftc_read_kernel, dim3(blocks), dim3(256), 0, s, a, b, n, 12345.678);
            else hipLaunchKernelGGL(ftc_write_kernel, dim3(blocks), dim3(256), 0, s, b, n);
        };
        launch();
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) launch();
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) == hipSuccess) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *ms_per_launch = ms / reps;
            status = FORT_SCOPE_OK;
        } else {
            status = FORT_SCOPE_EXECUTION;
        }
    }
    (void)hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return status;
}

// The FP64 matrix-core ceiling of this part: waves_per_simd workgroups of four waves per CU, each wave `iters`
// steps of eight independent 16x16x4 FP64 MFMAs (2048 flops each), reps timed launches after one warm-up.
// Writes the milliseconds and the flops of one launch.
int ftc_mfma_probe(int waves_per_simd, int iters, int reps, float *ms_per_launch, double *flops_per_launch)
{
    if (!ms_per_launch || !flops_per_launch || waves_per_simd <= 0 || waves_per_simd > 8 || iters <= 0 || reps <= 0)
        return FORT_SCOPE_ARGUMENT;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return FORT_SCOPE_RESOURCE;
    const unsigned blocks = (unsigned)cus * (unsigned)waves_per_simd;
    *flops_per_launch = (double)blocks * 4 * iters * 8 * 2048;
    double *out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int status = FORT_SCOPE_RESOURCE;
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(double)) == hipSuccess &&
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess) {
        hipLaunchKernelGGL(ftc_mfma_f64_kernel, dim3(blocks), dim3(256), 0, s, out, iters, 1.0);
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(ftc_mfma_f64_kernel, dim3(blocks), dim3(256), 0, s, out, iters, 1.0);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) == hipSuccess) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *ms_per_launch = ms / reps;
            status = FORT_SCOPE_OK;
        } else {
            status = FORT_SCOPE_EXECUTION;
        }
    }
    (void)hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (out) (void)hipFree(out);
    return status;
}

int ftc_event_create(void **event)
{
    hipEvent_t e = nullptr;
    if (!event || hipEventCreate(&e) != hipSuccess) return FORT_SCOPE_RESOURCE;
    *event = e;
    return FORT_SCOPE_OK;
}

int ftc_event_record(void *event, void *stream)
{
    return hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)) == hipSuccess
               ? FORT_SCOPE_OK
               : FORT_SCOPE_EXECUTION;
}

int ftc_event_elapsed_ms(void *start, void *stop, float *ms)
{
    if (!ms) return FORT_SCOPE_ARGUMENT;
    if (hipEventSynchronize(static_cast<hipEvent_t>(stop)) != hipSuccess) return FORT_SCOPE_EXECUTION;
    return hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)) == hipSuccess
               ? FORT_SCOPE_OK
               : FORT_SCOPE_EXECUTION;
}

int ftc_event_destroy(void *event)
{
    return hipEventDestroy(static_cast<hipEvent_t>(event)) == hipSuccess ? FORT_SCOPE_OK : FORT_SCOPE_ARGUMENT;
}

int ftc_tile_visit_probe_v1(int scheme, unsigned gx, unsigned gy, unsigned gz, unsigned workgroups, unsigned *counts)
{
    if (scheme < 0 || scheme > 3 || !counts || !gx || !gy || !gz || gy > 65535 || gz > 65535)
        return FORT_SCOPE_ARGUMENT;
    const unsigned long long total = (unsigned long long)gx * gy * gz;
    if (total > (1ull << 24)) return FORT_SCOPE_ARGUMENT;
    // The walk's step is workgroups / 8: fewer than eight workgroups, or a count that is not a multiple of
    // eight, is not a walk the kernels launch.
    if (scheme == 2 && (workgroups == 0 || workgroups % ftc::kXcds != 0 || workgroups > (1u << 20)))
        return FORT_SCOPE_ARGUMENT;
    unsigned *d = nullptr;
    const std::size_t bytes = (total + 1) * sizeof(unsigned);
    if (hipMalloc(&d, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return FORT_SCOPE_RESOURCE;
    }
    int status = FORT_SCOPE_OK;
    if (hipMemset(d, 0, bytes) != hipSuccess) status = FORT_SCOPE_EXECUTION;
    if (status == FORT_SCOPE_OK) {
        const dim3 grid = scheme == 2 ? dim3(workgroups) : dim3(gx, gy, gz);
        hipLaunchKernelGGL(ftc_tile_visit_kernel, grid, dim3(1), 0, nullptr, scheme, (unsigned)total, d);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(counts, d, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            status = FORT_SCOPE_EXECUTION;
    }
    (void)hipFree(d);
    return status;
}

// The number of visible devices; 0 without a GPU.
int ftc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

}  // extern "C"
