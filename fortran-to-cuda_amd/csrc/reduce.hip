// Fixed-order reductions (compiler/resident/reductions.py:1-45, SLOTS/FINAL/DIRECT at :66-72).
//
// A nest of n iterations (Fortran order over its loop box) with n > 65536 is split over 65536 slots, slot s
// taking iterations s, s + 65536, s + 2*65536, ... in order; 256 fold threads then take slot results
// f, f + 256, ... in order; the host folds the 256 values in index order into the caller's variable.
// A nest of at most 65536 iterations goes straight to 256 partial results (thread f takes iterations
// f, f + 256, ...). The order never depends on the launch, so real sums repeat exactly from run to run,
// and max/min/maxabs equal the serial result.
#include <cmath>
#include <mutex>

#include "entry.hpp"
#include "numeric.hpp"
#include "reduce.hpp"

namespace ftc {

// max and min fold as the reference's ordered maximum/minimum (compiler/runtime/numeric.hpp:48-58):
// max(acc, x) = acc > x ? acc : x, so a NaN operand or a signed zero decides by position exactly as on
// the host (the helpers keep the select from becoming v_max_f64 / v_min_f64).
template <int OP>
__device__ __forceinline__ double combine(double acc, double x)
{
    if constexpr (OP == kReduceSum) return acc + x;
    else if constexpr (OP == kReduceSumSquares) return acc + x * x;
    else if constexpr (OP == kReduceMax) return numeric::maximum(acc, x);
    else if constexpr (OP == kReduceMin) return numeric::minimum(acc, x);
    else return numeric::maximum(acc, numeric::chosen(x < 0.0) ? -x : x);
}

template <int OP>
__device__ __forceinline__ double identity()
{
    if constexpr (OP == kReduceSum || OP == kReduceSumSquares || OP == kReduceMaxAbs) return 0.0;
    else if constexpr (OP == kReduceMax) return -HUGE_VAL;
    else return HUGE_VAL;
}

// The box position (i, j, k) of an iteration, and the position step of `slots` iterations: walking a
// slot's iterations by adding the step (with carries) avoids 64-bit divisions in the loop.
struct BoxStep {
    long long di, dj, dk;
};

__device__ __forceinline__ void advance(const ReduceBox &b, const BoxStep &st, long long &i, long long &j, long long &k)
{
    i += st.di;
    if (i >= b.n1) {
        i -= b.n1;
        ++j;
    }
    j += st.dj;
    if (j >= b.n2) {
        j -= b.n2;
        ++k;
    }
    k += st.dk;
}

// Slot s folds iterations s, s + slots, ... in order. The loads of kUnroll consecutive iterations are
// issued together; the combines stay in iteration order.
template <int OP>
__device__ __forceinline__ void reduce_slots_body(const ReduceBox &b, long long n, int slots, const BoxStep &st,
                                                  double *out)
{
    constexpr int kUnroll = 8;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots) return;
    double acc = identity<OP>();
    long long i = 0, j = 0, k = 0, t = s;
    if (t < n) {  // n > 0, so the extents are positive
        const long long r = t / b.n1;
        i = t % b.n1;
        j = r % b.n2;
        k = r / b.n2;
    }
    for (; t + (long long)(kUnroll - 1) * slots < n; t += (long long)kUnroll * slots) {
        double v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            v[u] = b.x[i + j * b.ld1 + k * b.ld2];
            advance(b, st, i, j, k);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) acc = combine<OP>(acc, v[u]);
    }
    for (; t < n; t += slots) {
        acc = combine<OP>(acc, b.x[i + j * b.ld1 + k * b.ld2]);
        advance(b, st, i, j, k);
    }
    out[s] = acc;
}

template <int OP>
__global__ void reduce_slots(const ReduceBox b, long long n, int slots, BoxStep st, double *out)
{
    reduce_slots_body<OP>(b, n, slots, st, out);
}

// Folds `count` partial results into 256 (thread f takes f, f + 256, ... in order). For sums of squares
// the partial results are already squared, so the fold is a plain sum. The loads of a batch of 32 are all
// issued before the batch is combined, in order, so one memory latency covers 32 partials.
template <int OP>
__device__ __forceinline__ void reduce_fold_body(const double *partial, int count, double *out)
{
    constexpr int kBatch = 32;
    const int f = threadIdx.x;
    constexpr int FOP = OP == kReduceSumSquares ? kReduceSum : (OP == kReduceMaxAbs ? kReduceMax : OP);
    double acc = identity<FOP>();
    for (int t0 = f; t0 < count; t0 += kBatch * kReduceFinal) {
        double v[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
            const int t = t0 + q * kReduceFinal;
            v[q] = t < count ? partial[t] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kBatch; ++q)
            if (t0 + q * kReduceFinal < count) acc = combine<FOP>(acc, v[q]);
    }
    out[f] = acc;
}

template <int OP>
__global__ void reduce_fold(const double *partial, int count, double *out)
{
    reduce_fold_body<OP>(partial, count, out);
}

static BoxStep box_step(const ReduceBox &b, long long n, long long slots)
{
    if (n == 0) return BoxStep{0, 0, 0};
    const long long q = slots / b.n1;
    return BoxStep{slots % b.n1, q % b.n2, q / b.n2};
}

template <int OP>
static int launch(const ReduceBox &b, double *slots, double *final256, hipStream_t stream)
{
    const long long n = b.n1 * b.n2 * b.n3;
    if (n <= kReduceDirect) {
        hipLaunchKernelGGL(reduce_slots<OP>, dim3(1), dim3(kReduceFinal), 0, stream, b, n, kReduceFinal,
                           box_step(b, n, kReduceFinal), final256);
    } else {
        hipLaunchKernelGGL(reduce_slots<OP>, dim3(kReduceSlots / 256), dim3(256), 0, stream, b, n, kReduceSlots,
                           box_step(b, n, kReduceSlots), slots);
        hipLaunchKernelGGL(reduce_fold<OP>, dim3(1), dim3(kReduceFinal), 0, stream, slots, kReduceSlots, final256);
    }
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "reduction launch failed");
    return FORT_SCOPE_OK;
}

int reduce_launch(int op, const ReduceBox &b, double *slots, double *final256, hipStream_t stream)
{
    if (b.n1 < 0 || b.n2 < 0 || b.n3 < 0) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "negative extent");
    switch (op) {
    case kReduceSum: return launch<kReduceSum>(b, slots, final256, stream);
    case kReduceSumSquares: return launch<kReduceSumSquares>(b, slots, final256, stream);
    case kReduceMax: return launch<kReduceMax>(b, slots, final256, stream);
    case kReduceMin: return launch<kReduceMin>(b, slots, final256, stream);
    case kReduceMaxAbs: return launch<kReduceMaxAbs>(b, slots, final256, stream);
    default: return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unknown reduction");
    }
}

// The host's fold of the 256 values in index order into the initial value, done by one device thread so
// that the result can stay on the device.
template <int OP>
__global__ void reduce_final_fold(const double *final256, double initial, double *result)
{
    constexpr int HOP = OP == kReduceSumSquares ? kReduceSum : (OP == kReduceMaxAbs ? kReduceMax : OP);
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc = initial;
    for (int f = 0; f < kReduceFinal; ++f) acc = combine<HOP>(acc, final256[f]);
    *result = acc;
}

__global__ void reduce_store_initial(double initial, double *result)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *result = initial;
}

template <int OP>
static int launch_to_device(const ReduceBox &b, double *slots, double *final256, double initial, double *result,
                            hipStream_t stream)
{
    FTC_CHECK(launch<OP>(b, slots, final256, stream));
    hipLaunchKernelGGL(reduce_final_fold<OP>, dim3(1), dim3(64), 0, stream, final256, initial, result);
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "reduction launch failed");
    return FORT_SCOPE_OK;
}

int reduce_launch_to_device(int op, const ReduceBox &b, double *slots, double *final256, double initial,
                            double *result, hipStream_t stream)
{
    if (b.n1 < 0 || b.n2 < 0 || b.n3 < 0) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "negative extent");
    switch (op) {
    case kReduceSum: return launch_to_device<kReduceSum>(b, slots, final256, initial, result, stream);
    case kReduceSumSquares: return launch_to_device<kReduceSumSquares>(b, slots, final256, initial, result, stream);
    case kReduceMax: return launch_to_device<kReduceMax>(b, slots, final256, initial, result, stream);
    case kReduceMin: return launch_to_device<kReduceMin>(b, slots, final256, initial, result, stream);
    case kReduceMaxAbs: return launch_to_device<kReduceMaxAbs>(b, slots, final256, initial, result, stream);
    default: return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unknown reduction");
    }
}

double reduce_host_fold(int op, double initial, const double *final256)
{
    double acc = initial;
    for (int f = 0; f < kReduceFinal; ++f) {
        const double x = final256[f];
        switch (op) {
        case kReduceSum:
        case kReduceSumSquares: acc = acc + x; break;
        case kReduceMax:
        case kReduceMaxAbs: acc = acc > x ? acc : x; break;
        default: acc = acc < x ? acc : x; break;
        }
    }
    return acc;
}

// ------------------------------------------------------------------------ the queued form
// The three nests of a lowered reduction as tasks of the context's graph queue, the partial results in a
// registered work array as the reference keeps them (compiler/resident/reductions.py:18-31), and the final
// fold into the variable done on the device, so the value stays there until it is observed.
// No padding bytes: the runtime compares argument bytes to find a graph variant to replay.
struct SlotsTask {
    ReduceBox b;
    long long n;
    BoxStep st;
    double *out;
    int slots, zero;
};
struct FoldTask {
    const double *partial;
    double *out;
    int count, zero;
};
struct FinalTask {
    const double *final256;
    double initial;
    double *result;
};

template <int OP>
__global__ void reduce_slots_task(const SlotsTask a)
{
    reduce_slots_body<OP>(a.b, a.n, a.slots, a.st, a.out);
}

template <int OP>
__global__ void reduce_fold_task(const FoldTask a)
{
    reduce_fold_body<OP>(a.partial, a.count, a.out);
}

template <int OP>
__global__ void reduce_final_task(const FinalTask a)
{
    constexpr int HOP = OP == kReduceSumSquares ? kReduceSum : (OP == kReduceMaxAbs ? kReduceMax : OP);
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc = a.initial;
    for (int f = 0; f < kReduceFinal; ++f) acc = combine<HOP>(acc, a.final256[f]);
    *a.result = acc;
}

// A registered REAL64 work array: at least `count` elements along its first dimension, every other
// dimension of extent 1, so that its first `count` elements are one contiguous section.
static int work_array(fort_scope_t ctx, fort_buffer_t h, std::size_t count, uint32_t *rank)
{
    fort_scope_layout l{};
    FTC_CHECK(fort_scope_layout_get(ctx, h, &l));
    if (l.type != FORT_SCOPE_REAL64 || l.element_bytes != 8 || l.rank < 1 || l.rank > 3)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "reduction work arrays must be REAL64 of rank 1 to 3");
    for (uint32_t d = 1; d < l.rank; ++d)
        if (l.extents[d] != 1)
            return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "reduction work arrays extend along their first dimension");
    if (l.extents[0] < count) return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "reduction work array too small");
    *rank = l.rank;
    return FORT_SCOPE_OK;
}

// Elements [lo, hi) of a work array (its later dimensions have extent 1).
static Sec3 linear(std::size_t lo, std::size_t hi, uint32_t rank)
{
    (void)rank;
    return Sec3{{lo, 0, 0}, {hi, 1, 1}};
}

__global__ void reduce_store_task(const FinalTask a)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *a.result = a.initial;
}

template <int OP>
static int enqueue_reduction(fort_scope_t ctx, fort_buffer_t x, fort_buffer_t partial, fort_buffer_t result,
                             const ReduceBox &b, const Sec3 &box, uint32_t prank, uint32_t rrank, double initial,
                             double *pdev, double *rdev)
{
    const long long n = b.n1 * b.n2 * b.n3;
    const uint64_t key = 0xfd00000000000000ull | (uint64_t)OP;
    const Sec3 slots_sec = linear(0, kReduceSlots, prank), final_sec = linear(kReduceSlots, kReduceSlots + kReduceFinal, prank);
    double *final256 = pdev + kReduceSlots;
    if (n <= kReduceDirect) {
        SlotsTask a{b, n, box_step(b, n, kReduceFinal), final256, kReduceFinal, 0};
        FTC_CHECK(enqueue_kernel(ctx, key | 0x100, (const void *)reduce_slots_task<OP>, dim3(1), dim3(kReduceFinal), 0,
                                 &a, sizeof a, {{x, {box}, {}, {}}, {partial, {}, {}, {final_sec}}}));
    } else {
        SlotsTask a{b, n, box_step(b, n, kReduceSlots), pdev, kReduceSlots, 0};
        FTC_CHECK(enqueue_kernel(ctx, key | 0x200, (const void *)reduce_slots_task<OP>, dim3(kReduceSlots / 256),
                                 dim3(256), 0, &a, sizeof a, {{x, {box}, {}, {}}, {partial, {}, {}, {slots_sec}}}));
        FoldTask f{pdev, final256, kReduceSlots, 0};
        FTC_CHECK(enqueue_kernel(ctx, key | 0x300, (const void *)reduce_fold_task<OP>, dim3(1), dim3(kReduceFinal), 0,
                                 &f, sizeof f, {{partial, {slots_sec}, {}, {final_sec}}}));
    }
    FinalTask fin{final256, initial, rdev};
    return enqueue_kernel(ctx, key | 0x400, (const void *)reduce_final_task<OP>, dim3(1), dim3(64), 0, &fin, sizeof fin,
                          {{partial, {final_sec}, {}, {}}, {result, {}, {}, {linear(0, 1, rrank)}}});
}

// ------------------------------------------------------------------------ the entry that allocates nothing
// Its slot and fold values live in one file-scope device array, `scratch`, that every call on a device
// shares: 65,536 slot values followed by 256 fold values, 526,336 bytes, held by each device the library
// runs on until the process ends. Host code never uses the array's name as an address: the address is
// looked up once per device with hipGetSymbolAddress, with that device current, and kept in
// g_scratch_address.
constexpr std::size_t kStaticScratchDoubles = kReduceSlots + kReduceFinal;
static __device__ __align__(256) double scratch[kStaticScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static double *g_scratch_address[kMaxDevices];

static int static_scratch(double **address)
{
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices)
        return fort_scope_report_error(FORT_SCOPE_RESOURCE, "no usable current device");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return fort_scope_report_error(FORT_SCOPE_RESOURCE, "the reduction scratch has no device address");
        g_scratch_address[device] = static_cast<double *>(p);
    }
    *address = g_scratch_address[device];
    return FORT_SCOPE_OK;
}

}  // namespace ftc

using namespace ftc;

extern "C" {

// Reduces the section [lo, hi) (zero-based, rank 3) of a registered REAL64 array in the fixed order and
// folds it into *value (in: the variable's value before the nest, out: after). op: 0 sum, 1 sum of
// squares, 2 max, 3 min, 4 max of magnitudes. The result is back on the host when the call returns.
int cpp_fort_shared_fixed_order_reduction(fort_scope_t ctx, int mode, fort_buffer_t x, const int *op,
                                          const long long *lo3, const long long *hi3, double *value)
{
    FTC_CHECK(entry_preamble(mode));
    if (!op || !lo3 || !hi3 || !value) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null scalar argument");
    std::size_t e[3];
    FTC_CHECK(layout3(ctx, x, FORT_SCOPE_REAL64, 8, e));
    for (int d = 0; d < 3; ++d)
        if (lo3[d] < 0 || hi3[d] < lo3[d] || (std::size_t)hi3[d] > e[d])
            return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "reduction box outside the array");
    const long long n1 = hi3[0] - lo3[0], n2 = hi3[1] - lo3[1], n3 = hi3[2] - lo3[2];
    if (!n1 || !n2 || !n3) return FORT_SCOPE_OK;
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {x}, dev));
    fort_scope_scratch_lease_v1 lease{};
    FTC_CHECK(fort_scope_scratch_acquire_v1(ctx, (kReduceSlots + kReduceFinal) * sizeof(double), &lease));
    double *slots = static_cast<double *>(lease.device), *final256 = slots + kReduceSlots;
    // The reduction reads the device copy: make it current through a device access of the box, run on
    // the context stream, then bring the 256 values back.
    Sec3 box{{(std::size_t)lo3[0], (std::size_t)lo3[1], (std::size_t)lo3[2]},
             {(std::size_t)hi3[0], (std::size_t)hi3[1], (std::size_t)hi3[2]}};
    fort_scope_section sec{box.lo, box.hi};
    fort_scope_plan_binding bind{};
    bind.buffer = x;
    bind.access.read_count = 1;
    bind.access.reads = &sec;
    void *device = nullptr;
    int status = fort_scope_access_begin_v1(ctx, &bind, 1, 1, &device);
    int prev = 0;
    void *stream = nullptr;
    if (status == FORT_SCOPE_OK) status = fort_scope_gpu_enter(ctx, &prev, &stream);
    double host[kReduceFinal];
    if (status == FORT_SCOPE_OK) {
        ReduceBox b{static_cast<const double *>(device) + lo3[0] + lo3[1] * (long long)e[0] +
                        lo3[2] * (long long)e[0] * (long long)e[1],
                    n1, n2, n3, (long long)e[0], (long long)e[0] * (long long)e[1]};
        status = reduce_launch(*op, b, slots, final256, static_cast<hipStream_t>(stream));
        if (status == FORT_SCOPE_OK &&
            hipMemcpyAsync(host, final256, sizeof host, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)) !=
                hipSuccess)
            status = FORT_SCOPE_EXECUTION;
        if (status == FORT_SCOPE_OK && hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
            status = FORT_SCOPE_EXECUTION;
        (void)fort_scope_gpu_leave(ctx, prev);
        int ended = fort_scope_device_end(ctx, x);
        if (status == FORT_SCOPE_OK) status = ended;
        if (status == FORT_SCOPE_OK) *value = reduce_host_fold(*op, *value, host);
        else if (status == FORT_SCOPE_EXECUTION) (void)fort_scope_execution_error(ctx, "reduction failed after start");
    }
    int released = fort_scope_scratch_release_v1(ctx, lease.token);
    return status ? status : released;
}

// The fixed-order reduction on device pointers, with its partial results in the library's static scratch:
// no allocation, no host synchronisation, the result left in device memory. See ftc_reduce.h.
int ftc_fixed_order_reduction_static_v1(const double *x, long long n1, long long n2, long long n3, long long ld1,
                                        long long ld2, int op, double initial, double *result, void *stream)
{
    if (!x || !result) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null device pointer");
    if (op < kReduceSum || op > kReduceMaxAbs) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unknown reduction");
    if (n1 < 0 || n2 < 0 || n3 < 0) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "negative extent");
    // The partial results need 65,792 values whatever the box, so the scratch bounds no box; the iteration
    // count and the element offsets must fit 64-bit indices.
    const long long limit = 1ll << 62;
    auto beyond = [limit](long long a, long long b) { return a > 0 && b > limit / a; };
    if (beyond(n1, n2) || beyond(n1 * n2, n3) || beyond(n2, ld1) || beyond(n3, ld2))
        return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "reduction box beyond 2^62 elements");
    if (ld1 < n1 || ld2 < ld1 * n2)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "leading dimensions below the box extents");
    double *slots = nullptr;
    FTC_CHECK(static_scratch(&slots));
    double *final256 = slots + kReduceSlots;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!n1 || !n2 || !n3) {
        // No iterations: the variable keeps its value, bit for bit (a fold would turn -0.0 into +0.0).
        hipLaunchKernelGGL(reduce_store_initial, dim3(1), dim3(64), 0, s, initial, result);
        if (hipGetLastError() != hipSuccess)
            return fort_scope_report_error(FORT_SCOPE_EXECUTION, "reduction launch failed");
        return FORT_SCOPE_OK;
    }
    ReduceBox b{x, n1, n2, n3, ld1, ld2};
    return reduce_launch_to_device(op, b, slots, final256, initial, result, s);
}

// The reduction as queued tasks: partial results in `partial` (at least 65,792 elements), the variable's
// new value in result(1) on the device. See ftc_reduce.h.
int cpp_fort_shared_fixed_order_reduction_device(fort_scope_t ctx, int mode, fort_buffer_t x, fort_buffer_t partial,
                                                 fort_buffer_t result, const int *op, const long long *lo3,
                                                 const long long *hi3, const double *initial)
{
    FTC_CHECK(entry_preamble(mode));
    if (!op || !lo3 || !hi3 || !initial) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null scalar argument");
    if (*op < kReduceSum || *op > kReduceMaxAbs) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unknown reduction");
    FTC_CHECK(distinct_outputs({partial, result}, {x}));
    std::size_t e[3];
    FTC_CHECK(layout3(ctx, x, FORT_SCOPE_REAL64, 8, e));
    for (int d = 0; d < 3; ++d)
        if (lo3[d] < 0 || hi3[d] < lo3[d] || (std::size_t)hi3[d] > e[d])
            return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "reduction box outside the array");
    uint32_t prank = 0, rrank = 0;
    FTC_CHECK(work_array(ctx, partial, kReduceSlots + kReduceFinal, &prank));
    FTC_CHECK(work_array(ctx, result, 1, &rrank));
    const long long n1 = hi3[0] - lo3[0], n2 = hi3[1] - lo3[1], n3 = hi3[2] - lo3[2];
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {x, partial, result}, dev));
    const double *base = static_cast<const double *>(dev[0]);
    ReduceBox b{base + lo3[0] + lo3[1] * (long long)e[0] + lo3[2] * (long long)e[0] * (long long)e[1],
                n1 ? n1 : 1, n2 ? n2 : 1, (n1 && n2) ? n3 : 0, (long long)e[0], (long long)e[0] * (long long)e[1]};
    const Sec3 box{{(std::size_t)lo3[0], (std::size_t)lo3[1], (std::size_t)lo3[2]},
                   {(std::size_t)hi3[0], (std::size_t)hi3[1], (std::size_t)hi3[2]}};
    double *pdev = static_cast<double *>(dev[1]), *rdev = static_cast<double *>(dev[2]);
    if (!n1 || !n2 || !n3) {
        // No iterations: the variable keeps its value, bit for bit.
        FinalTask keep{nullptr, *initial, rdev};
        return enqueue_kernel(ctx, 0xfd00000000000500ull, (const void *)reduce_store_task, dim3(1), dim3(64), 0, &keep,
                              sizeof keep, {{result, {}, {}, {linear(0, 1, rrank)}}});
    }
    switch (*op) {
    case kReduceSum: return enqueue_reduction<kReduceSum>(ctx, x, partial, result, b, box, prank, rrank, *initial, pdev, rdev);
    case kReduceSumSquares:
        return enqueue_reduction<kReduceSumSquares>(ctx, x, partial, result, b, box, prank, rrank, *initial, pdev, rdev);
    case kReduceMax: return enqueue_reduction<kReduceMax>(ctx, x, partial, result, b, box, prank, rrank, *initial, pdev, rdev);
    case kReduceMin: return enqueue_reduction<kReduceMin>(ctx, x, partial, result, b, box, prank, rrank, *initial, pdev, rdev);
    default: return enqueue_reduction<kReduceMaxAbs>(ctx, x, partial, result, b, box, prank, rrank, *initial, pdev, rdev);
    }
}

}  // extern "C"
