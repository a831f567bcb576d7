// Entries for the loop nests of the reference's relax application (compiler/tests/apps/relax/relax.f90):
// the Jacobi sweep (relax.f90:29-35), the largest change of a sweep (:36-43) and the copy of the new
// interior (:44). The arrays are u(0:n+1, 0:n+1) and v(0:n+1, 0:n+1) of doubles; n is the program's.
//
// Numerical contract as everywhere in this library (-ffp-contract=off): the sweep evaluates
// v = 0.25 * (((u(i-1,j) + u(i+1,j)) + u(i,j-1)) + u(i,j+1)), the native build's order, so v is the native
// result bit for bit. The change is a maximum of |v - u|, which is the same value in any order of the cells
// (for fields without NaNs), so the device folds it in its own order.
#include <hip/hip_runtime.h>

#include "entry.hpp"

#pragma clang fp contract(off)

namespace ftc {

struct RelaxArgs {
    double *out;       // the sweep's v, the copy's u
    const double *in;  // the sweep's u, the copy's v
    const double *other;
    double *partial;   // the change's 256 block maxima
    long long ld;      // leading extent of both arrays
    long long origin;  // linear index of (0, 0)
    int n, pad;
};

// v(i, j) for 1 <= i, j <= n; one thread per cell, i fastest.
__global__ __launch_bounds__(256) void relax_sweep_kernel(const RelaxArgs s)
{
    const long long t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= (long long)s.n * s.n) return;
    const long long i = 1 + t % s.n, j = 1 + t / s.n;
    const long long p = s.origin + i + j * s.ld;
    const double *u = s.in;
    s.out[p] = 0.25 * (((u[p - 1] + u[p + 1]) + u[p - s.ld]) + u[p + s.ld]);
}

// u(1:n, 1:n) = v(1:n, 1:n).
__global__ __launch_bounds__(256) void relax_copy_kernel(const RelaxArgs s)
{
    const long long t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= (long long)s.n * s.n) return;
    const long long p = s.origin + (1 + t % s.n) + (1 + t / s.n) * s.ld;
    s.out[p] = s.in[p];
}

constexpr int kRelaxBlocks = 256;

// The largest |v - u| of the interior: each of 256 workgroups strides over the cells and leaves its maximum.
__global__ __launch_bounds__(256) void relax_change_kernel(const RelaxArgs s)
{
    __shared__ double best[256];
    const long long cells = (long long)s.n * s.n;
    double acc = 0.0;  // every |v - u| is >= +0, so +0 leaves the maximum of a nonempty set unchanged
    for (long long t = blockIdx.x * 256ll + threadIdx.x; t < cells; t += 256ll * kRelaxBlocks) {
        const long long p = s.origin + (1 + t % s.n) + (1 + t / s.n) * s.ld;
        const double d = fabs(s.in[p] - s.other[p]);
        acc = d > acc ? d : acc;
    }
    best[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = best[threadIdx.x + w];
            best[threadIdx.x] = o > best[threadIdx.x] ? o : best[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) s.partial[blockIdx.x] = best[0];
}

// The relax layout of an array: rank 2, REAL64, covering (0:n+1, 0:n+1). Its linear index of (0, 0) and its
// leading extent come back; an array that does not cover that range declines.
static int relax_layout(fort_scope_t ctx, fort_buffer_t h, int n, long long &ld, long long &origin,
                        std::size_t ext[2])
{
    fort_scope_layout l{};
    FTC_CHECK(fort_scope_layout_get(ctx, h, &l));
    if (l.rank != 2 || l.type != FORT_SCOPE_REAL64 || l.element_bytes != 8)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "entry layout/type mismatch");
    for (int d = 0; d < 2; ++d) {
        ext[d] = l.extents[d];
        const long long lo = l.lower_bounds ? l.lower_bounds[d] : 1;
        if (lo > 0 || lo + (long long)ext[d] - 1 < (long long)n + 1)
            return FORT_SCOPE_DECLINED;  // the original statements take any other shape
    }
    ld = (long long)ext[0];
    const long long lo0 = l.lower_bounds ? l.lower_bounds[0] : 1, lo1 = l.lower_bounds ? l.lower_bounds[1] : 1;
    origin = -lo0 - lo1 * ld;
    return FORT_SCOPE_OK;
}

// Zero-based section of the cells [i0, i1] x [j0, j1] of an array whose (0, 0) is at offsets (o0, o1).
static Sec3 rsec(long long o0, long long o1, long long i0, long long i1, long long j0, long long j1)
{
    return Sec3{{(std::size_t)(o0 + i0), (std::size_t)(o1 + j0), 0},
                {(std::size_t)(o0 + i1 + 1), (std::size_t)(o1 + j1 + 1), 1}};
}

// The pair of arrays of a relax entry, with the same layout; n >= 1.
static int relax_pair(fort_scope_t ctx, fort_buffer_t a, fort_buffer_t b, const int *n, long long &ld,
                      long long &origin, long long off[2])
{
    if (!n) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null shared scalar argument");
    if (*n < 1) return FORT_SCOPE_DECLINED;
    std::size_t ea[2], eb[2];
    long long lda, oa, ldb, ob;
    FTC_CHECK(relax_layout(ctx, a, *n, lda, oa, ea));
    FTC_CHECK(relax_layout(ctx, b, *n, ldb, ob, eb));
    if (ea[0] != eb[0] || ea[1] != eb[1] || oa != ob)
        return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "relax arrays of different layouts");
    ld = lda;
    origin = oa;
    off[0] = origin % ld;  // zero-based offsets of index 0 in each dimension
    off[1] = origin / ld;
    return FORT_SCOPE_OK;
}

static unsigned relax_blocks(int n) { return (unsigned)(((long long)n * n + 255) / 256); }

}  // namespace ftc

using namespace ftc;

extern "C" {

int cpp_fort_shared_relax_sweep(fort_scope_t ctx, int mode, fort_buffer_t v, fort_buffer_t u, const int *n)
{
    FTC_CHECK(entry_preamble(mode));
    long long ld, origin, o[2];
    FTC_CHECK(relax_pair(ctx, v, u, n, ld, origin, o));
    FTC_CHECK(distinct_outputs({v}, {u}));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {v, u}, dev));
    const int m = *n;
    RelaxArgs s{(double *)dev[0], (const double *)dev[1], nullptr, nullptr, ld, origin, m, 0};
    // u: the interior rows with both side columns, and the bottom and top rows of the interior columns.
    std::vector<Binding3> bind{{v, {}, {}, {rsec(o[0], o[1], 1, m, 1, m)}},
                               {u, {rsec(o[0], o[1], 0, m + 1, 1, m), rsec(o[0], o[1], 1, m, 0, 0),
                                    rsec(o[0], o[1], 1, m, m + 1, m + 1)}, {}, {}}};
    return enqueue_kernel(ctx, 0x7e1a000000000001ull, (const void *)relax_sweep_kernel, dim3(relax_blocks(m)),
                          dim3(256), 0, &s, sizeof s, bind);
}

int cpp_fort_shared_relax_copy(fort_scope_t ctx, int mode, fort_buffer_t u, fort_buffer_t v, const int *n)
{
    FTC_CHECK(entry_preamble(mode));
    long long ld, origin, o[2];
    FTC_CHECK(relax_pair(ctx, u, v, n, ld, origin, o));
    FTC_CHECK(distinct_outputs({u}, {v}));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {u, v}, dev));
    const int m = *n;
    RelaxArgs s{(double *)dev[0], (const double *)dev[1], nullptr, nullptr, ld, origin, m, 0};
    const Sec3 inner = rsec(o[0], o[1], 1, m, 1, m);
    std::vector<Binding3> bind{{u, {}, {}, {inner}}, {v, {inner}, {}, {}}};
    return enqueue_kernel(ctx, 0x7e1a000000000002ull, (const void *)relax_copy_kernel, dim3(relax_blocks(m)),
                          dim3(256), 0, &s, sizeof s, bind);
}

// change = max(change, max over the interior of |v - u|): the queued work that writes v or u runs first, then
// the 256 block maxima come back and the host folds them into change.
int cpp_fort_shared_relax_change(fort_scope_t ctx, int mode, fort_buffer_t v, fort_buffer_t u, const int *n,
                                 double *change)
{
    FTC_CHECK(entry_preamble(mode));
    if (!change) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null shared scalar argument");
    long long ld, origin, o[2];
    FTC_CHECK(relax_pair(ctx, v, u, n, ld, origin, o));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {v, u}, dev));
    const int m = *n;
    fort_scope_scratch_lease_v1 lease{};
    FTC_CHECK(fort_scope_scratch_acquire_v1(ctx, kRelaxBlocks * sizeof(double), &lease));
    const Sec3 inner = rsec(o[0], o[1], 1, m, 1, m);
    fort_scope_section sec{inner.lo, inner.hi};
    fort_scope_plan_binding bind[2]{};
    bind[0].buffer = v;
    bind[1].buffer = u;
    for (auto &b : bind) {
        b.access.read_count = 1;
        b.access.reads = &sec;
    }
    void *device[2] = {nullptr, nullptr};
    int status = fort_scope_access_begin_v1(ctx, bind, 2, 1, device);
    if (status != FORT_SCOPE_OK) {
        (void)fort_scope_scratch_release_v1(ctx, lease.token);
        return status;
    }
    int prev = 0;
    void *stream = nullptr;
    double host[kRelaxBlocks];
    status = fort_scope_gpu_enter(ctx, &prev, &stream);
    if (status == FORT_SCOPE_OK) {
        hipStream_t st = static_cast<hipStream_t>(stream);
        RelaxArgs s{nullptr, (const double *)device[0], (const double *)device[1], (double *)lease.device, ld, origin,
                    m, 0};
        hipLaunchKernelGGL(relax_change_kernel, dim3(kRelaxBlocks), dim3(256), 0, st, s);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(host, lease.device, sizeof host, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            status = FORT_SCOPE_EXECUTION;
        (void)fort_scope_gpu_leave(ctx, prev);
    }
    const int ended_v = fort_scope_device_end(ctx, v), ended_u = fort_scope_device_end(ctx, u);
    if (status == FORT_SCOPE_OK) status = ended_v ? ended_v : ended_u;
    if (status == FORT_SCOPE_OK) {
        double acc = *change;
        for (double h : host) acc = h > acc ? h : acc;
        *change = acc;
    } else if (status == FORT_SCOPE_EXECUTION) {
        (void)fort_scope_execution_error(ctx, "relax change failed after start");
    }
    const int released = fort_scope_scratch_release_v1(ctx, lease.token);
    return status ? status : released;
}

}  // extern "C"
