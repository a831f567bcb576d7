// Entries for the box application's device procedures (compiler/tests/apps/box): the two-layer periodic
// halo (boundaries.f90:6-41), the explicit diffusion step (stepper.f90:8-30) and the store of the new
// interior (stepper.f90:32-36). The arrays are a(-1:n+2, -1:n+2, -1:n+2) of doubles with two ghost layers,
// so the entries take n from the array layout; the step's dx, dt and nu are module variables of the
// application that the spliced call passes.
//
// Numerical contract as everywhere in this library (-ffp-contract=off): the diffusion evaluates
// b = a + dt*nu*lap with lap = (((((a(i-1) + a(i+1)) + a(j-1)) + a(j+1)) + a(k-1)) + a(k+1) - 6*a) / (dx*dx),
// the order the native build of stepper.f90 uses, so b is the native result bit for bit.
#include <hip/hip_runtime.h>

#include "entry.hpp"

#pragma clang fp contract(off)

namespace ftc {

struct BoxArgs {
    double *a;            // the halo fill's array, or the diffusion's input, or the store's target
    double *b;            // the diffusion's output, the store's source
    long long sj, sk;     // strides of j and k (extents n+4)
    double dx, dt, nu;
    int nx, ny, nz, pass; // pass of the halo fill: 0 x, 1 y, 2 z
};

// Linear index of Fortran (i, j, k) of an array with lower bounds -1.
__device__ __forceinline__ long long bidx(const BoxArgs &s, int i, int j, int k)
{
    return (long long)(i + 1) + (long long)(j + 1) * s.sj + (long long)(k + 1) * s.sk;
}

// One face pass of the periodic fill: x copies over the interior rows, y copies over the whole x range,
// z copies over the whole x-y range (boundaries.f90:6-41), so edges and corners follow from the order.
// One thread per (first, second) pair of the face; each thread copies both ghost layers of both sides.
__global__ __launch_bounds__(256) void box_periodic_kernel(const BoxArgs s)
{
    const int nx = s.nx, ny = s.ny, nz = s.nz;
    const long long t = blockIdx.x * 256ll + threadIdx.x;
    if (s.pass == 0) {
        if (t >= (long long)ny * nz) return;
        const int j = 1 + (int)(t % ny), k = 1 + (int)(t / ny);
        for (int i = 1; i <= 2; ++i) {
            s.a[bidx(s, nx + i, j, k)] = s.a[bidx(s, i, j, k)];
            s.a[bidx(s, i - 2, j, k)] = s.a[bidx(s, nx + i - 2, j, k)];
        }
    } else if (s.pass == 1) {
        if (t >= (long long)(nx + 4) * nz) return;
        const int i = -1 + (int)(t % (nx + 4)), k = 1 + (int)(t / (nx + 4));
        for (int j = 1; j <= 2; ++j) {
            s.a[bidx(s, i, ny + j, k)] = s.a[bidx(s, i, j, k)];
            s.a[bidx(s, i, j - 2, k)] = s.a[bidx(s, i, ny + j - 2, k)];
        }
    } else {
        if (t >= (long long)(nx + 4) * (ny + 4)) return;
        const int i = -1 + (int)(t % (nx + 4)), j = -1 + (int)(t / (nx + 4));
        for (int k = 1; k <= 2; ++k) {
            s.a[bidx(s, i, j, nz + k)] = s.a[bidx(s, i, j, k)];
            s.a[bidx(s, i, j, k - 2)] = s.a[bidx(s, i, j, nz + k - 2)];
        }
    }
}

// b = a + dt*nu*laplacian(a) on the interior; one thread per (i, j) column marching k.
__global__ __launch_bounds__(256) void box_diffuse_kernel(const BoxArgs s)
{
    const long long t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= (long long)s.nx * s.ny) return;
    const int i = 1 + (int)(t % s.nx), j = 1 + (int)(t / s.nx);
    const double *a = s.a;
    const double h2 = s.dx * s.dx, c = s.dt * s.nu;
    const long long sj = s.sj, sk = s.sk;
    for (int k = 1; k <= s.nz; ++k) {
        const long long p = bidx(s, i, j, k);
        const double lap = (((((((a[p - 1] + a[p + 1]) + a[p - sj]) + a[p + sj]) + a[p - sk]) + a[p + sk]) -
                             6.0 * a[p])) / h2;
        s.b[p] = a[p] + c * lap;
    }
}

// a(1:n, 1:n, 1:n) = b(1:n, 1:n, 1:n).
__global__ __launch_bounds__(256) void box_store_kernel(const BoxArgs s)
{
    const long long t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= (long long)s.nx * s.ny) return;
    const int i = 1 + (int)(t % s.nx), j = 1 + (int)(t / s.nx);
    for (int k = 1; k <= s.nz; ++k) {
        const long long p = bidx(s, i, j, k);
        s.a[p] = s.b[p];
    }
}

// Zero-based section of Fortran indices [f0, f1] x [g0, g1] x [h0, h1] (lower bounds -1).
static Sec3 fsec(int f0, int f1, int g0, int g1, int h0, int h1)
{
    return Sec3{{(std::size_t)(f0 + 1), (std::size_t)(g0 + 1), (std::size_t)(h0 + 1)},
                {(std::size_t)(f1 + 2), (std::size_t)(g1 + 2), (std::size_t)(h1 + 2)}};
}

// The box layout of an array: rank 3, REAL64, two ghost layers; n >= 2 in every direction (the copies of
// a periodic direction do not meet from two cells up). A smaller box declines, so the procedure runs on
// the host as written.
static int box_layout(fort_scope_t ctx, fort_buffer_t h, int n[3])
{
    std::size_t e[3];
    FTC_CHECK(layout3(ctx, h, FORT_SCOPE_REAL64, 8, e));
    for (int d = 0; d < 3; ++d) {
        if (e[d] < 6) return FORT_SCOPE_DECLINED;
        n[d] = (int)e[d] - 4;
    }
    return FORT_SCOPE_OK;
}

static unsigned blocks(long long threads) { return (unsigned)((threads + 255) / 256); }

}  // namespace ftc

using namespace ftc;

extern "C" {

int cpp_fort_shared_boundaries_periodic(fort_scope_t ctx, int mode, fort_buffer_t a)
{
    FTC_CHECK(entry_preamble(mode));
    int n[3];
    FTC_CHECK(box_layout(ctx, a, n));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {a}, dev));
    const int nx = n[0], ny = n[1], nz = n[2];
    BoxArgs s{(double *)dev[0], nullptr, (long long)nx + 4, ((long long)nx + 4) * ((long long)ny + 4), 0, 0, 0,
              nx, ny, nz, 0};
    // x: the first and last two interior columns of the interior rows into the ghost columns
    std::vector<Binding3> bx{{a, {fsec(1, 2, 1, ny, 1, nz), fsec(nx - 1, nx, 1, ny, 1, nz)}, {},
                              {fsec(nx + 1, nx + 2, 1, ny, 1, nz), fsec(-1, 0, 1, ny, 1, nz)}}};
    FTC_CHECK(enqueue_kernel(ctx, 0xb0c0000000000001ull, (const void *)box_periodic_kernel, dim3(blocks((long long)ny * nz)),
                             dim3(256), 0, &s, sizeof s, bx));
    // y: the first and last two rows, over the whole x range, into the ghost rows
    s.pass = 1;
    std::vector<Binding3> by{{a, {fsec(-1, nx + 2, 1, 2, 1, nz), fsec(-1, nx + 2, ny - 1, ny, 1, nz)}, {},
                              {fsec(-1, nx + 2, ny + 1, ny + 2, 1, nz), fsec(-1, nx + 2, -1, 0, 1, nz)}}};
    FTC_CHECK(enqueue_kernel(ctx, 0xb0c0000000000001ull, (const void *)box_periodic_kernel,
                             dim3(blocks((long long)(nx + 4) * nz)), dim3(256), 0, &s, sizeof s, by));
    // z: the first and last two planes, over the whole x-y range, into the ghost planes
    s.pass = 2;
    std::vector<Binding3> bz{{a, {fsec(-1, nx + 2, -1, ny + 2, 1, 2), fsec(-1, nx + 2, -1, ny + 2, nz - 1, nz)}, {},
                              {fsec(-1, nx + 2, -1, ny + 2, nz + 1, nz + 2), fsec(-1, nx + 2, -1, ny + 2, -1, 0)}}};
    return enqueue_kernel(ctx, 0xb0c0000000000001ull, (const void *)box_periodic_kernel,
                          dim3(blocks((long long)(nx + 4) * (ny + 4))), dim3(256), 0, &s, sizeof s, bz);
}

int cpp_fort_shared_stepper_diffuse(fort_scope_t ctx, int mode, fort_buffer_t a, fort_buffer_t b, const double *dx,
                                    const double *dt, const double *nu)
{
    FTC_CHECK(entry_preamble(mode));
    if (!dx || !dt || !nu) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null shared scalar argument");
    int n[3], m[3];
    FTC_CHECK(box_layout(ctx, a, n));
    FTC_CHECK(box_layout(ctx, b, m));
    if (n[0] != m[0] || n[1] != m[1] || n[2] != m[2])
        return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "diffusion arrays of different shapes");
    FTC_CHECK(distinct_outputs({b}, {a}));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {a, b}, dev));
    const int nx = n[0], ny = n[1], nz = n[2];
    BoxArgs s{(double *)dev[0], (double *)dev[1], (long long)nx + 4, ((long long)nx + 4) * ((long long)ny + 4), *dx,
              *dt, *nu, nx, ny, nz, 0};
    std::vector<Binding3> bind{{a, {fsec(0, nx + 1, 0, ny + 1, 0, nz + 1)}, {}, {}},
                               {b, {}, {}, {fsec(1, nx, 1, ny, 1, nz)}}};
    return enqueue_kernel(ctx, 0xb0c0000000000002ull, (const void *)box_diffuse_kernel, dim3(blocks((long long)nx * ny)),
                          dim3(256), 0, &s, sizeof s, bind);
}

int cpp_fort_shared_stepper_store(fort_scope_t ctx, int mode, fort_buffer_t a, fort_buffer_t b)
{
    FTC_CHECK(entry_preamble(mode));
    int n[3], m[3];
    FTC_CHECK(box_layout(ctx, a, n));
    FTC_CHECK(box_layout(ctx, b, m));
    if (n[0] != m[0] || n[1] != m[1] || n[2] != m[2])
        return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "store arrays of different shapes");
    FTC_CHECK(distinct_outputs({a}, {b}));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {a, b}, dev));
    const int nx = n[0], ny = n[1], nz = n[2];
    BoxArgs s{(double *)dev[0], (double *)dev[1], (long long)nx + 4, ((long long)nx + 4) * ((long long)ny + 4), 0, 0, 0,
              nx, ny, nz, 0};
    const Sec3 inner = fsec(1, nx, 1, ny, 1, nz);
    std::vector<Binding3> bind{{a, {}, {}, {inner}}, {b, {inner}, {}, {}}};
    return enqueue_kernel(ctx, 0xb0c0000000000003ull, (const void *)box_store_kernel, dim3(blocks((long long)nx * ny)),
                          dim3(256), 0, &s, sizeof s, bind);
}

}  // extern "C"
