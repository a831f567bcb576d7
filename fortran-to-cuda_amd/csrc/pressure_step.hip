// The nests around the pressure solve of one domain, as entries: ELMM's prepoisson (the divergence of
// the provisional velocity into the solver's right-hand side), postpoisson (the projection
// u -= dt grad(phi) with the halos) and fixpressuretoreference (phi shifted by its reference-plane mean,
// the mean being a fixed-order reduction), benchmarks/elmm/elmm.toml:212-225, 274-293. Their sources are absent from the
// reference, so the arithmetic is this repository's definition, the one oracle/step.c (orc_divergence,
// orc_projection) states and the ensemble's batched kernels use: the device results equal the oracle's
// bit for bit.
//
// Layouts are ELMM's (benchmarks/elmm/sources/gpu_pressure.h:8-12): U, V, W and RHS have one ghost layer
// (extents n+2), Phi two (extents n+4), column-major.
#include <hip/hip_runtime.h>

#include "entry.hpp"
#include "les_kernels.hpp"

namespace ftc {

struct PressureStepArgs {
    LesArgs g;                 // the domain (nx, ny, nz, strides of the n+2 arrays) for the halo images
    double *U, *V, *W;         // projection: updated in place
    const double *CU, *CV, *CW;  // divergence: read
    double *rhs;
    const double *phi;
    long long pj, pk;          // strides of phi (extents n+4)
    double rdx, rdy, rdz, rdt;  // 1/dx, 1/dy, 1/dz, 1/dt
    double gx, gy, gz;          // dt/dx, dt/dy, dt/dz
};

constexpr int kPTX = 64, kPTY = 4;

// rhs(a, b, c) = ((dU/dx + dV/dy) + dW/dz) / dt on interior cells, backward differences of the face
// velocities, in that order of operations (oracle/step.c:85-99). RHS shares the velocities' extents, so
// both use the linear index p.
__global__ __launch_bounds__(kPTX *kPTY) void pressure_divergence_kernel(const PressureStepArgs s)
{
    const int a = 1 + blockIdx.x * kPTX + threadIdx.x, b = 1 + blockIdx.y * kPTY + threadIdx.y;
    if (a > s.g.nx || b > s.g.ny) return;
    const long long sj = s.g.sj, sk = s.g.sk;
    for (int c = 1; c <= s.g.nz; ++c) {
        const long long p = a + b * sj + c * sk;
        s.rhs[p] = (((s.CU[p] - s.CU[p - 1]) * s.rdx + (s.CV[p] - s.CV[p - sj]) * s.rdy) +
                    (s.CW[p] - s.CW[p - sk]) * s.rdz) * s.rdt;
    }
}

// u -= dt/dx (phi(a+1) - phi(a)) with phi periodic in x and y (the solution's own cells, not its ghosts),
// likewise v; w on the interior faces c < nz; then the halo images (oracle/step.c:101-122).
__global__ __launch_bounds__(kPTX *kPTY) void pressure_projection_kernel(const PressureStepArgs s)
{
    const int a = 1 + blockIdx.x * kPTX + threadIdx.x, b = 1 + blockIdx.y * kPTY + threadIdx.y;
    const int nx = s.g.nx, ny = s.g.ny, nz = s.g.nz;
    if (a > nx || b > ny) return;
    const long long sj = s.g.sj, sk = s.g.sk, pj = s.pj, pk = s.pk;
    const ColumnImages ci = column_images(s.g, a + 1, b + 1);
    // phi's interior cell (a, b, c) (1-based) sits at (a + 1, b + 1, c + 1) of the array with two ghosts.
    const long long f0 = (a + 1) + (b + 1) * pj;
    const long long fx = ((a == nx ? 0 : a) + 2) + (b + 1) * pj;  // (a % nx) + 1, row b
    const long long fy = (a + 1) + ((b == ny ? 0 : b) + 2) * pj;  // column a, row (b % ny) + 1
    for (int c = 1; c <= nz; ++c) {
        const long long p = a + b * sj + c * sk;
        const long long o = (c + 1) * pk;
        const double f = s.phi[f0 + o];
        const PlaneImages zi = plane_images(s.g, c + 1);
        put_column(s.U, p, s.U[p] - s.gx * (s.phi[fx + o] - f), ci, zi, 0);
        put_column(s.V, p, s.V[p] - s.gy * (s.phi[fy + o] - f), ci, zi, 0);
        if (c < nz) put_column(s.W, p, s.W[p] - s.gz * (s.phi[f0 + o + pk] - f), ci, zi, 1);
    }
}

// phi -= ref on the interior cells, the ones the solve defines (rank-3 array with two ghost layers:
// cells 2 .. n+1 of each dimension, zero-based). The ghost layers keep what they held; nothing on the
// step reads them (the projection takes its x and y neighbours from the interior by periodicity), and
// a halo fill of phi refills them from the shifted interior. With ref_sum set, ref is ref_sum[0] / divisor,
// read on the device (the reference-plane sum a queued reduction left there).
struct ShiftArgs {  // no padding bytes: the runtime compares argument bytes to pick a graph variant
    double *phi;
    const double *ref_sum;
    long long pj, pk;  // strides of phi
    double ref, divisor;
    int nx, ny, nz, zero;
};

__global__ __launch_bounds__(256) void pressure_shift_kernel(const ShiftArgs a)
{
    const double ref = a.ref_sum ? a.ref_sum[0] / a.divisor : a.ref;
    const long long cells = (long long)a.nx * a.ny * a.nz;
    for (long long t = blockIdx.x * 256ll + threadIdx.x; t < cells; t += (long long)gridDim.x * 256) {
        const long long i = t % a.nx, r = t / a.nx;
        const long long p = (i + 2) + (r % a.ny + 2) * a.pj + (r / a.ny + 2) * a.pk;
        a.phi[p] = a.phi[p] - ref;
    }
}

static Sec3 pbox(std::size_t i0, std::size_t i1, std::size_t j0, std::size_t j1, std::size_t k0, std::size_t k1)
{
    return Sec3{{i0, j0, k0}, {i1, j1, k1}};
}

static int domain_args(fort_scope_t ctx, const fort_buffer_t *fields, int count, const double *dx, const double *dy,
                       const double *dz, const double *dt, const int *pnx, const int *pny, const int *pnz,
                       PressureStepArgs &s)
{
    if (!dx || !dy || !dz || !dt || !pnx || !pny || !pnz)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null shared scalar argument");
    const int nx = *pnx, ny = *pny, nz = *pnz;
    if (nx <= 0 || ny <= 0 || nz <= 0) return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "empty pressure domain");
    for (int n = 0; n < count; ++n) {
        std::size_t e[3];
        FTC_CHECK(layout3(ctx, fields[n], FORT_SCOPE_REAL64, 8, e));
        if (e[0] != (std::size_t)nx + 2 || e[1] != (std::size_t)ny + 2 || e[2] != (std::size_t)nz + 2)
            return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "pressure step arrays must have one ghost layer");
    }
    s = PressureStepArgs{};
    s.g.nx = nx;
    s.g.ny = ny;
    s.g.nz = nz;
    s.g.kw = nz;
    s.g.sj = (long long)nx + 2;
    s.g.sk = ((long long)nx + 2) * ((long long)ny + 2);
    s.pj = (long long)nx + 4;
    s.pk = ((long long)nx + 4) * ((long long)ny + 4);
    s.rdx = 1.0 / *dx;
    s.rdy = 1.0 / *dy;
    s.rdz = 1.0 / *dz;
    s.rdt = 1.0 / *dt;
    s.gx = *dt / *dx;
    s.gy = *dt / *dy;
    s.gz = *dt / *dz;
    return FORT_SCOPE_OK;
}

static int fix_reference(fort_scope_t ctx, int mode, fort_buffer_t phi, fort_buffer_t ref_sum, const double *phi_ref,
                         const int *nx, const int *ny, const int *nz)
{
    FTC_CHECK(entry_preamble(mode));
    if ((!ref_sum && !phi_ref) || !nx || !ny || !nz)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null shared scalar argument");
    if (*nx <= 0 || *ny <= 0 || *nz <= 0) return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "empty pressure domain");
    std::size_t e[3];
    FTC_CHECK(layout3(ctx, phi, FORT_SCOPE_REAL64, 8, e));
    if (e[0] != (std::size_t)*nx + 4 || e[1] != (std::size_t)*ny + 4 || e[2] != (std::size_t)*nz + 4)
        return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "phi must have two ghost layers");
    if (ref_sum) {
        fort_scope_layout l{};
        FTC_CHECK(fort_scope_layout_get(ctx, ref_sum, &l));
        if (l.type != FORT_SCOPE_REAL64 || l.element_bytes != 8 || l.rank < 1 || l.extents[0] < 1)
            return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "the reference sum is a REAL64 array");
        FTC_CHECK(distinct_outputs({phi}, {ref_sum}));
    }
    std::vector<void *> dev;
    FTC_CHECK(ref_sum ? reserve(ctx, {phi, ref_sum}, dev) : reserve(ctx, {phi}, dev));
    ShiftArgs args{(double *)dev[0], ref_sum ? (const double *)dev[1] : nullptr, (long long)e[0],
                   (long long)(e[0] * e[1]), phi_ref ? *phi_ref : 0.0, (double)*nx * (double)*ny, *nx, *ny, *nz, 0};
    const Sec3 inner = pbox(2, (std::size_t)*nx + 2, 2, (std::size_t)*ny + 2, 2, (std::size_t)*nz + 2);
    std::vector<Binding3> b{{phi, {inner}, {}, {inner}}};
    if (ref_sum) b.push_back({ref_sum, {Sec3{{0, 0, 0}, {1, 1, 1}}}, {}, {}});
    const long long cells = (long long)*nx * *ny * *nz;
    const unsigned blocks = (unsigned)std::min<long long>(4096, (cells + 255) / 256);
    return enqueue_kernel(ctx, ref_sum ? 0xd1e0000000000004ull : 0xd1e0000000000003ull,
                          (const void *)pressure_shift_kernel, dim3(blocks), dim3(256), 0, &args, sizeof args, b);
}

}  // namespace ftc

using namespace ftc;

extern "C" {

int cpp_fort_shared_pressure_divergence(fort_scope_t ctx, int mode, fort_buffer_t rhs, fort_buffer_t u,
                                        fort_buffer_t v, fort_buffer_t w, const double *dx, const double *dy,
                                        const double *dz, const double *dt, const int *nx, const int *ny,
                                        const int *nz)
{
    FTC_CHECK(entry_preamble(mode));
    PressureStepArgs s;
    const fort_buffer_t all[4] = {rhs, u, v, w};
    FTC_CHECK(domain_args(ctx, all, 4, dx, dy, dz, dt, nx, ny, nz, s));
    FTC_CHECK(distinct_outputs({rhs}, {u, v, w}));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {rhs, u, v, w}, dev));
    s.rhs = (double *)dev[0];
    s.CU = (const double *)dev[1];
    s.CV = (const double *)dev[2];
    s.CW = (const double *)dev[3];
    const std::size_t X = s.g.nx, Y = s.g.ny, Z = s.g.nz;
    std::vector<Binding3> b(4);
    b[0] = {rhs, {}, {}, {pbox(1, X + 1, 1, Y + 1, 1, Z + 1)}};
    b[1] = {u, {pbox(0, X + 1, 1, Y + 1, 1, Z + 1)}, {}, {}};
    b[2] = {v, {pbox(1, X + 1, 0, Y + 1, 1, Z + 1)}, {}, {}};
    b[3] = {w, {pbox(1, X + 1, 1, Y + 1, 0, Z + 1)}, {}, {}};
    const dim3 grid((unsigned)((X + kPTX - 1) / kPTX), (unsigned)((Y + kPTY - 1) / kPTY));
    return enqueue_kernel(ctx, 0xd1e0000000000001ull, (const void *)pressure_divergence_kernel, grid,
                          dim3(kPTX, kPTY), 0, &s, sizeof s, b);
}

int cpp_fort_shared_pressure_projection(fort_scope_t ctx, int mode, fort_buffer_t u, fort_buffer_t v,
                                        fort_buffer_t w, fort_buffer_t phi, const double *dx, const double *dy,
                                        const double *dz, const double *dt, const int *nx, const int *ny,
                                        const int *nz)
{
    FTC_CHECK(entry_preamble(mode));
    PressureStepArgs s;
    const fort_buffer_t vel[3] = {u, v, w};
    FTC_CHECK(domain_args(ctx, vel, 3, dx, dy, dz, dt, nx, ny, nz, s));
    std::size_t e[3];
    FTC_CHECK(layout3(ctx, phi, FORT_SCOPE_REAL64, 8, e));
    if (e[0] != (std::size_t)s.g.nx + 4 || e[1] != (std::size_t)s.g.ny + 4 || e[2] != (std::size_t)s.g.nz + 4)
        return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "phi must have two ghost layers");
    FTC_CHECK(distinct_outputs({u, v, w}, {phi}));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {u, v, w, phi}, dev));
    s.U = (double *)dev[0];
    s.V = (double *)dev[1];
    s.W = (double *)dev[2];
    s.phi = (const double *)dev[3];
    const std::size_t X = s.g.nx, Y = s.g.ny, Z = s.g.nz;
    const Sec3 interior = pbox(1, X + 1, 1, Y + 1, 1, Z + 1), full = pbox(0, X + 2, 0, Y + 2, 0, Z + 2);
    std::vector<Binding3> b(4);
    // Each thread reads its interior cells and writes them and their ghost images: the whole array for
    // U and V, and for W the faces 1..nz-1 with their x and y images (the walls are not written).
    b[0] = {u, {interior}, {}, {full}};
    b[1] = {v, {interior}, {}, {full}};
    b[2] = {w, {pbox(1, X + 1, 1, Y + 1, 1, Z)}, {}, {pbox(0, X + 2, 0, Y + 2, 1, Z)}};
    b[3] = {phi, {pbox(2, X + 2, 2, Y + 2, 2, Z + 2)}, {}, {}};
    const dim3 grid((unsigned)((X + kPTX - 1) / kPTX), (unsigned)((Y + kPTY - 1) / kPTY));
    return enqueue_kernel(ctx, 0xd1e0000000000002ull, (const void *)pressure_projection_kernel, grid,
                          dim3(kPTX, kPTY), 0, &s, sizeof s, b);
}

int cpp_fort_shared_pressure_projection_bc(fort_scope_t ctx, int mode, fort_buffer_t u, fort_buffer_t v,
                                           fort_buffer_t w, fort_buffer_t phi, const double *dx, const double *dy,
                                           const double *dz, const double *dt, const int *nx, const int *ny,
                                           const int *nz, const int *bc_x, const int *bc_y)
{
    FTC_CHECK(entry_preamble(mode));
    if (!bc_x || !bc_y) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null shared scalar argument");
    if (*bc_x != 0 || *bc_y != 0)
        return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "the projection implements periodic x and y only");
    return cpp_fort_shared_pressure_projection(ctx, mode, u, v, w, phi, dx, dy, dz, dt, nx, ny, nz);
}

int cpp_fort_shared_pressure_fix_reference(fort_scope_t ctx, int mode, fort_buffer_t phi, const double *phi_ref,
                                           const int *nx, const int *ny, const int *nz)
{
    return fix_reference(ctx, mode, phi, 0, phi_ref, nx, ny, nz);
}

int cpp_fort_shared_pressure_fix_reference_sum(fort_scope_t ctx, int mode, fort_buffer_t phi, fort_buffer_t ref_sum,
                                               const int *nx, const int *ny, const int *nz)
{
    return fix_reference(ctx, mode, phi, ref_sum, nullptr, nx, ny, nz);
}

}  // extern "C"
