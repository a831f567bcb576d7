// Staggered-grid LES stencils for gfx950: momentum advection (CDU, CDV, CDW), flux-form scalar advection,
// Smagorinsky viscosity and eddy-viscosity diffusion, fused into one k-marching kernel, and the C-ABI
// entries that submit it through the fort_scope queue.
//
// Numerical contract: this file is compiled with -ffp-contract=off and without fast-math, and every
// expression below keeps the operand order of the Fortran nest it implements, so each output is the same
// IEEE-754 binary64 result as the native build of those nests
// (benchmarks/cdu/src/cdu.f90:48-107, compiler/tests/fixtures/stencils/{cvd,cdw}.f90:44-98).
//
// Data layout in HBM: every field is one column-major (nx+2, ny+2, nz+2) array of doubles with one ghost
// layer, exactly the host layout the Fortran caller registered; i is the fastest index.
//
// Kernel shape: a workgroup owns a TX x TY tile of columns and marches a chunk of k planes. Each lane
// computes one (i, j) point per plane, so every wavefront reads 64 consecutive doubles of a row
// (coalesced 512-byte rows). The Smagorinsky viscosity of the planes k-1, k, k+1 over the tile plus a
// one-cell ring lives in an LDS ring buffer of three planes: each plane is computed once per chunk
// instead of once per consuming stencil point.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>

#include "entry.hpp"
#include "les_kernels.hpp"

#pragma clang fp contract(off)

namespace ftc {

// ------------------------------------------------------------------------------------------------ host
static void fill_coefficients(LesArgs &a, double dx, double dy, double dz, double cs)
{
    const double d[3] = {dx, dy, dz};
    for (int c = 0; c < 3; ++c) {
        a.c25[c] = 0.25 / d[c];
        a.c50[c] = 0.5 / d[c];
        a.c125[c] = 0.125 / d[c];
    }
    const double delta = std::cbrt(dx * dy * dz);
    a.sgs[0] = 1.0 / dx;
    a.sgs[1] = 1.0 / dy;
    a.sgs[2] = 1.0 / dz;
    a.sgs[3] = 0.25 / dx;
    a.sgs[4] = 0.25 / dy;
    a.sgs[5] = 0.25 / dz;
    a.sgs[6] = 1.0 / (dx * dx);
    a.sgs[7] = 1.0 / (dy * dy);
    a.sgs[8] = 1.0 / (dz * dz);
    a.sgs[9] = (cs * delta) * (cs * delta);
}

// Launch geometry: enough k chunks that the grid holds at least ~8 workgroups per CU of a 256-CU
// device, without chunks shorter than 4 planes (each chunk recomputes two viscosity planes).
// The k chunk of a launch over tiles of TX x (TY*rows) columns: about four rounds of the resident
// workgroups (`resident` per CU on 256 CUs).
static int choose_kchunk(int nx, int ny, int kplanes, int rows = 1, int resident = 2)
{
    const long long tiles = (long long)((nx + TX - 1) / TX) * ((ny + TY * rows - 1) / (TY * rows));
    const long long target = 4LL * 256 * resident;
    int chunks = (int)std::max<long long>(1, (target + tiles - 1) / tiles);
    int kchunk = (kplanes + chunks - 1) / chunks;
    if (const char *v = std::getenv("FTC_LES_KCHUNK")) kchunk = std::atoi(v);
    return std::max(std::min(kchunk, kplanes), std::min(4, kplanes));
}

// Compute units of the current device (256 on an MI355X), looked up once per device.
static int compute_units()
{
    static int cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        int n = 0;
        cached[dev] = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0 ? n : 256;
    }
    return cached[dev];
}

template <int MASK, bool SGS, bool STEP>
static const void *kernel_of()
{
    return reinterpret_cast<const void *>(&les_kernel<MASK, SGS, STEP>);
}

static dim3 grid_of(const LesArgs &a, int kplanes, int rows = 1, int width = TX)
{
    const int ty = NTH / width * rows;
    return dim3((a.nx + width - 1) / width, (a.ny + ty - 1) / ty, (kplanes + a.kchunk - 1) / a.kchunk);
}

template <bool STEP>
static const void *staged_kernel(int rows, int width, bool iso, bool two)
{
    if (rows == 2) return reinterpret_cast<const void *>(&les_sgs3_kernel<STEP, 2>);
    if (width == 32 && two)
        return iso ? reinterpret_cast<const void *>(&les_sgs3_kernel<STEP, 1, 32, true, true>)
                   : reinterpret_cast<const void *>(&les_sgs3_kernel<STEP, 1, 32, false, true>);
    if (width == 32)
        return iso ? reinterpret_cast<const void *>(&les_sgs3_kernel<STEP, 1, 32, true>)
                   : reinterpret_cast<const void *>(&les_sgs3_kernel<STEP, 1, 32>);
    if (width == 128)
        return iso ? reinterpret_cast<const void *>(&les_sgs3_kernel<STEP, 1, 128, true>)
                   : reinterpret_cast<const void *>(&les_sgs3_kernel<STEP, 1, 128>);
    return iso ? reinterpret_cast<const void *>(&les_sgs3_kernel<STEP, 1, TX, true>)
               : reinterpret_cast<const void *>(&les_sgs3_kernel<STEP, 1>);
}

// The buffer-addressed step kernels (les_step_kernel), by tile shape: 512 threads as 32 x 16 columns (the
// default) or 64 x 8 columns, four waves per SIMD, two workgroups per CU; or 256 threads as 64 x 4 columns,
// three workgroups per CU. The four-wave shapes load W and S after the update and read the stencil operands
// of each group of component updates (U and V | W | S: 32 x 16; each component: 64 x 8) from LDS again,
// which holds them within 128 registers. With the update at a raised wave priority the 32 x 16 tile is the
// faster (0.751 against 0.778-0.780 ms on one box; the same bits).
struct StepShape {
    int threads, width, per_cu;
    const void *iso, *general;
};

static StepShape step_shape(const char *name)
{
#define FTC_SHAPE(NT_, WX_, WPE_, FP_, SP_, PER_CU_)                                                        \
    StepShape{NT_, WX_, PER_CU_, reinterpret_cast<const void *>(&les_step_kernel<NT_, WX_, true, WPE_, FP_, SP_>), \
              reinterpret_cast<const void *>(&les_step_kernel<NT_, WX_, false, WPE_, FP_, SP_>)}
    if (name && std::strcmp(name, "64x8") == 0) return FTC_SHAPE(512, 64, 4, 3, 1, 2);
    if (name && std::strcmp(name, "64x4") == 0) return FTC_SHAPE(256, 64, 3, 0, 0, 3);
    return FTC_SHAPE(512, 32, 4, 3, 2, 2);
#undef FTC_SHAPE
}

// k chunks for the step kernel over `tiles` column tiles with `slots` workgroups resident at once: every chunk
// but the last is a multiple of six planes (its k loop is unrolled by six from k = 2). Takes the chunk count
// whose launch needs the fewest rounds of resident workgroups times the planes a workgroup then marches (its
// chunk plus about three planes of prologue). Returns the chunk length and sets the number of chunks.
static int step_chunks(long long tiles, int nz, long long slots, int &chunks)
{
    int best_k = nz, best_chunks = 1;
    long long best_cost = -1;
    for (int want = 1; want <= std::max(1, nz / 6); ++want) {
        const int k = std::max(6, ((nz + want - 1) / want + 5) / 6 * 6);
        const int n = (nz + k - 1) / k;
        const long long rounds = (tiles * n + slots - 1) / slots;
        const long long cost = rounds * (k + 3);
        if (best_cost < 0 || cost < best_cost) best_cost = cost, best_k = k, best_chunks = n;
    }
    if (const char *v = std::getenv("FTC_LES_KCHUNK")) {
        best_k = std::max(6, (std::max(1, std::atoi(v)) + 5) / 6 * 6);
        best_chunks = (nz + best_k - 1) / best_k;
    }
    chunks = best_chunks;
    return best_k;
}

static Sec3 box(std::size_t i0, std::size_t i1, std::size_t j0, std::size_t j1, std::size_t k0, std::size_t k1)
{
    return Sec3{{i0, j0, k0}, {i1, j1, k1}};
}

// One reference advection component (CDU, CDV or CDW) on arrays of any extents >= its loop extents + 2.
template <int MASK>
static int advection_component(fort_scope_t ctx, int mode, fort_buffer_t out, fort_buffer_t u, fort_buffer_t v,
                               fort_buffer_t w, const double *dxmin, const double *dymin, const double *dzmin,
                               const int *pnx, const int *pny, const int *pnz)
{
    FTC_CHECK(entry_preamble(mode));
    if (!dxmin || !dymin || !dzmin || !pnx || !pny || !pnz)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null shared scalar argument");
    const int nx = *pnx, ny = *pny, nz = *pnz;
    std::size_t eo[3], eu[3], ev[3], ew[3];
    FTC_CHECK(layout3(ctx, out, FORT_SCOPE_REAL64, 8, eo));
    FTC_CHECK(layout3(ctx, u, FORT_SCOPE_REAL64, 8, eu));
    FTC_CHECK(layout3(ctx, v, FORT_SCOPE_REAL64, 8, ev));
    FTC_CHECK(layout3(ctx, w, FORT_SCOPE_REAL64, 8, ew));
    for (int d = 0; d < 3; ++d)
        if (eo[d] != eu[d] || ev[d] != eu[d] || ew[d] != eu[d])
            return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "advection arrays of different shapes");
    FTC_CHECK(distinct_outputs({out}, {u, v, w}));
    if (nx <= 0 || ny <= 0 || nz <= 0) return FORT_SCOPE_OK;  // empty domains do not launch
    if ((std::size_t)nx + 2 > eu[0] || (std::size_t)ny + 2 > eu[1] || (std::size_t)nz + 2 > eu[2])
        return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "loop extents exceed the array extents");
    FTC_CHECK(fort_scope_forget_definition(ctx, out));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {out, u, v, w}, dev));
    LesArgs a{};
    a.U = (const double *)dev[1];
    a.V = (const double *)dev[2];
    a.W = (const double *)dev[3];
    a.S = nullptr;
    a.OU = (double *)dev[0];
    a.OV = (double *)dev[0];
    a.OW = (double *)dev[0];
    a.OS = nullptr;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.kw = nz + 1;
    a.sj = (long long)eu[0];
    a.sk = (long long)eu[0] * (long long)eu[1];
    fill_coefficients(a, *dxmin, *dymin, *dzmin, 0.0);
    a.kchunk = choose_kchunk(nx, ny, nz);
    const std::size_t X = nx, Y = ny, Z = nz;
    std::vector<Binding3> b(4);
    b[0] = {out, {}, {}, {box(1, X + 1, 1, Y + 1, 1, Z + 1)}};
    if (MASK == kU) {
        b[1] = {u, {box(0, X + 2, 0, Y + 2, 0, Z + 2)}, {}, {}};
        b[2] = {v, {box(1, X + 2, 0, Y + 1, 1, Z + 1)}, {}, {}};
        b[3] = {w, {box(1, X + 2, 1, Y + 1, 0, Z + 1)}, {}, {}};
    } else if (MASK == kV) {
        b[1] = {u, {box(0, X + 1, 1, Y + 2, 1, Z + 1)}, {}, {}};
        b[2] = {v, {box(0, X + 2, 0, Y + 2, 0, Z + 2)}, {}, {}};
        b[3] = {w, {box(1, X + 1, 1, Y + 2, 0, Z + 1)}, {}, {}};
    } else {
        b[1] = {u, {box(0, X + 1, 1, Y + 1, 1, Z + 2)}, {}, {}};
        b[2] = {v, {box(1, X + 1, 0, Y + 1, 1, Z + 2)}, {}, {}};
        b[3] = {w, {box(0, X + 2, 0, Y + 2, 0, Z + 2)}, {}, {}};
    }
    const uint64_t key = 0xadc0000000000000ull | MASK;
    return enqueue_kernel(ctx, key, kernel_of<MASK, false, false>(), grid_of(a, nz), dim3(TX, TY, 1), 0, &a,
                          sizeof a, b);
}

// The LES tendency (advection, optionally with Smagorinsky diffusion) or the explicit step built on it,
// over nx x ny x nz cells with one ghost layer: U, V and S on all interior cells, W on the interior
// faces k = 2..nz.
static int les_entry(bool step, fort_scope_t ctx, int mode, fort_buffer_t ou, fort_buffer_t ov, fort_buffer_t ow,
                     fort_buffer_t os, fort_buffer_t u, fort_buffer_t v, fort_buffer_t w, fort_buffer_t s,
                     const double *dx, const double *dy, const double *dz, const double *dt, const double *cs,
                     const int *pnx, const int *pny, const int *pnz, const int *psgs)
{
    FTC_CHECK(entry_preamble(mode));
    if (!dx || !dy || !dz || !cs || !pnx || !pny || !pnz || !psgs || (step && !dt))
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "null shared scalar argument");
    const int nx = *pnx, ny = *pny, nz = *pnz;
    const bool sgs = *psgs != 0;
    const fort_buffer_t all[8] = {ou, ov, ow, os, u, v, w, s};
    std::size_t e[8][3];
    for (int n = 0; n < 8; ++n) FTC_CHECK(layout3(ctx, all[n], FORT_SCOPE_REAL64, 8, e[n]));
    for (int n = 1; n < 8; ++n)
        for (int d = 0; d < 3; ++d)
            if (e[n][d] != e[0][d])
                return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "LES arrays of different shapes");
    FTC_CHECK(distinct_outputs({ou, ov, ow, os}, {u, v, w, s}));
    if (nx <= 0 || ny <= 0 || nz <= 0) return FORT_SCOPE_OK;
    if (e[0][0] != (std::size_t)nx + 2 || e[0][1] != (std::size_t)ny + 2 || e[0][2] != (std::size_t)nz + 2)
        return fort_scope_report_error(FORT_SCOPE_BOUNDARY, "LES arrays must have one ghost layer");
    if (!step)
        for (int n = 0; n < 4; ++n) FTC_CHECK(fort_scope_forget_definition(ctx, all[n]));
    std::vector<void *> dev;
    FTC_CHECK(reserve(ctx, {ou, ov, ow, os, u, v, w, s}, dev));
    LesArgs a{};
    a.OU = (double *)dev[0];
    a.OV = (double *)dev[1];
    a.OW = (double *)dev[2];
    a.OS = (double *)dev[3];
    a.U = (const double *)dev[4];
    a.V = (const double *)dev[5];
    a.W = (const double *)dev[6];
    a.S = (const double *)dev[7];
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.kw = nz;
    a.sj = nx + 2;
    a.sk = (long long)(nx + 2) * (ny + 2);
    a.dt = step ? *dt : 0.0;
    fill_coefficients(a, *dx, *dy, *dz, *cs);
    // The staged kernel: three-plane ring with one row per thread (three workgroups per CU) or two rows per
    // thread (FTC_LES_ROWS=2, two per CU); FTC_LES_UNSTAGED selects the unstaged kernel. Both variants are
    // bit-identical to the default and kept for comparisons.
    const bool staged = sgs && !std::getenv("FTC_LES_UNSTAGED");
    const char *rows_env = std::getenv("FTC_LES_ROWS");
    const int rows = staged && rows_env && std::atoi(rows_env) == 2 ? 2 : 1;
    // FTC_LES_WIDTH selects the 32 x 8 or 128 x 2 tile of the staged kernel (one row per thread).
    const char *width_env = std::getenv("FTC_LES_WIDTH");
    int width = staged && rows == 1 && width_env ? std::atoi(width_env) : TX;
    if (width != 32 && width != 128) width = TX;
    // FTC_LES_TWO_BARRIERS: the 32 x 8 tile with one more LDS slot per staged field and two barriers per
    // plane (51.6 KB, still three workgroups per CU).
    const bool two = width == 32 && std::getenv("FTC_LES_TWO_BARRIERS");
    a.kchunk = choose_kchunk(nx, ny, nz, rows, !staged ? 2 : (rows == 2 || width == 128 ? 2 : 3));
    // Tiles in launch order put each 64-column strip on one XCD, so the y neighbours that re-read a tile's
    // ghost rows share its L2. The XCD-contiguous order (FTC_LES_XCD_ORDER) also shares the x neighbours'
    // lines, but only between tiles resident at the same time; over four rounds of workgroups it read more.
    a.tile_order = std::getenv("FTC_LES_XCD_ORDER") ? 1 : 0;
    const bool persistent = staged && std::getenv("FTC_LES_PERSISTENT");
    if (persistent) a.tile_order = 2;
    const std::size_t X = nx, Y = ny, Z = nz;
    const Sec3 full = box(0, X + 2, 0, Y + 2, 0, Z + 2);
    std::vector<Binding3> b(8);
    if (step) {
        b[0] = {ou, {}, {}, {full}};
        b[1] = {ov, {}, {}, {full}};
        b[2] = {ow, {}, {}, {box(0, X + 2, 0, Y + 2, 1, Z)}};
        b[3] = {os, {}, {}, {full}};
    } else {
        const Sec3 interior = box(1, X + 1, 1, Y + 1, 1, Z + 1);
        b[0] = {ou, {}, {}, {interior}};
        b[1] = {ov, {}, {}, {interior}};
        b[2] = {ow, {}, {}, {box(1, X + 1, 1, Y + 1, 1, Z)}};
        b[3] = {os, {}, {}, {interior}};
    }
    b[4] = {u, {full}, {}, {}};
    b[5] = {v, {full}, {}, {}};
    b[6] = {w, {box(0, X + 2, 0, Y + 2, 0, Z + 1)}, {}, {}};
    b[7] = {s, {full}, {}, {}};
    const void *fn;
    // Equal spacings select the kernel that holds each per-direction coefficient once (same values, same
    // arithmetic); FTC_LES_ANISO keeps the general one for comparison.
    const bool iso = *dx == *dy && *dy == *dz && !std::getenv("FTC_LES_ANISO");
    if (staged) fn = step ? staged_kernel<true>(rows, width, iso, two) : staged_kernel<false>(rows, width, iso, two);
    else if (step) fn = sgs ? kernel_of<15, true, true>() : kernel_of<15, false, true>();
    else fn = sgs ? kernel_of<15, true, false>() : kernel_of<15, false, false>();
    const uint64_t key = 0x1e5000000000000full ^ ((uint64_t)step << 8) ^ ((uint64_t)sgs << 9) ^ ((uint64_t)iso << 10);
    dim3 grid = grid_of(a, nz, rows, width);
    if (persistent) {
        // One resident round: the workgroups the device holds at once (three, or two with two rows or the
        // wide tile, per CU), a multiple of eight, and no more than there are tiles (rounded up to eight).
        const unsigned tiles = grid.x * grid.y * grid.z;
        const unsigned per_cu = rows == 2 ? (unsigned)sgs3_occupancy<2, TX>()
                                : width == 128 ? (unsigned)sgs3_occupancy<1, 128>()
                                : width == 32  ? (two ? (unsigned)sgs3_occupancy<1, 32, true>()
                                                      : (unsigned)sgs3_occupancy<1, 32>())
                                               : (unsigned)sgs3_occupancy<1, TX>();
        const unsigned resident = (unsigned)compute_units() * per_cu / kXcds * kXcds;
        grid = dim3(std::max(kXcds, std::min(resident, (tiles + kXcds - 1) / kXcds * kXcds)), 1, 1);
    }
    dim3 block = staged ? dim3(width, NTH / width, 1) : dim3(TX, TY, 1);
    // The step with Smagorinsky diffusion runs on the buffer-addressed kernel when its arrays are below 2^31
    // bytes and no switch of the les_sgs3 kernels is set (FTC_LES_LEGACY selects those outright);
    // FTC_LES_SHAPE picks its tile shape. Tiles go to the XCDs in contiguous runs, so a resident round of
    // workgroups on one XCD covers a band of whole tile rows.
    const bool legacy_switch = std::getenv("FTC_LES_LEGACY") || rows != 1 || width != TX || persistent || two ||
                               std::getenv("FTC_LES_XCD_ORDER");
    if (staged && step && !legacy_switch && (long long)a.sk * 8 * (nz + 2) < (1LL << 31)) {
        const StepShape shape = step_shape(std::getenv("FTC_LES_SHAPE"));
        const int tyt = shape.threads / shape.width;
        const long long tiles = (long long)((nx + shape.width - 1) / shape.width) * ((ny + tyt - 1) / tyt);
        int chunks = 1;
        a.kchunk = step_chunks(tiles, nz, (long long)compute_units() * shape.per_cu, chunks);
        a.tile_order = 1;
        grid = dim3((nx + shape.width - 1) / shape.width, (ny + tyt - 1) / tyt, chunks);
        block = dim3(shape.threads, 1, 1);
        fn = iso ? shape.iso : shape.general;
    }
    return enqueue_kernel(ctx, key, fn, grid, block, 0, &a, sizeof a, b);
}

}  // namespace ftc

using namespace ftc;

extern "C" {

int cpp_fort_shared_momentumadvection_cdu(fort_scope_t ctx, int mode, fort_buffer_t u2, fort_buffer_t u,
                                          fort_buffer_t v, fort_buffer_t w, const double *dxmin,
                                          const double *dymin, const double *dzmin, const int *unx,
                                          const int *uny, const int *unz)
{
    return advection_component<kU>(ctx, mode, u2, u, v, w, dxmin, dymin, dzmin, unx, uny, unz);
}

int cpp_fort_shared_momentumadvection_cdv(fort_scope_t ctx, int mode, fort_buffer_t v2, fort_buffer_t u,
                                          fort_buffer_t v, fort_buffer_t w, const double *dxmin,
                                          const double *dymin, const double *dzmin, const int *vnx,
                                          const int *vny, const int *vnz)
{
    return advection_component<kV>(ctx, mode, v2, u, v, w, dxmin, dymin, dzmin, vnx, vny, vnz);
}

int cpp_fort_shared_momentumadvection_cdw(fort_scope_t ctx, int mode, fort_buffer_t w2, fort_buffer_t u,
                                          fort_buffer_t v, fort_buffer_t w, const double *dxmin,
                                          const double *dymin, const double *dzmin, const int *wnx,
                                          const int *wny, const int *wnz)
{
    return advection_component<kW>(ctx, mode, w2, u, v, w, dxmin, dymin, dzmin, wnx, wny, wnz);
}

int cpp_fort_shared_les_tendency(fort_scope_t ctx, int mode, fort_buffer_t tu, fort_buffer_t tv, fort_buffer_t tw,
                                 fort_buffer_t ts, fort_buffer_t u, fort_buffer_t v, fort_buffer_t w,
                                 fort_buffer_t s, const double *dx, const double *dy, const double *dz,
                                 const double *cs, const int *nx, const int *ny, const int *nz, const int *sgs)
{
    return les_entry(false, ctx, mode, tu, tv, tw, ts, u, v, w, s, dx, dy, dz, nullptr, cs, nx, ny, nz, sgs);
}

int cpp_fort_shared_les_step(fort_scope_t ctx, int mode, fort_buffer_t un, fort_buffer_t vn, fort_buffer_t wn,
                             fort_buffer_t sn, fort_buffer_t u, fort_buffer_t v, fort_buffer_t w, fort_buffer_t s,
                             const double *dx, const double *dy, const double *dz, const double *dt,
                             const double *cs, const int *nx, const int *ny, const int *nz, const int *sgs)
{
    return les_entry(true, ctx, mode, un, vn, wn, sn, u, v, w, s, dx, dy, dz, dt, cs, nx, ny, nz, sgs);
}

// The host-side coefficients every LES entry passes to its kernel, for callers that restate the step.
void ftc_les_coefficients(double dx, double dy, double dz, double cs, double *out19)
{
    LesArgs a{};
    fill_coefficients(a, dx, dy, dz, cs);
    for (int c = 0; c < 3; ++c) {
        out19[c] = a.c25[c];
        out19[3 + c] = a.c50[c];
        out19[6 + c] = a.c125[c];
    }
    for (int c = 0; c < 10; ++c) out19[9 + c] = a.sgs[c];
}

}  // extern "C"
