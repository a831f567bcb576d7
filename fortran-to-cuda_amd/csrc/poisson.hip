// Pressure Poisson solve on gfx950 (the provider declared in include/ftc_poisson.h) and the ELMM adapter
// that runs it inside a fort_scope external transaction (benchmarks/elmm/sources/gpu_pressure.cpp).
//
// Solve A phi = b, A the FD2 7-point Laplacian. Write Qx (nx x nx) and Qy (ny x ny) for the orthonormal
// real eigenvector bases of the 1D second differences in x and y (periodic: 1, cos, sin and the
// alternating vector; Neumann on the staggered faces: the DCT-II basis), with eigenvalues lx, ly. Then
//   B1 = Qx^T b          (per z plane: an nx x nx by nx x ny contraction)
//   B2 = B1 Qy           (per z plane: nx x ny by ny x ny)
//   (T_z + (lx_p + ly_q) I) Psi_pq = B2_pq   for every mode (p, q): FP64 Thomas along z
//   B3 = Psi Qy^T,  phi = Qx B3.
// For periodic x and y with power-of-two sizes the FP64 solve takes the orthonormal Hartley basis
// instead (cas(2 pi j k / n) / sqrt(n), its own inverse): the four transforms are O(n log n) FFT passes
// through LDS (fft_kernels.hpp), each one read and one write of the workspace. Otherwise the four
// contractions run on the MFMA engine (gemm.hpp): FP64 MFMA for the FP64 solve, split-bf16
// products with FP32 accumulation on FP32 workspaces for the mixed solve, whose FP64 residual
// r = b - A phi (fixed-order norms) drives FP64 iterative refinement phi += M^-1 r.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ftc_poisson.h"
#include "entry.hpp"
#include "fft_kernels.hpp"
#include "gemm.hpp"
#include "reduce.hpp"
#include "poisson_kernels.hpp"

#pragma clang fp contract(off)

namespace ftc {

static thread_local std::string g_poisson_error;

static int perr(int status, const std::string &message)
{
    g_poisson_error = message;
    return status;
}

struct View3 {
    double *x;  // interior origin
    long long ld1, ld2;
};

// Orthonormal eigenvector basis of the 1D second difference on n cells of width h, column m = mode m,
// and its eigenvalues. periodic: 1/sqrt(n); sqrt(2/n) cos / sin(2 pi f i / n); (-1)^i / sqrt(n) for even n.
// Neumann (staggered): DCT-II, c_m cos(pi m (i + 1/2) / n).
static void basis(int n, double h, bool periodic, std::vector<double> &Q, std::vector<double> &lam)
{
    Q.assign((std::size_t)n * n, 0.0);
    lam.assign(n, 0.0);
    const double pi = 3.14159265358979323846;
    const double r2 = std::sqrt(2.0 / n), r1 = 1.0 / std::sqrt((double)n);
    if (periodic) {
        for (int i = 0; i < n; ++i) Q[i] = r1;
        lam[0] = 0.0;
        for (int f = 1; 2 * f < n; ++f) {
            const double l = (2.0 * std::cos(2.0 * pi * f / n) - 2.0) / (h * h);
            for (int i = 0; i < n; ++i) {
                Q[(std::size_t)(2 * f - 1) * n + i] = r2 * std::cos(2.0 * pi * f * i / n);
                Q[(std::size_t)(2 * f) * n + i] = r2 * std::sin(2.0 * pi * f * i / n);
            }
            lam[2 * f - 1] = l;
            lam[2 * f] = l;
        }
        if (n % 2 == 0) {
            for (int i = 0; i < n; ++i) Q[(std::size_t)(n - 1) * n + i] = (i % 2 ? -r1 : r1);
            lam[n - 1] = -4.0 / (h * h);
        }
    } else {
        for (int m = 0; m < n; ++m) {
            const double c = m == 0 ? r1 : r2;
            for (int i = 0; i < n; ++i) Q[(std::size_t)m * n + i] = c * std::cos(pi * m * (i + 0.5) / n);
            lam[m] = (2.0 * std::cos(pi * m / n) - 2.0) / (h * h);
        }
    }
}

static std::vector<double> transpose(const std::vector<double> &Q, int n)
{
    std::vector<double> T(Q.size());
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) T[(std::size_t)b * n + a] = Q[(std::size_t)a * n + b];
    return T;
}

// The Hartley path: n a power of two in [4, 4096], periodic.
static bool hartley_size(int n) { return n >= 4 && n <= 4096 && (n & (n - 1)) == 0; }

static int log2_exact(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

// Eigenvalues of the periodic second difference for the Hartley index k, (2 cos(2 pi k / n) - 2) / h^2,
// and the FFT twiddles (cos, -sin)(2 pi k / n) for k < n/2.
static void hartley_tables(int n, double h, std::vector<double> &lam, std::vector<double> &tw)
{
    const double pi = 3.14159265358979323846;
    lam.assign(n, 0.0);
    for (int k = 1; k < n; ++k) lam[k] = (2.0 * std::cos(2.0 * pi * k / n) - 2.0) / (h * h);
    tw.assign(n, 0.0);
    for (int k = 0; k < n / 2; ++k) {
        tw[2 * k] = std::cos(2.0 * pi * k / n);
        tw[2 * k + 1] = -std::sin(2.0 * pi * k / n);
    }
}

}  // namespace ftc

using namespace ftc;

struct poisson_solver_handle_v1 {
    poisson_solver_config_v1 config{};
    int nx = 0, ny = 0, nz = 0;
    double hx = 0, hy = 0, hz = 0;
    bool perx = true, pery = true, mixed = false;
    bool hartley = false;  // FP64 solve through the FFT (Hartley) transforms
    double tol = 1e-12;
    int max_solves = 6, measure = 1;
    int device = 0;
    int cols = 64;
    // device storage (one allocation, carved)
    void *arena = nullptr;
    std::size_t arena_bytes = 0;
    double *Qx = nullptr, *QxT = nullptr, *Qy = nullptr, *QyT = nullptr, *lx = nullptr, *ly = nullptr;
    double2 *twx = nullptr, *twy = nullptr;  // Hartley path: FFT twiddles of x and y
    void *Qxs = nullptr, *QxTs = nullptr, *Qys = nullptr, *QyTs = nullptr;  // split bases (bf16 hi, lo planes)
    void *w0 = nullptr, *w1 = nullptr;
    double *res = nullptr, *slots = nullptr, *final256 = nullptr;
    double *partial = nullptr;     // mixed: per-block sums of squares of the residual
    double *phisum = nullptr;      // mixed: per-block sums of phi
    double *host_final = nullptr;  // pinned, 256 doubles
    poisson_solver_report_v1 report{};
    std::mutex mutex;
};

namespace ftc {

static int map_status(int scope_status)
{
    switch (scope_status) {
    case FORT_SCOPE_OK: return POISSON_SOLVER_OK;
    case FORT_SCOPE_ARGUMENT: return POISSON_SOLVER_ARGUMENT;
    case FORT_SCOPE_RESOURCE: return POISSON_SOLVER_RESOURCE;
    default: return POISSON_SOLVER_EXECUTION;
    }
}

// Complex points per workgroup of a pass: 4096 where the lines run across threads (the y transform:
// G = 2 * points / n lines make the 128-byte runs), 2048 where a line's own elements do.
static int hartley_points(bool econt, int n)
{
    static const int x = std::getenv("FTC_FFT_POINTS_X") ? std::atoi(std::getenv("FTC_FFT_POINTS_X")) : 2048;
    static const int y = std::getenv("FTC_FFT_POINTS_Y") ? std::atoi(std::getenv("FTC_FFT_POINTS_Y")) : 4096;
    const int p = econt ? x : y;
    return std::max(p == 2048 ? 2048 : 4096, n);  // at least two real lines (n <= 4096)
}

// Threads per workgroup of a pass of `points` points: 256 (8 or 16 points each), or 512 for the 4096-point
// workgroups (8 points each), whose 70 KB of LDS hold two workgroups per CU: 512 threads keep twice the
// waves in flight. FTC_FFT_THREADS=256 selects 256 everywhere.
static int hartley_threads(int points)
{
    const char *forced = std::getenv("FTC_FFT_THREADS");
    if (forced && std::atoi(forced) == 256) return 256;
    return points == 4096 ? 512 : 256;
}

template <class TI, class TO, int POINTS, int NT>
static void hartley_launch(bool econt, dim3 grid, std::size_t lds, hipStream_t stream, const HartleyPass &h)
{
    if (econt) hipLaunchKernelGGL((hartley_kernel<TI, TO, true, POINTS, NT>), grid, dim3(NT), lds, stream, h);
    else hipLaunchKernelGGL((hartley_kernel<TI, TO, false, POINTS, NT>), grid, dim3(NT), lds, stream, h);
}

template <class TI, class TO>
static int hartley_pass(const void *in, long long ie, long long ia, long long ic, void *out, long long oe, long long oa,
                        long long oc, int n, int A, int C, double beta, const double2 *tw, hipStream_t stream,
                        const double *shift = nullptr)
{
    const bool econt = ie == 1;
    const int points = hartley_points(econt, n);
    HartleyPass h{in, ie, ia, ic, out, oe, oa, oc, n, log2_exact(n), 2 * points / n, 0, A, C,
                  1.0 / std::sqrt((double)n), beta, tw, shift};
    h.logG = log2_exact(h.G);
    const std::size_t lds = (std::size_t)hartley_lds_bytes(n, points);
    const dim3 grid((unsigned)((A + h.G - 1) / h.G), (unsigned)C);
    if (points == 2048) hartley_launch<TI, TO, 2048, 256>(econt, grid, lds, stream, h);
    else if (hartley_threads(points) == 512) hartley_launch<TI, TO, 4096, 512>(econt, grid, lds, stream, h);
    else hartley_launch<TI, TO, 4096, 256>(econt, grid, lds, stream, h);
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "FFT launch failed");
    return FORT_SCOPE_OK;
}

// Raises the dynamic LDS limit of the Hartley kernels to what the largest pass needs.
template <class TI, class TO>
static bool hartley_attributes_of(int bytes)
{
    bool ok = true;
    for (const void *f : {reinterpret_cast<const void *>(&hartley_kernel<TI, TO, true, 2048, 256>),
                          reinterpret_cast<const void *>(&hartley_kernel<TI, TO, false, 2048, 256>),
                          reinterpret_cast<const void *>(&hartley_kernel<TI, TO, true, 4096, 256>),
                          reinterpret_cast<const void *>(&hartley_kernel<TI, TO, false, 4096, 256>),
                          reinterpret_cast<const void *>(&hartley_kernel<TI, TO, true, 4096, 512>),
                          reinterpret_cast<const void *>(&hartley_kernel<TI, TO, false, 4096, 512>)})
        ok = ok && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) == hipSuccess;
    return ok;
}

static bool hartley_attributes()
{
    int bytes = 0;
    for (int n = 4; n <= 4096; n *= 2)
        for (int points : {2048, 4096}) bytes = std::max(bytes, hartley_lds_bytes(n, std::max(points, n)));
    // FP64 throughout (the FP64 solve); FP64 or FP32 in, FP32 between passes, FP64 out (the mixed solve).
    return hartley_attributes_of<double, double>(bytes) && hartley_attributes_of<double, float>(bytes) &&
           hartley_attributes_of<float, float>(bytes) && hartley_attributes_of<float, double>(bytes);
}

// One FP64 solve through the Hartley transforms: phi_out = beta * phi_out + Hx Hy T^-1 Hy Hx in, with
// the workspaces W0, W1 holding the x index p fastest (W[p + q*nx + k*nx*ny]).
static int apply_inverse_hartley(poisson_solver_handle_v1 &s, const double *in, long long in1, long long in2,
                                 View3 out, double beta, hipStream_t stream)
{
    const int nx = s.nx, ny = s.ny, nz = s.nz;
    const long long plane = (long long)nx * ny;
    double *w0 = static_cast<double *>(s.w0), *w1 = static_cast<double *>(s.w1);
    FTC_CHECK((hartley_pass<double, double>(in, 1, in1, in2, w0, 1, nx, plane, nx, ny, nz, 0.0, s.twx, stream)));
    FTC_CHECK((hartley_pass<double, double>(w0, nx, 1, plane, w1, nx, 1, plane, ny, nx, nz, 0.0, s.twy, stream)));
    const unsigned blocks = (unsigned)((plane + s.cols - 1) / s.cols);
    const std::size_t lds = 2 * (std::size_t)tridiag_chunks(nz) * s.cols * sizeof(double);
    const double rz = 1.0 / (s.hz * s.hz);
    // Columns of 64, 96 or 128 levels are solved held in registers (the input read once); others, or with
    // FTC_TRIDIAG_CHUNKED, by the chunked sweep that re-reads each chunk for the back substitution.
    const bool resident = (nz == 64 || nz == 96 || nz == 128) && plane * nz * 8 < (1LL << 31) &&
                          !std::getenv("FTC_TRIDIAG_CHUNKED");
    const unsigned columns = (unsigned)((plane + 63) / 64);
    if (resident && nz == 128)
        hipLaunchKernelGGL(tridiag_z_resident_kernel<128>, dim3(columns), dim3(64), 0, stream,
                           static_cast<const double *>(w1), w0, nx, ny, s.lx, s.ly, rz);
    else if (resident && nz == 96)
        hipLaunchKernelGGL(tridiag_z_resident_kernel<96>, dim3(columns), dim3(64), 0, stream,
                           static_cast<const double *>(w1), w0, nx, ny, s.lx, s.ly, rz);
    else if (resident)
        hipLaunchKernelGGL(tridiag_z_resident_kernel<64>, dim3(columns), dim3(64), 0, stream,
                           static_cast<const double *>(w1), w0, nx, ny, s.lx, s.ly, rz);
    else
        hipLaunchKernelGGL((tridiag_z_kernel<double, double>), dim3(blocks), dim3(s.cols), lds, stream,
                           static_cast<const double *>(w1), TriOut<double>{w0, 0}, nx, ny, nz, s.lx, s.ly, rz, s.cols,
                           0);
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "tridiagonal launch failed");
    FTC_CHECK((hartley_pass<double, double>(w0, nx, 1, plane, w1, nx, 1, plane, ny, nx, nz, 0.0, s.twy, stream)));
    return hartley_pass<double, double>(w1, 1, nx, plane, out.x, 1, out.ld1, out.ld2, nx, ny, nz, beta, s.twx, stream);
}

// One application of the mixed solve's approximate inverse through the Hartley transforms: the same five
// passes as the FP64 solve, FP64 arithmetic inside each, the workspaces between them held in FP32. The
// first application reads the FP64 right-hand side, a correction the FP32 residual (in_float); the last
// pass adds into phi (beta 1 on corrections) and subtracts `shift` (phi's mean from the residual pass).
static int apply_inverse_hartley_mixed(poisson_solver_handle_v1 &s, const void *in, long long in1, long long in2,
                                       bool in_float, View3 out, double beta, const double *shift, hipStream_t stream)
{
    const int nx = s.nx, ny = s.ny, nz = s.nz;
    const long long plane = (long long)nx * ny;
    float *w0 = static_cast<float *>(s.w0), *w1 = static_cast<float *>(s.w1);
    if (in_float) FTC_CHECK((hartley_pass<float, float>(in, 1, in1, in2, w0, 1, nx, plane, nx, ny, nz, 0.0, s.twx, stream)));
    else FTC_CHECK((hartley_pass<double, float>(in, 1, in1, in2, w0, 1, nx, plane, nx, ny, nz, 0.0, s.twx, stream)));
    FTC_CHECK((hartley_pass<float, float>(w0, nx, 1, plane, w1, nx, 1, plane, ny, nx, nz, 0.0, s.twy, stream)));
    const unsigned blocks = (unsigned)((plane + s.cols - 1) / s.cols);
    const std::size_t lds = 2 * (std::size_t)tridiag_chunks(nz) * s.cols * sizeof(double);
    hipLaunchKernelGGL((tridiag_z_kernel<float, float>), dim3(blocks), dim3(s.cols), lds, stream,
                       static_cast<const float *>(w1), TriOut<float>{w0, 0}, nx, ny, nz, s.lx, s.ly,
                       1.0 / (s.hz * s.hz), s.cols, 0);
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "tridiagonal launch failed");
    FTC_CHECK((hartley_pass<float, float>(w0, nx, 1, plane, w1, nx, 1, plane, ny, nx, nz, 0.0, s.twy, stream)));
    return hartley_pass<float, double>(w1, 1, nx, plane, out.x, 1, out.ld1, out.ld2, nx, ny, nz, beta, s.twx, stream,
                                       shift);
}

// The contraction calls of one application of the approximate (mixed) or exact (FP64) inverse:
// phi_out = beta * phi_out + Qx [T^-1 ((Qx^T in) Qy)] Qy^T.
static int apply_inverse(poisson_solver_handle_v1 &s, const void *in, long long in1, long long in2, View3 out,
                         double beta, hipStream_t stream, bool in_split = false, const double *shift = nullptr)
{
    if (s.hartley && s.mixed) return apply_inverse_hartley_mixed(s, in, in1, in2, in_split, out, beta, shift, stream);
    if (s.hartley) return apply_inverse_hartley(s, static_cast<const double *>(in), in1, in2, out, beta, stream);
    const int nx = s.nx, ny = s.ny, nz = s.nz;
    const long long plane = (long long)nx * ny;
    const int prec = s.mixed ? kBf16x3 : kF64;
    const unsigned blocks = (unsigned)((plane + s.cols - 1) / s.cols);
    const std::size_t lds = 2 * (std::size_t)tridiag_chunks(nz) * s.cols * sizeof(double);
    const double rz = 1.0 / (s.hz * s.hz);
    if (s.mixed) {
        // Split-bf16 chain with every contraction operand K-contiguous: W0 = (Qx^T f)^T and W1 = W0 Qy keep
        // the mode index q fastest, so G2 and G3 read rows of W0 as vectors; G3 turns W1 back to i fastest.
        // The bases, W0 (written by G1 and by the z solve) and W1 between G3 and G4 are kept split (bf16 hi
        // and lo planes); W1 between G2 and the z solve is FP32.
        const int ak = kGemmAKContiguous, at = kGemmAKContiguous | kGemmTransposedC;
        const long long N = plane * nz, Lx = (long long)nx * nx, Ly = (long long)ny * ny;
        // The first application reads the FP64 right-hand side; the corrections read the split residual.
        GemmCall g1{s.Qxs, in, s.w0, kSplit, in_split ? kSplit : kDouble, kSplit, nx, ny, nx, nz, nx, in1, ny, 0, in2,
                    plane, 1.0, 0.0, at, Lx, in_split ? N : 0, N};
        FTC_CHECK(gemm_launch(g1, prec, stream));
        GemmCall g2{s.w0, s.Qys, s.w1, kSplit, kSplit, kFloat, nx, ny, ny, nz, ny, ny, ny, plane, 0, plane, 1.0, 0.0,
                    at, N, Ly, 0};
        FTC_CHECK(gemm_launch(g2, prec, stream));
        hipLaunchKernelGGL((tridiag_z_kernel<float, TriSplit>), dim3(blocks), dim3(s.cols), lds, stream,
                           static_cast<const float *>(s.w1), TriOut<TriSplit>{static_cast<__bf16 *>(s.w0), N}, nx,
                           ny, nz, s.lx, s.ly, rz, s.cols, 1);
        if (hipGetLastError() != hipSuccess)
            return fort_scope_report_error(FORT_SCOPE_EXECUTION, "tridiagonal launch failed");
        GemmCall g3{s.w0, s.QyTs, s.w1, kSplit, kSplit, kSplit, nx, ny, ny, nz, ny, ny, nx, plane, 0, plane, 1.0, 0.0,
                    ak, N, Ly, N};
        FTC_CHECK(gemm_launch(g3, prec, stream));
        GemmCall g4{s.QxTs, s.w1, out.x, kSplit, kSplit, kDouble, nx, ny, nx, nz, nx, nx, out.ld1, 0, plane, out.ld2,
                    1.0, beta, ak, Lx, N, 0, shift};
        return gemm_launch(g4, prec, stream);
    }
    GemmCall g1{s.QxT, in, s.w0, kDouble, kDouble, kDouble, nx, ny, nx, nz, nx, in1, nx, 0, in2, plane, 1.0, 0.0};
    FTC_CHECK(gemm_launch(g1, prec, stream));
    GemmCall g2{s.w0, s.Qy, s.w1, kDouble, kDouble, kDouble, nx, ny, ny, nz, nx, ny, nx, plane, 0, plane, 1.0, 0.0};
    FTC_CHECK(gemm_launch(g2, prec, stream));
    hipLaunchKernelGGL((tridiag_z_kernel<double, double>), dim3(blocks), dim3(s.cols), lds, stream,
                       static_cast<const double *>(s.w1), TriOut<double>{static_cast<double *>(s.w0), 0}, nx, ny, nz,
                       s.lx, s.ly, rz, s.cols, 0);
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "tridiagonal launch failed");
    GemmCall g3{s.w0, s.QyT, s.w1, kDouble, kDouble, kDouble, nx, ny, ny, nz, nx, ny, nx, plane, 0, plane, 1.0, 0.0};
    FTC_CHECK(gemm_launch(g3, prec, stream));
    GemmCall g4{s.Qx, s.w1, out.x, kDouble, kDouble, kDouble, nx, ny, nx, nz, nx, nx, out.ld1, 0, plane, out.ld2, 1.0,
                beta};
    return gemm_launch(g4, prec, stream);
}

// The mixed solve's residual, split for the next application; with `partial`, also the per-block sums of
// squares for its norm.
static dim3 residual_split_grid(const poisson_solver_handle_v1 &s)
{
    return dim3((s.nx + 63) / 64, (s.ny + 3) / 4, (s.nz + kResidualChunk - 1) / kResidualChunk);
}

static long long residual_blocks(const poisson_solver_handle_v1 &s)
{
    const dim3 g = residual_split_grid(s);
    return (long long)g.x * g.y * g.z;
}

// The mixed solve's residual, split for the next application, and the mean of phi (from the residual
// kernel's per-block sums, folded in fixed order) into mean_dev; with `partial`, also the per-block sums
// of squares of the residual for its norm.
static int residual_split(poisson_solver_handle_v1 &s, View3 phiv, View3 b, double *partial, double *mean_dev,
                          hipStream_t stream)
{
    const dim3 grid = residual_split_grid(s);
    const long long N = (long long)s.nx * s.ny * s.nz, nb = residual_blocks(s);
    // The Hartley passes read the residual as FP32, the split-bf16 contractions as bf16 hi and lo planes.
    if (s.hartley)
        hipLaunchKernelGGL(poisson_residual_split_kernel<false>, grid, dim3(64, 4), 0, stream, phiv.x, phiv.ld1,
                           phiv.ld2, b.x, b.ld1, b.ld2, (void *)s.res, N, partial, s.phisum, s.nx, s.ny, s.nz,
                           1.0 / (s.hx * s.hx), 1.0 / (s.hy * s.hy), 1.0 / (s.hz * s.hz), s.perx ? 1 : 0,
                           s.pery ? 1 : 0);
    else
        hipLaunchKernelGGL(poisson_residual_split_kernel<true>, grid, dim3(64, 4), 0, stream, phiv.x, phiv.ld1,
                           phiv.ld2, b.x, b.ld1, b.ld2, (void *)s.res, N, partial, s.phisum, s.nx, s.ny, s.nz,
                           1.0 / (s.hx * s.hx), 1.0 / (s.hy * s.hy), 1.0 / (s.hz * s.hz), s.perx ? 1 : 0,
                           s.pery ? 1 : 0);
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "residual launch failed");
    ReduceBox box{s.phisum, nb, 1, 1, nb, nb};
    FTC_CHECK(reduce_launch(kReduceSum, box, s.slots, s.final256, stream));
    hipLaunchKernelGGL(poisson_mean_kernel, dim3(1), dim3(1), 0, stream, s.final256, (double)N, mean_dev);
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "mean launch failed");
    return FORT_SCOPE_OK;
}

static int subtract_mean(poisson_solver_handle_v1 &s, View3 phi, const double *mean_dev, hipStream_t stream)
{
    const dim3 grid((s.nx + 63) / 64, (s.ny + 3) / 4, s.nz);
    hipLaunchKernelGGL(poisson_subtract_kernel, grid, dim3(64, 4), 0, stream, phi.x, phi.ld1, phi.ld2, s.nx, s.ny,
                       mean_dev);
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "mean removal failed");
    return FORT_SCOPE_OK;
}

static int residual(poisson_solver_handle_v1 &s, const double *phi, View3 phiv, View3 b, hipStream_t stream)
{
    const dim3 grid((s.nx + 63) / 64, (s.ny + 3) / 4, s.nz);
    hipLaunchKernelGGL(poisson_residual_kernel, grid, dim3(64, 4), 0, stream, phi, phiv.ld1, phiv.ld2, b.x, b.ld1,
                       b.ld2, s.res, s.nx, s.ny, s.nz, 1.0 / (s.hx * s.hx), 1.0 / (s.hy * s.hy), 1.0 / (s.hz * s.hz),
                       s.perx ? 1 : 0, s.pery ? 1 : 0);
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "residual launch failed");
    return FORT_SCOPE_OK;
}

// Fixed-order reduction `op` of a box, folded on the host; synchronises the stream.
static int reduce_to_host(poisson_solver_handle_v1 &s, int op, const ReduceBox &box, hipStream_t stream, double *out)
{
    FTC_CHECK(reduce_launch(op, box, s.slots, s.final256, stream));
    if (hipMemcpyAsync(s.host_final, s.final256, kReduceFinal * sizeof(double), hipMemcpyDeviceToHost, stream) !=
            hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return fort_scope_report_error(FORT_SCOPE_EXECUTION, "norm readback failed");
    *out = reduce_host_fold(op, 0.0, s.host_final);
    return FORT_SCOPE_OK;
}

// Fixed-order 2-norm of a box; synchronises the stream.
static int norm2(poisson_solver_handle_v1 &s, const ReduceBox &box, hipStream_t stream, double *out)
{
    FTC_CHECK(reduce_to_host(s, kReduceSumSquares, box, stream, out));
    *out = std::sqrt(*out);
    return FORT_SCOPE_OK;
}

static int sum_of(poisson_solver_handle_v1 &s, const ReduceBox &box, hipStream_t stream, double *out)
{
    return reduce_to_host(s, kReduceSum, box, stream, out);
}

// phi -= mean(phi) over the interior, in FP64 with a fixed-order sum, without a host synchronisation.
// The singular problem's solution is the zero-mean one; the approximate (split-bf16) transforms leak a
// small constant into phi that no residual sees, so the mixed solve removes it explicitly.
static int remove_mean(poisson_solver_handle_v1 &s, View3 phi, hipStream_t stream)
{
    ReduceBox box{phi.x, s.nx, s.ny, s.nz, phi.ld1, phi.ld2};
    FTC_CHECK(reduce_launch(kReduceSum, box, s.slots, s.final256, stream));
    double *mean = s.final256 + kReduceFinal;
    hipLaunchKernelGGL(poisson_mean_kernel, dim3(1), dim3(1), 0, stream, s.final256,
                       (double)s.nx * s.ny * s.nz, mean);
    const dim3 grid((s.nx + 63) / 64, (s.ny + 3) / 4, s.nz);
    hipLaunchKernelGGL(poisson_subtract_kernel, grid, dim3(64, 4), 0, stream, phi.x, phi.ld1, phi.ld2, s.nx, s.ny,
                       mean);
    if (hipGetLastError() != hipSuccess) return fort_scope_report_error(FORT_SCOPE_EXECUTION, "mean removal failed");
    return FORT_SCOPE_OK;
}

static std::size_t align_up(std::size_t x) { return (x + 255) & ~std::size_t(255); }

static std::size_t arena_size(const poisson_solver_handle_v1 &s)
{
    const std::size_t N = (std::size_t)s.nx * s.ny * s.nz, wbytes = s.mixed ? 4 : 8;
    std::size_t b = 0;
    b += 2 * align_up((std::size_t)s.nx * s.nx * 8) + 2 * align_up((std::size_t)s.ny * s.ny * 8);
    b += 2 * align_up((std::size_t)s.nx * s.nx * 4) + 2 * align_up((std::size_t)s.ny * s.ny * 4);
    b += align_up((std::size_t)s.nx * 8) + align_up((std::size_t)s.ny * 8);
    b += align_up((std::size_t)s.nx * 8) + align_up((std::size_t)s.ny * 8);  // Hartley twiddles
    b += 2 * align_up(N * wbytes) + align_up(N * 8);
    b += align_up((kReduceSlots + kReduceFinal + 1) * 8);  // slots, 256 partials, the mean
    b += 2 * align_up((std::size_t)residual_blocks(s) * 8);  // per-block residual sums of squares, of phi
    return b;
}

}  // namespace ftc

extern "C" {

const char *poisson_solver_last_error_v1(void) { return g_poisson_error.c_str(); }

int poisson_solver_create_v1(const poisson_solver_config_v1 *c, poisson_solver_handle_v1 **out)
{
    if (!c || !out) return perr(POISSON_SOLVER_ARGUMENT, "missing configuration or output");
    *out = nullptr;
    if (c->version != 1 || (c->flags & ~uint32_t(POISSON_SOLVER_REAL32 | POISSON_SOLVER_MIXED)))
        return perr(POISSON_SOLVER_ARGUMENT, "unknown solver configuration");
    if (c->flags & POISSON_SOLVER_REAL32)
        return perr(POISSON_SOLVER_UNSUPPORTED, "REAL32 fields are not supported (use POISSON_SOLVER_MIXED)");
    if (c->dimensions != 3) return perr(POISSON_SOLVER_UNSUPPORTED, "only three dimensions are supported");
    if (c->approximation != 2) return perr(POISSON_SOLVER_UNSUPPORTED, "only the FD2 7-point operator is supported");
    for (int d = 0; d < 3; ++d) {
        if (c->sizes[d] < 1 || c->sizes[d] > (1 << 20)) return perr(POISSON_SOLVER_ARGUMENT, "invalid size");
        if (!(c->lengths[d] > 0) || !std::isfinite(c->lengths[d]))
            return perr(POISSON_SOLVER_ARGUMENT, "invalid length");
        const int a = c->boundary_conditions[2 * d], b = c->boundary_conditions[2 * d + 1];
        if (a != b || (a != 0 && a != 4)) return perr(POISSON_SOLVER_UNSUPPORTED, "only periodic/Neumann face pairs");
    }
    if (c->boundary_conditions[4] != 4) return perr(POISSON_SOLVER_UNSUPPORTED, "z must be Neumann (staggered)");
    auto s = std::make_unique<poisson_solver_handle_v1>();
    s->config = *c;
    s->nx = c->sizes[0];
    s->ny = c->sizes[1];
    s->nz = c->sizes[2];
    s->hx = c->lengths[0] / s->nx;
    s->hy = c->lengths[1] / s->ny;
    s->hz = c->lengths[2] / s->nz;
    s->perx = c->boundary_conditions[0] == 0;
    s->pery = c->boundary_conditions[2] == 0;
    s->mixed = (c->flags & POISSON_SOLVER_MIXED) != 0;
    s->hartley = s->perx && s->pery && hartley_size(s->nx) && hartley_size(s->ny) && !std::getenv("FTC_POISSON_DENSE");
    s->max_solves = s->mixed ? 6 : 1;
    s->cols = 64;
    if (s->nz > kPoissonMaxNz) return perr(POISSON_SOLVER_UNSUPPORTED, "nz above 1024 exceeds the LDS of the z solve");
    s->arena_bytes = arena_size(*s);
    if (c->device_budget_bytes && s->arena_bytes > c->device_budget_bytes)
        return perr(POISSON_SOLVER_RESOURCE, "workspace exceeds device_budget_bytes");
    if (hipGetDevice(&s->device) != hipSuccess) return perr(POISSON_SOLVER_RESOURCE, "no device");
    if (hipMalloc(&s->arena, s->arena_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return perr(POISSON_SOLVER_RESOURCE, "workspace allocation failed");
    }
    if (hipHostMalloc(reinterpret_cast<void **>(&s->host_final), kReduceFinal * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(s->arena);
        return perr(POISSON_SOLVER_RESOURCE, "pinned allocation failed");
    }
    char *p = static_cast<char *>(s->arena);
    auto carve = [&](std::size_t bytes) {
        char *r = p;
        p += align_up(bytes);
        return r;
    };
    const int nx = s->nx, ny = s->ny;
    const std::size_t N = (std::size_t)nx * ny * s->nz, wbytes = s->mixed ? 4 : 8;
    s->Qx = (double *)carve((std::size_t)nx * nx * 8);
    s->QxT = (double *)carve((std::size_t)nx * nx * 8);
    s->Qy = (double *)carve((std::size_t)ny * ny * 8);
    s->QyT = (double *)carve((std::size_t)ny * ny * 8);
    s->Qxs = carve((std::size_t)nx * nx * 4);
    s->QxTs = carve((std::size_t)nx * nx * 4);
    s->Qys = carve((std::size_t)ny * ny * 4);
    s->QyTs = carve((std::size_t)ny * ny * 4);
    s->lx = (double *)carve((std::size_t)nx * 8);
    s->ly = (double *)carve((std::size_t)ny * 8);
    s->twx = (double2 *)carve((std::size_t)nx * 8);
    s->twy = (double2 *)carve((std::size_t)ny * 8);
    s->w0 = carve(N * wbytes);
    s->w1 = carve(N * wbytes);
    s->res = (double *)carve(N * 8);
    s->partial = (double *)carve((std::size_t)residual_blocks(*s) * 8);
    s->phisum = (double *)carve((std::size_t)residual_blocks(*s) * 8);
    s->slots = (double *)carve((kReduceSlots + kReduceFinal + 1) * 8);
    s->final256 = s->slots + kReduceSlots;
    std::vector<double> qx, lx, qy, ly, twx(2, 0.0), twy(2, 0.0);
    basis(nx, s->hx, s->perx, qx, lx);
    basis(ny, s->hy, s->pery, qy, ly);
    if (s->hartley) {
        // The z solve indexes the eigenvalues by the transforms' mode index: Hartley order here.
        hartley_tables(nx, s->hx, lx, twx);
        hartley_tables(ny, s->hy, ly, twy);
        if (!hartley_attributes()) {
            (void)hipGetLastError();
            (void)hipFree(s->arena);
            (void)hipHostFree(s->host_final);
            return perr(POISSON_SOLVER_RESOURCE, "FFT kernel attributes could not be set");
        }
    }
    std::vector<double> qxt = transpose(qx, nx), qyt = transpose(qy, ny);
    // The split-bf16 solve takes the bases pre-split (bf16 hi plane, then lo plane: the FP32 footprint).
    std::vector<uint16_t> qxf = split_planes(qx), qxtf = split_planes(qxt), qyf = split_planes(qy),
                          qytf = split_planes(qyt);
    bool ok = hipMemcpy(s->Qx, qx.data(), qx.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->QxT, qxt.data(), qxt.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->Qy, qy.data(), qy.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->QyT, qyt.data(), qyt.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->Qxs, qxf.data(), qxf.size() * 2, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->QxTs, qxtf.data(), qxtf.size() * 2, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->Qys, qyf.data(), qyf.size() * 2, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->QyTs, qytf.data(), qytf.size() * 2, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->lx, lx.data(), lx.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->ly, ly.data(), ly.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
              (!s->hartley || (hipMemcpy(s->twx, twx.data(), (std::size_t)nx * 8, hipMemcpyHostToDevice) == hipSuccess &&
                               hipMemcpy(s->twy, twy.data(), (std::size_t)ny * 8, hipMemcpyHostToDevice) == hipSuccess));
    if (!ok) {
        (void)hipGetLastError();
        (void)hipFree(s->arena);
        (void)hipHostFree(s->host_final);
        return perr(POISSON_SOLVER_RESOURCE, "basis upload failed");
    }
    s->report.version = 1;
    s->report.relative_residual = -1.0;
    *out = s.release();
    return POISSON_SOLVER_OK;
}

int poisson_solver_workspace_query_v1(poisson_solver_handle_v1 *s, poisson_solver_workspace_v1 *bytes)
{
    if (!s || !bytes) return perr(POISSON_SOLVER_ARGUMENT, "missing solver or output");
    bytes->version = 1;
    bytes->flags = POISSON_SOLVER_DEVICE_BUFFERS_EXACT | POISSON_SOLVER_DRIVER_OVERHEAD_EXCLUDED;
    bytes->persistent_device_bytes = s->arena_bytes;
    bytes->peak_device_bytes = s->arena_bytes;
    bytes->execution_extra_bytes = 0;
    return POISSON_SOLVER_OK;
}

int poisson_solver_set_refinement_v1(poisson_solver_handle_v1 *s, double tol, int32_t max_solves, int32_t measure)
{
    if (!s || !(tol >= 0) || max_solves < 1) return perr(POISSON_SOLVER_ARGUMENT, "invalid refinement controls");
    std::lock_guard<std::mutex> lock(s->mutex);
    s->tol = tol;
    s->max_solves = s->mixed ? max_solves : 1;
    s->measure = measure != 0;
    return POISSON_SOLVER_OK;
}

int poisson_solver_last_report_v1(poisson_solver_handle_v1 *s, poisson_solver_report_v1 *report)
{
    if (!s || !report) return perr(POISSON_SOLVER_ARGUMENT, "missing solver or output");
    *report = s->report;
    return POISSON_SOLVER_OK;
}

int poisson_solver_execute_device_v1(poisson_solver_handle_v1 *s, void *phi, size_t phi_bytes, const void *rhs,
                                     size_t rhs_bytes, const int32_t phi_ghosts[3], const int32_t rhs_ghosts[3],
                                     void *stream_ptr, int *started)
{
    if (started) *started = 0;
    if (!s || !phi || !rhs || !phi_ghosts || !rhs_ghosts || !started)
        return perr(POISSON_SOLVER_ARGUMENT, "missing execute arguments");
    std::lock_guard<std::mutex> lock(s->mutex);
    const int n[3] = {s->nx, s->ny, s->nz};
    long long pe[3], re[3];
    std::size_t pcount = 1, rcount = 1;
    for (int d = 0; d < 3; ++d) {
        if (phi_ghosts[d] < 0 || rhs_ghosts[d] < 0) return perr(POISSON_SOLVER_ARGUMENT, "negative ghost width");
        pe[d] = n[d] + 2LL * phi_ghosts[d];
        re[d] = n[d] + 2LL * rhs_ghosts[d];
        pcount *= (std::size_t)pe[d];
        rcount *= (std::size_t)re[d];
    }
    if (phi_bytes != pcount * 8 || rhs_bytes != rcount * 8)
        return perr(POISSON_SOLVER_ARGUMENT, "phi/rhs bytes differ from the configured layouts");
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    View3 pv{static_cast<double *>(phi) + phi_ghosts[0] + phi_ghosts[1] * pe[0] + phi_ghosts[2] * pe[0] * pe[1], pe[0],
             pe[0] * pe[1]};
    View3 bv{const_cast<double *>(static_cast<const double *>(rhs)) + rhs_ghosts[0] + rhs_ghosts[1] * re[0] +
                 rhs_ghosts[2] * re[0] * re[1],
             re[0], re[0] * re[1]};
    *started = 1;
    const long long plane = (long long)s->nx * s->ny;
    s->report.iterations = 0;
    s->report.converged = 0;
    s->report.relative_residual = -1.0;
    s->report.rhs_norm = -1.0;
    int st = FORT_SCOPE_OK;
    double bnorm = 0.0;
    double *mean_dev = s->final256 + kReduceFinal;  // device scalar: phi's mean from the last residual
    bool have_mean = false;
    if (s->measure) {
        ReduceBox box{bv.x, s->nx, s->ny, s->nz, bv.ld1, bv.ld2};
        st = norm2(*s, box, stream, &bnorm);
        s->report.rhs_norm = bnorm;
    }
    if (st == FORT_SCOPE_OK) st = apply_inverse(*s, bv.x, bv.ld1, bv.ld2, pv, 0.0, stream);
    s->report.iterations = 1;
    while (st == FORT_SCOPE_OK) {
        const bool last = s->report.iterations >= s->max_solves;
        if (!s->measure && last) break;
        // FP64 solve: the FP64 residual and its fixed-order norm. Mixed solve: the split residual, whose
        // norm is the fixed-order sum of the residual kernel's per-block sums of squares.
        st = s->mixed ? residual_split(*s, pv, bv, s->measure ? s->partial : nullptr, mean_dev, stream)
                      : residual(*s, pv.x, pv, bv, stream);
        if (st) break;
        have_mean = s->mixed;
        if (s->measure) {
            double rn = 0.0;
            if (s->mixed) {
                ReduceBox box{s->partial, residual_blocks(*s), 1, 1, residual_blocks(*s), residual_blocks(*s)};
                st = sum_of(*s, box, stream, &rn);
                rn = std::sqrt(rn);
            } else {
                ReduceBox box{s->res, s->nx, s->ny, s->nz, s->nx, plane};
                st = norm2(*s, box, stream, &rn);
            }
            if (st) break;
            s->report.relative_residual = bnorm > 0 ? rn / bnorm : rn;
            if (s->report.relative_residual <= s->tol) {
                s->report.converged = 1;
                break;
            }
            if (last) break;
        }
        // The mixed correction also subtracts phi's current mean: the Laplacian does not see a constant,
        // and the approximate transforms leak one (about the transform error times the correction).
        st = apply_inverse(*s, s->res, s->nx, plane, pv, 1.0, stream, s->mixed, s->mixed ? mean_dev : nullptr);
        have_mean = false;
        s->report.iterations++;
    }
    // Zero mean for the mixed solve: after a residual, subtract the mean it measured; after a single
    // application, remove the mean explicitly. After a correction, what is left is the last correction's
    // leak, as small as the field error that correction leaves.
    if (st == FORT_SCOPE_OK && s->mixed) {
        if (have_mean) st = subtract_mean(*s, pv, mean_dev, stream);
        else if (s->report.iterations == 1) st = remove_mean(*s, pv, stream);
    }
    if (st) return perr(map_status(st), fort_scope_error());
    return POISSON_SOLVER_OK;
}

int poisson_solver_destroy_v1(poisson_solver_handle_v1 **h)
{
    if (!h) return perr(POISSON_SOLVER_ARGUMENT, "missing solver pointer");
    poisson_solver_handle_v1 *s = *h;
    if (!s) return POISSON_SOLVER_OK;
    (void)hipDeviceSynchronize();
    if (s->arena) (void)hipFree(s->arena);
    if (s->host_final) (void)hipHostFree(s->host_final);
    delete s;
    *h = nullptr;
    return POISSON_SOLVER_OK;
}

}  // extern "C"

// ================================================================================================ ELMM adapter
struct elmm_gpu_pressure_handle_v1 {
    fort_scope_domain_t domain = 0;
    uint64_t lease = 0;
    std::size_t workspace = 0;
    int device = -1;
    poisson_solver_handle_v1 *solver = nullptr;
    elmm_gpu_pressure_config_v1 config{};
    bool ready = false;
    std::mutex mutex;
};

namespace ftc {
static thread_local std::string g_adapter_error;

static int adapter_error(int status, const std::string &m)
{
    g_adapter_error = m;
    return status;
}

static int provider_status(int status)
{
    if (!status) return FORT_SCOPE_OK;
    g_adapter_error = poisson_solver_last_error_v1();
    if (status == POISSON_SOLVER_ARGUMENT || status == POISSON_SOLVER_UNSUPPORTED) return FORT_SCOPE_ARGUMENT;
    if (status == POISSON_SOLVER_RESOURCE) return FORT_SCOPE_RESOURCE;
    return FORT_SCOPE_EXECUTION;
}

// The views the transaction hands over must be the dense REAL64 layouts the configuration promises.
static bool view_ok(const fort_scope_external_view_v1 &v, const elmm_gpu_pressure_handle_v1 &h, bool phi)
{
    if (v.type != FORT_SCOPE_REAL64 || v.rank != 3 || v.element_bytes != 8 || !v.extents || !v.strides ||
        !v.lower_bounds || !v.device)
        return false;
    const int g = phi ? 2 : 1;
    const int64_t *lower = phi ? h.config.phi_lower_bounds : h.config.rhs_lower_bounds;
    std::size_t bytes = 8;
    for (int d = 0; d < 3; ++d) {
        if (v.extents[d] != (std::size_t)h.config.solver.sizes[d] + 2 * g || v.strides[d] != bytes ||
            v.lower_bounds[d] != lower[d])
            return false;
        bytes *= v.extents[d];
    }
    return v.bytes == bytes;
}

static int solve_worker(void *opaque, const fort_scope_external_view_v1 *views, std::size_t count, void *stream,
                        int *started)
{
    *started = 0;
    auto &h = *static_cast<elmm_gpu_pressure_handle_v1 *>(opaque);
    if (count != 2 || !views) return adapter_error(FORT_SCOPE_ARGUMENT, "pressure transaction requires Phi and RHS");
    if (!view_ok(views[0], h, true) || !view_ok(views[1], h, false))
        return adapter_error(FORT_SCOPE_ARGUMENT, "pressure view differs from the checked dense layout");
    const int32_t pg[3] = {2, 2, 2}, rg[3] = {1, 1, 1};
    return provider_status(poisson_solver_execute_device_v1(h.solver, views[0].device, views[0].bytes, views[1].device,
                                                            views[1].bytes, pg, rg, stream, started));
}
}  // namespace ftc

extern "C" {

const char *elmm_gpu_pressure_error_v1(void) { return g_adapter_error.c_str(); }

poisson_solver_handle_v1 *elmm_gpu_pressure_provider_v1(elmm_gpu_pressure_handle_v1 *h) { return h ? h->solver : nullptr; }

int elmm_gpu_pressure_create_v1(fort_scope_domain_t domain, const elmm_gpu_pressure_config_v1 *c,
                                elmm_gpu_pressure_handle_v1 **out)
{
    if (!out) return adapter_error(FORT_SCOPE_ARGUMENT, "missing pressure provider output");
    *out = nullptr;
    if (!c || !domain) return adapter_error(FORT_SCOPE_ARGUMENT, "missing pressure config/domain");
    if (c->version != 1 || c->reserved || c->solver.version != 1)
        return adapter_error(FORT_SCOPE_ARGUMENT, "unknown pressure adapter configuration");
    if (!c->solver.device_budget_bytes) return adapter_error(FORT_SCOPE_ARGUMENT, "invalid provider workspace cap");
    for (int d = 0; d < 3; ++d)
        if (c->solver.sizes[d] < 2)
            return adapter_error(FORT_SCOPE_ARGUMENT, "two-layer pressure faces require extents >= 2");
    int device = -1;
    int st = fort_scope_domain_device_get_v1(domain, &device);
    if (st) return adapter_error(st, fort_scope_error());
    auto h = std::make_unique<elmm_gpu_pressure_handle_v1>();
    h->domain = domain;
    h->config = *c;
    h->device = device;
    st = fort_scope_domain_workspace_reserve_v1(domain, (std::size_t)c->solver.device_budget_bytes, &h->lease);
    if (st) return adapter_error(st, fort_scope_error());
    int previous = 0;
    (void)hipGetDevice(&previous);
    (void)hipSetDevice(device);
    st = provider_status(poisson_solver_create_v1(&c->solver, &h->solver));
    poisson_solver_workspace_v1 bytes{};
    if (!st) st = provider_status(poisson_solver_workspace_query_v1(h->solver, &bytes));
    if (!st) {
        h->workspace = (std::size_t)(bytes.persistent_device_bytes + bytes.execution_extra_bytes);
        st = fort_scope_domain_workspace_resize_v1(domain, h->lease, h->workspace);
        if (st) g_adapter_error = fort_scope_error();
    }
    (void)hipSetDevice(previous);
    if (st) {
        if (h->solver) (void)poisson_solver_destroy_v1(&h->solver);
        (void)fort_scope_domain_workspace_release_v1(domain, h->lease);
        return st;
    }
    h->ready = true;
    *out = h.release();
    return FORT_SCOPE_OK;
}

int elmm_gpu_pressure_solve_v1(fort_scope_t context, elmm_gpu_pressure_handle_v1 *h, fort_buffer_t phi,
                               fort_buffer_t rhs, fort_scope_external_report_v1 *report)
{
    if (report) *report = fort_scope_external_report_v1{1, FORT_SCOPE_ARGUMENT, 0, 0};
    if (!h || !report || !phi || !rhs || phi == rhs)
        return adapter_error(FORT_SCOPE_ARGUMENT, "missing/aliased pressure roots or provider");
    std::lock_guard<std::mutex> lock(h->mutex);
    fort_scope_domain_t domain = 0;
    int st = fort_scope_domain_get_v1(context, &domain);
    if (st) return adapter_error(st, fort_scope_error());
    if (domain != h->domain || !domain || !h->ready)
        return adapter_error(FORT_SCOPE_ARGUMENT, "pressure provider belongs to another domain or is not ready");
    std::size_t pl[3] = {2, 2, 2}, pu[3], rl[3] = {1, 1, 1}, ru[3];
    for (int d = 0; d < 3; ++d) {
        pu[d] = (std::size_t)h->config.solver.sizes[d] + 2;
        ru[d] = (std::size_t)h->config.solver.sizes[d] + 1;
    }
    fort_scope_section interior{pl, pu}, input{rl, ru};
    fort_scope_access write{0, 0, nullptr, 1, &interior, 1, &interior}, read{0, 1, &input, 0, nullptr, 0, nullptr};
    fort_scope_external_binding_v1 bindings[2] = {{phi, FORT_SCOPE_REAL64, 3, 8, write},
                                                  {rhs, FORT_SCOPE_REAL64, 3, 8, read}};
    fort_scope_external_call_v1 call{1, 0, bindings, 2, h->lease, h->workspace};
    st = fort_scope_external_execute_v1(context, &call, solve_worker, h, report);
    report->status = st;
    if (st) g_adapter_error = g_adapter_error.empty() ? fort_scope_error() : g_adapter_error;
    return st;
}

int elmm_gpu_pressure_destroy_v1(elmm_gpu_pressure_handle_v1 **out)
{
    if (!out) return adapter_error(FORT_SCOPE_ARGUMENT, "missing pressure provider pointer");
    elmm_gpu_pressure_handle_v1 *h = *out;
    if (!h) return FORT_SCOPE_OK;
    {
        std::lock_guard<std::mutex> lock(h->mutex);
        if (h->solver) {
            int st = provider_status(poisson_solver_destroy_v1(&h->solver));
            if (st) return st;
        }
        if (h->lease) {
            int st = fort_scope_domain_workspace_release_v1(h->domain, h->lease);
            if (st) return adapter_error(st, fort_scope_error());
            h->lease = 0;
        }
    }
    delete h;
    *out = nullptr;
    return FORT_SCOPE_OK;
}

}  // extern "C"
