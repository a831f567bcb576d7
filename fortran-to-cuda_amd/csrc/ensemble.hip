// Batched ensembles of uneven LES subdomains (include/ftc_ensemble.h): every stage of the projected LES
// step runs for all members in one persistent launch over a work list, the step is one captured hipGraph
// per buffer parity, and a small autotuner picks the LES k chunk, the contraction tile and the
// persistent grid. The per-member FP64 mode launches the very same kernels once per member per stage
// (the baseline the batched modes are measured against), so batched FP64 equals it bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ftc_ensemble.h"
#include "../../include/ftc_les.h"
#include "entry.hpp"
#include "gemm.hpp"
#include "gemm_kernels.hpp"
#include "les_kernels.hpp"
#include "poisson_kernels.hpp"

#pragma clang fp contract(off)

namespace ftc {

static thread_local std::string g_ens_error;

static int eerr(int status, const std::string &m)
{
    g_ens_error = m;
    return status;
}

// Device-side description of one member.
struct EnsMember {
    int nx, ny, nz, reserved;
    long long sj, sk;  // strides of the one-ghost field arrays
    double *f[2][4];   // U, V, W, S of buffer set 0 and 1
    double *phi, *rhs, *res;  // dense nx*ny*nz
    void *w0, *w1;            // transform workspaces (double or float)
    double *lx, *ly;
    double *part;             // 256 partial sums, then the mean
    double gx, gy, gz;        // dt/d
    double rdx, rdy, rdz, rdt;
    double rx, ry, rz;        // 1/h^2
};

struct TileItem {
    int member, bx, by, bz;
};
struct GemmItem {
    int call, mt, nt, b;
};
struct ColItem {
    int member, block;
};

constexpr int kCols = 64;

// ------------------------------------------------------------------------------------------------ kernels
// One launch per tile width over the LES tiles of the members that take that width, one workgroup per item
// (the dispatcher starts them in item order as slots free up): the buffer-addressed step body of the
// single-domain step kernel, 512 threads as 64 x 8 or 32 x 16 columns at four waves per SIMD (two workgroups
// per CU). An item marches the planes of its k chunk; every chunk but a member's last is a multiple of six
// planes and starts at a plane k = 2 (mod 6), as the body's unrolled k loop requires. args is __restrict__:
// no store of the step writes a member's argument block.
constexpr int kStepThreads = 512;

template <int WX>
constexpr int step_split() { return WX == 64 ? 1 : 2; }

template <int WX, bool ISO>
__global__ __attribute__((amdgpu_flat_work_group_size(kStepThreads, kStepThreads), amdgpu_waves_per_eu(4, 4))) void
ens_les_kernel(const LesArgs *__restrict__ args, const TileItem *__restrict__ items)
{
    __shared__ StepLds<kStepThreads, WX> L;
    const TileItem t = items[blockIdx.x];
    const LesArgs &a = args[t.member];
    const int k0 = 2 + t.bz * a.kchunk;
    const int k1 = min(k0 + a.kchunk - 1, a.nz + 1);
    les_step_body<kStepThreads, WX, ISO, 3, step_split<WX>()>(a, L, t.bx, t.by, k0, k1);
}

// The k chunk of a member of nz planes: the tuned length rounded up to a multiple of six, or the whole
// column.
static int les_chunk(int kchunk, int nz)
{
    const int k = (std::max(kchunk, 1) + 5) / 6 * 6;
    return std::min(k, nz);
}

// The LES tile width of a member under a tile policy: 0 every member 64 wide, 1 every member 32 wide,
// 2 per member the width that leaves fewer idle lanes in the last tile column (64 on a tie).
static int les_width(int policy, int nx)
{
    if (policy == 0) return 64;
    if (policy == 1) return 32;
    const int idle64 = (nx + 63) / 64 * 64 - nx, idle32 = (nx + 31) / 32 * 32 - nx;
    return idle32 < idle64 ? 32 : 64;
}

// rhs = div(U*) / dt on every cell of the member (U* = buffer set `src`). An item is a block of kFlat
// consecutive columns of a member's plane (x fastest), one thread per column, so every lane of a block but
// the member's last holds a column whatever the member's width.
constexpr int kFlat = 256;

__global__ __launch_bounds__(kFlat) void ens_divergence_kernel(const EnsMember *ms, const ColItem *items, int n, int src)
{
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const ColItem t = items[it];
        const EnsMember m = ms[t.member];
        const int col = t.block * kFlat + threadIdx.x;
        if (col >= m.nx * m.ny) continue;
        const int a = 1 + col % m.nx, b = 1 + col / m.nx;
        const double *U = src ? m.f[1][0] : m.f[0][0], *V = src ? m.f[1][1] : m.f[0][1];
        const double *W = src ? m.f[1][2] : m.f[0][2];
        // Planes in batches of four, every load of a batch before its stores (rhs is not read here).
        constexpr int kB = 4;
        const long long plane = (long long)m.nx * m.ny;
        for (int c0 = 1; c0 <= m.nz; c0 += kB) {
            double u0[kB], u1[kB], v0[kB], v1[kB], w0[kB], w1[kB];
#pragma unroll
            for (int d = 0; d < kB; ++d) {
                const long long p = a + b * m.sj + min(c0 + d, m.nz) * m.sk;
                u0[d] = U[p], u1[d] = U[p - 1], v0[d] = V[p], v1[d] = V[p - m.sj], w0[d] = W[p], w1[d] = W[p - m.sk];
            }
#pragma unroll
            for (int d = 0; d < kB; ++d) {
                const int c = c0 + d;
                if (c > m.nz) break;
                const long long q = (a - 1) + (long long)m.nx * (b - 1) + plane * (c - 1);
                m.rhs[q] = (((u0[d] - u1[d]) * m.rdx + (v0[d] - v1[d]) * m.rdy) + (w0[d] - w1[d]) * m.rdz) * m.rdt;
            }
        }
    }
}

// U -= dt grad(phi) on the interior faces of buffer set `dst`, in place, with the halo images (the same
// stores and images as put_with_images; the column's image offsets and phi's periodic neighbours are
// formed once per column). Items as for the divergence.
__global__ __launch_bounds__(kFlat) void ens_projection_kernel(const EnsMember *ms, const LesArgs *args,
                                                               const ColItem *items, int n, int dst)
{
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const ColItem t = items[it];
        const EnsMember m = ms[t.member];
        const LesArgs g = args[t.member];
        const int col = t.block * kFlat + threadIdx.x;
        if (col >= m.nx * m.ny) continue;
        const int a = 1 + col % m.nx, b = 1 + col / m.nx;
        double *U = dst ? m.f[1][0] : m.f[0][0], *V = dst ? m.f[1][1] : m.f[0][1];
        double *W = dst ? m.f[1][2] : m.f[0][2];
        const long long plane = (long long)m.nx * m.ny;
        const ColumnImages ci = column_images(g, a + 1, b + 1);
        const long long q0 = (a - 1) + (long long)m.nx * (b - 1);
        const long long qx0 = (a == m.nx ? 0 : a) + (long long)m.nx * (b - 1);  // (a % nx), row b-1
        const long long qy0 = (a - 1) + (long long)m.nx * (b == m.ny ? 0 : b);  // column a-1, row (b % ny)
        // Planes in batches of four (eight measured slower): every load of a batch before its stores, so that
        // they are in flight together (the stores write the batch's own cells and ghost images, which no load
        // of it reads).
        constexpr int kB = 4;
        for (int c0 = 1; c0 <= m.nz; c0 += kB) {
            double u[kB], v[kB], w[kB], f[kB], fx[kB], fy[kB], fz[kB];
#pragma unroll
            for (int d = 0; d < kB; ++d) {
                const int c = min(c0 + d, m.nz);
                const long long p = a + b * m.sj + c * m.sk;
                const long long o = plane * (c - 1);
                u[d] = U[p];
                v[d] = V[p];
                w[d] = W[p];
                f[d] = m.phi[q0 + o];
                fx[d] = m.phi[qx0 + o];
                fy[d] = m.phi[qy0 + o];
                fz[d] = m.phi[q0 + o + (c < m.nz ? plane : 0)];
            }
#pragma unroll
            for (int d = 0; d < kB; ++d) {
                const int c = c0 + d;
                if (c > m.nz) break;
                const long long p = a + b * m.sj + c * m.sk;
                const PlaneImages zi = plane_images(g, c + 1);
                put_column(U, p, u[d] - m.gx * (fx[d] - f[d]), ci, zi, 0);
                put_column(V, p, v[d] - m.gy * (fy[d] - f[d]), ci, zi, 0);
                if (c < m.nz) put_column(W, p, w[d] - m.gz * (fz[d] - f[d]), ci, zi, 1);
            }
        }
    }
}

// res = rhs - lap(phi): dense arrays, periodic x and y, Neumann z. Also the tile item's sum of phi, in the
// order of ens_mean_partial_kernel (each column over k, then the 256 columns in thread order), into
// part[base + it], for the mean the next correction subtracts.
__global__ __launch_bounds__(256) void ens_residual_kernel(const EnsMember *ms, const TileItem *items, int n,
                                                           double *part, int base)
{
    __shared__ double col[256];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const TileItem t = items[it];
        const EnsMember m = ms[t.member];
        const int i = t.bx * 64 + threadIdx.x, j = t.by * 4 + threadIdx.y;
        double acc = 0.0;
        if (i < m.nx && j < m.ny) {
            const int nx = m.nx, ny = m.ny, nz = m.nz;
            const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
            const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
            const long long plane = (long long)nx * ny;
            // phi at (i, j) over k is carried in registers (the k-1, k, k+1 window; Neumann ends repeat the
            // end plane), so only the x and y neighbours are read per plane.
#define P(x, y, z) m.phi[(x) + (long long)nx * (y) + plane * (z)]
            double c0 = P(i, j, 0), zm = c0;
#pragma unroll 4
            for (int k = 0; k < nz; ++k) {
                const double zp = k == nz - 1 ? c0 : P(i, j, k + 1);
                const double lap = (((P(im, j, k) + P(ip, j, k)) - 2.0 * c0) * m.rx +
                                    ((P(i, jm, k) + P(i, jp, k)) - 2.0 * c0) * m.ry) +
                                   ((zm + zp) - 2.0 * c0) * m.rz;
                const long long q = i + (long long)nx * j + plane * k;
                m.res[q] = m.rhs[q] - lap;
                acc = acc + c0;
                zm = c0;
                c0 = zp;
            }
#undef P
        }
        col[tid] = acc;
        __syncthreads();
        if (tid == 0) {
            double sum = 0.0;
            for (int c = 0; c < 256; ++c) sum = sum + col[c];
            part[base + it] = sum;
        }
        __syncthreads();
    }
}

// Partial sums of phi for the mean removal, one per tile item (64 x 4 columns over all planes): each column
// summed over k in order, then the 256 column sums in thread order. part[base + it] receives item it.
__global__ __launch_bounds__(256) void ens_mean_partial_kernel(const EnsMember *ms, const TileItem *items, int n,
                                                               double *part, int base)
{
    __shared__ double col[256];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const TileItem t = items[it];
        const EnsMember m = ms[t.member];
        const int i = t.bx * 64 + threadIdx.x, j = t.by * 4 + threadIdx.y;
        double acc = 0.0;
        if (i < m.nx && j < m.ny) {
            const long long plane = (long long)m.nx * m.ny;
            const double *ph = m.phi + i + (long long)m.nx * j;
#pragma unroll 8
            for (int k = 0; k < m.nz; ++k) acc = acc + ph[plane * k];
        }
        col[tid] = acc;
        __syncthreads();
        if (tid == 0) {
            double sum = 0.0;
            for (int c = 0; c < 256; ++c) sum = sum + col[c];
            part[base + it] = sum;
        }
        __syncthreads();
    }
}

// The mean of members m0 .. m0+count-1: each member's tile partials folded in tile order, stored in
// part[256] of the member.
__global__ __launch_bounds__(64) void ens_mean_fold_kernel(const EnsMember *ms, const int2 *tiles, const double *part,
                                                           int m0, int count)
{
    const int m = m0 + blockIdx.x * 64 + threadIdx.x;
    if (m >= m0 + count) return;
    const EnsMember mm = ms[m];
    const int2 r = tiles[m];
    double sum = 0.0;
    for (int q = 0; q < r.y; ++q) sum = sum + part[r.x + q];
    mm.part[256] = sum / ((double)mm.nx * mm.ny * mm.nz);
}

__global__ __launch_bounds__(256) void ens_subtract_kernel(const EnsMember *ms, const TileItem *items, int n)
{
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const TileItem t = items[it];
        const EnsMember m = ms[t.member];
        const int i = t.bx * 64 + threadIdx.x, j = t.by * 4 + threadIdx.y;
        if (i >= m.nx || j >= m.ny) continue;
        const double mean = m.part[256];
        const long long plane = (long long)m.nx * m.ny;
#pragma unroll 4
        for (int k = 0; k < m.nz; ++k) {
            double &v = m.phi[i + (long long)m.nx * j + plane * k];
            v = v - mean;
        }
    }
}

// QFAST: the mixed chain keeps mode (p, q) at column p*ny + q, the FP64 chain at p + q*nx.
template <class T, bool QFAST>
__global__ __launch_bounds__(kCols, 3) void ens_tridiag_kernel(const EnsMember *ms, const ColItem *items, int n)
{
    extern __shared__ double lds[];
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const ColItem t = items[it];
        const EnsMember m = ms[t.member];
        const long long plane = (long long)m.nx * m.ny;
        const long long col = (long long)t.block * kCols + threadIdx.x;
        if (col >= plane) continue;
        const double lam = QFAST ? m.lx[col / m.ny] + m.ly[col % m.ny] : m.lx[col % m.nx] + m.ly[col / m.nx];
        tridiag_column_prefetch<T, T>(static_cast<const T *>(m.w1), TriOut<T>{static_cast<T *>(m.w0), 0}, plane,
                                      m.nz, lam, m.rz, lds, threadIdx.x, kCols, col);
    }
}

template <int BM, int BN>
__global__ __launch_bounds__(kGemmThreads) __attribute__((amdgpu_waves_per_eu(gemm_small_waves(BM)))) void
ens_gemm_f64_kernel(const GemmCall *calls, const GemmItem *items, int n)
{
    __shared__ F64Lds<BM, BN> L;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const GemmItem t = items[it];
        gemm_f64_tile<BM, BN, 2, 2>(calls[t.call], L, t.mt, t.nt, t.b);
        __syncthreads();
    }
}

// The FP64 solve's transforms apply a small dense matrix (n <= kXformMax) along x or along y of a member's
// field. Side 0 (x): Y = alpha Q X + beta Y, X seen as n x (ny nz) with x fastest (column c = j + ny k is
// contiguous), Q column-major n x n. Side 1 (y): Y = alpha X R + beta Y, X seen as (nx nz) x n, row r =
// i + nx k at i + nx ny k, column m at stride nx; Q holds R transposed, so that an MFMA operand fragment is
// sixteen consecutive elements in both sides.
struct XformCall {
    const double *Q;
    const double *X;
    double *Y;
    int n, nx, ny, nz;
    double alpha, beta;
};
constexpr int kXformMax = 128;

// The transform of one block of 64 columns (x) or 64 rows (y) of one call per workgroup of 256 threads (four
// waves). The block of X is staged in LDS, the transform axis padded with zeros to a multiple of 16. Wave w
// takes the block's columns (rows) 16 w .. 16 w + 15 and every transform tile (n / 16 rounded up, a template
// argument so that the loop holds no branch and the MFMAs accumulate in place), and K runs in ascending steps
// of four, one v_mfma_f64_16x16x4f64 per step and tile. The operand roles and the order
// of the MFMAs over K are those of gemm_f64_tile on the same call, so the results are the same bits; the
// output index of a lane's first operand is the store's contiguous one. The Q fragments come from memory
// (the matrix is read by every block of the member, so from L2) one K slice ahead of their MFMAs.
constexpr int kXformBlock = 64;
constexpr int kQPad = 15;  // zero rows after each FP64 basis matrix: K runs to the next multiple of 16

template <int SIDE>
struct Xform {
    static constexpr int BB = kXformBlock;
    // The staged block as [block index][transform index] on both sides, so that a wave's K steps are 32 bytes
    // apart in LDS (all within one DS offset range of its base address).
    static constexpr int LD = kXformMax + 1;
    static constexpr int LDS = BB * LD;  // doubles of one staged block
    static constexpr int kOut = 0x7fff0000;  // a byte offset past every buffer below: such loads return zero
    static constexpr int MG = 256 / BB;      // y side: BB rows per pass, MG values of m at once
    static constexpr int NS = SIDE == 0 ? (BB / 4) * (kXformMax / 64) : kXformMax / MG;

    // The block in two halves, each with every load first (a buffer load past the block's bytes returns zero,
    // so the padding needs no branch), then its LDS stores.
    static __device__ __forceinline__ void stage(const XformCall &g, int blk, double *Xs, int t)
    {
        const int n = g.n, kp = (n + 15) & ~15, b0 = blk * BB;
        const int lane = t & 63, wave = t >> 6;
        const int plane = g.nx * g.ny;
        const int total = SIDE == 0 ? g.ny * g.nz : g.nx * g.nz;
        constexpr int H = NS / 2;
        double v[H];
        if (SIDE == 0) {
            // Column c = wave + 4 s, rows k = lane + 64 h: the half is h.
            const int cb = min(BB, total - b0);
            const __amdgpu_buffer_rsrc_t rX = field_rsrc(g.X + (long long)b0 * n, cb * n * 8);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = lane + 64 * h;
#pragma unroll
                for (int s = 0; s < BB / 4; ++s) v[s] = buffer_load(rX, k < n ? ((wave + 4 * s) * n + k) * 8 : kOut, 0);
#pragma unroll
                for (int s = 0; s < BB / 4; ++s)
                    if (k < kp) Xs[(wave + 4 * s) * LD + k] = v[s];
            }
        } else {
            // Row rl, m = mg + MG s: the half is s / H.
            const __amdgpu_buffer_rsrc_t rX = field_rsrc(g.X, plane * g.nz * 8);
            const int rl = t % BB, mg = t / BB;
            const int r = b0 + rl;
            const int base = r < total ? (r % g.nx) + (r / g.nx) * plane : -1;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int s = 0; s < H; ++s) {
                    const int m = mg + MG * (s + H * h);
                    v[s] = buffer_load(rX, base >= 0 && m < n ? (base + m * g.nx) * 8 : kOut, 0);
                }
#pragma unroll
                for (int s = 0; s < H; ++s) {
                    const int m = mg + MG * (s + H * h);
                    if (m < kp) Xs[rl * LD + m] = v[s];
                }
            }
        }
    }

    // The MFMAs over the staged block and the stores of the output block. Lane (q16, q4) of transform tile x
    // reads Q(16 x + q16, kk + q4) at the lane's byte offset (16 x + q16 + q4 n) 8 plus the step's scalar
    // offset kk n 8; rows past n read zero (their lane offset is out of range), and K past n reads the kQPad
    // zero rows that follow the matrix and meets zero rows of the staged block.
    // KS: K slices of 16 (n / 16 rounded up, also the number of transform tiles, all of them this wave's).
    template <int KS>
    static __device__ __forceinline__ void multiply(const XformCall &g, int blk, const double *Xs, int lane, int wave)
    {
        constexpr int XT = KS;
        const int n = g.n, b0 = blk * BB;
        const int plane = g.nx * g.ny;
        const int total = SIDE == 0 ? g.ny * g.nz : g.nx * g.nz;
        const int q16 = lane % 16, q4 = lane / 16;
        const __amdgpu_buffer_rsrc_t rQ = field_rsrc(g.Q, n * (n + kQPad) * 8);
        int qoff[XT];
#pragma unroll
        for (int x = 0; x < XT; ++x) {
            const int row = 16 * x + q16;
            qoff[x] = row < n ? (row + q4 * n) * 8 : kOut;
        }
        const int bcol = 16 * wave + q16;  // this lane's index within the block
        const double *xp = Xs + bcol * LD + q4;
        constexpr int XS = 4;  // LDS step of one K step
        f64x4 acc[XT];
#pragma unroll
        for (int x = 0; x < XT; ++x) acc[x] = f64x4{0.0, 0.0, 0.0, 0.0};
        // K runs over KS slices of 16 (the tiled contraction's slices, zeros past n), each slice's Q fragments
        // loaded while the previous slice multiplies; the trip counts are constants, so the loop unrolls into
        // straight code with exact waits on the loads.
        const int slice_bytes = 16 * n * 8;
        double qa[4][XT], qb[4][XT];
        auto load = [&](double (&q)[4][XT], int sl) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int x = 0; x < XT; ++x) q[u][x] = buffer_load(rQ, qoff[x], sl * slice_bytes + u * 4 * n * 8);
        };
        auto run = [&](const double (&q)[4][XT], int sl) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double xf = xp[(4 * sl + u) * XS];
#pragma unroll
                for (int x = 0; x < XT; ++x)
                    acc[x] = SIDE == 0 ? __builtin_amdgcn_mfma_f64_16x16x4f64(xf, q[u][x], acc[x], 0, 0, 0)
                                       : __builtin_amdgcn_mfma_f64_16x16x4f64(q[u][x], xf, acc[x], 0, 0, 0);
            }
        };
        load(qa, 0);
#pragma unroll
        for (int sl = 0; sl < KS; sl += 2) {
            if (sl + 1 < KS) load(qb, sl + 1);
            run(qa, sl);
            if (sl + 1 < KS) {
                if (sl + 2 < KS) load(qa, sl + 2);
                run(qb, sl + 1);
            }
        }
        // Lane l, element e of tile x: side 0 holds Y(transform index 16 x + q16, block index 16 w + q4 + 4e);
        // side 1 holds Y(block index 16 w + q16, transform index 16 x + q4 + 4e).
        long long rowbase = 0;
        bool rok = true;
        if (SIDE == 1) {
            const int r = b0 + bcol;
            rok = r < total;
            rowbase = rok ? (long long)(r % g.nx) + (long long)(r / g.nx) * plane : 0;
        }
#pragma unroll
        for (int x = 0; x < XT; ++x) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                long long idx;
                bool ok;
                if (SIDE == 0) {
                    const int m = 16 * x + q16;
                    const int c = b0 + 16 * wave + q4 + 4 * e;
                    ok = m < n && c < total;
                    idx = m + (long long)c * n;
                } else {
                    const int jj = 16 * x + q4 + 4 * e;
                    ok = rok && jj < n;
                    idx = rowbase + (long long)jj * g.nx;
                }
                if (ok) {
                    double v = g.alpha * acc[x][e];
                    if (g.beta != 0.0) v = v + g.beta * g.Y[idx];
                    g.Y[idx] = v;
                }
            }
        }
    }

    static __device__ __forceinline__ void compute(const XformCall &g, int blk, const double *Xs, int t)
    {
        const int lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
        switch ((g.n + 15) / 16) {
        case 1: multiply<1>(g, blk, Xs, lane, wave); break;
        case 2: multiply<2>(g, blk, Xs, lane, wave); break;
        case 3: multiply<3>(g, blk, Xs, lane, wave); break;
        case 4: multiply<4>(g, blk, Xs, lane, wave); break;
        case 5: multiply<5>(g, blk, Xs, lane, wave); break;
        case 6: multiply<6>(g, blk, Xs, lane, wave); break;
        case 7: multiply<7>(g, blk, Xs, lane, wave); break;
        default: multiply<8>(g, blk, Xs, lane, wave); break;
        }
    }
};

// One block per workgroup: stage, then multiply. The LDS block allows two workgroups per CU; declaring that
// occupancy holds a wave to 256 registers, where the MFMAs accumulate in place.
template <int SIDE>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(2, 2))) void
ens_xform_kernel(const XformCall *__restrict__ calls, const GemmItem *__restrict__ items)
{
    using X = Xform<SIDE>;
    __shared__ double Xs[X::LDS];
    const GemmItem it = items[blockIdx.x];
    const XformCall g = calls[it.call];
    X::stage(g, it.mt, Xs, threadIdx.x);
    __syncthreads();
    X::compute(g, it.mt, Xs, threadIdx.x);
}

template <class TA, class TB, class TO, int BM, int BN, bool TOUT>
__global__ __launch_bounds__(kGemmThreads) __attribute__((amdgpu_waves_per_eu(gemm_small_waves(BM)))) void
ens_gemm_bf16_kernel(const GemmCall *calls, const GemmItem *items, int n)
{
    __shared__ Bf16TnLds<BM, BN, 32> L;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const GemmItem t = items[it];
        gemm_bf16x3_tn_tile<TA, TB, TO, BM, BN, TOUT>(calls[t.call], L, t.mt, t.nt, t.b);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------ host
static std::size_t up256(std::size_t x) { return (x + 255) & ~std::size_t(255); }

// The basis of the periodic 1D second difference (same construction as the pressure provider).
// The periodic basis of an axis of n points as a table of 4n values: cosines r2 cos(2 pi j / n), sines
// r2 sin(2 pi j / n), the constant r1, and the alternating r1 (j = 0), -r1 (j = n / 2), with r1 = 1/sqrt(n) and
// r2 = sqrt(2/n). Mode 0 is the constant, modes 2f - 1 and 2f (0 < 2f < n) the cosine and sine of frequency
// f, and mode n - 1 of an even n the alternating vector: B(i, mode) = table[section n + (f i mod n)].
static void basis_table(int n, std::vector<double> &table)
{
    const double pi = 3.14159265358979323846;
    const double r2 = std::sqrt(2.0 / n), r1 = 1.0 / std::sqrt((double)n);
    table.assign((std::size_t)4 * n, 0.0);
    for (int j = 0; j < n; ++j) {
        table[j] = r2 * std::cos(2.0 * pi * j / n);
        table[n + j] = r2 * std::sin(2.0 * pi * j / n);
        table[2 * n + j] = r1;
        table[3 * n + j] = j == 0 ? r1 : -r1;
    }
}

// The basis matrix Q(i, mode) (column-major, n x n) from the table, and the eigenvalues of the periodic
// second difference over spacing h.
static void periodic_basis(int n, double h, const std::vector<double> &table, std::vector<double> &Q,
                           std::vector<double> &lam)
{
    Q.assign((std::size_t)n * n, 0.0);
    lam.assign(n, 0.0);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) Q[i] = table[2 * n];
    for (int f = 1; 2 * f < n; ++f) {
        const double l = (2.0 * std::cos(2.0 * pi * f / n) - 2.0) / (h * h);
        for (int i = 0; i < n; ++i) {
            const int j = (int)((long long)f * i % n);
            Q[(std::size_t)(2 * f - 1) * n + i] = table[j];
            Q[(std::size_t)(2 * f) * n + i] = table[n + j];
        }
        lam[2 * f - 1] = l;
        lam[2 * f] = l;
    }
    if (n % 2 == 0) {
        for (int i = 0; i < n; ++i) Q[(std::size_t)(n - 1) * n + i] = table[3 * n + (int)((long long)(n / 2) * i % n)];
        lam[n - 1] = -4.0 / (h * h);
    }
}

enum Stage { kLes, kDiv, kG1, kG2, kTri, kG3, kG4, kRes, kG1c, kG4c, kMean, kSub, kProj, kFold, kStageCount };

struct Range {
    int start, count;
};

}  // namespace ftc

using namespace ftc;

struct ftc_ensemble {
    int mode = 0, count = 0, solves = 3;
    bool projection = true;  // false: the step is the LES stage alone
    double dx = 0, dy = 0, dz = 0, dt = 0, cs = 0;
    std::vector<int> nx, ny, nz;
    // tuning
    int les_kchunk = 24, gemm_tile = 1, blocks_per_cu = 4, cus = 256;
    bool les_iso = true;         // equal spacings: the LES kernels hold one coefficient per kind
    int les_policy = 0;          // LES tile widths: see les_width
    std::vector<int> les_w;      // the width each member takes
    int n_les32 = 0;             // LES items of 32-wide members, after the n_les - n_les32 of 64-wide ones
    // device
    void *arena = nullptr;
    std::size_t arena_bytes = 0;
    hipStream_t stream = nullptr;
    std::vector<EnsMember> members;  // host copy
    EnsMember *d_members = nullptr;
    LesArgs *d_args[2] = {nullptr, nullptr};
    std::vector<LesArgs> args[2];
    GemmCall *d_calls = nullptr;  // kG1.. stage-major: calls[stage_slot * count + member]
    XformCall *d_xcalls = nullptr;  // the same calls as axis transforms (FP64 mode), same indexing
    bool xform[6] = {};             // transform slot runs on ens_xform_kernel (FP64, every member's n small)
    TileItem *d_les = nullptr, *d_tiles = nullptr;
    TileItem *d_les_all = nullptr;    // the same items for the all-member launches, longest first per width
    TileItem *d_tiles_all = nullptr;  // the tiles longest first, for the launches that keep no per-tile sums
    ColItem *d_cols_all = nullptr;    // the z-solve column blocks longest first
    GemmItem *d_gemm[6] = {};
    ColItem *d_cols = nullptr;
    double *d_tile_part = nullptr;  // one mean partial per tile item
    int2 *d_tile_ranges = nullptr;  // (first tile item, count) per member
    ColItem *d_flat = nullptr, *d_flat_all = nullptr;  // blocks of kFlat plane columns; all: longest first
    std::vector<Range> r_les, r_tiles, r_gemm[6], r_cols, r_flat;
    int n_les = 0, n_tiles = 0, n_gemm[6] = {}, n_cols = 0, n_flat = 0;
    void *items_arena = nullptr;
    int nzmax = 0;
    int parity = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int64_t launches = 0;
};

namespace ftc {

static bool mixed(const ftc_ensemble &e) { return e.mode == FTC_ENSEMBLE_MIXED_BATCHED; }

// The contraction calls of member m for the six transform stages (G1, G2, G3, G4, G1c, G4c).
static void member_calls(const ftc_ensemble &e, const EnsMember &m, const double *Qx, const double *QxT,
                         const double *Qy, const double *QyT, const float *Qxf, const float *QxTf, const float *Qyf,
                         const float *QyTf, GemmCall out[6])
{
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    const long long plane = (long long)nx * ny;
    if (mixed(e)) {
        // Every operand K-contiguous (the pressure provider's split-bf16 chain): W0 and W1 keep the mode
        // index q fastest between G1 and G3.
        const int ak = kGemmAKContiguous, at = kGemmAKContiguous | kGemmTransposedC;
        const int f = kFloat, d = kDouble;
        out[0] = GemmCall{Qxf, m.rhs, m.w0, f, d, f, nx, ny, nx, nz, nx, nx, ny, 0, plane, plane, 1.0, 0.0, at};
        out[1] = GemmCall{m.w0, Qyf, m.w1, f, f, f, nx, ny, ny, nz, ny, ny, ny, plane, 0, plane, 1.0, 0.0, at};
        out[2] = GemmCall{m.w0, QyTf, m.w1, f, f, f, nx, ny, ny, nz, ny, ny, nx, plane, 0, plane, 1.0, 0.0, ak};
        out[3] = GemmCall{QxTf, m.w1, m.phi, f, f, d, nx, ny, nx, nz, nx, nx, nx, 0, plane, plane, 1.0, 0.0, ak};
        out[4] = GemmCall{Qxf, m.res, m.w0, f, d, f, nx, ny, nx, nz, nx, nx, ny, 0, plane, plane, 1.0, 0.0, at};
        out[5] = GemmCall{QxTf, m.w1, m.phi, f, f, d, nx, ny, nx, nz, nx, nx, nx, 0, plane, plane, 1.0, 1.0, ak};
        out[5].shift = m.part + 256;  // the member's mean of phi from the preceding residual pass
        return;
    }
    const int d = kDouble;
    out[0] = GemmCall{QxT, m.rhs, m.w0, d, d, d, nx, ny, nx, nz, nx, nx, nx, 0, plane, plane, 1.0, 0.0};
    out[1] = GemmCall{m.w0, Qy, m.w1, d, d, d, nx, ny, ny, nz, nx, ny, nx, plane, 0, plane, 1.0, 0.0};
    out[2] = GemmCall{m.w0, QyT, m.w1, d, d, d, nx, ny, ny, nz, nx, ny, nx, plane, 0, plane, 1.0, 0.0};
    out[3] = GemmCall{Qx, m.w1, m.phi, d, d, d, nx, ny, nx, nz, nx, nx, nx, 0, plane, plane, 1.0, 0.0};
    out[4] = GemmCall{QxT, m.res, m.w0, d, d, d, nx, ny, nx, nz, nx, nx, nx, 0, plane, plane, 1.0, 0.0};
    out[5] = GemmCall{Qx, m.w1, m.phi, d, d, d, nx, ny, nx, nz, nx, nx, nx, 0, plane, plane, 1.0, 1.0};
}

// The FP64 mode's transform calls as axis transforms: slots G1, G4, G1c and G4c along x, G2 and G3 along y.
// G1 and G1c multiply by Qx^T, G4 and G4c by Qx (from the left); G2 by Qy and G3 by Qy^T (from the right, so
// their Q is Qy^T and Qy).
static void member_xforms(const EnsMember &m, const double *Qx, const double *QxT, const double *Qy,
                          const double *QyT, XformCall out[6])
{
    const double *rhs = m.rhs, *res = m.res;
    double *w0 = static_cast<double *>(m.w0), *w1 = static_cast<double *>(m.w1);
    auto x = [&](const double *Q, const double *X, double *Y, double beta) {
        return XformCall{Q, X, Y, m.nx, m.nx, m.ny, m.nz, 1.0, beta};
    };
    auto y = [&](const double *Qt, const double *X, double *Y) {
        return XformCall{Qt, X, Y, m.ny, m.nx, m.ny, m.nz, 1.0, 0.0};
    };
    out[0] = x(QxT, rhs, w0, 0.0);
    out[1] = y(QyT, w0, w1);
    out[2] = y(Qy, w0, w1);
    out[3] = x(Qx, w1, m.phi, 0.0);
    out[4] = x(QxT, res, w0, 0.0);
    out[5] = x(Qx, w1, m.phi, 1.0);
}

// Transform slot s runs along y (G2, G3) or along x (the others).
static bool slot_along_y(int s) { return s == 1 || s == 2; }

static int tile_bm(int tile) { return tile == 0 ? 64 : 32; }

// Builds (or rebuilds, after tuning) every work list and uploads it.
static int build_items(ftc_ensemble &e)
{
    std::vector<TileItem> les, tiles;
    std::vector<GemmItem> gemm[6];
    std::vector<ColItem> cols, flat;
    e.n_les32 = 0;
    e.r_tiles.clear();
    e.r_cols.clear();
    e.r_flat.clear();
    for (auto &r : e.r_gemm) r.clear();
    const int bm = tile_bm(e.gemm_tile);
    // LES tiles: the members of width 64 first, then those of width 32, each member's tiles contiguous.
    e.les_w.assign(e.count, 64);
    e.r_les.assign(e.count, Range{0, 0});
    for (int width : {64, 32})
        for (int m = 0; m < e.count; ++m) {
            if (les_width(e.les_policy, e.nx[m]) != width) continue;
            e.les_w[m] = width;
            const int nx = e.nx[m], ny = e.ny[m], nz = e.nz[m], rows = kStepThreads / width;
            const int kchunk = les_chunk(e.les_kchunk, nz);
            Range r{(int)les.size(), 0};
            for (int bz = 0; bz * kchunk < nz; ++bz)
                for (int by = 0; by < (ny + rows - 1) / rows; ++by)
                    for (int bx = 0; bx < (nx + width - 1) / width; ++bx) les.push_back({m, bx, by, bz});
            r.count = (int)les.size() - r.start;
            e.r_les[m] = r;
            if (width == 32) e.n_les32 += r.count;
        }
    for (int m = 0; m < e.count; ++m) {
        const int nx = e.nx[m], ny = e.ny[m], nz = e.nz[m];
        Range rt{(int)tiles.size(), 0};
        for (int by = 0; by < (ny + 3) / 4; ++by)
            for (int bx = 0; bx < (nx + 63) / 64; ++bx) tiles.push_back({m, bx, by, 0});
        rt.count = (int)tiles.size() - rt.start;
        e.r_tiles.push_back(rt);
        for (int s = 0; s < 6; ++s) {
            Range rg{(int)gemm[s].size(), 0};
            if (e.xform[s]) {
                // Axis transform: blocks of 64 columns (x) or 64 rows (y).
                const long long extent = slot_along_y(s) ? (long long)nx * nz : (long long)ny * nz;
                for (int b = 0; (long long)b * kXformBlock < extent; ++b) gemm[s].push_back({s * e.count + m, b, 0, 0});
            } else {
                const int mtiles = (nx + bm - 1) / bm, ntiles = (ny + bm - 1) / bm;
                for (int b = 0; b < nz; ++b)
                    for (int nt = 0; nt < ntiles; ++nt)
                        for (int mt = 0; mt < mtiles; ++mt) gemm[s].push_back({s * e.count + m, mt, nt, b});
            }
            rg.count = (int)gemm[s].size() - rg.start;
            e.r_gemm[s].push_back(rg);
        }
        Range rc{(int)cols.size(), 0};
        const long long plane = (long long)nx * ny;
        for (int b = 0; b * (long long)kCols < plane; ++b) cols.push_back({m, b});
        rc.count = (int)cols.size() - rc.start;
        e.r_cols.push_back(rc);
        Range rf{(int)flat.size(), 0};
        for (int b = 0; b * (long long)kFlat < plane; ++b) flat.push_back({m, b});
        rf.count = (int)flat.size() - rf.start;
        e.r_flat.push_back(rf);
    }
    e.n_les = (int)les.size();
    e.n_tiles = (int)tiles.size();
    e.n_cols = (int)cols.size();
    e.n_flat = (int)flat.size();
    // The all-member launches give every item a workgroup of its own, and the dispatcher starts them in
    // order as slots free up: so that the last ones to start are short, the items of each width are taken
    // longest first (most planes), each member's items staying together.
    std::vector<TileItem> les_all(les);
    auto planes = [&](const TileItem &t) {
        const int kc = les_chunk(e.les_kchunk, e.nz[t.member]);
        return std::min(kc, e.nz[t.member] - t.bz * kc);
    };
    const auto split = les_all.begin() + (e.n_les - e.n_les32);
    for (auto [lo, hi] : {std::make_pair(les_all.begin(), split), std::make_pair(split, les_all.end())})
        std::stable_sort(lo, hi, [&](const TileItem &a, const TileItem &b) { return planes(a) > planes(b); });
    std::vector<ColItem> cols_all(cols);
    std::stable_sort(cols_all.begin(), cols_all.end(),
                     [&](const ColItem &a, const ColItem &b) { return e.nz[a.member] > e.nz[b.member]; });
    std::vector<ColItem> flat_all(flat);
    std::stable_sort(flat_all.begin(), flat_all.end(),
                     [&](const ColItem &a, const ColItem &b) { return e.nz[a.member] > e.nz[b.member]; });
    std::vector<TileItem> tiles_all(tiles);
    std::stable_sort(tiles_all.begin(), tiles_all.end(),
                     [&](const TileItem &a, const TileItem &b) { return e.nz[a.member] > e.nz[b.member]; });
    std::size_t bytes = 2 * up256(les.size() * sizeof(TileItem)) + 2 * up256(tiles.size() * sizeof(TileItem)) +
                        2 * up256(cols.size() * sizeof(ColItem)) + 2 * up256(flat.size() * sizeof(ColItem)) +
                        up256(tiles.size() * sizeof(double)) + up256(e.count * sizeof(int2));
    for (int s = 0; s < 6; ++s) {
        e.n_gemm[s] = (int)gemm[s].size();
        bytes += up256(gemm[s].size() * sizeof(GemmItem));
    }
    if (e.items_arena) (void)hipFree(e.items_arena);
    e.items_arena = nullptr;
    if (hipMalloc(&e.items_arena, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return eerr(FORT_SCOPE_RESOURCE, "work-list allocation failed");
    }
    char *p = static_cast<char *>(e.items_arena);
    auto put = [&](const void *src, std::size_t n) {
        char *dst = p;
        if (n && hipMemcpy(dst, src, n, hipMemcpyHostToDevice) != hipSuccess) return (char *)nullptr;
        p += up256(n);
        return dst;
    };
    e.d_les = (TileItem *)put(les.data(), les.size() * sizeof(TileItem));
    e.d_les_all = (TileItem *)put(les_all.data(), les_all.size() * sizeof(TileItem));
    e.d_tiles = (TileItem *)put(tiles.data(), tiles.size() * sizeof(TileItem));
    e.d_tiles_all = (TileItem *)put(tiles_all.data(), tiles_all.size() * sizeof(TileItem));
    e.d_cols = (ColItem *)put(cols.data(), cols.size() * sizeof(ColItem));
    e.d_cols_all = (ColItem *)put(cols_all.data(), cols_all.size() * sizeof(ColItem));
    e.d_flat = (ColItem *)put(flat.data(), flat.size() * sizeof(ColItem));
    e.d_flat_all = (ColItem *)put(flat_all.data(), flat_all.size() * sizeof(ColItem));
    std::vector<int2> ranges(e.count);
    for (int m = 0; m < e.count; ++m) ranges[m] = make_int2(e.r_tiles[m].start, e.r_tiles[m].count);
    e.d_tile_ranges = (int2 *)put(ranges.data(), ranges.size() * sizeof(int2));
    e.d_tile_part = (double *)p;
    p += up256(tiles.size() * sizeof(double));
    for (int s = 0; s < 6; ++s) e.d_gemm[s] = (GemmItem *)put(gemm[s].data(), gemm[s].size() * sizeof(GemmItem));
    // LES arguments carry the k chunk.
    for (int par = 0; par < 2; ++par) {
        for (int m = 0; m < e.count; ++m) e.args[par][m].kchunk = les_chunk(e.les_kchunk, e.nz[m]);
        if (hipMemcpy(e.d_args[par], e.args[par].data(), e.count * sizeof(LesArgs), hipMemcpyHostToDevice) != hipSuccess)
            return eerr(FORT_SCOPE_RESOURCE, "argument upload failed");
    }
    return FORT_SCOPE_OK;
}

static unsigned grid_for(const ftc_ensemble &e, int items)
{
    const long long cap = (long long)e.cus * e.blocks_per_cu;
    return (unsigned)std::max<long long>(1, std::min<long long>(items, cap));
}

// Launches one stage, for all members (member < 0, persistent grid) or for one member (its own grid).
static void launch_stage(ftc_ensemble &e, int stage, int member, int parity)
{
    const bool all = member < 0;
    auto pick = [&](const std::vector<Range> &r, int total, int *start) {
        if (all) {
            *start = 0;
            return total;
        }
        *start = r[member].start;
        return r[member].count;
    };
    const bool mx = mixed(e);
    const int src = parity, dst = 1 - parity;
    int start = 0, n = 0;
    hipStream_t s = e.stream;
    switch (stage) {
    case kLes: {
        // One launch per width: the 64-wide members' tiles, then the 32-wide members'.
        auto launch = [&](int width, int first, int count) {
            if (!count) return;
            const dim3 grid(count);
            const TileItem *items = (all ? e.d_les_all : e.d_les) + first;
            const void *fn = width == 32 ? (e.les_iso ? reinterpret_cast<const void *>(&ens_les_kernel<32, true>)
                                                      : reinterpret_cast<const void *>(&ens_les_kernel<32, false>))
                                         : (e.les_iso ? reinterpret_cast<const void *>(&ens_les_kernel<64, true>)
                                                      : reinterpret_cast<const void *>(&ens_les_kernel<64, false>));
            const LesArgs *args = e.d_args[parity];
            void *kargs[] = {&args, &items};
            (void)hipLaunchKernel(fn, grid, dim3(kStepThreads), kargs, 0, s);
        };
        if (all) {
            launch(64, 0, e.n_les - e.n_les32);
            launch(32, e.n_les - e.n_les32, e.n_les32);
        } else {
            launch(e.les_w[member], e.r_les[member].start, e.r_les[member].count);
        }
        break;
    }
    case kDiv:
        n = pick(e.r_flat, e.n_flat, &start);
        if (n) hipLaunchKernelGGL(ens_divergence_kernel, dim3(all ? grid_for(e, n) : n), dim3(kFlat), 0, s,
                                  e.d_members, (all ? e.d_flat_all : e.d_flat + start), n, dst);
        break;
    case kProj:
        n = pick(e.r_flat, e.n_flat, &start);
        if (n) hipLaunchKernelGGL(ens_projection_kernel, dim3(all ? grid_for(e, n) : n), dim3(kFlat), 0, s,
                                  e.d_members, e.d_args[parity], (all ? e.d_flat_all : e.d_flat + start), n, dst);
        break;
    case kRes:
        n = pick(e.r_tiles, e.n_tiles, &start);
        if (n) hipLaunchKernelGGL(ens_residual_kernel, dim3(all ? grid_for(e, n) : n), dim3(64, 4), 0, s,
                                  e.d_members, e.d_tiles + start, n, e.d_tile_part, start);
        break;
    case kFold:
        hipLaunchKernelGGL(ens_mean_fold_kernel, dim3(all ? (e.count + 63) / 64 : 1), dim3(64), 0, s, e.d_members,
                           e.d_tile_ranges, e.d_tile_part, all ? 0 : member, all ? e.count : 1);
        break;
    case kSub:
        n = pick(e.r_tiles, e.n_tiles, &start);
        if (n) hipLaunchKernelGGL(ens_subtract_kernel, dim3(all ? grid_for(e, n) : n), dim3(64, 4), 0, s,
                                  e.d_members, (all ? e.d_tiles_all : e.d_tiles + start), n);
        break;
    case kMean:
        n = pick(e.r_tiles, e.n_tiles, &start);
        if (n) hipLaunchKernelGGL(ens_mean_partial_kernel, dim3(all ? grid_for(e, n) : n), dim3(64, 4), 0, s,
                                  e.d_members, e.d_tiles + start, n, e.d_tile_part, start);
        hipLaunchKernelGGL(ens_mean_fold_kernel, dim3(all ? (e.count + 63) / 64 : 1), dim3(64), 0, s, e.d_members,
                           e.d_tile_ranges, e.d_tile_part, all ? 0 : member, all ? e.count : 1);
        break;
    case kTri: {
        n = pick(e.r_cols, e.n_cols, &start);
        const std::size_t lds = 2 * (std::size_t)tridiag_chunks(e.nzmax) * kCols * sizeof(double);
        const ColItem *cols = all ? e.d_cols_all : e.d_cols + start;
        if (n) {
            if (mx) hipLaunchKernelGGL((ens_tridiag_kernel<float, true>), dim3(all ? grid_for(e, n) : n), dim3(kCols),
                                       lds, s, e.d_members, cols, n);
            else hipLaunchKernelGGL((ens_tridiag_kernel<double, false>), dim3(all ? grid_for(e, n) : n), dim3(kCols),
                                    lds, s, e.d_members, cols, n);
        }
        break;
    }
    default: {
        // Transform stages: slot 0..5 = G1, G2, G3, G4, G1c, G4c.
        const int slot = stage == kG1 ? 0 : stage == kG2 ? 1 : stage == kG3 ? 2 : stage == kG4 ? 3 : stage == kG1c ? 4 : 5;
        n = pick(e.r_gemm[slot], e.n_gemm[slot], &start);
        if (!n) break;
        const dim3 grid(all ? grid_for(e, n) : n), block(kGemmThreads);
        const GemmItem *items = e.d_gemm[slot] + start;
        const bool big = e.gemm_tile == 0;
        if (e.xform[slot]) {
            if (slot_along_y(slot)) hipLaunchKernelGGL(ens_xform_kernel<1>, dim3(n), block, 0, s, e.d_xcalls, items);
            else hipLaunchKernelGGL(ens_xform_kernel<0>, dim3(n), block, 0, s, e.d_xcalls, items);
        } else if (!mx) {
            if (big) hipLaunchKernelGGL((ens_gemm_f64_kernel<64, 64>), grid, block, 0, s, e.d_calls, items, n);
            else hipLaunchKernelGGL((ens_gemm_f64_kernel<32, 32>), grid, block, 0, s, e.d_calls, items, n);
        } else if (slot == 0 || slot == 4) {
            if (big) hipLaunchKernelGGL((ens_gemm_bf16_kernel<float, double, float, 64, 64, true>), grid, block, 0, s, e.d_calls, items, n);
            else hipLaunchKernelGGL((ens_gemm_bf16_kernel<float, double, float, 32, 32, true>), grid, block, 0, s, e.d_calls, items, n);
        } else if (slot == 1) {
            if (big) hipLaunchKernelGGL((ens_gemm_bf16_kernel<float, float, float, 64, 64, true>), grid, block, 0, s, e.d_calls, items, n);
            else hipLaunchKernelGGL((ens_gemm_bf16_kernel<float, float, float, 32, 32, true>), grid, block, 0, s, e.d_calls, items, n);
        } else if (slot == 2) {
            if (big) hipLaunchKernelGGL((ens_gemm_bf16_kernel<float, float, float, 64, 64, false>), grid, block, 0, s, e.d_calls, items, n);
            else hipLaunchKernelGGL((ens_gemm_bf16_kernel<float, float, float, 32, 32, false>), grid, block, 0, s, e.d_calls, items, n);
        } else {
            if (big) hipLaunchKernelGGL((ens_gemm_bf16_kernel<float, float, double, 64, 64, false>), grid, block, 0, s, e.d_calls, items, n);
            else hipLaunchKernelGGL((ens_gemm_bf16_kernel<float, float, double, 32, 32, false>), grid, block, 0, s, e.d_calls, items, n);
        }
    }
    }
}

// The stage sequence of one step.
static std::vector<int> step_stages(const ftc_ensemble &e)
{
    if (!e.projection) return {kLes};
    std::vector<int> st{kLes, kDiv, kG1, kG2, kTri, kG3, kG4};
    if (mixed(e)) {
        // Each correction's residual pass also sums phi per tile; the fold gives each member's mean, which
        // the correction's G4 subtracts (the Laplacian does not see a constant, and the split-bf16
        // transforms leak one). After the last correction only its own, much smaller leak is left. A single
        // application removes the mean explicitly.
        for (int c = 1; c < e.solves; ++c)
            for (int x : {kRes, kFold, kG1c, kG2, kTri, kG3, kG4c}) st.push_back(x);
        if (e.solves == 1) {
            st.push_back(kMean);
            st.push_back(kSub);
        }
    }
    st.push_back(kProj);
    return st;
}

static int enqueue_step(ftc_ensemble &e, int parity)
{
    const std::vector<int> st = step_stages(e);
    if (e.mode == FTC_ENSEMBLE_FP64_PER_MEMBER) {
        for (int m = 0; m < e.count; ++m)
            for (int stage : st) launch_stage(e, stage, m, parity);
        e.launches = (int64_t)st.size() * e.count;
    } else {
        for (int stage : st) launch_stage(e, stage, -1, parity);
        const bool two_widths = e.n_les32 > 0 && e.n_les32 < e.n_les;
        e.launches = (int64_t)st.size() + (two_widths ? 1 : 0);
    }
    if (hipGetLastError() != hipSuccess) return eerr(FORT_SCOPE_EXECUTION, "stage launch failed");
    return FORT_SCOPE_OK;
}

static void drop_graphs(ftc_ensemble &e)
{
    for (auto &g : e.graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// Captures one step per parity into an executable graph (batched modes).
static int capture(ftc_ensemble &e)
{
    drop_graphs(e);
    for (int par = 0; par < 2; ++par) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return eerr(FORT_SCOPE_RESOURCE, "graph capture failed to start");
        int st = enqueue_step(e, par);
        hipError_t end = hipStreamEndCapture(e.stream, &g);
        if (st || end != hipSuccess) {
            (void)hipGetLastError();
            if (g) (void)hipGraphDestroy(g);
            return st ? st : eerr(FORT_SCOPE_RESOURCE, "graph capture failed");
        }
        hipError_t inst = hipGraphInstantiate(&e.graph[par], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (inst != hipSuccess) {
            (void)hipGetLastError();
            return eerr(FORT_SCOPE_RESOURCE, "graph instantiation failed");
        }
    }
    return FORT_SCOPE_OK;
}

static float time_ms(ftc_ensemble &e, const std::function<void()> &work, int reps)
{
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    work();  // warm-up
    std::vector<float> t;
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(a, e.stream);
        work();
        (void)hipEventRecord(b, e.stream);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        t.push_back(ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

}  // namespace ftc

extern "C" {

const char *ftc_ensemble_error_v1(void) { return g_ens_error.c_str(); }

int ftc_ensemble_create_v1(int count, const int32_t *sizes, double dx, double dy, double dz, double dt, double cs,
                           int mode, ftc_ensemble **out)
{
    if (!out || !sizes || count <= 0) return eerr(FORT_SCOPE_ARGUMENT, "missing ensemble arguments");
    *out = nullptr;
    if (mode < 0 || mode > 2) return eerr(FORT_SCOPE_ARGUMENT, "unknown ensemble mode");
    if (!(dx > 0 && dy > 0 && dz > 0 && dt > 0)) return eerr(FORT_SCOPE_ARGUMENT, "spacing and dt must be positive");
    auto e = std::make_unique<ftc_ensemble>();
    e->mode = mode;
    e->count = count;
    e->dx = dx, e->dy = dy, e->dz = dz, e->dt = dt, e->cs = cs;
    e->les_iso = dx == dy && dy == dz;
    for (int m = 0; m < count; ++m) {
        const int a = sizes[3 * m], b = sizes[3 * m + 1], c = sizes[3 * m + 2];
        if (a < 2 || b < 2 || c < 2 || a > 4096 || b > 4096 || c > 150)
            return eerr(FORT_SCOPE_ARGUMENT, "member sizes must be 2..4096 (x, y) and 2..150 (z)");
        e->nx.push_back(a);
        e->ny.push_back(b);
        e->nz.push_back(c);
        e->nzmax = std::max(e->nzmax, c);
    }
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return eerr(FORT_SCOPE_RESOURCE, "no device");
    e->cus = prop.multiProcessorCount;
    const bool mx = mode == FTC_ENSEMBLE_MIXED_BATCHED;
    const std::size_t wbytes = mx ? 4 : 8;
    // Arena: per member fields (2 sets x 4), phi, rhs, res, w0, w1, bases, eigenvalues, partials.
    std::size_t bytes = 0;
    for (int m = 0; m < count; ++m) {
        const std::size_t P = (std::size_t)(e->nx[m] + 2) * (e->ny[m] + 2) * (e->nz[m] + 2);
        const std::size_t N = (std::size_t)e->nx[m] * e->ny[m] * e->nz[m];
        const std::size_t X = e->nx[m], Y = e->ny[m];
        bytes += 8 * up256(P * 8) + 3 * up256(N * 8) + 2 * up256(N * wbytes);
        bytes += 2 * up256(X * (X + kQPad) * 8) + 2 * up256(Y * (Y + kQPad) * 8) + 2 * up256(X * X * 4) +
                 2 * up256(Y * Y * 4);
        bytes += up256(X * 8) + up256(Y * 8) + up256(257 * 8);
    }
    bytes += up256(count * sizeof(EnsMember)) + 2 * up256(count * sizeof(LesArgs)) + up256(6 * count * sizeof(GemmCall)) +
             up256(6 * count * sizeof(XformCall));
    e->arena_bytes = bytes;
    if (hipMalloc(&e->arena, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return eerr(FORT_SCOPE_RESOURCE, "ensemble allocation failed");
    }
    if (hipMemset(e->arena, 0, bytes) != hipSuccess) return eerr(FORT_SCOPE_RESOURCE, "ensemble clear failed");
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return eerr(FORT_SCOPE_RESOURCE, "stream creation failed");
    char *p = static_cast<char *>(e->arena);
    auto carve = [&](std::size_t n) {
        char *r = p;
        p += up256(n);
        return r;
    };
    std::vector<GemmCall> calls(6 * count);
    std::vector<XformCall> xcalls(6 * count);
    // The FP64 mode applies its transforms along each axis with ens_xform_kernel when every member's extent
    // along that axis fits its LDS block; otherwise (or under FTC_ENS_TILED_TRANSFORMS) as contractions over
    // tiles, which give the same bits.
    const int max_nx = *std::max_element(e->nx.begin(), e->nx.end());
    const int max_ny = *std::max_element(e->ny.begin(), e->ny.end());
    const bool tiled = std::getenv("FTC_ENS_TILED_TRANSFORMS") != nullptr;
    for (int s = 0; s < 6; ++s)
        e->xform[s] = !mx && !tiled && (slot_along_y(s) ? max_ny : max_nx) <= kXformMax;
    e->args[0].resize(count);
    e->args[1].resize(count);
    for (int m = 0; m < count; ++m) {
        EnsMember mm{};
        mm.nx = e->nx[m], mm.ny = e->ny[m], mm.nz = e->nz[m];
        const std::size_t P = (std::size_t)(mm.nx + 2) * (mm.ny + 2) * (mm.nz + 2);
        const std::size_t N = (std::size_t)mm.nx * mm.ny * mm.nz;
        mm.sj = mm.nx + 2;
        mm.sk = (long long)(mm.nx + 2) * (mm.ny + 2);
        for (int s = 0; s < 2; ++s)
            for (int f = 0; f < 4; ++f) mm.f[s][f] = (double *)carve(P * 8);
        mm.phi = (double *)carve(N * 8);
        mm.rhs = (double *)carve(N * 8);
        mm.res = (double *)carve(N * 8);
        mm.w0 = carve(N * wbytes);
        mm.w1 = carve(N * wbytes);
        const int X = mm.nx, Y = mm.ny;
        // kQPad spare rows of zeros after each FP64 matrix: ens_xform_kernel reads K up to the next multiple of
        // 16 past n with the step in the scalar offset, which the buffer's range check does not cover.
        const std::size_t qxb = (std::size_t)X * (X + kQPad) * 8, qyb = (std::size_t)Y * (Y + kQPad) * 8;
        double *Qx = (double *)carve(qxb), *QxT = (double *)carve(qxb);
        double *Qy = (double *)carve(qyb), *QyT = (double *)carve(qyb);
        float *Qxf = (float *)carve((std::size_t)X * X * 4), *QxTf = (float *)carve((std::size_t)X * X * 4);
        float *Qyf = (float *)carve((std::size_t)Y * Y * 4), *QyTf = (float *)carve((std::size_t)Y * Y * 4);
        mm.lx = (double *)carve((std::size_t)X * 8);
        mm.ly = (double *)carve((std::size_t)Y * 8);
        mm.part = (double *)carve(257 * 8);
        mm.gx = dt / dx, mm.gy = dt / dy, mm.gz = dt / dz;
        mm.rdx = 1.0 / dx, mm.rdy = 1.0 / dy, mm.rdz = 1.0 / dz, mm.rdt = 1.0 / dt;
        mm.rx = 1.0 / (dx * dx), mm.ry = 1.0 / (dy * dy), mm.rz = 1.0 / (dz * dz);
        std::vector<double> qx, lx, qy, ly, tabx, taby;
        basis_table(X, tabx);
        basis_table(Y, taby);
        periodic_basis(X, dx, tabx, qx, lx);
        periodic_basis(Y, dy, taby, qy, ly);
        std::vector<double> qxt(qx.size()), qyt(qy.size());
        for (int a = 0; a < X; ++a)
            for (int b = 0; b < X; ++b) qxt[(std::size_t)b * X + a] = qx[(std::size_t)a * X + b];
        for (int a = 0; a < Y; ++a)
            for (int b = 0; b < Y; ++b) qyt[(std::size_t)b * Y + a] = qy[(std::size_t)a * Y + b];
        std::vector<float> fqx(qx.begin(), qx.end()), fqxt(qxt.begin(), qxt.end()), fqy(qy.begin(), qy.end()),
            fqyt(qyt.begin(), qyt.end());
        bool ok = hipMemcpy(Qx, qx.data(), qx.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(QxT, qxt.data(), qxt.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(Qy, qy.data(), qy.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(QyT, qyt.data(), qyt.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(Qxf, fqx.data(), fqx.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(QxTf, fqxt.data(), fqxt.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(Qyf, fqy.data(), fqy.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(QyTf, fqyt.data(), fqyt.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(mm.lx, lx.data(), lx.size() * 8, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(mm.ly, ly.data(), ly.size() * 8, hipMemcpyHostToDevice) == hipSuccess;
        if (!ok) return eerr(FORT_SCOPE_RESOURCE, "basis upload failed");
        GemmCall c6[6];
        member_calls(*e, mm, Qx, QxT, Qy, QyT, Qxf, QxTf, Qyf, QyTf, c6);
        for (int s = 0; s < 6; ++s) calls[s * count + m] = c6[s];
        XformCall x6[6];
        member_xforms(mm, Qx, QxT, Qy, QyT, x6);
        for (int s = 0; s < 6; ++s) xcalls[s * count + m] = x6[s];
        for (int par = 0; par < 2; ++par) {
            LesArgs a{};
            a.U = mm.f[par][0], a.V = mm.f[par][1], a.W = mm.f[par][2], a.S = mm.f[par][3];
            a.OU = mm.f[1 - par][0], a.OV = mm.f[1 - par][1], a.OW = mm.f[1 - par][2], a.OS = mm.f[1 - par][3];
            a.nx = mm.nx, a.ny = mm.ny, a.nz = mm.nz, a.kw = mm.nz;
            a.sj = mm.sj, a.sk = mm.sk;
            a.dt = dt;
            double c19[19];
            ftc_les_coefficients(dx, dy, dz, cs, c19);
            for (int c = 0; c < 3; ++c) a.c25[c] = c19[c], a.c50[c] = c19[3 + c], a.c125[c] = c19[6 + c];
            for (int c = 0; c < 10; ++c) a.sgs[c] = c19[9 + c];
            e->args[par][m] = a;
        }
        e->members.push_back(mm);
    }
    e->d_members = (EnsMember *)carve(count * sizeof(EnsMember));
    e->d_args[0] = (LesArgs *)carve(count * sizeof(LesArgs));
    e->d_args[1] = (LesArgs *)carve(count * sizeof(LesArgs));
    e->d_calls = (GemmCall *)carve(6 * count * sizeof(GemmCall));
    e->d_xcalls = (XformCall *)carve(6 * count * sizeof(XformCall));
    if (hipMemcpy(e->d_members, e->members.data(), count * sizeof(EnsMember), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->d_calls, calls.data(), calls.size() * sizeof(GemmCall), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->d_xcalls, xcalls.data(), xcalls.size() * sizeof(XformCall), hipMemcpyHostToDevice) != hipSuccess)
        return eerr(FORT_SCOPE_RESOURCE, "description upload failed");
    FTC_CHECK(build_items(*e));
    if (mode != FTC_ENSEMBLE_FP64_PER_MEMBER) FTC_CHECK(capture(*e));
    *out = e.release();
    return FORT_SCOPE_OK;
}

int ftc_ensemble_upload_v1(ftc_ensemble *e, int member, const double *u, const double *v, const double *w,
                           const double *s)
{
    if (!e || member < 0 || member >= e->count || !u || !v || !w || !s) return eerr(FORT_SCOPE_ARGUMENT, "bad upload");
    const EnsMember &m = e->members[member];
    const std::size_t bytes = (std::size_t)(m.nx + 2) * (m.ny + 2) * (m.nz + 2) * 8;
    const double *src[4] = {u, v, w, s};
    if (hipStreamSynchronize(e->stream) != hipSuccess) return eerr(FORT_SCOPE_EXECUTION, "stream failed");
    for (int set = 0; set < 2; ++set)
        for (int f = 0; f < 4; ++f)
            if (hipMemcpy(m.f[set][f], src[f], bytes, hipMemcpyHostToDevice) != hipSuccess)
                return eerr(FORT_SCOPE_EXECUTION, "upload failed");
    e->parity = 0;
    return FORT_SCOPE_OK;
}

int ftc_ensemble_download_v1(ftc_ensemble *e, int member, double *u, double *v, double *w, double *s)
{
    if (!e || member < 0 || member >= e->count || !u || !v || !w || !s) return eerr(FORT_SCOPE_ARGUMENT, "bad download");
    const EnsMember &m = e->members[member];
    const std::size_t bytes = (std::size_t)(m.nx + 2) * (m.ny + 2) * (m.nz + 2) * 8;
    double *dst[4] = {u, v, w, s};
    if (hipStreamSynchronize(e->stream) != hipSuccess) return eerr(FORT_SCOPE_EXECUTION, "stream failed");
    for (int f = 0; f < 4; ++f)
        if (hipMemcpy(dst[f], m.f[e->parity][f], bytes, hipMemcpyDeviceToHost) != hipSuccess)
            return eerr(FORT_SCOPE_EXECUTION, "download failed");
    return FORT_SCOPE_OK;
}

int ftc_ensemble_download_pressure_v1(ftc_ensemble *e, int member, double *phi)
{
    if (!e || member < 0 || member >= e->count || !phi) return eerr(FORT_SCOPE_ARGUMENT, "bad download");
    const EnsMember &m = e->members[member];
    if (hipStreamSynchronize(e->stream) != hipSuccess ||
        hipMemcpy(phi, m.phi, (std::size_t)m.nx * m.ny * m.nz * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return eerr(FORT_SCOPE_EXECUTION, "download failed");
    return FORT_SCOPE_OK;
}

int ftc_ensemble_set_solves_v1(ftc_ensemble *e, int solves)
{
    if (!e || solves < 1 || solves > 16) return eerr(FORT_SCOPE_ARGUMENT, "solves must be 1..16");
    e->solves = solves;
    if (e->mode != FTC_ENSEMBLE_FP64_PER_MEMBER) return capture(*e);
    return FORT_SCOPE_OK;
}

int ftc_ensemble_set_les_policy_v1(ftc_ensemble *e, int policy)
{
    if (!e || policy < 0 || policy > 2) return eerr(FORT_SCOPE_ARGUMENT, "LES tile policy is 0, 1 or 2");
    e->les_policy = policy;
    FTC_CHECK(build_items(*e));
    if (e->mode != FTC_ENSEMBLE_FP64_PER_MEMBER) return capture(*e);
    return FORT_SCOPE_OK;
}

int ftc_ensemble_set_projection_v1(ftc_ensemble *e, int on)
{
    if (!e || (on != 0 && on != 1)) return eerr(FORT_SCOPE_ARGUMENT, "projection is 0 or 1");
    e->projection = on == 1;
    if (e->mode != FTC_ENSEMBLE_FP64_PER_MEMBER) return capture(*e);
    return FORT_SCOPE_OK;
}

int ftc_ensemble_step_v1(ftc_ensemble *e, int steps)
{
    if (!e || steps < 0) return eerr(FORT_SCOPE_ARGUMENT, "bad step count");
    for (int n = 0; n < steps; ++n) {
        if (e->mode == FTC_ENSEMBLE_FP64_PER_MEMBER) {
            FTC_CHECK(enqueue_step(*e, e->parity));
        } else if (hipGraphLaunch(e->graph[e->parity], e->stream) != hipSuccess) {
            (void)hipGetLastError();
            return eerr(FORT_SCOPE_EXECUTION, "graph launch failed");
        }
        e->parity = 1 - e->parity;
    }
    return FORT_SCOPE_OK;
}

int ftc_ensemble_synchronize_v1(ftc_ensemble *e)
{
    if (!e) return eerr(FORT_SCOPE_ARGUMENT, "missing ensemble");
    return hipStreamSynchronize(e->stream) == hipSuccess ? FORT_SCOPE_OK : eerr(FORT_SCOPE_EXECUTION, "stream failed");
}

int ftc_ensemble_stream_v1(ftc_ensemble *e, void **stream)
{
    if (!e || !stream) return eerr(FORT_SCOPE_ARGUMENT, "missing ensemble");
    *stream = e->stream;
    return FORT_SCOPE_OK;
}

int ftc_ensemble_autotune_v1(ftc_ensemble *e, const char *cache_path)
{
    if (!e) return eerr(FORT_SCOPE_ARGUMENT, "missing ensemble");
    std::ostringstream key;
    key << "mode=" << e->mode << ";solves=" << e->solves << ";cus=" << e->cus << ";sizes=";
    for (int m = 0; m < e->count; ++m) key << e->nx[m] << "x" << e->ny[m] << "x" << e->nz[m] << ",";
    std::map<std::string, std::string> cache;
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            const auto tab = line.find('\t');
            if (tab != std::string::npos) cache[line.substr(0, tab)] = line.substr(tab + 1);
        }
        auto hit = cache.find(key.str());
        if (hit != cache.end() && std::sscanf(hit->second.c_str(), "%d %d %d %d", &e->les_kchunk, &e->gemm_tile,
                                              &e->blocks_per_cu, &e->les_policy) == 4) {
            FTC_CHECK(build_items(*e));
            return e->mode != FTC_ENSEMBLE_FP64_PER_MEMBER ? capture(*e) : FORT_SCOPE_OK;
        }
    }
    // Coordinate search in a fixed order: the LES k chunk on the LES stage, the contraction tile on the
    // transform stages, then the persistent grid on the whole step. The fields are scratch for the
    // search: the caller uploads (again) afterwards.
    const bool per_member = e->mode == FTC_ENSEMBLE_FP64_PER_MEMBER;
    auto run_stages = [&](std::initializer_list<int> stages) {
        for (int st : stages) {
            if (per_member)
                for (int m = 0; m < e->count; ++m) launch_stage(*e, st, m, 0);
            else launch_stage(*e, st, -1, 0);
        }
    };
    // The LES stage: tile-width policy (all 64 wide, widths per member by idle lanes, all 32 wide) and k chunk.
    float best = 1e30f;
    int best_k = e->les_kchunk, best_p = e->les_policy;
    for (int policy : {0, 2, 1})
        for (int k : {12, 24, 48, 96}) {
            e->les_policy = policy;
            e->les_kchunk = k;
            FTC_CHECK(build_items(*e));
            const float t = time_ms(*e, [&] { run_stages({kLes}); }, 3);
            if (t < best) best = t, best_k = k, best_p = policy;
        }
    e->les_kchunk = best_k;
    e->les_policy = best_p;
    best = 1e30f;
    int best_t = e->gemm_tile;
    for (int tile : {0, 1}) {
        e->gemm_tile = tile;
        FTC_CHECK(build_items(*e));
        const float t = time_ms(*e, [&] { run_stages({kG1, kG2, kG3, kG4}); }, 3);
        if (t < best) best = t, best_t = tile;
    }
    e->gemm_tile = best_t;
    FTC_CHECK(build_items(*e));
    if (!per_member) {
        best = 1e30f;
        int best_b = e->blocks_per_cu;
        for (int b : {2, 4, 8, 16}) {
            e->blocks_per_cu = b;
            const float t = time_ms(*e, [&] { (void)enqueue_step(*e, 0); }, 3);
            if (t < best) best = t, best_b = b;
        }
        e->blocks_per_cu = best_b;
    }
    if (hipStreamSynchronize(e->stream) != hipSuccess) return eerr(FORT_SCOPE_EXECUTION, "tuning run failed");
    if (cache_path) {
        // k chunk, contraction tile, workgroups per CU, LES tile policy, then the width each member takes.
        std::string widths = " widths=";
        for (int m = 0; m < e->count; ++m) widths += std::to_string(e->les_w[m]) + (m + 1 < e->count ? "," : "");
        cache[key.str()] = std::to_string(e->les_kchunk) + " " + std::to_string(e->gemm_tile) + " " +
                           std::to_string(e->blocks_per_cu) + " " + std::to_string(e->les_policy) + widths;
        std::ofstream out(cache_path, std::ios::trunc);
        for (auto &kv : cache) out << kv.first << '\t' << kv.second << '\n';
    }
    return per_member ? FORT_SCOPE_OK : capture(*e);
}

int ftc_ensemble_get_info_v1(ftc_ensemble *e, ftc_ensemble_info_v1 *info)
{
    if (!e || !info) return eerr(FORT_SCOPE_ARGUMENT, "missing ensemble or output");
    info->version = 1;
    info->mode = e->mode;
    info->members = e->count;
    info->solves = mixed(*e) ? e->solves : 1;
    info->les_kchunk = e->les_kchunk;
    info->gemm_tile = e->gemm_tile;
    info->blocks_per_cu = e->blocks_per_cu;
    info->les_policy = e->les_policy;
    // Batched: one launch per stage, two for the LES stage when both tile widths occur.
    const bool two_widths = e->n_les32 > 0 && e->n_les32 < e->n_les;
    info->launches_per_step = e->mode == FTC_ENSEMBLE_FP64_PER_MEMBER
                                  ? (int64_t)step_stages(*e).size() * e->count
                                  : (int64_t)step_stages(*e).size() + (two_widths ? 1 : 0);
    double bytes = 0;
    int64_t cells = 0;
    for (int m = 0; m < e->count; ++m) {
        const double P = (double)(e->nx[m] + 2) * (e->ny[m] + 2) * (e->nz[m] + 2);
        const double N = (double)e->nx[m] * e->ny[m] * e->nz[m];
        cells += (int64_t)N;
        // LES step 8(4P+4I); divergence 8(3P+N); pressure floor 8(2N); projection 8(3P+N+3I).
        bytes += 8 * (4 * P + 4 * N) + 8 * (3 * P + N) + 8 * 2 * N + 8 * (3 * P + N + 3 * N);
    }
    info->cells = cells;
    info->algorithmic_bytes = bytes;
    info->device_bytes = e->arena_bytes;
    return FORT_SCOPE_OK;
}

int ftc_ensemble_destroy_v1(ftc_ensemble **pe)
{
    if (!pe) return eerr(FORT_SCOPE_ARGUMENT, "missing ensemble pointer");
    ftc_ensemble *e = *pe;
    if (!e) return FORT_SCOPE_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(*e);
    if (e->items_arena) (void)hipFree(e->items_arena);
    if (e->arena) (void)hipFree(e->arena);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    *pe = nullptr;
    return FORT_SCOPE_OK;
}

}  // extern "C"
