// Device code of the pressure solve (poisson.hip holds the provider): residual, mean removal and the
// FP64 z-direction Thomas solves, also used by the ensemble's grouped launches.
#pragma once
#include <hip/hip_runtime.h>

#include "reduce.hpp"

#pragma clang fp contract(off)

namespace ftc {

// r = b - A phi over the interior (phi == nullptr: r = b). phi's neighbours wrap periodically or mirror
// (Neumann on the staggered face) instead of reading its ghost layers.
static __global__ void poisson_residual_kernel(const double *__restrict__ phi, long long p1, long long p2,
                                        const double *__restrict__ b, long long b1, long long b2,
                                        double *__restrict__ r, int nx, int ny, int nz, double rx, double ry,
                                        double rz, int perx, int pery)
{
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    const double bv = b[i + j * b1 + k * b2];
    const long long out = (long long)i + (long long)nx * ((long long)j + (long long)ny * k);
    if (!phi) {
        r[out] = bv;
        return;
    }
    const int im = i == 0 ? (perx ? nx - 1 : 0) : i - 1, ip = i == nx - 1 ? (perx ? 0 : nx - 1) : i + 1;
    const int jm = j == 0 ? (pery ? ny - 1 : 0) : j - 1, jp = j == ny - 1 ? (pery ? 0 : ny - 1) : j + 1;
    const int km = k == 0 ? 0 : k - 1, kp = k == nz - 1 ? nz - 1 : k + 1;
#define P(a, c, e) phi[(a) + (c) * p1 + (long long)(e) * p2]
    const double c0 = P(i, j, k);
    const double lap = (((P(im, j, k) + P(ip, j, k)) - 2.0 * c0) * rx + ((P(i, jm, k) + P(i, jp, k)) - 2.0 * c0) * ry) +
                       ((P(i, j, km) + P(i, j, kp)) - 2.0 * c0) * rz;
#undef P
    r[out] = bv - lap;
}

// The mixed solve's residual: r = b - lap(phi) in FP64, stored as the split-bf16 contractions read it
// (the value rounded to float, then bf16 hi and lo planes, lo `lo` elements after hi). A block covers
// 64 x 4 columns over kResidualChunk planes; phi at (i, j) is carried in registers over k (the k-1, k, k+1
// window, Neumann ends repeating the end plane). Per block (blocks in x, then y, then chunk order), in a
// fixed order (each thread over its k in order, then a tree over the 256 threads): the sum of phi into
// phisum[block] and, when `partial` is set, the sum of r*r into partial[block] for the residual norm.
constexpr int kResidualChunk = 16;

// SPLIT: r holds bf16 hi and lo planes (lo `lo` elements after hi) for the split-bf16 contractions;
// otherwise r holds FP32 values (the Hartley passes of the mixed solve read FP32).
template <bool SPLIT>
static __global__ __launch_bounds__(256) void poisson_residual_split_kernel(
    const double *__restrict__ phi, long long p1, long long p2, const double *__restrict__ b, long long b1,
    long long b2, void *__restrict__ r, long long lo, double *__restrict__ partial, double *__restrict__ phisum,
    int nx, int ny, int nz, double rx, double ry, double rz, int perx, int pery)
{
    __shared__ double sq[256], sp[256];
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int k0 = blockIdx.z * kResidualChunk, k1 = min(k0 + kResidualChunk, nz);
    const int t = threadIdx.y * 64 + threadIdx.x;
    double acc2 = 0.0, acc = 0.0;
    if (i < nx && j < ny) {
        const int im = i == 0 ? (perx ? nx - 1 : 0) : i - 1, ip = i == nx - 1 ? (perx ? 0 : nx - 1) : i + 1;
        const int jm = j == 0 ? (pery ? ny - 1 : 0) : j - 1, jp = j == ny - 1 ? (pery ? 0 : ny - 1) : j + 1;
#define P(a, c, e) phi[(a) + (c) * p1 + (long long)(e) * p2]
        double c0 = P(i, j, k0), zm = k0 == 0 ? c0 : P(i, j, k0 - 1);
        for (int k = k0; k < k1; ++k) {
            const double zp = k == nz - 1 ? c0 : P(i, j, k + 1);
            const double lap = (((P(im, j, k) + P(ip, j, k)) - 2.0 * c0) * rx +
                                ((P(i, jm, k) + P(i, jp, k)) - 2.0 * c0) * ry) +
                               ((zm + zp) - 2.0 * c0) * rz;
            const double v = b[i + j * b1 + k * b2] - lap;
            const long long out = (long long)i + (long long)nx * ((long long)j + (long long)ny * k);
            const float f = (float)v;
            if constexpr (SPLIT) {
                const __bf16 h = (__bf16)f;
                static_cast<__bf16 *>(r)[out] = h;
                static_cast<__bf16 *>(r)[out + lo] = (__bf16)(f - (float)h);
            } else {
                static_cast<float *>(r)[out] = f;
            }
            acc2 = acc2 + v * v;
            acc = acc + c0;
            zm = c0;
            c0 = zp;
        }
#undef P
    }
    sq[t] = acc2;
    sp[t] = acc;
    __syncthreads();
#pragma unroll
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            sq[t] = sq[t] + sq[t + w];
            sp[t] = sp[t] + sp[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const long long blk = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
        phisum[blk] = sp[0];
        if (partial) partial[blk] = sq[0];
    }
}

// The mean of phi from the 256 fixed-order partial sums (folded in index order), on the device.
static __global__ void poisson_mean_kernel(const double *__restrict__ final256, double count, double *__restrict__ mean)
{
    double acc = 0.0;
    for (int f = 0; f < kReduceFinal; ++f) acc = acc + final256[f];
    *mean = acc / count;
}

static __global__ void poisson_subtract_kernel(double *__restrict__ phi, long long p1, long long p2, int nx, int ny,
                                        const double *__restrict__ mean)
{
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    double &v = phi[i + j * p1 + (long long)k * p2];
    v = v - *mean;
}

// One FP64 Thomas solve per (p, q) column of the transformed right-hand side; a wavefront holds 64
// consecutive columns, so every z plane is read and written as one contiguous run.
//
// The elimination factors depend on lambda and k only, and the sweep is a dependent chain, so the solve
// keeps no column in LDS: the forward sweep stores just the state (c_{k-1}, d_{k-1}) entering each chunk
// of kTriChunk points (ck: 2 * chunks * cols doubles of LDS), and back substitution re-runs each chunk
// from its entry state into registers, re-reading that chunk of the input, before substituting through
// it. The re-run repeats the forward arithmetic operation for operation, so the result is the plain
// Thomas solve's, while LDS per column drops from 16*nz bytes to 16*chunks and many more waves fit a CU.
constexpr int kTriChunk = 16;
// Largest nz of the z solve: its checkpoints then take 64 KB of LDS per 64 columns.
constexpr int kPoissonMaxNz = 1024;

__host__ __device__ constexpr int tridiag_chunks(int nz) { return (nz + kTriChunk - 1) / kTriChunk; }

template <class T>
__device__ __forceinline__ void tridiag_load_chunk(const T *__restrict__ src, long long plane, int k0, int nz,
                                                   double (&v)[kTriChunk])
{
#pragma unroll
    for (int i = 0; i < kTriChunk; ++i) v[i] = k0 + i < nz ? (double)src[(long long)(k0 + i) * plane] : 0.0;
}

// Forward elimination over one chunk: c = rz*m, d = (f - rz*d)*m with m = 1/(diag_k - rz*c).
template <bool KEEP>
__device__ __forceinline__ void tridiag_forward_chunk(const double (&v)[kTriChunk], int k0, int nz, double lam,
                                                      double rz, double &c, double &d, double (&cl)[kTriChunk],
                                                      double (&dl)[kTriChunk])
{
#pragma unroll
    for (int i = 0; i < kTriChunk; ++i) {
        const int k = k0 + i;
        if (k < nz) {
            double diag = -2.0 * rz + lam;
            if (k == 0) diag = diag + rz;
            if (k == nz - 1) diag = diag + rz;
            const double m = 1.0 / (diag - rz * c);
            c = rz * m;
            d = (v[i] - rz * d) * m;
            if (KEEP) {
                cl[i] = c;
                dl[i] = d;
            }
        }
    }
}

// The constant null mode (lambda == 0): the solution of zero mean. psi_0 = 0 fixes the constant, the first
// nz-1 equations give the rest, and the mean is removed (the last equation holds when the right-hand side
// has zero mean). Two passes over the column: the sum, then the values again minus the mean.
// Output of the z solve: a plain array of TOUT, or (TOUT = TriSplit) bf16 hi and lo planes, lo `lo`
// elements after hi, holding the value rounded to float and split as the split-bf16 contractions split it.
struct TriSplit {};

template <class TOUT>
struct TriOut {
    TOUT *p;
    long long lo;
    __device__ __forceinline__ void put(long long e, double v) const { p[e] = (TOUT)v; }
};
template <>
struct TriOut<TriSplit> {
    __bf16 *p;
    long long lo;
    __device__ __forceinline__ void put(long long e, double v) const
    {
        const float f = (float)v;
        const __bf16 h = (__bf16)f;
        p[e] = h;
        p[e + lo] = (__bf16)(f - (float)h);
    }
};

template <class T, class TOUT>
__device__ __forceinline__ void tridiag_null_mode(const T *__restrict__ src, const TriOut<TOUT> &dst, long long plane,
                                                  int nz, double rz)
{
    double mean = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        double prev = 0.0, cur = 0.0, sum = 0.0;
        if (pass == 1) dst.put(0, 0.0 - mean);
        if (nz > 1) {
            cur = prev + (double)src[0] / rz;
            if (pass == 1) dst.put(plane, cur - mean);
            sum = cur;
            for (int k = 1; k < nz - 1; ++k) {
                const double next = ((double)src[(long long)k * plane] / rz - prev) + 2.0 * cur;
                prev = cur;
                cur = next;
                if (pass == 1) dst.put((long long)(k + 1) * plane, cur - mean);
                sum = sum + cur;
            }
        }
        mean = sum / nz;
    }
}

// Solves column `col` of a plane-major (plane = nx*ny values per z level) array; c is the thread's column
// within its block of `cols` and ck the block's checkpoint LDS.
template <class T, class TOUT = T>
__device__ __forceinline__ void tridiag_column(const T *__restrict__ in, TriOut<TOUT> out, long long plane, int nz,
                                               double lam, double rz, double *ck, int c, int cols, long long col)
{
    const T *src = in + col;
    TriOut<TOUT> dst = out;
    dst.p += col;
    if (lam == 0.0) {
        tridiag_null_mode<T, TOUT>(src, dst, plane, nz, rz);
        return;
    }
    const int chunks = tridiag_chunks(nz);
    double cs = 0.0, ds = 0.0, v[kTriChunk], cl[kTriChunk], dl[kTriChunk];
    for (int j = 0; j < chunks; ++j) {
        ck[(2 * j) * cols + c] = cs;
        ck[(2 * j + 1) * cols + c] = ds;
        tridiag_load_chunk<T>(src, plane, j * kTriChunk, nz, v);
        tridiag_forward_chunk<false>(v, j * kTriChunk, nz, lam, rz, cs, ds, cl, dl);
    }
    double psi = 0.0;
    for (int j = chunks - 1; j >= 0; --j) {
        const int k0 = j * kTriChunk;
        double cc = ck[(2 * j) * cols + c], dd = ck[(2 * j + 1) * cols + c];
        tridiag_load_chunk<T>(src, plane, k0, nz, v);
        tridiag_forward_chunk<true>(v, k0, nz, lam, rz, cc, dd, cl, dl);
#pragma unroll
        for (int i = kTriChunk - 1; i >= 0; --i) {
            const int k = k0 + i;
            if (k < nz) {
                psi = k == nz - 1 ? dl[i] : dl[i] - cl[i] * psi;
                dst.put((long long)k * plane, psi);
            }
        }
    }
}

// The same solve with the next chunk's loads in flight while the current chunk is eliminated: the raw
// input of chunk j +- 1 is fetched into a second register set before chunk j is widened and swept, in
// both the forward sweep and the back-substitution re-run. Same arithmetic, same result.
template <class T>
__device__ __forceinline__ void tridiag_fetch_chunk(const T *__restrict__ src, long long plane, int k0, int nz,
                                                    T (&raw)[kTriChunk])
{
#pragma unroll
    for (int i = 0; i < kTriChunk; ++i) raw[i] = k0 + i < nz ? src[(long long)(k0 + i) * plane] : T(0);
}

template <class T>
__device__ __forceinline__ void tridiag_widen(const T (&raw)[kTriChunk], double (&v)[kTriChunk])
{
#pragma unroll
    for (int i = 0; i < kTriChunk; ++i) v[i] = (double)raw[i];
}

template <class T, class TOUT = T>
__device__ __forceinline__ void tridiag_column_prefetch(const T *__restrict__ in, TriOut<TOUT> out, long long plane,
                                                        int nz, double lam, double rz, double *ck, int c, int cols,
                                                        long long col)
{
    const T *src = in + col;
    TriOut<TOUT> dst = out;
    dst.p += col;
    if (lam == 0.0) {
        tridiag_null_mode<T, TOUT>(src, dst, plane, nz, rz);
        return;
    }
    const int chunks = tridiag_chunks(nz);
    double cs = 0.0, ds = 0.0, v[kTriChunk], cl[kTriChunk], dl[kTriChunk];
    T ra[kTriChunk], rb[kTriChunk];
    auto sweep = [&](const T (&raw)[kTriChunk], int j) {
        ck[(2 * j) * cols + c] = cs;
        ck[(2 * j + 1) * cols + c] = ds;
        tridiag_widen<T>(raw, v);
        tridiag_forward_chunk<false>(v, j * kTriChunk, nz, lam, rz, cs, ds, cl, dl);
    };
    tridiag_fetch_chunk<T>(src, plane, 0, nz, ra);
    for (int j = 0; j < chunks; j += 2) {
        if (j + 1 < chunks) tridiag_fetch_chunk<T>(src, plane, (j + 1) * kTriChunk, nz, rb);
        sweep(ra, j);
        if (j + 1 < chunks) {
            if (j + 2 < chunks) tridiag_fetch_chunk<T>(src, plane, (j + 2) * kTriChunk, nz, ra);
            sweep(rb, j + 1);
        }
    }
    double psi = 0.0;
    auto substitute = [&](const T (&raw)[kTriChunk], int j) {
        const int k0 = j * kTriChunk;
        double cc = ck[(2 * j) * cols + c], dd = ck[(2 * j + 1) * cols + c];
        tridiag_widen<T>(raw, v);
        tridiag_forward_chunk<true>(v, k0, nz, lam, rz, cc, dd, cl, dl);
#pragma unroll
        for (int i = kTriChunk - 1; i >= 0; --i) {
            const int k = k0 + i;
            if (k < nz) {
                psi = k == nz - 1 ? dl[i] : dl[i] - cl[i] * psi;
                dst.put((long long)k * plane, psi);
            }
        }
    };
    tridiag_fetch_chunk<T>(src, plane, (chunks - 1) * kTriChunk, nz, ra);
    for (int j = chunks - 1; j >= 0; j -= 2) {
        if (j >= 1) tridiag_fetch_chunk<T>(src, plane, (j - 1) * kTriChunk, nz, rb);
        substitute(ra, j);
        if (j >= 1) {
            if (j >= 2) tridiag_fetch_chunk<T>(src, plane, (j - 2) * kTriChunk, nz, ra);
            substitute(rb, j - 1);
        }
    }
}

// One wave per 64 columns; the prefetching sweep needs up to 148 VGPRs for FP64 input (three waves per
// SIMD), 118 for FP32. Against the plain sweep it took the FP32-input solve at 512 x 512 x 128 from 160 to
// 124 us and the FP64 one (bandwidth-bound) from 186 to 182 us.
template <class T, class TOUT = T>
__global__ __launch_bounds__(64, 3) void tridiag_z_kernel(const T *__restrict__ in, TriOut<TOUT> out, int nx, int ny, int nz,
                                 const double *__restrict__ lx, const double *__restrict__ ly, double rz, int cols,
                                 int qfast)
{
    extern __shared__ double lds[];
    const int c = threadIdx.x;
    const long long col = (long long)blockIdx.x * cols + c;
    const long long plane = (long long)nx * ny;
    if (col >= plane) return;
    // Mode (p, q) sits at column p + q*nx, or at p*ny + q when the transforms keep q fastest.
    const double lam = qfast ? lx[col / ny] + ly[col % ny] : lx[col % nx] + ly[col / nx];
    tridiag_column_prefetch<T, TOUT>(in, out, plane, nz, lam, rz, lds, c, cols, col);
}

// The FP64 z solve with the whole column held in registers, so its input is read once: one wave per 64 columns
// and one wave per SIMD (up to 512 registers). The forward sweep keeps d' in place of the input and the c' value
// entering each chunk; the back substitution recomputes a chunk's c' from its checkpoint (c' depends on the
// eigenvalue and k only) with the same operations, so every value is the one tridiag_column computes, in the
// same order. Columns of exactly NZ points (a constant, so no level needs a bound check); arrays below 2^31
// bytes (the caller checks both), addressed through buffer resources so that the level offsets stay scalar. The
// null mode (lambda == 0) takes tridiag_null_mode.
template <int NZ>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 64), amdgpu_waves_per_eu(1, 1))) void
tridiag_z_resident_kernel(const double *__restrict__ in, double *__restrict__ out, int nx, int ny,
                          const double *__restrict__ lx, const double *__restrict__ ly, double rz)
{
    static_assert(NZ % kTriChunk == 0, "whole chunks");
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    const long long col = (long long)blockIdx.x * 64 + threadIdx.x;
    const long long plane = (long long)nx * ny;
    if (col >= plane) return;
    const double lam = lx[col % nx] + ly[col / nx];
    if (lam == 0.0) {
        tridiag_null_mode<double, double>(in + col, TriOut<double>{out + col, 0}, plane, NZ, rz);
        return;
    }
    const int bytes = (int)(plane * NZ * 8), pb = (int)(plane * 8), vo = (int)(col * 8);
    const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(in), 0, bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(out, 0, bytes, 0x00020000);
    constexpr int CH = NZ / kTriChunk;
    double d[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k)
        d[k] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rin, vo, k * pb, 0));
    auto diagonal = [&](int k) {
        double diag = -2.0 * rz + lam;
        if (k == 0) diag = diag + rz;
        if (k == NZ - 1) diag = diag + rz;
        return diag;
    };
    double entry[CH];
    double c = 0.0, dd = 0.0;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        entry[j] = c;
#pragma unroll
        for (int i = 0; i < kTriChunk; ++i) {
            const int k = j * kTriChunk + i;
            const double m = 1.0 / (diagonal(k) - rz * c);
            c = rz * m;
            dd = (d[k] - rz * dd) * m;
            d[k] = dd;
        }
    }
    double psi = 0.0;
#pragma unroll
    for (int j = CH - 1; j >= 0; --j) {
        // The chunk's level offsets come from an opaque base, so they are not the load offsets kept live in
        // scalar registers from the first loop to this one.
        int base = j * kTriChunk * pb;
        asm volatile("" : "+s"(base));
        double cl[kTriChunk];
        double cc = entry[j];
        // Opaque too: the compiler would otherwise reuse the forward sweep's c' values and keep all of them live.
        asm volatile("" : "+v"(cc));
#pragma unroll
        for (int i = 0; i < kTriChunk; ++i) {
            const double m = 1.0 / (diagonal(j * kTriChunk + i) - rz * cc);
            cc = rz * m;
            cl[i] = cc;
        }
#pragma unroll
        for (int i = kTriChunk - 1; i >= 0; --i) {
            const int k = j * kTriChunk + i;
            psi = k == NZ - 1 ? d[k] : d[k] - cl[i] * psi;
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, psi), rout, vo, base + i * pb, 0);
        }
    }
}

}  // namespace ftc
