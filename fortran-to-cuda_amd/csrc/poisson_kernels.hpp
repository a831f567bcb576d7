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

// One FP64 Thomas solve per (p, q) column of the transformed right-hand side (p fastest, so a wavefront
// reads 64 consecutive values of each z plane). The whole column is first staged into LDS with
// independent loads (the sweep itself is a dependent chain), then solved and written back from LDS.
// Solves column `col` (p + nx*q) of a plane-major (nx*ny per z plane) array; cp and dp are this block's
// LDS arrays of nz*cols doubles and c the thread's column within the block.
template <class T>
__device__ __forceinline__ void tridiag_column(const T *__restrict__ in, T *__restrict__ out, long long plane,
                                               int nz, double lam, double rz, double *cp, double *dp, int c,
                                               int cols, long long col)
{
    for (int k = 0; k < nz; ++k) dp[(long long)k * cols + c] = (double)in[col + (long long)k * plane];
#define IN(k) (dp[(long long)(k) * cols + c])
#define OUT(k, v) (dp[(long long)(k) * cols + c] = (v))
    double *result = dp;
    if (lam == 0.0) {
        // The constant null mode: the solution of zero mean. psi_0 = 0 fixes the constant, the first
        // nz-1 equations give the rest (kept in cp), then the mean is removed (the last equation holds
        // when the right-hand side has zero mean).
        double prev = 0.0, cur = 0.0, sum = 0.0;
        cp[c] = 0.0;
        if (nz > 1) {
            cur = prev + IN(0) / rz;
            cp[(long long)cols + c] = cur;
            sum = cur;
            for (int k = 1; k < nz - 1; ++k) {
                const double next = (IN(k) / rz - prev) + 2.0 * cur;
                prev = cur;
                cur = next;
                cp[(long long)(k + 1) * cols + c] = cur;
                sum = sum + cur;
            }
        }
        const double mean = sum / nz;
        for (int k = 0; k < nz; ++k) cp[(long long)k * cols + c] = cp[(long long)k * cols + c] - mean;
        result = cp;
    } else {
        double cprev = 0.0, dprev = 0.0;
        for (int k = 0; k < nz; ++k) {
            double diag = -2.0 * rz + lam;
            if (k == 0) diag = diag + rz;
            if (k == nz - 1) diag = diag + rz;
            const double m = 1.0 / (diag - rz * cprev);
            cprev = rz * m;
            dprev = (IN(k) - rz * dprev) * m;
            cp[(long long)k * cols + c] = cprev;
            OUT(k, dprev);
        }
        double psi = dprev;
        for (int k = nz - 2; k >= 0; --k) {
            psi = dp[(long long)k * cols + c] - cp[(long long)k * cols + c] * psi;
            OUT(k, psi);
        }
    }
    for (int k = 0; k < nz; ++k) out[col + (long long)k * plane] = (T)result[(long long)k * cols + c];
#undef IN
#undef OUT
}

template <class T>
__global__ void tridiag_z_kernel(const T *__restrict__ in, T *__restrict__ out, int nx, int ny, int nz,
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
    tridiag_column<T>(in, out, plane, nz, lam, rz, lds, lds + (long long)nz * cols, c, cols, col);
}

}  // namespace ftc
