// Batched orthonormal discrete Hartley transforms along one axis, for the periodic directions of the
// pressure solve.
//
// For a line f_0 .. f_{n-1} (n a power of two) the transform is
//   H_k = n^-1/2 * sum_j f_j cas(2 pi j k / n),   cas t = cos t + sin t.
// cas(2 pi j k / n) is an eigenvector of the periodic second difference f_{j-1} - 2 f_j + f_{j+1} with
// eigenvalue 2 cos(2 pi k / n) - 2 (cos and sin of the same frequency share it), the vectors are
// orthogonal with squared norm n, and the transform is its own inverse. So the solve transforms x and y
// forward, solves the z tridiagonal systems with lambda_pq = lx_p + ly_q, and applies the same transforms
// again to come back.
//
// Two real lines a and b travel as one complex line z = a + i b through an FFT
// (Z_k = sum_j z_j e^{-2 pi i j k / n}); with W = Z_{(n-k) mod n},
//   H(a)_k = (Re Z + Re W - Im Z + Im W) / 2,   H(b)_k = (Im Z + Im W + Re Z - Re W) / 2
// (times n^-1/2). All arithmetic is FP64; the inputs and outputs may be FP32 (the workspaces of the
// mixed solve) or FP64.
//
// The FFT is a Stockham autosort transform in LDS with radix-8 stages (one radix-4 or radix-2 stage
// takes the rest of log2 n): each thread holds POINTS / NT complex points in registers, applies the
// stage's twiddles and R-point DFTs there, and writes the results to their sorted places; the first stage
// reads its operands straight from global memory and needs no twiddles. A workgroup of NT threads (256
// or 512) takes POINTS complex points (2048 or 4096): G = 2 * POINTS / n real lines of one plane. The lines of plane c are a = 0 .. A-1 and element
// e of line (a, c) is at in[a*ia + c*ic + e*ie]. ECONT says that a line's elements are contiguous
// (ie == 1, the x transform): threads then walk the elements of a line; otherwise the lines are
// contiguous (ia == 1, the y transform) and threads walk across lines, so both read and write whole
// 128-byte runs.
#pragma once
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace ftc {

constexpr int kFftMaxPoints = 4096;  // complex points per workgroup: 2048 or 4096

struct HartleyPass {
    const void *in;
    long long ie, ia, ic;
    void *out;
    long long oe, oa, oc;
    int n, logn, G, logG;  // line length and real lines per workgroup (2 * POINTS / n), powers of two
    int A, C;              // lines per plane, planes
    double scale, beta;    // out = beta * out + scale * H(in) - shift; beta == 0 reads nothing of out
    const double2 *tw;     // tw[k] = (cos, -sin)(2 pi k / n), k < n/2
    const double *shift;   // a device scalar subtracted from every output value, or none
};

// LDS bytes of a pass over lines of n points: the complex lines (one padding point each, so that the
// lines of a y transform fall on different banks) and the n/2 twiddles.
__host__ __device__ constexpr int hartley_lds_bytes(int n, int points)
{
    return (points / n > 0 ? points / n : 1) * (n + 1) * 16 + (n / 2) * 16;
}

template <class T>
__device__ __forceinline__ double load_as_double(const T *p)
{
    return (double)*p;
}

template <class T>
__device__ __forceinline__ void store_from_double(T *p, double v)
{
    *p = (T)v;
}

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 w)
{
    return make_double2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}
__device__ __forceinline__ double2 mul_minus_i(double2 a) { return make_double2(a.y, -a.x); }

// In-place R-point DFT, X_k = sum_r v_r e^{-2 pi i r k / R}.
template <int R>
__device__ __forceinline__ void dft(double2 *v)
{
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const double2 t2 = cadd(v[1], v[3]), t3 = mul_minus_i(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else {
        double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        dft<4>(e);
        dft<4>(o);
        const double r = 0.70710678118654752440;  // sqrt(1/2)
        const double2 w1 = make_double2(r, -r), w3 = make_double2(-r, -r);
        const double2 o1 = cmul(o[1], w1), o2 = mul_minus_i(o[2]), o3 = cmul(o[3], w3);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
}

// Butterfly b of a stage of radix R: its complex line and its index j < n/R in the line.
template <bool ECONT>
__device__ __forceinline__ void butterfly_at(int b, int lognr, int logl, int &line, int &j)
{
    if constexpr (ECONT) {
        line = b >> lognr;
        j = b & ((1 << lognr) - 1);
    } else {
        line = b & ((1 << logl) - 1);
        j = b >> logl;
    }
}

// e^{-2 pi i m / n} for 0 <= m < n from the half table.
__device__ __forceinline__ double2 twiddle(const double2 *tw, int m, int n)
{
    const int h = n >> 1;
    if (m < h) return tw[m];
    const double2 w = tw[m - h];
    return make_double2(-w.x, -w.y);
}

// One Stockham stage of radix R on the LDS lines (not the first: reads LDS, applies twiddles).
template <int R, bool ECONT, int PT, int NT>
__device__ __forceinline__ void stockham_stage(double2 *z, const double2 *tw, int n, int logn, int logl, int stride,
                                               int ns, int logns, double2 (&v)[PT])
{
    constexpr int LOGR = R == 8 ? 3 : (R == 4 ? 2 : 1);
    const int lognr = logn - LOGR, nr = 1 << lognr;
#pragma unroll
    for (int t = 0; t < PT / R; ++t) {
        int line, j;
        butterfly_at<ECONT>(threadIdx.x + t * NT, lognr, logl, line, j);
        const double2 *zl = z + line * stride;
#pragma unroll
        for (int r = 0; r < R; ++r) v[t * R + r] = zl[j + r * nr];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PT / R; ++t) {
        int line, j;
        butterfly_at<ECONT>(threadIdx.x + t * NT, lognr, logl, line, j);
        const int k = j & (ns - 1);
        // W^(k r) with W = e^{-2 pi i / (ns R)}: index k r n / (ns R) of the n-point table.
#pragma unroll
        for (int r = 1; r < R; ++r) v[t * R + r] = cmul(v[t * R + r], twiddle(tw, (k * r) << (logn - logns - LOGR), n));
        dft<R>(&v[t * R]);
        double2 *zl = z + line * stride;
        const int d = ((j >> logns) << (logns + LOGR)) + k;
#pragma unroll
        for (int r = 0; r < R; ++r) zl[d + r * ns] = v[t * R + r];
    }
    __syncthreads();
}

// The first Stockham stage (ns = 1: no twiddles, results at j*R + r): operands straight from global
// memory, two real lines (a, a+1) per complex line; lines past the plane's last read as zero.
// The two real lines of a complex line side by side in memory (ia == 1, the y transform) as one vector.
template <class TI>
struct LinePair;
template <>
struct LinePair<double> {
    using V = double2;
};
template <>
struct LinePair<float> {
    using V = float2;
};

template <int R, class TI, bool ECONT, int PT, int NT>
__device__ __forceinline__ void first_stage(const HartleyPass &h, const TI *in, double2 *z, int a0, int logl,
                                            int stride, double2 (&v)[PT])
{
    constexpr int LOGR = R == 8 ? 3 : (R == 4 ? 2 : 1);
    const int lognr = h.logn - LOGR, nr = 1 << lognr;
    // Lines across threads with both lines of every pair present (A even) and adjacent (ia == 1): each pair
    // is one vector load, aligned because a, ie and ic are even (checked here), which covers a whole
    // 128-byte run per load instruction instead of every other element of it.
    const bool pairs = !ECONT && h.ia == 1 && (h.A & 1) == 0 && (h.ie & 1) == 0 && (h.ic & 1) == 0;
#pragma unroll
    for (int t = 0; t < PT / R; ++t) {
        int line, j;
        butterfly_at<ECONT>(threadIdx.x + t * NT, lognr, logl, line, j);
        const int a = a0 + 2 * line;
        const TI *p0 = in + (long long)a * h.ia, *p1 = p0 + h.ia;
        if (pairs) {
            using V = typename LinePair<TI>::V;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const long long e = (long long)(j + r * nr) * h.ie;
                const V pv = a < h.A ? *reinterpret_cast<const V *>(p0 + e) : V{0, 0};
                v[t * R + r] = make_double2((double)pv.x, (double)pv.y);
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const long long e = (long long)(j + r * nr) * h.ie;
                const double re = a < h.A ? load_as_double(p0 + e) : 0.0;
                const double im = a + 1 < h.A ? load_as_double(p1 + e) : 0.0;
                v[t * R + r] = make_double2(re, im);
            }
        }
        dft<R>(&v[t * R]);
        double2 *zl = z + line * stride;
#pragma unroll
        for (int r = 0; r < R; ++r) zl[j * R + r] = v[t * R + r];
    }
}

template <class TI, class TO, bool ECONT, int POINTS, int NT>
__global__ __launch_bounds__(NT) void hartley_kernel(const HartleyPass h)
{
    extern __shared__ double2 fft_lds[];
    const int n = h.n, logn = h.logn, G = h.G, logG = h.logG;
    const int logl = logG - 1;  // log2 of the complex lines per workgroup
    const int stride = n + 1;
    double2 *z = fft_lds;
    double2 *tw = fft_lds + (G >> 1) * stride;
    const int tid = threadIdx.x;
    const int a0 = blockIdx.x * G, c = blockIdx.y;
    const TI *in = static_cast<const TI *>(h.in) + (long long)c * h.ic;
    TO *out = static_cast<TO *>(h.out) + (long long)c * h.oc;
    for (int k = tid; k < (n >> 1); k += NT) tw[k] = h.tw[k];

    // The radices: radix 8 while at least three factors of two remain, then one radix-4 or radix-2 stage.
    const int eights = logn / 3, rest = logn - 3 * eights;
    constexpr int PT = POINTS / NT;
    double2 v[PT];
    // First stage (ns = 1, no twiddles), operands from global memory.
    if (eights) first_stage<8, TI, ECONT, PT, NT>(h, in, z, a0, logl, stride, v);
    else if (rest == 2) first_stage<4, TI, ECONT, PT, NT>(h, in, z, a0, logl, stride, v);
    else first_stage<2, TI, ECONT, PT, NT>(h, in, z, a0, logl, stride, v);
    __syncthreads();
    int ns = eights ? 8 : (1 << rest), logns = eights ? 3 : rest;
    for (int s = 1; s < eights; ++s) {
        stockham_stage<8, ECONT, PT, NT>(z, tw, n, logn, logl, stride, ns, logns, v);
        ns <<= 3;
        logns += 3;
    }
    if (eights && rest == 2) stockham_stage<4, ECONT, PT, NT>(z, tw, n, logn, logl, stride, ns, logns, v);
    if (eights && rest == 1) stockham_stage<2, ECONT, PT, NT>(z, tw, n, logn, logl, stride, ns, logns, v);

    const double sc = 0.5 * h.scale, beta = h.beta;
    for (int idx = tid; idx < G * n; idx += NT) {
        int g, k;
        if constexpr (ECONT) {
            g = idx >> logn;
            k = idx & (n - 1);
        } else {
            g = idx & (G - 1);
            k = idx >> logG;
        }
        const int a = a0 + g;
        if (a >= h.A) continue;
        const double2 *zl = z + (g >> 1) * stride;
        const double2 Z = zl[k], W = zl[(n - k) & (n - 1)];
        const double hv = (g & 1) ? ((Z.y + W.y) + (Z.x - W.x)) : ((Z.x + W.x) - (Z.y - W.y));
        TO *dst = out + (long long)a * h.oa + (long long)k * h.oe;
        const double val = sc * hv;
        double o = beta == 0.0 ? val : beta * load_as_double(dst) + val;
        if (h.shift) o = o - *h.shift;
        store_from_double(dst, o);
    }
}

}  // namespace ftc
