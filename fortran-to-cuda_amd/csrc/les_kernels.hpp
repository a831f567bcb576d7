// Device code of the staggered-grid LES stencils (see les.hip for the entries and the numerical
// contract): the fused k-marching kernels and, for persistent grouped launches, the per-tile body of the
// LDS-staged Smagorinsky kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "xcd.hpp"

#pragma clang fp contract(off)

namespace ftc {

constexpr int TX = 64;
constexpr int TY = 4;

struct LesArgs {
    const double *U, *V, *W, *S;
    double *OU, *OV, *OW, *OS;
    int nx, ny, nz;  // i loop 2..nx+1, j loop 2..ny+1, U/V/S k loop 2..nz+1
    int kw;          // W k loop 2..kw
    int kchunk;
    int tile_order;  // 0: tiles in launch order; 1: each XCD takes a contiguous run of tiles (xcd.hpp);
                     // 2: persistent banded walk (one resident round, les_sgs3_kernel)
    long long sj, sk;                   // strides of j and k in elements
    double c25[3], c50[3], c125[3];     // 0.25/d, 0.5/d, 0.125/d for d = dx, dy, dz
    double sgs[10];                     // 1/d, 0.25/d, 1/d^2 (x, y, z each), (Cs*Delta)^2
    double dt;
};

enum : int { kU = 1, kV = 2, kW = 4, kS = 8 };

__device__ __forceinline__ double cdu_point(const LesArgs &a, long long p)
{
    const double *U = a.U, *V = a.V, *W = a.W;
    const long long sj = a.sj, sk = a.sk;
    const double u0 = U[p], uxp = U[p + 1], uxm = U[p - 1], uyp = U[p + sj], uym = U[p - sj];
    const double uzp = U[p + sk], uzm = U[p - sk];
    const double v0 = V[p], vxp = V[p + 1], vym = V[p - sj], vxpym = V[p + 1 - sj];
    const double w0 = W[p], wxp = W[p + 1], wzm = W[p - sk], wxpzm = W[p + 1 - sk];
    const double Ax = a.c25[0], Ay = a.c25[1], Az = a.c25[2];
    double x = Ax * (uxp + u0) * (uxp + u0) - Ax * (u0 + uxm) * (u0 + uxm);
    double y = Ay * (uyp + u0) * (vxp + v0) - Ay * (u0 + uym) * (vxpym + vym);
    double z = Az * (uzp + u0) * (wxp + w0) - Az * (u0 + uzm) * (wxpzm + wzm);
    double t = -((x + y) + z);
    double Vadv = ((v0 + vxp) + vym) + vxpym;
    double Wadv = ((w0 + wxp) + wzm) + wxpzm;
    t = t - ((a.c50[0] * (uxp - uxm) * u0 + a.c125[1] * (uyp - uym) * Vadv) + a.c125[2] * (uzp - uzm) * Wadv);
    return t * 0.5;
}

__device__ __forceinline__ double cdv_point(const LesArgs &a, long long p)
{
    const double *U = a.U, *V = a.V, *W = a.W;
    const long long sj = a.sj, sk = a.sk;
    const double v0 = V[p], vxp = V[p + 1], vxm = V[p - 1], vyp = V[p + sj], vym = V[p - sj];
    const double vzp = V[p + sk], vzm = V[p - sk];
    const double u0 = U[p], uyp = U[p + sj], uxm = U[p - 1], uxmyp = U[p - 1 + sj];
    const double w0 = W[p], wyp = W[p + sj], wzm = W[p - sk], wypzm = W[p + sj - sk];
    const double Ax = a.c25[0], Ay = a.c25[1], Az = a.c25[2];
    double y = Ay * (vyp + v0) * (vyp + v0) - Ay * (v0 + vym) * (v0 + vym);
    double x = Ax * (vxp + v0) * (uyp + u0) - Ax * (v0 + vxm) * (uxmyp + uxm);
    double z = Az * (vzp + v0) * (wyp + w0) - Az * (v0 + vzm) * (wypzm + wzm);
    double t = -((y + x) + z);
    double Uadv = ((u0 + uyp) + uxm) + uxmyp;
    double Wadv = ((w0 + wyp) + wzm) + wypzm;
    t = t - ((a.c125[0] * (vxp - vxm) * Uadv + a.c50[1] * (vyp - vym) * v0) + a.c125[2] * (vzp - vzm) * Wadv);
    return t * 0.5;
}

__device__ __forceinline__ double cdw_point(const LesArgs &a, long long p)
{
    const double *U = a.U, *V = a.V, *W = a.W;
    const long long sj = a.sj, sk = a.sk;
    const double w0 = W[p], wxp = W[p + 1], wxm = W[p - 1], wyp = W[p + sj], wym = W[p - sj];
    const double wzp = W[p + sk], wzm = W[p - sk];
    const double u0 = U[p], uzp = U[p + sk], uxm = U[p - 1], uxmzp = U[p - 1 + sk];
    const double v0 = V[p], vzp = V[p + sk], vym = V[p - sj], vymzp = V[p - sj + sk];
    const double Ax = a.c25[0], Ay = a.c25[1], Az = a.c25[2];
    double z = Az * (wzp + w0) * (wzp + w0) - Az * (w0 + wzm) * (w0 + wzm);
    double y = Ay * (wyp + w0) * (vzp + v0) - Ay * (w0 + wym) * (vym + vymzp);
    double x = Ax * (wxp + w0) * (uzp + u0) - Ax * (w0 + wxm) * (uxmzp + uxm);
    double t = -((z + y) + x);
    double Uadv = ((u0 + uzp) + uxm) + uxmzp;
    double Vadv = ((v0 + vzp) + vym) + vymzp;
    t = t - ((a.c125[0] * (wxp - wxm) * Uadv + a.c125[1] * (wyp - wym) * Vadv) + a.c50[2] * (wzp - wzm) * w0);
    return t * 0.5;
}

__device__ __forceinline__ double sadv_point(const LesArgs &a, long long p)
{
    const double *U = a.U, *V = a.V, *W = a.W, *S = a.S;
    const long long sj = a.sj, sk = a.sk;
    const double s0 = S[p];
    const double Cx = a.c50[0], Cy = a.c50[1], Cz = a.c50[2];
    double x = Cx * U[p] * (S[p + 1] + s0) - Cx * U[p - 1] * (s0 + S[p - 1]);
    double y = Cy * V[p] * (S[p + sj] + s0) - Cy * V[p - sj] * (s0 + S[p - sj]);
    double z = Cz * W[p] * (S[p + sk] + s0) - Cz * W[p - sk] * (s0 + S[p - sk]);
    return -((x + y) + z);
}

// Smagorinsky viscosity at an interior cell whose linear index is p.
__device__ __forceinline__ double nu_point(const LesArgs &a, long long p)
{
    const double *U = a.U, *V = a.V, *W = a.W;
    const long long sj = a.sj, sk = a.sk;
    const double *q = a.sgs;
    double s11 = (U[p] - U[p - 1]) * q[0];
    double s22 = (V[p] - V[p - sj]) * q[1];
    double s33 = (W[p] - W[p - sk]) * q[2];
    double dudy = ((U[p + sj] + U[p - 1 + sj]) - (U[p - sj] + U[p - 1 - sj])) * q[4];
    double dvdx = ((V[p + 1] + V[p + 1 - sj]) - (V[p - 1] + V[p - 1 - sj])) * q[3];
    double dudz = ((U[p + sk] + U[p - 1 + sk]) - (U[p - sk] + U[p - 1 - sk])) * q[5];
    double dwdx = ((W[p + 1] + W[p + 1 - sk]) - (W[p - 1] + W[p - 1 - sk])) * q[3];
    double dvdz = ((V[p + sk] + V[p - sj + sk]) - (V[p - sk] + V[p - sj - sk])) * q[5];
    double dwdy = ((W[p + sj] + W[p + sj - sk]) - (W[p - sj] + W[p - sj - sk])) * q[4];
    double s12 = 0.5 * (dudy + dvdx);
    double s13 = 0.5 * (dudz + dwdx);
    double s23 = 0.5 * (dvdz + dwdy);
    double ss = 2.0 * ((s11 * s11 + s22 * s22) + s33 * s33) + 4.0 * ((s12 * s12 + s13 * s13) + s23 * s23);
    return q[9] * sqrt(ss);
}

template <int R>
using NuRing = double[R][TY + 2][TX + 2];

// Fills ring slot kp % 3 with nu at plane kp over the tile and its one-cell ring. A ghost cell takes the
// value of its image: periodic in x and y, zero-gradient in z.
__device__ __forceinline__ void nu_plane(const LesArgs &a, NuRing<3> &ring, int kp, int i0, int j0, int tid)
{
    const int kk = kp < 2 ? 2 : (kp > a.nz + 1 ? a.nz + 1 : kp);
    double(*plane)[TX + 2] = ring[kp % 3];
    for (int c = tid; c < (TX + 2) * (TY + 2); c += TX * TY) {
        const int li = c % (TX + 2), lj = c / (TX + 2);
        const int gi = i0 - 1 + li, gj = j0 - 1 + lj;
        double value = 0.0;
        if (gi <= a.nx + 2 && gj <= a.ny + 2) {
            const int ii = gi == 1 ? a.nx + 1 : (gi == a.nx + 2 ? 2 : gi);
            const int jj = gj == 1 ? a.ny + 1 : (gj == a.ny + 2 ? 2 : gj);
            value = nu_point(a, (long long)(ii - 1) + (long long)(jj - 1) * a.sj + (long long)(kk - 1) * a.sk);
        }
        plane[lj][li] = value;
    }
}

// Writes a new interior value and every ghost whose image it is (the result of the x, then y, then z
// halo passes). zmode: 0 tangential velocity (sign change at the walls), 1 wall-normal velocity (no z
// ghost), 2 scalar (zero gradient).
__device__ __forceinline__ void put_with_images(const LesArgs &a, double *F, long long p, double val, int i, int j,
                                                int k, int zmode)
{
    long long ox[3] = {0, 0, 0}, oy[3] = {0, 0, 0}, oz[2] = {0, 0};
    int nxo = 1, nyo = 1, nzo = 0;
    if (i == 2) ox[nxo++] = a.nx;
    if (i == a.nx + 1) ox[nxo++] = -a.nx;
    if (j == 2) oy[nyo++] = (long long)a.ny * a.sj;
    if (j == a.ny + 1) oy[nyo++] = -(long long)a.ny * a.sj;
    if (zmode != 1) {
        if (k == 2) oz[nzo++] = -a.sk;
        if (k == a.nz + 1) oz[nzo++] = a.sk;
    }
    const double zval = zmode == 0 ? -val : val;
    for (int x = 0; x < nxo; ++x)
        for (int y = 0; y < nyo; ++y) {
            const long long q = p + ox[x] + oy[y];
            F[q] = val;
            for (int z = 0; z < nzo; ++z) F[q + oz[z]] = zval;
        }
}

template <int MASK, bool SGS, bool STEP>
__global__ __launch_bounds__(TX *TY) void les_kernel(const LesArgs a)
{
    __shared__ NuRing<SGS ? 3 : 1> ring;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const int i0 = 2 + blockIdx.x * TX, j0 = 2 + blockIdx.y * TY;
    const int i = i0 + tx, j = j0 + ty;
    int klast = (MASK & (kU | kV | kS)) ? a.nz + 1 : 0;
    if ((MASK & kW) && a.kw > klast) klast = a.kw;
    const int k0 = 2 + (int)blockIdx.z * a.kchunk;
    if (k0 > klast) return;  // uniform over the workgroup
    const int k1 = min(k0 + a.kchunk - 1, klast);
    const bool inside = i <= a.nx + 1 && j <= a.ny + 1;
    if constexpr (SGS) {
        nu_plane(a, ring, k0 - 1, i0, j0, tid);
        nu_plane(a, ring, k0, i0, j0, tid);
    }
    const double dt = a.dt;
    const double *q = a.sgs;
    for (int k = k0; k <= k1; ++k) {
        if constexpr (SGS) {
            nu_plane(a, ring, k + 1, i0, j0, tid);
            __syncthreads();
        }
        if (inside) {
            const long long p = (long long)(i - 1) + (long long)(j - 1) * a.sj + (long long)(k - 1) * a.sk;
            const int lx = tx + 1, ly = ty + 1;
            const int slm = (k - 1) % 3, slz = k % 3, slp = (k + 1) % 3;
#define NU(di, dj, dk) ring[(dk) < 0 ? slm : ((dk) > 0 ? slp : slz)][ly + (dj)][lx + (di)]
            const long long sj = a.sj, sk = a.sk;
            if ((MASK & kU) && k <= a.nz + 1) {
                double t = cdu_point(a, p);
                if constexpr (SGS) {
                    const double *U = a.U;
                    const double u0 = U[p];
                    double fxp = NU(1, 0, 0) * (U[p + 1] - u0);
                    double fxm = NU(0, 0, 0) * (u0 - U[p - 1]);
                    double exy0 = 0.25 * (((NU(0, 0, 0) + NU(1, 0, 0)) + NU(0, 1, 0)) + NU(1, 1, 0));
                    double exym = 0.25 * (((NU(0, -1, 0) + NU(1, -1, 0)) + NU(0, 0, 0)) + NU(1, 0, 0));
                    double fyp = exy0 * (U[p + sj] - u0);
                    double fym = exym * (u0 - U[p - sj]);
                    double exz0 = 0.25 * (((NU(0, 0, 0) + NU(1, 0, 0)) + NU(0, 0, 1)) + NU(1, 0, 1));
                    double exzm = 0.25 * (((NU(0, 0, -1) + NU(1, 0, -1)) + NU(0, 0, 0)) + NU(1, 0, 0));
                    double fzp = exz0 * (U[p + sk] - u0);
                    double fzm = exzm * (u0 - U[p - sk]);
                    t = t + (((fxp - fxm) * q[6] + (fyp - fym) * q[7]) + (fzp - fzm) * q[8]);
                }
                if constexpr (STEP) put_with_images(a, a.OU, p, a.U[p] + dt * t, i, j, k, 0);
                else a.OU[p] = t;
            }
            if ((MASK & kV) && k <= a.nz + 1) {
                double t = cdv_point(a, p);
                if constexpr (SGS) {
                    const double *V = a.V;
                    const double v0 = V[p];
                    double exy0 = 0.25 * (((NU(0, 0, 0) + NU(1, 0, 0)) + NU(0, 1, 0)) + NU(1, 1, 0));
                    double exyxm = 0.25 * (((NU(-1, 0, 0) + NU(0, 0, 0)) + NU(-1, 1, 0)) + NU(0, 1, 0));
                    double fxp = exy0 * (V[p + 1] - v0);
                    double fxm = exyxm * (v0 - V[p - 1]);
                    double fyp = NU(0, 1, 0) * (V[p + sj] - v0);
                    double fym = NU(0, 0, 0) * (v0 - V[p - sj]);
                    double eyz0 = 0.25 * (((NU(0, 0, 0) + NU(0, 1, 0)) + NU(0, 0, 1)) + NU(0, 1, 1));
                    double eyzm = 0.25 * (((NU(0, 0, -1) + NU(0, 1, -1)) + NU(0, 0, 0)) + NU(0, 1, 0));
                    double fzp = eyz0 * (V[p + sk] - v0);
                    double fzm = eyzm * (v0 - V[p - sk]);
                    t = t + (((fxp - fxm) * q[6] + (fyp - fym) * q[7]) + (fzp - fzm) * q[8]);
                }
                if constexpr (STEP) put_with_images(a, a.OV, p, a.V[p] + dt * t, i, j, k, 0);
                else a.OV[p] = t;
            }
            if ((MASK & kW) && k <= a.kw) {
                double t = cdw_point(a, p);
                if constexpr (SGS) {
                    const double *W = a.W;
                    const double w0 = W[p];
                    double exz0 = 0.25 * (((NU(0, 0, 0) + NU(1, 0, 0)) + NU(0, 0, 1)) + NU(1, 0, 1));
                    double exzxm = 0.25 * (((NU(-1, 0, 0) + NU(0, 0, 0)) + NU(-1, 0, 1)) + NU(0, 0, 1));
                    double fxp = exz0 * (W[p + 1] - w0);
                    double fxm = exzxm * (w0 - W[p - 1]);
                    double eyz0 = 0.25 * (((NU(0, 0, 0) + NU(0, 1, 0)) + NU(0, 0, 1)) + NU(0, 1, 1));
                    double eyzym = 0.25 * (((NU(0, -1, 0) + NU(0, 0, 0)) + NU(0, -1, 1)) + NU(0, 0, 1));
                    double fyp = eyz0 * (W[p + sj] - w0);
                    double fym = eyzym * (w0 - W[p - sj]);
                    double fzp = NU(0, 0, 1) * (W[p + sk] - w0);
                    double fzm = NU(0, 0, 0) * (w0 - W[p - sk]);
                    t = t + (((fxp - fxm) * q[6] + (fyp - fym) * q[7]) + (fzp - fzm) * q[8]);
                }
                if constexpr (STEP) put_with_images(a, a.OW, p, a.W[p] + dt * t, i, j, k, 1);
                else a.OW[p] = t;
            }
            if ((MASK & kS) && k <= a.nz + 1) {
                double t = sadv_point(a, p);
                if constexpr (SGS) {
                    const double *S = a.S;
                    const double s0 = S[p];
                    double fxp = (0.5 * (NU(0, 0, 0) + NU(1, 0, 0))) * (S[p + 1] - s0);
                    double fxm = (0.5 * (NU(-1, 0, 0) + NU(0, 0, 0))) * (s0 - S[p - 1]);
                    double fyp = (0.5 * (NU(0, 0, 0) + NU(0, 1, 0))) * (S[p + sj] - s0);
                    double fym = (0.5 * (NU(0, -1, 0) + NU(0, 0, 0))) * (s0 - S[p - sj]);
                    double fzp = (0.5 * (NU(0, 0, 0) + NU(0, 0, 1))) * (S[p + sk] - s0);
                    double fzm = (0.5 * (NU(0, 0, -1) + NU(0, 0, 0))) * (s0 - S[p - sk]);
                    t = t + (((fxp - fxm) * q[6] + (fyp - fym) * q[7]) + (fzp - fzm) * q[8]);
                }
                if constexpr (STEP) put_with_images(a, a.OS, p, a.S[p] + dt * t, i, j, k, 2);
                else a.OS[p] = t;
            }
#undef NU
        }
        if constexpr (SGS) __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------ staging
// The Smagorinsky kernel stages every input plane through LDS once: a workgroup keeps its tile's velocity
// planes with a two-cell ring and its scalar and viscosity planes with a one-cell ring. The global loads
// of the next plane are issued into registers while the current one is computed, so each input value is
// read from HBM about once per workgroup and every stencil operand after that comes from LDS.
//
// Precondition (documented at the entries): the x and y ghost layers of U, V, W and S equal their
// periodic images, as every field written by the step or the halo fill is. The staged tile is gathered
// through periodic images, so a ghost cell's viscosity is the viscosity of its image.
constexpr int VX = TX + 4, VY = TY + 4;  // velocity tile with a two-cell ring
constexpr int CX = TX + 2, CY = TY + 2;  // scalar and viscosity tile with a one-cell ring
constexpr int NTH = TX * TY;
constexpr int VCELLS = VX * VY, CCELLS = CX * CY;
constexpr int VR = (VCELLS + NTH - 1) / NTH, CR = (CCELLS + NTH - 1) / NTH;

__device__ __forceinline__ int periodic_image(int c, int n)
{
    int r = (c - 2) % n;
    if (r < 0) r += n;
    return r + 2;
}

// ------------------------------------------------------------------------------------------------ step kernel
// put_with_images with the column's x/y image offsets computed once per thread (they do not depend on k)
// and the z images decided per plane (uniform over the workgroup): a cell off every boundary costs one
// store. The outputs are not read again by the launch, so they are written nontemporally.
struct ColumnImages {
    int ox0, ox1;            // x offsets of the ghost images of the column (nox of them)
    long long oy0, oy1;      // y offsets (noy of them)
    int nox, noy;
};

__device__ __forceinline__ ColumnImages column_images(const LesArgs &a, int i, int j)
{
    ColumnImages c{0, 0, 0, 0, 0, 0};
    if (i == 2) c.ox0 = a.nx, c.nox = 1;
    if (i == a.nx + 1) {
        if (c.nox) c.ox1 = -a.nx;
        else c.ox0 = -a.nx;
        c.nox++;
    }
    if (j == 2) c.oy0 = (long long)a.ny * a.sj, c.noy = 1;
    if (j == a.ny + 1) {
        if (c.noy) c.oy1 = -(long long)a.ny * a.sj;
        else c.oy0 = -(long long)a.ny * a.sj;
        c.noy++;
    }
    return c;
}

// The z ghost offsets of plane k for fields other than W (W has no z ghost written by the step).
struct PlaneImages {
    long long oz0, oz1;
    int noz;
};

__device__ __forceinline__ PlaneImages plane_images(const LesArgs &a, int k)
{
    PlaneImages z{0, 0, 0};
    if (k == 2) z.oz0 = -a.sk, z.noz = 1;
    if (k == a.nz + 1) {
        if (z.noz) z.oz1 = a.sk;
        else z.oz0 = a.sk;
        z.noz++;
    }
    return z;
}

__device__ __forceinline__ void put_value(double *F, long long q, double v) { __builtin_nontemporal_store(v, &F[q]); }

// Writes val at p and at every ghost image put_with_images writes, in the same cases: the x and y images
// (and their combinations), and for zmode != 1 the z images, negated for zmode 0.
__device__ __forceinline__ void put_column(double *F, long long p, double val, const ColumnImages &c,
                                           const PlaneImages &z, int zmode)
{
    const double zval = zmode == 0 ? -val : val;
    const int noz = zmode == 1 ? 0 : z.noz;
    auto with_z = [&](long long q) {
        put_value(F, q, val);
        if (noz > 0) put_value(F, q + z.oz0, zval);
        if (noz > 1) put_value(F, q + z.oz1, zval);
    };
    with_z(p);
    if (c.nox | c.noy) {
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int y = 0; y < 3; ++y) {
                if ((x == 0 && y == 0) || x > c.nox || y > c.noy) continue;
                const long long q = p + (x == 1 ? c.ox0 : (x == 2 ? c.ox1 : 0)) + (y == 1 ? c.oy0 : (y == 2 ? c.oy1 : 0));
                with_z(q);
            }
    }
}

// The v2 staging with the z-direction fluxes carried in registers. In flux form the lower z-face term of
// cell k is the upper z-face term of cell k-1, computed from the same operands in the same order, so a
// thread marching k keeps the upper terms of step k-1 (and its own column's k-1 velocities) and reads
// nothing of plane k-1 from LDS. The rings shrink to three velocity planes (k .. k+2), two scalar planes
// and two viscosity planes: 52 KB, three workgroups per CU. Every value is the one v2 and the oracle
// compute; only where it is kept changes.
// ROWS = 2 gives each thread two adjacent rows of a 64 x 8 tile (80 KB, two workgroups per CU): less halo
// per cell (viscosity ring 1.29x instead of 1.55x, velocity ring 1.59x instead of 2.13x) and the per-step
// bookkeeping shared by two cells.
// WX is the tile width: the 256 threads form WX x (256 / WX) (32 x 8, 64 x 4 or 128 x 2), each thread
// ROWS rows. Narrow tiles waste fewer lanes on narrow domains and recompute less ring (32 x 8: viscosity
// ring 1.33x, velocity ring 1.69x); wide tiles read longer rows.
template <int ROWS, int WX = TX>
struct Sgs3Tile {
    static_assert(WX == 32 || WX == 64 || WX == 128, "tile widths 32, 64 or 128");
    static constexpr int TXT = WX, TYB = NTH / WX;   // thread block
    static constexpr int TYT = TYB * ROWS;           // rows of the tile
    static constexpr int VXT = WX + 4, CXT = WX + 2;
    static constexpr int VYT = TYT + 4, CYT = TYT + 2;
    static constexpr int VC = VXT * VYT, CC = CXT * CYT;
    static constexpr int VRT = (VC + NTH - 1) / NTH, CRT = (CC + NTH - 1) / NTH;
};

// W is read at planes k and k+1 only while a step computes plane k (the viscosity of k+1 takes W at k+1 and
// k, the update of k takes W at k with its ring and at k+1 in the column), so it takes two slots where U and
// V, whose plane k+2 the viscosity of k+1 reads, take three; W is staged one plane behind them.
//
// TWO: two barriers per k-step instead of three. The third barrier keeps the next step's puts (U, V at k+3,
// W and S at k+2) off the slots the update of k still reads; with one more slot for each of them (4, 3, 3)
// those puts land on the planes of k-1, which no thread reads any more. The viscosity keeps two slots: the
// next step writes it only after its first barrier, which every update of k has reached.
template <int ROWS, int WX = TX, bool TWO = false>
struct Sgs3Lds {
    using T = Sgs3Tile<ROWS, WX>;
    static constexpr int SU = TWO ? 4 : 3, SW = TWO ? 3 : 2, SS = TWO ? 3 : 2;
    double uv[SU][2][T::VYT][T::VXT];  // U, V; slot = plane % SU
    double w[SW][T::VYT][T::VXT];      // W; slot = plane % SW
    double sca[SS][T::CYT][T::CXT];    // slot = plane % SS
    double nu[2][T::CYT][T::CXT];      // slot = plane & 1
};

// The upper z-face terms of a cell (advection and diffusion of u, v, w and s), which are the lower terms
// of the cell above.
struct ZCarry {
    double cu, du, cv, dv, cw, dw, cs, ds;  // flux-form terms
    double u, v, w, wxp, wyp;               // U, V, W at (i, j), W at (i+1, j) and (i, j+1)
};

// ISO: the grid spacings are equal (dx == dy == dz), so each per-direction coefficient is one value. The
// arithmetic is the same (the same values in the same order); the kernel only holds 7 coefficients instead of
// 20, which keeps them in scalar registers instead of spilling them. 0.25/d serves as both the advection's
// c25 and the viscosity's quarter-difference factor (the host computes both as 0.25/d).
template <bool STEP, int ROWS = 1, int WX = TX, bool ISO = false, bool TWO = false>
__device__ __forceinline__ void les_sgs3_body(const LesArgs &a, Sgs3Lds<ROWS, WX, TWO> &L, int bx, int by, int bz)
{
    using T = Sgs3Tile<ROWS, WX>;
    constexpr int SU = Sgs3Lds<ROWS, WX, TWO>::SU, SW = Sgs3Lds<ROWS, WX, TWO>::SW, SS = Sgs3Lds<ROWS, WX, TWO>::SS;
    constexpr int VR = T::VRT, CR = T::CRT, VCELLS = T::VC, CCELLS = T::CC;
    constexpr int VX = T::VXT, CX = T::CXT;
    const double *__restrict__ gU = a.U;
    const double *__restrict__ gV = a.V;
    const double *__restrict__ gW = a.W;
    const double *__restrict__ gS = a.S;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * WX + tx;
    const int i0 = 2 + bx * WX, j0 = 2 + by * T::TYT;
    const int i = i0 + tx;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = 2 + bz * a.kchunk;
    if (k0 > nz + 1) return;  // uniform over the workgroup
    const int k1 = min(k0 + a.kchunk - 1, nz + 1);
    const long long sj = a.sj, sk = a.sk;

    long long voff[VR], coff[CR];
    int vcell[VR], ccell[CR];
#pragma unroll
    for (int r = 0; r < VR; ++r) {
        const int c = tid + r * NTH;
        vcell[r] = c < VCELLS ? c : -1;
        const int cc = c < VCELLS ? c : 0;
        const int ii = periodic_image(i0 - 2 + cc % VX, nx), jj = periodic_image(j0 - 2 + cc / VX, ny);
        voff[r] = (long long)(ii - 1) + (long long)(jj - 1) * sj;
    }
#pragma unroll
    for (int r = 0; r < CR; ++r) {
        const int c = tid + r * NTH;
        ccell[r] = c < CCELLS ? c : -1;
        const int cc = c < CCELLS ? c : 0;
        const int ii = periodic_image(i0 - 1 + cc % CX, nx), jj = periodic_image(j0 - 1 + cc / CX, ny);
        coff[r] = (long long)(ii - 1) + (long long)(jj - 1) * sj;
    }
    const double c25x = a.c25[0], c25y = ISO ? c25x : a.c25[1], c25z = ISO ? c25x : a.c25[2];
    const double c50x = a.c50[0], c50y = ISO ? c50x : a.c50[1], c50z = ISO ? c50x : a.c50[2];
    const double c125x = a.c125[0], c125y = ISO ? c125x : a.c125[1], c125z = ISO ? c125x : a.c125[2];
    const double ix = a.sgs[0], iy = ISO ? ix : a.sgs[1], iz = ISO ? ix : a.sgs[2];  // 1/d
    const double hx = a.sgs[6], hy = ISO ? hx : a.sgs[7], hz = ISO ? hx : a.sgs[8];  // 1/d^2
    const double cs2 = a.sgs[9];                                                    // (Cs Delta)^2
    double rv[VR][2], rw[VR], rs[CR];
    auto fetch_uv = [&](int kp) {
        if (kp < 1 || kp > nz + 2) return;
        const long long base = (long long)(kp - 1) * sk;
#pragma unroll
        for (int r = 0; r < VR; ++r)
            if (vcell[r] >= 0) {
                rv[r][0] = gU[voff[r] + base];
                rv[r][1] = gV[voff[r] + base];
            }
    };
    auto put_uv = [&](int kp) {
        if (kp < 1 || kp > nz + 2) return;
        const int slot = kp % SU;
#pragma unroll
        for (int r = 0; r < VR; ++r)
            if (vcell[r] >= 0) {
                (&L.uv[slot][0][0][0])[vcell[r]] = rv[r][0];
                (&L.uv[slot][1][0][0])[vcell[r]] = rv[r][1];
            }
    };
    // W planes up to nz+1: the update reads W at k+1 only on the interior faces (k <= kw = nz), and the
    // clamped viscosity of plane nz+2 reads W at nz+1 and nz, which a plane nz+2 would overwrite.
    auto fetch_w = [&](int kp) {
        if (kp < 1 || kp > nz + 1) return;
        const long long base = (long long)(kp - 1) * sk;
#pragma unroll
        for (int r = 0; r < VR; ++r)
            if (vcell[r] >= 0) rw[r] = gW[voff[r] + base];
    };
    auto put_w = [&](int kp) {
        if (kp < 1 || kp > nz + 1) return;
        const int slot = kp % SW;
#pragma unroll
        for (int r = 0; r < VR; ++r)
            if (vcell[r] >= 0) (&L.w[slot][0][0])[vcell[r]] = rw[r];
    };
    auto fetch_sca = [&](int kp) {
        if (kp < 1 || kp > nz + 2) return;
        const long long base = (long long)(kp - 1) * sk;
#pragma unroll
        for (int r = 0; r < CR; ++r)
            if (ccell[r] >= 0) rs[r] = gS[coff[r] + base];
    };
    auto put_sca = [&](int kp) {
        if (kp < 1 || kp > nz + 2) return;
        const int slot = kp % SS;
#pragma unroll
        for (int r = 0; r < CR; ++r)
            if (ccell[r] >= 0) (&L.sca[slot][0][0])[ccell[r]] = rs[r];
    };
    // nu at plane qk (clamped to the interior planes: zero gradient in z) over the tile and its ring.
    auto nu_plane = [&](int qk) {
        const int kk = qk < 2 ? 2 : (qk > nz + 1 ? nz + 1 : qk);
        const int s0 = kk % SU, sm = (kk - 1) % SU, sp = (kk + 1) % SU;
        const int w0s = kk % SW, wms = (kk - 1) % SW;
        double(*nuo)[CX] = L.nu[qk & 1];
        for (int c = tid; c < CCELLS; c += NTH) {
            const int li = c % CX, lj = c / CX;
            const int vx = li + 1, vy = lj + 1;
#define QU(s, di, dj) L.uv[s][0][vy + (dj)][vx + (di)]
#define QV(s, di, dj) L.uv[s][1][vy + (dj)][vx + (di)]
#define QW(s, di, dj) L.w[s][vy + (dj)][vx + (di)]
            double s11 = (QU(s0, 0, 0) - QU(s0, -1, 0)) * ix;
            double s22 = (QV(s0, 0, 0) - QV(s0, 0, -1)) * iy;
            double s33 = (QW(w0s, 0, 0) - QW(wms, 0, 0)) * iz;
            double dudy = ((QU(s0, 0, 1) + QU(s0, -1, 1)) - (QU(s0, 0, -1) + QU(s0, -1, -1))) * c25y;
            double dvdx = ((QV(s0, 1, 0) + QV(s0, 1, -1)) - (QV(s0, -1, 0) + QV(s0, -1, -1))) * c25x;
            double dudz = ((QU(sp, 0, 0) + QU(sp, -1, 0)) - (QU(sm, 0, 0) + QU(sm, -1, 0))) * c25z;
            double dwdx = ((QW(w0s, 1, 0) + QW(wms, 1, 0)) - (QW(w0s, -1, 0) + QW(wms, -1, 0))) * c25x;
            double dvdz = ((QV(sp, 0, 0) + QV(sp, 0, -1)) - (QV(sm, 0, 0) + QV(sm, 0, -1))) * c25z;
            double dwdy = ((QW(w0s, 0, 1) + QW(wms, 0, 1)) - (QW(w0s, 0, -1) + QW(wms, 0, -1))) * c25y;
#undef QU
#undef QV
#undef QW
            double s12 = 0.5 * (dudy + dvdx);
            double s13 = 0.5 * (dudz + dwdx);
            double s23 = 0.5 * (dvdz + dwdy);
            double ss = 2.0 * ((s11 * s11 + s22 * s22) + s33 * s33) + 4.0 * ((s12 * s12 + s13 * s13) + s23 * s23);
            nuo[lj][li] = cs2 * sqrt(ss);
        }
    };

    const double dt = a.dt;
    const double Cz = c50z;
#define VU(s, di, dj) L.uv[s][0][ly + 2 + (dj)][tx + 2 + (di)]
#define VV(s, di, dj) L.uv[s][1][ly + 2 + (dj)][tx + 2 + (di)]
#define VW(s, di, dj) L.w[s][ly + 2 + (dj)][tx + 2 + (di)]
#define SS(s, di, dj) L.sca[s][ly + 1 + (dj)][tx + 1 + (di)]
#define NV(s, di, dj) L.nu[s][ly + 1 + (dj)][tx + 1 + (di)]

    // Prologue: U, V at planes k0-2 .. k0 and W at k0-2, k0-1 give nu(k0-1); U, V at k0+1 and W at k0 then
    // replace the planes k0-2 and give nu(k0); then the upper terms of plane k0-1 (the lower terms of k0)
    // are formed as a step forms them.
    for (int kp = k0 - 2; kp <= k0; ++kp) {
        fetch_uv(kp);
        put_uv(kp);
    }
    for (int kp = k0 - 2; kp <= k0 - 1; ++kp) {
        fetch_w(kp);
        put_w(kp);
    }
    for (int kp = k0 - 1; kp <= k0; ++kp) {
        fetch_sca(kp);
        put_sca(kp);
    }
    fetch_uv(k0 + 1);
    fetch_w(k0);
    // nu(k0-1) is clamped to plane 2 when k0 = 2 and then needs plane k0+1 (plane k0-2 does not exist).
    const bool clamped = k0 - 1 < 2;
    __syncthreads();
    if (!clamped) nu_plane(k0 - 1);
    __syncthreads();
    put_uv(k0 + 1);
    put_w(k0);
    __syncthreads();
    if (clamped) nu_plane(k0 - 1);
    nu_plane(k0);
    __syncthreads();
    ZCarry zcr[ROWS] = {};
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int ly = ty * ROWS + rr, j = j0 + ly;
        if (i <= nx + 1 && j <= ny + 1) {
        ZCarry &zc = zcr[rr];
        const int kv = k0 - 1;
        const int s0 = kv % SU, sp = (kv + 1) % SU, c0 = kv & 1, cp = (kv + 1) & 1;
        const int ws0 = kv % SW, wsp = (kv + 1) % SW, ss0 = kv % SS, ssp = (kv + 1) % SS;
        const double u0 = VU(s0, 0, 0), uzp = VU(sp, 0, 0);
        const double v0 = VV(s0, 0, 0), vzp = VV(sp, 0, 0);
        const double w0 = VW(ws0, 0, 0), wxp = VW(ws0, 1, 0), wyp = VW(ws0, 0, 1), wzp = VW(wsp, 0, 0);
        const double n000 = NV(c0, 0, 0), n100 = NV(c0, 1, 0), n010 = NV(c0, 0, 1), n001 = NV(cp, 0, 0);
        zc.cu = c25z * (uzp + u0) * (wxp + w0);
        zc.du = (0.25 * (((n000 + n100) + n001) + NV(cp, 1, 0))) * (uzp - u0);
        zc.cv = c25z * (vzp + v0) * (wyp + w0);
        zc.dv = (0.25 * (((n000 + n010) + n001) + NV(cp, 0, 1))) * (vzp - v0);
        zc.cw = c25z * (wzp + w0) * (wzp + w0);
        zc.dw = n001 * (wzp - w0);
        const double s0v = SS(ss0, 0, 0), szp = SS(ssp, 0, 0);
        zc.cs = Cz * w0 * (szp + s0v);
        zc.ds = (0.5 * (n000 + n001)) * (szp - s0v);
        zc.u = u0;
        zc.v = v0;
        zc.w = w0;
        zc.wxp = wxp;
        zc.wyp = wyp;
        }
    }
    fetch_uv(k0 + 2);
    fetch_w(k0 + 1);
    fetch_sca(k0 + 1);
    __syncthreads();

    for (int k = k0; k <= k1; ++k) {
        put_uv(k + 2);
        put_w(k + 1);
        put_sca(k + 1);
        fetch_uv(k + 3);
        fetch_w(k + 2);
        fetch_sca(k + 2);
        __syncthreads();
        nu_plane(k + 1);
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) {
            const int ly = ty * ROWS + rr, j = j0 + ly;
            if (i <= nx + 1 && j <= ny + 1) {
            ZCarry &zc = zcr[rr];
            const ColumnImages ci = column_images(a, i, j);
            const PlaneImages zi = plane_images(a, k);
            const int s0 = k % SU, sp = (k + 1) % SU;
            const int c0 = k & 1, cp = (k + 1) & 1;
            const int ws0 = k % SW, wsp = (k + 1) % SW, ss0 = k % SS, ssp = (k + 1) % SS;
            const double u0 = VU(s0, 0, 0), uxp = VU(s0, 1, 0), uxm = VU(s0, -1, 0), uyp = VU(s0, 0, 1);
            const double uym = VU(s0, 0, -1), uzp = VU(sp, 0, 0), uzm = zc.u;
            const double v0 = VV(s0, 0, 0), vxp = VV(s0, 1, 0), vxm = VV(s0, -1, 0), vyp = VV(s0, 0, 1);
            const double vym = VV(s0, 0, -1), vzp = VV(sp, 0, 0), vzm = zc.v;
            const double w0 = VW(ws0, 0, 0), wxp = VW(ws0, 1, 0), wxm = VW(ws0, -1, 0), wyp = VW(ws0, 0, 1);
            const double wym = VW(ws0, 0, -1), wzp = VW(wsp, 0, 0), wzm = zc.w;
            const double n000 = NV(c0, 0, 0), n100 = NV(c0, 1, 0), nm00 = NV(c0, -1, 0);
            const double n010 = NV(c0, 0, 1), n0m0 = NV(c0, 0, -1), n001 = NV(cp, 0, 0);
            const long long p = (long long)(i - 1) + (long long)(j - 1) * sj + (long long)(k - 1) * sk;
            ZCarry up;
            // U: CDU (cdu.f90:48-107) then the diffusion of u.
            {
                const double vxpym = VV(s0, 1, -1), wxpzm = zc.wxp;
                double x = c25x * (uxp + u0) * (uxp + u0) - c25x * (u0 + uxm) * (u0 + uxm);
                double y = c25y * (uyp + u0) * (vxp + v0) - c25y * (u0 + uym) * (vxpym + vym);
                up.cu = c25z * (uzp + u0) * (wxp + w0);
                double z = up.cu - zc.cu;
                double t = -((x + y) + z);
                double Vadv = ((v0 + vxp) + vym) + vxpym;
                double Wadv = ((w0 + wxp) + wzm) + wxpzm;
                t = t - ((c50x * (uxp - uxm) * u0 + c125y * (uyp - uym) * Vadv) + c125z * (uzp - uzm) * Wadv);
                t = t * 0.5;
                double fxp = n100 * (uxp - u0);
                double fxm = n000 * (u0 - uxm);
                double exy0 = 0.25 * (((n000 + n100) + n010) + NV(c0, 1, 1));
                double exym = 0.25 * (((n0m0 + NV(c0, 1, -1)) + n000) + n100);
                double fyp = exy0 * (uyp - u0);
                double fym = exym * (u0 - uym);
                double exz0 = 0.25 * (((n000 + n100) + n001) + NV(cp, 1, 0));
                up.du = exz0 * (uzp - u0);
                double fzp = up.du;
                double fzm = zc.du;
                t = t + (((fxp - fxm) * hx + (fyp - fym) * hy) + (fzp - fzm) * hz);
                if constexpr (STEP) put_column(a.OU, p, u0 + dt * t, ci, zi, 0);
                else a.OU[p] = t;
            }
            // V: CDV (cvd.f90:44-98) then the diffusion of v.
            {
                const double uxmyp = VU(s0, -1, 1), wypzm = zc.wyp;
                double y = c25y * (vyp + v0) * (vyp + v0) - c25y * (v0 + vym) * (v0 + vym);
                double x = c25x * (vxp + v0) * (uyp + u0) - c25x * (v0 + vxm) * (uxmyp + uxm);
                up.cv = c25z * (vzp + v0) * (wyp + w0);
                double z = up.cv - zc.cv;
                double t = -((y + x) + z);
                double Uadv = ((u0 + uyp) + uxm) + uxmyp;
                double Wadv = ((w0 + wyp) + wzm) + wypzm;
                t = t - ((c125x * (vxp - vxm) * Uadv + c50y * (vyp - vym) * v0) + c125z * (vzp - vzm) * Wadv);
                t = t * 0.5;
                double exy0 = 0.25 * (((n000 + n100) + n010) + NV(c0, 1, 1));
                double exyxm = 0.25 * (((nm00 + n000) + NV(c0, -1, 1)) + n010);
                double fxp = exy0 * (vxp - v0);
                double fxm = exyxm * (v0 - vxm);
                double fyp = n010 * (vyp - v0);
                double fym = n000 * (v0 - vym);
                double eyz0 = 0.25 * (((n000 + n010) + n001) + NV(cp, 0, 1));
                up.dv = eyz0 * (vzp - v0);
                double fzp = up.dv;
                double fzm = zc.dv;
                t = t + (((fxp - fxm) * hx + (fyp - fym) * hy) + (fzp - fzm) * hz);
                if constexpr (STEP) put_column(a.OV, p, v0 + dt * t, ci, zi, 0);
                else a.OV[p] = t;
            }
            // W on the interior faces: CDW (cdw.f90:44-98) then the diffusion of w.
            up.cw = zc.cw;
            up.dw = zc.dw;
            if (k <= a.kw) {
                const double uxmzp = VU(sp, -1, 0), vymzp = VV(sp, 0, -1);
                up.cw = c25z * (wzp + w0) * (wzp + w0);
                double z = up.cw - zc.cw;
                double y = c25y * (wyp + w0) * (vzp + v0) - c25y * (w0 + wym) * (vym + vymzp);
                double x = c25x * (wxp + w0) * (uzp + u0) - c25x * (w0 + wxm) * (uxmzp + uxm);
                double t = -((z + y) + x);
                double Uadv = ((u0 + uzp) + uxm) + uxmzp;
                double Vadv = ((v0 + vzp) + vym) + vymzp;
                t = t - ((c125x * (wxp - wxm) * Uadv + c125y * (wyp - wym) * Vadv) + c50z * (wzp - wzm) * w0);
                t = t * 0.5;
                double exz0 = 0.25 * (((n000 + n100) + n001) + NV(cp, 1, 0));
                double exzxm = 0.25 * (((nm00 + n000) + NV(cp, -1, 0)) + n001);
                double fxp = exz0 * (wxp - w0);
                double fxm = exzxm * (w0 - wxm);
                double eyz0 = 0.25 * (((n000 + n010) + n001) + NV(cp, 0, 1));
                double eyzym = 0.25 * (((n0m0 + n000) + NV(cp, 0, -1)) + n001);
                double fyp = eyz0 * (wyp - w0);
                double fym = eyzym * (w0 - wym);
                up.dw = n001 * (wzp - w0);
                double fzp = up.dw;
                double fzm = zc.dw;
                t = t + (((fxp - fxm) * hx + (fyp - fym) * hy) + (fzp - fzm) * hz);
                if constexpr (STEP) put_column(a.OW, p, w0 + dt * t, ci, zi, 1);
                else a.OW[p] = t;
            }
            // S: flux-form advection then the diffusion of s.
            {
                const double s0v = SS(ss0, 0, 0), sxp = SS(ss0, 1, 0), sxm = SS(ss0, -1, 0), syp = SS(ss0, 0, 1);
                const double sym = SS(ss0, 0, -1), szp = SS(ssp, 0, 0);
                const double Cx = c50x, Cy = c50y;
                double x = Cx * u0 * (sxp + s0v) - Cx * uxm * (s0v + sxm);
                double y = Cy * v0 * (syp + s0v) - Cy * vym * (s0v + sym);
                up.cs = Cz * w0 * (szp + s0v);
                double z = up.cs - zc.cs;
                double t = -((x + y) + z);
                double fxp = (0.5 * (n000 + n100)) * (sxp - s0v);
                double fxm = (0.5 * (nm00 + n000)) * (s0v - sxm);
                double fyp = (0.5 * (n000 + n010)) * (syp - s0v);
                double fym = (0.5 * (n0m0 + n000)) * (s0v - sym);
                up.ds = (0.5 * (n000 + n001)) * (szp - s0v);
                double fzp = up.ds;
                double fzm = zc.ds;
                t = t + (((fxp - fxm) * hx + (fyp - fym) * hy) + (fzp - fzm) * hz);
                if constexpr (STEP) put_column(a.OS, p, s0v + dt * t, ci, zi, 2);
                else a.OS[p] = t;
            }
            up.u = u0;
            up.v = v0;
            up.w = w0;
            up.wxp = wxp;
            up.wyp = wyp;
            zc = up;
            }
        }
        if constexpr (!TWO) __syncthreads();
    }
    // With two barriers per step the last update may still read its slots; the next tile of a persistent
    // walk starts by writing them.
    if constexpr (TWO) __syncthreads();
#undef VU
#undef VV
#undef VW
#undef SS
#undef NV
}

// Workgroups per CU the staged kernel is built for: three where its LDS allows (38.5 KB at 32 x 8, 47 KB at
// 64 x 4), else two (66-73 KB at 128 x 2 or with two rows per thread). Four would fit the 32 x 8 tile's LDS,
// but its registers do not fit four: held to 128 VGPRs it spills in the k loop and took 1.62 ms against
// 0.97 ms at three.
template <int ROWS, int WX, bool TWO = false>
constexpr int sgs3_occupancy()
{
    return sizeof(Sgs3Lds<ROWS, WX, TWO>) <= 160 * 1024 / 3 ? 3 : 2;
}

template <bool STEP, int ROWS, int WX = TX, bool ISO = false, bool TWO = false>
__global__ __launch_bounds__(NTH, (sgs3_occupancy<ROWS, WX, TWO>())) void les_sgs3_kernel(const LesArgs a)
{
    using T = Sgs3Tile<ROWS, WX>;
    __shared__ Sgs3Lds<ROWS, WX, TWO> L;
    if (a.tile_order == 2) {
        // Persistent banded walk (xcd.hpp) over the tiles in x-fastest order, by one resident round of
        // workgroups (a multiple of eight): the band resident on one XCD re-reads its x and y neighbours'
        // ring rows from the same L2.
        const unsigned gx = (a.nx + WX - 1) / WX, gy = (a.ny + T::TYT - 1) / T::TYT;
        const unsigned gz = (a.nz + a.kchunk - 1) / a.kchunk;
        const TileWalk walk = banded_walk(gx * gy * gz, gridDim.x, blockIdx.x);
        for (unsigned t = walk.first; t < walk.end; t += walk.step)
            les_sgs3_body<STEP, ROWS, WX, ISO, TWO>(a, L, t % gx, (t / gx) % gy, t / (gx * gy));
        return;
    }
    // Tiles x fastest, then y, then the k chunk: with the XCD-contiguous order the x and y neighbours
    // whose ghost rows and columns a tile re-reads run under the same L2 at about the same time.
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned t = launch_tile(a.tile_order);
    les_sgs3_body<STEP, ROWS, WX, ISO, TWO>(a, L, t % gx, (t / gx) % gy, t / (gx * gy));
}

// ------------------------------------------------------------------------------------------------ buffer-addressed step
// The step kernel the step entry launches. Same staging, same values and the same arithmetic in the same order as
// les_sgs3_body (three U/V planes, two W, scalar and viscosity planes in LDS, the z-face terms carried in
// registers), built for four waves per SIMD:
// - every global access goes through a buffer resource: a uniform per-plane byte offset in a scalar register and
//   a k-invariant 32-bit byte offset per thread, so no 64-bit address is rebuilt per load or store; ghost images
//   are offsets from the cell's, formed inside the k loop where a boundary column needs them;
// - the k loop is unrolled by six, so every LDS slot (plane % 3, plane % 2) is a constant and each LDS access a
//   fixed offset from the thread's base address;
// - the block is one-dimensional: NT threads forming WX x NT/WX columns.
// Arrays must be below 2^31 bytes (checked by the entry).
typedef unsigned int ftc_u32x2 __attribute__((ext_vector_type(2)));
template <int N>
using IC = std::integral_constant<int, N>;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const double *p, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(p), 0, bytes, 0x00020000);
}

__device__ __forceinline__ double buffer_load(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Plain store, no cache hint: the row ends that neighbouring tiles both write merge in L2.
__device__ __forceinline__ void buffer_store(__amdgpu_buffer_rsrc_t r, int voff, int soff, double v)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(ftc_u32x2, v), r, voff, soff, 0);
}

// The scalar and viscosity planes (one-cell ring, CWT = WX + 2 cells wide) are laid out with the velocity
// planes' row pitch, CXT = VXT = WX + 4: the ring cell at linear index c then reads the velocity tile at
// c + VXT + 1, so consecutive lanes read consecutive doubles across a row end as well. With the narrower
// pitch the index jumped by two doubles at every row end, and each 16-lane group that crossed one took a
// two-way bank conflict on every velocity read. The two extra columns hold duplicates of the last ring
// column and are never read.
template <int NT, int WX>
struct StepTile {
    static constexpr int TYT = NT / WX;
    static constexpr int VXT = WX + 4, VYT = TYT + 4, CXT = WX + 4, CWT = WX + 2, CYT = TYT + 2;
    static constexpr int VC = VXT * VYT, CC = CXT * CYT;
    static constexpr int VR = (VC + NT - 1) / NT, CR = (CC + NT - 1) / NT;
};

template <int NT, int WX>
struct StepLds {
    using T = StepTile<NT, WX>;
    double uv[3][2][T::VYT][T::VXT];  // U, V; slot = plane % 3
    double w[2][T::VYT][T::VXT];      // W; slot = plane & 1
    double sca[2][T::CYT][T::CXT];    // slot = plane & 1
    double nu[2][T::CYT][T::CXT];     // slot = plane & 1
};

// The stencil operands of the update of plane k (slots s0, sp of U/V, c0, cp of W, S and nu), read from LDS.
#define LES_STEP_OPERANDS                                                                                    \
    const double u0 = VU(s0, 0, 0), uxp = VU(s0, 1, 0), uxm = VU(s0, -1, 0), uyp = VU(s0, 0, 1);             \
    const double uym = VU(s0, 0, -1), uzp = VU(sp, 0, 0), uzm = zc.u;                                        \
    const double v0 = VV(s0, 0, 0), vxp = VV(s0, 1, 0), vxm = VV(s0, -1, 0), vyp = VV(s0, 0, 1);             \
    const double vym = VV(s0, 0, -1), vzp = VV(sp, 0, 0), vzm = zc.v;                                        \
    const double w0 = VW(c0, 0, 0), wxp = VW(c0, 1, 0), wxm = VW(c0, -1, 0), wyp = VW(c0, 0, 1);             \
    const double wym = VW(c0, 0, -1), wzp = VW(cp, 0, 0), wzm = zc.w;                                        \
    const double n000 = NV(c0, 0, 0), n100 = NV(c0, 1, 0), nm00 = NV(c0, -1, 0);                             \
    const double n010 = NV(c0, 0, 1), n0m0 = NV(c0, 0, -1), n001 = NV(cp, 0, 0);                             \
    (void)u0, (void)uxp, (void)uxm, (void)uyp, (void)uym, (void)uzp, (void)uzm, (void)v0, (void)vxp, (void)vxm, \
        (void)vyp, (void)vym, (void)vzp, (void)vzm, (void)w0, (void)wxp, (void)wxm, (void)wyp, (void)wym,   \
        (void)wzp, (void)wzm, (void)n000, (void)n100, (void)nm00, (void)n010, (void)n0m0, (void)n001
// SPLIT: the compiler may keep no operand across the component updates, so each group reads its own from LDS
// again; that holds fewer registers at a time (more LDS reads). SPLIT 1: U | V | W | S | carry; SPLIT 2:
// U and V | W | S and carry. Block b is 0 (U), 1 (V), 2 (W), 3 (S) or 4 (the carried values).
#define LES_STEP_FENCE(b)                                                                       \
    if constexpr ((SPLIT == 1 && (b) > 0) || (SPLIT == 2 && ((b) == 2 || (b) == 3)))          \
    asm volatile("" ::: "memory")

// One column tile over planes k0..k1 (k0 % 6 == 2). FP: where the global loads of the next planes are issued in
// a k-step: 0 all before the viscosity (in flight through the whole step), 3 U and V before the viscosity and W
// and S after the update (their registers are then free during it). SPLIT: see LES_STEP_FENCE.
template <int NT, int WX, bool ISO, int FP, int SPLIT>
__device__ __forceinline__ void les_step_body(const LesArgs &a, StepLds<NT, WX> &L, int bx, int by, int k0, int k1)
{
    using T = StepTile<NT, WX>;
    constexpr int VR = T::VR, CR = T::CR, VX = T::VXT, CX = T::CXT, VC = T::VC, CC = T::CC;
    const int tid = threadIdx.x;
    const int tx = tid % WX, ty = tid / WX;
    const int i0 = 2 + bx * WX, j0 = 2 + by * T::TYT;
    const int i = i0 + tx, j = j0 + ty;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int sj = (int)a.sj;
    const int skb = (int)(a.sk * 8);
    const int bytes = (int)(a.sk * 8 * (nz + 2));
    const __amdgpu_buffer_rsrc_t rU = field_rsrc(a.U, bytes), rV = field_rsrc(a.V, bytes);
    const __amdgpu_buffer_rsrc_t rW = field_rsrc(a.W, bytes), rS = field_rsrc(a.S, bytes);

    // Byte offsets within a plane of the staged cells this thread loads (periodic images of the ring), and the
    // velocity-tile index of each viscosity ring cell it computes.
    int voff[VR], coff[CR], nvi[CR];
#pragma unroll
    for (int r = 0; r < VR; ++r) {
        const int c = tid + r * NT;
        const int cc = c < VC ? c : 0;
        const int ii = periodic_image(i0 - 2 + cc % VX, nx), jj = periodic_image(j0 - 2 + cc / VX, ny);
        voff[r] = ((ii - 1) + (jj - 1) * sj) * 8;
    }
#pragma unroll
    for (int r = 0; r < CR; ++r) {
        const int c = tid + r * NT;
        const int cc = c < CC ? c : 0;
        const int li = min(cc % CX, T::CWT - 1), lj = cc / CX;  // a padding column repeats the last ring column
        const int ii = periodic_image(i0 - 1 + li, nx), jj = periodic_image(j0 - 1 + lj, ny);
        coff[r] = ((ii - 1) + (jj - 1) * sj) * 8;
        nvi[r] = (lj + 1) * VX + li + 1;
    }
    const double c25x = a.c25[0], c25y = ISO ? c25x : a.c25[1], c25z = ISO ? c25x : a.c25[2];
    const double c50x = a.c50[0], c50y = ISO ? c50x : a.c50[1], c50z = ISO ? c50x : a.c50[2];
    const double c125x = a.c125[0], c125y = ISO ? c125x : a.c125[1], c125z = ISO ? c125x : a.c125[2];
    const double ix = a.sgs[0], iy = ISO ? ix : a.sgs[1], iz = ISO ? ix : a.sgs[2];  // 1/d
    const double hx = a.sgs[6], hy = ISO ? hx : a.sgs[7], hz = ISO ? hx : a.sgs[8];  // 1/d^2
    const double cs2 = a.sgs[9];                                                    // (Cs Delta)^2
    double rv[VR][2], rw[VR], rs[CR];
#define VLIVE(r) ((r) < VR - 1 || tid + (r) * NT < VC)
#define CLIVE(r) ((r) < CR - 1 || tid + (r) * NT < CC)
    auto fetch_uv = [&](int kp) {
        if (kp < 1 || kp > nz + 2) return;
        const int so = (kp - 1) * skb;
#pragma unroll
        for (int r = 0; r < VR; ++r)
            if (VLIVE(r)) {
                rv[r][0] = buffer_load(rU, voff[r], so);
                rv[r][1] = buffer_load(rV, voff[r], so);
            }
    };
    auto put_uv = [&](int kp, int slot) {
        if (kp < 1 || kp > nz + 2) return;
#pragma unroll
        for (int r = 0; r < VR; ++r)
            if (VLIVE(r)) {
                (&L.uv[slot][0][0][0])[tid + r * NT] = rv[r][0];
                (&L.uv[slot][1][0][0])[tid + r * NT] = rv[r][1];
            }
    };
    // W planes up to nz+1 (see les_sgs3_body).
    auto fetch_w = [&](int kp) {
        if (kp < 1 || kp > nz + 1) return;
        const int so = (kp - 1) * skb;
#pragma unroll
        for (int r = 0; r < VR; ++r)
            if (VLIVE(r)) rw[r] = buffer_load(rW, voff[r], so);
    };
    auto put_w = [&](int kp, int slot) {
        if (kp < 1 || kp > nz + 1) return;
#pragma unroll
        for (int r = 0; r < VR; ++r)
            if (VLIVE(r)) (&L.w[slot][0][0])[tid + r * NT] = rw[r];
    };
    auto fetch_sca = [&](int kp) {
        if (kp < 1 || kp > nz + 2) return;
        const int so = (kp - 1) * skb;
#pragma unroll
        for (int r = 0; r < CR; ++r)
            if (CLIVE(r)) rs[r] = buffer_load(rS, coff[r], so);
    };
    auto put_sca = [&](int kp, int slot) {
        if (kp < 1 || kp > nz + 2) return;
#pragma unroll
        for (int r = 0; r < CR; ++r)
            if (CLIVE(r)) (&L.sca[slot][0][0])[tid + r * NT] = rs[r];
    };
    auto fetch_all = [&](int k) {
        fetch_uv(k + 3);
        fetch_w(k + 2);
        fetch_sca(k + 2);
    };
    // nu of one plane over the tile and its ring, from the U/V slots of that plane (s0) and the planes below
    // (sm) and above (sp), the W slots of the plane (w0s) and the one below (wms), into nu slot `oslot`.
    auto nu_slots = [&](int s0, int sm, int sp, int w0s, int wms, int oslot) __attribute__((always_inline)) {
        double *nuo = &L.nu[oslot][0][0];
#pragma unroll
        for (int r = 0; r < CR; ++r) {
            if (!CLIVE(r)) continue;
            const int vi = nvi[r];
#define QU(s, di, dj) (&L.uv[s][0][0][0])[vi + (dj) * VX + (di)]
#define QV(s, di, dj) (&L.uv[s][1][0][0])[vi + (dj) * VX + (di)]
#define QW(s, di, dj) (&L.w[s][0][0])[vi + (dj) * VX + (di)]
            double s11 = (QU(s0, 0, 0) - QU(s0, -1, 0)) * ix;
            double s22 = (QV(s0, 0, 0) - QV(s0, 0, -1)) * iy;
            double s33 = (QW(w0s, 0, 0) - QW(wms, 0, 0)) * iz;
            double dudy = ((QU(s0, 0, 1) + QU(s0, -1, 1)) - (QU(s0, 0, -1) + QU(s0, -1, -1))) * c25y;
            double dvdx = ((QV(s0, 1, 0) + QV(s0, 1, -1)) - (QV(s0, -1, 0) + QV(s0, -1, -1))) * c25x;
            double dudz = ((QU(sp, 0, 0) + QU(sp, -1, 0)) - (QU(sm, 0, 0) + QU(sm, -1, 0))) * c25z;
            double dwdx = ((QW(w0s, 1, 0) + QW(wms, 1, 0)) - (QW(w0s, -1, 0) + QW(wms, -1, 0))) * c25x;
            double dvdz = ((QV(sp, 0, 0) + QV(sp, 0, -1)) - (QV(sm, 0, 0) + QV(sm, 0, -1))) * c25z;
            double dwdy = ((QW(w0s, 0, 1) + QW(wms, 0, 1)) - (QW(w0s, 0, -1) + QW(wms, 0, -1))) * c25y;
#undef QU
#undef QV
#undef QW
            double s12 = 0.5 * (dudy + dvdx);
            double s13 = 0.5 * (dudz + dwdx);
            double s23 = 0.5 * (dvdz + dwdy);
            double ss = 2.0 * ((s11 * s11 + s22 * s22) + s33 * s33) + 4.0 * ((s12 * s12 + s13 * s13) + s23 * s23);
            nuo[tid + r * NT] = cs2 * sqrt(ss);
        }
    };
    // nu at plane qk, clamped to the interior planes (zero gradient in z).
    auto nu_plane = [&](int qk) {
        const int kk = qk < 2 ? 2 : (qk > nz + 1 ? nz + 1 : qk);
        nu_slots(kk % 3, (kk - 1) % 3, (kk + 1) % 3, kk & 1, (kk - 1) & 1, qk & 1);
    };

    const double dt = a.dt;
    const double Cz = c50z;
    const int vt = (ty + 2) * VX + tx + 2, ct = (ty + 1) * CX + tx + 1;
#define VU(s, di, dj) (&L.uv[s][0][0][0])[vt + (dj) * VX + (di)]
#define VV(s, di, dj) (&L.uv[s][1][0][0])[vt + (dj) * VX + (di)]
#define VW(s, di, dj) (&L.w[s][0][0])[vt + (dj) * VX + (di)]
#define SS(s, di, dj) (&L.sca[s][0][0])[ct + (dj) * CX + (di)]
#define NV(s, di, dj) (&L.nu[s][0][0])[ct + (dj) * CX + (di)]

    // Prologue as in les_sgs3_body.
    for (int kp = k0 - 2; kp <= k0; ++kp) {
        fetch_uv(kp);
        put_uv(kp, kp % 3);
    }
    for (int kp = k0 - 2; kp <= k0 - 1; ++kp) {
        fetch_w(kp);
        put_w(kp, kp & 1);
    }
    for (int kp = k0 - 1; kp <= k0; ++kp) {
        fetch_sca(kp);
        put_sca(kp, kp & 1);
    }
    fetch_uv(k0 + 1);
    fetch_w(k0);
    const bool clamped = k0 - 1 < 2;
    __syncthreads();
    if (!clamped) nu_plane(k0 - 1);
    __syncthreads();
    put_uv(k0 + 1, (k0 + 1) % 3);
    put_w(k0, k0 & 1);
    __syncthreads();
    if (clamped) nu_plane(k0 - 1);
    nu_plane(k0);
    __syncthreads();
    const bool inside = i <= nx + 1 && j <= ny + 1;
    ZCarry zc = {};
    if (inside) {
        const int kv = k0 - 1;
        const int s0 = kv % 3, sp = (kv + 1) % 3, c0 = kv & 1, cp = (kv + 1) & 1;
        const double u0 = VU(s0, 0, 0), uzp = VU(sp, 0, 0);
        const double v0 = VV(s0, 0, 0), vzp = VV(sp, 0, 0);
        const double w0 = VW(c0, 0, 0), wxp = VW(c0, 1, 0), wyp = VW(c0, 0, 1), wzp = VW(cp, 0, 0);
        const double n000 = NV(c0, 0, 0), n100 = NV(c0, 1, 0), n010 = NV(c0, 0, 1), n001 = NV(cp, 0, 0);
        zc.cu = c25z * (uzp + u0) * (wxp + w0);
        zc.du = (0.25 * (((n000 + n100) + n001) + NV(cp, 1, 0))) * (uzp - u0);
        zc.cv = c25z * (vzp + v0) * (wyp + w0);
        zc.dv = (0.25 * (((n000 + n010) + n001) + NV(cp, 0, 1))) * (vzp - v0);
        zc.cw = c25z * (wzp + w0) * (wzp + w0);
        zc.dw = n001 * (wzp - w0);
        const double s0v = SS(c0, 0, 0), szp = SS(cp, 0, 0);
        zc.cs = Cz * w0 * (szp + s0v);
        zc.ds = (0.5 * (n000 + n001)) * (szp - s0v);
        zc.u = u0;
        zc.v = v0;
        zc.w = w0;
        zc.wxp = wxp;
        zc.wyp = wyp;
    }
    fetch_uv(k0 + 2);
    fetch_w(k0 + 1);
    fetch_sca(k0 + 1);
    __syncthreads();

    const __amdgpu_buffer_rsrc_t rOU = field_rsrc(a.OU, bytes), rOV = field_rsrc(a.OV, bytes);
    const __amdgpu_buffer_rsrc_t rOW = field_rsrc(a.OW, bytes), rOS = field_rsrc(a.OS, bytes);
    const int vo = ((i - 1) + (j - 1) * sj) * 8;
    const bool edge = i == 2 || i == nx + 1 || j == 2 || j == ny + 1;
    // One k-step for a plane k with k % 6 == PH: every LDS slot index is a constant, so each LDS access is a
    // fixed offset from the thread's base address.
    auto step = [&](int k, auto phase) __attribute__((always_inline)) {
        constexpr int PH = decltype(phase)::value;
        put_uv(k + 2, (PH + 2) % 3);
        put_w(k + 1, (PH + 1) & 1);
        put_sca(k + 1, (PH + 1) & 1);
        if constexpr (FP == 0) fetch_all(k);
        if constexpr (FP == 3) fetch_uv(k + 3);
        __syncthreads();
        if (k + 1 <= nz + 1) {
            nu_slots((PH + 1) % 3, PH % 3, (PH + 2) % 3, (PH + 1) & 1, PH & 1, (PH + 1) & 1);
        } else {
            // Plane nz+2 takes the viscosity of plane nz+1 (zero gradient), computed by the previous step.
#pragma unroll
            for (int r = 0; r < CR; ++r)
                if (CLIVE(r)) (&L.nu[(PH + 1) & 1][0][0])[tid + r * NT] = (&L.nu[PH & 1][0][0])[tid + r * NT];
        }
        __syncthreads();
        // The update runs at a raised wave priority: the SIMD then issues the update's arithmetic ahead of the
        // other resident workgroup's viscosity and staging, so the two workgroups drift apart and one's loads
        // overlap the other's arithmetic (0.80-0.82 -> 0.76-0.77 ms at 512 x 512 x 128, the same bits).
        // Raised during the viscosity as well, it gains nothing.
        __builtin_amdgcn_s_setprio(1);
        const int so = (k - 1) * skb;
        const bool zlo = k == 2, zhi = k == nz + 1;
        // Writes val at the cell and at every ghost whose image it is, as put_with_images: the x and y images
        // and their combinations, each with the z images unless zmode is 1 (negated for zmode 0).
        auto put = [&](__amdgpu_buffer_rsrc_t F, double val, int zmode) {
            const double zval = zmode == 0 ? -val : val;
            const bool zi = zmode != 1;
            buffer_store(F, vo, so, val);
            if (zi && zlo) buffer_store(F, vo, so - skb, zval);
            if (zi && zhi) buffer_store(F, vo, so + skb, zval);
            if (edge) {
                // The image offsets are formed from the plane's offset, so they stay inside the k loop instead
                // of holding registers through every step for the few boundary columns that need them.
                const int base = vo + so;
                auto with_z = [&](int q) {
                    buffer_store(F, q, 0, val);
                    if (zi && zlo) buffer_store(F, q - skb, 0, zval);
                    if (zi && zhi) buffer_store(F, q + skb, 0, zval);
                };
                const bool hxa = i == 2, hxb = i == nx + 1, hya = j == 2, hyb = j == ny + 1;
                const int oxa = nx * 8, oya = ny * sj * 8;
                if (hxa) with_z(base + oxa);
                if (hxb) with_z(base - oxa);
                if (hya) with_z(base + oya);
                if (hyb) with_z(base - oya);
                if (hxa && hya) with_z(base + oxa + oya);
                if (hxa && hyb) with_z(base + oxa - oya);
                if (hxb && hya) with_z(base - oxa + oya);
                if (hxb && hyb) with_z(base - oxa - oya);
            }
        };
        if (inside) {
            constexpr int s0 = PH % 3, sp = (PH + 1) % 3;
            constexpr int c0 = PH & 1, cp = (PH + 1) & 1;
            ZCarry up;
            // U: CDU (cdu.f90:48-107) then the diffusion of u.
            {
                LES_STEP_FENCE(0);
                LES_STEP_OPERANDS;
                const double vxpym = VV(s0, 1, -1), wxpzm = zc.wxp;
                double x = c25x * (uxp + u0) * (uxp + u0) - c25x * (u0 + uxm) * (u0 + uxm);
                double y = c25y * (uyp + u0) * (vxp + v0) - c25y * (u0 + uym) * (vxpym + vym);
                up.cu = c25z * (uzp + u0) * (wxp + w0);
                double z = up.cu - zc.cu;
                double t = -((x + y) + z);
                double Vadv = ((v0 + vxp) + vym) + vxpym;
                double Wadv = ((w0 + wxp) + wzm) + wxpzm;
                t = t - ((c50x * (uxp - uxm) * u0 + c125y * (uyp - uym) * Vadv) + c125z * (uzp - uzm) * Wadv);
                t = t * 0.5;
                double fxp = n100 * (uxp - u0);
                double fxm = n000 * (u0 - uxm);
                double exy0 = 0.25 * (((n000 + n100) + n010) + NV(c0, 1, 1));
                double exym = 0.25 * (((n0m0 + NV(c0, 1, -1)) + n000) + n100);
                double fyp = exy0 * (uyp - u0);
                double fym = exym * (u0 - uym);
                double exz0 = 0.25 * (((n000 + n100) + n001) + NV(cp, 1, 0));
                up.du = exz0 * (uzp - u0);
                double fzp = up.du;
                double fzm = zc.du;
                t = t + (((fxp - fxm) * hx + (fyp - fym) * hy) + (fzp - fzm) * hz);
                put(rOU, u0 + dt * t, 0);
            }
            // V: CDV (cvd.f90:44-98) then the diffusion of v.
            {
                LES_STEP_FENCE(1);
                LES_STEP_OPERANDS;
                const double uxmyp = VU(s0, -1, 1), wypzm = zc.wyp;
                double y = c25y * (vyp + v0) * (vyp + v0) - c25y * (v0 + vym) * (v0 + vym);
                double x = c25x * (vxp + v0) * (uyp + u0) - c25x * (v0 + vxm) * (uxmyp + uxm);
                up.cv = c25z * (vzp + v0) * (wyp + w0);
                double z = up.cv - zc.cv;
                double t = -((y + x) + z);
                double Uadv = ((u0 + uyp) + uxm) + uxmyp;
                double Wadv = ((w0 + wyp) + wzm) + wypzm;
                t = t - ((c125x * (vxp - vxm) * Uadv + c50y * (vyp - vym) * v0) + c125z * (vzp - vzm) * Wadv);
                t = t * 0.5;
                double exy0 = 0.25 * (((n000 + n100) + n010) + NV(c0, 1, 1));
                double exyxm = 0.25 * (((nm00 + n000) + NV(c0, -1, 1)) + n010);
                double fxp = exy0 * (vxp - v0);
                double fxm = exyxm * (v0 - vxm);
                double fyp = n010 * (vyp - v0);
                double fym = n000 * (v0 - vym);
                double eyz0 = 0.25 * (((n000 + n010) + n001) + NV(cp, 0, 1));
                up.dv = eyz0 * (vzp - v0);
                double fzp = up.dv;
                double fzm = zc.dv;
                t = t + (((fxp - fxm) * hx + (fyp - fym) * hy) + (fzp - fzm) * hz);
                put(rOV, v0 + dt * t, 0);
            }
            // W on the interior faces: CDW (cdw.f90:44-98) then the diffusion of w.
            up.cw = zc.cw;
            up.dw = zc.dw;
            if (k <= a.kw) {
                LES_STEP_FENCE(2);
                LES_STEP_OPERANDS;
                const double uxmzp = VU(sp, -1, 0), vymzp = VV(sp, 0, -1);
                up.cw = c25z * (wzp + w0) * (wzp + w0);
                double z = up.cw - zc.cw;
                double y = c25y * (wyp + w0) * (vzp + v0) - c25y * (w0 + wym) * (vym + vymzp);
                double x = c25x * (wxp + w0) * (uzp + u0) - c25x * (w0 + wxm) * (uxmzp + uxm);
                double t = -((z + y) + x);
                double Uadv = ((u0 + uzp) + uxm) + uxmzp;
                double Vadv = ((v0 + vzp) + vym) + vymzp;
                t = t - ((c125x * (wxp - wxm) * Uadv + c125y * (wyp - wym) * Vadv) + c50z * (wzp - wzm) * w0);
                t = t * 0.5;
                double exz0 = 0.25 * (((n000 + n100) + n001) + NV(cp, 1, 0));
                double exzxm = 0.25 * (((nm00 + n000) + NV(cp, -1, 0)) + n001);
                double fxp = exz0 * (wxp - w0);
                double fxm = exzxm * (w0 - wxm);
                double eyz0 = 0.25 * (((n000 + n010) + n001) + NV(cp, 0, 1));
                double eyzym = 0.25 * (((n0m0 + n000) + NV(cp, 0, -1)) + n001);
                double fyp = eyz0 * (wyp - w0);
                double fym = eyzym * (w0 - wym);
                up.dw = n001 * (wzp - w0);
                double fzp = up.dw;
                double fzm = zc.dw;
                t = t + (((fxp - fxm) * hx + (fyp - fym) * hy) + (fzp - fzm) * hz);
                put(rOW, w0 + dt * t, 1);
            }
            // S: flux-form advection then the diffusion of s.
            {
                LES_STEP_FENCE(3);
                LES_STEP_OPERANDS;
                const double s0v = SS(c0, 0, 0), sxp = SS(c0, 1, 0), sxm = SS(c0, -1, 0), syp = SS(c0, 0, 1);
                const double sym = SS(c0, 0, -1), szp = SS(cp, 0, 0);
                const double Cx = c50x, Cy = c50y;
                double x = Cx * u0 * (sxp + s0v) - Cx * uxm * (s0v + sxm);
                double y = Cy * v0 * (syp + s0v) - Cy * vym * (s0v + sym);
                up.cs = Cz * w0 * (szp + s0v);
                double z = up.cs - zc.cs;
                double t = -((x + y) + z);
                double fxp = (0.5 * (n000 + n100)) * (sxp - s0v);
                double fxm = (0.5 * (nm00 + n000)) * (s0v - sxm);
                double fyp = (0.5 * (n000 + n010)) * (syp - s0v);
                double fym = (0.5 * (n0m0 + n000)) * (s0v - sym);
                up.ds = (0.5 * (n000 + n001)) * (szp - s0v);
                double fzp = up.ds;
                double fzm = zc.ds;
                t = t + (((fxp - fxm) * hx + (fyp - fym) * hy) + (fzp - fzm) * hz);
                put(rOS, s0v + dt * t, 2);
            }
            {
                LES_STEP_FENCE(4);
                LES_STEP_OPERANDS;
                up.u = u0;
                up.v = v0;
                up.w = w0;
                up.wxp = wxp;
                up.wyp = wyp;
            }
            zc = up;
        }
        __builtin_amdgcn_s_setprio(0);
        if constexpr (FP == 3) {
            fetch_w(k + 2);
            fetch_sca(k + 2);
        }
        __syncthreads();
    };
    // k0 % 6 == 2: every chunk but the last is a multiple of six planes (the entry chooses them so).
    for (int k = k0; k <= k1; k += 6) {
        step(k, IC<2>{});
        if (k + 1 > k1) break;
        step(k + 1, IC<3>{});
        if (k + 2 > k1) break;
        step(k + 2, IC<4>{});
        if (k + 3 > k1) break;
        step(k + 3, IC<5>{});
        if (k + 4 > k1) break;
        step(k + 4, IC<0>{});
        if (k + 5 > k1) break;
        step(k + 5, IC<1>{});
    }
#undef VU
#undef VV
#undef VW
#undef SS
#undef NV
#undef VLIVE
#undef CLIVE
#undef LES_STEP_OPERANDS
#undef LES_STEP_FENCE
}

// One tile of NT / WX rows by WX columns per workgroup, k chunks of a.kchunk planes, tiles x fastest, then y,
// then the chunk, in the order a.tile_order names (launch order or XCD-contiguous).
template <int NT, int WX, bool ISO, int WPE, int FP, int SPLIT>
__global__ __attribute__((amdgpu_flat_work_group_size(NT, NT), amdgpu_waves_per_eu(WPE, WPE))) void
les_step_kernel(const LesArgs a)
{
    __shared__ StepLds<NT, WX> L;
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned t = launch_tile(a.tile_order);
    const int bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    const int k0 = 2 + bz * a.kchunk;
    if (k0 > a.nz + 1) return;  // uniform over the workgroup
    // The last chunk takes the planes that remain.
    const int k1 = bz == (int)gridDim.z - 1 ? a.nz + 1 : min(k0 + a.kchunk - 1, a.nz + 1);
    les_step_body<NT, WX, ISO, FP, SPLIT>(a, L, bx, by, k0, k1);
}

}  // namespace ftc
