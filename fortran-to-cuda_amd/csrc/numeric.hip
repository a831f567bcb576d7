// Device checks of the intrinsic semantics in numeric.hpp: the exact power-of-two division against the
// plain division, and the boundary cases of the integer and real intrinsics, evaluated by gfx950 code.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>

#include "../../include/ftc_numeric.h"
#include "numeric.hpp"

namespace ftc {

template <class T>
__global__ void divide_probe_kernel(const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ ref,
                                    T *__restrict__ cand, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ref[i] = a[i] / b[i];
    cand[i] = numeric::divide(a[i], b[i]);
    ref[n + i] = (i - 41) / b[i];  // an INTEGER numerator promoted to the REAL kind of the divisor
    cand[n + i] = numeric::divide(i - 41, b[i]);
}

// Bit k of *out is set when check k fails.
__global__ void intrinsic_probe_kernel(int *out, volatile int *imin_src, volatile double *dzero_src)
{
    using namespace numeric;
    const int imin = *imin_src;          // INT_MIN, read at run time so nothing folds
    const double nz = -*dzero_src;       // -0.0
    int bad = 0, k = 0;
    auto check = [&](bool ok) {
        if (!ok) bad |= 1 << k;
        ++k;
    };
    check(mod(imin, -1) == 0);
    check(modulo(imin, -1) == 0);
    check(sign(imin, -1) == imin);
    check(dim(imin, 1) == 0);
    check(nint(std::nextafter(0.5, 0.0)) == 0);
    check(nint(std::nextafter(-0.5, 0.0)) == 0);
    check(nint(1.5) == 2 && nint(-1.5) == -2 && nint(2.5f) == 3);
    check(std::signbit(sign(1.0, nz)));
    check(std::signbit(modulo(0.0, -1.0)));
    check(modulo(-7, 3) == 2 && modulo(7, -3) == -2 && mod(-7, 3) == -1);
    check(modulo(-7.5, 2.0) == 0.5 && mod(-7.5, 2.0) == -1.5);
    const double qnan = std::nan("");
    check(std::isnan(minimum(1.0, qnan)) && minimum(qnan, 1.0) == 1.0);   // NaN decided by position
    check(std::isnan(maximum(1.0, qnan)) && maximum(qnan, 1.0) == 1.0);
    check(std::signbit(minimum(0.0, nz)) == true && std::signbit(minimum(nz, 0.0)) == false);  // ties keep the later
    check(minimum(3, 1, 2) == 1 && maximum(3.0, 7.0, 5.0) == 7.0);
    check(std::isnan(minimum_constant_first(1.0, qnan)) == false && minimum_constant_first(1.0, 0.5) == 0.5);
    check(floor_int(-1.5) == -2 && ceiling_int(-1.5) == -1 && floor_int(2.0f) == 2);
    check(merge(4, 5, true) == 4 && merge(4.0, 5.0, false) == 5.0);
    check(absolute(-2.5) == 2.5 && absolute(3) == 3);
    check(divide(1, 4.0) == 0.25 && divide(3.0f, 0.5f) == 6.0f && divide(3, 0.5f) == 6.0f);
    // The same orderings on operands known only at run time, where the backend chooses the instructions.
    const double rnan = nz / nz;
    const float nzf = static_cast<float>(nz), rnanf = nzf / nzf;
    check(std::signbit(maximum(nz, 0.0)) == false && std::signbit(maximum(0.0, nz)) == true);
    check(std::signbit(minimum(nzf, 0.0f)) == false && std::signbit(maximum(0.0f, nzf)) == true);
    check(std::signbit(minimum_constant_first(0.0, nz)) == false && std::signbit(maximum_constant_first(0.0, nz)) == false);
    check(std::signbit(absolute(nz)) && std::signbit(absolute(nzf)));
    check(std::isnan(minimum(1.0, rnan)) && minimum(rnan, 1.0) == 1.0 && std::isnan(maximum(2.0, 1.0, rnan)));
    check(std::isnan(minimum(1.0f, rnanf)) && maximum(rnanf, 1.0f) == 1.0f);
    check(minimum_constant_first(1.0, rnan) == 1.0 && maximum_constant_first(1.0f, rnanf) == 1.0f);
    *out = bad;
}

template <class T>
static int run_divide(const void *a, const void *b, void *ref, void *cand, int n)
{
    T *da = nullptr, *db = nullptr, *dr = nullptr, *dc = nullptr;
    const std::size_t bytes = (std::size_t)n * sizeof(T);
    int status = FORT_SCOPE_RESOURCE;
    if (hipMalloc(&da, bytes) == hipSuccess && hipMalloc(&db, bytes) == hipSuccess &&
        hipMalloc(&dr, 2 * bytes) == hipSuccess && hipMalloc(&dc, 2 * bytes) == hipSuccess &&
        hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) == hipSuccess) {
        hipLaunchKernelGGL(divide_probe_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, 0, da, db, dr, dc, n);
        status = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
                         hipMemcpy(ref, dr, 2 * bytes, hipMemcpyDeviceToHost) == hipSuccess &&
                         hipMemcpy(cand, dc, 2 * bytes, hipMemcpyDeviceToHost) == hipSuccess
                     ? FORT_SCOPE_OK
                     : FORT_SCOPE_EXECUTION;
    }
    (void)hipGetLastError();
    for (T *p : {da, db, dr, dc})
        if (p) (void)hipFree(p);
    return status;
}

}  // namespace ftc

extern "C" {

int ftc_divide_probe(int width, const void *a, const void *b, void *reference, void *candidate, int n)
{
    if (!a || !b || !reference || !candidate || n <= 0 || (width != 4 && width != 8)) return FORT_SCOPE_ARGUMENT;
    return width == 4 ? ftc::run_divide<float>(a, b, reference, candidate, n)
                      : ftc::run_divide<double>(a, b, reference, candidate, n);
}

int ftc_intrinsic_probe(int *failures)
{
    if (!failures) return FORT_SCOPE_ARGUMENT;
    int *d = nullptr, *imin = nullptr;
    double *zero = nullptr;
    int status = FORT_SCOPE_RESOURCE;
    const int host_min = INT_MIN;
    const double host_zero = 0.0;
    if (hipMalloc(&d, sizeof(int)) == hipSuccess && hipMalloc(&imin, sizeof(int)) == hipSuccess &&
        hipMalloc(&zero, sizeof(double)) == hipSuccess &&
        hipMemcpy(imin, &host_min, sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(zero, &host_zero, sizeof(double), hipMemcpyHostToDevice) == hipSuccess) {
        hipLaunchKernelGGL(ftc::intrinsic_probe_kernel, dim3(1), dim3(1), 0, 0, d, imin, zero);
        status = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
                         hipMemcpy(failures, d, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess
                     ? FORT_SCOPE_OK
                     : FORT_SCOPE_EXECUTION;
    }
    (void)hipGetLastError();
    if (d) (void)hipFree(d);
    if (imin) (void)hipFree(imin);
    if (zero) (void)hipFree(zero);
    return status;
}

}  // extern "C"
