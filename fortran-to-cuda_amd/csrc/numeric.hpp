// Scalar semantics of the Fortran intrinsics the generated kernels evaluate, identical on the host and on
// gfx950 (compiler/runtime/numeric.hpp:11-112 states the contract; compiler/emission/numerical_contract.py
// fixes the build: no contraction, no flush-to-zero).
//
//   divide            a / b; a normal power-of-two divisor with a normal reciprocal becomes a multiplication
//                     by that exact reciprocal (same real result, same single rounding, subnormals and
//                     signed zero included); every other divisor, and non-finite numerators, keep a / b
//   minimum/maximum   ordered folds value = value < x ? value : x (NaN and signed zero decided by position)
//   *_constant_first  the GNU Fortran form with a finite leading constant: x < c ? x : c
//   nint/floor/ceiling  to int, nint rounding half away from zero
//   mod               integer mod(a, -1) = 0 (no trap at INT_MIN); real mod is fmod
//   sign              |a| with the sign of b (copysign for reals, so sign(1, -0.0) = -1)
//   modulo            the remainder with the sign of p; a zero remainder takes the sign of p
//   dim, merge        a > b ? a - b : 0, and mask ? t : f (each argument evaluated once)
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FTC_HD __host__ __device__
#else
#define FTC_HD
#endif

namespace ftc::numeric {

FTC_HD inline uint64_t bits_of(double x)
{
    uint64_t u;
    std::memcpy(&u, &x, sizeof u);
    return u;
}
FTC_HD inline double double_of(uint64_t u)
{
    double x;
    std::memcpy(&x, &u, sizeof x);
    return x;
}
FTC_HD inline uint32_t bits_of(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, sizeof u);
    return u;
}
FTC_HD inline float float_of(uint32_t u)
{
    float x;
    std::memcpy(&x, &u, sizeof x);
    return x;
}

// b = ±2^(e - 1023) with a biased exponent 1 <= e <= 2045 has the exact reciprocal ±2^(1023 - e), whose
// biased exponent 2046 - e is again normal; a finite numerator then gives a * (1/b) == a / b bit for bit.
FTC_HD inline double power_of_two_divide(double a, double b)
{
    const uint64_t bb = bits_of(b);
    const unsigned e = (unsigned)((bb >> 52) & 0x7ffu);
    if ((bb & 0x000fffffffffffffull) == 0 && e > 0 && e < 2046 &&
        (bits_of(a) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull)
        return a * double_of((bb & 0x8000000000000000ull) | ((uint64_t)(2046 - e) << 52));
    return a / b;
}

FTC_HD inline float power_of_two_divide(float a, float b)
{
    const uint32_t bb = bits_of(b);
    const unsigned e = (bb >> 23) & 0xffu;
    if ((bb & 0x007fffffu) == 0 && e > 0 && e < 254 && (bits_of(a) & 0x7f800000u) != 0x7f800000u)
        return a * float_of((bb & 0x80000000u) | ((uint32_t)(254 - e) << 23));
    return a / b;
}

// The usual arithmetic conversion first (REAL/INTEGER and REAL32/REAL mixes keep their result type).
template <typename L, typename R>
FTC_HD inline auto divide(L a, R b) -> decltype(a / b)
{
    using T = decltype(a / b);
    return power_of_two_divide(static_cast<T>(a), static_cast<T>(b));
}

// The comparison result passes through an empty asm before it picks an operand. Without it the gfx950
// backend rewrites `x < c ? x : c` (c a constant) into v_min_f64 / v_max_f64, which order -0 below +0
// and so return the other zero than the select does; with it every fold below returns exactly the
// operand its comparison chose, on the host and on the device alike.
FTC_HD inline bool chosen(bool condition)
{
    int k = condition;
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(k));
#else
    asm("" : "+r"(k));
#endif
    return k != 0;
}

template <typename T, typename... Rest>
FTC_HD inline T minimum(T value, Rest... rest)
{
    ((value = chosen(value < rest) ? value : rest), ...);
    return value;
}

template <typename T, typename... Rest>
FTC_HD inline T maximum(T value, Rest... rest)
{
    ((value = chosen(value > rest) ? value : rest), ...);
    return value;
}

template <typename T>
FTC_HD inline T minimum_constant_first(T constant, T value)
{
    return chosen(value < constant) ? value : constant;
}
template <typename T>
FTC_HD inline T maximum_constant_first(T constant, T value)
{
    return chosen(value > constant) ? value : constant;
}

template <typename T>
FTC_HD inline T absolute(T value)
{
    return chosen(value < T{0}) ? -value : value;   // absolute(-0.0) is -0.0, as the select says
}

FTC_HD inline int nint(float v) { return static_cast<int>(::roundf(v)); }
FTC_HD inline int nint(double v) { return static_cast<int>(::round(v)); }
FTC_HD inline int floor_int(float v) { return static_cast<int>(::floorf(v)); }
FTC_HD inline int floor_int(double v) { return static_cast<int>(::floor(v)); }
FTC_HD inline int ceiling_int(float v) { return static_cast<int>(::ceilf(v)); }
FTC_HD inline int ceiling_int(double v) { return static_cast<int>(::ceil(v)); }

FTC_HD inline int mod(int a, int p) { return p == -1 ? 0 : a % p; }
FTC_HD inline float mod(float a, float p) { return ::fmodf(a, p); }
FTC_HD inline double mod(double a, double p) { return ::fmod(a, p); }

FTC_HD inline int sign(int a, int b)
{
    const long long m = a < 0 ? -static_cast<long long>(a) : static_cast<long long>(a);
    return static_cast<int>(b < 0 ? -m : m);
}
FTC_HD inline float sign(float a, float b) { return ::copysignf(a, b); }
FTC_HD inline double sign(double a, double b) { return ::copysign(a, b); }

template <typename T>
FTC_HD inline T modulo(T a, T p)
{
    const T r = mod(a, p);
    if (r == T{0}) return sign(T{0}, p);
    return (r < T{0}) != (p < T{0}) ? r + p : r;
}

template <typename T>
FTC_HD inline T dim(T a, T b)
{
    return a > b ? a - b : T{0};
}

template <typename T>
FTC_HD inline T merge(T t, T f, bool mask)
{
    return mask ? t : f;
}

}  // namespace ftc::numeric
