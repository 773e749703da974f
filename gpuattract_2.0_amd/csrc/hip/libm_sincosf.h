// Single-precision sine and cosine with the values of the host C library.
//
// The reference's euler2rotmat (model/matrixFunctions.h:51-60) calls
// std::cos/std::sin on floats; its release build turns each pair into one
// sincosf call of the C library (glibc 2.28 and later).  Those results are not
// correctly rounded (about 3 % of arguments differ by one ulp from the float
// nearest the true value), so reproducing the reference's rotation matrices
// bit for bit needs the library's own method, restated here for the device
// and the host alike.
//
// Method (Szabolcs Nagy's sincosf in Arm's optimized-routines, math/sincosf.c,
// adopted by glibc 2.28): everything in double precision, then one rounding
// to float.
//   * |x| < 0.75 (top 12 bits below those of pi/4): no reduction; below 2^-12
//     sin x = x and cos x = 1.
//   * |x| < 120: n = round(x * 2/pi) by a 2^24-scaled truncating shift, the
//     remainder r = x - n * pi/2 with one fused multiply-add.
//   * otherwise (finite): Payne-Hanek reduction of the 24-bit mantissa by a
//     64-bit window of the bits of 2/pi (kTwoOverPiWords), remainder scaled
//     by pi/2^63.
//   * the quadrant's sign is folded into r and, for quadrants 2 and 3, into
//     the cosine coefficients; sin and cos of r are odd/even polynomials
//     (degree 7 and 8) with the fused multiply-adds the library's FMA build
//     uses; quadrants 1 and 3 swap the two results.
// Checked against the host C library's sincosf, sinf and cosf: on the host
// for every finite float, on the device for a sweep of angles and every
// fixture angle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace att {

// Polynomial coefficients of the published method: cos r = C0 + C1 r^2 + ...
// + C4 r^8, sin r = r + S1 r^3 + S2 r^5 + S3 r^7.
constexpr double kLibmC0 = 0x1p0, kLibmC1 = -0x1.ffffffd0c621cp-2, kLibmC2 = 0x1.55553e1068f19p-5,
                 kLibmC3 = -0x1.6c087e89a359dp-10, kLibmC4 = 0x1.99343027bf8c3p-16;
constexpr double kLibmS1 = -0x1.555545995a603p-3, kLibmS2 = 0x1.1107605230bc4p-7,
                 kLibmS3 = -0x1.994eb3774cf24p-13;
constexpr double kLibmHpiInv24 = 0x1.45F306DC9C883p+23;  // 2/pi * 2^24
constexpr double kLibmHpi = 0x1.921FB54442D18p0;         // pi/2
constexpr double kLibmPi63 = 0x1.921FB54442D18p-62;      // pi * 2^-63

// Overlapping 32-bit windows of the binary expansion of 2/pi: word k holds
// bits [8k - 24, 8k + 8) after the binary point (leading words padded with
// zeros), i.e. floor(2/pi * 2^(8k+8)) mod 2^32.
__host__ __device__ inline uint32_t two_over_pi_word(unsigned k) {
    static constexpr uint32_t kTwoOverPiWords[24] = {
        0xa2,       0xa2f9,     0xa2f983,   0xa2f9836e, 0xf9836e4e, 0x836e4e44,
        0x6e4e4415, 0x4e441529, 0x441529fc, 0x1529fc27, 0x29fc2757, 0xfc2757d1,
        0x2757d1f5, 0x57d1f534, 0xd1f534dd, 0xf534ddc0, 0x34ddc0db, 0xddc0db62,
        0xc0db6295, 0xdb629599, 0x6295993c, 0x95993c43, 0x993c4390, 0x3c439041};
    return kTwoOverPiWords[k];
}

__host__ __device__ inline uint32_t libm_f2u(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    return v.u;
}

// sin and cos of the reduced argument x (x2 = x*x) for quadrant n; neg
// negates the cosine polynomial (the library negates its coefficients, which
// gives the same bits: every step rounds to nearest, symmetric in sign).
__host__ __device__ inline void libm_sincos_poly(double x, double x2, bool neg, int n, float* sinp,
                                                 float* cosp) {
    const double x4 = x2 * x2;
    const double x3 = x2 * x;
    const double c2 = __builtin_fma(x2, kLibmC4, kLibmC3);
    const double s1 = __builtin_fma(x2, kLibmS3, kLibmS2);
    const double c1 = __builtin_fma(x2, kLibmC1, kLibmC0);
    const double x5 = x3 * x2;
    const double x6 = x4 * x2;
    const double s = __builtin_fma(x3, kLibmS1, x);
    const double c = __builtin_fma(x4, kLibmC2, c1);
    const float vs = (float)__builtin_fma(x5, s1, s);
    float vc = (float)__builtin_fma(x6, c2, c);
    if (neg) vc = -vc;
    if (n & 1) { *sinp = vc; *cosp = vs; }
    else { *sinp = vs; *cosp = vc; }
}

// Sign of the quadrant: +1 for n = 0, 3 and -1 for n = 1, 2 (mod 4).
__host__ __device__ inline double libm_quadrant_sign(int n) {
    return ((n ^ (n >> 1)) & 1) ? -1.0 : 1.0;
}

// Payne-Hanek reduction of a float with |x| >= 120 (bits xi): returns the
// remainder in double and the quadrant in *np.
__host__ __device__ inline double libm_reduce_large(uint32_t xi, int* np) {
    const unsigned base = (xi >> 26) & 15u;
    const int shift = (xi >> 23) & 7;
    xi = (xi & 0xffffffu) | 0x800000u;
    xi <<= shift;
    uint64_t r0 = (uint64_t)(uint32_t)(xi * two_over_pi_word(base));
    const uint64_t r1 = (uint64_t)xi * two_over_pi_word(base + 4);
    const uint64_t r2 = (uint64_t)xi * two_over_pi_word(base + 8);
    r0 = (r2 >> 32) | (r0 << 32);
    r0 += r1;
    const uint64_t n = (r0 + (1ull << 61)) >> 62;
    r0 -= n << 62;
    *np = (int)n;
    return (double)(int64_t)r0 * kLibmPi63;
}

__host__ __device__ inline void sincosf_libm(float y, float* sinp, float* cosp) {
    double x = y;
    const uint32_t top = (libm_f2u(y) >> 20) & 0x7ffu;
    if (top < ((libm_f2u(0x1.921FB6p-1f) >> 20) & 0x7ffu)) {
        const double x2 = x * x;
        if (top < ((libm_f2u(0x1p-12f) >> 20) & 0x7ffu)) {
            *sinp = y;
            *cosp = 1.0f;
            return;
        }
        libm_sincos_poly(x, x2, false, 0, sinp, cosp);
    } else if (top < ((libm_f2u(120.0f) >> 20) & 0x7ffu)) {
        const double r = x * kLibmHpiInv24;
        const int n = ((int32_t)r + 0x800000) >> 24;
        x = __builtin_fma(-(double)n, kLibmHpi, x);
        const double s = libm_quadrant_sign(n);
        libm_sincos_poly(x * s, x * x, (n & 2) != 0, n, sinp, cosp);
    } else if (top < 0x7f8u) {
        const uint32_t xi = libm_f2u(y);
        const int sign = (int)(xi >> 31);
        int n;
        x = libm_reduce_large(xi, &n);
        const double s = libm_quadrant_sign(n + sign);
        libm_sincos_poly(x * s, x * x, ((n + sign) & 2) != 0, n, sinp, cosp);
    } else {
        *sinp = *cosp = y - y;  // infinity or NaN
    }
}

}  // namespace att
