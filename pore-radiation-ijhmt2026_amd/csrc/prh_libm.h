// sin and cos with the bits of the host C library (glibc 2.35, x86-64).
//
// generate_rays (porerad/view_factors.py:168-176) takes cos and sin of
// phi = 2*pi*u through NumPy, which calls the C library.  To give the device
// the same direction bits, this header restates the published dbl-64
// algorithm of glibc's sysdeps/ieee754/dbl-64/s_sin.c (the IBM Accurate
// Mathematical Library sin/cos, in the form it has had since glibc 2.28):
//
//   |x| < 2^-26 (sin) or 2^-27 (cos)   x, or 1
//   |x| < 0.855469                     table method at x
//   |x| < 2.426265                     table method at pi/2 - |x| as a double-double
//   |x| < 105414350                    Cody-Waite reduction by pi/2 in four parts,
//                                      then the table method in the quadrant
//
// The table method splits x = X_i + h with X_i = i/128 (rounded by adding
// 1.5*2^45), evaluates short odd/even polynomials in h and combines them with
// sin X_i and cos X_i held as double-doubles (prh_sincos_table.h, from
// gen_sincos_table.py); below |x| = 0.126 sin uses a degree-9 Taylor form
// with the low part folded in.
//
// x86-64 glibc runs a copy of that source built with -mfma when the CPU has
// FMA and AVX2; the compiler then contracts every product whose only uses are
// sums into a fused multiply-add.  Those contractions are written out here as
// fma() (this library is built with -ffp-contract=off), so the results are
// the FMA build's, which is the one NumPy reaches on any x86-64 host with
// FMA.  Checked bit for bit against the host's sin/cos on 3.6e8 arguments
// (random phi = 2*pi*k*2^-53 and dense runs around every branch point).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#include "prh_sincos_table.h"

namespace prh {
namespace libm {

// Polynomial coefficients of the table method and the Taylor form.
constexpr double SN3 = -1.66666666666664880952546298448555E-01;
constexpr double SN5 = 8.33333214285722277379541354343671E-03;
constexpr double CS2 = 4.99999999999999999999950396842453E-01;
constexpr double CS4 = -4.16666666666664434524222570944589E-02;
constexpr double CS6 = 1.38888874007937613028114285595617E-03;
constexpr double T1 = -0.16666666666666666;
constexpr double T2 = 0.0083333333333323288;
constexpr double T3 = -0.00019841269834414642;
constexpr double T4 = 2.755729806860771e-06;
constexpr double T5 = -2.5022014848318398e-08;
// 1.5 * 2^45: adding it rounds |x| to a multiple of 1/128 whose count sits
// in the low word.
constexpr double NODE_ROUND = 52776558133248.0;
// pi/2 as a double-double.
constexpr double HALF_PI_HI = 0x1.921fb54442d18p0;
constexpr double HALF_PI_LO = 0x1.1a62633145c07p-54;
// Reduction by pi/2: 2/pi, 1.5*2^52 (round to integer) and pi/2 in four parts.
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double ROUND_INT = 6755399441055744.0;
constexpr double PIO2_1 = 0x1.921fb58p0;
constexpr double PIO2_2 = -0x1.dde973cp-27;
constexpr double PIO2_3 = -0x1.cb3b398p-55;
constexpr double PIO2_4 = -0x1.d747f23e32ed7p-83;

__host__ __device__ inline uint32_t high_word(double x)
{
    uint64_t u;
    memcpy(&u, &x, sizeof u);
    return (uint32_t)(u >> 32);
}

__host__ __device__ inline uint32_t low_word(double x)
{
    uint64_t u;
    memcpy(&u, &x, sizeof u);
    return (uint32_t)u;
}

// sin(a + da) for |a| < 0.126.
__host__ __device__ inline double taylor_sin(double xx, double a, double da)
{
    const double p = fma(fma(fma(fma(T5, xx, T4), xx, T3), xx, T2), xx, T1);
    const double t = fma(fma(p, a, -(0.5 * da)), xx, da);
    return a + t;
}

// cos(x + dx) by the table method.
__host__ __device__ inline double table_cos(double x, double dx)
{
    if (x < 0.0) dx = -dx;
    const double u = NODE_ROUND + fabs(x);
    x = fabs(x) - (u - NODE_ROUND) + dx;
    const double xx = x * x;
    const double s = fma(x * xx, fma(xx, SN5, SN3), x);
    const double c = xx * fma(xx, fma(xx, CS6, CS4), CS2);
    const double *node = SINCOS_TABLE[low_word(u)];
    const double sn = node[0], ssn = node[1], cs = node[2], ccs = node[3];
    const double cor = fma(-sn, s, fma(-cs, c, fma(-s, ssn, ccs)));
    return cs + cor;
}

// sin(x + dx): Taylor form near zero, else the table method.
__host__ __device__ inline double table_sin(double x, double dx)
{
    const double x_in = x;
    if (fabs(x) < 0.126) return taylor_sin(x * x, x, dx);
    if (x <= 0.0) dx = -dx;
    const double u = NODE_ROUND + fabs(x);
    x = fabs(x) - (u - NODE_ROUND);
    const double xx = x * x;
    const double s = x + fma(x * xx, fma(xx, SN5, SN3), dx);
    const double c = fma(x, dx, xx * fma(xx, fma(xx, CS6, CS4), CS2));
    const double *node = SINCOS_TABLE[low_word(u)];
    const double sn = node[0], ssn = node[1], cs = node[2], ccs = node[3];
    const double cor = fma(cs, s, fma(-sn, c, fma(s, ccs, ssn)));
    return copysign(sn + cor, x_in);
}

// x = n*pi/2 + (a + da), |a| <= pi/4 (about); returns n mod 4.
__host__ __device__ inline int reduce_half_pi(double x, double &a, double &da)
{
    const double t = fma(x, TWO_OVER_PI, ROUND_INT);
    const double xn = t - ROUND_INT;
    const double y = fma(-xn, PIO2_2, fma(-xn, PIO2_1, x));
    const int n = (int)(low_word(t) & 3u);
    const double t2 = fma(-xn, PIO2_3, y);
    double db = fma(-xn, PIO2_3, y - t2);
    const double b = fma(-xn, PIO2_4, t2);
    db += fma(-xn, PIO2_4, t2 - b);
    a = b;
    da = db;
    return n;
}

// Largest |x| the reduction above serves (glibc switches to a multi-word
// reduction beyond it; generate_rays never gets there: phi < 2*pi).
constexpr double SINCOS_MAX_ARG = 105414350.0;

// sin(x) and cos(x) for finite |x| < SINCOS_MAX_ARG, each equal to the C
// library's sin(x) and cos(x).  Every path evaluates one table_sin and one
// table_cos, so a wavefront diverges only on the Taylor branch.
__host__ __device__ inline void sincos(double x, double &sin_x, double &cos_x)
{
    const uint32_t k = high_word(x) & 0x7fffffffu;
    // Arguments of the one table_sin and the one table_cos evaluation.
    double ts_arg, ts_low, tc_arg, tc_low;
    bool swapped = false;  // sin(x) comes from table_cos and cos(x) from table_sin
    bool middle = false;
    int quadrant = 0;      // sin(x) is in quadrant n, cos(x) in quadrant n + 1
    if (k < 0x3feb6000u) {  // |x| < 0.855469
        ts_arg = tc_arg = x;
        ts_low = tc_low = 0.0;
    } else if (k < 0x400368fdu) {  // |x| < 2.426265
        // sin(x) = sign(x) cos(y + pi/2 low) with the low part passed on
        // unsummed; cos(x) = sin(a + da) with the double-double summed first.
        const double y = HALF_PI_HI - fabs(x);
        const double a = y + HALF_PI_LO;
        tc_arg = y;
        tc_low = HALF_PI_LO;
        ts_arg = a;
        ts_low = (y - a) + HALF_PI_LO;
        swapped = true;
        middle = true;
    } else {
        double a, da;
        quadrant = reduce_half_pi(x, a, da);
        ts_arg = tc_arg = a;
        ts_low = tc_low = da;
        swapped = (quadrant & 1) != 0;
    }
    const double ts = table_sin(ts_arg, ts_low);
    const double tc = table_cos(tc_arg, tc_low);
    double s = swapped ? tc : ts;
    double c = swapped ? ts : tc;
    if (middle) s = copysign(s, x);
    if (quadrant & 2) s = -s;
    if ((quadrant + 1) & 2) c = -c;
    if (k < 0x3e500000u) s = x;    // |x| < 2^-26
    if (k < 0x3e400000u) c = 1.0;  // |x| < 2^-27
    sin_x = s;
    cos_x = c;
}

}  // namespace libm
}  // namespace prh
