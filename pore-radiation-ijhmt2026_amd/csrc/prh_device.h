// Device helpers shared by the tracer kernels: PCG64 stream, Lambertian ray
// construction and the FP64 traversal set-up, each following
// porerad/view_factors.py literally (same operations, same rounding order;
// the library is compiled with -ffp-contract=off).
#pragma once

#include "prh_internal.h"
#include "prh_libm.h"

namespace prh {

constexpr double TIE_TOLERANCE = 1.0e-12;   // view_factors.py:29
constexpr double FACE_EDGE_INSET = 1.0e-12; // view_factors.py:30
constexpr double INF_T = 1.0e300;           // view_factors.py:28
constexpr double TWO_PI = 2.0 * 3.141592653589793;

// np.remainder(a, b) for b > 0: fmod, plus b when the signs differ, +0.0 for
// an exact zero.  Origins lie within one period of [0, b), where the
// remainder is exact (Sterbenz) and needs no library fmod.
__host__ __device__ __forceinline__ double py_mod(double a, double b)
{
    double m;
    if (a >= 0.0 && a < b) m = a;
    else if (a >= b && a < 2.0 * b) m = a - b;
    else if (a < 0.0 && a > -b) m = a;
    else m = fmod(a, b);
    if (m != 0.0) {
        if (m < 0.0) m += b;
    } else {
        m = 0.0;
    }
    return m;
}

__device__ __forceinline__ int64_t py_imod64(int64_t a, int64_t b)
{
    int64_t m = a % b;
    if (m < 0) m += b;
    return m;
}

// PCG64 XSL-RR output after one LCG step.
__device__ __forceinline__ uint64_t pcg_next(u128 &s, const u128 inc)
{
    const u128 mult = make_u128(PCG_MULT_HI, PCG_MULT_LO);
    s = add128(mul128(s, mult), inc);
    const uint64_t x = s.hi ^ s.lo;
    const unsigned rot = (unsigned)(s.hi >> 58);
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

// Generator.random(): (next64 >> 11) * 2^-53.
__device__ __forceinline__ double pcg_uniform(u128 &s, const u128 inc)
{
    return (double)(pcg_next(s, inc) >> 11) * (1.0 / 9007199254740992.0);
}

// Advance the LCG by delta steps using the 2^b jump tables.
__device__ __forceinline__ u128 pcg_jump(u128 s, const u128 inc, uint64_t delta, const JumpTables &T)
{
    for (int b = 0; delta; ++b, delta >>= 1) {
        if (delta & 1u) s = add128(mul128(T.A[b], s), mul128(T.G[b], inc));
    }
    return s;
}

__device__ __forceinline__ double clip_uniform(double u)
{
    const double lo = FACE_EDGE_INSET, hi = 1.0 - FACE_EDGE_INSET;
    return u < lo ? lo : (u > hi ? hi : u);
}

// a / b and sqrt(x), correctly rounded, for operands that need none of the
// exponent scaling, class tests and fix-ups of the compiler's IEEE
// expansions: the expansions' own core sequences (reciprocal or reciprocal
// square root, Newton steps, one fused correction), which give the
// correctly rounded result whenever no intermediate leaves the normal
// range.  div_rn_plain: b finite, 2^-900 < |b| < 2^900, a == 0 or
// 2^-900 < |a| < 2^900, and |a / b| within the same range; sqrt_rn_plain:
// 2^-767 < x < 2^1000.  Generated rays stay far inside both: unit-vector
// components are 0 or above 1e-25 in magnitude, and a set-up numerator
// (a cell face minus an origin coordinate) is 0 or at least the smaller of
// 2^-52 and the face-edge inset.
__device__ __forceinline__ double div_rn_plain(double a, double b)
{
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    const double q = a * r;
    const double res = fma(-b, q, a);
    return fma(res, r, q);
}

__device__ __forceinline__ double sqrt_rn_plain(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    double d = fma(-g, g, x);
    h = fma(h, r, h);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

// One Lambertian ray of a patch (generate_rays, view_factors.py:147-181).
// normal_code: 0..5 = +x, -x, +y, -y, +z, -z (walls: hot 0, cold 1).
__device__ __forceinline__ void make_ray(u128 &s, const u128 inc, const double po0, const double po1, const double po2,
                                int normal_code, double fny, double fnz, double o[3], double d[3])
{
    const double u0 = pcg_uniform(s, inc);
    const double u1 = pcg_uniform(s, inc);
    const double u2 = pcg_uniform(s, inc);
    const double u3 = pcg_uniform(s, inc);
    const double ta = clip_uniform(u0) - 0.5;
    const double tb = clip_uniform(u1) - 0.5;
    const int axis = normal_code >> 1;
    const double sign = (normal_code & 1) ? -1.0 : 1.0;
    double x = po0, y = po1, z = po2;
    if (axis == 0) {
        y = y + ta;
        z = z + tb;
    } else if (axis == 1) {
        x = x + ta;
        z = z + tb;
    } else {
        x = x + ta;
        y = y + tb;
    }
    o[0] = x;
    o[1] = py_mod(y, fny);
    o[2] = py_mod(z, fnz);
    const double radius = u2 > 0.0 ? sqrt_rn_plain(u2) : 0.0;  // u2 is 0 or at least 2^-53
    const double phi = TWO_PI * u3;
    double sphi, cphi;
    libm::sincos(phi, sphi, cphi);  // the C library's bits, as NumPy's cos/sin give them
    const double la = radius * cphi;
    const double lb = radius * sphi;
    const double ct = sqrt_rn_plain(1.0 - u2);  // 1 - u2 >= 2^-53
    double d0, d1, d2;
    if (axis == 0) {
        d0 = sign * ct;
        d1 = la;
        d2 = sign * lb;
    } else if (axis == 1) {
        d1 = sign * ct;
        d0 = la;
        d2 = sign * lb;
    } else {
        d2 = sign * ct;
        d0 = la;
        d1 = sign * lb;
    }
    const double norm = sqrt_rn_plain((d0 * d0 + d1 * d1) + d2 * d2);  // about 1
    d[0] = div_rn_plain(d0, norm);
    d[1] = div_rn_plain(d1, norm);
    d[2] = div_rn_plain(d2, norm);
}

// _initial_cell, view_factors.py:185-190 (round() = round-half-even), as an
// integral double: exact for any origin the host accepts (|o| < 2^31).
__device__ __forceinline__ double initial_cell(double position, double direction)
{
    double index = floor(position);
    if (direction < 0.0 && fabs(position - rint(position)) <= TIE_TOLERANCE) index -= 1.0;
    return index;
}

// _axis_state, view_factors.py:193-199 (index + 1.0 is exact: index is integral).
// PLAIN: the operands are a generated ray's (div_rn_plain applies).
template <bool PLAIN = false>
__device__ __forceinline__ void axis_state(double position, double direction, double index, int &step,
                                           double &t_max, double &t_delta)
{
    // Branch-free: a branch per sign would run both division pairs in a
    // wavefront whose rays point both ways.  The numerator is the
    // reference's for the sign, and 1 / |d| is its -1.0 / d for d < 0
    // (division rounds symmetrically in sign).
    const bool up = direction > 0.0;
    const double num = up ? (index + 1.0) - position : index - position;
    const double tm = PLAIN ? div_rn_plain(num, direction) : num / direction;
    const double td = PLAIN ? div_rn_plain(1.0, fabs(direction)) : 1.0 / fabs(direction);
    step = up ? 1 : (direction < 0.0 ? -1 : 0);
    t_max = direction != 0.0 ? tm : INF_T;
    t_delta = direction != 0.0 ? td : INF_T;
}

// Python's f % n for an integral double f with |f| < 2^52, inv_n = 1 / n
// rounded: the quotient's floor may be one off, so one correction either way
// (every value is an exact integer).
__device__ __forceinline__ int floor_mod(double f, int n, double inv_n)
{
    const double dn = (double)n;
    double r = f - dn * floor(f * inv_n);
    r = r < 0.0 ? r + dn : r;
    r = r >= dn ? r - dn : r;
    return (int)r;
}

// Step cap of _trace_kernel, view_factors.py:223-224.
__device__ __forceinline__ uint32_t max_steps_of(double dx, double dy, double dz, int nx, int ny, int nz)
{
    const double crossing_scale = 8.0 * (double)nx * (1.0 + (fabs(dy) + fabs(dz)) / fabs(dx)) + 100.0;
    const double capped = crossing_scale < 100000000.0 ? crossing_scale : 100000000.0;
    const int64_t a = 10 * (int64_t)nx * ny * nz + 10;
    const int64_t b = (int64_t)capped;
    return (uint32_t)(a > b ? a : b);
}

}  // namespace prh
