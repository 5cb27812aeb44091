// xr.hpp -- extended-range ("Xrange") scalars for the perturbation kernels.
//
// A value is mantissa * 2^exp with a float64 (or complex128) mantissa and an
// int32 exponent, with the semantics of the reference's numba record type
// (src/fractalshades/numpy_utils/numba_xr.py:61-829): exponent co-alignment
// on add/sub/compare (_coexp_ufunc, 708-756), renormalisation of products
// whose mantissa exponent leaves +-100 (_need_renorm, 390-413), frexp with a
// mantissa in [1, 2) (_frexp, 674-685) and an exponent shift that clamps at
// the bottom of the exponent field (_exp2_shift, 695-706).
//
// All functions are __host__ __device__: the library computes the serial
// orbit-derivative recurrences on the host and iterates pixels on the GPU
// with the same code, so both sides round identically.  Everything here is
// plain IEEE binary64; the translation units that include it are compiled
// with -ffp-contract=off so that no multiply-add is fused.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define FSG_HD __host__ __device__ __forceinline__

namespace fsg {

struct cplx { double re, im; };
struct xf { double m; int32_t e; };        // real Xrange
struct xc { double re, im; int32_t e; };   // complex Xrange (shared exponent)

FSG_HD cplx mk(double re, double im) { cplx z; z.re = re; z.im = im; return z; }
FSG_HD xf mkf(double m, int32_t e) { xf r; r.m = m; r.e = e; return r; }
FSG_HD xc mkc(double re, double im, int32_t e) { xc r; r.re = re; r.im = im; r.e = e; return r; }

FSG_HD cplx operator+(cplx a, cplx b) { return mk(a.re + b.re, a.im + b.im); }
FSG_HD cplx operator-(cplx a, cplx b) { return mk(a.re - b.re, a.im - b.im); }
// (ac - bd) + (ad + bc) i, four products then one add each: numba's lowering
FSG_HD cplx operator*(cplx a, cplx b) {
    double ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im, bc = a.im * b.re;
    return mk(ac - bd, ad + bc);
}
FSG_HD cplx scale(double s, cplx a) { return mk(s * a.re, s * a.im); }

// ---- bit helpers on the high word of a double ----
FSG_HD uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
FSG_HD double from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }
FSG_HD int32_t biased_exp(double x) { return (int32_t)((bits(x) >> 52) & 0x7ff); }

// mantissa in [1, 2) with the sign kept, exponent = biased - 1023
FSG_HD double frexp1(double m, int32_t &e) {
    uint64_t b = bits(m);
    e = (int32_t)((b >> 52) & 0x7ff) - 1023;
    return from_bits((b & 0x800FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
}
// exponent-field shift clamped at 0 (shift <= 0 in every use below)
FSG_HD double exp2_shift(double m, int32_t shift) {
    uint64_t b = bits(m);
    int32_t ex = (int32_t)((b >> 52) & 0x7ff) + shift;
    ex = ex < 0 ? 0 : ex;
    return from_bits((b & 0x800FFFFFFFFFFFFFull) | ((uint64_t)(uint32_t)ex << 52));
}
FSG_HD bool need_renorm(double m) {
    int32_t d = biased_exp(m) - 1023;
    return (d < 0 ? -d : d) > 100;
}

// hypot with an exact power-of-two prescale: identical on host and device
FSG_HD double hypot2(double x, double y) {
    double a = fabs(x), b = fabs(y);
    if (a < b) { double t = a; a = b; b = t; }
    if (a == 0.) return 0.;
    if (!(a <= 1.7976931348623157e308)) return a;   // inf / nan
    int32_t e = biased_exp(a) - 1023;
    e = e < -1000 ? -1000 : (e > 1000 ? 1000 : e);
    double s = from_bits((uint64_t)(1023 - e) << 52);
    double as = a * s, bs = b * s;
    double r = sqrt(as * as + bs * bs);
    return r * from_bits((uint64_t)(1023 + e) << 52);
}

FSG_HD double ldexp_(double m, int32_t e) { return ldexp(m, e); }

// ---- normalisation ----
FSG_HD xf norm_f(double m, int32_t e) {
    if (m == 0.) return mkf(m, 0);
    int32_t ne;
    double nm = frexp1(m, ne);
    return mkf(nm, e + ne);
}
// co-exponent of two real mantissas: the mantissa of the smaller exponent
// is shifted to the larger one, unless either is zero.  Both candidates are
// formed and selected: a store through "whichever of m0 / m1 shifts" makes
// the compiler place both in private (scratch) memory.
FSG_HD void coexp(double &m0, int32_t e0, double &m1, int32_t e1, int32_t &e) {
    const int32_t d = e0 - e1;
    const bool z0 = m0 == 0., z1 = m1 == 0.;
    const double s0 = exp2_shift(m0, d < 0 ? d : 0), s1 = exp2_shift(m1, d > 0 ? -d : 0);
    m0 = (!z0 && !z1 && e1 > e0) ? s0 : m0;
    m1 = (!z0 && !z1 && e0 > e1) ? s1 : m1;
    e = z0 ? e1 : (z1 ? e0 : (e0 > e1 ? e0 : e1));
}
// co-exponent of two complex mantissas
FSG_HD void coexp(double &r0, double &i0, int32_t e0, double &r1, double &i1, int32_t e1, int32_t &e) {
    const int32_t d = e0 - e1;
    const bool z0 = r0 == 0. && i0 == 0., z1 = r1 == 0. && i1 == 0.;
    const int32_t d0 = d < 0 ? d : 0, d1 = d > 0 ? -d : 0;
    const double sr0 = exp2_shift(r0, d0), si0 = exp2_shift(i0, d0);
    const double sr1 = exp2_shift(r1, d1), si1 = exp2_shift(i1, d1);
    const bool sh0 = !z0 && !z1 && e1 > e0, sh1 = !z0 && !z1 && e0 > e1;
    r0 = sh0 ? sr0 : r0; i0 = sh0 ? si0 : i0;
    r1 = sh1 ? sr1 : r1; i1 = sh1 ? si1 : i1;
    e = z0 ? e1 : (z1 ? e0 : (e0 > e1 ? e0 : e1));
}
FSG_HD xc norm_c(double re, double im, int32_t e) {
    xf r = norm_f(re, e), i = norm_f(im, e);
    int32_t oe;
    coexp(r.m, r.e, i.m, i.e, oe);
    return mkc(r.m, i.m, oe);
}
FSG_HD xc to_xc(cplx z) { return norm_c(z.re, z.im, 0); }
FSG_HD xf to_xf(double x) { return norm_f(x, 0); }
FSG_HD double to_std(xf a) { return ldexp_(a.m, a.e); }
// (x + 1j * y) * s as numba evaluates it (to_standard, numba_xr.py:825-828):
// every float operand is promoted to a complex (v + 0j) and the products and
// sums are complex ones, (a+bi)(c+di) = (ac - bd) + (ad + bc)i.  The result
// is (x s, y s) unless a component is zero; the 0 * y, +0 and im * 0 terms
// then fix the sign of that zero.
FSG_HD cplx numba_cscale(double x, double y, double s) {
    const double re1 = 0. * y - 1. * 0.;           // (0 + 1j) * (y + 0j)
    const double im1 = 0. * 0. + 1. * y;
    const double re2 = x + re1, im2 = 0. + im1;    // (x + 0j) + ...
    return mk(re2 * s - im2 * 0., re2 * 0. + im2 * s);
}
FSG_HD cplx to_std(xc a) {
    xf r = norm_f(a.re, a.e), i = norm_f(a.im, a.e);
    int32_t oe;
    coexp(r.m, r.e, i.m, i.e, oe);
    return numba_cscale(r.m, i.m, ldexp_(1., oe));
}

// ---- arithmetic: complex ----
FSG_HD xc operator+(xc a, xc b) {
    int32_t e; coexp(a.re, a.im, a.e, b.re, b.im, b.e, e);
    return mkc(a.re + b.re, a.im + b.im, e);
}
FSG_HD xc operator-(xc a, xc b) {
    int32_t e; coexp(a.re, a.im, a.e, b.re, b.im, b.e, e);
    return mkc(a.re - b.re, a.im - b.im, e);
}
FSG_HD xc as_operand(cplx b) {
    if (need_renorm(b.re) || need_renorm(b.im)) return norm_c(b.re, b.im, 0);
    return mkc(b.re, b.im, 0);
}
FSG_HD xc operator+(xc a, cplx b) { return a + as_operand(b); }
FSG_HD xc operator+(xc a, xf b) { return a + mkc(b.m, 0., b.e); }
FSG_HD xc renorm_c(double re, double im, int32_t e) {
    if (need_renorm(re) || need_renorm(im)) return norm_c(re, im, e);
    return mkc(re, im, e);
}
FSG_HD xc operator*(xc a, xc b) {
    cplx m = mk(a.re, a.im) * mk(b.re, b.im);
    return renorm_c(m.re, m.im, a.e + b.e);
}
FSG_HD xc operator*(xc a, cplx b) {
    cplx m = mk(a.re, a.im) * b;
    return renorm_c(m.re, m.im, a.e);
}
FSG_HD xc operator*(double s, xc a) { return renorm_c(s * a.re, s * a.im, a.e); }
FSG_HD xf abs2(xc a) { return mkf(a.re * a.re + a.im * a.im, a.e + a.e); }

// ---- arithmetic: real ----
FSG_HD xf operator+(xf a, xf b) { int32_t e; coexp(a.m, a.e, b.m, b.e, e); return mkf(a.m + b.m, e); }
FSG_HD xf operator-(xf a, xf b) { int32_t e; coexp(a.m, a.e, b.m, b.e, e); return mkf(a.m - b.m, e); }
FSG_HD xf operator-(xf a) { return mkf(-a.m, a.e); }
FSG_HD xf xabs(xf a) { return mkf(fabs(a.m), a.e); }
FSG_HD xf renorm_f(double m, int32_t e) { return need_renorm(m) ? norm_f(m, e) : mkf(m, e); }
FSG_HD xf operator*(xf a, xf b) { return renorm_f(a.m * b.m, a.e + b.e); }
FSG_HD xf operator*(double s, xf a) { return renorm_f(s * a.m, a.e); }
FSG_HD xf operator+(xf a, double b) { return a + (need_renorm(b) ? norm_f(b, 0) : mkf(b, 0)); }

// comparisons with a standard double / another Xrange
FSG_HD bool lt(xf a, double b) { int32_t e; coexp(a.m, a.e, b, 0, e); return a.m < b; }
FSG_HD bool ge(xf a, double b) { int32_t e; coexp(a.m, a.e, b, 0, e); return a.m >= b; }
FSG_HD bool le(xf a, double b) { int32_t e; coexp(a.m, a.e, b, 0, e); return a.m <= b; }
FSG_HD bool le(xf a, xf b) { int32_t e; coexp(a.m, a.e, b.m, b.e, e); return a.m <= b.m; }

}  // namespace fsg
