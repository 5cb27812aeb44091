// xl.hpp -- extended-range binary64 with lazy, branch-free normalisation.
//
// The reference's Xrange record (numba_xr.py:61-829) renormalises products
// whose mantissa exponent leaves +-100 and aligns the operands of a sum with
// a branchy exponent shift.  Rescaling a mantissa by a power of two is exact,
// so where and how often a value is renormalised does not change any
// rounding: a sum rounds the exact sum of the two values once, a product
// rounds the exact product, whatever the split between mantissa and
// exponent -- provided no mantissa becomes subnormal or overflows.  This
// type keeps that provision by renormalising every result (three VALU
// instructions: v_frexp_exp_i32_f64 and two v_ldexp_f64) and aligns sums
// with v_ldexp_f64; a shifted operand that underflows is then below the
// rounding error of the other one, exactly as in the reference's shift.
// Zero carries the exponent XL_ZERO_E so that it never drags the other
// operand of a sum down.
//
// Conversions that the reference defines bit by bit (to_standard, the
// comparisons) are done with the reference's own procedure (xr.hpp).
#pragma once
#include "xr.hpp"

namespace fsg {

struct xl { double re, im; int32_t e; };
struct xlf { double m; int32_t e; };

// exponent of zero: below every exponent a value reaches (|e| < 2^17 down
// to 1e-20000), and small enough that the product of up to 31 zeros -- a
// power of a zero orbit point in the z^d + c kernels takes up to 7 -- still
// fits in int32
constexpr int32_t XL_ZERO_E = -(1 << 26);

#define FSG_DEV __device__ __forceinline__

FSG_DEV int32_t dexp(double v) { return __builtin_amdgcn_frexp_exp(v); }

// normalised: larger component in [1, 2), or (0, 0, XL_ZERO_E)
FSG_DEV xl xl_n(double re, double im, int32_t e) {
    const double mx = fmax(fabs(re), fabs(im));
    const int32_t k = dexp(mx) - 1;
    xl r;
    r.re = ldexp(re, -k);
    r.im = ldexp(im, -k);
    r.e = (mx == 0.) ? XL_ZERO_E : e + k;
    return r;
}
FSG_DEV xlf xlf_n(double m, int32_t e) {
    const int32_t k = dexp(m) - 1;
    xlf r;
    r.m = ldexp(m, -k);
    r.e = (m == 0.) ? XL_ZERO_E : e + k;
    return r;
}

FSG_DEV xl xl_of(cplx z) { return xl_n(z.re, z.im, 0); }
FSG_DEV xl xl_of(xc a) { return xl_n(a.re, a.im, a.e); }
FSG_DEV xl xl_of(double re, double im, int32_t e) { return xl_n(re, im, e); }
FSG_DEV xl xl_zero() { xl r; r.re = 0.; r.im = 0.; r.e = XL_ZERO_E; return r; }
FSG_DEV xlf xlf_of(double m, int32_t e) { return xlf_n(m, e); }
FSG_DEV xlf xlf_of(xf a) { return xlf_n(a.m, a.e); }
FSG_DEV xlf xlf_zero() { xlf r; r.m = 0.; r.e = XL_ZERO_E; return r; }

// back to the reference's record, for its bit-defined conversions
FSG_DEV xc xc_of(xl a) { return mkc(a.re, a.im, a.e == XL_ZERO_E ? 0 : a.e); }
FSG_DEV xf xf_of(xlf a) { return mkf(a.m, a.e == XL_ZERO_E ? 0 : a.e); }

// alignment shift of exponent e to E >= e; v_ldexp_f64 takes any int32
// exponent, and the zero sentinel keeps every difference within int32
FSG_DEV int32_t shift_of(int32_t e, int32_t E) { return e - E; }

// ---- complex ----
FSG_DEV xl operator+(xl a, xl b) {
    const int32_t E = max(a.e, b.e);
    const int32_t sa = shift_of(a.e, E), sb = shift_of(b.e, E);
    return xl_n(ldexp(a.re, sa) + ldexp(b.re, sb), ldexp(a.im, sa) + ldexp(b.im, sb), E);
}
FSG_DEV xl operator-(xl a, xl b) {
    const int32_t E = max(a.e, b.e);
    const int32_t sa = shift_of(a.e, E), sb = shift_of(b.e, E);
    return xl_n(ldexp(a.re, sa) - ldexp(b.re, sb), ldexp(a.im, sa) - ldexp(b.im, sb), E);
}
FSG_DEV xl operator*(xl a, xl b) {
    const double ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im, bc = a.im * b.re;
    return xl_n(ac - bd, ad + bc, a.e + b.e);
}
FSG_DEV xl operator*(xl a, cplx b) {
    const double ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im, bc = a.im * b.re;
    return xl_n(ac - bd, ad + bc, a.e);
}
FSG_DEV xl muls(double s, xl a) { return xl_n(s * a.re, s * a.im, a.e); }
FSG_DEV xl operator+(xl a, cplx b) { return a + xl_of(b); }
FSG_DEV xl operator+(xl a, xlf b) { xl c; c.re = b.m; c.im = 0.; c.e = b.e; return a + c; }

// |a|^2 as a real value (mantissa products rounded like extended_abs2)
FSG_DEV xlf abs2(xl a) { return xlf_n(a.re * a.re + a.im * a.im, 2 * max(a.e, -(1 << 27))); }

// ---- real ----
FSG_DEV xlf operator+(xlf a, xlf b) {
    const int32_t E = max(a.e, b.e);
    return xlf_n(ldexp(a.m, shift_of(a.e, E)) + ldexp(b.m, shift_of(b.e, E)), E);
}
FSG_DEV xlf operator-(xlf a, xlf b) {
    const int32_t E = max(a.e, b.e);
    return xlf_n(ldexp(a.m, shift_of(a.e, E)) - ldexp(b.m, shift_of(b.e, E)), E);
}
FSG_DEV xlf operator-(xlf a) { a.m = -a.m; return a; }
FSG_DEV xlf xabs(xlf a) { a.m = fabs(a.m); return a; }
FSG_DEV xlf operator*(xlf a, xlf b) { return xlf_n(a.m * b.m, a.e + b.e); }
FSG_DEV xlf operator*(double s, xlf a) { return xlf_n(s * a.m, a.e); }
FSG_DEV xlf mul_s(double s, xlf a) { return xlf_n(s * a.m, a.e); }
FSG_DEV xlf operator+(xlf a, double b) { return a + xlf_n(b, 0); }

// value comparisons (exact alignment; equal to the reference's co-exponent
// comparison for every operand pair the loop produces)
FSG_DEV bool lt(xlf a, double b) {
    const xlf bb = xlf_n(b, 0);
    const int32_t E = max(a.e, bb.e);
    return ldexp(a.m, shift_of(a.e, E)) < ldexp(bb.m, shift_of(bb.e, E));
}
FSG_DEV bool ge(xlf a, double b) { return !lt(a, b); }
FSG_DEV bool le(xlf a, double b) {
    const xlf bb = xlf_n(b, 0);
    const int32_t E = max(a.e, bb.e);
    return ldexp(a.m, shift_of(a.e, E)) <= ldexp(bb.m, shift_of(bb.e, E));
}
FSG_DEV bool le(xlf a, xlf b) {
    const int32_t E = max(a.e, b.e);
    return ldexp(a.m, shift_of(a.e, E)) <= ldexp(b.m, shift_of(b.e, E));
}

// Conversion to binary64 with the reference's procedure (to_standard,
// numba_xr.py:802-829: per-component frexp, co-exponent, product by
// 2^exp in complex arithmetic).  For a normalised value whose smaller
// component is zero or a normal number the co-exponent shift is exact and
// the procedure reduces to two products by 2^e, provided neither product is
// zero (numba_cscale sets the sign of a zero); otherwise it is run as
// written.
FSG_DEV cplx to_std(xl a) {
    const double mn = fmin(fabs(a.re), fabs(a.im));
    if (mn != 0. && mn < 0x1p-1022) return to_std(xc_of(a));
    const double s = ldexp(1., max(a.e, -1100));
    const cplx p = mk(a.re * s, a.im * s);
    if (p.re != 0. && p.im != 0.) return p;
    return numba_cscale(a.re, a.im, s);
}
FSG_DEV double to_std(xlf a) { return ldexp(a.m, max(a.e, -2200)); }

// ---- lazily normalised intermediates ---------------------------------------
// Within one iteration, values are combined without renormalising: their
// mantissas start in [1, 2) (state, orbit point) or within 2^+-100 (the
// reference's derivative paths) and a few products and sums cannot take
// them near the binary64 limits, so every operation still rounds exactly
// as the reference's.  Only a zero result needs care: it takes XL_ZERO_E so
// that it never sets the exponent of a later sum.  The state is
// renormalised (xl_of) once per iteration.
struct xz { double re, im; int32_t e; };

FSG_DEV xz xz_mk(double re, double im, int32_t e) {
    xz r;
    r.re = re; r.im = im;
    r.e = (fmax(fabs(re), fabs(im)) == 0.) ? XL_ZERO_E : e;
    return r;
}
FSG_DEV xz as_z(xl a) { xz r; r.re = a.re; r.im = a.im; r.e = a.e; return r; }
FSG_DEV xz xz_zero() { xz r; r.re = 0.; r.im = 0.; r.e = XL_ZERO_E; return r; }
FSG_DEV xl xl_of(xz a) { return xl_n(a.re, a.im, a.e); }

FSG_DEV xz operator+(xz a, xz b) {
    const int32_t E = max(a.e, b.e);
    const int32_t sa = shift_of(a.e, E), sb = shift_of(b.e, E);
    return xz_mk(ldexp(a.re, sa) + ldexp(b.re, sb), ldexp(a.im, sa) + ldexp(b.im, sb), E);
}
// A product of two values is zero only when a factor is (mantissas stay
// within 2^+-300, far from underflow), and a zero factor carries the
// sentinel, so the product's exponent is already below any other: no
// zero test is needed.  (At most 8 sentinels add up within one iteration --
// Z^7 times a zero path value -- far from the int32 limit; the state is
// renormalised after.)
FSG_DEV xz operator*(xz a, xz b) {
    const double ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im, bc = a.im * b.re;
    xz r;
    r.re = ac - bd; r.im = ad + bc; r.e = a.e + b.e;
    return r;
}
FSG_DEV xz operator*(xz a, cplx b) {
    const double ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im, bc = a.im * b.re;
    return xz_mk(ac - bd, ad + bc, a.e);
}
FSG_DEV xz muls(double s, xz a) { xz r; r.re = s * a.re; r.im = s * a.im; r.e = a.e; return r; }   // s != 0
// 2^f (a + b), normalised: the sum of the last operation of an iteration.
// The normalisation tests the result for zero, so the sum needs no zero
// test of its own; the factor 2^f goes to the exponent.
FSG_DEV xl add_n(xz a, xz b, int32_t f = 0) {
    const int32_t E = max(a.e, b.e);
    const int32_t sa = shift_of(a.e, E), sb = shift_of(b.e, E);
    return xl_n(ldexp(a.re, sa) + ldexp(b.re, sb), ldexp(a.im, sa) + ldexp(b.im, sb), E + f);
}
// a b, normalised (the normalisation's zero test replaces the product's)
FSG_DEV xl mul_n(xz a, cplx b) {
    const double ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im, bc = a.im * b.re;
    return xl_n(ac - bd, ad + bc, a.e);
}
// a + b for a sum that only a comparison reads: a zero result needs no
// exponent of its own there
FSG_DEV xz add_raw(xz a, xz b) {
    const int32_t E = max(a.e, b.e);
    const int32_t sa = shift_of(a.e, E), sb = shift_of(b.e, E);
    xz r;
    r.re = ldexp(a.re, sa) + ldexp(b.re, sb); r.im = ldexp(a.im, sa) + ldexp(b.im, sb); r.e = E;
    return r;
}
FSG_DEV xlf abs2(xz a) { return xlf_n(a.re * a.re + a.im * a.im, 2 * max(a.e, -(1 << 27))); }
// |a|^2 <= |b|^2 for normalised a, b: le(abs2(a), abs2(b)) without
// normalising the two squares -- each sum of squares lies in [1, 8) (or is
// zero), so the aligned larger one is at least 1 and a shift that takes the
// other below the normal range decides the comparison the same way
FSG_DEV bool le_abs2(xl a, xl b) {
    const double qa = a.re * a.re + a.im * a.im, qb = b.re * b.re + b.im * b.im;
    const int32_t ea = 2 * max(a.e, -(1 << 27)), eb = 2 * max(b.e, -(1 << 27));
    const int32_t E = max(ea, eb);
    return ldexp(qa, ea - E) <= ldexp(qb, eb - E);
}
// |a|^2 < b for a binary64 b > 0: the mantissa sum rounds as abs2's, and
// scaling it by 2^(2e) is exact unless the result leaves the normal range,
// where it lies far on one side of b whichever way it rounds
FSG_DEV bool abs2_lt(xz a, double b) {
    return ldexp(a.re * a.re + a.im * a.im, 2 * max(a.e, -(1 << 27))) < b;
}

// Lazily normalised reals for the burning-ship step (the same argument as
// xz: mantissas of the operands start in [1, 2), one iteration's products
// of at most three factors and sums of a few terms stay far inside the
// binary64 range, so each operation rounds as the reference's).  A sum that
// cancels to zero, or a product by a zero scalar (the folds' 0 / +-1 / +-2
// derivative factors), takes XL_ZERO_E; a product of two values is zero only
// when a factor already carries it.
struct xzf { double m; int32_t e; };

FSG_DEV xzf as_zf(xlf a) { xzf r; r.m = a.m; r.e = a.e; return r; }
FSG_DEV xlf xlf_of(xzf a) { return xlf_n(a.m, a.e); }
FSG_DEV xzf zf_mk(double m, int32_t e) { xzf r; r.m = m; r.e = (m == 0.) ? XL_ZERO_E : e; return r; }
FSG_DEV xzf operator+(xzf a, xzf b) {
    const int32_t E = max(a.e, b.e);
    return zf_mk(ldexp(a.m, shift_of(a.e, E)) + ldexp(b.m, shift_of(b.e, E)), E);
}
FSG_DEV xzf operator-(xzf a, xzf b) {
    const int32_t E = max(a.e, b.e);
    return zf_mk(ldexp(a.m, shift_of(a.e, E)) - ldexp(b.m, shift_of(b.e, E)), E);
}
FSG_DEV xzf operator-(xzf a) { a.m = -a.m; return a; }
FSG_DEV xzf operator*(xzf a, xzf b) { xzf r; r.m = a.m * b.m; r.e = a.e + b.e; return r; }
FSG_DEV xzf operator*(double s, xzf a) { xzf r; r.m = s * a.m; r.e = (s == 0.) ? XL_ZERO_E : a.e; return r; }
FSG_DEV xzf mul_s(double s, xzf a) { return s * a; }
// s * a for a binary64 s of any magnitude (BLA coefficients): s is split
// into a mantissa in [1, 2) and an exponent, so that sums of such products
// stay in range; the product rounds as s * a.m does
FSG_DEV xzf mul_w(double s, xzf a) {
    const int32_t k = dexp(s) - 1;
    xzf r;
    r.m = ldexp(s, -k) * a.m;
    r.e = (s == 0.) ? XL_ZERO_E : a.e + k;
    return r;
}
FSG_DEV xzf absv(xzf x) { x.m = fabs(x.m); return x; }
// comparisons with zero: the sign of the mantissa
FSG_DEV xzf diffabs(xzf X, xzf x) {
    const xzf Xx = X + x;
    if (X.m >= 0.) { if (Xx.m >= 0.) return x; return -(2. * X + x); }
    if (Xx.m <= 0.) return -x;
    return 2. * X + x;
}
FSG_DEV double ddX(xzf X, xzf x) {
    const xzf Xx = X + x;
    if (X.m >= 0.) return Xx.m >= 0. ? 0. : -2.;
    return Xx.m <= 0. ? 0. : 2.;
}
FSG_DEV double ddx(xzf X, xzf x) {
    const xzf Xx = X + x;
    if (X.m >= 0.) return Xx.m >= 0. ? 1. : -1.;
    return Xx.m <= 0. ? -1. : 1.;
}
FSG_DEV double sgn(xzf x) { return x.m < 0. ? -1. : 1.; }

// burning-ship helpers on Xrange reals (burning_ship.py:19-60)
FSG_DEV xlf diffabs(xlf X, xlf x) {
    if (ge(X, 0.)) { if (ge(X + x, 0.)) return x; return -(2. * X + x); }
    if (le(X + x, 0.)) return -x;
    return 2. * X + x;
}
FSG_DEV double ddX(xlf X, xlf x) {
    if (ge(X, 0.)) return ge(X + x, 0.) ? 0. : -2.;
    return le(X + x, 0.) ? 0. : 2.;
}
FSG_DEV double ddx(xlf X, xlf x) {
    if (ge(X, 0.)) return ge(X + x, 0.) ? 1. : -1.;
    return le(X + x, 0.) ? -1. : 1.;
}
FSG_DEV double sgn(xlf x) { return lt(x, 0.) ? -1. : 1.; }
FSG_DEV xlf absv(xlf x) { return xabs(x); }

}  // namespace fsg
