/*
 * fs_oracle.c -- CPU restatement of the Fractalshades perturbation loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the
 * smoke check of __graft_entry__ and the cpu_baseline leg of bench.py load it
 * (through oracle/_build/libfs_oracle.so) to check and time the product.  The
 * product (fractalshades_amd) never links, loads or calls it.
 *
 * Every function restates one reference function; citations are relative to
 * the reference repository root (src/fractalshades/...).  Arithmetic is done
 * in IEEE binary64 with -ffp-contract=off and in the same operation order as
 * the reference source.  The reference compiles these loops with numba
 * fastmath=True, which lets LLVM contract and reassociate: that part is not
 * reproducible and this restatement pins the un-contracted order.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - Xrange ops: exact vs float64 for add/sub on the reference test vectors
 *     (tests/test_numba_xr.py:23-46 seeds 100/800), <=2 ulp mul, exact
 *     to_standard / normalize round trips.
 *   - iterate (M2/Mn/BS, std and Xrange): escape counts vs brute-force mpmath
 *     direct iteration; Xrange mode forced on a shallow frame == std mode;
 *     BLA on == BLA off escape counts.
 *   - BLA build: merge == explicit composition of the halves; radii
 *     non-increasing under merge.
 *   - derivative paths: linear in the scale (dZndc(s) = s * dZndc(1)) and vs
 *     mpmath recurrences; the scale is dx, or dx * projection.scale
 *     (scale_deriv_xr, perturbation.py:459-467) for the exponential map.
 *
 * Parity unpinned: the exponential map's projected pixel offsets are an
 * input here (computed by the caller with numpy's complex exp); the
 * reference's numba exp may differ from it in the last bit of an offset.
 */
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* complex helpers: numba's complex128 ops are (a+bi)(c+di) = (ac-bd)+(ad+bc)i
 * (numba/cpython/numbers.py complex_mul_impl); real*complex is componentwise
 * once fastmath drops the 0*x terms. */
typedef struct { double re, im; } cpx;

static inline cpx C(double re, double im) { cpx z; z.re = re; z.im = im; return z; }
static inline cpx c_add(cpx a, cpx b) { return C(a.re + b.re, a.im + b.im); }
static inline cpx c_sub(cpx a, cpx b) { return C(a.re - b.re, a.im - b.im); }
static inline cpx c_mul(cpx a, cpx b) {
    double ac = a.re * b.re, bd = a.im * b.im, ad = a.re * b.im, bc = a.im * b.re;
    return C(ac - bd, ad + bc);
}
static inline cpx c_scale(double s, cpx a) { return C(s * a.re, s * a.im); }
static inline int c_is_zero(cpx a) { return (a.re == 0.) && (a.im == 0.); }

/* abs(complex): numba lowers it to hypot.  A fixed, power-of-two scaled
 * formula is used so that the product kernel can reproduce it bit for bit.
 * It differs from libm hypot by at most 1 ulp. */
static inline uint64_t d2u(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double u2d(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

static double fs_hypot(double x, double y) {
    double a = fabs(x), b = fabs(y);
    if (a < b) { double t = a; a = b; b = t; }
    if (a == 0.) return 0.;
    if (!isfinite(a)) return a;
    /* scale a into [1, 2) with an exact power of two */
    int32_t e = (int32_t)((d2u(a) >> 52) & 0x7ff) - 1023;
    if (e < -1000) e = -1000;  /* keep both 2^-e and 2^e normal numbers */
    if (e > 1000) e = 1000;
    double s = u2d((uint64_t)(1023 - e) << 52);       /* 2^-e */
    double as = a * s, bs = b * s;
    double r = sqrt(as * as + bs * bs);
    return r * u2d((uint64_t)(1023 + e) << 52);       /* * 2^e */
}
static inline double c_abs(cpx a) { return fs_hypot(a.re, a.im); }

/* ------------------------------------------------------------------------ */
/* Xrange scalars: numba_xr.py:61-116 (record {mantissa, exp int32}). */
typedef struct { double m; int32_t e; } xrf;   /* float64 mantissa */
typedef struct { cpx m; int32_t e; } xrc;      /* complex128 mantissa */

static inline xrf XF(double m, int32_t e) { xrf r; r.m = m; r.e = e; return r; }
static inline xrc XC(cpx m, int32_t e) { xrc r; r.m = m; r.e = e; return r; }

/* _frexp, numba_xr.py:674-685: mantissa in [1, 2), biased exponent - 1023 */
static inline double xr_frexp(double m, int32_t *e) {
    uint64_t bits = d2u(m);
    *e = (int32_t)(((int64_t)(bits >> 52) & 0x7ff) - 0x3ff);
    return u2d((bits & 0x8000000000000000ULL) + (0x3ffULL << 52)
               + (bits & 0xfffffffffffffULL));
}

/* _exp2_shift, numba_xr.py:695-706: ldexp on the exponent field, clamped at
 * 0 (no proper subnormals) */
static inline double xr_exp2_shift(double m, int32_t shift) {
    uint64_t bits = d2u(m);
    int64_t ex = (int64_t)((bits >> 52) & 0x7ff) + (int64_t)shift;
    if (ex < 0) ex = 0;
    return u2d((bits & 0x8000000000000000ULL) + ((uint64_t)ex << 52)
               + (bits & 0xfffffffffffffULL));
}

/* _need_renorm, numba_xr.py:390-413: |unbiased exponent| > 100 */
static inline int need_renorm_d(double m) {
    int64_t be = (int64_t)((d2u(m) >> 52) & 0x7ff) - 1023;
    return (be < 0 ? -be : be) > 100;
}
static inline int need_renorm_c(cpx m) {
    return need_renorm_d(m.re) || need_renorm_d(m.im);
}

/* _normalize_real, numba_xr.py:687-693 */
static inline xrf normalize_real(double m, int32_t e) {
    if (m == 0.) return XF(m, 0);
    int32_t ne;
    double nm = xr_frexp(m, &ne);
    return XF(nm, e + ne);
}

/* _coexp_ufunc real, numba_xr.py:713-731 */
static inline void coexp_real(double m0, int32_t e0, double m1, int32_t e1,
                              double *o0, double *o1, int32_t *oe) {
    double c0 = m0, c1 = m1;
    int32_t d = e0 - e1;
    int32_t e;
    if (m0 == 0.) e = e1;
    else if (m1 == 0.) e = e0;
    else if (e1 > e0) { c0 = xr_exp2_shift(c0, d); e = e1; }
    else if (e0 > e1) { c1 = xr_exp2_shift(c1, -d); e = e0; }
    else e = e0;
    *o0 = c0; *o1 = c1; *oe = e;
}

/* _coexp_ufunc complex, numba_xr.py:733-753 (a real operand has im = 0) */
static inline void coexp_cplx(cpx m0, int32_t e0, cpx m1, int32_t e1,
                              cpx *o0, cpx *o1, int32_t *oe) {
    cpx c0 = m0, c1 = m1;
    int32_t d = e0 - e1;
    int32_t e;
    if (c_is_zero(m0)) e = e1;
    else if (c_is_zero(m1)) e = e0;
    else if (e1 > e0) {
        c0 = C(xr_exp2_shift(c0.re, d), xr_exp2_shift(c0.im, d)); e = e1;
    } else if (e0 > e1) {
        c1 = C(xr_exp2_shift(c1.re, -d), xr_exp2_shift(c1.im, -d)); e = e0;
    } else e = e0;
    *o0 = c0; *o1 = c1; *oe = e;
}

/* _normalize complex, numba_xr.py:650-664 */
static inline xrc normalize_cplx(cpx m, int32_t e) {
    xrf r = normalize_real(m.re, e);
    xrf i = normalize_real(m.im, e);
    double o0, o1; int32_t oe;
    coexp_real(r.m, r.e, i.m, i.e, &o0, &o1, &oe);
    return XC(C(o0, o1), oe);
}

/* to_Xrange_scalar of a standard number, numba_xr.py:784-799 */
static inline xrc to_xrc(cpx z) { return normalize_cplx(z, 0); }
static inline xrf to_xrf(double x) { return normalize_real(x, 0); }

/* (x + 1j * y) * s as numba evaluates it: each float operand is promoted to
 * a complex (v + 0j) and the complex product and sum are taken as written,
 * (a+bi)(c+di) = (ac - bd) + (ad + bc)i.  The result equals (x s, y s)
 * except where a component is zero: 1j * y contributes a 0 * y term, the
 * promotion adds +0 to the imaginary part, and the final product adds
 * -(im * 0) and re * 0, which fixes the signs of zero results. */
static inline cpx numba_cscale(double x, double y, double s) {
    double re1 = 0. * y - 1. * 0.;          /* (0 + 1j) * (y + 0j) */
    double im1 = 0. * 0. + 1. * y;
    double re2 = x + re1, im2 = 0. + im1;   /* (x + 0j) + ... */
    return C(re2 * s - im2 * 0., re2 * 0. + im2 * s);   /* * (s + 0j) */
}

/* to_standard, numba_xr.py:802-829 */
static inline double xrf_std(xrf a) { return ldexp(a.m, a.e); }
static inline cpx xrc_std(xrc a) {
    xrf r = normalize_real(a.m.re, a.e);
    xrf i = normalize_real(a.m.im, a.e);
    double o0, o1; int32_t oe;
    coexp_real(r.m, r.e, i.m, i.e, &o0, &o1, &oe);
    return numba_cscale(o0, o1, ldexp(1., oe));
}

/* add / sub: numba_xr.py:318-388 */
static inline xrc xrc_add(xrc a, xrc b) {
    cpx o0, o1; int32_t oe;
    coexp_cplx(a.m, a.e, b.m, b.e, &o0, &o1, &oe);
    return XC(c_add(o0, o1), oe);
}
static inline xrc xrc_sub(xrc a, xrc b) {
    cpx o0, o1; int32_t oe;
    coexp_cplx(a.m, a.e, b.m, b.e, &o0, &o1, &oe);
    return XC(c_sub(o0, o1), oe);
}
/* Xrange + standard complex (numba_xr.py:329-337, 341-348) */
static inline xrc std_as_xr_operand(cpx b) {
    if (need_renorm_c(b)) return normalize_cplx(b, 0);
    return XC(b, 0);
}
static inline xrc xrc_add_std(xrc a, cpx b) { return xrc_add(a, std_as_xr_operand(b)); }
static inline xrc xrc_add_xrf(xrc a, xrf b) { return xrc_add(a, XC(C(b.m, 0.), b.e)); }

static inline xrf xrf_add(xrf a, xrf b) {
    double o0, o1; int32_t oe;
    coexp_real(a.m, a.e, b.m, b.e, &o0, &o1, &oe);
    return XF(o0 + o1, oe);
}
static inline xrf xrf_sub(xrf a, xrf b) {
    double o0, o1; int32_t oe;
    coexp_real(a.m, a.e, b.m, b.e, &o0, &o1, &oe);
    return XF(o0 - o1, oe);
}
static inline xrf std_as_xrf_operand(double b) {
    if (need_renorm_d(b)) return normalize_real(b, 0);
    return XF(b, 0);
}
static inline xrf xrf_add_std(xrf a, double b) { return xrf_add(a, std_as_xrf_operand(b)); }
static inline xrf xrf_neg(xrf a) { return XF(-a.m, a.e); }
static inline xrf xrf_abs(xrf a) { return XF(fabs(a.m), a.e); }

/* mul: numba_xr.py:415-448 */
static inline xrc xrc_mul(xrc a, xrc b) {
    cpx m = c_mul(a.m, b.m);
    int32_t e = a.e + b.e;
    if (need_renorm_c(m)) return normalize_cplx(m, e);
    return XC(m, e);
}
static inline xrc xrc_mul_std(xrc a, cpx b) {     /* xr * complex (and complex * xr) */
    cpx m = c_mul(a.m, b);
    if (need_renorm_c(m)) return normalize_cplx(m, a.e);
    return XC(m, a.e);
}
static inline xrc xrc_scale(double s, xrc a) {    /* float * xr */
    cpx m = c_scale(s, a.m);
    if (need_renorm_c(m)) return normalize_cplx(m, a.e);
    return XC(m, a.e);
}
static inline xrf xrf_mul(xrf a, xrf b) {
    double m = a.m * b.m;
    int32_t e = a.e + b.e;
    if (need_renorm_d(m)) return normalize_real(m, e);
    return XF(m, e);
}
static inline xrf xrf_scale(double s, xrf a) {
    double m = s * a.m;
    if (need_renorm_d(m)) return normalize_real(m, a.e);
    return XF(m, a.e);
}

/* extended_abs2, numba_xr.py:531-561 */
static inline xrf xrc_abs2(xrc a) {
    return XF(a.m.re * a.m.re + a.m.im * a.m.im, a.e + a.e);
}

/* comparisons, numba_xr.py:476-510 */
static inline int xrf_lt_std(xrf a, double b) {
    double o0, o1; int32_t oe;
    coexp_real(a.m, a.e, b, 0, &o0, &o1, &oe);
    return o0 < o1;
}
static inline int xrf_ge_std(xrf a, double b) {
    double o0, o1; int32_t oe;
    coexp_real(a.m, a.e, b, 0, &o0, &o1, &oe);
    return o0 >= o1;
}
static inline int xrf_le_std(xrf a, double b) {
    double o0, o1; int32_t oe;
    coexp_real(a.m, a.e, b, 0, &o0, &o1, &oe);
    return o0 <= o1;
}
static inline int xrf_le(xrf a, xrf b) {
    double o0, o1; int32_t oe;
    coexp_real(a.m, a.e, b.m, b.e, &o0, &o1, &oe);
    return o0 <= o1;
}

/* ------------------------------------------------------------------------ */
/* The oracle context: everything get_cycle_indep_args (perturbation.py:431-562)
 * hands to the tile loop, plus the calc_std_div flags baked into iterate. */
enum { OR_KIND_M2 = 0, OR_KIND_MN = 1, OR_KIND_BS = 2 };

typedef struct {
    int32_t kind, degree, flavor, xr;
    int64_t L, div_iter, order;
    const double *Zn;            /* [2L] interleaved complex128 */
    int32_t n_xr;
    const int32_t *xr_idx;       /* [n_xr] sorted */
    const double *xr_m;          /* holo: [2 n_xr] complex; BS: [2 n_xr] (x, y) */
    const int32_t *xr_e;         /* holo: [n_xr]; BS: [2 n_xr] (x, y) */
    double dx_m; int32_t dx_e;   /* dx_xr == lin_scale_xr */
    double sd_m; int32_t sd_e;   /* scale_deriv_xr = dx_xr * projection.scale (perturbation.py:459-467) */
    double drift_m[2]; int32_t drift_e[2];
    double lin_mat[4];
    double kc_m; int32_t kc_e;
    double bla_eps; int32_t bla_on;
    /* calc params */
    int64_t max_iter;
    double M2;                   /* M_divergence ** 2 */
    double eps2;                 /* epsilon_stationnary ** 2 */
    int32_t calc_dzndc, calc_dzndz;
    int32_t calc_orbit; int64_t backshift;   /* orbit point at exit - backshift */
    /* derived (filled by or_ctx_prepare) */
    double *dzndc_m; int32_t *dzndc_e;     /* holo: [2L] (+ [L] exps) */
    double *dzndz_m; int32_t *dzndz_e;     /* holo: [2(L+1)] (+ [L+1]) */
    double *dbs_m[4]; int32_t *dbs_e[4];   /* BS: [L] each */
    double *M_bla; double *r_bla; int64_t bla_len; int32_t stages_bla;
} or_ctx;

/* ------------------------------------------------------------------------ */
/* model kernels */
static inline double C_binom(int n, int k) {
    double r = 1.;  /* exact for the degrees used (n <= 16) */
    for (int i = 1; i <= k; i++) r = r * (double)(n - k + i) / (double)i;
    return r;
}

/* dfdz: mandelbrot_M2.py:599-602, mandelbrot_Mn.py:646-652 */
static inline cpx dfdz_std(const or_ctx *k, cpx z) {
    if (k->kind == OR_KIND_M2) return c_scale(2., z);
    cpx tmp = z;
    for (int i = 2; i < k->degree; i++) tmp = c_mul(tmp, z);
    return c_scale((double)k->degree, tmp);
}
static inline xrc dfdz_xr(const or_ctx *k, xrc z) {
    if (k->kind == OR_KIND_M2) return xrc_scale(2., z);
    xrc tmp = z;
    for (int i = 2; i < k->degree; i++) tmp = xrc_mul(tmp, z);
    return xrc_scale((double)k->degree, tmp);
}

/* BS Jacobian entries, burning_ship.py:440-532 */
static inline double sgn(double x) { return (x < 0.) ? -1. : 1.; }
static void bs_jac(int flavor, double x, double y, double J[4]) {
    double x2my2 = x * x - y * y;
    switch (flavor) {
    case 1: J[0] = 2. * x; J[1] = -2. * y; J[2] = 2. * sgn(x) * fabs(y); J[3] = 2. * sgn(y) * fabs(x); break;
    case 2: J[0] = 2. * x; J[1] = -2. * y; J[2] = 2. * fabs(y); J[3] = 2. * sgn(y) * x; break;
    case 3: J[0] = 2. * x; J[1] = -2. * fabs(y); J[2] = 2. * y; J[3] = 2. * x; break;
    case 4: J[0] = 2. * sgn(x2my2) * x; J[1] = -2. * sgn(x2my2) * y; J[2] = 2. * y; J[3] = 2. * x; break;
    default: J[0] = 2. * sgn(x2my2) * x; J[1] = -2. * sgn(x2my2) * y; J[2] = 2. * sgn(x) * fabs(y); J[3] = 2. * sgn(y) * fabs(x); break;
    }
}
/* the same on Xrange operands (incr_deriv_ref_BS, perturbation.py:2441-2463,
 * calls the model functions with Xrange records in Xrange mode) */
static inline xrf xrf_sgn(xrf x) { return XF(xrf_lt_std(x, 0.) ? -1. : 1., 0); }
static void bs_jac_xr(int flavor, xrf x, xrf y, xrf J[4]) {
    xrf x2my2 = xrf_sub(xrf_mul(x, x), xrf_mul(y, y));
    /* sgn() returns a float; products are float * xr */
    double sx = xrf_lt_std(x, 0.) ? -1. : 1.;
    double sy = xrf_lt_std(y, 0.) ? -1. : 1.;
    double sq = xrf_lt_std(x2my2, 0.) ? -1. : 1.;
    switch (flavor) {
    case 1: J[0] = xrf_scale(2., x); J[1] = xrf_scale(-2., y);
            J[2] = xrf_scale(2. * sx, xrf_abs(y)); J[3] = xrf_scale(2. * sy, xrf_abs(x)); break;
    case 2: J[0] = xrf_scale(2., x); J[1] = xrf_scale(-2., y);
            J[2] = xrf_scale(2., xrf_abs(y)); J[3] = xrf_scale(2. * sy, x); break;
    case 3: J[0] = xrf_scale(2., x); J[1] = xrf_scale(-2., xrf_abs(y));
            J[2] = xrf_scale(2., y); J[3] = xrf_scale(2., x); break;
    case 4: J[0] = xrf_scale(2. * sq, x); J[1] = xrf_scale(-2. * sq, y);
            J[2] = xrf_scale(2., y); J[3] = xrf_scale(2., x); break;
    default: J[0] = xrf_scale(2. * sq, x); J[1] = xrf_scale(-2. * sq, y);
            J[2] = xrf_scale(2. * sx, xrf_abs(y)); J[3] = xrf_scale(2. * sy, xrf_abs(x)); break;
    }
}

/* ------------------------------------------------------------------------ */
/* ref_path_get, perturbation.py:2519-2588: the cursor walk over the sorted
 * Xrange register is a membership lookup. */
static inline int xr_lookup(const or_ctx *k, int64_t idx) {
    int32_t lo = 0, hi = k->n_xr;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if ((int64_t)k->xr_idx[mid] < idx) lo = mid + 1; else hi = mid;
    }
    if (lo < k->n_xr && (int64_t)k->xr_idx[lo] == idx) return lo;
    return -1;
}
static inline cpx Zref(const or_ctx *k, int64_t i) { return C(k->Zn[2 * i], k->Zn[2 * i + 1]); }

/* ensure_xr (perturbation.py:2182-2193) of the value ref_path_get returns */
static inline xrc ref_xr(const or_ctx *k, int64_t i, int *is_xr) {
    int j = (k->n_xr > 0) ? xr_lookup(k, i) : -1;
    *is_xr = (j >= 0);
    if (j >= 0) return XC(C(k->xr_m[2 * j], k->xr_m[2 * j + 1]), k->xr_e[j]);
    return to_xrc(Zref(k, i));
}

/* ------------------------------------------------------------------------ */
/* numba_dZndc_path / numba_dZndz_path, perturbation.py:2282-2336, 2466-2516 */
static int holo_paths(or_ctx *k) {
    int64_t L = k->L;
    int64_t valid = L < k->div_iter ? L : k->div_iter;
    xrf scale = XF(k->sd_m, k->sd_e);
    double scale_std = xrf_std(scale);
    int dummy;
    if (k->calc_dzndc) {
        k->dzndc_m = (double *)calloc(2 * L, sizeof(double));
        k->dzndc_e = (int32_t *)calloc(L, sizeof(int32_t));
        if (valid < 2) return -1;
        int64_t i;
        if (!k->xr) {
            cpx *D = (cpx *)k->dzndc_m;
            for (i = 1; i < valid; i++)
                D[i] = c_add(c_mul(dfdz_std(k, Zref(k, i - 1)), D[i - 1]), C(scale_std, 0.));
            i = valid - 1;
            if (i == k->order - 1)
                D[0] = c_add(c_mul(dfdz_std(k, Zref(k, i)), D[i]), C(scale_std, 0.));
        } else {
            xrc *D = (xrc *)malloc(L * sizeof(xrc));
            D[0] = XC(C(0., 0.), 0);
            for (i = 1; i < valid; i++) {
                xrc r = ref_xr(k, i - 1, &dummy);
                D[i] = xrc_add_xrf(xrc_mul(dfdz_xr(k, r), D[i - 1]), scale);
            }
            for (i = valid; i < L; i++) D[i] = XC(C(0., 0.), 0);
            i = valid - 1;
            if (i == k->order - 1)  /* standard dfdz(Zn_path[i]) here */
                D[0] = xrc_add_xrf(xrc_mul_std(D[i], dfdz_std(k, Zref(k, i))), scale);
            for (i = 0; i < L; i++) {
                k->dzndc_m[2 * i] = D[i].m.re; k->dzndc_m[2 * i + 1] = D[i].m.im;
                k->dzndc_e[i] = D[i].e;
            }
            free(D);
        }
    }
    if (k->calc_dzndz) {
        k->dzndz_m = (double *)calloc(2 * (L + 1), sizeof(double));
        k->dzndz_e = (int32_t *)calloc(L + 1, sizeof(int32_t));
        if (valid < 3) return -1;
        int64_t i;
        if (!k->xr) {
            cpx *D = (cpx *)k->dzndz_m;
            D[1] = C(1., 0.);
            for (i = 2; i < valid; i++)
                D[i] = c_mul(dfdz_std(k, Zref(k, i - 1)), D[i - 1]);
            i = valid - 1;
            D[L] = c_mul(dfdz_std(k, Zref(k, i)), D[i]);
        } else {
            xrc *D = (xrc *)malloc((L + 1) * sizeof(xrc));
            for (i = 0; i <= L; i++) D[i] = XC(C(0., 0.), 0);
            D[1] = XC(C(1., 0.), 0);
            for (i = 2; i < valid; i++) {
                xrc r = ref_xr(k, i - 1, &dummy);
                D[i] = xrc_mul(dfdz_xr(k, r), D[i - 1]);
            }
            i = valid - 1;
            xrc r = ref_xr(k, i, &dummy);
            D[L] = xrc_mul(dfdz_xr(k, r), D[i]);
            for (i = 0; i <= L; i++) {
                k->dzndz_m[2 * i] = D[i].m.re; k->dzndz_m[2 * i + 1] = D[i].m.im;
                k->dzndz_e[i] = D[i].e;
            }
            free(D);
        }
    }
    return 0;
}

/* numba_dZndc_path_BS + incr_deriv_ref_BS, perturbation.py:2339-2463 */
static int bs_paths(or_ctx *k) {
    int64_t L = k->L;
    int64_t valid = L < k->div_iter ? L : k->div_iter;
    xrf scale = XF(k->sd_m, k->sd_e);
    double s = xrf_std(scale);
    for (int q = 0; q < 4; q++) {
        k->dbs_m[q] = (double *)calloc(L, sizeof(double));
        k->dbs_e[q] = (int32_t *)calloc(L, sizeof(int32_t));
    }
    if (!k->calc_dzndc) return 0;
    if (valid < 2) return -1;
    int64_t i;
    if (!k->xr) {
        double *a = k->dbs_m[0], *b = k->dbs_m[1], *c = k->dbs_m[2], *d = k->dbs_m[3];
        int64_t from_i, to_i;
        for (i = 1; i <= valid; i++) {
            if (i < valid) { from_i = i - 1; to_i = i; }
            else { i = valid - 1; if (i != k->order - 1) break; from_i = i; to_i = 0; }
            double J[4];
            bs_jac(k->flavor, Zref(k, from_i).re, Zref(k, from_i).im, J);
            double xa = a[from_i], xb = b[from_i], ya = c[from_i], yb = d[from_i];
            a[to_i] = J[0] * xa + J[1] * ya + s;
            b[to_i] = J[0] * xb + J[1] * yb;
            c[to_i] = J[2] * xa + J[3] * ya;
            d[to_i] = J[2] * xb + J[3] * yb - s;
            if (to_i == 0) break;
        }
    } else {
        xrf *P[4];
        for (int q = 0; q < 4; q++) {
            P[q] = (xrf *)malloc(L * sizeof(xrf));
            for (i = 0; i < L; i++) P[q][i] = XF(0., 0);
        }
        int64_t from_i, to_i;
        for (i = 1; i <= valid; i++) {
            if (i < valid) { from_i = i - 1; to_i = i; }
            else { i = valid - 1; if (i != k->order - 1) break; from_i = i; to_i = 0; }
            /* ref_path_get_BS + ensure_xr_BS */
            int j = (k->n_xr > 0) ? xr_lookup(k, from_i) : -1;
            xrf X, Y;
            if (j >= 0) { X = XF(k->xr_m[2 * j], k->xr_e[2 * j]); Y = XF(k->xr_m[2 * j + 1], k->xr_e[2 * j + 1]); }
            else { X = to_xrf(Zref(k, from_i).re); Y = to_xrf(Zref(k, from_i).im); }
            xrf J[4];
            bs_jac_xr(k->flavor, X, Y, J);
            xrf xa = P[0][from_i], xb = P[1][from_i], ya = P[2][from_i], yb = P[3][from_i];
            P[0][to_i] = xrf_add(xrf_add(xrf_mul(J[0], xa), xrf_mul(J[1], ya)), scale);
            P[1][to_i] = xrf_add(xrf_mul(J[0], xb), xrf_mul(J[1], yb));
            P[2][to_i] = xrf_add(xrf_mul(J[2], xa), xrf_mul(J[3], ya));
            P[3][to_i] = xrf_sub(xrf_add(xrf_mul(J[2], xb), xrf_mul(J[3], yb)), scale);
            if (to_i == 0) break;
        }
        for (int q = 0; q < 4; q++) {
            for (i = 0; i < L; i++) { k->dbs_m[q][i] = P[q][i].m; k->dbs_e[q][i] = P[q][i].e; }
            free(P[q]);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* BLA tree: numba_make_BLA / combine_BLA / BLA_index,
 * perturbation.py:1819-1881, 1976-2024, 2108-2114.  M holds 2 complex per
 * entry (A, B) for holomorphic models, 8 float64 for BS. */
#define STG_COMPRESSED 3

static inline int64_t bla_index(int64_t i, int stg) { return 2 * i + ((1LL << stg) - 1); }

static void combine_holo(cpx *M, double *r, double kc, int stg, int64_t len, double eps) {
    int64_t step = 1LL << stg;
    for (int64_t i = 0; i < len - step + 1; i += step) {
        int64_t ii = i + step / 2;
        if (ii >= len) break;
        int64_t i1 = bla_index(i, stg - 1), i2 = bla_index(ii, stg - 1), ir = bla_index(i, stg);
        M[2 * ir] = c_mul(M[2 * i2], M[2 * i1]);
        M[2 * ir + 1] = c_add(c_mul(M[2 * i2], M[2 * i1 + 1]), M[2 * i2 + 1]);
        double r1 = r[i1], r2 = r[i2];
        double mA1 = c_abs(M[2 * i1]);
        double mB1 = c_abs(M[2 * i1 + 1]);
        double num = r2 - mB1 * kc;
        double den = mA1 > eps ? mA1 : eps;
        double q = num / den;
        double r2b = 0.95 * (q > 0. ? q : 0.);
        r[ir] = (r2b < r1) ? r2b : r1;   /* python min(r1, r2b) */
    }
}

static void combine_bs(double *M, double *r, double kc, int stg, int64_t len, double eps) {
    int64_t step = 1LL << stg;
    for (int64_t i = 0; i < len - step + 1; i += step) {
        int64_t ii = i + step / 2;
        if (ii >= len) break;
        int64_t i1 = bla_index(i, stg - 1), i2 = bla_index(ii, stg - 1), ir = bla_index(i, stg);
        const double *a = M + 8 * i1, *b = M + 8 * i2;
        double *o = M + 8 * ir;
        double t[8];
        t[0] = b[0] * a[0] + b[1] * a[2];
        t[1] = b[0] * a[1] + b[1] * a[3];
        t[2] = b[2] * a[0] + b[3] * a[2];
        t[3] = b[2] * a[1] + b[3] * a[3];
        t[4] = b[0] * a[4] + b[1] * a[6] + b[4];
        t[5] = b[0] * a[5] + b[1] * a[7] + b[5];
        t[6] = b[2] * a[4] + b[3] * a[6] + b[6];
        t[7] = b[2] * a[5] + b[3] * a[7] + b[7];
        double r1 = r[i1], r2 = r[i2];
        double mA1 = fabs(a[0]);
        if (fabs(a[1]) > mA1) mA1 = fabs(a[1]);
        if (fabs(a[2]) > mA1) mA1 = fabs(a[2]);
        if (fabs(a[3]) > mA1) mA1 = fabs(a[3]);
        double mB1 = fabs(a[4]);
        if (fabs(a[5]) > mB1) mB1 = fabs(a[5]);
        if (fabs(a[6]) > mB1) mB1 = fabs(a[6]);
        if (fabs(a[7]) > mB1) mB1 = fabs(a[7]);
        for (int q = 0; q < 8; q++) o[q] = t[q];
        double num = r2 - mB1 * kc;
        double den = mA1 > eps ? mA1 : eps;
        double qq = num / den;
        double r2b = 0.95 * (qq > 0. ? qq : 0.);
        r[ir] = (r2b < r1) ? r2b : r1;   /* python min(r1, r2b) */
    }
}

static int stages_bla(int64_t L) {   /* int(ceil(log2(L))), perturbation.py:1976-1981 */
    int s = 0;
    while ((1LL << s) < L) s++;
    return s;
}

static int make_bla(or_ctx *k) {
    int64_t L = k->L;
    double kc = xrf_std(XF(k->kc_m, k->kc_e));
    double eps = k->bla_eps;
    int64_t comp = L >> STG_COMPRESSED;
    int dim = (k->kind == OR_KIND_BS) ? 8 : 4;   /* doubles per entry */
    k->bla_len = 2 * comp;
    k->stages_bla = stages_bla(L);
    k->M_bla = (double *)calloc((size_t)(2 * comp * dim + 8), sizeof(double));
    k->r_bla = (double *)calloc((size_t)(2 * comp + 1), sizeof(double));
    double tM[16 * 8]; double tr[16];
    for (int64_t i = 0; i < comp; i++) {
        memset(tM, 0, sizeof(tM)); memset(tr, 0, sizeof(tr));
        for (int j = 0; j < 8; j++) {
            cpx Z = Zref(k, i * 8 + j);
            if (dim == 4) {
                cpx A = dfdz_std(k, Z);
                ((cpx *)tM)[2 * (2 * j)] = A;
                ((cpx *)tM)[2 * (2 * j) + 1] = C(1., 0.);
                tr[2 * j] = eps * c_abs(A);
            } else {
                double J[4];
                bs_jac(k->flavor, Z.re, Z.im, J);
                double *o = tM + 8 * (2 * j);
                o[0] = J[0]; o[1] = J[1]; o[2] = J[2]; o[3] = J[3];
                o[4] = 1.; o[5] = 0.; o[6] = 0.; o[7] = -1.;
                double ax = fabs(Z.re), ay = fabs(Z.im);
                tr[2 * j] = eps * (ax < ay ? ax : ay);
            }
        }
        for (int stg = 1; stg <= STG_COMPRESSED; stg++) {
            if (dim == 4) combine_holo((cpx *)tM, tr, kc, stg, 8, eps);
            else combine_bs(tM, tr, kc, stg, 8, eps);
        }
        memcpy(k->M_bla + (2 * i) * dim, tM + 7 * dim, dim * sizeof(double));
        k->r_bla[2 * i] = tr[7];
    }
    for (int stg = 1; stg < k->stages_bla - STG_COMPRESSED; stg++) {
        if (dim == 4) combine_holo((cpx *)k->M_bla, k->r_bla, kc, stg, comp, eps);
        else combine_bs(k->M_bla, k->r_bla, kc, stg, comp, eps);
    }
    return 0;
}

/* ref_BLA_get, perturbation.py:2116-2170: the first stage from the top with
 * |z| < r.  |z| = az * 2^ez: the binary64 loops pass hypot(z) and ez = 0;
 * the Xrange loop passes the modulus of its mantissa and its exponent, and
 * the radius is scaled instead (exact), so that |z| keeps its precision where
 * the binary64 copy of z (the reference's Z[zn]) is subnormal or zero.  With
 * a normal copy the two give the same decision. */
static inline int64_t bla_get_scaled(const or_ctx *k, double az, int32_t ez, int64_t n,
                                     int64_t first_invalid, const double **Mout) {
    int64_t it = n >> STG_COMPRESSED;
    int stages;
    for (stages = STG_COMPRESSED; stages < k->stages_bla; stages++) {
        if (it & 1) break;
        it >>= 1;
    }
    if (stages == k->stages_bla) stages = k->stages_bla - 1;
    int64_t invalid_step = first_invalid - n;
    int dim = (k->kind == OR_KIND_BS) ? 8 : 4;
    for (int stg = stages; stg >= STG_COMPRESSED; stg--) {
        int64_t step = 1LL << stg;
        if (step >= invalid_step) continue;
        int64_t idx = bla_index(n >> STG_COMPRESSED, stg - STG_COMPRESSED);
        if (az < ldexp(k->r_bla[idx], -ez)) {
            *Mout = k->M_bla + idx * dim;
            return step;
        }
    }
    return 0;
}
static inline int64_t bla_get(const or_ctx *k, double zre, double zim, int64_t n,
                              int64_t first_invalid, const double **Mout) {
    return bla_get_scaled(k, fs_hypot(zre, zim), 0, n, first_invalid, Mout);
}

/* ------------------------------------------------------------------------ */
/* per-iteration model kernels, holomorphic (mandelbrot_M2.py:607-622,
 * mandelbrot_Mn.py:659-729) -- standard precision */
static inline cpx p_zn_std(const or_ctx *k, cpx z, cpx Z, cpx c) {
    if (k->kind == OR_KIND_M2)
        return c_add(c_mul(z, c_add(z, c_scale(2., Z))), c);
    int n = k->degree;
    cpx tmp = c_mul(z, c_add(z, c_scale(C_binom(n, 1), Z)));
    cpx Zpk = Z;
    for (int q = 2; q < n; q++) {
        Zpk = c_mul(Zpk, Z);
        tmp = c_mul(z, c_add(tmp, c_scale(C_binom(n, q), Zpk)));
    }
    return c_add(tmp, c);
}
/* dzndc and dzndz share the formula (with the matching orbit path) */
static inline cpx p_deriv_std(const or_ctx *k, cpx z, cpx d, cpx Z, cpx Zd) {
    if (k->kind == OR_KIND_M2)
        return c_scale(2., c_add(c_mul(c_add(Z, z), d), c_mul(Zd, z)));
    int n = k->degree;
    cpx mul = c_add(z, c_scale(C_binom(n, 1), Z));
    cpx tmp = c_mul(z, mul);
    cpx dt = c_add(c_mul(d, mul), c_mul(z, c_add(d, c_scale(C_binom(n, 1), Zd))));
    cpx Zpk = Z, dZpk;
    for (int q = 2; q < n; q++) {
        dZpk = c_mul(c_scale((double)q, Zpk), Zd);
        Zpk = c_mul(Zpk, Z);
        mul = c_add(tmp, c_scale(C_binom(n, q), Zpk));
        dt = c_add(c_mul(d, mul), c_mul(z, c_add(dt, c_scale(C_binom(n, q), dZpk))));
        tmp = c_mul(z, mul);
    }
    return dt;
}
/* the same, Xrange operands */
static inline xrc p_zn_xr(const or_ctx *k, xrc z, xrc Z, xrc c) {
    if (k->kind == OR_KIND_M2)
        return xrc_add(xrc_mul(z, xrc_add(z, xrc_scale(2., Z))), c);
    int n = k->degree;
    xrc tmp = xrc_mul(z, xrc_add(z, xrc_scale(C_binom(n, 1), Z)));
    xrc Zpk = Z;
    for (int q = 2; q < n; q++) {
        Zpk = xrc_mul(Zpk, Z);
        tmp = xrc_mul(z, xrc_add(tmp, xrc_scale(C_binom(n, q), Zpk)));
    }
    return xrc_add(tmp, c);
}
static inline xrc p_deriv_xr(const or_ctx *k, xrc z, xrc d, xrc Z, xrc Zd) {
    if (k->kind == OR_KIND_M2)
        return xrc_scale(2., xrc_add(xrc_mul(xrc_add(Z, z), d), xrc_mul(Zd, z)));
    int n = k->degree;
    xrc mul = xrc_add(z, xrc_scale(C_binom(n, 1), Z));
    xrc tmp = xrc_mul(z, mul);
    xrc dt = xrc_add(xrc_mul(d, mul), xrc_mul(z, xrc_add(d, xrc_scale(C_binom(n, 1), Zd))));
    xrc Zpk = Z, dZpk;
    for (int q = 2; q < n; q++) {
        dZpk = xrc_mul(xrc_scale((double)q, Zpk), Zd);
        Zpk = xrc_mul(Zpk, Z);
        mul = xrc_add(tmp, xrc_scale(C_binom(n, q), Zpk));
        dt = xrc_add(xrc_mul(d, mul), xrc_mul(z, xrc_add(dt, xrc_scale(C_binom(n, q), dZpk))));
        tmp = xrc_mul(z, mul);
    }
    return dt;
}

/* ------------------------------------------------------------------------ */
/* ref_path_c_from_pix + lin_proj_impl, perturbation.py:2214-2230, 2665-2672 */
static inline xrc c_from_pix_holo(const or_ctx *k, double px, double py) {
    double x1 = k->lin_mat[0] * px + k->lin_mat[1] * py;
    double y1 = k->lin_mat[2] * px + k->lin_mat[3] * py;
    cpx m = C(k->dx_m * x1, k->dx_m * y1);
    xrc cx = need_renorm_c(m) ? normalize_cplx(m, k->dx_e) : XC(m, k->dx_e);
    return xrc_add(cx, XC(C(k->drift_m[0], k->drift_m[1]), k->drift_e[0]));
}

/* numba_iterate(...).numba_impl, perturbation.py:1065-1400 (holomorphic).
 * Z fields: zn, [dzndz], [dzndc] (mandelbrot_M2.py:482-506). */
/* calc_orbit: perturbation.py:1096-1101 (init), 1266-1273 (record at the
 * escape test, after plain steps only), 1390-1396 (finish: iterate the
 * full recurrence in binary64 from the last-but-one record to
 * n - backshift, with c + Zn_path[1]). */
typedef struct { int64_t div_shift, i1, i2; cpx z1, z2; } orbit_rec;

static inline void orbit_record(orbit_rec *o, int64_t n, int64_t backshift, cpx ZZ) {
    int64_t div = n / backshift;
    if (div > o->div_shift) {
        o->div_shift = div;
        o->i2 = o->i1; o->z2 = o->z1;
        o->i1 = n; o->z1 = ZZ;
    }
}

/* zn ** d + c for the orbit finish: M2_iterate (mandelbrot_M2.py:14-15);
 * Mn_iterate (mandelbrot_Mn.py:13-17) as binary powering */
static cpx zn_iterate(const or_ctx *k, cpx z, cpx c) {
    if (k->kind == OR_KIND_M2) return c_add(c_mul(z, z), c);
    cpx r = C(1., 0.), p = z;
    for (int64_t m = k->degree; m > 0; m >>= 1) {
        if (m & 1) r = c_mul(r, p);
        if (m > 1) p = c_mul(p, p);
    }
    return c_add(r, c);
}

static cpx orbit_finish(const or_ctx *k, orbit_rec *o, int64_t n, cpx c) {
    cpx zo = o->z2;
    const cpx CC = c_add(c, Zref(k, 1));
    while (o->i2 < n - k->backshift) { zo = zn_iterate(k, zo, CC); o->i2++; }
    return zo;
}

static int64_t iterate_holo(const or_ctx *k, xrc c_xr, cpx *out_z, cpx *out_dzndz,
                            cpx *out_dzndc, cpx *out_orbit, int32_t *U, int8_t *stop) {
    orbit_rec orb = {0, 0, 0, {0., 0.}, {0., 0.}};
    const int64_t L = k->L, order = k->order, div_iter = k->div_iter;
    const int dc = k->calc_dzndc, dz = k->calc_dzndz;
    int64_t w = 0, n = 0;
    int64_t first_invalid = L;
    if (div_iter < first_invalid) first_invalid = div_iter;
    if (order < first_invalid) first_invalid = order;
    int nullify = 0;
    const int64_t w_wraped = L;
    int dyn = 1;
    const double *Mb;
    int64_t step;

    if (!k->xr) {
        cpx c = xrc_std(c_xr);
        const cpx *D = (const cpx *)k->dzndc_m, *Dz = (const cpx *)k->dzndz_m;
        cpx z = C(0., 0.), d = C(0., 0.), h = C(0., 0.);
        for (;;) {
            if (k->bla_on && (w & 7) == 0) {
                step = bla_get(k, z.re, z.im, w, first_invalid, &Mb);
                if (step != 0) {
                    cpx A = C(Mb[0], Mb[1]), B = C(Mb[2], Mb[3]);
                    n += step;
                    w = (w + step) % order;
                    z = c_add(c_mul(A, z), c_mul(B, c));
                    if (dc) d = c_mul(A, d);
                    if (dz) h = c_mul(A, h);
                    continue;
                }
            }
            n += 1;
            cpx Zr = Zref(k, w);
            cpx nd = d, nh = h;
            if (dc) nd = p_deriv_std(k, z, d, Zr, dyn ? C(0., 0.) : D[w]);
            if (dz) nh = p_deriv_std(k, z, h, Zr, nullify ? Dz[0] : Dz[w]);
            d = nd; h = nh;
            z = p_zn_std(k, z, Zr, c);
            w += 1;
            if (w >= order) w = w % order;
            if (n >= k->max_iter) { *stop = 0; break; }
            if (dz) {
                cpx nx = nullify ? Dz[0] : (n == order ? Dz[w_wraped] : Dz[w]);
                cpx ZdZ = c_add(h, nx);
                if (ZdZ.re * ZdZ.re + ZdZ.im * ZdZ.im < k->eps2) { *stop = 2; break; }
            }
            cpx ZZ = c_add(z, Zref(k, w));
            double fsq = ZZ.re * ZZ.re + ZZ.im * ZZ.im;
            if (k->calc_orbit) orbit_record(&orb, n, k->backshift, ZZ);
            if (fsq > k->M2) { *stop = 1; break; }
            if (w >= div_iter - 1) {
                z = ZZ;
                if (dc) d = c_add(d, D[w]);
                if (dz) {
                    if (!nullify) h = c_add(h, n == order ? Dz[w_wraped] : Dz[w]);
                    nullify = 1;
                }
                w = 0;
                continue;
            }
            dyn = (fabs(ZZ.re) <= fabs(z.re)) && (fabs(ZZ.im) <= fabs(z.im));
            if (dyn) {
                z = ZZ;
                if (dc) d = c_add(d, D[w]);
                if (dz) {
                    if (!nullify) h = c_add(h, n == order ? Dz[w_wraped] : Dz[w]);
                    nullify = 1;
                }
                w = 0;
                continue;
            }
        }
        *U = (int32_t)w;
        *out_z = c_add(z, Zref(k, w));
        if (dc) *out_dzndc = c_add(d, D[w]);
        if (dz) *out_dzndz = h;
        if (k->calc_orbit) *out_orbit = orbit_finish(k, &orb, n, c);
        return n;
    }

    /* Xrange branch */
#define DXR(i) XC(C(k->dzndc_m[2 * (i)], k->dzndc_m[2 * (i) + 1]), k->dzndc_e[i])
#define DZXR(i) XC(C(k->dzndz_m[2 * (i)], k->dzndz_m[2 * (i) + 1]), k->dzndz_e[i])
    const xrc xzero = XC(C(0., 0.), 0);
    xrc zx = xzero, dx = xzero, hx = xzero;
    cpx zs = C(0., 0.);          /* Z[zn]: standard copy */
    int out_is_xr = 0;
    xrc out_xr = xzero;
    for (;;) {
        if (k->bla_on && (w & 7) == 0) {
            step = bla_get_scaled(k, sqrt(zx.m.re * zx.m.re + zx.m.im * zx.m.im), zx.e, w,
                                  first_invalid, &Mb);
            if (step != 0) {
                cpx A = C(Mb[0], Mb[1]), B = C(Mb[2], Mb[3]);
                n += step;
                w = (w + step) % order;
                zx = xrc_add(xrc_mul_std(zx, A), xrc_mul_std(c_xr, B));
                zs = xrc_std(zx);
                if (dc) dx = xrc_mul_std(dx, A);
                if (dz) hx = xrc_mul_std(hx, A);
                continue;
            }
        }
        n += 1;
        int isx;
        xrc Zr = ref_xr(k, w, &isx);
        if (isx) { out_is_xr = 1; out_xr = Zr; } else out_is_xr = 0;
        xrc ndx = dx, nhx = hx;
        if (dc) ndx = p_deriv_xr(k, zx, dx, Zr, dyn ? xzero : DXR(w));
        if (dz) nhx = p_deriv_xr(k, zx, hx, Zr, nullify ? DZXR(0) : DZXR(w));
        dx = ndx; hx = nhx;
        zx = p_zn_xr(k, zx, Zr, c_xr);
        zs = xrc_std(zx);
        w += 1;
        if (w >= order) w = w % order;
        if (n >= k->max_iter) { *stop = 0; break; }
        if (dz) {
            xrc nx = nullify ? DZXR(0) : (n == order ? DZXR(w_wraped) : DZXR(w));
            xrc ZdZ = xrc_add(hx, nx);
            if (xrf_lt_std(xrc_abs2(ZdZ), k->eps2)) { *stop = 2; break; }
        }
        /* ref_path_get at w (sets out_is_xr / out_xr) */
        {
            int j = (k->n_xr > 0) ? xr_lookup(k, w) : -1;
            out_is_xr = (j >= 0);
            if (j >= 0) out_xr = XC(C(k->xr_m[2 * j], k->xr_m[2 * j + 1]), k->xr_e[j]);
        }
        cpx Zn1 = Zref(k, w);
        cpx ZZ = c_add(zs, Zn1);
        double fsq = ZZ.re * ZZ.re + ZZ.im * ZZ.im;
        if (k->calc_orbit) orbit_record(&orb, n, k->backshift, ZZ);
        if (fsq > k->M2) { *stop = 1; break; }
        if (w >= div_iter - 1) {
            zs = ZZ;
            zx = to_xrc(ZZ);
            if (dc) dx = xrc_add(dx, DXR(w));
            if (dz) {
                if (!nullify) hx = xrc_add(hx, n == order ? DZXR(w_wraped) : DZXR(w));
                nullify = 1;
            }
            w = 0;
            continue;
        }
        dyn = (fabs(ZZ.re) <= fabs(zs.re)) && (fabs(ZZ.im) <= fabs(zs.im));
        if (dyn) {
            xrc ZZx = out_is_xr ? xrc_add(zx, out_xr) : xrc_add_std(zx, Zn1);
            if (xrf_le(xrc_abs2(ZZx), xrc_abs2(zx))) {
                zx = ZZx;
                zs = xrc_std(ZZx);
                if (dc) dx = xrc_add(dx, DXR(w));
                if (dz) {
                    if (!nullify) hx = xrc_add(hx, n == order ? DZXR(w_wraped) : DZXR(w));
                    nullify = 1;
                }
                w = 0;
                continue;
            }
        }
    }
    *U = (int32_t)w;
    *out_z = c_add(xrc_std(zx), Zref(k, w));
    if (dc) *out_dzndc = xrc_std(xrc_add(dx, DXR(w)));
    /* perturbation.py:1377-1380 writes back zn and dzndc only: in Xrange mode
     * the standard dzndz field keeps its initial 0 (numba_initialize). */
    if (dz) *out_dzndz = C(0., 0.);
    if (k->calc_orbit) *out_orbit = orbit_finish(k, &orb, n, xrc_std(c_xr));
    return n;
#undef DXR
#undef DZXR
}

/* ------------------------------------------------------------------------ */
/* Burning ship family: burning_ship.py:12-60 (diffabs helpers), 535-859
 * (_p_iter_zn, _p_iter_hessian); loop perturbation.py:1468-1791. */
static inline double diffabs(double X, double x) {
    if (X >= 0.) { if (X + x >= 0.) return x; return -(2. * X + x); }
    if (X + x <= 0.) return -x;
    return 2. * X + x;
}
static inline double ddiffabsdX(double X, double x) {
    if (X >= 0.) return (X + x >= 0.) ? 0. : -2.;
    return (X + x <= 0.) ? 0. : 2.;
}
static inline double ddiffabsdx(double X, double x) {
    if (X >= 0.) return (X + x >= 0.) ? 1. : -1.;
    return (X + x <= 0.) ? -1. : 1.;
}

typedef struct { double x, y, dxa, dxb, dya, dyb; } bs_std;

static inline void bs_zn_std(int fl, double *x, double *y, double X, double Y, double a, double b) {
    double nx, ny;
    switch (fl) {
    case 1: {
        double XY = X * Y;
        nx = *x * (*x + 2. * X) - *y * (*y + 2. * Y) + a;
        ny = 2. * diffabs(XY, *x * *y + *x * Y + *y * X) - b;
    } break;
    case 2:
        nx = *x * (*x + 2. * X) - *y * (*y + 2. * Y) + a;
        ny = 2. * (X * diffabs(Y, *y) + *x * fabs(Y + *y)) - b;
        break;
    case 3:
        nx = *x * (*x + 2. * X) - Y * diffabs(Y, *y) - *y * fabs(Y + *y) + a;
        ny = 2. * (X * *y + Y * *x + *x * *y) - b;
        break;
    case 4: {
        double X2mY2 = X * X - Y * Y;
        nx = diffabs(X2mY2, *x * (*x + 2. * X) - *y * (*y + 2. * Y)) + a;
        ny = 2. * (X * *y + Y * *x + *x * *y) - b;
    } break;
    default: {
        double XY = X * Y;
        double X2mY2 = X * X - Y * Y;
        nx = diffabs(X2mY2, *x * (*x + 2. * X) - *y * (*y + 2. * Y)) + a;
        ny = 2. * diffabs(XY, *x * *y + *x * Y + *y * X) - b;
    } break;
    }
    *x = nx; *y = ny;
}

/* Jacobian ("hessian") perturbation: Z = (x, y, dxa, dxb, dya, dyb) */
static inline void bs_hess_std(int fl, bs_std *Z, double X, double Y,
                               double Xa, double Xb, double Ya, double Yb) {
    double x = Z->x, y = Z->y, xa = Z->dxa, xb = Z->dxb, ya = Z->dya, yb = Z->dyb;
    double nxa, nxb, nya, nyb;
    if (fl == 1 || fl == 5) {
        double opX = X * Y;
        double dXa = Xa * Y + X * Ya, dXb = Xb * Y + X * Yb;
        double opx = x * y + x * Y + y * X;
        double dxa_ = xa * y + x * ya + xa * Y + x * Ya + ya * X + y * Xa;
        double dxb_ = xb * y + x * yb + xb * Y + x * Yb + yb * X + y * Xb;
        double gX = ddiffabsdX(opX, opx), gx = ddiffabsdx(opX, opx);
        nya = 2. * (gX * dXa + gx * dxa_);
        nyb = 2. * (gX * dXb + gx * dxb_);
        if (fl == 1) {
            nxa = 2. * ((X + x) * xa + Xa * x) - 2. * ((Y + y) * ya + Ya * y);
            nxb = 2. * ((X + x) * xb + Xb * x) - 2. * ((Y + y) * yb + Yb * y);
        } else {
            double oX = X * X - Y * Y;
            double oXa = 2. * (X * Xa - Y * Ya), oXb = 2. * (X * Xb - Y * Yb);
            double ox = x * (x + 2. * X) - y * (y + 2. * Y);
            double oxa = xa * (x + 2. * X) + x * (xa + 2. * Xa) - ya * (y + 2. * Y) - y * (ya + 2. * Ya);
            double oxb = xb * (x + 2. * X) + x * (xb + 2. * Xb) - yb * (y + 2. * Y) - y * (yb + 2. * Yb);
            double hX = ddiffabsdX(oX, ox), hx = ddiffabsdx(oX, ox);
            nxa = hX * oXa + hx * oxa;
            nxb = hX * oXb + hx * oxb;
        }
    } else if (fl == 2) {
        double da = diffabs(Y, y), gX = ddiffabsdX(Y, y), gx = ddiffabsdx(Y, y);
        double Yy = Y + y, ab = fabs(Yy), sg = sgn(Yy);
        nxa = 2. * (((X + x) * xa + Xa * x) - ((Y + y) * ya + Ya * y));
        nxb = 2. * (((X + x) * xb + Xb * x) - ((Y + y) * yb + Yb * y));
        nya = 2. * (Xa * da + X * (gX * Ya + gx * ya) + xa * ab + x * sg * (Ya + ya));
        nyb = 2. * (Xb * da + X * (gX * Yb + gx * yb) + xb * ab + x * sg * (Yb + yb));
    } else if (fl == 3) {
        double da = diffabs(Y, y), gX = ddiffabsdX(Y, y), gx = ddiffabsdx(Y, y);
        double Yy = Y + y, ab = fabs(Yy), sg = sgn(Yy);
        nxa = xa * (x + 2. * X) + x * (xa + 2. * Xa) - Ya * da - Y * (Ya * gX + ya * gx) - ya * ab - y * sg * (Ya + ya);
        nxb = xb * (x + 2. * X) + x * (xb + 2. * Xb) - Yb * da - Y * (Yb * gX + yb * gx) - yb * ab - y * sg * (Yb + yb);
        nya = 2. * (Xa * y + X * ya + Ya * x + Y * xa + xa * y + x * ya);
        nyb = 2. * (Xb * y + X * yb + Yb * x + Y * xb + xb * y + x * yb);
    } else { /* 4: Celtic */
        double oX = X * X - Y * Y;
        double oXa = 2. * (X * Xa - Y * Ya), oXb = 2. * (X * Xb - Y * Yb);
        double ox = x * (x + 2. * X) - y * (y + 2. * Y);
        double oxa = xa * (x + 2. * X) + x * (xa + 2. * Xa) - ya * (y + 2. * Y) - y * (ya + 2. * Ya);
        double oxb = xb * (x + 2. * X) + x * (xb + 2. * Xb) - yb * (y + 2. * Y) - y * (yb + 2. * Yb);
        double hX = ddiffabsdX(oX, ox), hx = ddiffabsdx(oX, ox);
        nxa = hX * oXa + hx * oxa;
        nxb = hX * oXb + hx * oxb;
        nya = 2. * (Xa * y + X * ya + Ya * x + Y * xa + xa * y + x * ya);
        nyb = 2. * (Xb * y + X * yb + Yb * x + Y * xb + xb * y + x * yb);
    }
    Z->dxa = nxa; Z->dxb = nxb; Z->dya = nya; Z->dyb = nyb;
}

/* ---- Xrange versions of the BS kernels ---- */
static inline xrf xr_diffabs(xrf X, xrf x) {
    if (xrf_ge_std(X, 0.)) {
        if (xrf_ge_std(xrf_add(X, x), 0.)) return xrf_scale(1., x);
        return xrf_neg(xrf_add(xrf_scale(2., X), x));
    }
    if (xrf_le_std(xrf_add(X, x), 0.)) return xrf_neg(x);
    return xrf_add(xrf_scale(2., X), x);
}
static inline double xr_ddiffabsdX(xrf X, xrf x) {
    if (xrf_ge_std(X, 0.)) return xrf_ge_std(xrf_add(X, x), 0.) ? 0. : -2.;
    return xrf_le_std(xrf_add(X, x), 0.) ? 0. : 2.;
}
static inline double xr_ddiffabsdx(xrf X, xrf x) {
    if (xrf_ge_std(X, 0.)) return xrf_ge_std(xrf_add(X, x), 0.) ? 1. : -1.;
    return xrf_le_std(xrf_add(X, x), 0.) ? -1. : 1.;
}
static inline double xr_sgn(xrf x) { return xrf_lt_std(x, 0.) ? -1. : 1.; }

#define XA xrf_add
#define XS xrf_sub
#define XM xrf_mul
#define XK xrf_scale

static inline void bs_zn_xr(int fl, xrf *x, xrf *y, xrf X, xrf Y, xrf a, xrf b) {
    xrf nx, ny;
    /* x * (x + 2 X) - y * (y + 2 Y) */
    #define QUAD() XS(XM(*x, XA(*x, XK(2., X))), XM(*y, XA(*y, XK(2., Y))))
    switch (fl) {
    case 1: {
        xrf XY = XM(X, Y);
        nx = XA(QUAD(), a);
        ny = XS(XK(2., xr_diffabs(XY, XA(XA(XM(*x, *y), XM(*x, Y)), XM(*y, X)))), b);
    } break;
    case 2:
        nx = XA(QUAD(), a);
        ny = XS(XK(2., XA(XM(X, xr_diffabs(Y, *y)), XM(*x, xrf_abs(XA(Y, *y))))), b);
        break;
    case 3:
        nx = XA(XS(XS(XM(*x, XA(*x, XK(2., X))), XM(Y, xr_diffabs(Y, *y))), XM(*y, xrf_abs(XA(Y, *y)))), a);
        ny = XS(XK(2., XA(XA(XM(X, *y), XM(Y, *x)), XM(*x, *y))), b);
        break;
    case 4: {
        xrf X2mY2 = XS(XM(X, X), XM(Y, Y));
        nx = XA(xr_diffabs(X2mY2, QUAD()), a);
        ny = XS(XK(2., XA(XA(XM(X, *y), XM(Y, *x)), XM(*x, *y))), b);
    } break;
    default: {
        xrf XY = XM(X, Y);
        xrf X2mY2 = XS(XM(X, X), XM(Y, Y));
        nx = XA(xr_diffabs(X2mY2, QUAD()), a);
        ny = XS(XK(2., xr_diffabs(XY, XA(XA(XM(*x, *y), XM(*x, Y)), XM(*y, X)))), b);
    } break;
    }
    #undef QUAD
    *x = nx; *y = ny;
}

typedef struct { xrf x, y, dxa, dxb, dya, dyb; } bs_xr;

static inline void bs_hess_xr(int fl, bs_xr *Z, xrf X, xrf Y, xrf Xa, xrf Xb, xrf Ya, xrf Yb) {
    xrf x = Z->x, y = Z->y, xa = Z->dxa, xb = Z->dxb, ya = Z->dya, yb = Z->dyb;
    xrf nxa, nxb, nya, nyb;
    if (fl == 1 || fl == 5) {
        xrf opX = XM(X, Y);
        xrf dXa = XA(XM(Xa, Y), XM(X, Ya)), dXb = XA(XM(Xb, Y), XM(X, Yb));
        xrf opx = XA(XA(XM(x, y), XM(x, Y)), XM(y, X));
        xrf dxa_ = XA(XA(XA(XA(XA(XM(xa, y), XM(x, ya)), XM(xa, Y)), XM(x, Ya)), XM(ya, X)), XM(y, Xa));
        xrf dxb_ = XA(XA(XA(XA(XA(XM(xb, y), XM(x, yb)), XM(xb, Y)), XM(x, Yb)), XM(yb, X)), XM(y, Xb));
        double gX = xr_ddiffabsdX(opX, opx), gx = xr_ddiffabsdx(opX, opx);
        nya = XK(2., XA(XK(gX, dXa), XK(gx, dxa_)));
        nyb = XK(2., XA(XK(gX, dXb), XK(gx, dxb_)));
        if (fl == 1) {
            nxa = XS(XK(2., XA(XM(XA(X, x), xa), XM(Xa, x))), XK(2., XA(XM(XA(Y, y), ya), XM(Ya, y))));
            nxb = XS(XK(2., XA(XM(XA(X, x), xb), XM(Xb, x))), XK(2., XA(XM(XA(Y, y), yb), XM(Yb, y))));
        } else {
            xrf oX = XS(XM(X, X), XM(Y, Y));
            xrf oXa = XK(2., XS(XM(X, Xa), XM(Y, Ya))), oXb = XK(2., XS(XM(X, Xb), XM(Y, Yb)));
            xrf ox = XS(XM(x, XA(x, XK(2., X))), XM(y, XA(y, XK(2., Y))));
            xrf oxa = XS(XS(XA(XM(xa, XA(x, XK(2., X))), XM(x, XA(xa, XK(2., Xa)))), XM(ya, XA(y, XK(2., Y)))), XM(y, XA(ya, XK(2., Ya))));
            xrf oxb = XS(XS(XA(XM(xb, XA(x, XK(2., X))), XM(x, XA(xb, XK(2., Xb)))), XM(yb, XA(y, XK(2., Y)))), XM(y, XA(yb, XK(2., Yb))));
            double hX = xr_ddiffabsdX(oX, ox), hx = xr_ddiffabsdx(oX, ox);
            nxa = XA(XK(hX, oXa), XK(hx, oxa));
            nxb = XA(XK(hX, oXb), XK(hx, oxb));
        }
    } else if (fl == 2) {
        xrf da = xr_diffabs(Y, y);
        double gX = xr_ddiffabsdX(Y, y), gx = xr_ddiffabsdx(Y, y);
        xrf Yy = XA(Y, y), ab = xrf_abs(Yy);
        double sg = xr_sgn(Yy);
        nxa = XK(2., XS(XA(XM(XA(X, x), xa), XM(Xa, x)), XA(XM(XA(Y, y), ya), XM(Ya, y))));
        nxb = XK(2., XS(XA(XM(XA(X, x), xb), XM(Xb, x)), XA(XM(XA(Y, y), yb), XM(Yb, y))));
        nya = XK(2., XA(XA(XA(XM(Xa, da), XM(X, XA(XK(gX, Ya), XK(gx, ya)))), XM(xa, ab)), XM(XK(sg, x), XA(Ya, ya))));
        nyb = XK(2., XA(XA(XA(XM(Xb, da), XM(X, XA(XK(gX, Yb), XK(gx, yb)))), XM(xb, ab)), XM(XK(sg, x), XA(Yb, yb))));
    } else if (fl == 3) {
        xrf da = xr_diffabs(Y, y);
        double gX = xr_ddiffabsdX(Y, y), gx = xr_ddiffabsdx(Y, y);
        xrf Yy = XA(Y, y), ab = xrf_abs(Yy);
        double sg = xr_sgn(Yy);
        nxa = XS(XS(XS(XS(XA(XM(xa, XA(x, XK(2., X))), XM(x, XA(xa, XK(2., Xa)))), XM(Ya, da)), XM(Y, XA(XK(gX, Ya), XK(gx, ya)))), XM(ya, ab)), XM(XK(sg, y), XA(Ya, ya)));
        nxb = XS(XS(XS(XS(XA(XM(xb, XA(x, XK(2., X))), XM(x, XA(xb, XK(2., Xb)))), XM(Yb, da)), XM(Y, XA(XK(gX, Yb), XK(gx, yb)))), XM(yb, ab)), XM(XK(sg, y), XA(Yb, yb)));
        nya = XK(2., XA(XA(XA(XA(XA(XM(Xa, y), XM(X, ya)), XM(Ya, x)), XM(Y, xa)), XM(xa, y)), XM(x, ya)));
        nyb = XK(2., XA(XA(XA(XA(XA(XM(Xb, y), XM(X, yb)), XM(Yb, x)), XM(Y, xb)), XM(xb, y)), XM(x, yb)));
    } else {
        xrf oX = XS(XM(X, X), XM(Y, Y));
        xrf oXa = XK(2., XS(XM(X, Xa), XM(Y, Ya))), oXb = XK(2., XS(XM(X, Xb), XM(Y, Yb)));
        xrf ox = XS(XM(x, XA(x, XK(2., X))), XM(y, XA(y, XK(2., Y))));
        xrf oxa = XS(XS(XA(XM(xa, XA(x, XK(2., X))), XM(x, XA(xa, XK(2., Xa)))), XM(ya, XA(y, XK(2., Y)))), XM(y, XA(ya, XK(2., Ya))));
        xrf oxb = XS(XS(XA(XM(xb, XA(x, XK(2., X))), XM(x, XA(xb, XK(2., Xb)))), XM(yb, XA(y, XK(2., Y)))), XM(y, XA(yb, XK(2., Yb))));
        double hX = xr_ddiffabsdX(oX, ox), hx = xr_ddiffabsdx(oX, ox);
        nxa = XA(XK(hX, oXa), XK(hx, oxa));
        nxb = XA(XK(hX, oXb), XK(hx, oxb));
        nya = XK(2., XA(XA(XA(XA(XA(XM(Xa, y), XM(X, ya)), XM(Ya, x)), XM(Y, xa)), XM(xa, y)), XM(x, ya)));
        nyb = XK(2., XA(XA(XA(XA(XA(XM(Xb, y), XM(X, yb)), XM(Yb, x)), XM(Y, xb)), XM(xb, y)), XM(x, yb)));
    }
    Z->dxa = nxa; Z->dxb = nxb; Z->dya = nya; Z->dyb = nyb;
}

/* numba_iterate_BS(...).numba_impl, perturbation.py:1468-1791 */
/* BS_iterate (burning_ship.py:82-123): the full recurrence per flavour */
static void bs_iterate_full(int fl, double *x, double *y, double A, double B) {
    const double X = *x, Y = *y;
    double nx, ny;
    switch (fl) {
        case 1: nx = X * X - Y * Y + A; ny = 2. * fabs(X * Y) - B; break;
        case 2: nx = X * X - Y * Y + A; ny = 2. * X * fabs(Y) - B; break;
        case 3: nx = X * X - Y * fabs(Y) + A; ny = 2. * X * Y - B; break;
        case 4: nx = fabs(X * X - Y * Y) + A; ny = 2. * X * Y - B; break;
        default: nx = fabs(X * X - Y * Y) + A; ny = 2. * fabs(X * Y) - B; break;
    }
    *x = nx; *y = ny;
}

/* orbit finish of the BS loop, perturbation.py:1778-1787 */
static void bs_orbit_finish(const or_ctx *k, orbit_rec *o, int64_t n, double a, double b, double *out) {
    double x = o->z2.re, y = o->z2.im;
    const cpx R1 = Zref(k, 1);
    const double AA = a + R1.re, BB = b + R1.im;
    while (o->i2 < n - k->backshift) { bs_iterate_full(k->flavor, &x, &y, AA, BB); o->i2++; }
    out[0] = x; out[1] = y;
}

static int64_t iterate_bs(const or_ctx *k, double px, double py, double *out6,
                          int32_t *U, int8_t *stop) {
    orbit_rec orb = {0, 0, 0, {0., 0.}, {0., 0.}};
    const int64_t L = k->L, order = k->order, div_iter = k->div_iter;
    const int hs = k->calc_dzndc;
    const int fl = k->flavor;
    int64_t w = 0, n = 0;
    int64_t first_invalid = L;
    if (div_iter < first_invalid) first_invalid = div_iter;
    if (order < first_invalid) first_invalid = order;
    int dyn = 1;
    const double *Mb;
    int64_t step;

    /* ref_path_c_from_pix_BS, perturbation.py:2260-2280 */
    double x1 = k->lin_mat[0] * px + k->lin_mat[1] * py;
    double y1 = k->lin_mat[2] * px + k->lin_mat[3] * py;
    cpx m = C(k->dx_m * x1, k->dx_m * y1);
    xrc cx = need_renorm_c(m) ? normalize_cplx(m, k->dx_e) : XC(m, k->dx_e);
    xrf a_xr = xrf_add(XF(cx.m.re, cx.e), XF(k->drift_m[0], k->drift_e[0]));
    xrf b_xr = xrf_add(XF(cx.m.im, cx.e), XF(k->drift_m[1], k->drift_e[1]));
    double a = xrf_std(a_xr), b = xrf_std(b_xr);
    const double *PA = k->dbs_m[0], *PB = k->dbs_m[1], *PC = k->dbs_m[2], *PD = k->dbs_m[3];

    if (!k->xr) {
        bs_std Z = {0., 0., 0., 0., 0., 0.};
        for (;;) {
            if (k->bla_on && (w & 7) == 0) {
                step = bla_get(k, Z.x, Z.y, w, first_invalid, &Mb);
                if (step != 0) {
                    n += step;
                    w = (w + step) % order;
                    double nx = Mb[0] * Z.x + Mb[1] * Z.y + Mb[4] * a + Mb[5] * b;
                    double ny = Mb[2] * Z.x + Mb[3] * Z.y + Mb[6] * a + Mb[7] * b;
                    Z.x = nx; Z.y = ny;
                    if (hs) {
                        double na = Mb[0] * Z.dxa + Mb[1] * Z.dya;
                        double nb = Mb[0] * Z.dxb + Mb[1] * Z.dyb;
                        double nc = Mb[2] * Z.dxa + Mb[3] * Z.dya;
                        double nd = Mb[2] * Z.dxb + Mb[3] * Z.dyb;
                        Z.dxa = na; Z.dxb = nb; Z.dya = nc; Z.dyb = nd;
                    }
                    continue;
                }
            }
            n += 1;
            cpx R = Zref(k, w);
            if (hs) {
                double ra = PA[w], rb = PB[w], rc = PC[w], rd = PD[w];
                if (dyn) { ra = 0.; rb = 0.; rc = 0.; rd = 0.; }
                bs_hess_std(fl, &Z, R.re, R.im, ra, rb, rc, rd);
            }
            bs_zn_std(fl, &Z.x, &Z.y, R.re, R.im, a, b);
            if (n >= k->max_iter) { *stop = 0; break; }
            w += 1;
            if (w >= order) w = w % order;
            cpx R1 = Zref(k, w);
            double XX = Z.x + R1.re, YY = Z.y + R1.im;
            double fsq = XX * XX + YY * YY;
            if (k->calc_orbit) orbit_record(&orb, n, k->backshift, C(XX, YY));
            if (fsq > k->M2) { *stop = 1; break; }
            if (w >= div_iter - 1) {
                Z.x = XX; Z.y = YY;
                if (hs) { Z.dxa += PA[w]; Z.dxb += PB[w]; Z.dya += PC[w]; Z.dyb += PD[w]; }
                w = 0;
                continue;
            }
            dyn = (fabs(XX) <= fabs(Z.x)) && (fabs(YY) <= fabs(Z.y));
            if (dyn) {
                Z.x = XX; Z.y = YY;
                if (hs) { Z.dxa += PA[w]; Z.dxb += PB[w]; Z.dya += PC[w]; Z.dyb += PD[w]; }
                w = 0;
                continue;
            }
        }
        *U = (int32_t)w;
        cpx R = Zref(k, w);
        out6[0] = Z.x + R.re; out6[1] = Z.y + R.im;
        if (hs) {
            out6[2] = Z.dxa + PA[w]; out6[3] = Z.dxb + PB[w];
            out6[4] = Z.dya + PC[w]; out6[5] = Z.dyb + PD[w];
        }
        if (k->calc_orbit) bs_orbit_finish(k, &orb, n, a, b, out6 + 6);
        return n;
    }

#define PX(q, i) XF(k->dbs_m[q][i], k->dbs_e[q][i])
    const xrf zf = XF(0., 0);
    bs_xr Z = {zf, zf, zf, zf, zf, zf};
    double xs = 0., ys = 0.;
    int out_is_xr = 0;
    xrf oxx = zf, oxy = zf;
    for (;;) {
        if (k->bla_on && (w & 7) == 0) {
            step = bla_get(k, xs, ys, w, first_invalid, &Mb);
            if (step != 0) {
                n += step;
                w = (w + step) % order;
                /* apply_BLA_BS, perturbation.py:1793-1798 */
                xrf nx = XA(XA(XA(XK(Mb[0], Z.x), XK(Mb[1], Z.y)), XK(Mb[4], a_xr)), XK(Mb[5], b_xr));
                xrf ny = XA(XA(XA(XK(Mb[2], Z.x), XK(Mb[3], Z.y)), XK(Mb[6], a_xr)), XK(Mb[7], b_xr));
                Z.x = nx; Z.y = ny;
                xs = xrf_std(Z.x); ys = xrf_std(Z.y);
                if (hs) {
                    xrf na = XA(XK(Mb[0], Z.dxa), XK(Mb[1], Z.dya));
                    xrf nb = XA(XK(Mb[0], Z.dxb), XK(Mb[1], Z.dyb));
                    xrf nc = XA(XK(Mb[2], Z.dxa), XK(Mb[3], Z.dya));
                    xrf nd = XA(XK(Mb[2], Z.dxb), XK(Mb[3], Z.dyb));
                    Z.dxa = na; Z.dxb = nb; Z.dya = nc; Z.dyb = nd;
                }
                continue;
            }
        }
        n += 1;
        xrf X, Y;
        {
            int j = (k->n_xr > 0) ? xr_lookup(k, w) : -1;
            out_is_xr = (j >= 0);
            if (j >= 0) {
                oxx = XF(k->xr_m[2 * j], k->xr_e[2 * j]); oxy = XF(k->xr_m[2 * j + 1], k->xr_e[2 * j + 1]);
                X = oxx; Y = oxy;
            } else { X = to_xrf(Zref(k, w).re); Y = to_xrf(Zref(k, w).im); }
        }
        if (hs) {
            xrf ra = PX(0, w), rb = PX(1, w), rc = PX(2, w), rd = PX(3, w);
            if (dyn) { ra = zf; rb = zf; rc = zf; rd = zf; }
            bs_hess_xr(fl, &Z, X, Y, ra, rb, rc, rd);
        }
        bs_zn_xr(fl, &Z.x, &Z.y, X, Y, a_xr, b_xr);
        xs = xrf_std(Z.x); ys = xrf_std(Z.y);
        if (n >= k->max_iter) { *stop = 0; break; }
        w += 1;
        if (w >= order) w = w % order;
        {
            int j = (k->n_xr > 0) ? xr_lookup(k, w) : -1;
            out_is_xr = (j >= 0);
            if (j >= 0) { oxx = XF(k->xr_m[2 * j], k->xr_e[2 * j]); oxy = XF(k->xr_m[2 * j + 1], k->xr_e[2 * j + 1]); }
        }
        cpx R1 = Zref(k, w);
        double XX = xs + R1.re, YY = ys + R1.im;
        double fsq = XX * XX + YY * YY;
        if (k->calc_orbit) orbit_record(&orb, n, k->backshift, C(XX, YY));
        if (fsq > k->M2) { *stop = 1; break; }
        if (w >= div_iter - 1) {
            xs = XX; ys = YY;
            Z.x = to_xrf(XX); Z.y = to_xrf(YY);
            if (hs) {
                Z.dxa = XA(Z.dxa, PX(0, w)); Z.dxb = XA(Z.dxb, PX(1, w));
                Z.dya = XA(Z.dya, PX(2, w)); Z.dyb = XA(Z.dyb, PX(3, w));
            }
            w = 0;
            continue;
        }
        dyn = (fabs(XX) <= fabs(xs)) && (fabs(YY) <= fabs(ys));
        if (dyn) {
            xrf XXx, YYx;
            if (out_is_xr) { XXx = XA(Z.x, oxx); YYx = XA(Z.y, oxy); }
            else { XXx = xrf_add_std(Z.x, R1.re); YYx = xrf_add_std(Z.y, R1.im); }
            xrf lhs = XA(XM(XXx, XXx), XM(YYx, YYx));
            xrf rhs = XA(XM(Z.x, Z.x), XM(Z.y, Z.y));
            if (xrf_le(lhs, rhs)) {
                Z.x = XXx; Z.y = YYx;
                xs = xrf_std(XXx); ys = xrf_std(YYx);
                if (hs) {
                    Z.dxa = XA(Z.dxa, PX(0, w)); Z.dxb = XA(Z.dxb, PX(1, w));
                    Z.dya = XA(Z.dya, PX(2, w)); Z.dyb = XA(Z.dyb, PX(3, w));
                }
                w = 0;
                continue;
            }
        }
    }
    *U = (int32_t)w;
    cpx R = Zref(k, w);
    out6[0] = xrf_std(xrf_add_std(Z.x, R.re));
    out6[1] = xrf_std(xrf_add_std(Z.y, R.im));
    if (hs) {
        out6[2] = xrf_std(XA(Z.dxa, PX(0, w))); out6[3] = xrf_std(XA(Z.dxb, PX(1, w)));
        out6[4] = xrf_std(XA(Z.dya, PX(2, w))); out6[5] = xrf_std(XA(Z.dyb, PX(3, w)));
    }
    if (k->calc_orbit) bs_orbit_finish(k, &orb, n, a, b, out6 + 6);
    return n;
#undef PX
}

/* ======================================================================== */
/* exported API (ctypes) */

typedef struct {
    int32_t kind, degree, flavor, xr;
    int64_t L, div_iter, order;
    const double *Zn;
    int32_t n_xr; int32_t pad0;
    const int32_t *xr_idx; const double *xr_m; const int32_t *xr_e;
    double dx_m; int32_t dx_e; int32_t pad1;
    double drift_m[2]; int32_t drift_e[2];
    double lin_mat[4];
    double kc_m; int32_t kc_e; int32_t bla_on;
    double bla_eps;
    int64_t max_iter;
    double M2, eps2;
    int32_t calc_dzndc, calc_dzndz;
    int32_t calc_orbit, backshift;
    double sd_m; int32_t sd_e; int32_t pad3;   /* derivative scale; sd_m == 0: dx */
} or_desc;

void *or_ctx_create(const or_desc *d) {
    or_ctx *k = (or_ctx *)calloc(1, sizeof(or_ctx));
    k->kind = d->kind; k->degree = d->degree; k->flavor = d->flavor; k->xr = d->xr;
    k->L = d->L; k->div_iter = d->div_iter; k->order = d->order;
    k->Zn = d->Zn; k->n_xr = d->n_xr; k->xr_idx = d->xr_idx; k->xr_m = d->xr_m; k->xr_e = d->xr_e;
    k->dx_m = d->dx_m; k->dx_e = d->dx_e;
    k->sd_m = d->sd_m != 0. ? d->sd_m : d->dx_m; k->sd_e = d->sd_m != 0. ? d->sd_e : d->dx_e;
    k->drift_m[0] = d->drift_m[0]; k->drift_m[1] = d->drift_m[1];
    k->drift_e[0] = d->drift_e[0]; k->drift_e[1] = d->drift_e[1];
    memcpy(k->lin_mat, d->lin_mat, sizeof(k->lin_mat));
    k->kc_m = d->kc_m; k->kc_e = d->kc_e;
    k->bla_on = d->bla_on; k->bla_eps = d->bla_eps;
    k->max_iter = d->max_iter; k->M2 = d->M2; k->eps2 = d->eps2;
    k->calc_dzndc = d->calc_dzndc; k->calc_dzndz = (k->kind == OR_KIND_BS) ? 0 : d->calc_dzndz;
    k->calc_orbit = d->calc_orbit; k->backshift = d->backshift;
    if (k->calc_orbit && k->backshift <= 0) { free(k); return NULL; }
    int rc = (k->kind == OR_KIND_BS) ? bs_paths(k) : holo_paths(k);
    if (rc != 0) { free(k); return NULL; }
    if (k->bla_on) {
        if (stages_bla(k->L) <= STG_COMPRESSED) { free(k); return NULL; }
        make_bla(k);
    }
    return k;
}

void or_ctx_destroy(void *p) {
    or_ctx *k = (or_ctx *)p;
    if (!k) return;
    free(k->dzndc_m); free(k->dzndc_e); free(k->dzndz_m); free(k->dzndz_e);
    for (int q = 0; q < 4; q++) { free(k->dbs_m[q]); free(k->dbs_e[q]); }
    free(k->M_bla); free(k->r_bla);
    free(k);
}

/* copies of the derived tables for cross-checks */
int64_t or_bla_info(void *p, int32_t *stages) {
    or_ctx *k = (or_ctx *)p; *stages = k->stages_bla; return k->bla_len;
}
void or_bla_copy(void *p, double *M, double *r) {
    or_ctx *k = (or_ctx *)p;
    int dim = (k->kind == OR_KIND_BS) ? 8 : 4;
    memcpy(M, k->M_bla, (size_t)(k->bla_len * dim) * sizeof(double));
    memcpy(r, k->r_bla, (size_t)k->bla_len * sizeof(double));
}
/* which = 0: dzndc (holo, [2L] + [L]); 1: dzndz ([2(L+1)] + [L+1]); 2..5: BS paths */
void or_path_copy(void *p, int which, double *m, int32_t *e) {
    or_ctx *k = (or_ctx *)p;
    int64_t L = k->L;
    if (which == 0) { memcpy(m, k->dzndc_m, 2 * L * sizeof(double)); memcpy(e, k->dzndc_e, L * sizeof(int32_t)); }
    else if (which == 1) { memcpy(m, k->dzndz_m, 2 * (L + 1) * sizeof(double)); memcpy(e, k->dzndz_e, (L + 1) * sizeof(int32_t)); }
    else { memcpy(m, k->dbs_m[which - 2], L * sizeof(double)); memcpy(e, k->dbs_e[which - 2], L * sizeof(int32_t)); }
}

/* numba_cycles_perturb / numba_cycles_perturb_BS, perturbation.py:988-1045,
 * 1406-1451.  Z is the reference's [n_Z, npts] layout: complex128 for the
 * holomorphic models, float64 for BS.  Pixels are distributed with OpenMP,
 * dynamic schedule (the reference runs one thread per 200x200 tile). */
int or_cycles(void *p, const double *c_pix, int64_t npts, double *Z, int32_t *U,
              int8_t *stop_reason, int32_t *stop_iter, int nthreads) {
    const or_ctx *k = (const or_ctx *)p;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    if (k->kind == OR_KIND_BS) {
        const int nj = 2 + (k->calc_dzndc ? 4 : 0);
        #pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < npts; i++) {
            double out8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            int8_t st = -1;
            int64_t n = iterate_bs(k, c_pix[2 * i], c_pix[2 * i + 1], out8, &U[i], &st);
            for (int f = 0; f < nj; f++) Z[f * npts + i] = out8[f];
            if (k->calc_orbit) { Z[nj * npts + i] = out8[6]; Z[(nj + 1) * npts + i] = out8[7]; }
            stop_iter[i] = (int32_t)n;
            stop_reason[i] = st;
        }
        return 0;
    }
    int f_dz = k->calc_dzndz ? 1 : -1;
    int f_dc = k->calc_dzndc ? (k->calc_dzndz ? 2 : 1) : -1;
    int f_orb = k->calc_orbit ? 1 + (k->calc_dzndz ? 1 : 0) + (k->calc_dzndc ? 1 : 0) : -1;
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < npts; i++) {
        xrc cx = c_from_pix_holo(k, c_pix[2 * i], c_pix[2 * i + 1]);
        cpx oz = C(0, 0), oh = C(0, 0), od = C(0, 0), oo = C(0, 0);
        int8_t st = -1;
        int64_t n = iterate_holo(k, cx, &oz, &oh, &od, &oo, &U[i], &st);
        Z[2 * i] = oz.re; Z[2 * i + 1] = oz.im;
        if (f_dz > 0) { Z[2 * (f_dz * npts + i)] = oh.re; Z[2 * (f_dz * npts + i) + 1] = oh.im; }
        if (f_dc > 0) { Z[2 * (f_dc * npts + i)] = od.re; Z[2 * (f_dc * npts + i) + 1] = od.im; }
        if (f_orb > 0) { Z[2 * (f_orb * npts + i)] = oo.re; Z[2 * (f_orb * npts + i) + 1] = oo.im; }
        stop_iter[i] = (int32_t)n;
        stop_reason[i] = st;
    }
    return 0;
}

/* ---- Xrange primitive probes, for the closed-form tests ---- */
/* op: 0 add, 1 sub, 2 mul; complex operands (mantissa re/im + exp) */
void or_xrc_binop(int op, int64_t n, const double *am, const int32_t *ae,
                  const double *bm, const int32_t *be, double *om, int32_t *oe) {
    for (int64_t i = 0; i < n; i++) {
        xrc a = XC(C(am[2 * i], am[2 * i + 1]), ae[i]);
        xrc b = XC(C(bm[2 * i], bm[2 * i + 1]), be[i]);
        xrc r = (op == 0) ? xrc_add(a, b) : (op == 1) ? xrc_sub(a, b) : xrc_mul(a, b);
        om[2 * i] = r.m.re; om[2 * i + 1] = r.m.im; oe[i] = r.e;
    }
}
void or_xrf_binop(int op, int64_t n, const double *am, const int32_t *ae,
                  const double *bm, const int32_t *be, double *om, int32_t *oe) {
    for (int64_t i = 0; i < n; i++) {
        xrf a = XF(am[i], ae[i]), b = XF(bm[i], be[i]);
        xrf r = (op == 0) ? xrf_add(a, b) : (op == 1) ? xrf_sub(a, b) : xrf_mul(a, b);
        om[i] = r.m; oe[i] = r.e;
    }
}
/* complex to_standard and normalize */
void or_xrc_to_std(int64_t n, const double *am, const int32_t *ae, double *out) {
    for (int64_t i = 0; i < n; i++) {
        cpx r = xrc_std(XC(C(am[2 * i], am[2 * i + 1]), ae[i]));
        out[2 * i] = r.re; out[2 * i + 1] = r.im;
    }
}
void or_xrc_normalize(int64_t n, const double *am, const int32_t *ae, double *om, int32_t *oe) {
    for (int64_t i = 0; i < n; i++) {
        xrc r = normalize_cplx(C(am[2 * i], am[2 * i + 1]), ae[i]);
        om[2 * i] = r.m.re; om[2 * i + 1] = r.m.im; oe[i] = r.e;
    }
}
double or_hypot(double x, double y) { return fs_hypot(x, y); }
