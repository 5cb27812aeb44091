// fsg_fp32x.hip -- the binary32-mantissa + int32-exponent form of the tile
// kernel (FSG_PREC_FP32X).  Not the drop-in: it trades the reference's
// binary64 / Xrange rounding for speed and reports its error against the
// binary64 path.
//
// Every complex quantity of the delta loop (z, dz/dc, dz/dz, c, the orbit
// and its derivative paths, the BLA coefficients) is a pair of floats whose
// larger component is kept in [1, 2) plus an int32 exponent.  Normalising is
// three instructions (v_frexp_exp_i32_f32, two v_ldexp_f32), with no
// data-dependent branch, so the exponent range is unbounded for the cost of
// binary32 arithmetic.  The escape test and the rebase tests that compare z
// with the reference orbit run in binary64, on the binary64 orbit.
//
// The loop structure, BLA lookup and rebase rules are those of the binary64
// kernels (numba_iterate, src/fractalshades/perturbation.py:1065-1400).
#include <cstring>
#include <cstdlib>
#include "models.hpp"
#include "tile_loop.hpp"

namespace fsg {

// ---------------------------------------------------------------------------
// extended-exponent binary32 complex
struct xs { float x, y; int32_t e; };

constexpr int32_t XS_ZERO_E = -(1 << 28);   // exponent of 0: aligns to nothing

__device__ __forceinline__ int32_t fexp(float v) { return __builtin_amdgcn_frexp_expf(v); }

__device__ __forceinline__ xs xs_norm(float x, float y, int32_t e) {
    float mx = fmaxf(fabsf(x), fabsf(y));
    if (mx == 0.f) return xs{0.f, 0.f, XS_ZERO_E};
    int32_t k = fexp(mx) - 1;                  // mx in [2^k, 2^(k+1))
    return xs{ldexpf(x, -k), ldexpf(y, -k), e + k};
}

__device__ __forceinline__ xs operator+(xs a, xs b) {
    int32_t E = max(a.e, b.e);
    int32_t sa = max(a.e - E, -200), sb = max(b.e - E, -200);
    return xs_norm(ldexpf(a.x, sa) + ldexpf(b.x, sb), ldexpf(a.y, sa) + ldexpf(b.y, sb), E);
}
__device__ __forceinline__ xs operator*(xs a, xs b) {
    return xs_norm(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x, a.e + b.e);
}
__device__ __forceinline__ xs muls(double s, xs a) {
    float f = (float)s;
    return xs_norm(f * a.x, f * a.y, a.e);
}
__device__ __forceinline__ xs xs_zero() { return xs{0.f, 0.f, XS_ZERO_E}; }

// sum of two unnormalised terms m1 2^e1 + m2 2^e2, normalised
__device__ __forceinline__ xs xs_sum(float x1, float y1, int32_t e1, float x2, float y2, int32_t e2) {
    const int32_t E = max(e1, e2);
    const int32_t s1 = max(e1 - E, -200), s2 = max(e2 - E, -200);
    return xs_norm(ldexpf(x1, s1) + ldexpf(x2, s2), ldexpf(y1, s1) + ldexpf(y2, s2), E);
}
// mantissa product
__device__ __forceinline__ void cmul(float ax, float ay, float bx, float by, float &rx, float &ry) {
    rx = ax * bx - ay * by;
    ry = ax * by + ay * bx;
}

// binary64 value <-> xs
__device__ __forceinline__ xs xs_from(double re, double im, int32_t e) {
    double mx = fmax(fabs(re), fabs(im));
    if (mx == 0.) return xs_zero();
    int32_t k = __builtin_amdgcn_frexp_exp(mx) - 1;
    return xs_norm((float)ldexp(re, -k), (float)ldexp(im, -k), e + k);
}
__device__ __forceinline__ cplx xs_to_d(xs a) {
    int32_t e = max(a.e, -2000);
    return mk(ldexp((double)a.x, e), ldexp((double)a.y, e));
}
// |a|^2 as (binary64 mantissa, exponent)
__device__ __forceinline__ double xs_abs2_m(xs a) { return (double)a.x * a.x + (double)a.y * a.y; }
__device__ __forceinline__ bool xs_abs2_le(xs a, xs b) {
    // |a|^2 <= |b|^2 with exponents 2 a.e and 2 b.e
    double ma = xs_abs2_m(a), mb = xs_abs2_m(b);
    if (mb == 0.) return ma == 0.;
    if (ma == 0.) return true;
    int32_t d = 2 * (a.e - b.e);
    d = max(min(d, 1100), -1100);
    return ldexp(ma, d) <= mb;
}

// lazily normalised intermediates of one step: products add exponents,
// sums align on the larger one, nothing is renormalised until the step's
// result (mantissas of a z^d + c step stay far inside the binary32 range:
// at most d factors in [1, 2) and binomial coefficients <= 70).  Zero takes
// an exponent low enough to align to nothing, small enough that the
// exponents of the 8 zero factors a step can multiply still fit in int32.
struct xt { float x, y; int32_t e; };
constexpr int32_t XT_ZERO_E = -(1 << 25);

__device__ __forceinline__ xt as_t(xs a) { return xt{a.x, a.y, a.e == XS_ZERO_E ? XT_ZERO_E : a.e}; }
__device__ __forceinline__ xs of_t(xt a) { return xs_norm(a.x, a.y, a.e); }
__device__ __forceinline__ xt operator+(xt a, xt b) {
    const int32_t E = max(a.e, b.e);
    const int32_t sa = max(a.e - E, -200), sb = max(b.e - E, -200);
    const float x = ldexpf(a.x, sa) + ldexpf(b.x, sb), y = ldexpf(a.y, sa) + ldexpf(b.y, sb);
    return xt{x, y, (x == 0.f && y == 0.f) ? XT_ZERO_E : E};
}
__device__ __forceinline__ xt operator*(xt a, xt b) {
    return xt{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x, a.e + b.e};
}
__device__ __forceinline__ xt muls(double s, xt a) { const float f = (float)s; return xt{f * a.x, f * a.y, a.e}; }

// ---------------------------------------------------------------------------
// per-reference tables in the xs format, built once per context
__global__ void k_xs_from_path(const double *m, const int32_t *e, int32_t n, xs *out) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = xs_from(m[2 * i], m[2 * i + 1], e ? e[i] : 0);
}

// orbit: binary64 values, with the Xrange register taking over below 1e-300
__global__ void k_xs_orbit(const cplx *Zn, int32_t L, const int32_t *xr_idx, const double *xr_m,
                           const int32_t *xr_e, int32_t n_xr, xs *out) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    int32_t lo = 0, hi = n_xr;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (xr_idx[mid] < i) lo = mid + 1; else hi = mid;
    }
    if (lo < n_xr && xr_idx[lo] == i) out[i] = xs_from(xr_m[2 * lo], xr_m[2 * lo + 1], xr_e[lo]);
    else out[i] = xs_from(Zn[i].re, Zn[i].im, 0);
}

// BLA (A, B) pairs
__global__ void k_xs_bla(const double *M, int64_t n, xs *A, xs *B) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    A[i] = xs_from(M[4 * i], M[4 * i + 1], 0);
    B[i] = xs_from(M[4 * i + 2], M[4 * i + 3], 0);
}

hipError_t build_xs_tables(const Tables &t, int64_t bla_len, TablesX &x, hipStream_t stream) {
    const int tb = 256;
    hipLaunchKernelGGL(k_xs_orbit, dim3((t.L + tb - 1) / tb), dim3(tb), 0, stream, t.Zn, t.L,
                       t.xr_idx, t.xr_m, t.xr_e, t.n_xr, (xs *)x.Zx);
    if (t.dc) hipLaunchKernelGGL(k_xs_from_path, dim3((t.L + tb - 1) / tb), dim3(tb), 0, stream,
                                 (const double *)t.dc, t.dc_e, t.L, (xs *)x.DCx);
    if (t.dz) hipLaunchKernelGGL(k_xs_from_path, dim3((t.L + tb) / tb), dim3(tb), 0, stream,
                                 (const double *)t.dz, t.dz_e, t.L + 1, (xs *)x.DZx);
    if (t.bla_on && bla_len > 0)
        hipLaunchKernelGGL(k_xs_bla, dim3((unsigned)((bla_len + tb - 1) / tb)), dim3(tb), 0, stream,
                           t.M, bla_len, (xs *)x.Ax, (xs *)x.Bx);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Holomorphic models on scaled binary32 mantissas, with a running error
// bound.
//
// z, dz/dc and dz/dz are each a pair of floats times 2^e, with one exponent
// per quantity that stays fixed between renormalisations: the mantissas are
// renormalised at every BLA point (w a multiple of 8, before the lookup),
// after a BLA jump, at a rebase, and when one grows past 2^32.  In between,
// a z^2 + c step is plain binary32 arithmetic on the mantissas (c is held in
// z's scale), with fused multiply-adds; z's value in absolute terms, needed
// next to the orbit point, is one ldexp.  The escape and rebase tests run in
// binary64 on the binary64 orbit, as the drop-in's do.
//
// e2 estimates the variance of z - z_64 (in z's scale squared), z_64 being
// the binary64 loop's delta at the same step: each step multiplies it by
// |f'(Z + z)|^2 and adds the square of the step's own rounding (KE x the
// magnitudes of the terms, c and the orbit point being rounded to binary32
// too).  Rounding errors of successive steps are independent, so they add
// in quadrature; a worst-case sum, linear in the step count, would exceed
// the actual error by orders of magnitude after a thousand steps and leave
// no decision certain.  The error bound is ERR_SIGMAS standard deviations.
// The multiplier is the exact squared modulus: an over-estimate compounds
// over thousands of steps while z itself grows by the modulus only.  Every decision the loop takes
// on z -- the BLA stage (|z| against the radii), the escape, the rebase --
// is checked against that bound: when the binary64 loop could have decided
// otherwise, the pixel stops as REASON_UNSURE and a binary64 pass
// recomputes it.  The interior test is checked against a relative margin.
constexpr float KE = 4.f * 0x1p-24f;          // fresh error per step, x (term magnitudes)
constexpr double ERR_SIGMAS = 8.;              // error bound, in standard deviations
constexpr float RENORM_AT = 0x1p32f;           // mantissa 1-norm that forces a renormalisation
constexpr double INTERIOR_MARGIN = 1e-3;       // relative margin of the stationarity test
// Steps (plain steps, each with the BLA jumps before it) after which a pixel
// still iterating stops as REASON_UNSURE and is left to the binary64 pass.
// The last pixels of a launch run alone in their waves, where a step's time
// is its latency, and the binary32 step, with its error estimate and scaled
// exponents, has the longer one: a pixel of 5e6 plain steps took 15 s in
// the binary32 pass against 2.5 s in the binary64 one.  Below the cap the
// pass loses nothing; above it a pixel costs at most the cap's binary32
// steps more than in the binary64 loop alone.
constexpr int32_t FP32X_STEP_CAP = 1 << 16;

// BLA stage selection of bla_find with its certainty: |z| is known to
// within eb, so the binary64 loop takes the same stage unless a radius lies
// within eb of |z|; only the accepted stage's radius and the next stage's
// can (the radii do not grow with the stage for one start).
__device__ __forceinline__ int32_t bla_find_m(const Tables &t, double az, double eb, int32_t w,
                                              int32_t &idx, bool &unsure) {
    const int32_t it = w >> 3;
    const int32_t invalid_step = t.first_invalid - w;
    if (invalid_step <= 8) return 0;
    const int32_t tz = (it == 0) ? 31 : __builtin_ctz((uint32_t)it);
    const int32_t s_inv = 31 - __builtin_clz((uint32_t)(invalid_step - 1));
    const int32_t cap = min(min(3 + tz, t.stages_bla - 1), s_inv);
    int32_t hi = cap, lo = 3, best = 0;
    double rb = 0.;
    while (lo <= hi) {
        const int32_t mid = (lo + hi) >> 1;
        const double r = t.r[2 * it + (1 << (mid - 3)) - 1];
        if (az < r) { best = mid; rb = r; lo = mid + 1; }
        else hi = mid - 1;
    }
    if (best > 0 && rb - az <= eb) unsure = true;
    const int32_t nxt = best > 0 ? best + 1 : 3;
    if (nxt <= cap && az - t.r[2 * it + (1 << (nxt - 3)) - 1] < eb) unsure = true;
    if (best == 0) return 0;
    idx = 2 * it + (1 << (best - 3)) - 1;
    return 1 << best;
}

__device__ __forceinline__ float l1(float x, float y) { return fabsf(x) + fabsf(y); }
__device__ __forceinline__ float l2(float x, float y) { return __builtin_sqrtf(x * x + y * y); }

template <int KIND, int DEG>
struct HoloS32 {
    struct State { xs z, c, d, h; float cx, cy, e2; int32_t w, n, steps; bool dyn, nul; };

    // c in z's scale, after z's exponent moved
    __device__ static void sync_c(State &s) {
        const int32_t sc = min(s.c.e - s.z.e, 100);
        s.cx = ldexpf(s.c.x, sc); s.cy = ldexpf(s.c.y, sc);
    }
    __device__ static void renorm_z(State &s) {
        const float mz = fmaxf(fabsf(s.z.x), fabsf(s.z.y));
        const int32_t k = (mz == 0.f) ? 0 : fexp(mz) - 1;
        s.z.x = ldexpf(s.z.x, -k); s.z.y = ldexpf(s.z.y, -k); s.z.e += k;
        s.e2 = ldexpf(s.e2, -2 * k);
        sync_c(s);
    }
    __device__ static void renorm(State &s) {
        renorm_z(s);
        s.d = xs_norm(s.d.x, s.d.y, s.d.e);
        s.h = xs_norm(s.h.x, s.h.y, s.h.e);
    }
    // a scaled value as a lazily normalised record (a zero mantissa takes
    // the zero exponent, so that it never sets the exponent of a sum)
    __device__ static xt tz(const xs &v) {
        return (v.x == 0.f && v.y == 0.f) ? xt{0.f, 0.f, XT_ZERO_E} : as_t(v);
    }
    // z as a binary64 value
    __device__ static cplx zd(const State &s) {
        const int32_t e = max(s.z.e, -2000);
        return mk(ldexp((double)s.z.x, e), ldexp((double)s.z.y, e));
    }
    // a new z from a binary64 value (rebase), its error carried over
    __device__ static void set_z(State &s, double re, double im) {
        const int32_t e0 = s.z.e;
        s.z = xs_from(re, im, 0);
        if (s.z.e == XS_ZERO_E) s.z.e = s.c.e;
        const float f = KE * l1(s.z.x, s.z.y);
        s.e2 = (float)ldexp((double)s.e2, max(min(2 * (e0 - s.z.e), 1000), -1100)) + f * f;
        sync_c(s);
    }
    // the error bound of z, absolute
    __device__ static double err_abs(const State &s) {
        return ERR_SIGMAS * ldexp((double)__builtin_sqrtf(s.e2), max(s.z.e, -2000));
    }

    __device__ static void init(State &s, const KArgs &a, double px, double py) {
        const double x1 = a.lin[0] * px + a.lin[1] * py;
        const double y1 = a.lin[2] * px + a.lin[3] * py;
        s.c = xs_from(a.dx_m * x1, a.dx_m * y1, a.dx_e) + xs_from(a.drift_m[0], a.drift_m[1], a.drift_e[0]);
        s.z = xs{0.f, 0.f, s.c.e == XS_ZERO_E ? 0 : s.c.e};
        s.d = xs_zero(); s.h = xs_zero();
        s.e2 = 0.f;
        sync_c(s);
        s.w = 0; s.n = 0; s.steps = 0; s.dyn = true; s.nul = false;
    }

    // one plain step on mantissas: the orbit point Zf (absolute, binary32),
    // the path values Zd (dz/dc) and Zh (dz/dz) as normalised records
    __device__ static void fast_step(State &s, const KArgs &a, float Zfx, float Zfy, const xs &Zd, const xs &Zh) {
        const float ux = ldexpf(s.z.x, s.z.e), uy = ldexpf(s.z.y, s.z.e);     // z, absolute
        const float sx = Zfx + ux, sy = Zfy + uy;                               // Z + z
        float fp1;                                                              // |f'(Z + z)|^2
        if constexpr (KIND == KIND_M2) {
            fp1 = 4.f * (sx * sx + sy * sy);                                  // |f'|^2
            // 2 ((Z + z) d + Zd z): both terms on the larger exponent
            if (a.calc_dc) {
                const int32_t E = max(s.d.e, Zd.e + s.z.e);
                const float k1 = ldexpf(1.f, max(s.d.e - E, -200)), k2 = ldexpf(1.f, max(Zd.e + s.z.e - E, -200));
                const float ax = k1 * s.d.x, ay = k1 * s.d.y, bx = k2 * Zd.x, by = k2 * Zd.y;
                s.d.x = ax * sx - ay * sy + (bx * s.z.x - by * s.z.y);
                s.d.y = ax * sy + ay * sx + (bx * s.z.y + by * s.z.x);
                s.d.e = E + 1;
            }
            if (a.calc_dz) {
                const int32_t E = max(s.h.e, Zh.e + s.z.e);
                const float k1 = ldexpf(1.f, max(s.h.e - E, -200)), k2 = ldexpf(1.f, max(Zh.e + s.z.e - E, -200));
                const float ax = k1 * s.h.x, ay = k1 * s.h.y, bx = k2 * Zh.x, by = k2 * Zh.y;
                s.h.x = ax * sx - ay * sy + (bx * s.z.x - by * s.z.y);
                s.h.y = ax * sy + ay * sx + (bx * s.z.y + by * s.z.x);
                s.h.e = E + 1;
            }
            // z (z + 2 Z) + c
            const float tx = 2.f * Zfx + ux, ty = 2.f * Zfy + uy;
            const float nx = s.z.x * tx - s.z.y * ty + s.cx;
            const float ny = s.z.x * ty + s.z.y * tx + s.cy;
            const float fr = KE * (l1(s.z.x, s.z.y) * l1(tx, ty) + l1(s.cx, s.cy));
            s.e2 = fp1 * s.e2 + fr * fr;
            s.z.x = nx; s.z.y = ny;
        } else {
            // (Z + z)^d - Z^d + c by the binomial Horner scheme of p_zn, the
            // running term in z's scale (its absolute value one ldexp away)
            float mx = ux + (float)cbinom(DEG, 1) * Zfx, my = uy + (float)cbinom(DEG, 1) * Zfy;
            float tx = s.z.x * mx - s.z.y * my, ty = s.z.x * my + s.z.y * mx;
            float px = Zfx, py = Zfy;                      // Z^(q-1)
#pragma unroll
            for (int q = 2; q < DEG; q++) {
                const float qx = px * Zfx - py * Zfy, qy = px * Zfy + py * Zfx;
                px = qx; py = qy;
                mx = ldexpf(tx, s.z.e) + (float)cbinom(DEG, q) * px;
                my = ldexpf(ty, s.z.e) + (float)cbinom(DEG, q) * py;
                tx = s.z.x * mx - s.z.y * my; ty = s.z.x * my + s.z.y * mx;
            }
            // |f'(Z + z)| = d |Z + z|^(d-1); the step's terms are bounded over
            // |z| by d (|Z| + |z|)^(d-1) ((Z + z)^d - Z^d, mean value theorem)
            float g = 1.f, gt = 1.f;
            const float ls = l2(sx, sy), lt = l1(Zfx, Zfy) + l1(ux, uy);
#pragma unroll
            for (int q = 1; q < DEG; q++) { g *= ls; gt *= lt; }
            fp1 = (float)DEG * g;
            fp1 *= fp1;
            if (a.calc_dc) s.d = dstep(s, Zd, s.d, ux, uy, Zfx, Zfy);
            if (a.calc_dz) s.h = dstep(s, Zh, s.h, ux, uy, Zfx, Zfy);
            const float fr = KE * ((float)(DEG * DEG) * gt * l1(s.z.x, s.z.y) + l1(s.cx, s.cy));
            s.e2 = fp1 * s.e2 + fr * fr;
            s.z.x = tx + s.cx; s.z.y = ty + s.cy;
        }
    }

    // derivative delta of z^d + c (p_deriv of models.hpp) on the scaled
    // mantissas: the result's exponent is T = max(e_d, e_Zd + e_z)
    __device__ static xs dstep(const State &s, const xs &Zd, const xs &d, float ux, float uy,
                               float Zfx, float Zfy) {
        const int32_t T = max(d.e, Zd.e + s.z.e);
        const float kd = ldexpf(1.f, max(d.e - T, -200)), kz = ldexpf(1.f, max(Zd.e + s.z.e - T, -200));
        const float dx = kd * d.x, dy = kd * d.y;                    // d in scale T
        const float zdx = kz * Zd.x, zdy = kz * Zd.y;
        const float gx = zdx * s.z.x - zdy * s.z.y, gy = zdx * s.z.y + zdy * s.z.x;   // Zd z in scale T
        const float C1 = (float)cbinom(DEG, 1);
        float mx = ux + C1 * Zfx, my = uy + C1 * Zfy;                // z + C1 Z
        float tmx = ux * mx - uy * my, tmy = ux * my + uy * mx;      // tmp = z mul, absolute
        // dt = d mul + z d + C1 Zd z
        float ex = dx * mx - dy * my + (ux * dx - uy * dy) + C1 * gx;
        float ey = dx * my + dy * mx + (ux * dy + uy * dx) + C1 * gy;
        float px = Zfx, py = Zfy;                                    // Z^(q-1)
#pragma unroll
        for (int q = 2; q < DEG; q++) {
            // z dZpk = q Z^(q-1) Zd z
            const float hx = (float)q * (px * gx - py * gy), hy = (float)q * (px * gy + py * gx);
            const float qx = px * Zfx - py * Zfy, qy = px * Zfy + py * Zfx;
            px = qx; py = qy;
            const float Cq = (float)cbinom(DEG, q);
            mx = tmx + Cq * px; my = tmy + Cq * py;
            const float nx = dx * mx - dy * my + (ux * ex - uy * ey) + Cq * hx;
            const float ny = dx * my + dy * mx + (ux * ey + uy * ex) + Cq * hy;
            ex = nx; ey = ny;
            tmx = ux * mx - uy * my; tmy = ux * my + uy * mx;
        }
        return xs{ex, ey, T};
    }

    __device__ static int step(State &s, const KArgs &a, Counters &k) {
        const Tables &t = a.t;
        const xs *Zx = (const xs *)a.tx.Zx, *DC = (const xs *)a.tx.DCx, *DZ = (const xs *)a.tx.DZx;
        if (++s.steps > FP32X_STEP_CAP) return REASON_UNSURE;
        if ((s.w & 7) == 0) {
            renorm(s);
            while (t.bla_on && (s.w & 7) == 0) {
                const double az = ldexp((double)hypotf(s.z.x, s.z.y), max(s.z.e, -2000));
                const double eb = err_abs(s) + 0x1p-21 * az;
                int32_t idx;
                bool unsure = false;
                const int32_t st = bla_find_m(t, az, eb, s.w, idx, unsure);
                k.tries++;
                if (unsure) return REASON_UNSURE;
                if (st == 0) break;
                // z <- A z + B c, d <- A d, h <- A h (A, B normalised records)
                const xs A = ((const xs *)a.tx.Ax)[idx], B = ((const xs *)a.tx.Bx)[idx];
                const int32_t E = max(A.e, B.e);
                const float ka = ldexpf(1.f, max(A.e - E, -200)), kb = ldexpf(1.f, max(B.e - E, -200));
                const float Ax = ka * A.x, Ay = ka * A.y, Bx = kb * B.x, By = kb * B.y;
                const float nx = Ax * s.z.x - Ay * s.z.y + (Bx * s.cx - By * s.cy);
                const float ny = Ax * s.z.y + Ay * s.z.x + (Bx * s.cy + By * s.cx);
                const float fr = KE * (l1(Ax, Ay) * l1(s.z.x, s.z.y) + l1(Bx, By) * l1(s.cx, s.cy));
                s.e2 = (Ax * Ax + Ay * Ay) * s.e2 + fr * fr;
                s.z.x = nx; s.z.y = ny; s.z.e += E;
                if (a.calc_dc) { float x, y; cmul(A.x, A.y, s.d.x, s.d.y, x, y); s.d = xs{x, y, s.d.e + A.e}; }
                if (a.calc_dz) { float x, y; cmul(A.x, A.y, s.h.x, s.h.y, x, y); s.h = xs{x, y, s.h.e + A.e}; }
                s.n += st;
                s.w += st;                      // w + st < first_invalid <= order
                renorm(s);
                k.bla++; k.bla_it += st;
            }
        }
        s.n += 1; k.plain++;
        const xs Zr = Zx[s.w];
        const xs Zd = (a.calc_dc && !s.dyn) ? DC[s.w] : xs_zero();
        const xs Zh = a.calc_dz ? (s.nul ? DZ[0] : DZ[s.w]) : xs_zero();
        constexpr int32_t LOW = (KIND == KIND_M2) ? -60 : -100 / (DEG - 1);
        if (Zr.e >= LOW) {
            fast_step(s, a, ldexpf(Zr.x, Zr.e), ldexpf(Zr.y, Zr.e), Zd, Zh);
            if (l1(s.z.x, s.z.y) > RENORM_AT) renorm_z(s);
            if (l1(s.d.x, s.d.y) > RENORM_AT) s.d = xs_norm(s.d.x, s.d.y, s.d.e);
            if (l1(s.h.x, s.h.y) > RENORM_AT) s.h = xs_norm(s.h.x, s.h.y, s.h.e);
        } else {
            // the orbit point is below binary32's comfortable range: one step
            // on lazily normalised records (products add exponents, sums align)
            const xt z0 = tz(s.z), Zt = as_t(Zr);
            const int32_t e0 = s.z.e;
            if (a.calc_dc) s.d = of_t(p_deriv<KIND, DEG>(z0, tz(s.d), Zt, as_t(Zd)));
            if (a.calc_dz) s.h = of_t(p_deriv<KIND, DEG>(z0, tz(s.h), Zt, as_t(Zh)));
            const xs nz = of_t(p_zn<KIND, DEG>(z0, Zt, as_t(s.c)));
            // |f'(Z + z)| <= d (|Z| + |z|)^(d-1), the step's terms <= |z| (|z| + d |Z|)^(d-1):
            // magnitudes in units of 2^E, E the larger of |z|'s and |Z|'s exponents
            const float mz = l2(s.z.x, s.z.y), mZ = l2(Zr.x, Zr.y);
            const int32_t ez = mz == 0.f ? XS_ZERO_E : e0 + fexp(mz);
            const int32_t eZ = mZ == 0.f ? XS_ZERO_E : Zr.e + fexp(mZ);
            const int32_t E = max(max(ez, eZ), -(1 << 24));
            const double lz = ldexp((double)mz, max(e0 - E, -1100)), lZ = ldexp((double)mZ, max(Zr.e - E, -1100));
            double g = 1., gt = 1.;
            for (int q = 1; q < DEG; q++) { g *= lz + lZ; gt *= lz + DEG * lZ; }
            s.z = (nz.e == XS_ZERO_E) ? xs{0.f, 0.f, e0} : nz;
            sync_c(s);
            // in units of z's new scale
            const double prop = ldexp(DEG * DEG * g * g * (double)s.e2,
                                      max(min(2 * ((DEG - 1) * E + e0 - s.z.e), 1000), -1100));
            const double fresh = ldexp(KE * lz * gt, max(min(DEG * E - s.z.e, 500), -1100));
            const float fr = KE * (l1(s.z.x, s.z.y) + l1(s.cx, s.cy));
            s.e2 = (float)(prop + fresh * fresh) + fr * fr;
        }
        s.w += 1;
        if (s.w >= t.order) s.w = 0;
        if (s.n >= a.max_iter) return 0;
        if (a.calc_dz) {
            const xs q = s.h + (s.nul ? DZ[0] : (s.n == t.order ? DZ[t.L] : DZ[s.w]));
            const double q2 = ldexp(xs_abs2_m(q), 2 * max(q.e, -2000));
            if (fabs(q2 - a.eps2) <= INTERIOR_MARGIN * a.eps2) return REASON_UNSURE;
            if (q2 < a.eps2) return 2;
        }
        // escape and rebase tests in binary64, on the binary64 orbit
        const cplx z64 = zd(s);
        const cplx Zn1 = t.Zn[s.w];
        const cplx ZZ = z64 + Zn1;
        const double eabs = err_abs(s);
        const double r2 = ZZ.re * ZZ.re + ZZ.im * ZZ.im;
        // an error as large as a quarter of the escape radius leaves no decision certain
        if (eabs * eabs > 0.0625 * a.M2) return REASON_UNSURE;
        if (r2 > 0.25 * a.M2) {
            // |ZZ| within its error bound of M: the binary64 loop may escape one step apart
            const double m = sqrt(a.M2);
            if (fabs(sqrt(r2) - m) <= eabs + 0x1p-50 * m) return REASON_UNSURE;
            if (r2 > a.M2) return 1;
        }
        if (s.w >= t.div_iter - 1) {
            set_z(s, ZZ.re, ZZ.im);
            if (a.calc_dc) s.d = s.d + DC[s.w];
            if (a.calc_dz) {
                if (!s.nul) s.h = s.h + (s.n == t.order ? DZ[t.L] : DZ[s.w]);
                s.nul = true;
            }
            s.w = 0;
            k.reb++;
            return -1;
        }
        // dynamic rebase: |ZZ| <= |z| component-wise; uncertain when a
        // component lies within twice the error bound of its threshold
        // (a zero orbit point leaves ZZ = z in both loops: certain)
        const double dre = fabs(ZZ.re) - fabs(z64.re), dim = fabs(ZZ.im) - fabs(z64.im);
        const double mg = 2. * eabs;
        const bool zero_ref = (Zn1.re == 0. && Zn1.im == 0.);
        if (!zero_ref && dre <= mg && dim <= mg && !(dre < -mg && dim < -mg)) return REASON_UNSURE;
        s.dyn = (dre <= 0.) && (dim <= 0.);
        if (s.dyn) {
            // the binary64 loop confirms with the extended-range |ZZ|^2 <= |z|^2,
            // which can decide otherwise only where z or the orbit point
            // underflows binary64: left to that loop
            if (fabs(z64.re) + fabs(z64.im) < 1e-290 || (zero_ref && Zx[s.w].e != XS_ZERO_E))
                return REASON_UNSURE;
            const xs ZZx = s.z + Zx[s.w];
            if (xs_abs2_le(ZZx, s.z)) {
                set_z(s, ZZ.re, ZZ.im);
                if (a.calc_dc) s.d = s.d + DC[s.w];
                if (a.calc_dz) {
                    if (!s.nul) s.h = s.h + (s.n == t.order ? DZ[t.L] : DZ[s.w]);
                    s.nul = true;
                }
                s.w = 0;
                k.reb++;
            }
        }
        return -1;
    }

    __device__ static void finish(const State &s, const KArgs &a, int64_t i, int reason) {
        a.stop_reason[i] = (int8_t)reason;
        if (reason == REASON_UNSURE) return;        // the binary64 pass writes this pixel
        const int64_t n = a.ld;
        const cplx z = zd(s) + a.t.Zn[s.w];
        a.Z[2 * i] = z.re; a.Z[2 * i + 1] = z.im;
        int f = 1;
        if (a.calc_dz) {
            const cplx h = xs_to_d(xs_norm(s.h.x, s.h.y, s.h.e));
            a.Z[2 * (n + i)] = h.re; a.Z[2 * (n + i) + 1] = h.im;
            f = 2;
        }
        if (a.calc_dc) {
            const cplx d = xs_to_d(xs_norm(s.d.x, s.d.y, s.d.e) + ((const xs *)a.tx.DCx)[s.w]);
            a.Z[2 * (f * n + i)] = d.re; a.Z[2 * (f * n + i) + 1] = d.im;
        }
        a.U[i] = s.w; a.stop_iter[i] = s.n;
    }
};

// ---------------------------------------------------------------------------
// burning-ship family: real quantities as a binary32 mantissa in [1, 2)
// (or 0) and an int32 exponent; the step and Jacobian formulas are the
// binary64 kernels' own templates (models.hpp bs_zn / bs_hess,
// burning_ship.py:535-859) instantiated on this type
struct xsr { float m; int32_t e; };

__device__ __forceinline__ xsr xsr_n(float m, int32_t e) {
    const int32_t k = fexp(m) - 1;
    xsr r;
    r.m = ldexpf(m, -k);
    r.e = (m == 0.f) ? XS_ZERO_E : e + k;
    return r;
}
__device__ __forceinline__ xsr xsr_from(double m, int32_t e) {
    const int32_t k = __builtin_amdgcn_frexp_exp(m) - 1;
    return xsr_n((float)ldexp(m, -k), (m == 0.) ? XS_ZERO_E : e + k);
}
__device__ __forceinline__ double xsr_to_d(xsr a) { return ldexp((double)a.m, max(a.e, -2000)); }
__device__ __forceinline__ xsr operator+(xsr a, xsr b) {
    const int32_t E = max(a.e, b.e);
    return xsr_n(ldexpf(a.m, max(a.e - E, -200)) + ldexpf(b.m, max(b.e - E, -200)), E);
}
__device__ __forceinline__ xsr operator-(xsr a) { a.m = -a.m; return a; }
__device__ __forceinline__ xsr operator-(xsr a, xsr b) { return a + (-b); }
__device__ __forceinline__ xsr operator*(xsr a, xsr b) { return xsr_n(a.m * b.m, a.e + b.e); }
__device__ __forceinline__ xsr operator*(double s, xsr a) { return xsr_n((float)s * a.m, a.e); }
__device__ __forceinline__ xsr mul_s(double s, xsr a) { return xsr_n((float)s * a.m, a.e); }
// s * a for a binary64 s of any magnitude: the BLA coefficients of deep
// frames leave the binary32 range (1e-800: |M| up to ~1e300), so s is split
// into a binary32 mantissa and an exponent first
__device__ __forceinline__ xsr mul_w(double s, xsr a) {
    const int32_t k = __builtin_amdgcn_frexp_exp(s) - 1;
    return xsr_n((float)ldexp(s, -k) * a.m, a.e + k);
}
__device__ __forceinline__ bool xsr_le(xsr a, xsr b) {   // a <= b
    const int32_t E = max(a.e, b.e);
    return ldexpf(a.m, max(a.e - E, -200)) <= ldexpf(b.m, max(b.e - E, -200));
}
__device__ __forceinline__ xsr absv(xsr x) { x.m = fabsf(x.m); return x; }
__device__ __forceinline__ double sgn(xsr x) { return x.m < 0.f ? -1. : 1.; }
// |X + x| - |X| and its partial derivatives (burning_ship.py:19-60)
__device__ __forceinline__ xsr diffabs(xsr X, xsr x) {
    if (X.m >= 0.f) { if ((X + x).m >= 0.f) return x; return -(2. * X + x); }
    if ((X + x).m <= 0.f) return -x;
    return 2. * X + x;
}
__device__ __forceinline__ double ddX(xsr X, xsr x) {
    if (X.m >= 0.f) return (X + x).m >= 0.f ? 0. : -2.;
    return (X + x).m <= 0.f ? 0. : 2.;
}
__device__ __forceinline__ double ddx(xsr X, xsr x) {
    if (X.m >= 0.f) return (X + x).m >= 0.f ? 1. : -1.;
    return (X + x).m <= 0.f ? -1. : 1.;
}

// Lazily normalised form of xsr for one iteration (as xzf for binary64):
// the operands start normalised, the step's products of at most three
// factors and sums of a few terms stay far inside the binary32 range, and
// the state is renormalised once per iteration.  A cancelled sum or a
// product by a zero scalar takes XS_ZERO_E.
struct xzs { float m; int32_t e; };

__device__ __forceinline__ xzs as_zs(xsr a) { xzs r; r.m = a.m; r.e = a.e; return r; }
__device__ __forceinline__ xsr xsr_of(xzs a) { return xsr_n(a.m, a.e); }
__device__ __forceinline__ xzs zs_mk(float m, int32_t e) { xzs r; r.m = m; r.e = (m == 0.f) ? XS_ZERO_E : e; return r; }
__device__ __forceinline__ xzs operator+(xzs a, xzs b) {
    const int32_t E = max(a.e, b.e);
    return zs_mk(ldexpf(a.m, max(a.e - E, -200)) + ldexpf(b.m, max(b.e - E, -200)), E);
}
__device__ __forceinline__ xzs operator-(xzs a) { a.m = -a.m; return a; }
__device__ __forceinline__ xzs operator-(xzs a, xzs b) { return a + (-b); }
__device__ __forceinline__ xzs operator*(xzs a, xzs b) { xzs r; r.m = a.m * b.m; r.e = a.e + b.e; return r; }
__device__ __forceinline__ xzs operator*(double s, xzs a) {
    xzs r; r.m = (float)s * a.m; r.e = (s == 0.) ? XS_ZERO_E : a.e; return r;
}
__device__ __forceinline__ xzs mul_s(double s, xzs a) { return s * a; }
__device__ __forceinline__ xzs mul_w(double s, xzs a) {
    const int32_t k = __builtin_amdgcn_frexp_exp(s) - 1;
    xzs r; r.m = (float)ldexp(s, -k) * a.m; r.e = (s == 0.) ? XS_ZERO_E : a.e + k; return r;
}
__device__ __forceinline__ xzs absv(xzs x) { x.m = fabsf(x.m); return x; }
__device__ __forceinline__ double sgn(xzs x) { return x.m < 0.f ? -1. : 1.; }
__device__ __forceinline__ xzs diffabs(xzs X, xzs x) {
    const xzs Xx = X + x;
    if (X.m >= 0.f) { if (Xx.m >= 0.f) return x; return -(2. * X + x); }
    if (Xx.m <= 0.f) return -x;
    return 2. * X + x;
}
__device__ __forceinline__ double ddX(xzs X, xzs x) {
    const xzs Xx = X + x;
    if (X.m >= 0.f) return Xx.m >= 0.f ? 0. : -2.;
    return Xx.m <= 0.f ? 0. : 2.;
}
__device__ __forceinline__ double ddx(xzs X, xzs x) {
    const xzs Xx = X + x;
    if (X.m >= 0.f) return Xx.m >= 0.f ? 1. : -1.;
    return Xx.m <= 0.f ? -1. : 1.;
}

// per-reference burning-ship tables in the xsr format, built once per
// context: the orbit (the Xrange register below 1e-300) and the Jacobian paths
__global__ void k_xsr_orbit(const cplx *Zn, int32_t L, const int32_t *xr_idx, const double *xr_m,
                            const int32_t *xr_e, int32_t n_xr, xsr *out) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    int32_t lo = 0, hi = n_xr;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (xr_idx[mid] < i) lo = mid + 1; else hi = mid;
    }
    if (lo < n_xr && xr_idx[lo] == i) {
        out[2 * i] = xsr_from(xr_m[2 * lo], xr_e[2 * lo]);
        out[2 * i + 1] = xsr_from(xr_m[2 * lo + 1], xr_e[2 * lo + 1]);
    } else {
        out[2 * i] = xsr_from(Zn[i].re, 0);
        out[2 * i + 1] = xsr_from(Zn[i].im, 0);
    }
}

__global__ void k_xsr_path(const double *m, const int32_t *e, int32_t n, xsr *out) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = xsr_from(m[i], e ? e[i] : 0);
}

hipError_t build_xsr_tables(const Tables &t, TablesX &x, hipStream_t stream) {
    const int tb = 256;
    hipLaunchKernelGGL(k_xsr_orbit, dim3((t.L + tb - 1) / tb), dim3(tb), 0, stream, t.Zn, t.L,
                       t.xr_idx, t.xr_m, t.xr_e, t.n_xr, (xsr *)x.ZRx);
    for (int q = 0; q < 4; q++)
        if (t.bs[q] && x.BSx[q])
            hipLaunchKernelGGL(k_xsr_path, dim3((t.L + tb - 1) / tb), dim3(tb), 0, stream,
                               t.bs[q], t.bs_e[q], t.L, (xsr *)x.BSx[q]);
    return hipGetLastError();
}

template <int FL>
struct BsF32 {
    struct State { xsr x, y, P[4], a, b; double xs, ys; int32_t w, n; bool dyn; };

    // the tables hold xsr_from of the binary64 values (build_xsr_tables)
    __device__ static xsr PATH(const KArgs &a, int q, int32_t i) { return static_cast<const xsr *>(a.tx.BSx[q])[i]; }
    __device__ static void ZR(const KArgs &a, int32_t i, xsr &X, xsr &Y) {
        const xsr *z = static_cast<const xsr *>(a.tx.ZRx);
        X = z[2 * i]; Y = z[2 * i + 1];
    }

    __device__ static void init(State &s, const KArgs &a, double px, double py) {
        s.a = xsr_from(a.dx_m * (a.lin[0] * px + a.lin[1] * py), a.dx_e) + xsr_from(a.drift_m[0], a.drift_e[0]);
        s.b = xsr_from(a.dx_m * (a.lin[2] * px + a.lin[3] * py), a.dx_e) + xsr_from(a.drift_m[1], a.drift_e[1]);
        const xsr z0 = xsr{0.f, XS_ZERO_E};
        s.x = z0; s.y = z0; s.P[0] = s.P[1] = s.P[2] = s.P[3] = z0;
        s.xs = 0.; s.ys = 0.;
        s.w = 0; s.n = 0; s.dyn = true;
    }

    __device__ static int step(State &s, const KArgs &a, Counters &k) {
        const Tables &t = a.t;
        while (t.bla_on && (s.w & 7) == 0) {
            int32_t idx;
            const int32_t st = bla_find(t, hypot2(s.xs, s.ys), s.w, idx);
            k.tries++;
            if (st == 0) break;
            {
                const double *M = t.M + 8 * idx;
                s.n += st; s.w += st;
                const xzs x = as_zs(s.x), y = as_zs(s.y), ca = as_zs(s.a), cb = as_zs(s.b);
                s.x = xsr_of(mul_w(M[0], x) + mul_w(M[1], y) + mul_w(M[4], ca) + mul_w(M[5], cb));
                s.y = xsr_of(mul_w(M[2], x) + mul_w(M[3], y) + mul_w(M[6], ca) + mul_w(M[7], cb));
                s.xs = xsr_to_d(s.x); s.ys = xsr_to_d(s.y);
                if (a.calc_dc) {
                    const xzs P0 = as_zs(s.P[0]), P1 = as_zs(s.P[1]), P2 = as_zs(s.P[2]), P3 = as_zs(s.P[3]);
                    s.P[0] = xsr_of(mul_w(M[0], P0) + mul_w(M[1], P2));
                    s.P[1] = xsr_of(mul_w(M[0], P1) + mul_w(M[1], P3));
                    s.P[2] = xsr_of(mul_w(M[2], P0) + mul_w(M[3], P2));
                    s.P[3] = xsr_of(mul_w(M[2], P1) + mul_w(M[3], P3));
                }
                k.bla++; k.bla_it += st;
            }
        }
        s.n += 1; k.plain++;
        xsr X, Y;
        ZR(a, s.w, X, Y);
        // one iteration on lazily normalised values, renormalised once
        const xzs x = as_zs(s.x), y = as_zs(s.y), Xz = as_zs(X), Yz = as_zs(Y);
        if (a.calc_dc) {
            const xzs z0 = xzs{0.f, XS_ZERO_E};
            xzs Rd[4] = {z0, z0, z0, z0};
            if (!s.dyn) for (int q = 0; q < 4; q++) Rd[q] = as_zs(PATH(a, q, s.w));
            xzs P[4] = {as_zs(s.P[0]), as_zs(s.P[1]), as_zs(s.P[2]), as_zs(s.P[3])};
            bs_hess<FL>(x, y, P, Xz, Yz, Rd);
            for (int q = 0; q < 4; q++) s.P[q] = xsr_of(P[q]);
        }
        xzs nx = x, ny = y;
        bs_zn<FL>(nx, ny, Xz, Yz, as_zs(s.a), as_zs(s.b));
        s.x = xsr_of(nx); s.y = xsr_of(ny);
        s.xs = xsr_to_d(s.x); s.ys = xsr_to_d(s.y);
        if (s.n >= a.max_iter) return 0;
        s.w += 1;
        if (s.w >= t.order) s.w = 0;
        // escape and rebase tests in binary64, on the binary64 orbit
        const cplx R1 = t.Zn[s.w];
        const double XX = s.xs + R1.re, YY = s.ys + R1.im;
        if (XX * XX + YY * YY > a.M2) return 1;
        if (s.w >= t.div_iter - 1) {
            s.xs = XX; s.ys = YY;
            s.x = xsr_from(XX, 0); s.y = xsr_from(YY, 0);
            if (a.calc_dc) for (int q = 0; q < 4; q++) s.P[q] = s.P[q] + PATH(a, q, s.w);
            s.w = 0;
            k.reb++;
            return -1;
        }
        s.dyn = (fabs(XX) <= fabs(s.xs)) && (fabs(YY) <= fabs(s.ys));
        if (s.dyn) {
            xsr X1, Y1;
            ZR(a, s.w, X1, Y1);
            const xsr XXx = s.x + X1, YYx = s.y + Y1;
            if (xsr_le(XXx * XXx + YYx * YYx, s.x * s.x + s.y * s.y)) {
                s.x = XXx; s.y = YYx;
                s.xs = xsr_to_d(XXx); s.ys = xsr_to_d(YYx);
                if (a.calc_dc) for (int q = 0; q < 4; q++) s.P[q] = s.P[q] + PATH(a, q, s.w);
                s.w = 0;
                k.reb++;
            }
        }
        return -1;
    }

    __device__ static void finish(const State &s, const KArgs &a, int64_t i, int reason) {
        const int64_t n = a.ld;
        const cplx R = a.t.Zn[s.w];
        a.Z[i] = s.xs + R.re; a.Z[n + i] = s.ys + R.im;
        if (a.calc_dc)
            for (int q = 0; q < 4; q++) a.Z[(2 + q) * n + i] = xsr_to_d(s.P[q] + PATH(a, q, s.w));
        a.U[i] = s.w; a.stop_iter[i] = s.n; a.stop_reason[i] = (int8_t)reason;
    }
};

typedef void (*tile_fn)(const KArgs);
template <class Model> static tile_fn fn_of() { return k_tile<Model>; }

// The burning-ship binary32 loop with the tables its step reads most staged
// in LDS: the xsr orbit pairs, the binary64 orbit of the escape tests and
// the BLA radii and coefficients (the Jacobian paths stay in global memory).
// Every staged pointer is an LDS address, every other one global.
template <int FL>
__global__ void __launch_bounds__(1024) k_tile_lds_bsx(const KArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const LdsPlanX &p = a.ldsx;
    KArgs al = a;
    const int64_t K = p.K, KB = p.KB;
    al.t.Zn = stage<true>(lds, p.zn, a.t.Zn, K);
    al.t.r = stage<true>(lds, p.r, a.t.r, KB);
    al.t.M = stage<true>(lds, p.ax, a.t.M, KB * 8);     // the plan's ax slot: binary64 coefficients
    al.tx.ZRx = stage<true>(lds, p.zx, static_cast<const xsr *>(a.tx.ZRx), 2 * K);
    __syncthreads();
    tile_loop<BsF32<FL>>(al);
}

static LdsPlanX plan_lds_bsx(const Tables &t, int max_bytes) {
    LdsPlanX p;
    std::memset(&p, 0xff, sizeof(p));
    int64_t K = std::min<int64_t>(t.L, std::min<int64_t>((int64_t)t.div_iter + 1, t.order));
    K = std::max<int64_t>(K, std::min<int64_t>(t.L, 2));
    const int64_t bla_len = 2 * (int64_t)(t.L >> 3);
    const int64_t KB = t.bla_on ? std::min<int64_t>(bla_len, K / 4 + 2) : 0;
    int64_t off = 0;
    const auto take = [&](int64_t bytes) { int64_t o = off; off += (bytes + 15) & ~(int64_t)15; return (int32_t)o; };
    p.zx = take((int64_t)XSR_BYTES * 2 * K);
    p.zn = take(16 * K);
    if (KB > 0) { p.r = take(8 * KB); p.ax = take(64 * KB); }
    p.K = (int32_t)K; p.KB = (int32_t)KB; p.full = 0;
    p.bytes = (off <= max_bytes && (!t.bla_on || KB > 0)) ? (int32_t)off : 0;
    return p;
}

static tile_fn select_bsx_lds(const LaunchCfg &c) {
    switch (c.flavor) {
        case 1: return k_tile_lds_bsx<1>;
        case 2: return k_tile_lds_bsx<2>;
        case 3: return k_tile_lds_bsx<3>;
        case 4: return k_tile_lds_bsx<4>;
        case 5: return k_tile_lds_bsx<5>;
        default: return nullptr;
    }
}

static tile_fn select_f32(const LaunchCfg &c) {
    if (c.kind == KIND_M2) return fn_of<HoloS32<KIND_M2, 2>>();
    if (c.kind == KIND_BS) {
        switch (c.flavor) {
            case 1: return fn_of<BsF32<1>>();
            case 2: return fn_of<BsF32<2>>();
            case 3: return fn_of<BsF32<3>>();
            case 4: return fn_of<BsF32<4>>();
            case 5: return fn_of<BsF32<5>>();
            default: return nullptr;
        }
    }
    if (c.kind == KIND_MN) {
        switch (c.degree) {
#define FSG_MN(D) case D: return fn_of<HoloS32<KIND_MN, D>>();
            FSG_MN(2) FSG_MN(3) FSG_MN(4) FSG_MN(5) FSG_MN(6) FSG_MN(7) FSG_MN(8)
#undef FSG_MN
            default: return nullptr;
        }
    }
    return nullptr;
}

// The holomorphic binary32 loop with every table it reads staged in LDS
// (the binary64 k_tile_lds<Model, true> for this form's tables): binary64
// orbit and BLA radii for the tests, xs orbit, derivative paths (the
// wrapped dZndz value at [K]) and BLA coefficients.
template <class Model, bool FULL>
__global__ void __launch_bounds__(1024) k_tile_lds_x(const KArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const LdsPlanX &p = a.ldsx;
    KArgs al = a;
    Tables &t = al.t;
    TablesX &x = al.tx;
    const int64_t K = p.K, KB = p.KB;
    t.Zn = stage<FULL>(lds, p.zn, a.t.Zn, K);
    t.r = stage<FULL>(lds, p.r, a.t.r, KB);
    x.Zx = stage<FULL>(lds, p.zx, static_cast<const xs *>(a.tx.Zx), K);
    x.DCx = stage<FULL>(lds, p.dcx, static_cast<const xs *>(a.tx.DCx), K);
    x.Ax = stage<FULL>(lds, p.ax, static_cast<const xs *>(a.tx.Ax), KB);
    x.Bx = stage<FULL>(lds, p.bx, static_cast<const xs *>(a.tx.Bx), KB);
    if (!FULL && p.dzx < 0) {
        // read from global memory
    } else if (a.tx.DZx) {
        const xs *src = static_cast<const xs *>(a.tx.DZx);
        xs *dz = reinterpret_cast<xs *>(lds + p.dzx);
        for (int64_t i = threadIdx.x; i < K; i += blockDim.x) dz[i] = src[i];
        if (threadIdx.x == 0) dz[K] = src[a.t.L];
        x.DZx = dz;
        t.L = (int32_t)K;
    } else {
        x.DZx = nullptr;
    }
    __syncthreads();
    tile_loop<Model>(al);
}

static LdsPlanX plan_lds_x(const Tables &t, const TablesX &x, int max_bytes) {
    LdsPlanX p;
    std::memset(&p, 0xff, sizeof(p));
    int64_t K = std::min<int64_t>(t.L, std::min<int64_t>((int64_t)t.div_iter + 1, t.order));
    K = std::max<int64_t>(K, std::min<int64_t>(t.L, 2));
    const int64_t bla_len = 2 * (int64_t)(t.L >> 3);
    const int64_t KB = (t.bla_on && x.Ax) ? std::min<int64_t>(bla_len, K / 4 + 2) : 0;
    int64_t off = 0;
    const auto take = [&](int64_t bytes) { int64_t o = off; off += (bytes + 15) & ~(int64_t)15; return (int32_t)o; };
    p.zn = take(16 * K);
    p.zx = take((int64_t)XS_BYTES * K);
    if (x.DCx) p.dcx = take((int64_t)XS_BYTES * K);
    if (x.DZx) p.dzx = take((int64_t)XS_BYTES * (K + 1));
    if (KB > 0) { p.r = take(8 * KB); p.ax = take((int64_t)XS_BYTES * KB); p.bx = take((int64_t)XS_BYTES * KB); }
    p.K = (int32_t)K; p.KB = (int32_t)KB;
    if (off <= max_bytes && (!t.bla_on || KB > 0)) { p.bytes = (int32_t)off; p.full = 1; return p; }
    // partial: the xs orbit and the BLA tables, the rest from global memory
    std::memset(&p, 0xff, sizeof(p));
    off = 0;
    p.zx = take((int64_t)XS_BYTES * K);
    if (KB > 0) { p.r = take(8 * KB); p.ax = take((int64_t)XS_BYTES * KB); p.bx = take((int64_t)XS_BYTES * KB); }
    p.K = (int32_t)K; p.KB = (int32_t)KB;
    p.bytes = (off <= max_bytes) ? (int32_t)off : 0;
    p.full = 0;
    return p;
}

template <class Model> static tile_fn fn_lds_x(bool full) {
    return full ? k_tile_lds_x<Model, true> : k_tile_lds_x<Model, false>;
}

static tile_fn select_f32_lds(const LaunchCfg &c, bool full) {
    if (c.kind == KIND_M2) return fn_lds_x<HoloS32<KIND_M2, 2>>(full);
    if (c.kind == KIND_MN) {
        switch (c.degree) {
#define FSG_MN(D) case D: return fn_lds_x<HoloS32<KIND_MN, D>>(full);
            FSG_MN(2) FSG_MN(3) FSG_MN(4) FSG_MN(5) FSG_MN(6) FSG_MN(7) FSG_MN(8)
#undef FSG_MN
            default: return nullptr;
        }
    }
    return nullptr;
}

hipError_t launch_tile_f32(const LaunchCfg &cfg, const KArgs &a, hipStream_t stream) {
    // FSG_TABLES=global keeps the tables in global memory (for comparisons)
    const char *tables = std::getenv("FSG_TABLES");
    const bool no_lds = tables && !std::strcmp(tables, "global");
    if (!cfg.orbit && !no_lds && (cfg.kind != KIND_BS || a.tx.ZRx)) {
        KArgs al = a;
        const bool bs = cfg.kind == KIND_BS;
        al.ldsx = bs ? plan_lds_bsx(a.t, max_lds_bytes()) : plan_lds_x(a.t, a.tx, max_lds_bytes());
        tile_fn fl = al.ldsx.bytes <= 0 ? nullptr
                     : bs ? select_bsx_lds(cfg) : select_f32_lds(cfg, al.ldsx.full != 0);
        const int fill = fl ? fill_blocks(reinterpret_cast<const void *>(fl), 1024, al.ldsx.bytes) : -1;
        if (fill > 0) {
            hipLaunchKernelGGL(fl, dim3(launch_blocks(fill, al, 1024)), dim3(1024),
                               (size_t)al.ldsx.bytes, stream, al);
            return hipGetLastError();
        }
    }
    tile_fn f = select_f32(cfg);
    if (!f) return hipErrorInvalidValue;
    const int fill = grid_fill_blocks(reinterpret_cast<const void *>(f));
    if (fill <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(f, dim3(launch_blocks(fill, a)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

bool f32_supported(const LaunchCfg &cfg) { return select_f32(cfg) != nullptr; }

}  // namespace fsg
