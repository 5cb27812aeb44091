// fsg_kernels.hip -- gfx950 kernels for the per-pixel perturbation loop.
//
// k_tile: persistent, work-stealing tile kernel.  Each lane owns one pixel at
// a time and runs the reference's per-pixel loop (BLA jump, perturbation step,
// max_iter / interior / escape tests, rebasing); when its pixel stops it
// writes the results and takes the next pixel index from a global cursor.
// Refill is wave-aggregated: one atomic per wave hands out consecutive indices
// to all lanes that need work (ballot + mbcnt).  The arithmetic is scalar
// complex recurrences (no MFMA); the kernel is bound by the FP64 VALU rate.
//
// Reference loops restated here:
//   holomorphic: numba_iterate, src/fractalshades/perturbation.py:1065-1400
//   burning ship: numba_iterate_BS, perturbation.py:1468-1791
//   model kernels: models/mandelbrot_M2.py:599-622, mandelbrot_Mn.py:646-729,
//                  burning_ship.py:12-60, 440-859
//   BLA: numba_make_BLA* / combine_BLA* / ref_BLA_get, perturbation.py:1819-2170
#include "models.hpp"
#include "tile_loop.hpp"
#include "xl.hpp"

#include <mutex>
#include <unordered_map>

namespace fsg {

// BLA lookup on binary64 |z| (ref_BLA_get, perturbation.py:2116-2170)
__device__ __forceinline__ int32_t bla_lookup(const Tables &t, double zre, double zim,
                                              int32_t w, int32_t &idx) {
    return bla_find(t, hypot2(zre, zim), w, idx);
}

// ---- pixel -> c (ref_path_c_from_pix + lin_proj_impl, perturbation.py:2214-2230, 2665-2672)
FSG_HD xc c_from_pix(const KArgs &a, double px, double py) {
    double x1 = a.lin[0] * px + a.lin[1] * py;
    double y1 = a.lin[2] * px + a.lin[3] * py;
    xc cx = renorm_c(a.dx_m * x1, a.dx_m * y1, a.dx_e);
    return cx + mkc(a.drift_m[0], a.drift_m[1], a.drift_e[0]);
}

FSG_HD cplx zc() { return mk(0., 0.); }
FSG_HD xc zx() { return mkc(0., 0., 0); }

__device__ __forceinline__ xl c_from_pix_xl(const KArgs &a, double px, double py) {
    const double x1 = a.lin[0] * px + a.lin[1] * py;
    const double y1 = a.lin[2] * px + a.lin[3] * py;
    return xl_of(a.dx_m * x1, a.dx_m * y1, a.dx_e) + xl_of(a.drift_m[0], a.drift_m[1], a.drift_e[0]);
}

// --------------------------------------------------------------------------
// Holomorphic models, binary64 deltas
template <int KIND, int DEG>
struct HoloStd {
    struct State { cplx z, d, h, c; int32_t w, n; bool dyn, nul; };

    __device__ static void init(State &s, const KArgs &a, int64_t i) {
        s.c = to_std(c_from_pix(a, a.pix[2 * i], a.pix[2 * i + 1]));
        s.z = zc(); s.d = zc(); s.h = zc();
        s.w = 0; s.n = 0; s.dyn = true; s.nul = false;
    }

    __device__ static int step(State &s, const KArgs &a, Counters &k) {
        const Tables &t = a.t;
        while (t.bla_on && (s.w & 7) == 0) {
            int32_t idx;
            int32_t st = bla_lookup(t, s.z.re, s.z.im, s.w, idx);
            k.tries++;
            if (st == 0) break;
            {
                const double *Mb = t.M + 4 * idx;
                cplx A = mk(Mb[0], Mb[1]), B = mk(Mb[2], Mb[3]);
                s.n += st;
                s.w += st;                      // w + st < first_invalid <= order
                s.z = A * s.z + B * s.c;
                if (a.calc_dc) s.d = A * s.d;
                if (a.calc_dz) s.h = A * s.h;
                k.bla++; k.bla_it += st;
            }
        }
        s.n += 1; k.plain++;
        const cplx Zr = t.Zn[s.w];
        if (a.calc_dc) s.d = p_deriv<KIND, DEG>(s.z, s.d, Zr, s.dyn ? zc() : t.dc[s.w]);
        if (a.calc_dz) s.h = p_deriv<KIND, DEG>(s.z, s.h, Zr, s.nul ? t.dz[0] : t.dz[s.w]);
        s.z = p_zn<KIND, DEG>(s.z, Zr, s.c);
        s.w += 1;
        if (s.w >= t.order) s.w = 0;
        if (s.n >= a.max_iter) return 0;
        if (a.calc_dz) {
            cplx nx = s.nul ? t.dz[0] : (s.n == t.order ? t.dz[t.L] : t.dz[s.w]);
            cplx q = s.h + nx;
            if (q.re * q.re + q.im * q.im < a.eps2) return 2;
        }
        const cplx ZZ = s.z + t.Zn[s.w];
        if (ZZ.re * ZZ.re + ZZ.im * ZZ.im > a.M2) return 1;
        bool reb = (s.w >= t.div_iter - 1);
        if (!reb) {
            s.dyn = (fabs(ZZ.re) <= fabs(s.z.re)) && (fabs(ZZ.im) <= fabs(s.z.im));
            reb = s.dyn;
        }
        if (reb) {
            s.z = ZZ;
            if (a.calc_dc) s.d = s.d + t.dc[s.w];
            if (a.calc_dz) {
                if (!s.nul) s.h = s.h + (s.n == t.order ? t.dz[t.L] : t.dz[s.w]);
                s.nul = true;
            }
            s.w = 0;
            k.reb++;
        }
        return -1;
    }

    __device__ static void finish(const State &s, const KArgs &a, int64_t i, int reason) {
        const int64_t n = a.ld;
        cplx z = s.z + a.t.Zn[s.w];
        a.Z[2 * i] = z.re; a.Z[2 * i + 1] = z.im;
        int f = 1;
        if (a.calc_dz) { a.Z[2 * (n + i)] = s.h.re; a.Z[2 * (n + i) + 1] = s.h.im; f = 2; }
        if (a.calc_dc) {
            cplx d = s.d + a.t.dc[s.w];
            a.Z[2 * (f * n + i)] = d.re; a.Z[2 * (f * n + i) + 1] = d.im;
        }
        a.U[i] = s.w; a.stop_iter[i] = s.n; a.stop_reason[i] = (int8_t)reason;
    }
};

// Holomorphic models, Xrange deltas (dx < 1e-300).  The deltas are kept in
// the lazily normalised extended-range form of xl.hpp: every result is
// renormalised without branches, which rounds exactly like the reference's
// Xrange record (a rescaling by a power of two is exact).
template <int KIND, int DEG>
struct HoloXr {
    struct State { xl z, d, h, c; cplx zs; int32_t w, n; bool dyn, nul; };

    __device__ static xl DC(const Tables &t, int32_t i) { const cplx m = t.dc[i]; return xl_of(m.re, m.im, t.dc_e[i]); }
    __device__ static xl DZ(const Tables &t, int32_t i) { const cplx m = t.dz[i]; return xl_of(m.re, m.im, t.dz_e[i]); }
    // orbit point i: the register's Xrange value where the binary64 orbit
    // underflows, the binary64 value elsewhere
    __device__ static xl ZR(const Tables &t, int32_t i) {
        const int32_t j = t.xr_pos ? t.xr_pos[i] : -1;
        if (j >= 0) return xl_of(t.xr_m[2 * j], t.xr_m[2 * j + 1], t.xr_e[j]);
        return xl_of(t.Zn[i]);
    }

    __device__ static void init(State &s, const KArgs &a, int64_t i) {
        s.c = c_from_pix_xl(a, a.pix[2 * i], a.pix[2 * i + 1]);
        s.z = xl_zero(); s.d = xl_zero(); s.h = xl_zero();
        s.zs = zc();
        s.w = 0; s.n = 0; s.dyn = true; s.nul = false;
    }

    __device__ static int step(State &s, const KArgs &a, Counters &k) {
        const Tables &t = a.t;
        while (t.bla_on && (s.w & 7) == 0) {
            int32_t idx;
            int32_t st = bla_lookup(t, s.zs.re, s.zs.im, s.w, idx);
            k.tries++;
            if (st == 0) break;
            {
                const double *Mb = t.M + 4 * idx;
                cplx A = mk(Mb[0], Mb[1]), B = mk(Mb[2], Mb[3]);
                s.n += st;
                s.w += st;
                s.z = s.z * A + s.c * B;
                s.zs = to_std(s.z);
                if (a.calc_dc) s.d = s.d * A;
                if (a.calc_dz) s.h = s.h * A;
                k.bla++; k.bla_it += st;
            }
        }
        s.n += 1; k.plain++;
        const xl Zr = ZR(t, s.w);
        if (a.calc_dc) s.d = p_deriv<KIND, DEG>(s.z, s.d, Zr, s.dyn ? xl_zero() : DC(t, s.w));
        if (a.calc_dz) s.h = p_deriv<KIND, DEG>(s.z, s.h, Zr, s.nul ? DZ(t, 0) : DZ(t, s.w));
        s.z = p_zn<KIND, DEG>(s.z, Zr, s.c);
        s.zs = to_std(s.z);
        s.w += 1;
        if (s.w >= t.order) s.w = 0;
        if (s.n >= a.max_iter) return 0;
        if (a.calc_dz) {
            xl nx = s.nul ? DZ(t, 0) : (s.n == t.order ? DZ(t, t.L) : DZ(t, s.w));
            if (lt(abs2(s.h + nx), a.eps2)) return 2;
        }
        const cplx ZZ = s.zs + t.Zn[s.w];
        if (ZZ.re * ZZ.re + ZZ.im * ZZ.im > a.M2) return 1;
        if (s.w >= t.div_iter - 1) {
            s.zs = ZZ;
            s.z = xl_of(ZZ);
            if (a.calc_dc) s.d = s.d + DC(t, s.w);
            if (a.calc_dz) {
                if (!s.nul) s.h = s.h + (s.n == t.order ? DZ(t, t.L) : DZ(t, s.w));
                s.nul = true;
            }
            s.w = 0;
            k.reb++;
            return -1;
        }
        s.dyn = (fabs(ZZ.re) <= fabs(s.zs.re)) && (fabs(ZZ.im) <= fabs(s.zs.im));
        if (s.dyn) {
            const xl ZZx = s.z + ZR(t, s.w);
            if (le(abs2(ZZx), abs2(s.z))) {
                s.z = ZZx;
                s.zs = to_std(ZZx);
                if (a.calc_dc) s.d = s.d + DC(t, s.w);
                if (a.calc_dz) {
                    if (!s.nul) s.h = s.h + (s.n == t.order ? DZ(t, t.L) : DZ(t, s.w));
                    s.nul = true;
                }
                s.w = 0;
                k.reb++;
            }
        }
        return -1;
    }

    __device__ static void finish(const State &s, const KArgs &a, int64_t i, int reason) {
        const int64_t n = a.ld;
        cplx z = to_std(s.z) + a.t.Zn[s.w];
        a.Z[2 * i] = z.re; a.Z[2 * i + 1] = z.im;
        int f = 1;
        // perturbation.py:1377-1380 writes zn and dzndc back only; the
        // standard dzndz field keeps the 0 written by numba_initialize
        if (a.calc_dz) { a.Z[2 * (n + i)] = 0.; a.Z[2 * (n + i) + 1] = 0.; f = 2; }
        if (a.calc_dc) {
            cplx d = to_std(s.d + DC(a.t, s.w));
            a.Z[2 * (f * n + i)] = d.re; a.Z[2 * (f * n + i) + 1] = d.im;
        }
        a.U[i] = s.w; a.stop_iter[i] = s.n; a.stop_reason[i] = (int8_t)reason;
    }
};

// --------------------------------------------------------------------------
// Burning-ship family, binary64 deltas
template <int FL>
struct BsStd {
    struct State { double x, y, P[4], a, b; int32_t w, n; bool dyn; };

    __device__ static void init(State &s, const KArgs &a, int64_t i) {
        xc cx = renorm_c(a.dx_m * (a.lin[0] * a.pix[2 * i] + a.lin[1] * a.pix[2 * i + 1]),
                         a.dx_m * (a.lin[2] * a.pix[2 * i] + a.lin[3] * a.pix[2 * i + 1]), a.dx_e);
        s.a = to_std(mkf(cx.re, cx.e) + mkf(a.drift_m[0], a.drift_e[0]));
        s.b = to_std(mkf(cx.im, cx.e) + mkf(a.drift_m[1], a.drift_e[1]));
        s.x = 0.; s.y = 0.; s.P[0] = s.P[1] = s.P[2] = s.P[3] = 0.;
        s.w = 0; s.n = 0; s.dyn = true;
    }

    __device__ static int step(State &s, const KArgs &a, Counters &k) {
        const Tables &t = a.t;
        while (t.bla_on && (s.w & 7) == 0) {
            int32_t idx;
            int32_t st = bla_lookup(t, s.x, s.y, s.w, idx);
            k.tries++;
            if (st == 0) break;
            {
                const double *M = t.M + 8 * idx;
                s.n += st; s.w += st;
                double nx = M[0] * s.x + M[1] * s.y + M[4] * s.a + M[5] * s.b;
                double ny = M[2] * s.x + M[3] * s.y + M[6] * s.a + M[7] * s.b;
                s.x = nx; s.y = ny;
                if (a.calc_dc) {
                    double p0 = M[0] * s.P[0] + M[1] * s.P[2];
                    double p1 = M[0] * s.P[1] + M[1] * s.P[3];
                    double p2 = M[2] * s.P[0] + M[3] * s.P[2];
                    double p3 = M[2] * s.P[1] + M[3] * s.P[3];
                    s.P[0] = p0; s.P[1] = p1; s.P[2] = p2; s.P[3] = p3;
                }
                k.bla++; k.bla_it += st;
            }
        }
        s.n += 1; k.plain++;
        const cplx R = t.Zn[s.w];
        if (a.calc_dc) {
            double Rd[4] = {0., 0., 0., 0.};
            if (!s.dyn) { Rd[0] = t.bs[0][s.w]; Rd[1] = t.bs[1][s.w]; Rd[2] = t.bs[2][s.w]; Rd[3] = t.bs[3][s.w]; }
            bs_hess<FL>(s.x, s.y, s.P, R.re, R.im, Rd);
        }
        bs_zn<FL>(s.x, s.y, R.re, R.im, s.a, s.b);
        if (s.n >= a.max_iter) return 0;
        s.w += 1;
        if (s.w >= t.order) s.w = 0;
        const cplx R1 = t.Zn[s.w];
        double XX = s.x + R1.re, YY = s.y + R1.im;
        if (XX * XX + YY * YY > a.M2) return 1;
        bool reb = (s.w >= t.div_iter - 1);
        if (!reb) {
            s.dyn = (fabs(XX) <= fabs(s.x)) && (fabs(YY) <= fabs(s.y));
            reb = s.dyn;
        }
        if (reb) {
            s.x = XX; s.y = YY;
            if (a.calc_dc)
                for (int q = 0; q < 4; q++) s.P[q] += t.bs[q][s.w];
            s.w = 0;
            k.reb++;
        }
        return -1;
    }

    __device__ static void finish(const State &s, const KArgs &a, int64_t i, int reason) {
        const int64_t n = a.ld;
        const cplx R = a.t.Zn[s.w];
        a.Z[i] = s.x + R.re; a.Z[n + i] = s.y + R.im;
        if (a.calc_dc)
            for (int q = 0; q < 4; q++) a.Z[(2 + q) * n + i] = s.P[q] + a.t.bs[q][s.w];
        a.U[i] = s.w; a.stop_iter[i] = s.n; a.stop_reason[i] = (int8_t)reason;
    }
};

// Burning-ship family, Xrange deltas (xl.hpp reals)
template <int FL>
struct BsXr {
    struct State { xlf x, y, P[4], a, b; double xs, ys; int32_t w, n; bool dyn; };

    __device__ static xlf PATH(const Tables &t, int q, int32_t i) { return xlf_of(t.bs[q][i], t.bs_e[q][i]); }
    __device__ static void ZR(const Tables &t, int32_t i, xlf &X, xlf &Y) {
        const int32_t j = t.xr_pos ? t.xr_pos[i] : -1;
        if (j >= 0) { X = xlf_of(t.xr_m[2 * j], t.xr_e[2 * j]); Y = xlf_of(t.xr_m[2 * j + 1], t.xr_e[2 * j + 1]); }
        else { const cplx R = t.Zn[i]; X = xlf_of(R.re, 0); Y = xlf_of(R.im, 0); }
    }

    __device__ static void init(State &s, const KArgs &a, int64_t i) {
        const double px = a.pix[2 * i], py = a.pix[2 * i + 1];
        s.a = xlf_of(a.dx_m * (a.lin[0] * px + a.lin[1] * py), a.dx_e) + xlf_of(a.drift_m[0], a.drift_e[0]);
        s.b = xlf_of(a.dx_m * (a.lin[2] * px + a.lin[3] * py), a.dx_e) + xlf_of(a.drift_m[1], a.drift_e[1]);
        const xlf z0 = xlf_zero();
        s.x = z0; s.y = z0; s.P[0] = s.P[1] = s.P[2] = s.P[3] = z0;
        s.xs = 0.; s.ys = 0.;
        s.w = 0; s.n = 0; s.dyn = true;
    }

    __device__ static int step(State &s, const KArgs &a, Counters &k) {
        const Tables &t = a.t;
        while (t.bla_on && (s.w & 7) == 0) {
            int32_t idx;
            int32_t st = bla_lookup(t, s.xs, s.ys, s.w, idx);
            k.tries++;
            if (st == 0) break;
            {
                const double *M = t.M + 8 * idx;
                s.n += st; s.w += st;
                xlf nx = M[0] * s.x + M[1] * s.y + M[4] * s.a + M[5] * s.b;
                xlf ny = M[2] * s.x + M[3] * s.y + M[6] * s.a + M[7] * s.b;
                s.x = nx; s.y = ny;
                s.xs = to_std(s.x); s.ys = to_std(s.y);
                if (a.calc_dc) {
                    xlf p0 = M[0] * s.P[0] + M[1] * s.P[2];
                    xlf p1 = M[0] * s.P[1] + M[1] * s.P[3];
                    xlf p2 = M[2] * s.P[0] + M[3] * s.P[2];
                    xlf p3 = M[2] * s.P[1] + M[3] * s.P[3];
                    s.P[0] = p0; s.P[1] = p1; s.P[2] = p2; s.P[3] = p3;
                }
                k.bla++; k.bla_it += st;
            }
        }
        s.n += 1; k.plain++;
        xlf X, Y;
        ZR(t, s.w, X, Y);
        if (a.calc_dc) {
            const xlf z0 = xlf_zero();
            xlf Rd[4] = {z0, z0, z0, z0};
            if (!s.dyn) for (int q = 0; q < 4; q++) Rd[q] = PATH(t, q, s.w);
            bs_hess<FL>(s.x, s.y, s.P, X, Y, Rd);
        }
        bs_zn<FL>(s.x, s.y, X, Y, s.a, s.b);
        s.xs = to_std(s.x); s.ys = to_std(s.y);
        if (s.n >= a.max_iter) return 0;
        s.w += 1;
        if (s.w >= t.order) s.w = 0;
        const cplx R1 = t.Zn[s.w];
        double XX = s.xs + R1.re, YY = s.ys + R1.im;
        if (XX * XX + YY * YY > a.M2) return 1;
        if (s.w >= t.div_iter - 1) {
            s.xs = XX; s.ys = YY;
            s.x = xlf_of(XX, 0); s.y = xlf_of(YY, 0);
            if (a.calc_dc) for (int q = 0; q < 4; q++) s.P[q] = s.P[q] + PATH(t, q, s.w);
            s.w = 0;
            k.reb++;
            return -1;
        }
        s.dyn = (fabs(XX) <= fabs(s.xs)) && (fabs(YY) <= fabs(s.ys));
        if (s.dyn) {
            xlf X1, Y1;
            ZR(t, s.w, X1, Y1);
            const xlf XXx = s.x + X1, YYx = s.y + Y1;
            if (le(XXx * XXx + YYx * YYx, s.x * s.x + s.y * s.y)) {
                s.x = XXx; s.y = YYx;
                s.xs = to_std(XXx); s.ys = to_std(YYx);
                if (a.calc_dc) for (int q = 0; q < 4; q++) s.P[q] = s.P[q] + PATH(t, q, s.w);
                s.w = 0;
                k.reb++;
            }
        }
        return -1;
    }

    __device__ static void finish(const State &s, const KArgs &a, int64_t i, int reason) {
        const int64_t n = a.ld;
        const cplx R = a.t.Zn[s.w];
        a.Z[i] = to_std(s.x + R.re); a.Z[n + i] = to_std(s.y + R.im);
        if (a.calc_dc)
            for (int q = 0; q < 4; q++) a.Z[(2 + q) * n + i] = to_std(s.P[q] + PATH(a.t, q, s.w));
        a.U[i] = s.w; a.stop_iter[i] = s.n; a.stop_reason[i] = (int8_t)reason;
    }
};

// --------------------------------------------------------------------------
// BLA build: level 0 + three compressed levels per block of 8 iterations,
// then one launch per stored level; every merge of a level is independent.

struct HEntry { cplx A, B; double r; };

// combine_BLA, perturbation.py:1983-2024: e1 covers [i, ii), e2 [ii, i + 2 step)
FSG_HD HEntry merge(const HEntry &e1, const HEntry &e2, double kc, double eps) {
    HEntry o;
    o.A = e2.A * e1.A;
    o.B = e2.A * e1.B + e2.B;
    double mA1 = hypot2(e1.A.re, e1.A.im), mB1 = hypot2(e1.B.re, e1.B.im);
    double num = e2.r - mB1 * kc;
    double den = (eps > mA1) ? eps : mA1;
    double q = num / den;
    double r2b = 0.95 * (q > 0. ? q : 0.);
    o.r = (r2b < e1.r) ? r2b : e1.r;
    return o;
}

struct BEntry { double M[8]; double r; };

// combine_BLA_BS, perturbation.py:2027-2105
FSG_HD BEntry merge(const BEntry &e1, const BEntry &e2, double kc, double eps) {
    const double *a = e1.M, *b = e2.M;
    BEntry o;
    o.M[0] = b[0] * a[0] + b[1] * a[2];
    o.M[1] = b[0] * a[1] + b[1] * a[3];
    o.M[2] = b[2] * a[0] + b[3] * a[2];
    o.M[3] = b[2] * a[1] + b[3] * a[3];
    o.M[4] = b[0] * a[4] + b[1] * a[6] + b[4];
    o.M[5] = b[0] * a[5] + b[1] * a[7] + b[5];
    o.M[6] = b[2] * a[4] + b[3] * a[6] + b[6];
    o.M[7] = b[2] * a[5] + b[3] * a[7] + b[7];
    double mA1 = fabs(a[0]), mB1 = fabs(a[4]);
    for (int q = 1; q < 4; q++) {
        mA1 = fabs(a[q]) > mA1 ? fabs(a[q]) : mA1;
        mB1 = fabs(a[4 + q]) > mB1 ? fabs(a[4 + q]) : mB1;
    }
    double num = e2.r - mB1 * kc;
    double den = (eps > mA1) ? eps : mA1;
    double q = num / den;
    double r2b = 0.95 * (q > 0. ? q : 0.);
    o.r = (r2b < e1.r) ? r2b : e1.r;
    return o;
}

__device__ __forceinline__ HEntry load_h(const double *M, const double *r, int64_t i) {
    HEntry e; e.A = mk(M[4 * i], M[4 * i + 1]); e.B = mk(M[4 * i + 2], M[4 * i + 3]); e.r = r[i];
    return e;
}
__device__ __forceinline__ void store_h(double *M, double *r, int64_t i, const HEntry &e) {
    M[4 * i] = e.A.re; M[4 * i + 1] = e.A.im; M[4 * i + 2] = e.B.re; M[4 * i + 3] = e.B.im; r[i] = e.r;
}
__device__ __forceinline__ BEntry load_b(const double *M, const double *r, int64_t i) {
    BEntry e;
    for (int q = 0; q < 8; q++) e.M[q] = M[8 * i + q];
    e.r = r[i];
    return e;
}
__device__ __forceinline__ void store_b(double *M, double *r, int64_t i, const BEntry &e) {
    for (int q = 0; q < 8; q++) M[8 * i + q] = e.M[q];
    r[i] = e.r;
}

// RADII_ONLY: the coefficients M are already in place (a context sharing
// another's orbit and BLA coefficients); only the radii, which depend on kc
// and eps, are written
template <int KIND, int DEG, bool RADII_ONLY>
__global__ void k_bla0_holo(const cplx *Zn, int32_t comp, double kc, double eps, double *M, double *r) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= comp) return;
    HEntry e[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        cplx A = dfdz<KIND, DEG>(Zn[8 * i + j]);
        e[j].A = A; e[j].B = mk(1., 0.); e[j].r = eps * hypot2(A.re, A.im);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) e[j] = merge(e[2 * j], e[2 * j + 1], kc, eps);
#pragma unroll
    for (int j = 0; j < 2; j++) e[j] = merge(e[2 * j], e[2 * j + 1], kc, eps);
    const HEntry o = merge(e[0], e[1], kc, eps);
    if constexpr (RADII_ONLY) r[2 * (int64_t)i] = o.r;
    else store_h(M, r, 2 * (int64_t)i, o);
}

template <int FL, bool RADII_ONLY>
__global__ void k_bla0_bs(const cplx *Zn, int32_t comp, double kc, double eps, double *M, double *r) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= comp) return;
    BEntry e[8];
    for (int j = 0; j < 8; j++) {
        cplx Z = Zn[8 * i + j];
        bs_jac<FL>(Z.re, Z.im, e[j].M);
        e[j].M[4] = 1.; e[j].M[5] = 0.; e[j].M[6] = 0.; e[j].M[7] = -1.;
        double ax = fabs(Z.re), ay = fabs(Z.im);
        e[j].r = eps * (ax < ay ? ax : ay);
    }
    for (int j = 0; j < 4; j++) e[j] = merge(e[2 * j], e[2 * j + 1], kc, eps);
    for (int j = 0; j < 2; j++) e[j] = merge(e[2 * j], e[2 * j + 1], kc, eps);
    const BEntry o = merge(e[0], e[1], kc, eps);
    if constexpr (RADII_ONLY) r[2 * (int64_t)i] = o.r;
    else store_b(M, r, 2 * (int64_t)i, o);
}

// one stored level: merge i in [0, floor(comp / 2^stg)) of stage stg
template <bool BS, bool RADII_ONLY>
__global__ void k_bla_merge(int32_t comp, int32_t stg, double kc, double eps, double *M, double *r) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t step = 1LL << stg;
    if (t >= comp / step) return;
    const int64_t i = t * step, ii = i + step / 2;
    const int64_t i1 = 2 * i + (step / 2 - 1), i2 = 2 * ii + (step / 2 - 1), ir = 2 * i + (step - 1);
    if constexpr (BS) {
        const BEntry o = merge(load_b(M, r, i1), load_b(M, r, i2), kc, eps);
        if constexpr (RADII_ONLY) r[ir] = o.r; else store_b(M, r, ir, o);
    } else {
        const HEntry o = merge(load_h(M, r, i1), load_h(M, r, i2), kc, eps);
        if constexpr (RADII_ONLY) r[ir] = o.r; else store_h(M, r, ir, o);
    }
}

// --------------------------------------------------------------------------
// dispatch
typedef void (*tile_fn)(const KArgs);

template <class Model> static tile_fn fn_of() { return k_tile<Model>; }

static tile_fn select_tile(const LaunchCfg &c) {
    if (c.kind == KIND_M2) return c.xr ? fn_of<HoloXr<KIND_M2, 2>>() : fn_of<HoloStd<KIND_M2, 2>>();
    if (c.kind == KIND_MN) {
        switch (c.degree) {
#define FSG_MN(D) case D: return c.xr ? fn_of<HoloXr<KIND_MN, D>>() : fn_of<HoloStd<KIND_MN, D>>();
            FSG_MN(2) FSG_MN(3) FSG_MN(4) FSG_MN(5) FSG_MN(6) FSG_MN(7) FSG_MN(8)
#undef FSG_MN
            default: return nullptr;
        }
    }
    if (c.kind == KIND_BS) {
        switch (c.flavor) {
#define FSG_BS(F) case F: return c.xr ? fn_of<BsXr<F>>() : fn_of<BsStd<F>>();
            FSG_BS(1) FSG_BS(2) FSG_BS(3) FSG_BS(4) FSG_BS(5)
#undef FSG_BS
            default: return nullptr;
        }
    }
    return nullptr;
}

// blocks that fill the GPU at the kernel's occupancy, cached per kernel
int grid_fill_blocks(const void *fn) {
    static std::mutex mu;
    static std::unordered_map<long long, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    const long long key = (long long)(uintptr_t)fn * 64 + dev;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    int ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess) return -1;
    const int blocks = ncu * (per_cu < 1 ? 1 : per_cu);
    std::lock_guard<std::mutex> g(mu);
    cache[key] = blocks;
    return blocks;
}

int tile_kernel_blocks(const LaunchCfg &cfg) {
    tile_fn f = select_tile(cfg);
    if (!f) return -1;
    return grid_fill_blocks(reinterpret_cast<const void *>(f));
}

int launch_blocks(int fill, const KArgs &a) {
    int blocks = fill;
    if (a.max_blocks > 0 && a.max_blocks < blocks) blocks = a.max_blocks;
    if (a.max_blocks < 0) blocks = (fill + (-a.max_blocks) - 1) / (-a.max_blocks);   // a share of the GPU
    const int64_t need = (a.npts + 255) / 256;
    if (need < blocks) blocks = (int)(need > 0 ? need : 1);
    return blocks;
}

hipError_t launch_tile(const LaunchCfg &cfg, const KArgs &a, hipStream_t stream) {
    tile_fn f = select_tile(cfg);
    if (!f) return hipErrorInvalidValue;
    const int fill = grid_fill_blocks(reinterpret_cast<const void *>(f));
    if (fill <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(f, dim3(launch_blocks(fill, a)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

template <bool RO>
static hipError_t bla_build_t(int kind, int degree, int flavor, const cplx *Zn, int32_t L,
                              double kc, double eps, double *M, double *r, int32_t stages,
                              hipStream_t stream) {
    const int32_t comp = L >> 3;
    if (comp <= 0) return hipErrorInvalidValue;
    const int tb = 128;
    dim3 g0((comp + tb - 1) / tb);
    if (kind == KIND_M2) hipLaunchKernelGGL((k_bla0_holo<KIND_M2, 2, RO>), g0, dim3(tb), 0, stream, Zn, comp, kc, eps, M, r);
    else if (kind == KIND_MN) {
        switch (degree) {
#define FSG_MN(D) case D: hipLaunchKernelGGL((k_bla0_holo<KIND_MN, D, RO>), g0, dim3(tb), 0, stream, Zn, comp, kc, eps, M, r); break;
            FSG_MN(2) FSG_MN(3) FSG_MN(4) FSG_MN(5) FSG_MN(6) FSG_MN(7) FSG_MN(8)
#undef FSG_MN
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (flavor) {
#define FSG_BS(F) case F: hipLaunchKernelGGL((k_bla0_bs<F, RO>), g0, dim3(tb), 0, stream, Zn, comp, kc, eps, M, r); break;
            FSG_BS(1) FSG_BS(2) FSG_BS(3) FSG_BS(4) FSG_BS(5)
#undef FSG_BS
            default: return hipErrorInvalidValue;
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int32_t stg = 1; stg < stages - 3; stg++) {
        int64_t nm = comp >> stg;
        if (nm <= 0) break;
        dim3 g((unsigned)((nm + tb - 1) / tb));
        if (kind == KIND_BS) hipLaunchKernelGGL((k_bla_merge<true, RO>), g, dim3(tb), 0, stream, comp, stg, kc, eps, M, r);
        else hipLaunchKernelGGL((k_bla_merge<false, RO>), g, dim3(tb), 0, stream, comp, stg, kc, eps, M, r);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_bla_build(int kind, int degree, int flavor, const cplx *Zn, int32_t L,
                            double kc, double eps, double *M, double *r, int32_t stages,
                            hipStream_t stream, bool radii_only) {
    return radii_only ? bla_build_t<true>(kind, degree, flavor, Zn, L, kc, eps, M, r, stages, stream)
                      : bla_build_t<false>(kind, degree, flavor, Zn, L, kc, eps, M, r, stages, stream);
}

}  // namespace fsg
