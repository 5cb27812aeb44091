// tile_models.hpp -- the per-pixel models of the tile kernel (init / step /
// finish of numba_iterate and numba_iterate_BS), shared by the translation
// units that instantiate them.
//
// Reference loops restated here:
//   holomorphic: numba_iterate, src/fractalshades/perturbation.py:1065-1400
//   burning ship: numba_iterate_BS, perturbation.py:1468-1791
//   model kernels: models/mandelbrot_M2.py:599-622, mandelbrot_Mn.py:646-729,
//                  burning_ship.py:12-60, 440-859
// The ORB parameter adds calc_orbit (perturbation.py:1096-1101, 1266-1273,
// 1390-1396): the orbit point at exit - backshift, in an extra output field.
#pragma once
#include "models.hpp"
#include "tile_loop.hpp"
#include "xl.hpp"

namespace fsg {

// calc_orbit recorder: the last two full orbit points (Z + z), taken at the
// escape test of plain steps when n reaches the next multiple of backshift
// (perturbation.py:1266-1273; n // backshift > div_shift is n >= next).
template <bool ORB> struct OrbitRec {
    __device__ void reset(int32_t) {}
    __device__ void record(int32_t, int32_t, double, double) {}
};
template <> struct OrbitRec<true> {
    int32_t next, i1, i2;
    double x1, y1, x2, y2;
    __device__ void reset(int32_t b) { next = b; i1 = 0; i2 = 0; x1 = y1 = x2 = y2 = 0.; }
    __device__ void record(int32_t n, int32_t b, double x, double y) {
        if (n >= next) {
            next = (n / b + 1) * b;
            i2 = i1; x2 = x1; y2 = y1;
            i1 = n; x1 = x; y1 = y;
        }
    }
};

// the full recurrence, for the orbit finish: M2_iterate (mandelbrot_M2.py:
// 14-15), Mn_iterate (mandelbrot_Mn.py:13-17) by binary powering, and
// BS_iterate (burning_ship.py:82-123)
template <int KIND, int DEG>
__device__ __forceinline__ cplx zn_full(cplx z, cplx c) {
    if constexpr (KIND == KIND_M2) {
        return z * z + c;
    } else {
        cplx r = mk(1., 0.), p = z;
#pragma unroll
        for (int m = DEG; m > 0; m >>= 1) {
            if (m & 1) r = r * p;
            if (m > 1) p = p * p;
        }
        return r + c;
    }
}
template <int FL>
__device__ __forceinline__ void bs_full(double &x, double &y, double A, double B) {
    const double X = x, Y = y;
    if constexpr (FL == 1) { x = X * X - Y * Y + A; y = 2. * fabs(X * Y) - B; }
    else if constexpr (FL == 2) { x = X * X - Y * Y + A; y = 2. * X * fabs(Y) - B; }
    else if constexpr (FL == 3) { x = X * X - Y * fabs(Y) + A; y = 2. * X * Y - B; }
    else if constexpr (FL == 4) { x = fabs(X * X - Y * Y) + A; y = 2. * X * Y - B; }
    else { x = fabs(X * X - Y * Y) + A; y = 2. * fabs(X * Y) - B; }
}

// orbit finish (perturbation.py:1390-1396): from the last-but-one record to
// n - backshift with c + Zn_path[1]
template <int KIND, int DEG>
__device__ __forceinline__ void orbit_finish_holo(const OrbitRec<true> &o, const KArgs &a, int32_t n,
                                                  cplx c, int64_t i) {
    cplx zo = mk(o.x2, o.y2);
    const cplx CC = c + a.t.Zn[1];
    for (int32_t k = o.i2; k < n - a.backshift; k++) zo = zn_full<KIND, DEG>(zo, CC);
    const int64_t fo = 1 + a.calc_dz + a.calc_dc;
    a.Z[2 * (fo * a.ld + i)] = zo.re; a.Z[2 * (fo * a.ld + i) + 1] = zo.im;
}
// the same for the burning ship (perturbation.py:1778-1787); the second
// constant is b + Im Zn_path[1], as the reference has it
template <int FL>
__device__ __forceinline__ void orbit_finish_bs(const OrbitRec<true> &o, const KArgs &a, int32_t n,
                                                double ca, double cb, int64_t i) {
    double x = o.x2, y = o.y2;
    const cplx R1 = a.t.Zn[1];
    const double AA = ca + R1.re, BB = cb + R1.im;
    for (int32_t k = o.i2; k < n - a.backshift; k++) bs_full<FL>(x, y, AA, BB);
    const int64_t fo = 2 + (a.calc_dc ? 4 : 0);
    a.Z[fo * a.ld + i] = x; a.Z[(fo + 1) * a.ld + i] = y;
}

// BLA lookup on binary64 |z| (ref_BLA_get, perturbation.py:2116-2170)
__device__ __forceinline__ int32_t bla_lookup(const Tables &t, double zre, double zim,
                                              int32_t w, int32_t &idx) {
    return bla_find(t, hypot2(zre, zim), w, idx);
}
// The same on an extended-range z: the modulus of its mantissa against the
// radii scaled by 2^-e.  Where the binary64 copy of z is a normal number
// this takes the decision hypot2 would (the prescale of hypot2 is the same
// power of two); where that copy is subnormal or zero it keeps |z|'s
// precision.
__device__ __forceinline__ int32_t bla_lookup(const Tables &t, xl z, int32_t w, int32_t &idx) {
    return bla_find<true>(t, sqrt(z.re * z.re + z.im * z.im), w, idx, z.e);
}

// The z^d + c step with the derivatives the launch computes (dz/dc when dc,
// dz/dz when dz): one pass over the shared partial products
// (p_zn_derivs); Zd / Zh are read only for the derivatives computed.
// add(a, b) takes the last sum of each result and returns it as an R
// (binary64: the sum; extended range: the normalised sum).
template <int DEG, class T, class R, class Add>
__device__ __forceinline__ R mn_step(T z, T Z, T c, bool dc, bool dz, R &d, R &h, T d0, T h0, T Zd, T Zh,
                                     Add add) {
    if (dc && dz) {
        const T v[2] = {d0, h0};
        const T Zv[2] = {Zd, Zh};
        R o[2];
        const R r = p_zn_derivs<DEG, 2>(z, Z, c, v, Zv, o, add);
        d = o[0]; h = o[1];
        return r;
    }
    if (dc || dz) {
        const T v[1] = {dc ? d0 : h0};
        const T Zv[1] = {dc ? Zd : Zh};
        R o[1];
        const R r = p_zn_derivs<DEG, 1>(z, Z, c, v, Zv, o, add);
        (dc ? d : h) = o[0];
        return r;
    }
    const T zz[1] = {z};
    R o[1];
    return p_zn_derivs<DEG, 0>(z, Z, c, zz, zz, o, add);
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
// Holomorphic models, binary64 deltas.  DER >= 0 fixes the derivatives at
// compile time (bit 0: dz/dc, bit 1: dz/dz); DER = -1 reads them from the
// launch arguments.
template <int KIND, int DEG, bool ORB = false, int DER = -1>
struct HoloStd {
    __device__ static bool DC_ON(const KArgs &a) { if constexpr (DER >= 0) return (DER & 1) != 0; else return a.calc_dc != 0; }
    __device__ static bool DZ_ON(const KArgs &a) { if constexpr (DER >= 0) return (DER & 2) != 0; else return a.calc_dz != 0; }
    struct State { cplx z, d, h, c; int32_t w, n; bool dyn, nul; OrbitRec<ORB> orb; };

    __device__ static void init(State &s, const KArgs &a, double px, double py) {
        s.c = to_std(c_from_pix(a, px, py));
        s.z = zc(); s.d = zc(); s.h = zc();
        s.w = 0; s.n = 0; s.dyn = true; s.nul = false;
        s.orb.reset(a.backshift);
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
                if (DC_ON(a)) s.d = A * s.d;
                if (DZ_ON(a)) s.h = A * s.h;
                k.bla++; k.bla_it += st;
            }
        }
        s.n += 1; k.plain++;
        // Every table value the step reads is loaded here, before any of it
        // is used: the orbit point and path values at w, the next orbit point
        // (escape and rebase tests) and the dZndz value of the interior test,
        // whose indexes are known at the start of the step.  A pixel whose
        // wave has no other work left (the last, longest pixels of a launch)
        // then waits for one table round trip per step instead of one per
        // test.
        const int32_t w1 = (s.w + 1 >= t.order) ? 0 : s.w + 1;
        const cplx Zr = t.Zn[s.w];
        const cplx Z1 = t.Zn[w1];
        cplx Zd = zc(), Zh = zc(), H1 = zc();
        if (DC_ON(a)) { const cplx v = t.dc[s.w]; Zd = s.dyn ? zc() : v; }
        if (DZ_ON(a)) {
            Zh = t.dz[s.nul ? 0 : s.w];
            H1 = t.dz[s.nul ? 0 : (s.n == t.order ? t.L : w1)];
        }
        if constexpr (KIND == KIND_MN) {
            s.z = mn_step<DEG>(s.z, Zr, s.c, DC_ON(a), DZ_ON(a), s.d, s.h, s.d, s.h, Zd, Zh,
                               [](cplx u, cplx v) { return u + v; });
        } else {
            if (DC_ON(a)) s.d = p_deriv<KIND, DEG>(s.z, s.d, Zr, Zd);
            if (DZ_ON(a)) s.h = p_deriv<KIND, DEG>(s.z, s.h, Zr, Zh);
            s.z = p_zn<KIND, DEG>(s.z, Zr, s.c);
        }
        s.w = w1;
        // The stop and rebase tests are evaluated together: a step that meets
        // none of them (nearly every step) leaves by one branch, and the
        // tests' own order (max_iter, interior, escape, rebase) is kept on
        // the other path.  (The orbit recorder runs between the tests: its
        // kernels take that path on every step.)
        bool t_int = false;
        if (DZ_ON(a)) {
            const cplx q = s.h + H1;
            t_int = q.re * q.re + q.im * q.im < a.eps2;
        }
        const cplx ZZ = s.z + Z1;
        const bool t_esc = ZZ.re * ZZ.re + ZZ.im * ZZ.im > a.M2;
        const bool reb_div = s.w >= t.div_iter - 1;
        const bool dyn = (fabs(ZZ.re) <= fabs(s.z.re)) && (fabs(ZZ.im) <= fabs(s.z.im));
        if constexpr (!ORB) {
            if (!(s.n >= a.max_iter || t_int || t_esc || reb_div || dyn)) { s.dyn = false; return -1; }
        }
        if (s.n >= a.max_iter) return 0;
        if (t_int) return 2;
        s.orb.record(s.n, a.backshift, ZZ.re, ZZ.im);
        if (t_esc) return 1;
        bool reb = reb_div;
        if (!reb) {
            s.dyn = dyn;
            reb = dyn;
        }
        if (reb) {
            s.z = ZZ;
            if (DC_ON(a)) s.d = s.d + t.dc[s.w];
            if (DZ_ON(a)) {
                if (!s.nul) s.h = s.h + H1;     // dZndz at w (at L after a full period)
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
        if (DZ_ON(a)) { a.Z[2 * (n + i)] = s.h.re; a.Z[2 * (n + i) + 1] = s.h.im; f = 2; }
        if (DC_ON(a)) {
            cplx d = s.d + a.t.dc[s.w];
            a.Z[2 * (f * n + i)] = d.re; a.Z[2 * (f * n + i) + 1] = d.im;
        }
        if constexpr (ORB) orbit_finish_holo<KIND, DEG>(s.orb, a, s.n, s.c, i);
        a.U[i] = s.w; a.stop_iter[i] = s.n; a.stop_reason[i] = (int8_t)reason;
    }
};

// Holomorphic models, Xrange deltas (dx < 1e-300).  The deltas are kept in
// the lazily normalised extended-range form of xl.hpp: every result is
// renormalised without branches, which rounds exactly like the reference's
// Xrange record (a rescaling by a power of two is exact).
// DER >= 0 fixes the derivatives at compile time (bit 0: dz/dc, bit 1:
// dz/dz): the step is then one block of straight-line code, which the
// compiler schedules across the z, dz/dc and dz/dz recurrences; DER = -1
// reads them from the launch arguments.
template <int KIND, int DEG, bool ORB = false, int DER = -1>
struct HoloXr {
    __device__ static bool DC_ON(const KArgs &a) { if constexpr (DER >= 0) return (DER & 1) != 0; else return a.calc_dc != 0; }
    __device__ static bool DZ_ON(const KArgs &a) { if constexpr (DER >= 0) return (DER & 2) != 0; else return a.calc_dz != 0; }
    struct State { xl z, d, h, c; cplx zs; int32_t w, n; bool dyn, nul; OrbitRec<ORB> orb; };

    // derivative path values, unnormalised: the host's Xrange records keep
    // |mantissa| within 2^+-100, and a zero entry carries the zero exponent
    // code (the tables are built so), so a value is used as it is loaded
    __device__ static xz DCz(const Tables &t, int32_t i) { const cplx m = t.dc[i]; xz r; r.re = m.re; r.im = m.im; r.e = t.dc_e[i]; return r; }
    __device__ static xz DZz(const Tables &t, int32_t i) { const cplx m = t.dz[i]; xz r; r.re = m.re; r.im = m.im; r.e = t.dz_e[i]; return r; }
    // an orbit point, normalised, from its code k and binary64 value z: the
    // register's Xrange value where the binary64 orbit underflows, z
    // elsewhere, scaled by the exponent k (xl_of of the point, without its
    // exponent extraction)
    __device__ static xl ZRv(const Tables &t, int32_t k, cplx z) {
        if (k < ZK_REGISTER) {
            const int32_t j = k - INT32_MIN;
            return xl_of(t.xr_m[2 * j], t.xr_m[2 * j + 1], t.xr_e[j]);
        }
        xl r;
        r.re = ldexp(z.re, -k); r.im = ldexp(z.im, -k); r.e = k;
        return r;
    }

    __device__ static void init(State &s, const KArgs &a, double px, double py) {
        s.c = c_from_pix_xl(a, px, py);
        s.z = xl_zero(); s.d = xl_zero(); s.h = xl_zero();
        s.zs = zc();
        s.w = 0; s.n = 0; s.dyn = true; s.nul = false;
        s.orb.reset(a.backshift);
    }

    __device__ static int step(State &s, const KArgs &a, Counters &k) {
        const Tables &t = a.t;
        while (t.bla_on && (s.w & 7) == 0) {
            int32_t idx;
            int32_t st = bla_lookup(t, s.z, s.w, idx);
            k.tries++;
            if (st == 0) break;
            {
                const double *Mb = t.M + 4 * idx;
                cplx A = mk(Mb[0], Mb[1]), B = mk(Mb[2], Mb[3]);
                s.n += st;
                s.w += st;
                s.z = add_n(as_z(s.z) * A, as_z(s.c) * B);
                s.zs = to_std(s.z);
                if (DC_ON(a)) s.d = mul_n(as_z(s.d), A);
                if (DZ_ON(a)) s.h = mul_n(as_z(s.h), A);
                k.bla++; k.bla_it += st;
            }
        }
        s.n += 1; k.plain++;
        // Every table value the step reads is loaded here, as in HoloStd: the
        // orbit point, its code and the path values at w, the next orbit
        // point and its code (escape and rebase tests) and the dZndz value of
        // the interior test (z^d + c: after the step, whose partial products
        // need the registers; loaded first, it spilled at d = 3)
        const int32_t w1 = (s.w + 1 >= t.order) ? 0 : s.w + 1;
        const int32_t k0 = t.zn_k[s.w], k1 = t.zn_k[w1];
        const cplx Z0 = t.Zn[s.w], Z1 = t.Zn[w1];
        xz Zd = xz_zero(), Zh = xz_zero(), H1 = xz_zero();
        if (DC_ON(a)) { const xz v = DCz(t, s.w); Zd = s.dyn ? xz_zero() : v; }
        if (DZ_ON(a)) {
            Zh = DZz(t, s.nul ? 0 : s.w);
            if constexpr (KIND == KIND_M2) H1 = DZz(t, s.nul ? 0 : (s.n == t.order ? t.L : w1));
        }
        const xz Zr = as_z(ZRv(t, k0, Z0));
        const xz z0 = as_z(s.z);
        if constexpr (KIND == KIND_M2) {
            // mandelbrot_M2.py:607-622 with Z + z formed once for both
            // derivatives and the factors 2 taken into the exponents (exact)
            const xz Zz = Zr + z0;
            if (DC_ON(a)) s.d = add_n(Zz * as_z(s.d), Zd * z0, 1);
            if (DZ_ON(a)) s.h = add_n(Zz * as_z(s.h), Zh * z0, 1);
            s.z = add_n(z0 * (z0 + muls(2., Zr)), as_z(s.c));
        } else {
            // the last sums normalised at once (add_n: no zero test of their own)
            s.z = mn_step<DEG>(z0, Zr, as_z(s.c), DC_ON(a), DZ_ON(a), s.d, s.h, as_z(s.d), as_z(s.h),
                               Zd, Zh, [](xz u, xz v) { return add_n(u, v); });
        }
        if constexpr (KIND == KIND_MN) {
            if (DZ_ON(a)) H1 = DZz(t, s.nul ? 0 : (s.n == t.order ? t.L : w1));
        }
        s.zs = to_std(s.z);
        s.w = w1;
        // the stop and rebase tests evaluated together, as in HoloStd
        const bool t_int = DZ_ON(a) && abs2_lt(add_raw(as_z(s.h), H1), a.eps2);
        const cplx ZZ = s.zs + Z1;
        const bool t_esc = ZZ.re * ZZ.re + ZZ.im * ZZ.im > a.M2;
        const bool reb_div = s.w >= t.div_iter - 1;
        const bool dyn = (fabs(ZZ.re) <= fabs(s.zs.re)) && (fabs(ZZ.im) <= fabs(s.zs.im));
        if constexpr (!ORB) {
            if (!(s.n >= a.max_iter || t_int || t_esc || reb_div || dyn)) { s.dyn = false; return -1; }
        }
        if (s.n >= a.max_iter) return 0;
        if (t_int) return 2;
        s.orb.record(s.n, a.backshift, ZZ.re, ZZ.im);
        if (t_esc) return 1;
        // both rebases (perturbation.py:1283-1372) share the derivative
        // updates: one block, whichever of them a lane takes
        bool reb = reb_div;
        xl ZZx;
        if (!reb_div) {
            s.dyn = dyn;
            if (s.dyn) {
                ZZx = s.z + ZRv(t, k1, Z1);
                reb = le_abs2(ZZx, s.z);
            }
        }
        if (reb) {
            if (reb_div) { s.zs = ZZ; s.z = xl_of(ZZ); }
            else { s.z = ZZx; s.zs = to_std(ZZx); }
            // the path values unnormalised: the sum's normalisation covers them
            if (DC_ON(a)) s.d = add_n(as_z(s.d), DCz(t, s.w));
            if (DZ_ON(a)) {
                if (!s.nul) s.h = add_n(as_z(s.h), H1);   // dZndz at w (at L after a full period)
                s.nul = true;
            }
            s.w = 0;
            k.reb++;
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
        if (DZ_ON(a)) { a.Z[2 * (n + i)] = 0.; a.Z[2 * (n + i) + 1] = 0.; f = 2; }
        if (DC_ON(a)) {
            cplx d = to_std(add_n(as_z(s.d), DCz(a.t, s.w)));
            a.Z[2 * (f * n + i)] = d.re; a.Z[2 * (f * n + i) + 1] = d.im;
        }
        if constexpr (ORB) orbit_finish_holo<KIND, DEG>(s.orb, a, s.n, to_std(s.c), i);
        a.U[i] = s.w; a.stop_iter[i] = s.n; a.stop_reason[i] = (int8_t)reason;
    }
};

// --------------------------------------------------------------------------
// Burning-ship family, binary64 deltas
template <int FL, bool ORB = false>
struct BsStd {
    struct State { double x, y, P[4], a, b; int32_t w, n; bool dyn; OrbitRec<ORB> orb; };

    __device__ static void init(State &s, const KArgs &a, double px, double py) {
        xc cx = renorm_c(a.dx_m * (a.lin[0] * px + a.lin[1] * py),
                         a.dx_m * (a.lin[2] * px + a.lin[3] * py), a.dx_e);
        s.a = to_std(mkf(cx.re, cx.e) + mkf(a.drift_m[0], a.drift_e[0]));
        s.b = to_std(mkf(cx.im, cx.e) + mkf(a.drift_m[1], a.drift_e[1]));
        s.x = 0.; s.y = 0.; s.P[0] = s.P[1] = s.P[2] = s.P[3] = 0.;
        s.w = 0; s.n = 0; s.dyn = true;
        s.orb.reset(a.backshift);
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
        // the step's table reads first, its tests together (as in HoloStd)
        const int32_t w1 = (s.w + 1 >= t.order) ? 0 : s.w + 1;
        const cplx R = t.Zn[s.w];
        const cplx R1 = t.Zn[w1];
        if (a.calc_dc) {
            double Rd[4] = {0., 0., 0., 0.};
            if (!s.dyn) { Rd[0] = t.bs[0][s.w]; Rd[1] = t.bs[1][s.w]; Rd[2] = t.bs[2][s.w]; Rd[3] = t.bs[3][s.w]; }
            bs_hess<FL>(s.x, s.y, s.P, R.re, R.im, Rd);
        }
        bs_zn<FL>(s.x, s.y, R.re, R.im, s.a, s.b);
        const double XX = s.x + R1.re, YY = s.y + R1.im;
        const bool t_esc = XX * XX + YY * YY > a.M2;
        const bool reb_div = w1 >= t.div_iter - 1;
        const bool dyn = (fabs(XX) <= fabs(s.x)) && (fabs(YY) <= fabs(s.y));
        if constexpr (!ORB) {
            if (!(s.n >= a.max_iter || t_esc || reb_div || dyn)) { s.w = w1; s.dyn = false; return -1; }
        }
        if (s.n >= a.max_iter) return 0;
        s.w = w1;
        s.orb.record(s.n, a.backshift, XX, YY);
        if (t_esc) return 1;
        bool reb = reb_div;
        if (!reb) {
            s.dyn = dyn;
            reb = dyn;
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
        if constexpr (ORB) orbit_finish_bs<FL>(s.orb, a, s.n, s.a, s.b, i);
        a.U[i] = s.w; a.stop_iter[i] = s.n; a.stop_reason[i] = (int8_t)reason;
    }
};

// Burning-ship family, Xrange deltas (xl.hpp reals)
template <int FL, bool ORB = false>
struct BsXr {
    struct State { xlf x, y, P[4], a, b; double xs, ys; int32_t w, n; bool dyn; OrbitRec<ORB> orb; };

    __device__ static xlf PATH(const Tables &t, int q, int32_t i) { return xlf_of(t.bs[q][i], t.bs_e[q][i]); }
    // orbit point i, normalised per coordinate: the register's Xrange
    // values where the binary64 orbit underflows, the binary64 values
    // elsewhere, scaled by the exponents the tables hold for them
    // (from the point's codes kx, ky and binary64 value R, loaded beforehand)
    __device__ static void ZRv(const Tables &t, int32_t kx, int32_t ky, cplx R, xlf &X, xlf &Y) {
        if (kx < ZK_REGISTER) {
            const int32_t j = kx - INT32_MIN;
            X = xlf_of(t.xr_m[2 * j], t.xr_e[2 * j]); Y = xlf_of(t.xr_m[2 * j + 1], t.xr_e[2 * j + 1]);
        } else {
            X.m = ldexp(R.re, -kx); X.e = kx;
            Y.m = ldexp(R.im, -ky); Y.e = ky;
        }
    }

    __device__ static void init(State &s, const KArgs &a, double px, double py) {
        s.a = xlf_of(a.dx_m * (a.lin[0] * px + a.lin[1] * py), a.dx_e) + xlf_of(a.drift_m[0], a.drift_e[0]);
        s.b = xlf_of(a.dx_m * (a.lin[2] * px + a.lin[3] * py), a.dx_e) + xlf_of(a.drift_m[1], a.drift_e[1]);
        const xlf z0 = xlf_zero();
        s.x = z0; s.y = z0; s.P[0] = s.P[1] = s.P[2] = s.P[3] = z0;
        s.xs = 0.; s.ys = 0.;
        s.w = 0; s.n = 0; s.dyn = true;
        s.orb.reset(a.backshift);
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
                const xzf x = as_zf(s.x), y = as_zf(s.y);
                const xzf ca = as_zf(s.a), cb = as_zf(s.b);
                const xzf nx = mul_w(M[0], x) + mul_w(M[1], y) + mul_w(M[4], ca) + mul_w(M[5], cb);
                const xzf ny = mul_w(M[2], x) + mul_w(M[3], y) + mul_w(M[6], ca) + mul_w(M[7], cb);
                s.x = xlf_of(nx); s.y = xlf_of(ny);
                s.xs = to_std(s.x); s.ys = to_std(s.y);
                if (a.calc_dc) {
                    const xzf P0 = as_zf(s.P[0]), P1 = as_zf(s.P[1]), P2 = as_zf(s.P[2]), P3 = as_zf(s.P[3]);
                    s.P[0] = xlf_of(mul_w(M[0], P0) + mul_w(M[1], P2));
                    s.P[1] = xlf_of(mul_w(M[0], P1) + mul_w(M[1], P3));
                    s.P[2] = xlf_of(mul_w(M[2], P0) + mul_w(M[3], P2));
                    s.P[3] = xlf_of(mul_w(M[2], P1) + mul_w(M[3], P3));
                }
                k.bla++; k.bla_it += st;
            }
        }
        s.n += 1; k.plain++;
        // the step's table reads first, its tests together (as in HoloStd)
        const int32_t w1 = (s.w + 1 >= t.order) ? 0 : s.w + 1;
        const int32_t kx0 = t.zn_k[2 * s.w], ky0 = t.zn_k[2 * s.w + 1];
        const int32_t kx1 = t.zn_k[2 * w1], ky1 = t.zn_k[2 * w1 + 1];
        const cplx R0 = t.Zn[s.w], R1 = t.Zn[w1];
        xlf X, Y;
        ZRv(t, kx0, ky0, R0, X, Y);
        // one iteration on lazily normalised values, renormalised once
        const xzf x = as_zf(s.x), y = as_zf(s.y), Xz = as_zf(X), Yz = as_zf(Y);
        if (a.calc_dc) {
            const xzf z0 = as_zf(xlf_zero());
            xzf Rd[4] = {z0, z0, z0, z0};
            // the host's Xrange path records keep |mantissa| within 2^+-100 (a
            // zero entry with the zero exponent code): used as loaded, as
            // HoloXr's DCz
            if (!s.dyn) for (int q = 0; q < 4; q++) { Rd[q].m = t.bs[q][s.w]; Rd[q].e = t.bs_e[q][s.w]; }
            xzf P[4] = {as_zf(s.P[0]), as_zf(s.P[1]), as_zf(s.P[2]), as_zf(s.P[3])};
            bs_hess<FL>(x, y, P, Xz, Yz, Rd);
            for (int q = 0; q < 4; q++) s.P[q] = xlf_of(P[q]);
        }
        xzf nx = x, ny = y;
        bs_zn<FL>(nx, ny, Xz, Yz, as_zf(s.a), as_zf(s.b));
        s.x = xlf_of(nx); s.y = xlf_of(ny);
        s.xs = to_std(s.x); s.ys = to_std(s.y);
        const double XX = s.xs + R1.re, YY = s.ys + R1.im;
        const bool t_esc = XX * XX + YY * YY > a.M2;
        const bool reb_div = w1 >= t.div_iter - 1;
        const bool dyn = (fabs(XX) <= fabs(s.xs)) && (fabs(YY) <= fabs(s.ys));
        if constexpr (!ORB) {
            if (!(s.n >= a.max_iter || t_esc || reb_div || dyn)) { s.w = w1; s.dyn = false; return -1; }
        }
        if (s.n >= a.max_iter) return 0;
        s.w = w1;
        s.orb.record(s.n, a.backshift, XX, YY);
        if (t_esc) return 1;
        if (reb_div) {
            s.xs = XX; s.ys = YY;
            s.x = xlf_of(XX, 0); s.y = xlf_of(YY, 0);
            if (a.calc_dc) for (int q = 0; q < 4; q++) s.P[q] = s.P[q] + PATH(t, q, s.w);
            s.w = 0;
            k.reb++;
            return -1;
        }
        s.dyn = dyn;
        if (s.dyn) {
            xlf X1, Y1;
            ZRv(t, kx1, ky1, R1, X1, Y1);
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
        if constexpr (ORB) orbit_finish_bs<FL>(s.orb, a, s.n, to_std(s.a), to_std(s.b), i);
        a.U[i] = s.w; a.stop_iter[i] = s.n; a.stop_reason[i] = (int8_t)reason;
    }
};

}  // namespace fsg
