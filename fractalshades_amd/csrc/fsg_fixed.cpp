// fsg_fixed.cpp -- reference orbits in multi-limb fixed point, on the host.
//
// The reference computes the orbit of its reference point in MPFR/MPC, one
// correctly rounded operation at a time, and keeps each iterate rounded to
// binary64 plus an Xrange copy of those below 1e-300:
//   perturbation_mandelbrotN_select_FP_loop (mpmath_utils/FP_loop.pyx:274-421)
//   perturbation_nonholomorphic_FP_loop     (FP_loop.pyx:1828-1979)
// The iterates of these loops stay bounded (the loop stops at |z| > M), so a
// fixed-point number -- a sign and a magnitude of n 64-bit limbs, of which
// nf are fraction limbs -- holds them at a constant absolute resolution.
// With 64*nf >= prec + 128 fraction bits each operation's error (one unit of
// 2^-64nf) is far below MPFR's (half an ulp of a prec-bit value), so the
// binary64 and Xrange values read off the orbit are the ones that MPFR's
// orbit rounds to, except when an iterate lies within 2^-prec (relative) of
// a binary64 rounding boundary.  Sums and differences are exact; each
// product is rounded once.  A step of z^2 + c costs two n-limb products:
// x^2 - y^2 = (x + y)(x - y) and xy.
#include "fsgpu.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace fsg {
int set_error(int code, const std::string &msg);
}

using namespace fsg;

namespace {

typedef uint64_t u64;
typedef unsigned __int128 u128;

constexpr double XR_TSHOLD = 1e-300;   // FP_loop.pyx: XR_TSHOLD
constexpr int GUARD = 2;               // limbs kept below the product's unit

// Threads that each form one chunk of a product's columns.  They spin
// between products (a product of a few thousand limb multiplications takes a
// few microseconds, too short to sleep and wake), and leave when the team
// is destroyed at the end of the orbit.
class Team {
  public:
    explicit Team(int T) : T(T) {
        for (int t = 1; t < T; t++) th.emplace_back([this, t] { loop(t); });
    }
    ~Team() {
        quit.store(true, std::memory_order_release);
        gen.fetch_add(1, std::memory_order_acq_rel);
        for (auto &x : th) x.join();
    }
    const int T;
    // fn(ctx, t) on every member t = 0..T-1 (0 is the caller); returns when all are done
    void run(void (*fn)(void *, int), void *ctx) {
        job = fn;
        arg = ctx;
        done.store(0, std::memory_order_relaxed);
        gen.fetch_add(1, std::memory_order_acq_rel);
        fn(ctx, 0);
        for (int spin = 0; done.load(std::memory_order_acquire) != T - 1; spin++) relax(spin);
    }

  private:
    // spin briefly, then give the core away (when the team has more threads
    // than free cores, a spinning member would hold up a descheduled one)
    static void relax(int spin) {
        if (spin < 4096) __builtin_ia32_pause();
        else std::this_thread::yield();
    }
    void loop(int t) {
        uint64_t seen = 0;
        for (;;) {
            uint64_t g;
            for (int spin = 0; (g = gen.load(std::memory_order_acquire)) == seen; spin++) relax(spin);
            seen = g;
            if (quit.load(std::memory_order_acquire)) return;
            job(arg, t);
            done.fetch_add(1, std::memory_order_acq_rel);
        }
    }
    std::vector<std::thread> th;
    std::atomic<uint64_t> gen{0};
    std::atomic<int> done{0};
    std::atomic<bool> quit{false};
    void (*job)(void *, int) = nullptr;
    void *arg = nullptr;
};

// A fixed-point value: (-1)^neg * sum m[k] 2^(64k) * 2^(-64 nf)
struct Fx {
    std::vector<u64> m;
    bool neg = false;
};

class Arith {
  public:
    // team: threads for the products (null: the caller forms every column)
    Arith(int n, int nf, Team *team = nullptr)
        : n(n), nf(nf), lo(nf - GUARD > 0 ? nf - GUARD : 0), acc(2 * n + 2), team(team) {
        // split the columns lo .. 2n-2 into chunks of about equal work
        const int T = team ? team->T : 1;
        std::vector<double> work;
        double total = 0.;
        for (int k = lo; k <= 2 * n - 2; k++) {
            const int i0 = k - (n - 1) > 0 ? k - (n - 1) : 0, i1 = k < n - 1 ? k : n - 1;
            total += i1 - i0 + 1;
            work.push_back(total);
        }
        cut.assign(T + 1, 2 * n - 1);
        cut[0] = lo;
        for (int t = 1; t < T; t++) {
            const double goal = total * t / T;
            cut[t] = lo + (int)(std::lower_bound(work.begin(), work.end(), goal) - work.begin());
            if (cut[t] < cut[t - 1]) cut[t] = cut[t - 1];
        }
        carry.assign(2 * T, 0);
    }
    const int n, nf, lo;

    Fx make() const { Fx r; r.m.assign(n, 0); return r; }

    static bool is_zero(const Fx &a) {
        for (u64 w : a.m) if (w) return false;
        return true;
    }
    int cmp_mag(const Fx &a, const Fx &b) const {
        for (int k = n - 1; k >= 0; k--)
            if (a.m[k] != b.m[k]) return a.m[k] < b.m[k] ? -1 : 1;
        return 0;
    }
    // r = |a| + |b|
    void add_mag(Fx &r, const Fx &a, const Fx &b) const {
        u64 carry = 0;
        for (int k = 0; k < n; k++) {
            const u128 t = (u128)a.m[k] + b.m[k] + carry;
            r.m[k] = (u64)t;
            carry = (u64)(t >> 64);
        }
        overflow_ |= carry != 0;
    }
    // r = |a| - |b|, |a| >= |b|
    void sub_mag(Fx &r, const Fx &a, const Fx &b) const {
        u64 borrow = 0;
        for (int k = 0; k < n; k++) {
            const u64 x = a.m[k], y = b.m[k];
            const u64 d = x - y - borrow;
            borrow = (x < y) || (x - y < borrow);
            r.m[k] = d;
        }
    }
    // r = a + (bneg ? -|b| : |b|); r may alias a or b
    void add_signed(Fx &r, const Fx &a, const Fx &b, bool bneg) const {
        if (a.neg == bneg) {
            const bool s = a.neg;
            add_mag(r, a, b);
            r.neg = s;
        } else if (cmp_mag(a, b) >= 0) {
            const bool s = a.neg;
            sub_mag(r, a, b);
            r.neg = s;
        } else {
            sub_mag(r, b, a);
            r.neg = bneg;
        }
        if (r.neg && is_zero(r)) r.neg = false;
    }
    void add(Fx &r, const Fx &a, const Fx &b) const { add_signed(r, a, b, b.neg); }
    void sub(Fx &r, const Fx &a, const Fx &b) const { add_signed(r, a, b, !b.neg); }

    // r = a * b rounded to the fixed-point grid (half up on the magnitude).
    // Column by column (product scanning, a 192-bit accumulator): only the
    // columns from nf - GUARD up are formed; the ones below add up to less
    // than n^2 2^-64 units of the last place.  With a team, each member
    // forms one chunk of columns and the chunks' carries are added after.
    void mul(Fx &r, const Fx &a, const Fx &b) {
        pa = a.m.data();
        pb = b.m.data();
        const int T = team ? team->T : 1;
        if (T > 1) team->run(&Arith::chunk_job, this);
        else columns(0);
        const int top = 2 * n - lo;
        for (int t = 0; t + 1 < T; t++) add_at(cut[t + 1] - lo, carry[2 * t], carry[2 * t + 1], top);
        // round: add half a unit (bit 63 of limb nf - 1)
        const int half = nf - 1 - lo;
        if (half >= 0) add_at(half, (u64)1 << 63, 0, top);
        const int off = nf - lo;
        for (int k = 0; k < n; k++) r.m[k] = acc[off + k];
        for (int k = off + n; k <= top; k++) overflow_ |= acc[k] != 0;
        r.neg = (a.neg != b.neg) && !is_zero(r);
    }
    // r = 2a (exact)
    void twice(Fx &r, const Fx &a) const {
        u64 carry = 0;
        for (int k = 0; k < n; k++) {
            const u64 w = a.m[k];
            r.m[k] = (w << 1) | carry;
            carry = w >> 63;
        }
        overflow_ |= carry != 0;
        r.neg = a.neg;
    }

    // index of the highest set bit of |a| (-1 for zero)
    int msb(const Fx &a) const {
        for (int k = n - 1; k >= 0; k--)
            if (a.m[k]) return 64 * k + 63 - __builtin_clzll(a.m[k]);
        return -1;
    }
    // bits [p, p + 64) of |a| (p may be negative: zeros shift in)
    u64 bits_at(const Fx &a, int p) const {
        if (p >= 64 * n) return 0;
        if (p < 0) {
            if (p <= -64) return 0;
            return a.m[0] << (-p);
        }
        const int k = p >> 6, s = p & 63;
        u64 w = a.m[k] >> s;
        if (s && k + 1 < n) w |= a.m[k + 1] << (64 - s);
        return w;
    }
    // any set bit below bit p of |a|
    bool sticky_below(const Fx &a, int p) const {
        if (p <= 0) return false;
        const int k = p >> 6, s = p & 63;
        for (int q = 0; q < k && q < n; q++) if (a.m[q]) return true;
        if (k < n && s && (a.m[k] & (((u64)1 << s) - 1))) return true;
        return false;
    }
    // |a| >> p rounded to nearest, ties to even (the result fits 54 bits)
    u64 round_at(const Fx &a, int p) const {
        if (p <= 0) return bits_at(a, p);
        const u64 q = bits_at(a, p);
        const bool half = (bits_at(a, p - 1) & 1) != 0;
        const bool rest = sticky_below(a, p - 1);
        if (half && (rest || (q & 1))) return q + 1;
        return q;
    }
    // binary64 nearest to a (ties to even), subnormals on the 2^-1074 grid
    double to_double(const Fx &a) const {
        const int F = 64 * nf;
        const int b = msb(a);
        if (b < 0) return 0.;
        int p = b - 52;
        if (p < F - 1074) p = F - 1074;
        const u64 q = round_at(a, p);
        const double v = std::ldexp((double)q, p - F);
        return a.neg ? -v : v;
    }
    // (m, e): 0.5 <= |m| < 1, m * 2^e = a rounded to 53 bits, (0, 0) for 0
    // (xrange mpf_to_xr / FP_loop.pyx:443 mpfr_t_to_Xrange)
    void to_xr(const Fx &a, double &m, int32_t &e) const {
        const int F = 64 * nf;
        const int b = msb(a);
        if (b < 0) { m = 0.; e = 0; return; }
        u64 q = round_at(a, b - 52);
        int32_t ex = b + 1 - F;
        if (q >> 53) { q >>= 1; ex += 1; }
        m = std::ldexp((double)q, -53);
        if (a.neg) m = -m;
        e = ex;
    }
    bool overflowed() const { return overflow_; }
    void flag_overflow() const { overflow_ = true; }

  private:
    static void chunk_job(void *self, int t) { static_cast<Arith *>(self)->columns(t); }
    // columns cut[t] .. cut[t+1]-1 of pa * pb into acc; the carry out of the
    // chunk (two limbs, into columns cut[t+1] and cut[t+1]+1) into carry[2t..]
    void columns(int t) {
        const u64 *A = pa, *B = pb;
        const int k0 = cut[t], k1 = cut[t + 1];
        u64 c0 = 0, c1 = 0, c2 = 0;
        for (int k = k0; k < k1; k++) {
            const int i0 = k - (n - 1) > 0 ? k - (n - 1) : 0;
            const int i1 = k < n - 1 ? k : n - 1;
            for (int i = i0; i <= i1; i++) {
                const u128 p = (u128)A[i] * B[k - i];
                const u128 s = (u128)c0 + (u64)p;
                c0 = (u64)s;
                const u128 q = (u128)c1 + (u64)(p >> 64) + (u64)(s >> 64);
                c1 = (u64)q;
                c2 += (u64)(q >> 64);
            }
            acc[k - lo] = c0;
            c0 = c1; c1 = c2; c2 = 0;
        }
        if (k1 == 2 * n - 1) {        // the last chunk ends the product
            acc[k1 - lo] = c0;
            acc[k1 + 1 - lo] = c1;
        } else {
            carry[2 * t] = c0;
            carry[2 * t + 1] = c1;
        }
    }
    // acc += (v0 + v1 2^64) at limb k, carrying up to limb top
    void add_at(int k, u64 v0, u64 v1, int top) {
        u64 c = 0;
        for (; k <= top; k++) {
            const u128 s = (u128)acc[k] + v0 + c;
            acc[k] = (u64)s;
            c = (u64)(s >> 64);
            v0 = v1;
            v1 = 0;
            if (!c && !v0) return;
        }
        overflow_ |= (c | v0) != 0;
    }

    std::vector<u64> acc;
    Team *team;
    std::vector<int> cut;
    std::vector<u64> carry;
    const u64 *pa = nullptr, *pb = nullptr;
    mutable bool overflow_ = false;
};

struct Cx { Fx re, im; };

// (x + iy)^2
void csqr(Arith &A, Cx &r, const Cx &z, Fx &t1, Fx &t2, Fx &t3) {
    A.add(t1, z.re, z.im);
    A.sub(t2, z.re, z.im);
    A.mul(t3, z.re, z.im);
    A.mul(r.re, t1, t2);
    A.twice(r.im, t3);
}
// a * b with three rounded products (sums are exact)
void cmul(Arith &A, Cx &r, const Cx &a, const Cx &b, Fx &t1, Fx &t2, Fx &t3, Fx &k1) {
    A.add(t1, a.re, a.im);
    A.mul(k1, t1, b.re);            // k1 = c (a + b)
    A.sub(t1, b.im, b.re);
    A.mul(t2, a.re, t1);            // k2 = a (d - c)
    A.add(t1, b.re, b.im);
    A.mul(t3, a.im, t1);            // k3 = b (c + d)
    A.sub(r.re, k1, t3);
    A.add(r.im, k1, t2);
}

struct Work {
    Fx t1, t2, t3, k1;
    Cx p, b;
    explicit Work(const Arith &A) : t1(A.make()), t2(A.make()), t3(A.make()), k1(A.make()) {
        p.re = A.make(); p.im = A.make(); b.re = A.make(); b.im = A.make();
    }
};

// z^d, d >= 2, by binary powering
void cpow(Arith &A, Cx &r, const Cx &z, int d, Work &w) {
    if (d == 2) { csqr(A, r, z, w.t1, w.t2, w.t3); return; }
    w.b = z;
    bool have = false;
    for (int e = d;;) {
        if (e & 1) {
            if (have) { cmul(A, w.p, r, w.b, w.t1, w.t2, w.t3, w.k1); r = w.p; }
            else { r = w.b; have = true; }
        }
        e >>= 1;
        if (!e) break;
        csqr(A, w.p, w.b, w.t1, w.t2, w.t3);
        w.b = w.p;
    }
}

// burning-ship family (burning_ship.py:81-123): (x, y) -> (X + a, Y - b)
void bs_step(Arith &A, int flavor, Fx &x, Fx &y, const Fx &a, const Fx &b, Work &w) {
    Fx &s = w.t1, &d = w.t2, &xy = w.t3, &X = w.k1;
    switch (flavor) {
    case 3: {   // x^2 - y|y|, 2xy
        A.mul(s, x, x);
        Fx ay = y; ay.neg = false;
        A.mul(d, y, ay);
        A.sub(X, s, d);
        A.mul(xy, x, y);
        break;
    }
    default: {
        A.add(s, x, y);
        A.sub(d, x, y);
        A.mul(X, s, d);                 // x^2 - y^2
        if (flavor == 4 || flavor == 5) X.neg = false;
        if (flavor == 2) { Fx ay = y; ay.neg = false; A.mul(xy, x, ay); }   // x|y|
        else A.mul(xy, x, y);
        if (flavor == 1 || flavor == 5) xy.neg = false;                    // |xy|
        break;
    }
    }
    A.add(x, X, a);
    A.twice(xy, xy);
    A.sub(y, xy, b);
}

Fx load(const Arith &A, const uint64_t *limbs, int32_t neg) {
    Fx r = A.make();
    std::memcpy(r.m.data(), limbs, sizeof(u64) * A.n);
    r.neg = neg && !Arith::is_zero(r);
    return r;
}

// descriptor checks shared by the entry points; extra_bits: integer bits
// needed beyond M^degree
int check_desc(const fsg_orbit_desc *d, double extra_bits) {
    if (!d) return set_error(FSG_E_INVALID, "null orbit descriptor");
    if (d->kind != FSG_KIND_M2 && d->kind != FSG_KIND_MN && d->kind != FSG_KIND_BS)
        return set_error(FSG_E_INVALID, "unknown model kind");
    const bool holo = d->kind != FSG_KIND_BS;
    const int deg = d->kind == FSG_KIND_M2 ? 2 : d->degree;
    if (holo && (deg < 2 || deg > 64)) return set_error(FSG_E_INVALID, "degree must be in 2..64");
    if (!holo && (d->flavor < 1 || d->flavor > 5)) return set_error(FSG_E_INVALID, "burning-ship flavor must be 1..5");
    if (d->frac_limbs < 1 || d->n_limbs <= d->frac_limbs || d->n_limbs > (1 << 16))
        return set_error(FSG_E_INVALID, "bad limb counts");
    if (!d->c_re || !d->c_im || d->max_iter < 0 || !(d->M > 0.))
        return set_error(FSG_E_INVALID, "bad orbit arguments");
    // the integer part must hold M^d + |c|: the last iterate before the escape test
    const double int_bits = (holo ? deg : 2) * std::log2(d->M + 2.) + 3. + extra_bits;
    if (int_bits > 64. * (d->n_limbs - d->frac_limbs))
        return set_error(FSG_E_INVALID, "too few integer limbs for the escape radius");
    return 0;
}

// threads for the products: worth their hand-off above ~2500 bits
std::unique_ptr<Team> make_team(const fsg_orbit_desc *d) {
    int T = d->n_threads;
    if (T <= 0) {
        const int hw = (int)std::thread::hardware_concurrency();
        T = d->n_limbs >= 40 ? std::min(4, hw > 1 ? hw : 1) : 1;
    }
    T = std::min(T, 64);
    std::unique_ptr<Team> team;
    if (T > 1) team.reset(new Team(T));
    return team;
}

// ---- derivatives: block floating point -------------------------------------
// The derivatives of the orbit with respect to c grow without bound, so they
// are kept as K fixed-point mantissas sharing one binary exponent:
// value_k = v[k] * 2^e, with the largest |v[k]| in [1, 2) (all zero: e = 0).
// A product with a fixed-point iterate is rounded on the fixed-point grid,
// so the relative error of a derivative stays near 2^-(64 nf) times the
// inverse size of the iterate it was multiplied by.
struct Blk {
    std::vector<Fx> v;
    int64_t e = 0;
};

// shift |a| right by s >= 0 bits (truncating)
void shr(const Arith &A, Fx &a, int64_t s) {
    const int n = A.n;
    if (s >= 64LL * n) { std::fill(a.m.begin(), a.m.end(), 0); a.neg = false; return; }
    const int q = (int)(s >> 6), b = (int)(s & 63);
    for (int k = 0; k < n; k++) {
        const int src = k + q;
        u64 w = src < n ? a.m[src] >> b : 0;
        if (b && src + 1 < n) w |= a.m[src + 1] << (64 - b);
        a.m[k] = w;
    }
    if (Arith::is_zero(a)) a.neg = false;
}
// shift |a| left by s >= 0 bits (the caller keeps it within the limbs)
void shl(const Arith &A, Fx &a, int64_t s) {
    const int n = A.n;
    const int q = (int)(s >> 6), b = (int)(s & 63);
    for (int k = n - 1; k >= 0; k--) {
        const int src = k - q;
        u64 w = src >= 0 ? a.m[src] << b : 0;
        if (b && src - 1 >= 0) w |= a.m[src - 1] >> (64 - b);
        a.m[k] = w;
    }
}
// bring the largest mantissa to [1, 2)
void normalise(const Arith &A, Blk &x) {
    int top = -1;
    for (const Fx &f : x.v) top = std::max(top, A.msb(f));
    if (top < 0) { x.e = 0; for (Fx &f : x.v) f.neg = false; return; }
    const int64_t s = (int64_t)top - 64LL * A.nf;
    if (s > 0) for (Fx &f : x.v) shr(A, f, s);
    else if (s < 0) for (Fx &f : x.v) shl(A, f, -s);
    x.e += s;
}
bool blk_zero(const Blk &x) {
    for (const Fx &f : x.v) if (!Arith::is_zero(f)) return false;
    return true;
}
// x += y (same K), aligned on the larger exponent, then normalised
void blk_add(const Arith &A, Blk &x, const Blk &y, Blk &tmp) {
    if (blk_zero(y)) return;
    if (blk_zero(x)) { x = y; return; }
    tmp = y;
    if (x.e >= tmp.e) { for (Fx &f : tmp.v) shr(A, f, x.e - tmp.e); }
    else { for (Fx &f : x.v) shr(A, f, tmp.e - x.e); x.e = tmp.e; }
    for (size_t k = 0; k < x.v.size(); k++) A.add(x.v[k], x.v[k], tmp.v[k]);
    normalise(A, x);
}
// |a| * k (exact), k small
void mul_small(const Arith &A, Fx &a, u64 k) {
    u64 carry = 0;
    for (int i = 0; i < A.n; i++) {
        const u128 t = (u128)a.m[i] * k + carry;
        a.m[i] = (u64)t;
        carry = (u64)(t >> 64);
    }
    if (carry) A.flag_overflow();
}

// binary64 mantissa with a wide exponent, for the ball-method test
struct XE { double m; int64_t e; };
XE xe(double m, int64_t e) {
    if (m == 0.) return {0., 0};
    int k;
    const double f = std::frexp(m, &k);
    return {f, e + k};
}
XE xe_of(const Arith &A, const Fx &a, int64_t e) {
    double m;
    int32_t k;
    A.to_xr(a, m, k);
    return xe(m, (int64_t)k + e);
}
XE xe_mul(XE a, XE b) { return xe(a.m * b.m, a.e + b.e); }
XE xe_div(XE a, XE b) { return xe(a.m / b.m, a.e - b.e); }
XE xe_add(XE a, XE b) {
    if (a.m == 0.) return b;
    if (b.m == 0.) return a;
    const int64_t E = std::max(a.e, b.e);
    const auto sh = [](double m, int64_t s) { return s < -2000 ? 0. : std::ldexp(m, (int)s); };
    return xe(sh(a.m, a.e - E) + sh(b.m, b.e - E), E);
}
XE xe_neg(XE a) { a.m = -a.m; return a; }
XE xe_hypot(XE a, XE b) {
    if (a.m == 0.) return {std::fabs(b.m), b.e};
    if (b.m == 0.) return {std::fabs(a.m), a.e};
    const int64_t E = std::max(a.e, b.e);
    const auto sh = [](double m, int64_t s) { return s < -2000 ? 0. : std::ldexp(m, (int)s); };
    return xe(std::hypot(sh(a.m, a.e - E), sh(b.m, b.e - E)), E);
}
bool xe_lt1(XE a) { return a.m == 0. || a.e <= 0; }   // |a| < 1 (m in [0.5, 1))
double xe_double(XE a) { return a.e < -1100 ? 0. : std::ldexp(a.m, (int)std::min<int64_t>(a.e, 1100)); }

// One step of the holomorphic derivative, dz <- d z^(d-1) dz + 1
// (FP_loop.pyx iter_Mn_deriv, with k = 1)
struct HoloDeriv {
    Blk dz, one, tmp;
    Cx zp, prod;
    explicit HoloDeriv(const Arith &A) {
        dz.v = {A.make(), A.make()};
        one.v = {A.make(), A.make()};
        one.v[0].m[A.nf] = 1;         // 1 = 2^(64 nf) on the fixed-point grid
        tmp.v = {A.make(), A.make()};
        zp.re = A.make(); zp.im = A.make(); prod.re = A.make(); prod.im = A.make();
    }
    void step(Arith &A, const Cx &z, int deg, Work &w) {
        if (!blk_zero(dz)) {
            const Cx *base = &z;
            if (deg > 2) { cpow(A, zp, z, deg - 1, w); base = &zp; }
            Cx m;
            m.re = dz.v[0]; m.im = dz.v[1];
            cmul(A, prod, *base, m, w.t1, w.t2, w.t3, w.k1);
            dz.v[0] = prod.re; dz.v[1] = prod.im;
            mul_small(A, dz.v[0], (u64)deg);
            mul_small(A, dz.v[1], (u64)deg);
            normalise(A, dz);
        }
        blk_add(A, dz, one, tmp);
    }
};

// Burning-ship Jacobian of (x, y) (burning_ship.py:440-532), entries
// (dfx/dx, dfx/dy, dfy/dx, dfy/dy) as fixed-point values
void bs_jac(const Arith &A, int flavor, const Fx &x, const Fx &y, Fx J[4]) {
    const auto sgn_neg = [](const Fx &v) { return v.neg; };    // sgn(v) = -1 iff v < 0
    Fx ax = x, ay = y;
    ax.neg = false; ay.neg = false;
    // q = sgn(x^2 - y^2) = sgn((x + y)(x - y)), from the magnitudes
    const int c = A.cmp_mag(x, y);
    const bool q_neg = c < 0;
    const auto twice_signed = [&](Fx &r, const Fx &v, bool flip) {
        A.twice(r, v);
        if (flip && !Arith::is_zero(r)) r.neg = !r.neg;
    };
    switch (flavor) {
    case 1:
        twice_signed(J[0], x, false); twice_signed(J[1], y, true);
        twice_signed(J[2], ay, sgn_neg(x)); twice_signed(J[3], ax, sgn_neg(y));
        break;
    case 2:
        twice_signed(J[0], x, false); twice_signed(J[1], y, true);
        twice_signed(J[2], ay, false); twice_signed(J[3], x, sgn_neg(y));
        break;
    case 3:
        twice_signed(J[0], x, false); twice_signed(J[1], ay, true);
        twice_signed(J[2], y, false); twice_signed(J[3], x, false);
        break;
    case 4:
        twice_signed(J[0], x, q_neg); twice_signed(J[1], y, !q_neg);
        twice_signed(J[2], y, false); twice_signed(J[3], x, false);
        break;
    default:
        twice_signed(J[0], x, q_neg); twice_signed(J[1], y, !q_neg);
        twice_signed(J[2], ay, sgn_neg(x)); twice_signed(J[3], ax, sgn_neg(y));
        break;
    }
}

// D <- J D + diag(1, -1), D = d(x, y)/d(a, b) of the burning ship
struct BsDeriv {
    Blk D, one, tmp;
    Fx J[4], p, q;
    explicit BsDeriv(const Arith &A) {
        D.v = {A.make(), A.make(), A.make(), A.make()};
        one.v = {A.make(), A.make(), A.make(), A.make()};
        one.v[0].m[A.nf] = 1;
        one.v[3].m[A.nf] = 1;
        one.v[3].neg = true;
        tmp.v = D.v;
        for (Fx &f : J) f = A.make();
        p = A.make(); q = A.make();
    }
    void step(Arith &A, int flavor, const Fx &x, const Fx &y) {
        if (!blk_zero(D)) {
            bs_jac(A, flavor, x, y, J);
            Blk r = D;
            for (int row = 0; row < 2; row++)
                for (int col = 0; col < 2; col++) {
                    A.mul(p, J[2 * row], D.v[col]);
                    A.mul(q, J[2 * row + 1], D.v[2 + col]);
                    A.add(r.v[2 * row + col], p, q);
                }
            D = r;
            normalise(A, D);
        }
        blk_add(A, D, one, tmp);
    }
};

int fsg_ref_orbit_impl(const fsg_orbit_desc *d, fsg_orbit_out *o) {
    int rc = check_desc(d, 0.);
    if (rc) return rc;
    if (!o || !o->Zn) return set_error(FSG_E_INVALID, "bad orbit output");
    if (o->xr_cap < 0 || (o->xr_cap > 0 && (!o->xr_index || !o->xr_mant || !o->xr_exp)))
        return set_error(FSG_E_INVALID, "bad Xrange register buffers");
    const bool holo = d->kind != FSG_KIND_BS;
    const int deg = d->kind == FSG_KIND_M2 ? 2 : d->degree;
    std::unique_ptr<Team> team = make_team(d);
    Arith A(d->n_limbs, d->frac_limbs, team.get());
    Work w(A);
    const Fx cre = load(A, d->c_re, d->c_re_neg), cim = load(A, d->c_im, d->c_im_neg);
    Cx z, zp;
    z.re = A.make(); z.im = A.make(); zp.re = A.make(); zp.im = A.make();

    double *Zn = o->Zn;
    Zn[0] = 0.; Zn[1] = 0.;
    int64_t n_xr = 0;
    int64_t i = 0;
    double abs_i = 0.;
    const int64_t max_iter = d->max_iter;
    for (i = 1; i <= max_iter; i++) {
        if (holo) {
            cpow(A, zp, z, deg, w);
            A.add(z.re, zp.re, cre);
            A.add(z.im, zp.im, cim);
        } else {
            bs_step(A, d->flavor, z.re, z.im, cre, cim, w);
        }
        if (A.overflowed()) return set_error(FSG_E_INVALID, "orbit overflow: escape radius too large");
        const double x = A.to_double(z.re), y = A.to_double(z.im);
        Zn[2 * i] = x;
        Zn[2 * i + 1] = y;
        abs_i = std::hypot(x, y);
        if (abs_i > d->M) break;
        const bool small = holo ? abs_i < XR_TSHOLD : (std::fabs(x) < XR_TSHOLD || std::fabs(y) < XR_TSHOLD);
        if (d->need_xr && small) {
            if (n_xr < o->xr_cap) {
                double rm, im;
                int32_t re_e, im_e;
                A.to_xr(z.re, rm, re_e);
                A.to_xr(z.im, im, im_e);
                if (holo) {
                    // complex Xrange: both parts on the larger exponent
                    // (FP_loop.pyx:424 mpc_t_to_Xrange)
                    int32_t e = 0;
                    if (rm == 0. && im == 0.) e = 0;
                    else if (rm == 0.) e = im_e;
                    else if (im == 0.) e = re_e;
                    else {
                        e = re_e > im_e ? re_e : im_e;
                        rm = std::ldexp(rm, re_e - e);
                        im = std::ldexp(im, im_e - e);
                    }
                    re_e = im_e = e;
                }
                o->xr_index[n_xr] = i;
                o->xr_mant[2 * n_xr] = rm;
                o->xr_mant[2 * n_xr + 1] = im;
                o->xr_exp[2 * n_xr] = re_e;
                o->xr_exp[2 * n_xr + 1] = im_e;
            }
            n_xr++;
        }
    }
    // first invalid index (FP_loop.pyx:398-402)
    if (i > max_iter) i = max_iter + 1;
    for (int64_t k = i + 1; k <= max_iter; k++) Zn[2 * k] = Zn[2 * k + 1] = 0.;
    o->div_iter = i;
    o->n_xr = n_xr;
    if (n_xr > o->xr_cap) return set_error(FSG_E_RANGE, "Xrange register larger than xr_cap");
    return 0;
}


int fsg_ball_period_impl(const fsg_orbit_desc *d, double px_mant, int64_t px_exp, int64_t *order) {
    int rc = check_desc(d, 8.);
    if (rc) return rc;
    if (!order || !(px_mant > 0.) || !std::isfinite(px_mant))
        return set_error(FSG_E_INVALID, "bad ball-method arguments");
    const bool holo = d->kind != FSG_KIND_BS;
    const int deg = d->kind == FSG_KIND_M2 ? 2 : d->degree;
    std::unique_ptr<Team> team = make_team(d);
    Arith A(d->n_limbs, d->frac_limbs, team.get());
    Work w(A);
    const Fx cre = load(A, d->c_re, d->c_re_neg), cim = load(A, d->c_im, d->c_im_neg);
    const XE px = xe(px_mant, px_exp);
    *order = -1;
    if (holo) {
        Cx z, zn;
        z.re = A.make(); z.im = A.make(); zn.re = A.make(); zn.im = A.make();
        HoloDeriv D(A);
        for (int64_t i = 1; i <= d->max_iter; i++) {
            D.step(A, z, deg, w);
            cpow(A, zn, z, deg, w);
            A.add(z.re, zn.re, cre);
            A.add(z.im, zn.im, cim);
            if (A.overflowed()) return set_error(FSG_E_INVALID, "overflow: escape radius too large");
            if (std::hypot(A.to_double(z.re), A.to_double(z.im)) > d->M) break;
            const XE az = xe_hypot(xe_of(A, z.re, 0), xe_of(A, z.im, 0));
            const XE adz = xe_hypot(xe_of(A, D.dz.v[0], D.dz.e), xe_of(A, D.dz.v[1], D.dz.e));
            if (xe_lt1(xe_div(xe_div(az, adz), px))) { *order = i; break; }
        }
    } else {
        Fx x = A.make(), y = A.make();
        BsDeriv D(A);
        for (int64_t i = 1; i <= d->max_iter; i++) {
            D.step(A, d->flavor, x, y);
            bs_step(A, d->flavor, x, y, cre, cim, w);
            if (A.overflowed()) return set_error(FSG_E_INVALID, "overflow: escape radius too large");
            if (std::hypot(A.to_double(x), A.to_double(y)) > d->M) break;
            XE m[4];
            for (int k = 0; k < 4; k++) m[k] = xe_of(A, D.D.v[k], D.D.e);
            const XE det = xe_add(xe_mul(m[0], m[3]), xe_neg(xe_mul(m[2], m[1])));
            if (det.m == 0.) continue;
            const XE X = xe_of(A, x, 0), Y = xe_of(A, y, 0);
            const XE rx = xe_div(xe_div(xe_add(xe_mul(m[3], X), xe_neg(xe_mul(m[1], Y))), det), px);
            const XE ry = xe_div(xe_div(xe_add(xe_mul(m[0], Y), xe_neg(xe_mul(m[2], X))), det), px);
            if (xe_lt1(rx) && xe_lt1(ry) && std::hypot(xe_double(rx), xe_double(ry)) < 1.) { *order = i; break; }
        }
    }
    return 0;
}

int fsg_newton_eval_impl(const fsg_orbit_desc *d, int64_t order, int32_t divisors, fsg_mp_values *out) {
    int rc = check_desc(d, 8.);
    if (rc) return rc;
    if (!out || order < 1 || out->cap < 0 || (out->cap > 0 && (!out->limbs || !out->neg || !out->exp)))
        return set_error(FSG_E_INVALID, "bad Newton arguments");
    const bool holo = d->kind != FSG_KIND_BS;
    const int deg = d->kind == FSG_KIND_M2 ? 2 : d->degree;
    std::unique_ptr<Team> team = make_team(d);
    Arith A(d->n_limbs, d->frac_limbs, team.get());
    Work w(A);
    const Fx cre = load(A, d->c_re, d->c_re_neg), cim = load(A, d->c_im, d->c_im_neg);
    int64_t n = 0;
    const auto put = [&](const Fx &v, int64_t e) {
        if (n < out->cap) {
            std::memcpy(out->limbs + n * A.n, v.m.data(), sizeof(u64) * A.n);
            out->neg[n] = v.neg ? 1 : 0;
            out->exp[n] = e;
        }
        n++;
    };
    if (holo) {
        Cx z, zn;
        z.re = A.make(); z.im = A.make(); zn.re = A.make(); zn.im = A.make();
        HoloDeriv D(A);
        for (int64_t i = 1; i <= order; i++) {
            D.step(A, z, deg, w);
            cpow(A, zn, z, deg, w);
            A.add(z.re, zn.re, cre);
            A.add(z.im, zn.im, cim);
            if (A.overflowed()) return set_error(FSG_E_RANGE, "orbit left the fixed-point range");
            if (i == order || (divisors && order % i == 0)) {
                put(z.re, 0); put(z.im, 0); put(D.dz.v[0], D.dz.e); put(D.dz.v[1], D.dz.e);
            }
        }
    } else {
        Fx x = A.make(), y = A.make();
        BsDeriv D(A);
        for (int64_t i = 1; i <= order; i++) {
            D.step(A, d->flavor, x, y);
            bs_step(A, d->flavor, x, y, cre, cim, w);
            if (A.overflowed()) return set_error(FSG_E_RANGE, "orbit left the fixed-point range");
        }
        put(x, 0); put(y, 0);
        for (int k = 0; k < 4; k++) put(D.D.v[k], D.D.e);
    }
    out->n = n;
    if (n > out->cap) return set_error(FSG_E_INVALID, "value buffer too small");
    return 0;
}

// C entry points: exceptions (allocation, thread creation) become error codes
template <class F> int guarded(F &&f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return set_error(FSG_E_NOMEM, "out of host memory");
    } catch (const std::exception &e) {
        return set_error(FSG_E_INVALID, e.what());
    }
}

}  // namespace

extern "C" {

int fsg_ref_orbit(const fsg_orbit_desc *d, fsg_orbit_out *o) {
    return guarded([&] { return fsg_ref_orbit_impl(d, o); });
}

int fsg_ball_period(const fsg_orbit_desc *d, double px_mant, int64_t px_exp, int64_t *order) {
    return guarded([&] { return fsg_ball_period_impl(d, px_mant, px_exp, order); });
}

int fsg_newton_eval(const fsg_orbit_desc *d, int64_t order, int32_t divisors, fsg_mp_values *out) {
    return guarded([&] { return fsg_newton_eval_impl(d, order, divisors, out); });
}

}  // extern "C"
