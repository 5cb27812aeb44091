// fsg_api.hip -- C ABI of libfsgpu.so (declared in include/fsgpu.h).
//
// A context owns the device copy of one reference: orbit, Xrange register,
// orbit-derivative paths (computed here on the host, serially, with the same
// Xrange code as the kernels) and the BLA tree (built on the GPU).
#include "fsgpu.h"
#include "models.hpp"
#include "xl.hpp"

#include <memory>
#include <mutex>
#include <string>
#include <vector>
#include <cstring>
#include <climits>
#include <chrono>
#include <thread>

using namespace fsg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) { g_err = msg; return code; }

#define FSG_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(FSG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

__global__ void k_zero_u64(unsigned long long *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// stop_reason = -1 (not computed), the reference's initial value (core.py:2072)
__global__ void k_fill_i8(int8_t *p, int64_t n, int8_t v) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}


template <class T> T *dalloc(size_t n, hipError_t &err) {
    void *p = nullptr;
    if (n == 0) n = 1;
    err = hipMalloc(&p, n * sizeof(T));
    return static_cast<T *>(p);
}

// Tables::zn_k: per orbit point, the exponent k that xl_n / xlf_n (xl.hpp)
// take to normalise its binary64 value (larger component, or the one
// component, scaled into [1, 2); the exponent of zero for a zero value), or
// INT32_MIN + its Xrange register slot.  Holomorphic: one code per point
// (the complex value); burning ship: two (x, then y).
void point_codes(const fsg_ref_desc *d, int w, std::vector<int32_t> &code) {
    const int64_t L = d->ref_len;
    code.assign((size_t)(w * L), 0);
    const auto norm_exp = [](double mx) {
        int k0 = 0;
        if (mx != 0.) (void)std::frexp(mx, &k0);
        return (mx == 0.) ? XL_ZERO_E : k0 - 1;
    };
    for (int64_t i = 0; i < L; i++) {
        const double x = std::fabs(d->Zn_path[2 * i]), y = std::fabs(d->Zn_path[2 * i + 1]);
        if (w == 1) {
            code[(size_t)i] = norm_exp(std::max(x, y));
        } else {
            code[(size_t)(2 * i)] = norm_exp(x);
            code[(size_t)(2 * i + 1)] = norm_exp(y);
        }
    }
    for (int32_t j = 0; j < d->n_xr; j++) {
        const int64_t i = d->xr_index[j];
        if (i >= 0 && i < L) code[(size_t)(w * i)] = INT32_MIN + j;
    }
}

// ---- host-side orbit derivative paths (perturbation.py:2282-2516) ----

cplx dfdz_rt(int kind, int deg, cplx z) {
    if (kind == KIND_M2) return scale(2., z);
    cplx tmp = z;
    for (int q = 2; q < deg; q++) tmp = tmp * z;
    return scale((double)deg, tmp);
}
xc dfdz_rt(int kind, int deg, xc z) {
    if (kind == KIND_M2) return 2. * z;
    xc tmp = z;
    for (int q = 2; q < deg; q++) tmp = tmp * z;
    return (double)deg * tmp;
}

template <class T> void jac_rt(int fl, T x, T y, T J[4]) {
    switch (fl) {
        case 1: bs_jac<1>(x, y, J); break;
        case 2: bs_jac<2>(x, y, J); break;
        case 3: bs_jac<3>(x, y, J); break;
        case 4: bs_jac<4>(x, y, J); break;
        default: bs_jac<5>(x, y, J); break;
    }
}

struct HostRef {
    const fsg_ref_desc *d;
    int32_t L, div_iter, order;
    cplx Z(int64_t i) const { return mk(d->Zn_path[2 * i], d->Zn_path[2 * i + 1]); }
    int32_t find(int64_t i) const {
        int32_t lo = 0, hi = d->n_xr;
        while (lo < hi) {
            int32_t mid = (lo + hi) >> 1;
            if (d->xr_index[mid] < i) lo = mid + 1; else hi = mid;
        }
        return (lo < d->n_xr && d->xr_index[lo] == i) ? lo : -1;
    }
    // ref_path_get + ensure_xr (perturbation.py:2182-2193, 2519-2588)
    xc Zxr(int64_t i) const {
        int32_t j = d->n_xr ? find(i) : -1;
        if (j >= 0) return mkc(d->xr_mant[2 * j], d->xr_mant[2 * j + 1], d->xr_exp[j]);
        return to_xc(Z(i));
    }
    void ZxrBS(int64_t i, xf &X, xf &Y) const {
        int32_t j = d->n_xr ? find(i) : -1;
        if (j >= 0) {
            X = mkf(d->xr_mant[2 * j], d->xr_exp[2 * j]);
            Y = mkf(d->xr_mant[2 * j + 1], d->xr_exp[2 * j + 1]);
        } else {
            X = to_xf(Z(i).re); Y = to_xf(Z(i).im);
        }
    }
};

// the scale of the dz/dc paths: dx, times the projection's scale
xf deriv_scale(const fsg_ref_desc *d) {
    return d->deriv_mant != 0. ? mkf(d->deriv_mant, d->deriv_exp) : mkf(d->dx_mant, d->dx_exp);
}

// Entries of the Xrange path tables whose mantissa is zero carry the zero
// exponent code XL_ZERO_E, so that the kernels load a path value as it is,
// without a zero test per load (the value and its scaling are unchanged: a
// zero mantissa is zero at any exponent)
void zero_codes(const std::vector<cplx> &m, std::vector<int32_t> &e) {
    for (size_t k = 0; k < m.size(); k++)
        if (m[k].re == 0. && m[k].im == 0.) e[k] = XL_ZERO_E;
}
void zero_codes(const std::vector<double> &m, std::vector<int32_t> &e) {
    for (size_t k = 0; k < m.size(); k++)
        if (m[k] == 0.) e[k] = XL_ZERO_E;
}

// dZndc: D_0 = 0, D_i = f'(Z_{i-1}) D_{i-1} + s, wrapped value at [0] for cycles
void holo_dzndc(const HostRef &h, std::vector<cplx> &m, std::vector<int32_t> &e) {
    const fsg_ref_desc *d = h.d;
    const int64_t L = h.L, valid = std::min<int64_t>(h.L, h.div_iter);
    m.assign(L, mk(0., 0.)); e.assign(L, 0);
    const xf s = deriv_scale(d);
    if (!d->xr_mode) {
        const double ss = to_std(s);
        for (int64_t i = 1; i < valid; i++) m[i] = dfdz_rt(d->kind, d->degree, h.Z(i - 1)) * m[i - 1] + mk(ss, 0.);
        int64_t i = valid - 1;
        if (i == (int64_t)d->ref_order - 1) m[0] = dfdz_rt(d->kind, d->degree, h.Z(i)) * m[i] + mk(ss, 0.);
    } else {
        std::vector<xc> D(L, mkc(0., 0., 0));
        for (int64_t i = 1; i < valid; i++) D[i] = dfdz_rt(d->kind, d->degree, h.Zxr(i - 1)) * D[i - 1] + s;
        int64_t i = valid - 1;
        if (i == (int64_t)d->ref_order - 1) D[0] = D[i] * dfdz_rt(d->kind, d->degree, h.Z(i)) + s;
        for (int64_t k = 0; k < L; k++) { m[k] = mk(D[k].re, D[k].im); e[k] = D[k].e; }
        zero_codes(m, e);
    }
}

// dZndz: [1] = 1, D_i = f'(Z_{i-1}) D_{i-1}, wrapped value stored at [L]
void holo_dzndz(const HostRef &h, std::vector<cplx> &m, std::vector<int32_t> &e) {
    const fsg_ref_desc *d = h.d;
    const int64_t L = h.L, valid = std::min<int64_t>(h.L, h.div_iter);
    m.assign(L + 1, mk(0., 0.)); e.assign(L + 1, 0);
    if (!d->xr_mode) {
        m[1] = mk(1., 0.);
        for (int64_t i = 2; i < valid; i++) m[i] = dfdz_rt(d->kind, d->degree, h.Z(i - 1)) * m[i - 1];
        int64_t i = valid - 1;
        m[L] = dfdz_rt(d->kind, d->degree, h.Z(i)) * m[i];
    } else {
        std::vector<xc> D(L + 1, mkc(0., 0., 0));
        D[1] = mkc(1., 0., 0);
        for (int64_t i = 2; i < valid; i++) D[i] = dfdz_rt(d->kind, d->degree, h.Zxr(i - 1)) * D[i - 1];
        int64_t i = valid - 1;
        D[L] = dfdz_rt(d->kind, d->degree, h.Zxr(i)) * D[i];
        for (int64_t k = 0; k <= L; k++) { m[k] = mk(D[k].re, D[k].im); e[k] = D[k].e; }
        zero_codes(m, e);
    }
}

// Burning-ship Jacobian paths (numba_dZndc_path_BS + incr_deriv_ref_BS)
void bs_paths(const HostRef &h, std::vector<double> P[4], std::vector<int32_t> E[4]) {
    const fsg_ref_desc *d = h.d;
    const int64_t L = h.L, valid = std::min<int64_t>(h.L, h.div_iter);
    for (int q = 0; q < 4; q++) { P[q].assign(L, 0.); E[q].assign(L, 0); }
    if (!d->calc_dzndc) return;
    const xf s = deriv_scale(d);
    auto incr_std = [&](int64_t from, int64_t to) {
        double J[4];
        cplx R = h.Z(from);
        jac_rt<double>(d->flavor, R.re, R.im, J);
        double xa = P[0][from], xb = P[1][from], ya = P[2][from], yb = P[3][from];
        const double ss = to_std(s);
        P[0][to] = J[0] * xa + J[1] * ya + ss;
        P[1][to] = J[0] * xb + J[1] * yb;
        P[2][to] = J[2] * xa + J[3] * ya;
        P[3][to] = J[2] * xb + J[3] * yb - ss;
    };
    if (!d->xr_mode) {
        for (int64_t i = 1; i < valid; i++) incr_std(i - 1, i);
        if (valid - 1 == (int64_t)d->ref_order - 1) incr_std(valid - 1, 0);
        return;
    }
    std::vector<xf> X[4];
    for (int q = 0; q < 4; q++) X[q].assign(L, mkf(0., 0));
    auto incr_xr = [&](int64_t from, int64_t to) {
        xf Rx, Ry, J[4];
        h.ZxrBS(from, Rx, Ry);
        jac_rt<xf>(d->flavor, Rx, Ry, J);
        xf xa = X[0][from], xb = X[1][from], ya = X[2][from], yb = X[3][from];
        X[0][to] = J[0] * xa + J[1] * ya + s;
        X[1][to] = J[0] * xb + J[1] * yb;
        X[2][to] = J[2] * xa + J[3] * ya;
        X[3][to] = J[2] * xb + J[3] * yb - s;
    };
    for (int64_t i = 1; i < valid; i++) incr_xr(i - 1, i);
    if (valid - 1 == (int64_t)d->ref_order - 1) incr_xr(valid - 1, 0);
    for (int q = 0; q < 4; q++) {
        for (int64_t k = 0; k < L; k++) { P[q][k] = X[q][k].m; E[q][k] = X[q][k].e; }
        zero_codes(P[q], E[q]);
    }
}

int stages_of(int64_t L) {
    int s = 0;
    while ((1LL << s) < L) s++;
    return s;
}

}  // namespace

struct fsg_ctx {
    fsg_ref_desc d;
    int32_t L, div_iter, order, first_invalid;
    int device;
    // device tables
    cplx *Zn = nullptr, *dc = nullptr, *dz = nullptr;
    int32_t *dc_e = nullptr, *dz_e = nullptr;
    double *bs[4] = {nullptr, nullptr, nullptr, nullptr};
    int32_t *bs_e[4] = {nullptr, nullptr, nullptr, nullptr};
    int32_t *xr_idx = nullptr, *xr_e = nullptr, *zn_k = nullptr;
    int32_t zn_k_w = 0;
    double *xr_m = nullptr;
    double *M = nullptr, *r = nullptr;
    int64_t bla_len = 0;
    TablesX tx = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int32_t stages = 0;
    // staging for fsg_tile
    std::mutex mu;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // interrupt flag the tile kernel reads at each refill: pinned host
    // memory, written by the host while it waits for a tile
    uint32_t *stop_h = nullptr;
    const uint32_t *stop_d = nullptr;
    char *buf = nullptr;
    size_t buf_bytes = 0;
    unsigned long long *work = nullptr;
    std::vector<void *> owned;
    // orbit, Xrange register and BLA coefficients: owned jointly by every
    // context created on the same reference (fsg_ctx_create_shared)
    struct Shared {
        std::vector<void *> p;
        ~Shared() { for (void *q : p) if (q) (void)hipFree(q); }
    };
    std::shared_ptr<Shared> shared;
    uint64_t ref_hash = 0;

    ~fsg_ctx() {
        for (void *p : owned) if (p) (void)hipFree(p);
        if (buf) (void)hipFree(buf);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stop_h) (void)hipHostFree(stop_h);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// FNV-1a of the reference data a shared context must agree on
// (64-bit words, then the tail bytes)
uint64_t fnv(uint64_t h, const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        h ^= w; h *= 1099511628211ull;
    }
    for (; i < n; i++) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}
uint64_t ref_hash_of(const fsg_ref_desc *d) {
    uint64_t h = 1469598103934665603ull;
    const int32_t key[4] = {d->kind, d->kind == FSG_KIND_M2 ? 2 : d->degree, d->flavor, d->n_xr};
    h = fnv(h, key, sizeof(key));
    h = fnv(h, d->Zn_path, (size_t)d->ref_len * 16);
    if (d->n_xr > 0) {
        const bool holo = d->kind != FSG_KIND_BS;
        h = fnv(h, d->xr_index, (size_t)d->n_xr * 4);
        h = fnv(h, d->xr_mant, (size_t)d->n_xr * 16);
        h = fnv(h, d->xr_exp, (size_t)d->n_xr * (holo ? 4 : 8));
    }
    return h;
}

template <class T> int upload(fsg_ctx *c, T *&dst, const T *src, size_t n, bool shared = false) {
    hipError_t err;
    dst = dalloc<T>(n, err);
    if (err != hipSuccess) return fail(FSG_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(err));
    (shared ? c->shared->p : c->owned).push_back(dst);
    if (n) FSG_HIP(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

int validate(const fsg_ref_desc *d) {
    if (!d) return fail(FSG_E_INVALID, "null descriptor");
    if (d->kind < FSG_KIND_M2 || d->kind > FSG_KIND_BS) return fail(FSG_E_INVALID, "unknown kind");
    if (d->kind == FSG_KIND_MN && (d->degree < 2 || d->degree > 8))
        return fail(FSG_E_UNSUPPORTED, "Mandelbrot_N degree must be in 2..8");
    if (d->kind == FSG_KIND_BS && (d->flavor < 1 || d->flavor > 5))
        return fail(FSG_E_INVALID, "burning-ship flavour must be in 1..5");
    if (d->ref_len < 2 || d->ref_len > INT32_MAX / 4) return fail(FSG_E_INVALID, "ref_len out of range");
    if (!d->Zn_path) return fail(FSG_E_INVALID, "null Zn_path");
    if (d->n_xr < 0 || (d->n_xr > 0 && (!d->xr_index || !d->xr_mant || !d->xr_exp)))
        return fail(FSG_E_INVALID, "bad Xrange register");
    for (int32_t j = 1; j < d->n_xr; j++)
        if (d->xr_index[j] <= d->xr_index[j - 1]) return fail(FSG_E_INVALID, "Xrange register not sorted");
    if (d->max_iter < 1 || d->max_iter + d->ref_len >= (int64_t)INT32_MAX)
        return fail(FSG_E_INVALID, "max_iter + ref_len must stay below 2^31");
    if (d->ref_div_iter < 1 || d->ref_div_iter > (int64_t)INT32_MAX)
        return fail(FSG_E_INVALID, "ref_div_iter out of range");
    if (d->ref_order < 1) return fail(FSG_E_INVALID, "ref_order must be >= 1");
    if (d->precision != FSG_PREC_FP64 && d->precision != FSG_PREC_FP32X)
        return fail(FSG_E_INVALID, "unknown precision mode");
    if (d->calc_orbit && (d->backshift < 1 || d->backshift > (1 << 20)))
        return fail(FSG_E_INVALID, "calc_orbit needs backshift in 1..2^20");
    if (d->calc_orbit && d->precision != FSG_PREC_FP64)
        return fail(FSG_E_UNSUPPORTED, "calc_orbit is computed by the FP64 kernels only");
    if (d->bla_on && stages_of(d->ref_len) <= 3)
        return fail(FSG_E_UNSUPPORTED, "BLA needs a reference orbit longer than 8");
    int64_t valid = std::min(d->ref_len, d->ref_div_iter);
    if (d->kind != FSG_KIND_BS && d->calc_dzndc && valid < 2) return fail(FSG_E_INVALID, "orbit too short for dzndc");
    if (d->kind != FSG_KIND_BS && d->interior_detect && valid < 3) return fail(FSG_E_INVALID, "orbit too short for dzndz");
    if (d->kind == FSG_KIND_BS && d->calc_dzndc && valid < 2) return fail(FSG_E_INVALID, "orbit too short for hessian");
    return 0;
}

int fill_args(fsg_ctx *c, KArgs &a) {
    std::memset(&a, 0, sizeof(a));
    Tables &t = a.t;
    t.Zn = c->Zn; t.L = c->L; t.div_iter = c->div_iter; t.order = c->order; t.first_invalid = c->first_invalid;
    t.dc = c->dc; t.dc_e = c->dc_e; t.dz = c->dz; t.dz_e = c->dz_e;
    for (int q = 0; q < 4; q++) { t.bs[q] = c->bs[q]; t.bs_e[q] = c->bs_e[q]; }
    t.n_xr = c->d.n_xr; t.xr_idx = c->xr_idx; t.xr_m = c->xr_m; t.xr_e = c->xr_e;
    t.zn_k = c->zn_k; t.zn_k_w = c->zn_k_w;
    t.M = c->M; t.r = c->r; t.stages_bla = c->stages; t.bla_on = c->d.bla_on;
    a.tx = c->tx;
    for (int q = 0; q < 4; q++) a.lin[q] = c->d.lin_mat[q];
    a.dx_m = c->d.dx_mant; a.dx_e = c->d.dx_exp;
    a.drift_m[0] = c->d.drift_mant[0]; a.drift_m[1] = c->d.drift_mant[1];
    a.drift_e[0] = c->d.drift_exp[0]; a.drift_e[1] = c->d.drift_exp[1];
    a.max_iter = (int32_t)c->d.max_iter;
    a.calc_dc = c->d.calc_dzndc ? 1 : 0;
    a.calc_dz = (c->d.kind != FSG_KIND_BS && c->d.interior_detect) ? 1 : 0;
    a.M2 = c->d.M_divergence * c->d.M_divergence;
    a.eps2 = c->d.epsilon_stationnary * c->d.epsilon_stationnary;
    a.calc_orbit = c->d.calc_orbit ? 1 : 0;
    a.backshift = c->d.calc_orbit ? c->d.backshift : 1;
    if (c->d.kind == FSG_KIND_BS) a.nz = 2 + (a.calc_dc ? 4 : 0) + 2 * a.calc_orbit;
    else a.nz = 1 + a.calc_dc + a.calc_dz + a.calc_orbit;
    return 0;
}

LaunchCfg cfg_of(const fsg_ctx *c) {
    LaunchCfg g;
    g.kind = c->d.kind; g.degree = c->d.degree; g.flavor = c->d.flavor; g.xr = c->d.xr_mode ? 1 : 0;
    g.precision = c->d.precision;
    g.orbit = c->d.calc_orbit ? 1 : 0;
    return g;
}

int n_fields(const fsg_ctx *c) {
    const int orb = c->d.calc_orbit ? 1 : 0;
    return (c->d.kind == FSG_KIND_BS) ? 2 + (c->d.calc_dzndc ? 4 : 0) + 2 * orb
                                      : 1 + (c->d.calc_dzndc ? 1 : 0) + (c->d.interior_detect ? 1 : 0) + orb;
}

size_t z_bytes(const fsg_ctx *c, int64_t n) {
    int nz = n_fields(c);
    return (size_t)nz * (size_t)n * (c->d.kind == FSG_KIND_BS ? 8 : 16);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int launch(fsg_ctx *c, const double *pix, int64_t n, double *Z, int32_t *U, int8_t *sr, int32_t *si,
           unsigned long long *work, hipStream_t stream, bool timed, int64_t ld = -1, int max_blocks = 0,
           int64_t grid_w = 0, const uint32_t *stop = nullptr) {
    KArgs a;
    fill_args(c, a);
    a.pix = pix; a.npts = n; a.ld = ld < 0 ? n : ld; a.Z = Z; a.U = U; a.stop_reason = sr; a.stop_iter = si;
    a.work = work; a.max_blocks = max_blocks; a.grid_w = grid_w; a.stop = stop;
    hipLaunchKernelGGL(k_zero_u64, dim3((WORK_WORDS + 255) / 256), dim3(256), 0, stream, work, WORK_WORDS);
    FSG_HIP(hipGetLastError());
    if (n == 0) return 0;
    if (timed) FSG_HIP(hipEventRecord(c->ev0, stream));
    if (c->d.precision == FSG_PREC_FP32X) {
        FSG_HIP(launch_tile_f32(cfg_of(c), a, stream));
        if (c->d.kind != FSG_KIND_BS) {
            // the pixels the binary32 pass left unsure, again in binary64
            KArgs r = a;
            r.redo = 1;
            LaunchCfg cf = cfg_of(c);
            cf.precision = FSG_PREC_FP64;
            FSG_HIP(launch_tile(cf, r, stream));
        }
    } else {
        FSG_HIP(launch_tile(cfg_of(c), a, stream));
    }
    if (timed) FSG_HIP(hipEventRecord(c->ev1, stream));
    return 0;
}

}  // namespace

// ---- internal entry points of the serving loop (fsg_serve.hip) ----
namespace fsg {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
int ctx_holomorphic(const fsg_ctx *c) { return c->d.kind != FSG_KIND_BS; }
int ctx_nz(const fsg_ctx *c) { return n_fields(c); }
size_t ctx_z_bytes(const fsg_ctx *c, int64_t n) { return z_bytes(c, n); }
// pixels [0, n) of pix into columns [0, n) of outputs with row stride ld
int ctx_launch_range(fsg_ctx *c, const double *pix, int64_t n, int64_t ld, double *Z, int32_t *U,
                     int8_t *sr, int32_t *si, unsigned long long *work, hipStream_t stream, int max_blocks,
                     int64_t grid_w) {
    return launch(c, pix, n, Z, U, sr, si, work, stream, false, ld, max_blocks, grid_w);
}
size_t work_words() { return WORK_WORDS; }
}  // namespace fsg

namespace {

// ---- device staging of the host-buffer entry points ----
// bytes of staging for n pixels: c_pix, Z, U, stop_iter, stop_reason
size_t staging_bytes(const fsg_ctx *c, int64_t n) {
    const size_t pb = (size_t)n * 16, zb = z_bytes(c, n), ub = (size_t)n * 4, rb = (size_t)n;
    return align256(pb) + align256(zb) + 2 * align256(ub) + align256(rb);
}

// one tile through device staging at b (staging_bytes(c, n) bytes) with the
// cursor and counters at work; the caller holds c->mu
int tile_staged(fsg_ctx *c, const double *c_pix, int64_t n, int64_t grid_w, double *Z, int32_t *U,
                int8_t *stop_reason, int32_t *stop_iter, const volatile uint8_t *interrupted,
                fsg_stats *stats, char *b, unsigned long long *work) {
    const size_t pb = (size_t)n * 16, zb = z_bytes(c, n), ub = (size_t)n * 4, rb = (size_t)n;
    const size_t off_z = align256(pb), off_u = off_z + align256(zb), off_i = off_u + align256(ub);
    const size_t off_r = off_i + align256(ub);
    if (n > 0) FSG_HIP(hipMemcpyAsync(b, c_pix, pb, hipMemcpyHostToDevice, c->stream));
    const uint32_t *stop = nullptr;
    if (interrupted && n > 0) {
        // an interrupted tile leaves its pixels not taken at stop_reason -1
        *c->stop_h = 0u;
        stop = c->stop_d;
        hipLaunchKernelGGL(k_fill_i8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           reinterpret_cast<int8_t *>(b + off_r), n, (int8_t)-1);
        FSG_HIP(hipGetLastError());
    }
    int rc = launch(c, reinterpret_cast<double *>(b), n, reinterpret_cast<double *>(b + off_z),
                    reinterpret_cast<int32_t *>(b + off_u), reinterpret_cast<int8_t *>(b + off_r),
                    reinterpret_cast<int32_t *>(b + off_i), work, c->stream, true, -1, 0, grid_w, stop);
    if (rc) return rc;
    if (stop) {
        // wait for the kernel (its end event), passing the caller's interrupt
        // flag on to it (the reference polls it after every pixel,
        // perturbation.py:1042-1043); the copies to pageable host memory are
        // enqueued after, since enqueueing them can block this thread
        int spins = 0;
        hipError_t q;
        while ((q = hipEventQuery(c->ev1)) == hipErrorNotReady) {
            if (*interrupted) *c->stop_h = 1u;
            if (++spins < 200) std::this_thread::yield();
            else std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
        if (q != hipSuccess) return fail(FSG_E_HIP, std::string("tile: ") + hipGetErrorString(q));
    }
    if (n > 0) {
        FSG_HIP(hipMemcpyAsync(Z, b + off_z, zb, hipMemcpyDeviceToHost, c->stream));
        FSG_HIP(hipMemcpyAsync(U, b + off_u, ub, hipMemcpyDeviceToHost, c->stream));
        FSG_HIP(hipMemcpyAsync(stop_iter, b + off_i, ub, hipMemcpyDeviceToHost, c->stream));
        FSG_HIP(hipMemcpyAsync(stop_reason, b + off_r, rb, hipMemcpyDeviceToHost, c->stream));
    }
    FSG_HIP(hipStreamSynchronize(c->stream));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        unsigned long long w[WORK_WORDS];
        FSG_HIP(hipMemcpy(w, work, sizeof(w), hipMemcpyDeviceToHost));
        stats->plain_steps = w[1]; stats->bla_steps = w[2]; stats->bla_iters = w[3]; stats->rebases = w[4];
        stats->bla_lookups = w[5]; stats->redo_pixels = w[6];
        float ms = 0.f;
        if (n > 0 && hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) stats->kernel_ms = ms;
    }
    if (interrupted && *interrupted) return FSG_INTERRUPTED;
    return 0;
}

// ---- the staging of fsg_tile_grid_noalloc -------------------------------
// A file-scope device array: every device the library runs on holds its own
// copy for the life of the process, and no call allocates.  Its address on a
// device comes from hipGetSymbolAddress, taken once per device with that
// device current; host code uses that address, never the array's name.
constexpr size_t NOALLOC_WORK_BYTES = 4096;      // cursors and counters, then the tile staging
static_assert(NOALLOC_WORK_BYTES >= WORK_WORDS * sizeof(unsigned long long), "workspace does not fit");
static __device__ __align__(256) unsigned char scratch[FSG_NOALLOC_SCRATCH_BYTES];

struct ScratchDev {
    std::once_flag once;
    hipError_t err = hipSuccess;
    char *ptr = nullptr;
    std::mutex mu;                               // one tile at a time per device
};
constexpr int MAX_DEVICES = 64;
ScratchDev g_scratch[MAX_DEVICES];

int scratch_of(int dev, char **out) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(FSG_E_INVALID, "device id out of range");
    ScratchDev &s = g_scratch[dev];
    std::call_once(s.once, [&] {
        int cur = -1;
        s.err = hipGetDevice(&cur);
        if (s.err != hipSuccess) return;
        if (cur != dev && (s.err = hipSetDevice(dev)) != hipSuccess) return;
        void *p = nullptr;
        s.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (s.err == hipSuccess) s.ptr = static_cast<char *>(p);
        if (cur != dev) (void)hipSetDevice(cur);
    });
    if (s.err != hipSuccess || !s.ptr)
        return fail(FSG_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(s.err));
    *out = s.ptr;
    return 0;
}

}  // namespace

extern "C" {

int fsg_abi_version(void) { return FSG_ABI_VERSION; }

const char *fsg_last_error(void) { return g_err.c_str(); }

size_t fsg_work_bytes(void) { return WORK_WORDS * sizeof(unsigned long long); }

}  // extern "C"

namespace {

int ctx_create(const fsg_ref_desc *desc, fsg_ctx *base, fsg_ctx **out) {
    if (!out) return fail(FSG_E_INVALID, "null output pointer");
    *out = nullptr;
    int rc = validate(desc);
    if (rc) return rc;
    const uint64_t h_ref = ref_hash_of(desc);
    if (base) {
        if (base->ref_hash != h_ref || base->L != desc->ref_len)
            return fail(FSG_E_INVALID, "shared context: the orbit or its Xrange register differs from the base's");
        int dev = -1;
        if (hipGetDevice(&dev) != hipSuccess || dev != base->device)
            return fail(FSG_E_INVALID, "shared context: the base lives on another device");
    }
    fsg_ctx *c = new (std::nothrow) fsg_ctx();
    if (!c) return fail(FSG_E_NOMEM, "out of host memory");
    c->d = *desc;
    c->d.Zn_path = nullptr; c->d.xr_index = nullptr; c->d.xr_mant = nullptr; c->d.xr_exp = nullptr;
    if (c->d.kind == FSG_KIND_M2) c->d.degree = 2;
    if (c->d.kind == FSG_KIND_BS) c->d.interior_detect = 0;
    c->L = (int32_t)desc->ref_len;
    c->div_iter = (int32_t)desc->ref_div_iter;
    c->order = desc->ref_order > (int64_t)INT32_MAX ? INT32_MAX : (int32_t)desc->ref_order;
    c->first_invalid = std::min(c->L, std::min(c->div_iter, c->order));
    auto cleanup = [&](int code) { delete c; return code; };

    if (hipGetDevice(&c->device) != hipSuccess) return cleanup(fail(FSG_E_HIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return cleanup(fail(FSG_E_HIP, "stream/event creation failed"));
    {
        void *hp = nullptr, *dp = nullptr;
        if (hipHostMalloc(&hp, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
            return cleanup(fail(FSG_E_NOMEM, "hipHostMalloc interrupt flag"));
        c->stop_h = static_cast<uint32_t *>(hp);
        *c->stop_h = 0u;
        if (hipHostGetDevicePointer(&dp, hp, 0) != hipSuccess)
            return cleanup(fail(FSG_E_HIP, "hipHostGetDevicePointer interrupt flag"));
        c->stop_d = static_cast<const uint32_t *>(dp);
    }

    const int64_t L = c->L;
    const int32_t nx = desc->n_xr;
    const int holo = desc->kind != FSG_KIND_BS;
    c->ref_hash = h_ref;
    if (base) {
        c->shared = base->shared;
        c->Zn = base->Zn; c->xr_idx = base->xr_idx; c->xr_m = base->xr_m; c->xr_e = base->xr_e;
    } else {
        c->shared = std::make_shared<fsg_ctx::Shared>();
        if ((rc = upload(c, c->Zn, reinterpret_cast<const cplx *>(desc->Zn_path), (size_t)L, true))) return cleanup(rc);
        if ((rc = upload(c, c->xr_idx, desc->xr_index, (size_t)nx, true))) return cleanup(rc);
        if ((rc = upload(c, c->xr_m, desc->xr_mant, (size_t)2 * nx, true))) return cleanup(rc);
        if ((rc = upload(c, c->xr_e, desc->xr_exp, (size_t)(holo ? nx : 2 * nx), true))) return cleanup(rc);
    }
    // the normalisation codes: one load per orbit point, also where the
    // point comes from the register (no search of the sorted indices)
    c->zn_k_w = holo ? 1 : 2;
    if (base && base->zn_k && base->zn_k_w == c->zn_k_w) {
        c->zn_k = base->zn_k;
    } else {
        std::vector<int32_t> code;
        point_codes(desc, c->zn_k_w, code);
        if ((rc = upload(c, c->zn_k, code.data(), code.size(), true))) return cleanup(rc);
    }

    HostRef h{&c->d, c->L, c->div_iter, c->order};
    // the derivative recurrences read the host orbit and register
    fsg_ref_desc dh = c->d;
    dh.Zn_path = desc->Zn_path; dh.xr_index = desc->xr_index; dh.xr_mant = desc->xr_mant; dh.xr_exp = desc->xr_exp;
    h.d = &dh;
    if (holo) {
        std::vector<cplx> m; std::vector<int32_t> e;
        if (c->d.calc_dzndc) {
            holo_dzndc(h, m, e);
            if ((rc = upload(c, c->dc, m.data(), m.size()))) return cleanup(rc);
            if ((rc = upload(c, c->dc_e, e.data(), e.size()))) return cleanup(rc);
        }
        if (c->d.interior_detect) {
            holo_dzndz(h, m, e);
            if ((rc = upload(c, c->dz, m.data(), m.size()))) return cleanup(rc);
            if ((rc = upload(c, c->dz_e, e.data(), e.size()))) return cleanup(rc);
        }
    } else if (c->d.calc_dzndc) {
        std::vector<double> P[4]; std::vector<int32_t> E[4];
        bs_paths(h, P, E);
        for (int q = 0; q < 4; q++) {
            if ((rc = upload(c, c->bs[q], P[q].data(), P[q].size()))) return cleanup(rc);
            if ((rc = upload(c, c->bs_e[q], E[q].data(), E[q].size()))) return cleanup(rc);
        }
    }

    if (c->d.bla_on) {
        c->stages = stages_of(L);
        int64_t comp = L >> 3;
        c->bla_len = 2 * comp;
        int dim = holo ? 4 : 8;
        hipError_t err;
        // the coefficients depend on the orbit only: shared with the base
        // when it has them; the radii depend on kc and eps: per context
        const bool reuse_M = base && base->M && base->bla_len == c->bla_len;
        if (reuse_M) {
            c->M = base->M;
        } else {
            c->M = dalloc<double>((size_t)(c->bla_len * dim), err);
            if (err != hipSuccess) return cleanup(fail(FSG_E_NOMEM, "hipMalloc BLA"));
            (base ? c->owned : c->shared->p).push_back(c->M);
            if (hipMemsetAsync(c->M, 0, (size_t)(c->bla_len * dim) * sizeof(double), c->stream) != hipSuccess)
                return cleanup(fail(FSG_E_HIP, "hipMemsetAsync BLA"));
        }
        c->r = dalloc<double>((size_t)c->bla_len, err);
        if (err != hipSuccess) return cleanup(fail(FSG_E_NOMEM, "hipMalloc BLA"));
        c->owned.push_back(c->r);
        if (hipMemsetAsync(c->r, 0, (size_t)c->bla_len * sizeof(double), c->stream) != hipSuccess)
            return cleanup(fail(FSG_E_HIP, "hipMemsetAsync BLA"));
        double kc = to_std(mkf(desc->kc_mant, desc->kc_exp));
        hipError_t be = launch_bla_build(desc->kind, c->d.degree, desc->flavor, c->Zn, c->L, kc,
                                         desc->bla_eps, c->M, c->r, c->stages, c->stream, reuse_M);
        if (be != hipSuccess) return cleanup(fail(FSG_E_HIP, std::string("BLA build: ") + hipGetErrorString(be)));
        if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(fail(FSG_E_HIP, "BLA build sync"));
    }
    if (c->d.precision == FSG_PREC_FP32X && !f32_supported(cfg_of(c)))
        return cleanup(fail(FSG_E_UNSUPPORTED, "no FP32X kernel for this model"));
    // binary32 copies of the burning-ship orbit and Jacobian paths
    if (c->d.precision == FSG_PREC_FP32X && c->d.kind == FSG_KIND_BS) {
        hipError_t xe;
        void *p = dalloc<char>((size_t)L * 2 * XSR_BYTES, xe);
        c->owned.push_back(p); c->tx.ZRx = p;
        if (xe != hipSuccess) return cleanup(fail(FSG_E_NOMEM, "hipMalloc FP32X tables"));
        for (int q = 0; q < 4; q++) {
            if (!c->bs[q]) continue;
            p = dalloc<char>((size_t)L * XSR_BYTES, xe);
            c->owned.push_back(p); c->tx.BSx[q] = p;
            if (xe != hipSuccess) return cleanup(fail(FSG_E_NOMEM, "hipMalloc FP32X tables"));
        }
        KArgs ka;
        fill_args(c, ka);
        hipError_t be = build_xsr_tables(ka.t, c->tx, c->stream);
        if (be != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            return cleanup(fail(FSG_E_HIP, "FP32X table build failed"));
    }
    // binary32 copies of the holomorphic tables
    if (c->d.precision == FSG_PREC_FP32X && c->d.kind != FSG_KIND_BS) {
        hipError_t xe;
        void *p;
        const size_t L1 = (size_t)L + 1;
        p = dalloc<char>(L1 * XS_BYTES, xe); c->owned.push_back(p); c->tx.Zx = p;
        if (xe != hipSuccess) return cleanup(fail(FSG_E_NOMEM, "hipMalloc FP32X tables"));
        p = dalloc<char>(L1 * XS_BYTES, xe); c->owned.push_back(p); c->tx.DCx = c->dc ? p : nullptr;
        if (xe != hipSuccess) return cleanup(fail(FSG_E_NOMEM, "hipMalloc FP32X tables"));
        p = dalloc<char>((L1 + 1) * XS_BYTES, xe); c->owned.push_back(p); c->tx.DZx = c->dz ? p : nullptr;
        if (xe != hipSuccess) return cleanup(fail(FSG_E_NOMEM, "hipMalloc FP32X tables"));
        p = dalloc<char>((size_t)(c->bla_len + 1) * XS_BYTES, xe); c->owned.push_back(p); c->tx.Ax = p;
        if (xe != hipSuccess) return cleanup(fail(FSG_E_NOMEM, "hipMalloc FP32X tables"));
        p = dalloc<char>((size_t)(c->bla_len + 1) * XS_BYTES, xe); c->owned.push_back(p); c->tx.Bx = p;
        if (xe != hipSuccess) return cleanup(fail(FSG_E_NOMEM, "hipMalloc FP32X tables"));
        KArgs ka;
        fill_args(c, ka);
        hipError_t be = build_xs_tables(ka.t, c->bla_len, c->tx, c->stream);
        if (be != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            return cleanup(fail(FSG_E_HIP, "FP32X table build failed"));
    }
    hipError_t err;
    c->work = dalloc<unsigned long long>(WORK_WORDS, err);
    if (err != hipSuccess) return cleanup(fail(FSG_E_NOMEM, "hipMalloc work"));
    c->owned.push_back(c->work);
    *out = c;
    return 0;
}

}  // namespace

extern "C" {

int fsg_ctx_create(const fsg_ref_desc *desc, fsg_ctx **out) { return ctx_create(desc, nullptr, out); }

int fsg_ctx_create_shared(const fsg_ref_desc *desc, fsg_ctx *base, fsg_ctx **out) {
    if (!base) return fail(FSG_E_INVALID, "null base context");
    return ctx_create(desc, base, out);
}

int fsg_ctx_destroy(fsg_ctx *ctx) {
    if (!ctx) return 0;
    { std::lock_guard<std::mutex> lk(ctx->mu); (void)hipStreamSynchronize(ctx->stream); }
    delete ctx;
    return 0;
}

int fsg_ctx_bla_shape(fsg_ctx *c, int64_t *bla_len, int32_t *stages) {
    if (!c) return fail(FSG_E_INVALID, "null context");
    if (bla_len) *bla_len = c->bla_len;
    if (stages) *stages = c->stages;
    return 0;
}

int fsg_ctx_read_table(fsg_ctx *c, int which, void *out, size_t bytes) {
    if (!c || !out) return fail(FSG_E_INVALID, "null argument");
    const void *src = nullptr;
    size_t n = 0;
    const size_t L = (size_t)c->L;
    const int dim = c->d.kind == FSG_KIND_BS ? 8 : 4;
    switch (which) {
        case 0: src = c->M; n = (size_t)c->bla_len * dim * 8; break;
        case 1: src = c->r; n = (size_t)c->bla_len * 8; break;
        case 2: src = c->dc; n = L * 16; break;
        case 3: src = c->dc_e; n = L * 4; break;
        case 4: src = c->dz; n = (L + 1) * 16; break;
        case 5: src = c->dz_e; n = (L + 1) * 4; break;
        case 6: case 7: case 8: case 9: src = c->bs[which - 6]; n = L * 8; break;
        case 10: case 11: case 12: case 13: src = c->bs_e[which - 10]; n = L * 4; break;
        default: return fail(FSG_E_INVALID, "unknown table");
    }
    if (!src) return fail(FSG_E_INVALID, "table not computed for this context");
    if (bytes < n) return fail(FSG_E_INVALID, "output buffer too small");
    std::lock_guard<std::mutex> lk(c->mu);
    FSG_HIP(hipMemcpy(out, src, n, hipMemcpyDeviceToHost));
    return 0;
}

int fsg_tile(fsg_ctx *c, const double *c_pix, int64_t n, double *Z, int32_t *U, int8_t *stop_reason,
             int32_t *stop_iter, const volatile uint8_t *interrupted, fsg_stats *stats) {
    return fsg_tile_grid(c, c_pix, n, 0, Z, U, stop_reason, stop_iter, interrupted, stats);
}

int fsg_tile_grid(fsg_ctx *c, const double *c_pix, int64_t n, int64_t grid_w, double *Z, int32_t *U,
                  int8_t *stop_reason, int32_t *stop_iter, const volatile uint8_t *interrupted,
                  fsg_stats *stats) {
    if (!c) return fail(FSG_E_INVALID, "null context");
    if (n < 0 || (n > 0 && (!c_pix || !Z || !U || !stop_reason || !stop_iter)))
        return fail(FSG_E_INVALID, "bad tile buffers");
    if (grid_w < 0) return fail(FSG_E_INVALID, "negative grid width");
    if (interrupted && *interrupted) return FSG_INTERRUPTED;
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t total = staging_bytes(c, n);
    if (total > c->buf_bytes) {
        if (c->buf) { (void)hipFree(c->buf); c->buf = nullptr; c->buf_bytes = 0; }
        if (hipMalloc(&c->buf, total) != hipSuccess) { c->buf = nullptr; return fail(FSG_E_NOMEM, "hipMalloc tile buffers"); }
        c->buf_bytes = total;
    }
    return tile_staged(c, c_pix, n, grid_w, Z, U, stop_reason, stop_iter, interrupted, stats, c->buf, c->work);
}

int fsg_tile_grid_noalloc(fsg_ctx *c, const double *c_pix, int64_t n, int64_t grid_w, double *Z,
                          int32_t *U, int8_t *stop_reason, int32_t *stop_iter,
                          const volatile uint8_t *interrupted, fsg_stats *stats) {
    if (!c) return fail(FSG_E_INVALID, "null context");
    if (n < 0 || (n > 0 && (!c_pix || !Z || !U || !stop_reason || !stop_iter)))
        return fail(FSG_E_INVALID, "bad tile buffers");
    if (grid_w < 0) return fail(FSG_E_INVALID, "negative grid width");
    const size_t need = NOALLOC_WORK_BYTES + staging_bytes(c, n);
    if (need > FSG_NOALLOC_SCRATCH_BYTES)
        return fail(FSG_E_RANGE, "fsg_tile_grid_noalloc: a tile of " + std::to_string(n) + " pixels needs " +
                                 std::to_string(need) + " bytes of device staging, above the " +
                                 std::to_string((size_t)FSG_NOALLOC_SCRATCH_BYTES) +
                                 " bytes this entry point holds per device; use fsg_tile_grid or smaller tiles");
    if (interrupted && *interrupted) return FSG_INTERRUPTED;
    char *base = nullptr;
    int rc = scratch_of(c->device, &base);
    if (rc) return rc;
    // one staging area per device: calls on contexts of the same device
    // take turns (the call is synchronous)
    std::lock_guard<std::mutex> lk_dev(g_scratch[c->device].mu);
    std::lock_guard<std::mutex> lk(c->mu);
    return tile_staged(c, c_pix, n, grid_w, Z, U, stop_reason, stop_iter, interrupted, stats,
                       base + NOALLOC_WORK_BYTES, reinterpret_cast<unsigned long long *>(base));
}

int fsg_tile_device(fsg_ctx *c, const double *d_c_pix, int64_t n, double *d_Z, int32_t *d_U,
                    int8_t *d_stop_reason, int32_t *d_stop_iter, void *d_work, void *stream) {
    return fsg_tile_grid_device(c, d_c_pix, n, 0, d_Z, d_U, d_stop_reason, d_stop_iter, d_work, stream);
}

int fsg_tile_grid_device(fsg_ctx *c, const double *d_c_pix, int64_t n, int64_t grid_w, double *d_Z,
                         int32_t *d_U, int8_t *d_stop_reason, int32_t *d_stop_iter, void *d_work,
                         void *stream) {
    if (!c) return fail(FSG_E_INVALID, "null context");
    if (!d_work) return fail(FSG_E_INVALID, "null work buffer");
    if (n < 0 || (n > 0 && (!d_c_pix || !d_Z || !d_U || !d_stop_reason || !d_stop_iter)))
        return fail(FSG_E_INVALID, "bad tile buffers");
    if (grid_w < 0) return fail(FSG_E_INVALID, "negative grid width");
    return launch(c, d_c_pix, n, d_Z, d_U, d_stop_reason, d_stop_iter,
                  static_cast<unsigned long long *>(d_work), static_cast<hipStream_t>(stream), false,
                  -1, 0, grid_w);
}

int fsg_work_stats(const void *d_work, fsg_stats *stats) {
    if (!d_work || !stats) return fail(FSG_E_INVALID, "null argument");
    unsigned long long w[WORK_WORDS];
    FSG_HIP(hipMemcpy(w, d_work, sizeof(w), hipMemcpyDeviceToHost));
    stats->plain_steps = w[1]; stats->bla_steps = w[2]; stats->bla_iters = w[3]; stats->rebases = w[4];
    stats->bla_lookups = w[5]; stats->redo_pixels = w[6];
    return 0;
}

}  // extern "C"
