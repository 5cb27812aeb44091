// fsg_post.hip -- the potential and distance-estimate layers of a tile,
// computed on the GPU from the raw results while they are still in HBM.
//
// Restates, per pixel:
//   Continuous_iter_pp with kind "infinity" (src/fractalshades/postproc.py:
//   352-406, formula 1001-1009):
//     k = |a_d|^(1/(d-1)),  nu_frac = -log(log|z_n k| / log(M k)) / log d,
//     (div, mod) = divmod(-nu_frac, 1),  nu = (n - div - floor_iter) - mod
//   DEM_pp (postproc.py:684-731): |z_n| ln|z_n| / |dz_n/dc|, where for the
//   burning ship |dz_n/dc| is the larger singular value of the Jacobian,
//   (hypot(a + d, b - c) + hypot(a - d, b + c)) / 2, and values below
//   px_snap are snapped to 0.
//   Attr_normal_pp with kind "potential" / "infinity" (postproc.py:572-628):
//     holomorphic z_n / (dz_n/dc); burning ship J^T z_n with J the Jacobian
//     d(x, y)/d(a, b); then the transposed skew (apply_unskew_1d,
//     core.py:3147-3158) and division by the modulus.  The direction of
//     z_n / (dz_n/dc) is that of z_n conj(dz_n/dc), which is formed instead
//     (after a power-of-two scaling of the derivative, so that no product
//     overflows): the unit vector is the same.
// Computed in binary64, stored as binary32 (the reference's default
// post-processing dtype, settings.py:82).  Pixels that did not diverge get
// NaN in every layer: the reference's formulas have no meaning there and its
// layers mask them by stop reason.
#include "fsgpu.h"
#include "fsg_internal.hpp"

#include <cmath>
#include <string>

namespace fsg {
int set_error(int code, const std::string &msg);
}

using namespace fsg;

namespace {

struct PostArgs {
    const double *Z;
    int64_t n, ld;
    const int32_t *si;
    const int8_t *sr;
    float *nu, *dem, *nrm;
    int32_t holo, row_dc;
    double k, log_mk, log_d, floor_iter, px_snap;
    double s00, s01, s10, s11;
};

__global__ void __launch_bounds__(256) k_postproc(const PostArgs p) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += stride) {
        if (p.sr[i] != 1) {
            if (p.nu) p.nu[i] = __builtin_nanf("");
            if (p.dem) p.dem[i] = __builtin_nanf("");
            if (p.nrm) { p.nrm[2 * i] = __builtin_nanf(""); p.nrm[2 * i + 1] = __builtin_nanf(""); }
            continue;
        }
        double zr, zi;
        if (p.holo) { zr = p.Z[2 * i]; zi = p.Z[2 * i + 1]; }
        else { zr = p.Z[i]; zi = p.Z[p.ld + i]; }
        if (p.nu) {
            const double azk = hypot(zr * p.k, zi * p.k);
            const double nu_frac = -(log(log(azk) / p.log_mk) / p.log_d);
            // numpy divmod(x, 1.): floor and the non-negative remainder
            const double x = -nu_frac;
            const double dv = floor(x);
            const double md = x - dv;
            const double n = (double)p.si[i] - dv;
            p.nu[i] = (float)((n - p.floor_iter) - md);
        }
        if (p.dem) {
            const double az = hypot(zr, zi);
            double ad;
            if (p.holo) {
                const int64_t r = (int64_t)p.row_dc * p.ld + i;
                ad = hypot(p.Z[2 * r], p.Z[2 * r + 1]);
            } else {
                const double a = p.Z[(int64_t)p.row_dc * p.ld + i], b = p.Z[(int64_t)(p.row_dc + 1) * p.ld + i];
                const double c = p.Z[(int64_t)(p.row_dc + 2) * p.ld + i], d = p.Z[(int64_t)(p.row_dc + 3) * p.ld + i];
                ad = 0.5 * (hypot(a + d, b - c) + hypot(a - d, b + c));
            }
            double v = az * log(az) / ad;
            if (p.px_snap > 0. && v < p.px_snap) v = 0.;
            p.dem[i] = (float)v;
        }
        if (p.nrm) {
            double nx, ny;
            if (p.holo) {
                const int64_t r = (int64_t)p.row_dc * p.ld + i;
                double dr = p.Z[2 * r], di = p.Z[2 * r + 1];
                const int sh = -__builtin_amdgcn_frexp_exp(fmax(fabs(dr), fabs(di)));
                dr = ldexp(dr, sh); di = ldexp(di, sh);
                nx = zr * dr + zi * di;          // z conj(dz/dc)
                ny = zi * dr - zr * di;
            } else {
                const double *J = p.Z + (int64_t)p.row_dc * p.ld + i;
                double a = J[0], b = J[p.ld], c = J[2 * p.ld], d = J[3 * p.ld];
                const int sh = -__builtin_amdgcn_frexp_exp(fmax(fmax(fabs(a), fabs(b)), fmax(fabs(c), fabs(d))));
                a = ldexp(a, sh); b = ldexp(b, sh); c = ldexp(c, sh); d = ldexp(d, sh);
                nx = a * zr + c * zi;            // J^T z
                ny = b * zr + d * zi;
            }
            const double ux = p.s00 * nx + p.s10 * ny, uy = p.s01 * nx + p.s11 * ny;
            const double m = hypot(ux, uy);
            p.nrm[2 * i] = (float)(ux / m);
            p.nrm[2 * i + 1] = (float)(uy / m);
        }
    }
}

int fill_post(const fsg_post_desc *pp, PostArgs &a) {
    if (!pp) return set_error(FSG_E_INVALID, "null post-processing descriptor");
    if (!(pp->d > 1.) || !(pp->M > 0.)) return set_error(FSG_E_INVALID, "potential needs d > 1 and M > 0");
    a.holo = pp->holomorphic ? 1 : 0;
    a.row_dc = pp->dzndc_row;
    a.k = std::pow(std::fabs(pp->a_d), 1. / (pp->d - 1.));
    a.log_mk = std::log(pp->M * a.k);
    a.log_d = std::log(pp->d);
    a.floor_iter = pp->floor_iter;
    a.px_snap = pp->px_snap;
    const bool skew = pp->skew[0] != 0. || pp->skew[1] != 0. || pp->skew[2] != 0. || pp->skew[3] != 0.;
    a.s00 = skew ? pp->skew[0] : 1.; a.s01 = skew ? pp->skew[1] : 0.;
    a.s10 = skew ? pp->skew[2] : 0.; a.s11 = skew ? pp->skew[3] : 1.;
    return 0;
}

}  // namespace

namespace fsg {

// enqueue the layers of n pixels whose raw results sit in Z (row stride ld)
int post_launch(const fsg_post_desc *pp, const double *Z, int64_t n, int64_t ld, const int32_t *si,
                const int8_t *sr, float *nu, float *dem, float *nrm, hipStream_t stream) {
    PostArgs a;
    int rc = fill_post(pp, a);
    if (rc) return rc;
    if ((dem || nrm) && pp->dzndc_row < 1)
        return set_error(FSG_E_INVALID, "DEM and normal need the dz/dc (Jacobian) rows");
    a.Z = Z; a.n = n; a.ld = ld; a.si = si; a.sr = sr; a.nu = nu; a.dem = dem; a.nrm = nrm;
    if (n <= 0) return 0;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_postproc, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    if (hipGetLastError() != hipSuccess) return set_error(FSG_E_HIP, "post-processing launch failed");
    return 0;
}

}  // namespace fsg

extern "C" {

int fsg_postproc_device(const fsg_post_desc *pp, const double *d_Z, int64_t n, int64_t ld,
                        const int32_t *d_stop_iter, const int8_t *d_stop_reason, float *d_nu, float *d_dem,
                        float *d_normal, void *stream) {
    if (n < 0 || ld < n || (n > 0 && (!d_Z || !d_stop_iter || !d_stop_reason)))
        return set_error(FSG_E_INVALID, "bad post-processing buffers");
    return post_launch(pp, d_Z, n, ld, d_stop_iter, d_stop_reason, d_nu, d_dem, d_normal,
                       static_cast<hipStream_t>(stream));
}

}  // extern "C"
