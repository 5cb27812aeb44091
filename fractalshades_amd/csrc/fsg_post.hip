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
//   Before the DEM and the normal, the derivatives take the loop's
//   dzndc_modifier (perturbation.py:1387-1388, which the tile kernels leave
//   to the host) and the projection's map at the pixel (get_dzndc,
//   postproc.py:184-206): dz/dc times df (apply_df, 972-978), the Jacobian
//   times the dfBS matrix (apply_dfBS, 980-998), the maps of the exponential
//   projection (projection.py:375-453).
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
    const double *pix;
    int64_t n, ld;
    const int32_t *si;
    const int8_t *sr;
    float *nu, *dem, *nrm;
    int32_t holo, row_dc;
    double k, log_mk, log_d, floor_iter, px_snap;
    double s00, s01, s10, s11;
    int32_t proj, mod;
    double w_scale, mod_scale;
};

// The projection's map at pixel i, from its projected offset w = exp(h + t)
// (h real: the map's h_moy, or h_moy - h_band by bands): exp(t) = w e^-h,
// exp(i Im t) = w / |w|.  Holomorphic factor (fr, fi); burning-ship matrix
// (m00, m01, m10, m11) as projection.py:412-453 writes them.
__device__ void proj_map(const PostArgs &p, int64_t i, double &fr, double &fi, double m[4]) {
    const double wx = p.pix[2 * i], wy = p.pix[2 * i + 1];
    if (p.proj == 1) {                 // exp(t); (-Im e^t, -Re e^t, Re e^t, -Im e^t)
        const double cr = wx * p.w_scale, sr = wy * p.w_scale;
        fr = cr; fi = sr;
        m[0] = -sr; m[1] = -cr; m[2] = cr; m[3] = -sr;
    } else if (p.proj == 2) {          // exp(Re t)
        const double r = hypot(wx, wy) * p.w_scale;
        fr = r; fi = 0.;
        m[0] = r; m[1] = 0.; m[2] = 0.; m[3] = r;
    } else if (p.proj == 3) {          // exp(i Im t)
        const double a = hypot(wx, wy), c = wx / a, s = wy / a;
        fr = c; fi = s;
        m[0] = c; m[1] = -s; m[2] = s; m[3] = c;
    } else {
        fr = 1.; fi = 0.;
        m[0] = 1.; m[1] = 0.; m[2] = 0.; m[3] = 1.;
    }
}

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
        if (!p.dem && !p.nrm) continue;
        // the derivatives: dz/dc, or the Jacobian (a b; c d) = (dX/dA dX/dB; dY/dA dY/dB)
        double dr = 0., di = 0., a = 0., b = 0., c = 0., d = 0.;
        if (p.holo) {
            const int64_t r = (int64_t)p.row_dc * p.ld + i;
            dr = p.Z[2 * r]; di = p.Z[2 * r + 1];
        } else {
            const double *J = p.Z + (int64_t)p.row_dc * p.ld + i;
            a = J[0]; b = J[p.ld]; c = J[2 * p.ld]; d = J[3 * p.ld];
        }
        if (p.mod != 0) {
            // the loop's dzndc_modifier at the pixel
            const double wx = p.pix[2 * i], wy = p.pix[2 * i + 1];
            const double f = (p.mod == 1 ? hypot(wx, wy) : hypot(wx + 1.e-6, wy)) * p.mod_scale;
            dr *= f; di *= f; a *= f; b *= f; c *= f; d *= f;
        }
        if (p.proj != 0) {
            double fr, fi, m[4];
            proj_map(p, i, fr, fi, m);
            if (p.holo) {
                const double xr = fr * dr - fi * di, xi = fr * di + fi * dr;
                dr = xr; di = xi;
            } else {
                const double a2 = a * m[0] + b * m[2], b2 = a * m[1] + b * m[3];
                const double c2 = c * m[0] + d * m[2], d2 = c * m[1] + d * m[3];
                a = a2; b = b2; c = c2; d = d2;
            }
        }
        if (p.dem) {
            const double az = hypot(zr, zi);
            const double ad = p.holo ? hypot(dr, di) : 0.5 * (hypot(a + d, b - c) + hypot(a - d, b + c));
            double v = az * log(az) / ad;
            if (p.px_snap > 0. && v < p.px_snap) v = 0.;
            p.dem[i] = (float)v;
        }
        if (p.nrm) {
            double nx, ny;
            if (p.holo) {
                const int sh = -__builtin_amdgcn_frexp_exp(fmax(fabs(dr), fabs(di)));
                const double er = ldexp(dr, sh), ei = ldexp(di, sh);
                nx = zr * er + zi * ei;          // z conj(dz/dc)
                ny = zi * er - zr * ei;
            } else {
                const int sh = -__builtin_amdgcn_frexp_exp(fmax(fmax(fabs(a), fabs(b)), fmax(fabs(c), fabs(d))));
                const double ea = ldexp(a, sh), eb = ldexp(b, sh), ec = ldexp(c, sh), ed = ldexp(d, sh);
                nx = ea * zr + ec * zi;          // J^T z
                ny = eb * zr + ed * zi;
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
    if (pp->proj < 0 || pp->proj > 4) return set_error(FSG_E_INVALID, "unknown projection map");
    a.proj = pp->proj == 4 ? 0 : pp->proj;      // the map "one" changes nothing
    a.w_scale = pp->w_scale;
    if (pp->mod < 0 || pp->mod > 2) return set_error(FSG_E_INVALID, "unknown derivative modifier");
    a.mod = pp->mod; a.mod_scale = pp->mod_scale;
    return 0;
}

}  // namespace

namespace fsg {

// enqueue the layers of n pixels whose raw results sit in Z (row stride ld)
int post_launch(const fsg_post_desc *pp, const double *Z, int64_t n, int64_t ld, const double *pix,
                const int32_t *si, const int8_t *sr, float *nu, float *dem, float *nrm, hipStream_t stream) {
    PostArgs a;
    int rc = fill_post(pp, a);
    if (rc) return rc;
    if ((dem || nrm) && pp->dzndc_row < 1)
        return set_error(FSG_E_INVALID, "DEM and normal need the dz/dc (Jacobian) rows");
    if ((a.proj != 0 || a.mod != 0) && (dem || nrm) && n > 0 && !pix)
        return set_error(FSG_E_INVALID, "the projection's map needs the pixels");
    a.Z = Z; a.pix = pix; a.n = n; a.ld = ld; a.si = si; a.sr = sr; a.nu = nu; a.dem = dem; a.nrm = nrm;
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
                        const double *d_c_pix, const int32_t *d_stop_iter, const int8_t *d_stop_reason,
                        float *d_nu, float *d_dem, float *d_normal, void *stream) {
    if (n < 0 || ld < n || (n > 0 && (!d_Z || !d_stop_iter || !d_stop_reason)))
        return set_error(FSG_E_INVALID, "bad post-processing buffers");
    return post_launch(pp, d_Z, n, ld, d_c_pix, d_stop_iter, d_stop_reason, d_nu, d_dem, d_normal,
                       static_cast<hipStream_t>(stream));
}

}  // extern "C"
