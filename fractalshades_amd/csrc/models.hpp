// models.hpp -- per-iteration formulas of the three model families, generic
// over binary64 (cplx / double) and Xrange (xc / xf) operands.  Used by the
// GPU kernels and by the library's host-side orbit-derivative recurrences.
//   M2: models/mandelbrot_M2.py:599-622       Mn: mandelbrot_Mn.py:628-729
//   burning ship: burning_ship.py:12-60 (diffabs), 440-532 (Jacobian),
//                 535-619 (delta iteration), 622-859 (Jacobian perturbation)
#pragma once
#include "fsg_internal.hpp"

namespace fsg {

// --------------------------------------------------------------------------
// model formulas, generic over standard (cplx) and Xrange (xc) operands

FSG_HD cplx muls(double s, cplx a) { return scale(s, a); }
FSG_HD xc muls(double s, xc a) { return s * a; }

constexpr double cbinom(int n, int k) {
    double r = 1.;
    for (int i = 1; i <= k; i++) r = r * (double)(n - k + i) / (double)i;
    return r;
}

// z_{n+1} - Z_{n+1} for z^2 + c (M2) and z^d + c (Mn, binomial expansion)
template <int KIND, int DEG, class T>
FSG_HD T p_zn(T z, T Z, T c) {
    if constexpr (KIND == KIND_M2) {
        return z * (z + muls(2., Z)) + c;
    } else {
        T tmp = z * (z + muls(cbinom(DEG, 1), Z));
        T Zpk = Z;
#pragma unroll
        for (int q = 2; q < DEG; q++) {
            Zpk = Zpk * Z;
            tmp = z * (tmp + muls(cbinom(DEG, q), Zpk));
        }
        return tmp + c;
    }
}

// perturbation of a derivative (dz/dc or dz/dz): d is the pixel's delta, Zd
// the reference orbit's value of the same derivative
template <int KIND, int DEG, class T>
FSG_HD T p_deriv(T z, T d, T Z, T Zd) {
    if constexpr (KIND == KIND_M2) {
        return muls(2., (Z + z) * d + Zd * z);
    } else {
        T mul = z + muls(cbinom(DEG, 1), Z);
        T tmp = z * mul;
        T dt = d * mul + z * (d + muls(cbinom(DEG, 1), Zd));
        T Zpk = Z;
#pragma unroll
        for (int q = 2; q < DEG; q++) {
            T dZpk = muls((double)q, Zpk) * Zd;
            Zpk = Zpk * Z;
            mul = tmp + muls(cbinom(DEG, q), Zpk);
            dt = d * mul + z * (dt + muls(cbinom(DEG, q), dZpk));
            tmp = z * mul;
        }
        return dt;
    }
}

// z^d + c and the perturbation of ND derivatives (dz/dc and/or dz/dz) in
// one pass: p_zn's partial products are the factors "tmp" and "mul" of
// p_deriv, and the powers of Z are the same, so they are formed once.  Every
// value is the same expression, in the same order, as in p_zn and p_deriv.
// dv[i]: the pixel's derivative delta, Zdv[i]: the reference orbit's value
// of the same derivative, out[i]: the next value of dv[i].  The last sum of
// each result is taken by add(a, b), which may return another type R (a
// normalised value).  Returns the next z.
template <int DEG, int ND, class T, class R, class Add>
FSG_HD R p_zn_derivs(T z, T Z, T c, const T *dv, const T *Zdv, R *out, Add add) {
    T mul = z + muls(cbinom(DEG, 1), Z);
    T tmp = z * mul;
    T dt[ND > 0 ? ND : 1];
    if constexpr (DEG == 2) {
#pragma unroll
        for (int i = 0; i < ND; i++) out[i] = add(dv[i] * mul, z * (dv[i] + muls(cbinom(DEG, 1), Zdv[i])));
    } else {
#pragma unroll
        for (int i = 0; i < ND; i++) dt[i] = dv[i] * mul + z * (dv[i] + muls(cbinom(DEG, 1), Zdv[i]));
    }
    T Zpk = Z;
#pragma unroll
    for (int q = 2; q < DEG; q++) {
        T dZpk[ND > 0 ? ND : 1];
#pragma unroll
        for (int i = 0; i < ND; i++) dZpk[i] = muls((double)q, Zpk) * Zdv[i];
        Zpk = Zpk * Z;
        mul = tmp + muls(cbinom(DEG, q), Zpk);
        if (q + 1 < DEG) {
#pragma unroll
            for (int i = 0; i < ND; i++) dt[i] = dv[i] * mul + z * (dt[i] + muls(cbinom(DEG, q), dZpk[i]));
        } else {
#pragma unroll
            for (int i = 0; i < ND; i++) out[i] = add(dv[i] * mul, z * (dt[i] + muls(cbinom(DEG, q), dZpk[i])));
        }
        tmp = z * mul;
    }
    return add(tmp, c);
}

template <int KIND, int DEG, class T>
FSG_HD T dfdz(T z) {
    if constexpr (KIND == KIND_M2) {
        return muls(2., z);
    } else {
        T tmp = z;
#pragma unroll
        for (int q = 2; q < DEG; q++) tmp = tmp * z;
        return muls((double)DEG, tmp);
    }
}

// Burning-ship flavours: Jacobian of f at (x, y), burning_ship.py:440-532
FSG_HD double sgn(double x) { return x < 0. ? -1. : 1.; }
template <int FL>
FSG_HD void bs_jac(double x, double y, double J[4]) {
    double q = x * x - y * y;
    if constexpr (FL == 1) { J[0] = 2. * x; J[1] = -2. * y; J[2] = 2. * sgn(x) * fabs(y); J[3] = 2. * sgn(y) * fabs(x); }
    else if constexpr (FL == 2) { J[0] = 2. * x; J[1] = -2. * y; J[2] = 2. * fabs(y); J[3] = 2. * sgn(y) * x; }
    else if constexpr (FL == 3) { J[0] = 2. * x; J[1] = -2. * fabs(y); J[2] = 2. * y; J[3] = 2. * x; }
    else if constexpr (FL == 4) { J[0] = 2. * sgn(q) * x; J[1] = -2. * sgn(q) * y; J[2] = 2. * y; J[3] = 2. * x; }
    else { J[0] = 2. * sgn(q) * x; J[1] = -2. * sgn(q) * y; J[2] = 2. * sgn(x) * fabs(y); J[3] = 2. * sgn(y) * fabs(x); }
}
template <int FL>
FSG_HD void bs_jac(xf x, xf y, xf J[4]) {
    xf q = x * x - y * y;
    double sx = lt(x, 0.) ? -1. : 1., sy = lt(y, 0.) ? -1. : 1., sq = lt(q, 0.) ? -1. : 1.;
    if constexpr (FL == 1) { J[0] = 2. * x; J[1] = -2. * y; J[2] = (2. * sx) * xabs(y); J[3] = (2. * sy) * xabs(x); }
    else if constexpr (FL == 2) { J[0] = 2. * x; J[1] = -2. * y; J[2] = 2. * xabs(y); J[3] = (2. * sy) * x; }
    else if constexpr (FL == 3) { J[0] = 2. * x; J[1] = -2. * xabs(y); J[2] = 2. * y; J[3] = 2. * x; }
    else if constexpr (FL == 4) { J[0] = (2. * sq) * x; J[1] = (-2. * sq) * y; J[2] = 2. * y; J[3] = 2. * x; }
    else { J[0] = (2. * sq) * x; J[1] = (-2. * sq) * y; J[2] = (2. * sx) * xabs(y); J[3] = (2. * sy) * xabs(x); }
}

// |X + x| - |X| and its partial derivatives, burning_ship.py:19-60
FSG_HD double diffabs(double X, double x) {
    if (X >= 0.) return (X + x >= 0.) ? x : -(2. * X + x);
    return (X + x <= 0.) ? -x : 2. * X + x;
}
FSG_HD double ddX(double X, double x) {
    if (X >= 0.) return (X + x >= 0.) ? 0. : -2.;
    return (X + x <= 0.) ? 0. : 2.;
}
FSG_HD double ddx(double X, double x) {
    if (X >= 0.) return (X + x >= 0.) ? 1. : -1.;
    return (X + x <= 0.) ? -1. : 1.;
}
FSG_HD xf diffabs(xf X, xf x) {
    if (ge(X, 0.)) { if (ge(X + x, 0.)) return 1. * x; return -(2. * X + x); }
    if (le(X + x, 0.)) return -x;
    return 2. * X + x;
}
FSG_HD double ddX(xf X, xf x) {
    if (ge(X, 0.)) return ge(X + x, 0.) ? 0. : -2.;
    return le(X + x, 0.) ? 0. : 2.;
}
FSG_HD double ddx(xf X, xf x) {
    if (ge(X, 0.)) return ge(X + x, 0.) ? 1. : -1.;
    return le(X + x, 0.) ? -1. : 1.;
}
FSG_HD double sgn(xf x) { return lt(x, 0.) ? -1. : 1.; }
FSG_HD double absv(double x) { return fabs(x); }
FSG_HD xf absv(xf x) { return xabs(x); }
FSG_HD double mul_s(double s, double x) { return s * x; }
FSG_HD xf mul_s(double s, xf x) { return s * x; }

// delta iteration of (x, y), burning_ship.py:535-619
template <int FL, class T>
FSG_HD void bs_zn(T &x, T &y, T X, T Y, T a, T b) {
    T nx, ny;
    if constexpr (FL == 1) {
        T XY = X * Y;
        nx = x * (x + mul_s(2., X)) - y * (y + mul_s(2., Y)) + a;
        ny = mul_s(2., diffabs(XY, x * y + x * Y + y * X)) - b;
    } else if constexpr (FL == 2) {
        nx = x * (x + mul_s(2., X)) - y * (y + mul_s(2., Y)) + a;
        ny = mul_s(2., X * diffabs(Y, y) + x * absv(Y + y)) - b;
    } else if constexpr (FL == 3) {
        nx = x * (x + mul_s(2., X)) - Y * diffabs(Y, y) - y * absv(Y + y) + a;
        ny = mul_s(2., X * y + Y * x + x * y) - b;
    } else if constexpr (FL == 4) {
        T q = X * X - Y * Y;
        nx = diffabs(q, x * (x + mul_s(2., X)) - y * (y + mul_s(2., Y))) + a;
        ny = mul_s(2., X * y + Y * x + x * y) - b;
    } else {
        T XY = X * Y;
        T q = X * X - Y * Y;
        nx = diffabs(q, x * (x + mul_s(2., X)) - y * (y + mul_s(2., Y))) + a;
        ny = mul_s(2., diffabs(XY, x * y + x * Y + y * X)) - b;
    }
    x = nx; y = ny;
}

// Jacobian perturbation, burning_ship.py:622-859.  P = (xa, xb, ya, yb) are
// the pixel's deltas of dx/da, dx/db, dy/da, dy/db; R = the reference's.
template <int FL, class T>
FSG_HD void bs_hess(T x, T y, T P[4], T X, T Y, const T R[4]) {
    const T xa = P[0], xb = P[1], ya = P[2], yb = P[3];
    const T Xa = R[0], Xb = R[1], Ya = R[2], Yb = R[3];
    T nxa, nxb, nya, nyb;
    if constexpr (FL == 1 || FL == 5) {
        T opX = X * Y;
        T dXa = Xa * Y + X * Ya, dXb = Xb * Y + X * Yb;
        T opx = x * y + x * Y + y * X;
        T dxa_ = xa * y + x * ya + xa * Y + x * Ya + ya * X + y * Xa;
        T dxb_ = xb * y + x * yb + xb * Y + x * Yb + yb * X + y * Xb;
        double gX = ddX(opX, opx), gx = ddx(opX, opx);
        nya = mul_s(2., mul_s(gX, dXa) + mul_s(gx, dxa_));
        nyb = mul_s(2., mul_s(gX, dXb) + mul_s(gx, dxb_));
        if constexpr (FL == 1) {
            nxa = mul_s(2., (X + x) * xa + Xa * x) - mul_s(2., (Y + y) * ya + Ya * y);
            nxb = mul_s(2., (X + x) * xb + Xb * x) - mul_s(2., (Y + y) * yb + Yb * y);
        } else {
            T oX = X * X - Y * Y;
            T oXa = mul_s(2., X * Xa - Y * Ya), oXb = mul_s(2., X * Xb - Y * Yb);
            T ox = x * (x + mul_s(2., X)) - y * (y + mul_s(2., Y));
            T oxa = xa * (x + mul_s(2., X)) + x * (xa + mul_s(2., Xa)) - ya * (y + mul_s(2., Y)) - y * (ya + mul_s(2., Ya));
            T oxb = xb * (x + mul_s(2., X)) + x * (xb + mul_s(2., Xb)) - yb * (y + mul_s(2., Y)) - y * (yb + mul_s(2., Yb));
            double hX = ddX(oX, ox), hx = ddx(oX, ox);
            nxa = mul_s(hX, oXa) + mul_s(hx, oxa);
            nxb = mul_s(hX, oXb) + mul_s(hx, oxb);
        }
    } else if constexpr (FL == 2) {
        T da = diffabs(Y, y);
        double gX = ddX(Y, y), gx = ddx(Y, y);
        T Yy = Y + y, ab = absv(Yy);
        double sg = sgn(Yy);
        nxa = mul_s(2., ((X + x) * xa + Xa * x) - ((Y + y) * ya + Ya * y));
        nxb = mul_s(2., ((X + x) * xb + Xb * x) - ((Y + y) * yb + Yb * y));
        nya = mul_s(2., Xa * da + X * (mul_s(gX, Ya) + mul_s(gx, ya)) + xa * ab + mul_s(sg, x) * (Ya + ya));
        nyb = mul_s(2., Xb * da + X * (mul_s(gX, Yb) + mul_s(gx, yb)) + xb * ab + mul_s(sg, x) * (Yb + yb));
    } else if constexpr (FL == 3) {
        T da = diffabs(Y, y);
        double gX = ddX(Y, y), gx = ddx(Y, y);
        T Yy = Y + y, ab = absv(Yy);
        double sg = sgn(Yy);
        nxa = xa * (x + mul_s(2., X)) + x * (xa + mul_s(2., Xa)) - Ya * da - Y * (mul_s(gX, Ya) + mul_s(gx, ya)) - ya * ab - mul_s(sg, y) * (Ya + ya);
        nxb = xb * (x + mul_s(2., X)) + x * (xb + mul_s(2., Xb)) - Yb * da - Y * (mul_s(gX, Yb) + mul_s(gx, yb)) - yb * ab - mul_s(sg, y) * (Yb + yb);
        nya = mul_s(2., Xa * y + X * ya + Ya * x + Y * xa + xa * y + x * ya);
        nyb = mul_s(2., Xb * y + X * yb + Yb * x + Y * xb + xb * y + x * yb);
    } else {
        T oX = X * X - Y * Y;
        T oXa = mul_s(2., X * Xa - Y * Ya), oXb = mul_s(2., X * Xb - Y * Yb);
        T ox = x * (x + mul_s(2., X)) - y * (y + mul_s(2., Y));
        T oxa = xa * (x + mul_s(2., X)) + x * (xa + mul_s(2., Xa)) - ya * (y + mul_s(2., Y)) - y * (ya + mul_s(2., Ya));
        T oxb = xb * (x + mul_s(2., X)) + x * (xb + mul_s(2., Xb)) - yb * (y + mul_s(2., Y)) - y * (yb + mul_s(2., Yb));
        double hX = ddX(oX, ox), hx = ddx(oX, ox);
        nxa = mul_s(hX, oXa) + mul_s(hx, oxa);
        nxb = mul_s(hX, oXb) + mul_s(hx, oxb);
        nya = mul_s(2., Xa * y + X * ya + Ya * x + Y * xa + xa * y + x * ya);
        nyb = mul_s(2., Xb * y + X * yb + Yb * x + Y * xb + xb * y + x * yb);
    }
    P[0] = nxa; P[1] = nxb; P[2] = nya; P[3] = nyb;
}

}  // namespace fsg
