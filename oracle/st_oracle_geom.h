// Ray-element intersection of the CPU oracle.  TEST INFRASTRUCTURE ONLY.
// Restates coretrace/determineelementintersectionnew.cpp, intersect.cpp,
// quadricsurfaceclosedform.cpp, torusclosedform.cpp, root432.cpp, surface.cpp,
// surfacezatxypair.cpp and newzstartforcubicsplinesurf.cpp, keeping their
// floating-point evaluation order.
#pragma once
#include <algorithm>
#include <cctype>

#include "st_oracle_model.h"

namespace oracle {

static const double kSlope60 = 1.7320508075688767;  // tan(60 deg)

// coretrace/intersect.cpp:61,119-138
inline int sgn_nonneg(double x) { return x >= 0; }
inline bool in_triangle(double x1, double y1, double x2, double y2, double x3, double y3, double xt, double yt) {
    double a = (x1 - xt) * (y2 - yt) - (x2 - xt) * (y1 - yt);
    double b = (x2 - xt) * (y3 - yt) - (x3 - xt) * (y2 - yt);
    double c = (x3 - xt) * (y1 - yt) - (x1 - xt) * (y3 - yt);
    return sgn_nonneg(a) == sgn_nonneg(b) && sgn_nonneg(b) == sgn_nonneg(c);
}
inline bool in_quad(const double *p, double xt, double yt) {
    return in_triangle(p[0], p[1], p[2], p[3], p[4], p[5], xt, yt) ||
           in_triangle(p[0], p[1], p[4], p[5], p[6], p[7], xt, yt);
}

// Aperture membership of a surface point (x, y).  mode_plane selects the
// variant used on the aperture-plane crossing inside the Newton start logic
// (coretrace/intersect.cpp:272-583), where the rectangle test can never fail
// (:402,408) and the annulus does not bypass its radial test; otherwise it is
// the final test of coretrace/determineelementintersectionnew.cpp:104-623.
inline bool aperture_contains(const Element &e, double x, double y, bool mode_plane) {
    const double A = e.p[0], Bp = e.p[1], C = e.p[2];
    double r = sqrt(x * x + y * y);
    switch (e.shape) {
    case 'c': case 'C':
        return !(r > A / 2.0);
    case 'h': case 'H': {
        double Ro = A / 2.0;
        if (r > Ro) return false;
        double Ri = Ro * cos(30.0 * (kDegToRad));
        if (r <= Ri) return true;
        double XL = sqrt(Ro * Ro - Ri * Ri);
        if ((x <= Ro) && (x > XL)) {
            double Y1 = kSlope60 * (x - Ro);
            double Y2 = -Y1;
            return (y >= Y1) && (y <= Y2);
        }
        if ((x <= XL) && (x >= -XL)) return (y >= -Ri) && (y <= Ri);
        if ((x < -XL) && (x >= -Ro)) {
            double Y3 = kSlope60 * (x + Ro);
            double Y4 = -Y3;
            return (y >= Y4) && (y <= Y3);
        }
        return false;
    }
    case 't': case 'T': {
        double Ro = A / 2.0;
        if (r > Ro) return false;
        double Ri = Ro * sin(30.0 * (kDegToRad));
        if (r <= Ri) return true;
        if ((x <= Ro) && (x > 0.0)) {
            double Y1 = -kSlope60 * (x - Ri / cos(30.0 * (kDegToRad)));
            double Y2 = -Ri;
            return (y <= Y1) && (y >= Y2);
        }
        if ((x >= -Ro) && (x <= 0.0)) {
            double Y3 = kSlope60 * (x + Ri / cos(30.0 * (kDegToRad)));
            double Y4 = -Ri;
            return (y >= Y4) && (y <= Y3);
        }
        return false;
    }
    case 'r': case 'R':
        if (mode_plane) return true;  // intersect.cpp:402,408 conditions are never true
        if ((x > A / 2.0) || (x < -A / 2.0)) return false;
        if ((y > Bp / 2.0) || (y < -Bp / 2.0)) return false;
        return true;
    case 'a': case 'A': {
        bool radial = mode_plane || !((A == 0.0) && (Bp == 0.0));
        if (radial && ((r < A) || (r > Bp))) return false;
        double half = C * (kDegToRad) / 2.0;
        if (x >= 0.0) {
            return !((asin(y / r) > half) || (asin(y / r) < -half));
        }
        if (x < 0.0) {
            if ((y >= 0) && ((acos(y / r) + M_PI / 2.0) > half)) return false;
            if ((y < 0) && ((-acos(-y / r) - M_PI / 2.0) < -half)) return false;
            return true;
        }
        return false;  // x is NaN
    }
    case 'l': case 'L':
        if (!((A == 0.0) && (Bp == 0.0))) {
            if ((x < A) || (x > Bp)) return false;
        }
        if ((y < -C / 2.0) || (y > C / 2.0)) return false;
        return true;
    case 'i': case 'I':
        return in_triangle(e.p[0], e.p[1], e.p[2], e.p[3], e.p[4], e.p[5], x, y);
    case 'q': case 'Q':
        return in_quad(e.p, x, y);
    default:
        return false;
    }
}

// (x*x) with x = 2*K*d/a, expanded the way the reference's parenthesis-free
// sqr macro expands (coretrace/quadricsurfaceclosedform.cpp:55,297-306).
inline double sqr_grad(double K, double d, double a2) { return (((((2.0 * K) * d) / a2) * 2.0) * K) * d / a2; }

// coretrace/quadricsurfaceclosedform.cpp:58-308
inline void quadric_closed_form(const Element &e, const double P[3], const double D[3], double X[3], double DF[3],
                                double *path, int *err) {
    double Xc = 0, Yc = 0, Zc = 0, Kx = 0, Ky = 0, Kz = 0, r = 0, r2 = 0, a2 = 0, b2 = 0, c2 = 0;
    double A = 0, B = 0, C = 0, t1, t2;
    *err = 0;
    switch (e.surf) {
    case 's': case 'S': {
        a2 = 1; b2 = 1; c2 = 1; Kx = 1;
        Ky = (e.stype == 7) ? 0 : 1;
        Kz = 1;
        r = 1.0 / e.vcx;
        r2 = r * r;
        Zc = r;
        double xd = P[0] - Xc, yd = P[1] - Yc, zd = P[2] - Zc;
        A = D[0] * D[0] * Kx / a2 + D[1] * D[1] * Ky / b2 + D[2] * D[2] * Kz / c2;
        B = 2.0 * (Kx * xd * D[0] / a2 + Ky * yd * D[1] / b2 + Kz * zd * D[2] / c2);
        C = Kx * xd * xd / a2 + Ky * yd * yd / b2 + Kz * zd * zd / c2 - r2;
        break;
    }
    case 'p': case 'P': {
        a2 = 2.0 / e.vcx;
        b2 = e.vcy > 0.0 ? 2.0 / e.vcy : 1e10;
        c2 = 1.0;
        Kx = 1;
        Ky = (e.stype == 7) ? 0 : 1;
        Kz = 0;
        double xd = P[0] - Xc, yd = P[1] - Yc, zd = P[2] - Zc;
        A = D[0] * D[0] * Kx / a2 + D[1] * D[1] * Ky / b2;
        B = 2.0 * (Kx * xd * D[0] / a2 + Ky * yd * D[1] / b2) - D[2] / c2;
        C = Kx * xd * xd / a2 + Ky * yd * yd / b2 - zd / c2;
        break;
    }
    case 'o': case 'O': {
        a2 = 1; b2 = 1; c2 = 1; Kx = 1; Ky = 1;
        Kz = e.kappa;
        Zc = 1.0 / e.kappa / e.vcx;
        double xd = P[0] - Xc, yd = P[1] - Yc, zd = P[2] - Zc;
        A = D[0] * D[0] * Kx / a2 + D[1] * D[1] * Ky / b2 + D[2] * D[2] * Kz / c2;
        B = 2.0 * (Kx * xd * D[0] / a2 + Ky * yd * D[1] / b2 + Kz * zd * D[2] / c2);
        C = Kx * xd * xd / a2 + Ky * yd * yd / b2 + Kz * zd * zd / c2 - 1.0 / e.kappa / (e.vcx * e.vcx);
        break;
    }
    case 't': case 'T': {
        a2 = 1; b2 = 1; c2 = 1; Kx = 1; Ky = 0; Kz = 1;
        r = 1.0 / e.curv_rev;
        r2 = r * r;
        Zc = r;
        double xd = P[0] - Xc, yd = P[1] - Yc, zd = P[2] - Zc;
        A = D[0] * D[0] * Kx / a2 + D[1] * D[1] * Ky / b2 + D[2] * D[2] * Kz / c2;
        B = 2.0 * (Kx * xd * D[0] / a2 + Ky * yd * D[1] / b2 + Kz * zd * D[2] / c2);
        C = Kx * xd * xd / a2 + Ky * yd * yd / a2 + Kz * zd * zd / c2 - r2;
        break;
    }
    default:
        break;
    }

    bool have = false;
    if (fabs(A) < 1e-12) {
        t1 = -C / B;
        if (t1 > 0) {
            for (int i = 0; i < 3; i++) X[i] = P[i] + t1 * D[i];
            *path = t1;
            have = true;
        } else {
            *path = 0.0;
            *err = 1;
            return;
        }
    }
    if (!have) {
        if ((B * B) > 4.0 * A * C) {
            t1 = (-B + sqrt((B * B) - 4.0 * A * C)) / (2.0 * A);
            t2 = (-B - sqrt((B * B) - 4.0 * A * C)) / (2.0 * A);
            if (t2 > 0) {
                for (int i = 0; i < 3; i++) X[i] = P[i] + t2 * D[i];
                *path = t2;
                if (X[2] > e.zap) {
                    for (int i = 0; i < 3; i++) X[i] = P[i] + t1 * D[i];
                    *path = t1;
                }
                bool len_check = (e.surf == 't' || e.surf == 'T') ||
                                 ((e.surf == 's' || e.surf == 'S') && e.stype == 7);
                if (len_check && ((X[1] < -e.p[2] / 2.0) || (X[1] > e.p[2] / 2.0))) {
                    for (int i = 0; i < 3; i++) X[i] = P[i] + t1 * D[i];
                    *path = t1;
                }
            } else if (t2 == 0) {
                for (int i = 0; i < 3; i++) X[i] = P[i] + t1 * D[i];
                *path = t1;
            } else if (t2 < 0 && t1 > 0) {
                for (int i = 0; i < 3; i++) X[i] = P[i] + t1 * D[i];
                *path = t1;
            } else if (t1 <= 0) {
                *path = t1;
                *err = 1;
                return;
            }
            // any other outcome (NaN roots) falls through with X left as is
        } else {
            *path = 0.0;
            *err = 1;
            return;
        }
    }

    if (e.surf == 'p' || e.surf == 'P') {
        double sm = sqrt(sqr_grad(Kx, X[0] - Xc, a2) + sqr_grad(Ky, X[1] - Yc, b2) + 1.0);
        DF[0] = -(2.0 * Kx * (X[0] - Xc) / a2) / sm;
        DF[1] = -(2.0 * Ky * (X[1] - Yc) / b2) / sm;
        DF[2] = 1.0 / sm;
    } else {
        double sm = sqrt(sqr_grad(Kx, X[0] - Xc, a2) + sqr_grad(Ky, X[1] - Yc, b2) + sqr_grad(Kz, X[2] - Zc, c2));
        DF[0] = -(2.0 * Kx * (X[0] - Xc) / a2) / sm;
        DF[1] = -(2.0 * Ky * (X[1] - Yc) / b2) / sm;
        DF[2] = -(2.0 * Kz * (X[2] - Zc) / c2) / sm;
    }
}

// Ferrari quartic via a bisected resolvent cubic, coretrace/root432.cpp:53-284.
// State lives in one struct per call, as in the reference's CR432 object.
struct Quartic {
    double a[5][5];
    double r[5];
    double aa = 0, b = 0, c = 0, d = 0, ii = 0, im = 0, k = 0, l = 0, m = 0, q = 0, rr = 0, s = 0, sw = 0;
    double f(double x) const { return x * x * x + a[3][2] * x * x + a[3][1] * x + a[3][0]; }
    void root2() {
        double w = a[2][1] * a[2][1] - 4 * a[2][0];
        double q1 = -a[2][1] / 2.0;
        if (w < 0) {
            im = sqrt(-w) / 2.0;
            r[1] = q1;
            r[2] = q1;
        } else if (w == 0) {
            r[1] = q1;
            r[2] = q1;
            im = 0.0;
        } else if (w > 0) {
            double q2 = sqrt(w) / 2.0;
            im = 0.0;
            r[1] = q1 + q2;
            r[2] = q1 - q2;
        }
    }
    void root3() {
        double xc;
        if (a[3][0] == 0) {
            xc = 0;
        } else {
            double am = fabs(a[3][0]);
            for (int i = 1; i <= 3; i++) {
                double tt = fabs(a[3][i]);
                if (am < tt) am = tt;
            }
            if (a[3][0] > 0) {
                aa = -am - 1;
                b = 0;
            } else {
                aa = 0;
                b = am + 1;
            }
            double xa = aa, xb = b, y1 = f(aa), y2 = f(b);
            const double er = 0.000001;
            if (y1 == 0) {
                xc = aa;
            } else if (y2 == 0) {
                xc = b;
            } else {
                for (;;) {
                    xc = (xa + xb) / 2.0;
                    double te = f(xc);
                    if (te == 0) break;
                    if (xb - xa < er) break;
                    if (f(xa) * te > 0)
                        xa = xc;
                    else
                        xb = xc;
                }
            }
        }
        r[3] = xc;
        if (sw == -1) return;
        a[2][1] = a[3][2] + xc;
        a[2][0] = a[3][1] + a[3][2] * xc + xc * xc;
        root2();
    }
    void solve4() {
        const int n = 4;
        for (int i = 0; i < n; i++) a[n][i] = a[n][i] / a[n][n];
        aa = a[4][3];
        b = a[4][2];
        c = a[4][1];
        d = a[4][0];
        q = b - (3.0 * aa * aa / 8.0);
        rr = c - (aa * b / 2) + (aa * aa * aa / 8.0);
        s = d - (aa * c / 4) + (aa * aa * b / 16.0) - (3 * aa * aa * aa * aa / 256.0);
        a[3][2] = q / 2.0;
        a[3][1] = (q * q - 4 * s) / 16.0;
        a[3][0] = -(rr * rr / 64.0);
        if ((rr != 0) || (a[3][1] >= 0)) {
            sw = -1;
            root3();
        } else {
            a[2][1] = a[3][2];
            a[2][0] = a[3][1];
            root2();
            r[3] = r[1];
        }
        k = sqrt(r[3]);
        if (k == 0) {
            rr = sqrt(q * q - 4 * s);
        } else {
            q = q + (4 * r[3]);
            rr = rr / (2 * k);
        }
        l = (q - rr) / 2.0;
        m = (q + rr) / 2.0;
        a[2][1] = 2 * k;
        a[2][0] = l;
        root2();
        r[3] = r[1] - (a[4][3] / 4);
        r[4] = r[2] - (a[4][3] / 4.0);
        ii = im;
        a[2][1] = -2 * k;
        a[2][0] = m;
        root2();
        r[2] = r[2] - (a[4][3] / 4.0);
        r[1] = r[1] - (a[4][3] / 4.0);
    }
};

// coretrace/torusclosedform.cpp:57-188
inline void torus_closed_form(const Element &e, const double P[3], const double D[3], double X[3], double DF[3],
                              double *path, int *err) {
    const double Rs = e.cross_r, Ra = e.annular_r;
    const double Xo = P[0], Yo = P[1], Zo = P[2];
    const double Ep = D[0], Et = D[1], Rh = D[2];
    *err = 0;
    Quartic Q;
    for (int i = 0; i < 5; i++) {
        Q.r[i] = 0.0;
        for (int j = 0; j < 5; j++) Q.a[i][j] = 0.0;
    }
    Q.a[4][4] = pow(Ep, 4) + 2.0 * Ep * Ep * (Et * Et + Rh * Rh) + pow(Et, 4) + 2.0 * Et * Et * Rh * Rh + pow(Rh, 4);
    Q.a[4][3] = 4.0 * (Ep * Ep + Et * Et + Rh * Rh) * (Ep * Xo + Et * Yo + Rh * Zo - Rh * Rs);
    Q.a[4][2] = Xo * Xo * (6.0 * Ep * Ep + 2.0 * Et * Et + 2.0 * Rh * Rh) + 8.0 * Ep * Xo * (Et * Yo + Rh * Zo - Rh * Rs) +
                2.0 * Yo * Yo * (Ep * Ep + 3.0 * Et * Et + Rh * Rh) + 8.0 * Et * Rh * Yo * (Zo - Rs) +
                (Ep * Ep + Et * Et + 3.0 * Rh * Rh) * (2.0 * Zo * Zo - 4.0 * Rs * Zo) -
                2.0 * Ra * Ra * (Ep * Ep + Et * Et - Rh * Rh) + 4.0 * Rh * Rh * Rs * Rs;
    Q.a[4][1] = 4.0 * (Xo * Xo * (Ep * Xo + Et * Yo + Rh * Zo - Rh * Rs) + Yo * Yo * (Ep * Xo + Et * Yo + Rh * Zo - Rh * Rs) +
                       Zo * Zo * (Ep * Xo + Et * Yo + Rh * Zo - 3.0 * Rh * Rs) - 2.0 * Ep * Rs * Xo * Zo - Ep * Ra * Ra * Xo -
                       2.0 * Et * Rs * Yo * Zo - Et * Ra * Ra * Yo + Rh * Ra * Ra * (Zo - Rs) + 2.0 * Rh * Rs * Rs * Zo);
    Q.a[4][0] = pow(Xo, 4) + 2.0 * Xo * Xo * (Yo * Yo + Zo * Zo - 2.0 * Rs * Zo - Ra * Ra) + pow(Yo, 4) +
                2.0 * Yo * Yo * (Zo * Zo - 2.0 * Rs * Zo - Ra * Ra) + pow(Zo, 4) - 4.0 * Rs * Zo * Zo * Zo +
                2.0 * Ra * Ra * Zo * Zo + 4.0 * Rs * Rs * Zo * Zo - 4.0 * Ra * Ra * Rs * Zo + pow(Ra, 4);

    // Root_432(4, ...): copy, solve, pass results (root432.cpp:250-284)
    Quartic S;
    for (int i = 0; i < 5; i++) {
        S.r[i] = 0.0;
        for (int j = 0; j < 5; j++) S.a[i][j] = Q.a[i][j];
    }
    S.solve4();
    double rv[5];
    for (int i = 0; i < 5; i++) rv[i] = S.r[i];
    double im1 = S.im, im2 = S.ii;

    double L = 0.0;
    if (im1 == 0.0 && im2 == 0.0) {
        piksrt_ref(4, rv);
        if (rv[4] <= 0.0) {
            *path = rv[4];
            *err = 1;
            return;
        }
        L = rv[4];
        if (!(rv[3] <= 0.0)) {
            L = rv[3];
            if (!(rv[2] <= 0.0)) {
                L = rv[2];
                if (!(rv[1] <= 0.0)) L = rv[1];
            }
        }
    } else if (im1 == 0.0 && im2 != 0.0) {
        L = (rv[1] <= rv[2]) ? rv[1] : rv[2];
    } else if (im1 != 0.0 && im2 == 0.0) {
        L = (rv[3] <= rv[4]) ? rv[3] : rv[4];
    } else {
        L = 0.0;
    }
    *path = L;
    if (L == 0.0) {
        *err = 1;
        return;
    }
    double x = Xo + L * Ep, y = Yo + L * Et, z = Zo + L * Rh;
    double Fx = -2.0 * x * (Ra - sqrt(x * x + y * y)) / sqrt(x * x + y * y);
    double Fy = -2.0 * y * (Ra - sqrt(x * x + y * y)) / sqrt(x * x + y * y);
    double Fz = 2.0 * (z - Rs);
    X[0] = x; X[1] = y; X[2] = z;
    DF[0] = -Fx; DF[1] = -Fy; DF[2] = -Fz;
}

// Measured slopes at (x, y) by the modified Shepard weighting of the VSHOT
// points within the region of interest (coretrace/vshot.cpp:290-367).
inline void vshot_shepard(const Element &e, double x, double y, double *zx, double *zy) {
    const size_t n = e.vshot.size() / 5;
    const double density = (double)n / e.area;
    const double roi2 = 30.0 / density / 3.14159;
    double s = 0.0, wx = 0.0, wy = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double *v = &e.vshot[i * 5];
        const double xx = v[0] - x, yy = v[1] - y;
        if (xx == 0.0 && yy == 0.0) {
            *zx = v[2];
            *zy = v[3];
            return;
        }
        const double r2 = xx * xx + yy * yy;
        if (r2 < roi2) {
            s = s + 1.0 / r2;
            wx = wx + v[2] / r2;
            wy = wy + v[3] / r2;
        }
    }
    if (s == 0.0) {
        *zx = 0.0;
        *zy = 0.0;
        return;
    }
    *zx = wx / s;
    *zy = wy / s;
}

// Local kriging model of a finite-element surface at (x, y)
// (coretrace/vshot.cpp:166-238, FEInterpKD, with the GaussMarkov model and
// LU solve of interpolate.cpp): the nodes of the hash-tree zone holding the
// point and of its neighbours, sorted by distance; the nearest seven fit a
// power variogram (beta 1.999, no nugget); z there and one-sided
// differences `step` away give the slopes.
struct FeFailure {};
inline void fe_interp(const Element &e, double x, double y, double step, double *zr, double *dzdx, double *dzdy) {
    std::vector<int> near;
    if (!e.fe_tree.data_at(x, y, near) || near.size() < 3) throw FeFailure();
    struct RE { double r; int i; };
    std::vector<RE> nd;
    for (int i : near) {
        const double *v = &e.fe_nodes[3 * (size_t)i];
        nd.push_back(RE{std::sqrt(std::pow(x - v[0], 2) + std::pow(y - v[1], 2)), i});
    }
    std::sort(nd.begin(), nd.end(), [](const RE &a, const RE &b) { return a.r < b.r; });
    const int m = (int)nd.size() > 7 ? 7 : (int)nd.size();
    std::vector<double> px(m), py(m), pz(m + 1);
    for (int k = 0; k < m; k++) {
        const double *v = &e.fe_nodes[3 * (size_t)nd[k].i];
        px[k] = v[0]; py[k] = v[1]; pz[k] = v[2];
    }
    const double bet = 1.999;
    double num = 0., den = 0.;
    for (int i = 0; i < m; i++)
        for (int j = i + 1; j < m; j++) {
            double rb = 0.;
            rb += (px[i] - px[j]) * (px[i] - px[j]);
            rb += (py[i] - py[j]) * (py[i] - py[j]);
            rb = std::pow(rb, 0.5 * bet);
            num += rb * (0.5 * ((pz[i] - pz[j]) * (pz[i] - pz[j])) - 0.0);
            den += rb * rb;
        }
    const double alph = num / den;
    auto vg = [&](double r) { return 0.0 + alph * std::pow(r, bet); };
    auto rd = [&](double ax, double ay, double bx, double by) {
        double d = 0.;
        d += (ax - bx) * (ax - bx);
        d += (ay - by) * (ay - by);
        return std::sqrt(d);
    };
    const int n = m + 1;
    std::vector<std::vector<double>> lu(n, std::vector<double>(n, 0.0));
    for (int i = 0; i < m; i++) {
        for (int j = i; j < m; j++) lu[i][j] = lu[j][i] = vg(rd(px[i], py[i], px[j], py[j]));
        lu[i][m] = lu[m][i] = 1.;
    }
    lu[m][m] = pz[m] = 0.;
    // LU decomposition with implicit partial pivoting (Numerical Recipes)
    std::vector<int> indx(n);
    std::vector<double> vv(n);
    for (int i = 0; i < n; i++) {
        double big = 0.0, temp;
        for (int j = 0; j < n; j++)
            if ((temp = std::fabs(lu[i][j])) > big) big = temp;
        if (big == 0.0) throw FeFailure();
        vv[i] = 1.0 / big;
    }
    // imax keeps its value from the previous column when no pivot candidate
    // compares greater than zero (a zero or NaN column, as duplicate nodes
    // give), as in the reference, which declares it once for the whole loop
    int imax = 0;
    for (int k = 0; k < n; k++) {
        double big = 0.0, temp;
        for (int i = k; i < n; i++) {
            temp = vv[i] * std::fabs(lu[i][k]);
            if (temp > big) { big = temp; imax = i; }
        }
        if (k != imax) {
            for (int j = 0; j < n; j++) std::swap(lu[imax][j], lu[k][j]);
            vv[imax] = vv[k];
        }
        indx[k] = imax;
        if (lu[k][k] == 0.0) lu[k][k] = 1.0e-40;
        for (int i = k + 1; i < n; i++) {
            temp = lu[i][k] /= lu[k][k];
            for (int j = k + 1; j < n; j++) lu[i][j] -= temp * lu[k][j];
        }
    }
    std::vector<double> yv(pz);
    {
        int ii = 0;
        for (int i = 0; i < n; i++) {
            const int ip = indx[i];
            double sum = yv[ip];
            yv[ip] = yv[i];
            if (ii != 0)
                for (int j = ii - 1; j < i; j++) sum -= lu[i][j] * yv[j];
            else if (sum != 0.0)
                ii = i + 1;
            yv[i] = sum;
        }
        for (int i = n - 1; i >= 0; i--) {
            double sum = yv[i];
            for (int j = i + 1; j < n; j++) sum -= lu[i][j] * yv[j];
            yv[i] = sum / lu[i][i];
        }
    }
    auto interp = [&](double qx, double qy) {
        double v = 0.;
        for (int i = 0; i < m; i++) v += yv[i] * vg(rd(qx, qy, px[i], py[i]));
        v += yv[m] * 1.;
        return v;
    };
    double xs = x, ys = y;
    *zr = interp(xs, ys);
    xs += step;
    const double za = interp(xs, ys);
    xs -= step;
    ys += step;
    const double zb = interp(xs, ys);
    *dzdx = (*zr - za) / step;
    *dzdy = (*zr - zb) / step;
}

// coretrace/surface.cpp:57-413 (types 1/7 incl. cone, 3, 4, 5, 6, 8, 9)
inline void surface_eval(double X[3], const Element &e, double *F, double DF[3], int *err) {
    double x = X[0], y = X[1], z = X[2];
    double dfx = 0, dfy = 0, dfz = 0;
    *err = 0;
    if (e.stype == 1 || e.stype == 7) {
        double rho2 = (e.stype == 1) ? x * x + y * y : x * x;
        if (e.cone_half != 0.0) {
            *F = z - sqrt(rho2) / tan(e.cone_half * (kDegToRad));
            dfx = -x / (sqrt(rho2) * tan(e.cone_half * (kDegToRad)));
            dfy = -y / (sqrt(rho2) * tan(e.cone_half * (kDegToRad)));
            dfz = 1.0;
        } else {
            const double sum1 = 0.0, sum2 = 0.0;  // the Alpha sums are disabled in the reference
            if (1.0 - e.kappa * (e.vcx * e.vcx * x * x + e.vcy * e.vcy * y * y) < 0) *err = 1;
            double term = sqrt(1.0 - e.kappa * (e.vcx * e.vcx * x * x + e.vcy * e.vcy * y * y));
            *F = z - (e.vcx * x * x + e.vcy * y * y) / (1.0 + term) - sum2;
            dfx = -x * (e.vcx / term + 2.0 * sum1);
            dfy = -y * (e.vcy / term + 2.0 * sum1);
            dfz = 1.0;
        }
    } else if (e.stype == 3) {
        dfx = e.alpha[0];
        dfy = e.alpha[1];
        dfz = e.alpha[2];
        *F = dfx * x + dfy * y + dfz * z - e.alpha[3];
    } else if (e.stype == 4) {
        // finite element (coretrace/surface.cpp:218-242)
        if (x * x + y * y == 0.0) {
            *F = z;
            DF[0] = 0.0;
            DF[1] = 0.0;
            DF[2] = 1.0;
            return;
        }
        const double density = (double)(e.fe_nodes.size() / 3) / e.area;
        const double delta = 0.001 / sqrt(density);
        double zr = 0.0, dzx = 0.0, dzy = 0.0;
        fe_interp(e, x, y, delta, &zr, &dzx, &dzy);
        X[2] = zr;
        *F = z - zr;
        dfx = dzx;
        dfy = dzy;
        dfz = 1.0;
    } else if (e.stype == 5) {
        // VSHOT (coretrace/surface.cpp:244-298): z from the monomial fit, the
        // normal from the interpolated measured slopes (angles, rad)
        if (x * x + y * y == 0.0) {
            *F = z;
            DF[0] = 0.0;
            DF[1] = 0.0;
            DF[2] = 1.0;
            return;
        }
        double zm = eval_mono(x, y, e);
        double zx = 0.0, zy = 0.0;
        vshot_shepard(e, x, y, &zx, &zy);
        X[2] = zm;
        *F = z - zm;
        dfx = -tan(zx);
        dfy = -tan(zy);
        dfz = 1.0;
    } else if (e.stype == 6) {
        double zz = eval_mono(x, y, e);
        mono_slope(e, x, y, &dfx, &dfy);
        X[2] = zz;
        *F = z - zz;
        dfx = -dfx;
        dfy = -dfy;
        dfz = 1.0;
    } else if (e.stype == 8) {
        double yval = (e.shape == 'l' || e.shape == 'L') ? 0.0 : y;
        double zz = eval_poly(x, yval, e);
        poly_slope(e, x, yval, &dfx, &dfy);
        X[2] = zz;
        *F = z - zz;
        dfx = -dfx;
        dfy = -dfy;
        dfz = 1.0;
    } else if (e.stype == 9) {
        double rho = sqrt(x * x + y * y);
        double drdx = x / rho, drdy = y / rho;
        if (e.shape == 'l' || e.shape == 'L') {
            rho = x;
            drdx = 1.0;
            drdy = 0.0;
        }
        double zz = 0.0, dzdr = 0.0;
        if (!splint(e.sx, e.sy, e.sy2, (int)e.sx.size(), rho, &zz, &dzdr)) {
            *err = 3;
            return;
        }
        dfx = dzdr * drdx;
        dfy = dzdr * drdy;
        X[2] = zz;
        *F = z - zz;
        dfx = -dfx;
        dfy = -dfy;
        dfz = 1.0;
    }
    DF[0] = dfx;
    DF[1] = dfy;
    DF[2] = dfz;
}

// coretrace/surfacezatxypair.cpp:56-280; *F is left untouched for types the
// reference does not handle there.
inline void surface_z_at_xy(const double X[3], const Element &e, double *F) {
    double x = X[0], y = X[1];
    if (e.stype == 1 || e.stype == 7) {
        double rho2 = (e.stype == 1) ? x * x + y * y : x * x;
        double rho = sqrt(rho2);
        (void)rho;
        if (e.cone_half != 0.0) {
            *F = sqrt(rho2) / tan(e.cone_half * (kDegToRad));
            return;
        }
        if (e.kappa * (e.vcx * e.vcx * x * x + e.vcy * e.vcy * y * y) > 1.0) {
            *F = 0.0;
            return;
        }
        *F = (e.vcx * x * x + e.vcy * y * y) /
             (1.0 + sqrt(1.0 - e.kappa * (e.vcx * e.vcx * x * x + e.vcy * e.vcy * y * y)));
        return;
    }
    if (e.stype == 6) {
        *F = eval_mono(x, y, e);
        return;
    }
    if (e.stype == 5) {  // coretrace/surfacezatxypair.cpp:195-207
        *F = (x * x + y * y == 0.0) ? 0.0 : eval_mono(x, y, e);
        return;
    }
    if (e.stype == 8) {
        double yval = (e.shape == 'l' || e.shape == 'L') ? 0.0 : y;
        *F = eval_poly(x, yval, e);
        return;
    }
}

// coretrace/newzstartforcubicsplinesurf.cpp:57-142
inline void spline_zstart(double CR, const double P[3], const double D[3], char shape, double *zs, double *pl,
                          int *ef) {
    *ef = 0;
    if (shape == 'a' || shape == 'A') {
        double A = D[0] * D[0] + D[1] * D[1];
        double B = 2.0 * (P[0] * D[0] + P[1] * D[1]);
        double C = P[0] * P[0] + P[1] * P[1] - CR * CR;
        if (B * B > 4.0 * A * C) {
            double t1 = (-B + sqrt(B * B - 4.0 * A * C)) / (2.0 * A);
            double t2 = (-B - sqrt(B * B - 4.0 * A * C)) / (2.0 * A);
            if (t2 > 0) {
                *zs = P[2] + t2 * D[2];
                *pl = t2;
                return;
            }
            if (t2 == 0) {
                *zs = P[2] + t1 * D[2];
                *pl = t1;
                return;
            }
            if (t2 < 0 && t1 > 0) {
                *zs = P[2] + t1 * D[2];
                *pl = t1;
                return;
            }
            if (t1 <= 0) {
                *pl = t1;
                *ef = 1;
                return;
            }
        } else {
            *pl = 0.0;
            *ef = 2;
            return;
        }
    } else if (shape == 'l' || shape == 'L') {
        double t1 = (CR - P[0]) / D[0];
        if (t1 >= 0.0) {
            *zs = P[2] + t1 * D[2];
            *pl = t1;
        } else {
            *pl = 0.0;
            *ef = 1;
        }
    }
}

// Newton start plane for cubic-spline surfaces, coretrace/intersect.cpp:637-946.
// Returns false when splint fails (ErrorFlag 3).
inline bool spline_start_plane(const Element &e, const double X[3], const double D[3], const double P0[3],
                               double *ZStart) {
    const double OuterR = e.sx[e.sx.size() - 1];
    const double InnerR = e.sx[0];
    const char ap = e.shape;
    const bool isA = (ap == 'a' || ap == 'A'), isL = (ap == 'l' || ap == 'L');
    const double ZA = e.sy[e.sy.size() - 1];
    double S00 = -X[2] / (D[2] + 0.00000000001);
    double X10 = X[0] + D[0] * S00, Y10 = X[1] + D[1] * S00;
    double R10 = sqrt(X10 * X10 + Y10 * Y10);
    double S0A = (ZA - X[2]) / (D[2] + 0.00000000001);
    double X1A = X[0] + D[0] * S0A, Y1A = X[1] + D[1] * S0A;
    double R1A = sqrt(X1A * X1A + Y1A * Y1A);
    double zcs = 0, plcs = 0;
    int efcs = 0;
    auto try_cyl = [&](double CR, const double *from) {
        spline_zstart(CR, from, D, ap, &zcs, &plcs, &efcs);
        if (efcs == 0) *ZStart = zcs;
    };

    if (X[2] >= ZA && D[2] < 0.0) {
        double PA[3] = {X1A, Y1A, ZA};
        if ((isA && (R1A > OuterR)) || (isL && (X1A > OuterR))) {
            try_cyl(OuterR * 0.999999, PA);
            return true;
        }
        if ((isA && (R1A <= OuterR)) || (isL && (X1A <= OuterR))) {
            if ((isA && (R10 >= OuterR)) || (isL && (X10 >= OuterR))) {
                try_cyl(OuterR * 0.999999, PA);
                return true;
            }
            *ZStart = 0.0;
            return true;
        }
    }
    if (X[2] <= 0.0 && D[2] > 0.0) {
        double PZ[3] = {X10, Y10, 0.0};
        if ((isA && (R10 >= OuterR)) || (isL && (X10 >= OuterR))) {
            try_cyl(OuterR * 0.999999, PZ);
            return true;
        }
        if ((isA && ((R10 < OuterR) && (R10 > InnerR))) || (isL && ((X10 < OuterR) && (X10 > InnerR)))) {
            *ZStart = 0.0;
            return true;
        }
        if ((isA && (R10 <= InnerR)) || (isL && (X10 <= InnerR))) {
            try_cyl(InnerR / 0.999999, PZ);
            return true;
        }
    }
    if (X[2] < ZA && X[2] > 0.0 && D[2] < 0.0) {
        double R1 = sqrt(X[0] * X[0] + X[1] * X[1]);
        if ((isA && (R1 >= OuterR)) || (isL && (X[0] >= OuterR))) {
            try_cyl(OuterR * 0.999999, P0);
            return true;
        }
        if ((isA && ((R1 < OuterR) && (R1 > InnerR))) || (isL && ((X[0] < OuterR) && (X[0] > InnerR)))) {
            double in = isA ? R1 : X[0];
            double Z1, dz;
            if (!splint(e.sx, e.sy, e.sy2, (int)e.sx.size(), in, &Z1, &dz)) return false;
            if (Z1 < X[2]) {
                if ((isA && (R10 >= OuterR)) || (isL && (X10 >= OuterR))) {
                    try_cyl(OuterR * 0.999999, P0);
                    return true;
                }
                *ZStart = 0.0;
                return true;
            }
            if (Z1 >= X[2]) {
                *ZStart = X[2];
                return true;
            }
            return true;
        }
        if ((isA && (R1 <= InnerR)) || (isL && (X[0] <= InnerR))) {
            try_cyl(InnerR / 0.999999, P0);
            return true;
        }
    }
    if (X[2] < ZA && X[2] > 0.0 && D[2] > 0.0) {
        double R1 = sqrt(X[0] * X[0] + X[1] * X[1]);
        if ((isA && (R1 >= OuterR)) || (isL && (X[0] >= OuterR))) {
            try_cyl(OuterR * 0.999999, P0);
            return true;
        }
        if ((isA && ((R1 < OuterR) && (R1 > InnerR))) || (isL && ((X[0] < OuterR) && (X[0] > InnerR)))) {
            double in = isA ? R1 : X[0];
            double Z1, dz;
            if (!splint(e.sx, e.sy, e.sy2, (int)e.sx.size(), in, &Z1, &dz)) return false;
            if (Z1 < X[2]) {
                if ((isA && (R1A >= OuterR)) || (isL && (X1A >= OuterR))) {
                    try_cyl(OuterR * 0.999999, P0);
                    return true;
                }
                return true;
            }
            if (Z1 >= X[2]) {
                if ((isA && (R10 < OuterR)) || (isL && (X10 < OuterR))) {
                    *ZStart = 0.0;
                    return true;
                }
                if ((isA && (R10 >= OuterR)) || (isL && (X10 >= OuterR))) {
                    double PD[3] = {X10, Y10, 0.0};
                    try_cyl(OuterR * 0.999999, PD);
                    return true;
                }
            }
        }
        if ((isA && (R1 <= InnerR)) || (isL && (X[0] <= InnerR))) {
            try_cyl(OuterR * 0.999999, P0);
            return true;
        }
    }
    return true;
}

// Optional torus start point (off unless st_oracle_set_torus_start is
// called).  From an origin kilometres away the reference quartic is
// ill-conditioned: a 1-ulp change of the origin flips the outcome of a
// sizeable fraction of rays, so no other build can reproduce those answers.
// The HIP library therefore starts the solve on the ray where it enters
// the sphere of radius 2r+1 about the element's hit-bounding sphere (c, r)
// when the origin lies beyond it.  This restates that sphere for a torus
// (x, y from the aperture as the library bounds them, z in [0, 2 Rs], padded
// by 1e-4 + 1e-7 max|x,y|) so the oracle can run the same reference solve
// from the same start point.
inline void torus_start_sphere(Element &e) {
    const double A = e.p[0], B = e.p[1], C = e.p[2];
    double xl, xh, yl, yh;
    switch (tolower(e.shape)) {
    case 'c': case 'h': case 't':
        xl = yl = -A / 2.0; xh = yh = A / 2.0;
        break;
    case 'r':
        xl = -A / 2.0; xh = A / 2.0; yl = -B / 2.0; yh = B / 2.0;
        break;
    case 'a':
        if (A == 0.0 && B == 0.0) {
            double r = e.annular_r + e.cross_r;
            xl = yl = -r; xh = yh = r;
        } else {
            xl = yl = -fabs(B); xh = yh = fabs(B);
        }
        break;
    case 'l':
        if (A == 0.0 && B == 0.0) { xl = -1e3; xh = 1e3; }
        else { xl = std::min(A, B); xh = std::max(A, B); }
        yl = -C / 2.0; yh = C / 2.0;
        break;
    default: {
        int nv = (tolower(e.shape) == 'q') ? 4 : 3;
        xl = yl = 1e300; xh = yh = -1e300;
        for (int k = 0; k < nv; k++) {
            xl = std::min(xl, e.p[2 * k]); xh = std::max(xh, e.p[2 * k]);
            yl = std::min(yl, e.p[2 * k + 1]); yh = std::max(yh, e.p[2 * k + 1]);
        }
        break;
    }
    }
    const double zl = 0.0, zh = 2.0 * e.cross_r;
    const double pad = 1e-4 + 1e-7 * std::max(std::max(fabs(xl), fabs(xh)), std::max(fabs(yl), fabs(yh)));
    const double lo[3] = {xl - pad, yl - pad, zl - pad}, hi[3] = {xh + pad, yh + pad, zh + pad};
    for (int k = 0; k < 3; k++) e.torus_start_c[k] = (lo[k] + hi[k]) / 2;
    e.torus_start_r = 0.5 * sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                 (hi[2] - lo[2]) * (hi[2] - lo[2]));
}

// The torus of torus_closed_form lies in the shell |z - Rs| <= |Rs|,
// rho <= |Ra| + |Rs| and, for a ring torus, rho >= |Ra| - |Rs|; a line that
// stays 1e-3 (|Ra| + |Rs|) outside it has no real root, and the solve is
// skipped (the no-root outcome: error 1, path 0).  Only with the start
// sphere (the GPU library's arithmetic): the reference's own solve from a
// far origin is ill-conditioned enough to return roots off the torus, which
// the reference-stream mode keeps.
inline bool torus_shell(const Element &e, const double P[3], const double D[3]) {
    const double Rs = e.cross_r, Ra = e.annular_r;
    const double m = 1e-3 * (fabs(Ra) + fabs(Rs));
    const double z0 = Rs - fabs(Rs) - m, z1 = Rs + fabs(Rs) + m;
    double ta = -1e300, tb = 1e300;
    if (D[2] != 0.0) {
        const double u = (z0 - P[2]) / D[2], v = (z1 - P[2]) / D[2];
        ta = fmin(u, v);
        tb = fmax(u, v);
    } else if (P[2] < z0 || P[2] > z1) {
        return false;
    }
    const double a = D[0] * D[0] + D[1] * D[1], b = P[0] * D[0] + P[1] * D[1], c = P[0] * P[0] + P[1] * P[1];
    const double tv = fmin(fmax(a > 0.0 ? -b / a : ta, ta), tb);
    const double hi = fabs(Ra) + fabs(Rs) + m;
    if ((a * tv + 2.0 * b) * tv + c > hi * hi) return false;
    const double lo = fabs(Ra) - fabs(Rs) - m;
    if (lo > 0.0 && ta > -1e300 && tb < 1e300) {
        const double ra = (a * ta + 2.0 * b) * ta + c, rb = (a * tb + 2.0 * b) * tb + c;
        if (fmax(ra, rb) < lo * lo) return false;
    }
    return true;
}

// coretrace/intersect.cpp:140-1011
inline void intersect(const double P[3], const double D[3], const Element &e, double X[3], double Cout[3], double DF[3],
                      double *path, int *err) {
    *err = 0;
    for (int i = 0; i < 3; i++) {
        X[i] = P[i];
        Cout[i] = D[i];
    }
    const char s = e.surf;
    if (e.stype == 2) {
        quadric_closed_form(e, P, D, X, DF, path, err);
        return;
    }
    if ((e.stype == 1 || e.stype == 7) && (s == 's' || s == 'S')) {
        quadric_closed_form(e, P, D, X, DF, path, err);
        return;
    }
    if (e.stype == 1 && (s == 'o' || s == 'O')) {
        quadric_closed_form(e, P, D, X, DF, path, err);
        return;
    }
    if (s == 'p' || s == 'P') {
        quadric_closed_form(e, P, D, X, DF, path, err);
        return;
    }
    if (e.stype == 10) {
        double t0 = 0.0;
        if (e.torus_start_r > 0.0) {
            const double d[3] = {P[0] - e.torus_start_c[0], P[1] - e.torus_start_c[1], P[2] - e.torus_start_c[2]};
            const double tc = -(d[0] * D[0] + d[1] * D[1] + d[2] * D[2]);
            const double dist2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double rr = 2.0 * e.torus_start_r + 1.0;
            if (dist2 > rr * rr && tc > rr) t0 = tc - rr;
        }
        const double Ps[3] = {P[0] + t0 * D[0], P[1] + t0 * D[1], P[2] + t0 * D[2]};
        if (e.torus_start_r > 0.0 && !torus_shell(e, t0 == 0.0 ? P : Ps, D)) {
            *path = 0.0;
            *err = 1;
        } else if (t0 == 0.0) {
            torus_closed_form(e, P, D, X, DF, path, err);
        } else {
            torus_closed_form(e, Ps, D, X, DF, path, err);
            *path = t0 + *path;
        }
        return;
    }

    // Newton-Raphson start plane (intersect.cpp:272-633)
    double S0Ap;
    if (e.zap - X[2] == 0.0)
        S0Ap = 0.0;
    else
        S0Ap = (e.zap - X[2]) / (Cout[2] + 0.00000000001);
    double xa = X[0] + Cout[0] * S0Ap;
    double ya = X[1] + Cout[1] * S0Ap;
    bool za_inside = aperture_contains(e, xa, ya, true);

    double ZStart = 0.0;
    if (e.stype != 3 && e.stype != 4 && e.stype != 9) {
        double FXY = 0.0;
        surface_z_at_xy(X, e, &FXY);
        if (X[2] <= 0.0 && Cout[2] > 0.0)
            ZStart = 0.0;
        else if (X[2] <= FXY && Cout[2] > 0.0)
            ZStart = 0.0;
        else if (X[2] <= FXY && Cout[2] < 0.0 && X[2] > e.zap && za_inside)
            ZStart = 0.0;
        else if (X[2] <= FXY && Cout[2] < 0.0)
            ZStart = e.zap;
        else if (X[2] > FXY && Cout[2] < 0.0)
            ZStart = 0.0;
        else if (X[2] > FXY && Cout[2] > 0.0)
            ZStart = e.zap;
    }
    if (e.stype == 9) {
        if (!spline_start_plane(e, X, Cout, P, &ZStart)) {
            *err = 3;
            return;
        }
    }

    double S0;
    if (ZStart - X[2] == 0.0)
        S0 = 0.0;
    else
        S0 = (ZStart - X[2]) / (Cout[2] + 0.00000000001);
    double X1 = X[0] + Cout[0] * S0;
    double Y1 = X[1] + Cout[1] * S0;
    double SJ1 = 0.0, SJ = 0.0;
    double lower = -1e10, upper = 1e10;
    if (e.stype == 9) {
        double s_to_xmin = (e.sx[0] - X1) / (Cout[0] + 0.00000000001);
        double s_to_xmax = (e.sx[e.sx.size() - 1] - X1) / (Cout[0] + 0.00000000001);
        lower = fmin(s_to_xmin, s_to_xmax);
        upper = fmax(s_to_xmin, s_to_xmax);
    }
    const double Eps = 0.000001;
    int it = 0;
    bool done = false;
    while (it++ < 20) {
        SJ = SJ1;
        X[0] = X1 + Cout[0] * SJ;
        X[1] = Y1 + Cout[1] * SJ;
        X[2] = ZStart + Cout[2] * SJ;
        double F;
        int ok = 0;
        surface_eval(X, e, &F, DF, &ok);
        if (ok != 0) {
            *err = 2;
            done = true;
            break;
        }
        double dfd = dot3(DF, Cout);
        if (fabs(F) <= Eps * fabs(dfd)) {
            done = true;
            break;
        }
        SJ1 = SJ - F / dfd;
        if (e.stype == 9) {
            if ((F < 0 && Cout[2] > 0) || (F > 0 && Cout[2] < 0))
                lower = fmax(SJ, lower);
            else if ((F < 0 && Cout[2] < 0) || (F > 0 && Cout[2] > 0))
                upper = fmin(SJ, upper);
            if (SJ1 < lower || SJ1 > upper) SJ1 = 0.5 * (lower + upper);
            if (upper <= lower) break;
        }
    }
    if (!done) *err = 1;
    *path = S0 + SJ;
}

// coretrace/determineelementintersectionnew.cpp:58-633
inline void element_intersection(const Element &e, const double P[3], const double D[3], double X[3], double Cout[3],
                                 double DF[3], double *path, int *err, int *hit, int *back) {
    *err = 0;
    intersect(P, D, e, X, Cout, DF, path, err);
    bool ok = !(*err > 0 || *path < 0);
    if (ok) ok = aperture_contains(e, X[0], X[1], false);
    if (!ok) {
        *hit = 0;
        for (int i = 0; i < 3; i++) X[i] = Cout[i] = DF[i] = 0.0;
        *back = 0;
        *path = 0.0;
        return;
    }
    *back = (dot3(D, DF) < 0) ? 0 : 1;
    *hit = 1;
    if (*back) {
        DF[0] = -DF[0];
        DF[1] = -DF[1];
        DF[2] = -DF[2];
    }
}

}  // namespace oracle
