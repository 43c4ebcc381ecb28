// Optics, sun sampling and scene set-up of the CPU oracle.  TEST
// INFRASTRUCTURE ONLY.  Restates coretrace/errors.cpp,
// surfacenormalerrors.cpp, interaction.cpp, generateray.cpp,
// suntoprimarystage.cpp, apertureplane.cpp and input.cpp:153-509.
#pragma once
#include <strings.h>

#include "st_oracle_geom.h"
#include "st_oracle_rng.h"

namespace oracle {

// Local ray frame shared by Errors and SurfaceNormalErrors
// (coretrace/errors.cpp:98-121, surfacenormalerrors.cpp:87-106).
inline void ray_frame(const double c[3], double RT[3][3]) {
    double e[3] = {0, 0, 0};
    if (c[2] == 0.0) {
        if (c[0] == 0.0) {
            e[0] = 0.0;
            e[1] = M_PI / 2.0;
        } else {
            e[0] = M_PI / 2.0;
            e[1] = atan2(c[1], sqrt(c[0] * c[0] + c[2] * c[2]));
        }
    } else {
        e[0] = atan2(c[0], c[2]);
        e[1] = atan2(c[1], sqrt(c[0] * c[0] + c[2] * c[2]));
    }
    e[2] = 0.0;
    double R[3][3];
    euler_matrices(e, R, RT);
}

inline void cone_direction(double theta2, double phi, const double RT[3][3], double out[3]) {
    double theta = sqrt(theta2);
    double loc[3] = {sin(theta) * cos(phi), sin(theta) * sin(phi), cos(theta)};
    mat_vec(RT, loc, out);
}

// Counter-mode sampler of the 'd' sunshape's polar angle.  The law is the
// reference's rejection loop's (coretrace/errors.cpp:160-185): (tx, ty)
// uniform on the square of half-side max_angle, kept with probability
// I(theta)/I_max inside the circle, I the linear interpolation of
// errors.cpp:167-178, so theta has density I(theta) theta on [0, max_angle].
// The counter streams draw it by inverting that law's CDF instead, from one
// 53-bit uniform: on each table segment [lo, lo + w] the density is
// (a + b x)(lo + x) in x = theta - lo and its CDF from lo the cubic
// G(x) = x (p0 + x (p1 + x p2)).  The segment holding u * total comes from a
// bisection over the masses before each segment, x from Newton steps kept in
// a shrinking bracket (bisecting when a step leaves it) until a step moves x
// by at most 1e-15 of the segment's far end.  Tables that are not ascending, or hold negative intensities,
// keep the rejection loop.
inline void sun_icdf_build(Sun &s) {
    s.icdf.clear();
    const size_t n = s.angle.size();
    bool ok = s.max_angle > 0 && s.max_intensity > 0;
    for (size_t i = 1; i < n; i++) ok = ok && s.angle[i] > s.angle[i - 1];
    for (double v : s.intensity) ok = ok && v >= 0.0;
    if (!ok) return;
    double before = 0.0;
    for (size_t i = 1; i < n; i++) {
        const double t0 = s.angle[i - 1], t1 = s.angle[i];
        const double lo = t0 > 0.0 ? t0 : 0.0;
        const double hi = t1 < s.max_angle ? t1 : s.max_angle;
        if (!(hi > lo)) continue;
        const double slope = (s.intensity[i] - s.intensity[i - 1]) / (t1 - t0);
        const double at_lo = s.intensity[i - 1] + (s.intensity[i] - s.intensity[i - 1]) * (lo - t0) / (t1 - t0);
        const double w = hi - lo;
        const double p0 = at_lo * lo, p1 = (at_lo + slope * lo) / 2.0, p2 = slope / 3.0;
        const double seg[8] = {lo, w, at_lo, slope, p0, p1, p2, before};
        s.icdf.insert(s.icdf.end(), seg, seg + 8);
        before = before + w * (p0 + w * (p1 + w * p2));
    }
    if (s.icdf.empty() || !(before > 0.0)) {
        s.icdf.clear();
        return;
    }
    s.icdf.push_back(before);
}

// theta^2 (mrad^2) from the stream's next two 32-bit words w0, w1.
inline double sun_icdf_theta2(const Sun &s, uint32_t w0, uint32_t w1) {
    const size_t nseg = s.icdf.size() / 8;
    const double *t = s.icdf.data();
    const double total = t[8 * nseg];
    const double u = ((double)w0 * 2097152.0 + (double)(w1 >> 11)) * 1.1102230246251565e-16;  // [0, 1), 2^-53 steps
    const double m = u * total;
    size_t k = 0;
    for (size_t a = 0, b = nseg - 1; a < b;) {  // the last segment with mass-before <= m
        const size_t mid = (a + b + 1) / 2;
        if (t[8 * mid + 7] <= m) a = mid;
        else b = mid - 1;
        k = a;
    }
    const double *g = t + 8 * k;
    const double lo = g[0], w = g[1], av = g[2], bv = g[3], p0 = g[4], p1 = g[5], p2 = g[6];
    double r = m - g[7];
    const double mass = w * (p0 + w * (p1 + w * p2));
    if (!(r < mass)) r = mass;
    if (!(r > 0.0)) r = 0.0;
    double xl = 0.0, xh = w;
    double x = mass > 0.0 ? w * (r / mass) : 0.0;
    const double tol = 1e-15 * (lo + w);
    for (int it = 0; it < 60; ++it) {
        const double f = x * (p0 + x * (p1 + x * p2)) - r;
        if (f > 0.0) xh = x;
        else xl = x;
        const double d = (av + bv * x) * (lo + x);
        double xn;
        if (d > 0.0) {
            xn = x - f / d;
            if (fabs(xn - x) <= tol) {
                x = xn < 0.0 ? 0.0 : (xn > w ? w : xn);
                break;
            }
        } else {
            xn = 0.5 * (xl + xh);
        }
        if (!(xn > xl && xn < xh)) xn = 0.5 * (xl + xh);
        if (!(xh > xl)) break;
        x = xn;
    }
    const double th = lo + x;
    return th * th;
}

// Sunshape (source 1) or specularity (source 2) perturbation,
// coretrace/errors.cpp:60-226.
inline void errors(Rng &rng, const double cin[3], int source, const Sun &sun, const Element &el, const OpticFace *opt,
                   double cout[3], const double DF[3]) {
    double RT[3][3];
    ray_frame(cin, RT);
    char dist = 'g';
    double delop = 0;
    if (source == 1) {
        dist = sun.shape;
        delop = sun.sigma / 1000.0;
    }
    if (source == 2) {
        dist = opt->dist;
        delop = opt->spec / 1000.0;
    }
    unsigned int maxcall = 0;
    double theta2 = 0.0;
    for (;;) {
        switch (dist) {
        case 'g': case 'G': {
            double tx = rng.normal(0., delop);
            double ty = rng.normal(0., delop);
            theta2 = tx * tx + ty * ty;
            break;
        }
        case 'p': case 'P': {
            double tx, ty;
            do {
                tx = 2.0 * delop * rng() - delop;
                ty = 2.0 * delop * rng() - delop;
                theta2 = tx * tx + ty * ty;
            } while (theta2 > (delop * delop));
            break;
        }
        case 'd': case 'D': {
            ST_OPS_PUSH(3);
            if (rng.mode() == 1 && !sun.icdf.empty()) {
                const uint32_t w0 = (uint32_t)(rng() * 4294967296.0);
                const uint32_t w1 = (uint32_t)(rng() * 4294967296.0);
                theta2 = sun_icdf_theta2(sun, w0, w1) / 1000000.0;
                ST_OPS_POP();
                break;
            }
            const std::vector<double> &ang = sun.angle;
            const std::vector<double> &inten = sun.intensity;
            for (;;) {
                double tx = 2.0 * sun.max_angle * rng() - sun.max_angle;
                double ty = 2.0 * sun.max_angle * rng() - sun.max_angle;
                theta2 = tx * tx + ty * ty;
                double theta = sqrt(theta2);
                size_t i = 0;
                while (i < ang.size() - 1 && ang[i] < theta) i++;
                double stest;
                if (i == 0)
                    stest = inten[0];
                else
                    stest = inten[i - 1] + (inten[i] - inten[i - 1]) * (theta - ang[i - 1]) / (ang[i] - ang[i - 1]);
                if (rng() > (stest / sun.max_intensity)) continue;
                if (theta2 > (sun.max_angle * sun.max_angle)) continue;
                break;
            }
            theta2 = theta2 / 1000000.0;
            ST_OPS_POP();
            break;
        }
        case 'f': case 'F':
            theta2 = pow(asin(sqrt(rng())), 2);
            break;
        default:
            break;
        }
        double phi = rng() * 2.0 * 3.1415926535897932385;
        cone_direction(theta2, phi, RT, cout);
        if ((source == 2) && (el.interaction == 2) && (dot3(cout, DF) < 0) && maxcall++ < 50000) continue;
        break;
    }
}

// Slope error applied to the surface normal,
// coretrace/surfacenormalerrors.cpp:58-167.
inline void surface_normal_errors(Rng &rng, const double cin[3], const OpticFace *opt, double cout[3]) {
    double RT[3][3];
    ray_frame(cin, RT);
    char dist = opt->dist;
    double delop = opt->slope / 1000.0;
    double theta2 = 0.0;
    switch (dist) {
    case 'g': case 'G': {
        double tx = rng.normal(0., delop);
        double ty = rng.normal(0., delop);
        theta2 = tx * tx + ty * ty;
        break;
    }
    case 'p': case 'P': {
        double tx, ty;
        do {
            tx = 2.0 * delop * rng() - delop;
            ty = 2.0 * delop * rng() - delop;
            theta2 = tx * tx + ty * ty;
        } while (theta2 > (delop * delop));
        break;
    }
    }
    double phi = rng() * 2.0 * 3.1415926535897932385;
    cone_direction(theta2, phi, RT, cout);
}

// Reflection (type 2) and refraction with Fresnel/TIR (type 1),
// coretrace/interaction.cpp:58-201.  Returns the error flag (12 when the
// refraction Newton loop does not converge; cout is then left untouched).
inline int interaction(Rng &rng, const double cin[3], const double DF[3], int type, const OpticFace *opt,
                       double cout[3]) {
    if (type == 1) {
        double n1 = opt->refr[0], n2 = opt->refr[2];
        double mu = n1 / n2;
        double D2 = dot3(DF, DF);
        double B = (mu * mu - 1.0) / D2;
        double A = mu * dot3(cin, DF) / D2;
        double A2 = A * A;
        if (B > A2) {
            A = dot3(cin, DF) / dot3(DF, DF);
            for (int i = 0; i < 3; i++) cout[i] = cin[i] - 2.0 * A * DF[i];
            return 0;
        }
        double u[3];
        u[0] = -DF[0] / sqrt(dot3(DF, DF));
        u[1] = -DF[1] / sqrt(dot3(DF, DF));
        u[2] = -DF[2] / sqrt(dot3(DF, DF));
        double ia = acos(dot3(cin, u));
        double st = n1 * sin(ia) / n2;
        double ct = sqrt(1 - st * st);
        double rs_n = (n1 * cos(ia) - n2 * ct);
        double rs_d = (n1 * cos(ia) + n2 * ct);
        double Rs = (rs_n / rs_d) * (rs_n / rs_d);
        double rp_n = (n1 * ct) - n2 * cos(ia);
        double rp_d = (n1 * ct) + n2 * cos(ia);
        double Rp = (rp_n / rp_d) * (rp_n / rp_d);
        double Rave = (Rp + Rs) / 2.0;
        if (Rave < rng()) {
            double g = -B / (2.0 * A), g1 = 0.0;
            for (int i = 1; i < 10; i++) {
                g1 = (g * g - B) / (2.0 * (g + A));
                if (fabs(g - g1) < 0.000005) {
                    for (int j = 0; j < 3; j++) cout[j] = mu * cin[j] + g1 * DF[j];
                    return 0;
                }
                g = g1;
            }
            return 12;
        }
        A = dot3(cin, DF) / dot3(DF, DF);
        for (int i = 0; i < 3; i++) cout[i] = cin[i] - 2.0 * A * DF[i];
        return 0;
    }
    if (type == 2) {
        double A = dot3(cin, DF) / dot3(DF, DF);
        for (int i = 0; i < 3; i++) cout[i] = cin[i] - 2.0 * A * DF[i];
    }
    return 0;
}

// ------------------------------------------------------- scene set-up ------
// coretrace/apertureplane.cpp:57-491
inline void aperture_plane(Element &e) {
    auto sq = [](double v) { return (v) * (v); };
    const double A = e.p[0], Bp = e.p[1];
    double maxr = 0.0;
    if (e.shape == 'i' || e.shape == 'I' || e.shape == 'q' || e.shape == 'Q') {
        maxr = sqrt(sq(e.p[0]) + sq(e.p[1]));
        double t = sqrt(sq(e.p[2]) + sq(e.p[3]));
        if (t > maxr) maxr = t;
        t = sqrt(sq(e.p[4]) + sq(e.p[5]));
        if (t > maxr) maxr = t;
        if (e.shape == 'q' || e.shape == 'Q') {
            t = sqrt(sq(e.p[6]) + sq(e.p[7]));
            if (t > maxr) maxr = t;
        }
    }
    const char sh = (char)tolower(e.shape);
    const bool cht = (sh == 'c' || sh == 'h' || sh == 't');
    const bool al = (sh == 'a' || sh == 'l');
    const bool iq = (sh == 'i' || sh == 'q');
    auto r2_al = [&]() { return (fabs(Bp) >= fabs(A)) ? sq(Bp) : sq(A); };
    switch (e.surf) {
    case 'g': case 'G': {
        double radius2 = 0, sum = 0;
        if (cht) {
            radius2 = sq(A / 2.0);
            double rho2i = 1.0;
            for (int i = 0; i < 5; i++) {
                rho2i = rho2i * radius2;
                sum += e.alpha[i] * rho2i;
            }
            e.zap = e.vcx * radius2 / (1 + sqrt(1 - e.kappa * e.vcx * e.vcx * radius2)) + sum;
        } else if (sh == 'r') {
            radius2 = sq(A / 2.0) + sq(Bp / 2.0);
            for (int i = 0; i < 5; i++) sum += e.alpha[i] * pow(radius2, (int)(i + 1));
            e.zap = e.vcx * radius2 / (1 + sqrt(1 - e.kappa * sq(e.vcx) * radius2)) + sum;
        } else if (al) {
            radius2 = r2_al();
            for (int i = 0; i < 5; i++) sum = sum + e.alpha[i] * pow(radius2, (int)(i + 1));
            e.zap = e.vcx * radius2 / (1 + sqrt(1 - e.kappa * sq(e.vcx) * radius2)) + sum;
        } else if (iq) {
            radius2 = sq(maxr);
            e.zap = e.vcx * radius2 / (1 + sqrt(1 - e.kappa * sq(e.vcx) * radius2)) + 0.0;
        }
        break;
    }
    case 's': case 'S': {
        double F = 1.0 / (2.0 * e.vcx);
        if (cht)
            e.zap = 2.0 * F - sqrt(sq(2.0 * F) - sq(A / 2.0));
        else if (sh == 'r')
            e.zap = 2.0 * F - sqrt(sq(2.0 * F) - (sq(A / 2.0) + sq(Bp / 2.0)));
        else if (al)
            e.zap = 2.0 * F - sqrt(sq(2.0 * F) - r2_al());
        else if (iq)
            e.zap = 2.0 * F - sqrt(sq(2.0 * F) - sq(maxr));
        break;
    }
    case 'p': case 'P': {
        double F = 1.0 / (2.0 * e.vcx);
        if (cht)
            e.zap = sq(A / 2.0) / (4.0 * F);
        else if (sh == 'r')
            e.zap = (sq(A / 2.0) + sq(Bp / 2.0)) / (4.0 * F);
        else if (al)
            e.zap = r2_al() / (4.0 * F);
        else if (iq)
            e.zap = sq(maxr) / (4.0 * F);
        break;
    }
    case 'f': case 'F':
        e.zap = 0.0;
        break;
    case 'o': case 'O': {
        if (cht)
            e.zap = (1 - sqrt(1 - e.kappa * sq(e.vcx) * sq(A / 2.0))) / (e.kappa * e.vcx);
        else if (sh == 'r')
            e.zap = (1 - sqrt(1 - e.kappa * sq(e.vcx) * (sq(A / 2.0) + sq(Bp / 2.0)))) / (e.kappa * e.vcx);
        else if (al)
            e.zap = (1 - sqrt(1 - e.kappa * sq(e.vcx) * r2_al())) / (e.kappa * e.vcx);
        else if (iq)
            e.zap = (1 - sqrt(1 - e.kappa * sq(e.vcx) * sq(maxr))) / (e.kappa * e.vcx);
        break;
    }
    case 'c': case 'C': {
        double t = tan(e.cone_half * (kDegToRad));
        if (cht)
            e.zap = A / (2.0 * t);
        else if (sh == 'r')
            e.zap = sqrt(sq(A / 2.0) + sq(Bp / 2.0)) / (2.0 * t);
        else if (sh == 'a')
            e.zap = Bp / (t);
        else if (iq)
            e.zap = sqrt(sq(maxr)) / (2.0 * t);
        break;
    }
    case 't': case 'T':
        e.zap = 2.0 / e.curv_rev;
        break;
    case 'd': case 'D':
        e.zap = 2.0 * e.cross_r;
        break;
    case 'm': case 'M': {
        if (!(e.bc_n > 2)) return;
        double F = (1.0 / (4.0 * e.B(2, 0)) + 1.0 / (4.0 * e.B(2, 2))) / 2.0;
        if (cht)
            e.zap = sq(A / 2.0) / (4.0 * F);
        else if (sh == 'r')
            e.zap = (sq(A / 2.0) + sq(Bp / 2.0)) / (4.0 * F);
        else if (al)
            e.zap = r2_al() / (4.0 * F);
        else if (iq)
            e.zap = sq(maxr) / (4.0 * F);
        break;
    }
    case 'r': case 'R': {
        if (cht)
            e.zap = eval_poly(A / 2.0, 0, e);
        else if (sh == 'r')
            e.zap = eval_poly(sqrt(sq(A / 2.0) + sq(Bp / 2.0)), 0, e);
        else if (al)
            e.zap = eval_poly(sqrt(r2_al()), 0, e);
        else if (iq)
            e.zap = eval_poly(sqrt(sq(maxr)), 0, e);
        break;
    }
    case 'i': case 'I':
        e.zap = e.sy[e.sy.size() - 1];
        break;
    case 'e': case 'E': {
        // the largest node *y* (column 1), at least 0, as the reference has it
        // (coretrace/apertureplane.cpp:458-471)
        double lastz = 0.0;
        for (size_t i = 0; i < e.fe_nodes.size() / 3; i++) {
            double newz = e.fe_nodes[i * 3 + 1];
            if (newz > lastz) lastz = newz;
        }
        e.zap = lastz;
        break;
    }
    case 'v': case 'V': {
        // the largest fitted z over the data points, at least 0
        // (coretrace/apertureplane.cpp:474-488)
        double lastz = 0.0;
        for (size_t i = 0; i < e.vshot.size() / 5; i++) {
            double zm = eval_mono(e.vshot[i * 5 + 0], e.vshot[i * 5 + 1], e);
            if (zm > lastz) lastz = zm;
        }
        e.zap = lastz;
        break;
    }
    default:
        break;
    }
}

// coretrace/input.cpp:288-366
inline bool translate_surface_params(System &sys, Element &e, const double p[8]) {
    switch (e.surf) {
    case 's': case 'S':
        e.vcx = p[0];
        e.vcy = (e.shape == 'l' || e.shape == 'L') ? 0.0 : e.vcx;
        e.kappa = 1.0;
        e.stype = (e.vcy == 0.0) ? 7 : 1;
        return true;
    case 'p': case 'P':
        e.vcx = p[0];
        e.vcy = p[1];
        e.kappa = 0.0;
        e.stype = (e.vcy == 0.0) ? 7 : 1;
        return true;
    case 'o': case 'O':
        e.vcx = p[0];
        e.vcy = p[0];
        e.kappa = p[1];
        e.stype = (e.vcy == 0.0) ? 7 : 1;
        return true;
    case 'g': case 'G':
        e.vcx = p[0];
        e.vcy = p[1];
        e.kappa = p[2];
        for (int i = 0; i < 5; i++) e.alpha[i] = p[3 + i];
        e.stype = (e.vcy == 0.0) ? 7 : 1;
        return true;
    case 'f': case 'F':
        e.alpha[0] = 0; e.alpha[1] = 0; e.alpha[2] = 1; e.alpha[3] = 0; e.alpha[4] = 0;
        e.stype = 3;
        return true;
    case 'c': case 'C':
        e.cone_half = p[0];
        e.stype = 1;
        return true;
    case 't': case 'T':
        if ((e.shape != 'l' && e.shape != 'L') || e.p[0] != 0.0 || e.p[1] != 0.0) {
            sys.errlog("Bad aperture or parameters of element not consistent with cylindrical surface\n");
            return false;
        }
        e.curv_rev = p[0];
        e.vcx = e.curv_rev;
        e.stype = 2;
        return true;
    case 'd': case 'D':
        if ((e.shape != 'a' && e.shape != 'A') || e.p[0] != 0.0 || e.p[1] != 0.0) {
            sys.errlog("Bad aperture or parameters of element not consistent with toroidal surface\n");
            return false;
        }
        e.annular_r = p[0];
        e.cross_r = p[1];
        e.stype = 10;
        return true;
    default:
        sys.errlog("invalid surface index: %c\n", e.surf);
        return false;
    }
}

// Surface files .mon (Zernike), .sht (VSHOT), .csi (cubic spline), .ply
// (polynomial) and .fed (finite element), coretrace/input.cpp:368-561.
inline bool read_surface_file(const char *file, Element &e, System &sys) {
    const char *dot = strrchr(file, '.');
    if (!dot) {
        sys.errlog("Could not determine surface file type: '%s'\n", file);
        return false;
    }
    std::string ext(dot + 1);
    for (auto &ch : ext) ch = (char)tolower(ch);
    FILE *fp = fopen(file, "r");
    if (!fp) {
        sys.errlog("Could not open surface file for read: %s\n", file);
        return false;
    }
    char line[1024];
    auto readln = [&]() {
        line[0] = 0;
        if (!fgets(line, sizeof line, fp)) line[0] = 0;
        size_t n = strlen(line);
        if (n > 0 && line[n - 1] == '\n') line[n - 1] = 0;
    };
    if (ext == "mon") {
        readln();
        int order = atoi(line);
        e.bc_n = order + 1;
        e.bc.assign((size_t)e.bc_n * e.bc_n, 0.0);
        for (int k = 0; k <= order; k++)
            for (int m = 0; m <= k; m++) {
                readln();
                e.bc[(size_t)k * e.bc_n + m] = atof(line);
            }
        e.fit_order = order;
        e.surf = 'm';
        e.stype = 6;
    } else if (ext == "sht") {
        // coretrace/input.cpp:415-452: title, radius/focal length/target
        // distance, (idum, order, points), RMS slope/scale, the lower
        // triangle of B, then five columns per point
        readln();
        readln();
        int idum = 0, order = 0, npts = 0;
        readln();
        sscanf(line, "%d %d %d", &idum, &order, &npts);
        readln();
        e.fit_order = order;
        e.bc_n = order + 1;
        e.bc.assign((size_t)e.bc_n * e.bc_n, 0.0);
        for (int k = 0; k <= order; k++)
            for (int m = 0; m <= k; m++) {
                readln();
                e.bc[(size_t)k * e.bc_n + m] = atof(line);
            }
        e.vshot.assign((size_t)npts * 5, 0.0);
        for (int i = 0; i < npts; i++) {
            double v[5] = {0, 0, 0, 0, 0};
            readln();
            sscanf(line, "%lg %lg %lg %lg %lg", &v[0], &v[1], &v[2], &v[3], &v[4]);
            for (int c = 0; c < 5; c++) e.vshot[(size_t)i * 5 + c] = v[c];
        }
        e.surf = 'v';
        e.stype = 5;
    } else if (ext == "fed") {
        // coretrace/input.cpp:510-561: title, node count, x y z per node,
        // then the spatial hash tree over the nodes
        readln();
        readln();
        int n = atoi(line);
        e.fe_nodes.assign((size_t)std::max(n, 0) * 3, 0.0);
        for (int i = 0; i < n; i++) {
            double x = 0, y = 0, z = 0;
            readln();
            sscanf(line, "%lg %lg %lg", &x, &y, &z);
            e.fe_nodes[3 * (size_t)i] = x;
            e.fe_nodes[3 * (size_t)i + 1] = y;
            e.fe_nodes[3 * (size_t)i + 2] = z;
        }
        st_fe::build(e.fe_tree, e.fe_nodes);
        e.surf = 'e';
        e.stype = 4;
    } else if (ext == "ply") {
        readln();
        int order = atoi(line);
        e.poly.assign(order + 1, 0.0);
        for (int k = 0; k <= order; k++) {
            readln();
            e.poly[k] = atof(line);
        }
        e.fit_order = order;
        e.surf = 'r';
        e.stype = 8;
    } else if (ext == "csi") {
        readln();
        int n = atoi(line);
        e.sx.assign(n, 0.0);
        e.sy.assign(n, 0.0);
        e.sy2.assign(n, 0.0);
        double x = 0, y = 0;
        for (int k = 0; k < n; k++) {
            readln();
            sscanf(line, "%lg %lg", &x, &y);
            e.sx[k] = x;
            e.sy[k] = y;
        }
        readln();
        sscanf(line, "%lg %lg", &x, &y);
        spline(e.sx, e.sy, n, x, y, e.sy2);
        if (((e.shape != 'a') && (e.shape != 'A') && (e.shape != 'l') && (e.shape != 'L')) || (e.p[0] < e.sx[0]) ||
            (e.p[1] > e.sx[e.sx.size() - 1])) {
            sys.errlog("Error: Element uses cubic spline interpolation: make sure aperture is type 'a' or 'l'.\n");
            fclose(fp);
            return false;
        }
        e.surf = 'i';
        e.stype = 9;
    } else {
        fclose(fp);
        sys.errlog("Surface file type extension unknown: '%s'\n", ext.c_str());
        return false;
    }
    fclose(fp);
    return true;
}

// coretrace/input.cpp:87-150
inline bool check_inputs(System &sys) {
    for (size_t i = 0; i < sys.stages.size(); i++) {
        for (size_t j = 0; j < sys.stages[i]->elements.size(); j++) {
            Element &e = *sys.stages[i]->elements[j];
            char s = e.surf, a = e.shape;
            if ((s == 'p' || s == 'P') && (a == 'l' || a == 'L') && e.vcy != 0.0) {
                sys.errlog("Parabolic surface combined with single axis curvature section aperture must have Cy = 0.");
                return false;
            }
            if (s == 's' || s == 'S') {
                if (e.vcx == 0.0) {
                    sys.errlog("Spherical surface must have a curvature greater than 0.0.");
                    return false;
                }
                if (e.vcy != 0.0 && e.vcy != e.vcx) {
                    sys.errlog("Spherical surface must have Cy=0 or Cy=Cx.");
                    return false;
                }
            }
            if ((s == 'o' || s == 'O') && (e.kappa == 1.0 || e.kappa == 0.0)) {
                sys.errlog("Kappa = 0.0 or 1.0 defines either a parabola or a sphere.");
                return false;
            }
        }
    }
    return true;
}

// coretrace/input.cpp:153-286
inline bool init_geometries(System &sys) {
    auto sq = [](double v) { return (v) * (v); };
    for (size_t i = 0; i < sys.stages.size(); i++) {
        Stage &st = *sys.stages[i];
        if (!euler_from_aim(st.origin, st.aim, st.zrot, st.euler)) {
            sys.errlog("Stage %d geometry as defined not possible.\n", (int)i);
            return false;
        }
        euler_matrices(st.euler, st.R, st.RT);
        for (size_t j = 0; j < st.elements.size(); j++) {
            Element &e = *st.elements[j];
            if (!euler_from_aim(e.origin, e.aim, e.zrot, e.euler)) {
                sys.errlog("Stage %d element %d geometry as defined not possible.\n", (int)i, (int)j);
                return false;
            }
            euler_matrices(e.euler, e.R, e.RT);
            const double *P = e.p;
            switch (e.shape) {
            case 'c': case 'C': e.area = M_PI * sq(P[0] / 2); break;
            case 'h': case 'H':
                e.area = 5.0 * sq(P[0] / 2.0) * cos(30.0 * (kDegToRad)) * sin(30.0 * (kDegToRad));
                break;
            case 'r': case 'R': e.area = P[0] * P[1]; break;
            case 't': case 'T': e.area = 0.25 * sq(P[0] * cos(30.0 * (kDegToRad))) * sqrt(3.0); break;
            case 'a': case 'A': e.area = P[2] * (kDegToRad) * (P[1] - P[0]); break;
            case 'l': case 'L': e.area = (P[1] - P[0]) * P[2]; break;
            case 'i': case 'I': {
                double v11 = P[0] - P[2], v12 = P[1] - P[3], v21 = P[4] - P[2], v22 = P[5] - P[3];
                double v1m = sqrt(v11 * v11 + v12 * v12), v2m = sqrt(v21 * v21 + v22 * v22);
                double th = acos((v11 * v21 + v12 * v22) / (v1m * v2m));
                e.area = 0.5 * v1m * v2m * sin(th);
                break;
            }
            case 'q': case 'Q': {
                double v11 = P[0] - P[2], v12 = P[1] - P[3], v21 = P[4] - P[2], v22 = P[5] - P[3];
                double v31 = P[4] - P[6], v32 = P[5] - P[7], v41 = P[0] - P[6], v42 = P[1] - P[7];
                double v1m = sqrt(v11 * v11 + v12 * v12), v2m = sqrt(v21 * v21 + v22 * v22);
                double v3m = sqrt(v31 * v31 + v32 * v32), v4m = sqrt(v41 * v41 + v42 * v42);
                double t1 = acos((v11 * v21 + v12 * v22) / (v1m * v2m));
                double t2 = acos((v31 * v41 + v32 * v42) / (v3m * v4m));
                e.area = 0.5 * (v1m * v2m * sin(t1) + v3m * v4m * sin(t2));
                break;
            }
            default:
                sys.errlog("invalid shape '%c' for aperture, element %d\n", e.shape, (int)j);
                return false;
            }
            if (!st.virt) {
                OpticSet *set = nullptr;
                for (size_t k = 0; k < sys.optics.size(); k++)
                    if (strcasecmp(sys.optics[k]->name.c_str(), e.optic_name.c_str()) == 0) set = sys.optics[k];
                if (!set) {
                    sys.errlog("Stage %d element %d invalid optical property set name: '%s'\n", (int)i, (int)j,
                               e.optic_name.c_str());
                    return false;
                }
                e.optics = set;
                set->front.refr[2] = set->back.refr[0];
                set->front.refr[3] = set->back.refr[1];
                set->back.refr[2] = set->front.refr[0];
                set->back.refr[3] = set->front.refr[1];
            } else {
                e.optics = nullptr;
            }
            aperture_plane(e);
        }
    }
    return check_inputs(sys);
}

// Sun frame and sun-plane box of the first stage,
// coretrace/suntoprimarystage.cpp:59-293.
inline bool sun_to_primary_stage(System &sys, Stage &st, double pos_sun_stage[3]) {
    Sun &sun = sys.sun;
    double dx = 0.0 - sun.origin[0], dy = 0.0 - sun.origin[1], dz = 0.0 - sun.origin[2];
    double dtot = sqrt(dx * dx + dy * dy + dz * dz);
    if (dtot == 0.0) {
        sys.errlog("error calculating sun position in primary stage, dtot = 0.0\n");
        return false;
    }
    double cg[3] = {dx / dtot, dy / dtot, dz / dtot};
    double cs[3];
    to_local(st.origin, cg, st.origin, st.R, pos_sun_stage, cs);
    sun.euler[0] = atan2(cs[0], cs[2]);
    sun.euler[1] = asin(cs[1]);
    sun.euler[2] = 0.0;
    double R[3][3];
    euler_matrices(sun.euler, R, sun.RT);
    const double zero[3] = {0, 0, 0}, cdum[3] = {0, 0, 1};
    sun.maxrad = 0.0;
    sun.xcm = 0.0;
    sun.ycm = 0.0;
    sun.maxx = -1.0e20;
    sun.minx = 1.0e20;
    sun.maxy = -1.0e20;
    sun.miny = 1.0e20;
    double xs = 0.0, ys = 0.0;
    for (Element *e : st.elements) {
        if (!e->enabled) continue;
        double pl[3], cl[3];
        to_local(e->origin, cdum, zero, R, pl, cl);
        xs = xs + pl[0];
        ys = ys + pl[1];
    }
    sun.xcm = xs / st.elements.size();
    sun.ycm = ys / st.elements.size();
    size_t n = 0;
    for (Element *ep : st.elements) {
        Element &e = *ep;
        if (!e.enabled) continue;
        double pl[3], cl[3];
        to_local(e.origin, cdum, zero, R, pl, cl);
        double x = pl[0] - sun.xcm, y = pl[1] - sun.ycm;
        double radius = sqrt(x * x + y * y);
        double xmin = pl[0], xmax = pl[0], ymin = pl[1], ymax = pl[1];
        e.pos_sun[0] = pl[0];
        e.pos_sun[1] = pl[1];
        e.pos_sun[2] = pl[2];
        double pad = 0.0;
        bool padded = true;
        const double A = e.p[0], Bp = e.p[1], C = e.p[2];
        switch (e.shape) {
        case 'c': case 'C': case 'h': case 'H': pad = A / 2.0; break;
        case 'r': case 'R': pad = sqrt(((A * A) + (Bp * Bp)) / 4.0); break;
        case 't': case 'T': pad = A / (2.0 * cos(30.0 * (kDegToRad))); break;
        case 'a': case 'A': pad = Bp; break;
        case 'l': case 'L': {
            double xl = (fabs(Bp) >= fabs(A)) ? Bp : A;
            pad = sqrt((xl * xl) + 0.25 * C * C);
            break;
        }
        case 'i': case 'I': case 'q': case 'Q': {
            double r1 = sqrt((e.p[0] * e.p[0]) + (e.p[1] * e.p[1]));
            double r2 = sqrt((e.p[2] * e.p[2]) + (e.p[3] * e.p[3]));
            double r3 = sqrt((e.p[4] * e.p[4]) + (e.p[5] * e.p[5]));
            double rt = r1;
            if (r2 > rt) rt = r2;
            if (r3 > rt) rt = r3;
            if (e.shape == 'q' || e.shape == 'Q') {
                double r4 = sqrt((e.p[6] * e.p[6]) + (e.p[7] * e.p[7]));
                if (r4 > rt) rt = r4;
            }
            pad = rt;
            break;
        }
        default: padded = false; break;
        }
        if (padded) {
            radius = radius + pad;
            xmin = xmin - pad;
            ymin = ymin - pad;
            xmax = xmax + pad;
            ymax = ymax + pad;
        }
        if (radius > sun.maxrad) sun.maxrad = radius;
        if (xmin < sun.minx) sun.minx = xmin;
        if (xmax > sun.maxx) sun.maxx = xmax;
        if (ymin < sun.miny) sun.miny = ymin;
        if (ymax > sun.maxy) sun.maxy = ymax;
        n++;
    }
    if (n == 0) sys.errlog("error calculating sun position in primary stage because no elements were enabled");
    return n > 0;
}

// coretrace/generateray.cpp:58-161
inline void generate_ray(Rng &rng, const double pos_sun_stage[3], const Stage &st, const Sun &sun, double pg[3],
                         double cg[3], double psun[3]) {
    psun[0] = psun[1] = psun[2] = 0.0;
    if (sun.point_source) {
        pg[0] = sun.origin[0];
        pg[1] = sun.origin[1];
        pg[2] = sun.origin[2];
        for (int i = 0; i < 3; i++) {
            int s = (rng() <= 0.5) ? -1 : 1;
            cg[i] = s * rng();
        }
        double m = sqrt(cg[0] * cg[0] + cg[1] * cg[1] + cg[2] * cg[2]);
        cg[0] = cg[0] / m;
        cg[1] = cg[1] / m;
        cg[2] = cg[2] / m;
        return;
    }
    double x = sun.minx + (sun.maxx - sun.minx) * rng();
    double y = sun.miny + (sun.maxy - sun.miny) * rng();
    psun[0] = x;
    psun[1] = y;
    psun[2] = -10000.0;
    double csun[3] = {0.0, 0.0, 1.0};
    double ps[3], cs[3];
    to_ref(psun, csun, pos_sun_stage, sun.RT, ps, cs);
    to_ref(ps, cs, st.origin, st.RT, pg, cg);
}

}  // namespace oracle
