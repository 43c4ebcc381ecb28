// Scene compiler of the HIP backend (see st_scene.h).
#include "st_scene.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <strings.h>

namespace st {

static const double kDegRad = 0.017453292519943295;  // coretrace/types.h:64

// coretrace/mathproc.cpp:175-217
void euler_matrix(const double e[3], double R[9]) {
    double ca = cos(e[0]), cb = cos(e[1]), cg = cos(e[2]);
    double sa = sin(e[0]), sb = sin(e[1]), sg = sin(e[2]);
    R[0] = ca * cg + sa * sb * sg;
    R[1] = -cb * sg;
    R[2] = -sa * cg + ca * sb * sg;
    R[3] = ca * sg - sa * sb * cg;
    R[4] = cb * cg;
    R[5] = -sa * sg - ca * sb * cg;
    R[6] = sa * cb;
    R[7] = sb;
    R[8] = ca * cb;
}

// coretrace/input.cpp:158-176
bool euler_from_aim(const double org[3], const double aim[3], double zrot, double e[3]) {
    double dx = aim[0] - org[0], dy = aim[1] - org[1], dz = aim[2] - org[2];
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d == 0.0) return false;
    dx = dx / d;
    dy = dy / d;
    dz = dz / d;
    e[0] = atan2(dx, dz);
    e[1] = asin(dy);
    e[2] = zrot * kDegRad;
    return true;
}

static void mv(const double *M, const double v[3], double o[3]) {
    o[0] = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
    o[1] = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
    o[2] = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
}
static void mtv(const double *M, const double v[3], double o[3]) {
    o[0] = M[0] * v[0] + M[3] * v[1] + M[6] * v[2];
    o[1] = M[1] * v[0] + M[4] * v[1] + M[7] * v[2];
    o[2] = M[2] * v[0] + M[5] * v[1] + M[8] * v[2];
}

// ------------------------------------------------- parameter translation --
bool translate_surface_params(HElement &e, const double p[8], std::string &msg) {
    const bool line = (e.shape == 'l' || e.shape == 'L');
    switch (e.surf) {
    case 's': case 'S':
        e.vcx = p[0];
        e.vcy = line ? 0.0 : e.vcx;
        e.kappa = 1.0;
        e.stype = (e.vcy == 0.0) ? 7 : 1;
        return true;
    case 'p': case 'P':
        e.vcx = p[0]; e.vcy = p[1]; e.kappa = 0.0;
        e.stype = (e.vcy == 0.0) ? 7 : 1;
        return true;
    case 'o': case 'O':
        e.vcx = p[0]; e.vcy = p[0]; e.kappa = p[1];
        e.stype = (e.vcy == 0.0) ? 7 : 1;
        return true;
    case 'g': case 'G':
        e.vcx = p[0]; e.vcy = p[1]; e.kappa = p[2];
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
        if (!line || e.p[0] != 0.0 || e.p[1] != 0.0) {
            msg = "Bad aperture or parameters of element not consistent with cylindrical surface\n";
            return false;
        }
        e.curv_rev = p[0];
        e.vcx = e.curv_rev;
        e.stype = 2;
        return true;
    case 'd': case 'D':
        if ((e.shape != 'a' && e.shape != 'A') || e.p[0] != 0.0 || e.p[1] != 0.0) {
            msg = "Bad aperture or parameters of element not consistent with toroidal surface\n";
            return false;
        }
        e.ann_r = p[0];
        e.cross_r = p[1];
        e.stype = 10;
        return true;
    default: {
        char b[64];
        snprintf(b, sizeof b, "invalid surface index: %c\n", e.surf);
        msg = b;
        return false;
    }
    }
}

// Numerical-Recipes natural/clamped cubic spline second derivatives
// (coretrace/mathproc.cpp:278-330).
static void spline_y2(const std::vector<double> &x, const std::vector<double> &y, double yp1, double ypn,
                      std::vector<double> &y2) {
    const int n = (int)x.size();
    std::vector<double> u(n);
    y2.assign(n, 0.0);
    if (yp1 > 0.99e30) {
        y2[0] = 0.0;
        u[0] = 0.0;
    } else {
        y2[0] = -0.5;
        u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
    }
    for (int i = 1; i < n - 1; i++) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    double qn = 0.0, un = 0.0;
    if (!(ypn > 0.99e30)) {
        qn = 0.5;
        un = (3.0 / (x[n - 1] - x[n - 2])) * (ypn - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]));
    }
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}

bool read_surface_file(HElement &e, const char *file, std::string &msg) {
    const char *dot = strrchr(file, '.');
    char b[600];
    if (!dot) {
        snprintf(b, sizeof b, "Could not determine surface file type: '%s'\n", file);
        msg = b;
        return false;
    }
    std::string ext(dot + 1);
    for (auto &c : ext) c = (char)tolower(c);
    FILE *fp = fopen(file, "r");
    if (!fp) {
        snprintf(b, sizeof b, "Could not open surface file for read: %s\n", file);
        msg = b;
        return false;
    }
    char line[1024];
    auto next = [&]() {
        line[0] = 0;
        if (!fgets(line, sizeof line, fp)) line[0] = 0;
    };
    bool ok = true;
    if (ext == "mon") {
        next();
        int order = atoi(line);
        e.fit_order = order;
        e.zern.assign((size_t)(order + 1) * (order + 1), 0.0);
        for (int k = 0; k <= order; k++)
            for (int m = 0; m <= k; m++) {
                next();
                e.zern[(size_t)k * (order + 1) + m] = atof(line);
            }
        e.surf = 'm';
        e.stype = 6;
    } else if (ext == "sht") {
        // coretrace/input.cpp:415-452: title; radius, focal length, target
        // distance; (unused, order, points); RMS slope and scale; the lower
        // triangle of B; five columns per point (the fifth unused)
        next();
        next();
        int idum = 0, order = 0, npts = 0;
        next();
        sscanf(line, "%d %d %d", &idum, &order, &npts);
        next();
        e.fit_order = order;
        e.zern.assign((size_t)(order + 1) * (order + 1), 0.0);
        for (int k = 0; k <= order; k++)
            for (int m = 0; m <= k; m++) {
                next();
                e.zern[(size_t)k * (order + 1) + m] = atof(line);
            }
        e.vshot.assign((size_t)std::max(npts, 0) * 4, 0.0);
        for (int i = 0; i < npts; i++) {
            double v[5] = {0, 0, 0, 0, 0};
            next();
            sscanf(line, "%lg %lg %lg %lg %lg", &v[0], &v[1], &v[2], &v[3], &v[4]);
            for (int c = 0; c < 4; c++) e.vshot[(size_t)i * 4 + c] = v[c];
        }
        e.surf = 'v';
        e.stype = 5;
    } else if (ext == "fed") {
        // coretrace/input.cpp:510-561: title, node count, x y z per node,
        // then the spatial hash tree over the nodes
        next();
        next();
        const int n = atoi(line);
        e.fe_nodes.assign((size_t)std::max(n, 0) * 3, 0.0);
        for (int i = 0; i < n; i++) {
            double x = 0, y = 0, z = 0;
            next();
            sscanf(line, "%lg %lg %lg", &x, &y, &z);
            e.fe_nodes[3 * (size_t)i] = x;
            e.fe_nodes[3 * (size_t)i + 1] = y;
            e.fe_nodes[3 * (size_t)i + 2] = z;
        }
        st_fe::build(e.fe_tree, e.fe_nodes);
        e.surf = 'e';
        e.stype = 4;
    } else if (ext == "ply") {
        next();
        int order = atoi(line);
        e.fit_order = order;
        e.poly.assign(order + 1, 0.0);
        for (int k = 0; k <= order; k++) {
            next();
            e.poly[k] = atof(line);
        }
        e.surf = 'r';
        e.stype = 8;
    } else if (ext == "csi") {
        next();
        int n = atoi(line);
        e.sx.assign(n, 0.0);
        e.sy.assign(n, 0.0);
        double x = 0, y = 0;
        for (int k = 0; k < n; k++) {
            next();
            sscanf(line, "%lg %lg", &x, &y);
            e.sx[k] = x;
            e.sy[k] = y;
        }
        next();
        sscanf(line, "%lg %lg", &x, &y);
        spline_y2(e.sx, e.sy, x, y, e.sy2);
        const bool al = (e.shape == 'a' || e.shape == 'A' || e.shape == 'l' || e.shape == 'L');
        if (!al || e.p[0] < e.sx[0] || e.p[1] > e.sx[n - 1]) {
            msg = "Error: Element uses cubic spline interpolation: make sure aperture is type 'a' or 'l'.\n";
            ok = false;
        } else {
            e.surf = 'i';
            e.stype = 9;
        }
    } else {
        snprintf(b, sizeof b, "Surface file type extension unknown: '%s'\n", ext.c_str());
        msg = b;
        ok = false;
    }
    fclose(fp);
    return ok;
}

// -------------------------------------------------------- aperture plane --
// coretrace/apertureplane.cpp:57-491
static void aperture_plane(HElement &e) {
    auto sq = [](double v) { return v * v; };
    const double A = e.p[0], B = e.p[1];
    const char sh = (char)tolower(e.shape);
    double maxr = 0.0;
    if (sh == 'i' || sh == 'q') {
        maxr = sqrt(sq(e.p[0]) + sq(e.p[1]));
        maxr = std::max(maxr, sqrt(sq(e.p[2]) + sq(e.p[3])));
        maxr = std::max(maxr, sqrt(sq(e.p[4]) + sq(e.p[5])));
        if (sh == 'q') maxr = std::max(maxr, sqrt(sq(e.p[6]) + sq(e.p[7])));
    }
    const bool cht = (sh == 'c' || sh == 'h' || sh == 't');
    const bool al = (sh == 'a' || sh == 'l');
    const bool iq = (sh == 'i' || sh == 'q');
    const double r2al = (fabs(B) >= fabs(A)) ? sq(B) : sq(A);
    switch (tolower(e.surf)) {
    case 'g': {
        double r2 = 0, sum = 0;
        if (cht) {
            r2 = sq(A / 2.0);
            double rho = 1.0;
            for (int i = 0; i < 5; i++) { rho = rho * r2; sum += e.alpha[i] * rho; }
            e.zap = e.vcx * r2 / (1 + sqrt(1 - e.kappa * e.vcx * e.vcx * r2)) + sum;
        } else if (sh == 'r') {
            r2 = sq(A / 2.0) + sq(B / 2.0);
            for (int i = 0; i < 5; i++) sum += e.alpha[i] * pow(r2, i + 1);
            e.zap = e.vcx * r2 / (1 + sqrt(1 - e.kappa * sq(e.vcx) * r2)) + sum;
        } else if (al) {
            for (int i = 0; i < 5; i++) sum = sum + e.alpha[i] * pow(r2al, i + 1);
            e.zap = e.vcx * r2al / (1 + sqrt(1 - e.kappa * sq(e.vcx) * r2al)) + sum;
        } else if (iq) {
            r2 = sq(maxr);
            e.zap = e.vcx * r2 / (1 + sqrt(1 - e.kappa * sq(e.vcx) * r2)) + 0.0;
        }
        break;
    }
    case 's': {
        double F = 1.0 / (2.0 * e.vcx);
        if (cht) e.zap = 2.0 * F - sqrt(sq(2.0 * F) - sq(A / 2.0));
        else if (sh == 'r') e.zap = 2.0 * F - sqrt(sq(2.0 * F) - (sq(A / 2.0) + sq(B / 2.0)));
        else if (al) e.zap = 2.0 * F - sqrt(sq(2.0 * F) - r2al);
        else if (iq) e.zap = 2.0 * F - sqrt(sq(2.0 * F) - sq(maxr));
        break;
    }
    case 'p': {
        double F = 1.0 / (2.0 * e.vcx);
        if (cht) e.zap = sq(A / 2.0) / (4.0 * F);
        else if (sh == 'r') e.zap = (sq(A / 2.0) + sq(B / 2.0)) / (4.0 * F);
        else if (al) e.zap = r2al / (4.0 * F);
        else if (iq) e.zap = sq(maxr) / (4.0 * F);
        break;
    }
    case 'f':
        e.zap = 0.0;
        break;
    case 'o': {
        auto z = [&](double r2) { return (1 - sqrt(1 - e.kappa * sq(e.vcx) * r2)) / (e.kappa * e.vcx); };
        if (cht) e.zap = z(sq(A / 2.0));
        else if (sh == 'r') e.zap = z(sq(A / 2.0) + sq(B / 2.0));
        else if (al) e.zap = z(r2al);
        else if (iq) e.zap = z(sq(maxr));
        break;
    }
    case 'c': {
        double t = tan(e.cone_half * kDegRad);
        if (cht) e.zap = A / (2.0 * t);
        else if (sh == 'r') e.zap = sqrt(sq(A / 2.0) + sq(B / 2.0)) / (2.0 * t);
        else if (sh == 'a') e.zap = B / t;
        else if (iq) e.zap = sqrt(sq(maxr)) / (2.0 * t);
        break;
    }
    case 't':
        e.zap = 2.0 / e.curv_rev;
        break;
    case 'd':
        e.zap = 2.0 * e.cross_r;
        break;
    case 'm': {
        int n = e.fit_order + 1;
        if (!(n > 2)) return;
        double F = (1.0 / (4.0 * e.zern[2 * n + 0]) + 1.0 / (4.0 * e.zern[2 * n + 2])) / 2.0;
        if (cht) e.zap = sq(A / 2.0) / (4.0 * F);
        else if (sh == 'r') e.zap = (sq(A / 2.0) + sq(B / 2.0)) / (4.0 * F);
        else if (al) e.zap = r2al / (4.0 * F);
        else if (iq) e.zap = sq(maxr) / (4.0 * F);
        break;
    }
    case 'r': {
        auto ev = [&](double r) {
            double z = 0.0;
            for (int i = 0; i <= e.fit_order; i++) z = z + e.poly[i] * pow(r, i);
            return z;
        };
        if (cht) e.zap = ev(A / 2.0);
        else if (sh == 'r') e.zap = ev(sqrt(sq(A / 2.0) + sq(B / 2.0)));
        else if (al) e.zap = ev(sqrt(r2al));
        else if (iq) e.zap = ev(maxr);
        break;
    }
    case 'i':
        e.zap = e.sy.empty() ? 0.0 : e.sy.back();
        break;
    case 'e': {
        // the largest node *y* (column 1), at least 0, as the reference has it
        // (coretrace/apertureplane.cpp:458-471)
        double lastz = 0.0;
        for (size_t i = 0; i < e.fe_nodes.size() / 3; i++)
            if (e.fe_nodes[i * 3 + 1] > lastz) lastz = e.fe_nodes[i * 3 + 1];
        e.zap = lastz;
        break;
    }
    case 'v': {
        // the largest fitted z over the data points, at least 0
        // (coretrace/apertureplane.cpp:474-488)
        const int n = e.fit_order + 1;
        double lastz = 0.0;
        for (size_t i = 0; i < e.vshot.size() / 4; i++) {
            const double x = e.vshot[i * 4], y = e.vshot[i * 4 + 1];
            double zm = 0.0;
            for (int k = 0; k <= e.fit_order; k++)
                for (int m = 0; m <= k; m++) zm += e.zern[(size_t)k * n + m] * pow(x, m) * pow(y, k - m);
            if (zm > lastz) lastz = zm;
        }
        e.zap = lastz;
        break;
    }
    default:
        break;
    }
}

// Aperture area (coretrace/input.cpp:204-254).
static double aperture_area(const HElement &e) {
    auto sq = [](double v) { return v * v; };
    const double *P = e.p;
    switch (tolower(e.shape)) {
    case 'c': return M_PI * sq(P[0] / 2);
    case 'h': return 5.0 * sq(P[0] / 2.0) * cos(30.0 * kDegRad) * sin(30.0 * kDegRad);
    case 'r': return P[0] * P[1];
    case 't': return 0.25 * sq(P[0] * cos(30.0 * kDegRad)) * sqrt(3.0);
    case 'a': return P[2] * kDegRad * (P[1] - P[0]);
    case 'l': return (P[1] - P[0]) * P[2];
    case 'i': {
        const double v11 = P[0] - P[2], v12 = P[1] - P[3], v21 = P[4] - P[2], v22 = P[5] - P[3];
        const double v1m = sqrt(v11 * v11 + v12 * v12), v2m = sqrt(v21 * v21 + v22 * v22);
        return 0.5 * v1m * v2m * sin(acos((v11 * v21 + v12 * v22) / (v1m * v2m)));
    }
    case 'q': {
        const double v11 = P[0] - P[2], v12 = P[1] - P[3], v21 = P[4] - P[2], v22 = P[5] - P[3];
        const double v31 = P[4] - P[6], v32 = P[5] - P[7], v41 = P[0] - P[6], v42 = P[1] - P[7];
        const double v1m = sqrt(v11 * v11 + v12 * v12), v2m = sqrt(v21 * v21 + v22 * v22);
        const double v3m = sqrt(v31 * v31 + v32 * v32), v4m = sqrt(v41 * v41 + v42 * v42);
        const double t1 = acos((v11 * v21 + v12 * v22) / (v1m * v2m));
        const double t2 = acos((v31 * v41 + v32 * v42) / (v3m * v4m));
        return 0.5 * (v1m * v2m * sin(t1) + v3m * v4m * sin(t2));
    }
    default: return 0.0;
    }
}

static int ap_kind(char c) {
    switch (tolower(c)) {
    case 'c': return AP_CIRCLE;
    case 'h': return AP_HEX;
    case 't': return AP_TRI;
    case 'r': return AP_RECT;
    case 'a': return AP_ANNULUS;
    case 'l': return AP_LINE;
    case 'i': return AP_ITRI;
    case 'q': return AP_QUAD;
    default: return 0;
    }
}

static int surf_kind(const HElement &e) {
    switch (tolower(e.surf)) {
    case 's': return SURF_SPHERE;
    case 'p': return SURF_PARABOLA;
    case 'o': return e.stype == 1 ? SURF_CONIC : SURF_GENERAL;
    case 't': return SURF_CYLINDER;
    case 'd': return SURF_TORUS;
    case 'f': return SURF_FLAT;
    case 'm': return SURF_ZERNIKE;
    case 'v': return SURF_VSHOT;
    case 'e': return SURF_FE;
    case 'i': return SURF_SPLINE;
    case 'r': return SURF_POLY;
    default: return SURF_GENERAL;  // 'g', 'c'
    }
}

// Surface height used only to bound where accepted hits can lie.
static double bound_z(const HElement &e, double x, double y) {
    switch (tolower(e.surf)) {
    case 's': {
        double c = e.vcx;
        double r2 = x * x + (e.stype == 7 ? 0.0 : y * y);
        double q = 1.0 - c * c * r2;
        if (q < 0) q = 0;
        return c * r2 / (1.0 + sqrt(q));
    }
    case 'p':
        return e.vcx * x * x / 2.0 + (e.stype == 7 ? 0.0 : e.vcy * y * y / 2.0);
    case 'o': case 'g': {
        if (e.cone_half != 0.0) return sqrt(x * x + y * y) / tan(e.cone_half * kDegRad);
        double r2 = e.vcx * x * x + e.vcy * y * y;
        double q = 1.0 - e.kappa * (e.vcx * e.vcx * x * x + e.vcy * e.vcy * y * y);
        if (q < 0) q = 0;
        return r2 / (1.0 + sqrt(q));
    }
    case 'c':
        return sqrt(x * x + y * y) / tan(e.cone_half * kDegRad);
    case 'm': case 'v': {
        int n = e.fit_order + 1;
        double z = 0.0;
        for (int i = 0; i <= e.fit_order; i++)
            for (int j = 0; j <= i; j++) z += e.zern[(size_t)i * n + j] * pow(x, j) * pow(y, i - j);
        return z;
    }
    case 'r': {
        double r = (e.shape == 'l' || e.shape == 'L') ? fabs(x) : sqrt(x * x + y * y);
        double z = 0.0;
        for (int i = 0; i <= e.fit_order; i++) z += e.poly[i] * pow(r, i);
        return z;
    }
    case 'i': {
        double r = (e.shape == 'l' || e.shape == 'L') ? x : sqrt(x * x + y * y);
        const int n = (int)e.sx.size();
        if (n < 2) return 0.0;
        if (r <= e.sx[0]) return e.sy[0];
        if (r >= e.sx[n - 1]) return e.sy[n - 1];
        int lo = 0, hi = n - 1;
        while (hi - lo > 1) {
            int k = (lo + hi) / 2;
            if (e.sx[k] > r) hi = k; else lo = k;
        }
        double h = e.sx[hi] - e.sx[lo];
        double a = (e.sx[hi] - r) / h, b = (r - e.sx[lo]) / h;
        return a * e.sy[lo] + b * e.sy[hi] + ((a * a * a - a) * e.sy2[lo] + (b * b * b - b) * e.sy2[hi]) * (h * h) / 6.0;
    }
    default:
        return 0.0;
    }
}

// Element-frame boxes holding every hit the element's intersection can
// accept, as lo[b], hi[b].  Returns how many boxes bound the hits: 1, or 2
// for a closed-form sphere or conic, whose solve takes the far root when the
// near one lies above the z aperture (coretrace/quadricsurfaceclosedform.cpp
// root selection) -- the other side of a sphere, the other sheet of a
// hyperboloid -- and 0 for the torus, whose quartic solution
// (coretrace/torusclosedform.cpp, root432.cpp) can land off the surface, far
// outside the torus, so culling must always test it (box 0 is still its
// nominal box); 0 also for finite-element surfaces, whose kriging
// interpolant can overshoot its nodes.
static int element_boxes(const HElement &e, double lo[2][3], double hi[2][3]) {
    const double A = e.p[0], B = e.p[1], C = e.p[2];
    double xl, xh, yl, yh;
    const char sh = (char)tolower(e.shape);
    const char sf = (char)tolower(e.surf);
    switch (sh) {
    case 'c': case 'h': case 't':
        xl = yl = -A / 2.0; xh = yh = A / 2.0;
        break;
    case 'r':
        xl = -A / 2.0; xh = A / 2.0; yl = -B / 2.0; yh = B / 2.0;
        break;
    case 'a':
        if (A == 0.0 && B == 0.0) {
            double r = (sf == 'd') ? e.ann_r + e.cross_r : 1e3;
            xl = yl = -r; xh = yh = r;
        } else {
            xl = yl = -fabs(B); xh = yh = fabs(B);
        }
        break;
    case 'l':
        if (A == 0.0 && B == 0.0) {
            double r = (sf == 't') ? 1.0 / fabs(e.curv_rev) : 1e3;
            xl = -r; xh = r;
        } else {
            xl = std::min(A, B); xh = std::max(A, B);
        }
        yl = -C / 2.0; yh = C / 2.0;
        break;
    default: {  // i, q
        int nv = (sh == 'q') ? 4 : 3;
        xl = yl = 1e300; xh = yh = -1e300;
        for (int k = 0; k < nv; k++) {
            xl = std::min(xl, e.p[2 * k]); xh = std::max(xh, e.p[2 * k]);
            yl = std::min(yl, e.p[2 * k + 1]); yh = std::max(yh, e.p[2 * k + 1]);
        }
        break;
    }
    }
    double zl[2] = {std::min(0.0, e.zap), HUGE_VAL}, zh[2] = {std::max(0.0, e.zap), -HUGE_VAL};
    int nb = 1;
    if (sf == 't') {
        zl[0] = 0.0; zh[0] = 2.0 / fabs(e.curv_rev);
    } else if (sf == 'd') {
        zl[0] = 0.0; zh[0] = 2.0 * e.cross_r;
    } else if (sf != 'f') {
        // Near and far roots lie on the branches
        // z = (1 -/+ sqrt(1 - k c^2 r^2)) / (k c), k = 1 for the sphere; both
        // are monotonic in r^2, so the grid samples (corners included), the
        // point nearest the axis and, where the box reaches past the
        // surface's rim (q = 0), the rim height 1/(k c) bound them.
        const bool two = sf == 's' || (sf == 'o' && e.stype == 1 && e.kappa != 0.0);
        const double c = e.vcx, k = sf == 's' ? 1.0 : e.kappa;
        const bool line = sf == 's' && e.stype == 7;
        auto put = [&](int b, double z) {
            if (std::isfinite(z)) { zl[b] = std::min(zl[b], z); zh[b] = std::max(zh[b], z); }
        };
        auto sample = [&](double x, double y) {
            put(0, bound_z(e, x, y));
            if (two && c != 0.0) {
                const double q = 1.0 - k * c * c * (x * x + (line ? 0.0 : y * y));
                if (q >= 0.0) put(1, (1.0 + sqrt(q)) / (k * c));
            }
        };
        const int G = 64;
        for (int i = 0; i <= G; i++)
            for (int j = 0; j <= G; j++) sample(xl + (xh - xl) * i / G, yl + (yh - yl) * j / G);
        sample(std::min(std::max(0.0, xl), xh), std::min(std::max(0.0, yl), yh));
        if (two && c != 0.0 && k > 0.0) {
            const double mx = std::max(fabs(xl), fabs(xh)), my = line ? 0.0 : std::max(fabs(yl), fabs(yh));
            if (k * c * c * (mx * mx + my * my) >= 1.0) { put(0, 1.0 / (k * c)); put(1, 1.0 / (k * c)); }
        }
        for (int b = 0; b < 2; b++) {
            if (!(zl[b] <= zh[b])) continue;
            const double pad = 0.05 * (zh[b] - zl[b]) + 1e-3;
            zl[b] -= pad; zh[b] += pad;
        }
        // The trace refuses a hit above the aperture plane, z > zap, of
        // these surfaces (coretrace/raytrace.cpp:747-751): a far branch
        // wholly above it, the far cap of a concave mirror, bounds no hit.
        if (zl[1] <= zh[1] && zl[1] > e.zap) {
            zl[1] = HUGE_VAL;
            zh[1] = -HUGE_VAL;
        }
        if (zl[1] <= zh[1]) {
            if (zl[1] <= zh[0] && zl[0] <= zh[1]) {  // the branches meet: one box
                zl[0] = std::min(zl[0], zl[1]);
                zh[0] = std::max(zh[0], zh[1]);
            } else {
                nb = 2;
            }
        }
    }
    const double pad = 1e-4 + 1e-7 * std::max({fabs(xl), fabs(xh), fabs(yl), fabs(yh)});
    for (int b = 0; b < 2; b++) {
        const int src = b < nb ? b : 0;
        lo[b][0] = xl - pad; hi[b][0] = xh + pad;
        lo[b][1] = yl - pad; hi[b][1] = yh + pad;
        lo[b][2] = zl[src] - pad; hi[b][2] = zh[src] + pad;
    }
    return (sf == 'd' || sf == 'e') ? 0 : nb;
}

// ------------------------------------------------------------------- BVH --
struct BuildItem {
    float lo[3], hi[3];
    float c[3];
    int32_t idx;
};

static double box_area(const float lo[3], const float hi[3]) {
    const double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

// Binned surface-area-heuristic build: split where the expected cost of a
// ray entering the node, node test (1) plus element tests (2 each) weighted
// by the children's surface areas, is least; leaves hold at most 7 items
// (the node's 3-bit count).
static int build_bvh(std::vector<BuildItem> &items, int begin, int end, std::vector<DevBvhNode> &nodes,
                     std::vector<int32_t> &out_items) {
    const int me = (int)nodes.size();
    nodes.push_back(DevBvhNode{});
    float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
    float clo[3] = {3e38f, 3e38f, 3e38f}, chi[3] = {-3e38f, -3e38f, -3e38f};
    for (int i = begin; i < end; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], items[i].lo[k]); hi[k] = std::max(hi[k], items[i].hi[k]);
            clo[k] = std::min(clo[k], items[i].c[k]); chi[k] = std::max(chi[k], items[i].c[k]);
        }
    DevBvhNode nd{};
    for (int k = 0; k < 3; k++) { nd.lo[k] = lo[k]; nd.hi[k] = hi[k]; }
    const int n = end - begin;
    auto make_leaf = [&]() {
        nd.leaf = ((int)out_items.size() << 3) | n;
        for (int i = begin; i < end; i++) out_items.push_back(items[i].idx);
        nd.skip = me + 1;
        nodes[me] = nd;
        return me;
    };
    if (n <= 2) return make_leaf();
    constexpr int NB = 16;
    const double area = std::max(box_area(lo, hi), 1e-30);
    double best = HUGE_VAL;
    int best_ax = -1, best_b = 0;
    for (int ax = 0; ax < 3; ax++) {
        const double ext = (double)chi[ax] - clo[ax];
        if (!(ext > 0)) continue;
        int cnt[NB] = {};
        float blo[NB][3], bhi[NB][3];
        for (int b = 0; b < NB; b++)
            for (int k = 0; k < 3; k++) { blo[b][k] = 3e38f; bhi[b][k] = -3e38f; }
        for (int i = begin; i < end; i++) {
            const int b = std::min(NB - 1, (int)((items[i].c[ax] - clo[ax]) / ext * NB));
            cnt[b]++;
            for (int k = 0; k < 3; k++) {
                blo[b][k] = std::min(blo[b][k], items[i].lo[k]);
                bhi[b][k] = std::max(bhi[b][k], items[i].hi[k]);
            }
        }
        double right_area[NB];
        int right_cnt[NB];
        float rlo[3] = {3e38f, 3e38f, 3e38f}, rhi[3] = {-3e38f, -3e38f, -3e38f};
        int rc = 0;
        for (int b = NB - 1; b >= 1; b--) {
            rc += cnt[b];
            for (int k = 0; k < 3; k++) { rlo[k] = std::min(rlo[k], blo[b][k]); rhi[k] = std::max(rhi[k], bhi[b][k]); }
            right_cnt[b] = rc;
            right_area[b] = rc ? box_area(rlo, rhi) : 0.0;
        }
        float llo[3] = {3e38f, 3e38f, 3e38f}, lhi[3] = {-3e38f, -3e38f, -3e38f};
        int lc = 0;
        for (int b = 1; b < NB; b++) {  // split before bin b
            lc += cnt[b - 1];
            for (int k = 0; k < 3; k++) {
                llo[k] = std::min(llo[k], blo[b - 1][k]);
                lhi[k] = std::max(lhi[k], bhi[b - 1][k]);
            }
            if (lc == 0 || right_cnt[b] == 0) continue;
            const double c = 1.0 + 2.0 * (box_area(llo, lhi) * lc + right_area[b] * right_cnt[b]) / area;
            if (c < best) { best = c; best_ax = ax; best_b = b; }
        }
    }
    if (n <= 7 && !(best < 2.0 * n)) return make_leaf();
    int mid;
    if (best_ax >= 0) {
        const int ax = best_ax;
        const double ext = (double)chi[ax] - clo[ax];
        auto it = std::partition(items.begin() + begin, items.begin() + end, [&](const BuildItem &a) {
            return std::min(NB - 1, (int)((a.c[ax] - clo[ax]) / ext * NB)) < best_b;
        });
        mid = (int)(it - items.begin());
    } else {
        // every centroid equal: halve the range
        mid = begin + n / 2;
    }
    build_bvh(items, begin, mid, nodes, out_items);   // left child = me + 1
    build_bvh(items, mid, end, nodes, out_items);
    nd.leaf = 0;
    nd.skip = (int)nodes.size();                      // first node after this subtree
    nodes[me] = nd;
    return me;
}

static float round_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -3e38f);
    return f;
}
static float round_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, 3e38f);
    return f;
}

// ------------------------------------------------------------- compile ----
bool compile_scene(SceneModel &m, bool use_cull, CompiledScene &out, std::string &msg) {
    char b[600];
    out = CompiledScene();
    // InitGeometries (coretrace/input.cpp:153-286)
    for (size_t i = 0; i < m.stages.size(); i++) {
        HStage &st = m.stages[i];
        if (!euler_from_aim(st.origin, st.aim, st.zrot, st.euler)) {
            snprintf(b, sizeof b, "Stage %d geometry as defined not possible.\n", (int)i);
            msg = b;
            return false;
        }
        euler_matrix(st.euler, st.R);
        for (size_t j = 0; j < st.elements.size(); j++) {
            HElement &e = st.elements[j];
            if (!euler_from_aim(e.origin, e.aim, e.zrot, e.euler)) {
                snprintf(b, sizeof b, "Stage %d element %d geometry as defined not possible.\n", (int)i, (int)j);
                msg = b;
                return false;
            }
            euler_matrix(e.euler, e.R);
            if (ap_kind(e.shape) == 0) {
                snprintf(b, sizeof b, "invalid shape '%c' for aperture, element %d\n", e.shape, (int)j);
                msg = b;
                return false;
            }
            if (!st.virt) {
                const HOptic *set = nullptr;
                for (const HOptic &o : m.optics)
                    if (strcasecmp(o.name.c_str(), e.optic_name.c_str()) == 0) set = &o;
                if (!set) {
                    snprintf(b, sizeof b, "Stage %d element %d invalid optical property set name: '%s'\n", (int)i,
                             (int)j, e.optic_name.c_str());
                    msg = b;
                    return false;
                }
            }
            aperture_plane(e);
        }
    }
    // refractive index wiring (coretrace/input.cpp:269-275), applied to
    // every optic that any non-virtual element uses
    for (HOptic &o : m.optics) {
        bool used = false;
        for (const HStage &st : m.stages)
            if (!st.virt)
                for (const HElement &e : st.elements)
                    if (strcasecmp(o.name.c_str(), e.optic_name.c_str()) == 0) used = true;
        if (!used) continue;
        o.front.refr[2] = o.back.refr[0];
        o.front.refr[3] = o.back.refr[1];
        o.back.refr[2] = o.front.refr[0];
        o.back.refr[3] = o.front.refr[1];
    }
    // CheckInputs (coretrace/input.cpp:87-150)
    for (size_t i = 0; i < m.stages.size(); i++)
        for (size_t j = 0; j < m.stages[i].elements.size(); j++) {
            const HElement &e = m.stages[i].elements[j];
            char s = (char)tolower(e.surf), a = (char)tolower(e.shape);
            const char *err = nullptr;
            if (s == 'p' && a == 'l' && e.vcy != 0.0)
                err = "Parabolic surface combined with single axis curvature section aperture must have Cy = 0.";
            else if (s == 's' && e.vcx == 0.0)
                err = "Spherical surface must have a curvature greater than 0.0.";
            else if (s == 's' && e.vcy != 0.0 && e.vcy != e.vcx)
                err = "Spherical surface must have Cy=0 or Cy=Cx.";
            else if (s == 'o' && (e.kappa == 1.0 || e.kappa == 0.0))
                err = "Kappa = 0.0 or 1.0 defines either a parabola or a sphere.";
            if (err) {
                snprintf(b, sizeof b, "%s  Element %d, Stage %d.", err, (int)j + 1, (int)i + 1);
                msg = b;
                return false;
            }
        }

    // --------------------------------------------------- device tables ----
    auto add_table = [&](const std::vector<double> &a, const std::vector<double> &v) {
        int off = (int)out.aux.size();
        out.aux.insert(out.aux.end(), a.begin(), a.end());
        out.aux.insert(out.aux.end(), v.begin(), v.end());
        return off;
    };
    std::vector<int> optic_face0(m.optics.size(), -1);
    for (size_t k = 0; k < m.optics.size(); k++) {
        optic_face0[k] = (int)out.faces.size();
        for (int fb = 0; fb < 2; fb++) {
            const HFace &h = fb ? m.optics[k].back : m.optics[k].front;
            DevFace f{};
            f.dist = (unsigned char)h.dist;
            f.reflectivity = h.reflectivity;
            f.transmissivity = h.transmissivity;
            f.slope = h.slope;
            f.spec = h.spec;
            f.n_in = h.refr[0];
            f.n_out = h.refr[2];
            f.refl_n = h.use_refl ? (int)h.refl_ang.size() : 0;
            f.trans_n = h.use_trans ? (int)h.trans_ang.size() : 0;
            if (f.refl_n) f.refl_tab = add_table(h.refl_ang, h.refl_val);
            if (f.trans_n) f.trans_tab = add_table(h.trans_ang, h.trans_val);
            out.faces.push_back(f);
        }
    }
    for (size_t i = 0; i < m.stages.size(); i++) {
        HStage &st = m.stages[i];
        DevStage ds{};
        for (int k = 0; k < 3; k++) ds.org[k] = st.origin[k];
        for (int k = 0; k < 9; k++) ds.R[k] = st.R[k];
        ds.elem0 = (int)out.elems.size();
        ds.nelem = (int)st.elements.size();
        ds.virt = st.virt;
        ds.multihit = st.multihit;
        ds.tracethrough = st.tracethrough;
        std::vector<BuildItem> items;
        std::vector<int32_t> unbounded;
        for (size_t j = 0; j < st.elements.size(); j++) {
            HElement &e = st.elements[j];
            DevElem d{};
            for (int k = 0; k < 3; k++) d.org[k] = e.origin[k];
            for (int k = 0; k < 9; k++) d.R[k] = e.R[k];
            for (int k = 0; k < 8; k++) d.p[k] = e.p[k];
            d.zap = e.zap;
            d.kappa = e.kappa;
            d.vcx = e.vcx;
            d.vcy = e.vcy;
            d.ann_r = e.ann_r;
            d.cross_r = e.cross_r;
            d.curv_rev = e.curv_rev;
            d.cone_tan = (e.cone_half != 0.0) ? tan(e.cone_half * kDegRad) : 0.0;
            d.ap = ap_kind(e.shape);
            d.surf = surf_kind(e);
            d.stype = e.stype;
            d.interaction = e.interaction;
            d.enabled = e.enabled;
            d.fit_order = e.fit_order;
            d.zap_filter = !(tolower(e.surf) == 'm' || tolower(e.surf) == 'r');
            d.face_front = d.face_back = -1;
            if (!st.virt) {
                for (size_t k = 0; k < m.optics.size(); k++)
                    if (strcasecmp(m.optics[k].name.c_str(), e.optic_name.c_str()) == 0) {
                        d.face_front = optic_face0[k];
                        d.face_back = optic_face0[k] + 1;
                    }
            }
            if (d.surf == SURF_FE) {
                // the node tree flattened into aux (doubles): a header
                // {x0, x1, y0, y1, address bits, derivative step, zones,
                // nodes}, then per zone {m0, m1, terminal, first candidate,
                // candidates}, then x y z per node, then the candidate lists
                // (the zone's nodes, then its neighbours' nodes)
                const st_fe::Tree &t = e.fe_tree;
                const int nn = (int)(e.fe_nodes.size() / 3), nz = (int)t.zones.size();
                const double density = (double)nn / aperture_area(e);
                d.aux = (int)out.aux.size();
                out.aux.insert(out.aux.end(), {t.xlim[0], t.xlim[1], t.ylim[0], t.ylim[1], (double)t.key_length(),
                                               0.001 / sqrt(density), (double)nz, (double)nn});
                const size_t zone0 = out.aux.size();
                out.aux.resize(zone0 + 5 * (size_t)nz);
                out.aux.insert(out.aux.end(), e.fe_nodes.begin(), e.fe_nodes.end());
                for (int z = 0; z < nz; z++) {
                    const st_fe::Zone &zn = t.zones[z];
                    double *rec = &out.aux[zone0 + 5 * (size_t)z];
                    rec[0] = zn.m0;
                    rec[1] = zn.m1;
                    rec[2] = zn.terminal ? 1.0 : 0.0;
                    rec[3] = (double)out.aux.size();
                    rec[4] = (double)(zn.data.size() + zn.neighbors.size());
                    for (int i : zn.data) out.aux.push_back(i);
                    for (int i : zn.neighbors) out.aux.push_back(i);
                }
            } else if (d.surf == SURF_ZERNIKE || d.surf == SURF_VSHOT) {
                d.aux = (int)out.aux.size();
                d.aux_n = e.fit_order + 1;
                out.aux.insert(out.aux.end(), e.zern.begin(), e.zern.end());
                if (d.surf == SURF_VSHOT) {
                    // density = points / aperture area, and the region of
                    // interest of the Shepard weights (coretrace/surface.cpp:
                    // 258, vshot.cpp:317)
                    d.aux2 = (int)out.aux.size();
                    d.aux2_n = (int)(e.vshot.size() / 4);
                    out.aux.insert(out.aux.end(), e.vshot.begin(), e.vshot.end());
                    const double density = (double)d.aux2_n / aperture_area(e);
                    d.roi2 = 30.0 / density / 3.14159;
                }
            } else if (d.surf == SURF_SPLINE) {
                d.aux = (int)out.aux.size();
                d.aux_n = (int)e.sx.size();
                out.aux.insert(out.aux.end(), e.sx.begin(), e.sx.end());
                out.aux.insert(out.aux.end(), e.sy.begin(), e.sy.end());
                out.aux.insert(out.aux.end(), e.sy2.begin(), e.sy2.end());
            } else if (d.surf == SURF_POLY) {
                d.aux = (int)out.aux.size();
                d.aux_n = (int)e.poly.size();
                out.aux.insert(out.aux.end(), e.poly.begin(), e.poly.end());
            }
            double lo[2][3], hi[2][3];
            const int nb = element_boxes(e, lo, hi);
            out.enbox.push_back((int8_t)nb);
            for (int q = 0; q < 2; q++)
                out.ebox.insert(out.ebox.end(), {lo[q][0], lo[q][1], lo[q][2], hi[q][0], hi[q][1], hi[q][2]});
            // sphere around box 0 (the torus solve starts on it for far origins)
            double c[3] = {(lo[0][0] + hi[0][0]) / 2, (lo[0][1] + hi[0][1]) / 2, (lo[0][2] + hi[0][2]) / 2};
            double r = 0.5 * sqrt((hi[0][0] - lo[0][0]) * (hi[0][0] - lo[0][0]) +
                                  (hi[0][1] - lo[0][1]) * (hi[0][1] - lo[0][1]) +
                                  (hi[0][2] - lo[0][2]) * (hi[0][2] - lo[0][2]));
            for (int k = 0; k < 3; k++) d.bsphere_c[k] = c[k];
            d.bsphere_r = r;
            if (d.surf != SURF_FLAT && d.surf > SURF_CYLINDER) ds.generic = std::max(ds.generic, 1);
            if (d.surf == SURF_FE) ds.generic = 2;
            out.elems.push_back(d);
            if (e.enabled && nb == 0) unbounded.push_back((int32_t)j);
            for (int q = 0; e.enabled && q < nb; q++) {
                // stage-frame box from the 8 corners of an element-frame box
                BuildItem it;
                double slo[3] = {1e300, 1e300, 1e300}, shi[3] = {-1e300, -1e300, -1e300};
                for (int v8 = 0; v8 < 8; v8++) {
                    double v[3] = {(v8 & 1) ? hi[q][0] : lo[q][0], (v8 & 2) ? hi[q][1] : lo[q][1],
                                   (v8 & 4) ? hi[q][2] : lo[q][2]};
                    double w[3];
                    mtv(e.R, v, w);
                    for (int k = 0; k < 3; k++) {
                        w[k] += e.origin[k];
                        slo[k] = std::min(slo[k], w[k]);
                        shi[k] = std::max(shi[k], w[k]);
                    }
                }
                for (int k = 0; k < 3; k++) {
                    double pad = 1e-2 + 1e-5 * std::max(fabs(slo[k]), fabs(shi[k]));
                    it.lo[k] = round_down(slo[k] - pad);
                    it.hi[k] = round_up(shi[k] + pad);
                    it.c[k] = 0.5f * (it.lo[k] + it.hi[k]);
                }
                it.idx = (int32_t)j;
                items.push_back(it);
                if (i == 0) {
                    DevBeamItem bi{};
                    for (int k = 0; k < 3; k++) { bi.lo[k] = it.lo[k]; bi.hi[k] = it.hi[k]; }
                    bi.j = (int32_t)j;
                    bi.box = q;
                    out.s0_boxes.push_back(bi);
                }
            }
        }
        // the BVH holds the bounded elements; the others (torus, finite
        // element) are tested before every walk
        if (use_cull && items.size() + unbounded.size() > 8 && !items.empty()) {
            ds.bvh0 = (int)out.bvh.size();
            ds.always0 = (int)out.always_items.size();
            ds.nalways = (int)unbounded.size();
            out.always_items.insert(out.always_items.end(), unbounded.begin(), unbounded.end());
            std::vector<DevBvhNode> nodes;
            build_bvh(items, 0, (int)items.size(), nodes, out.bvh_items);
            ds.nbvh = (int)nodes.size();
            // child indices are stage-local (the kernel offsets the node base)
            out.bvh.insert(out.bvh.end(), nodes.begin(), nodes.end());
        }
        out.stages.push_back(ds);
    }
    return true;
}

}  // namespace st

namespace st {

// Convex hull of sun-plane points as outward half-planes n.p <= off,
// relative to the vertices' mean and padded by pad (monotone chain,
// counter-clockwise).  False when degenerate or above 6 edges.
static bool foot_hull(std::vector<std::pair<double, double>> pts, double pad, DevFoot &f) {
    using P = std::pair<double, double>;
    std::sort(pts.begin(), pts.end());
    auto cross = [](const P &o, const P &a, const P &b) {
        return (a.first - o.first) * (b.second - o.second) - (a.second - o.second) * (b.first - o.first);
    };
    std::vector<P> h(2 * pts.size());
    int k = 0;
    for (size_t i = 0; i < pts.size(); i++) {
        while (k >= 2 && cross(h[k - 2], h[k - 1], pts[i]) <= 0) k--;
        h[k++] = pts[i];
    }
    for (int i = (int)pts.size() - 2, t = k + 1; i >= 0; i--) {
        while (k >= t && cross(h[k - 2], h[k - 1], pts[i]) <= 0) k--;
        h[k++] = pts[i];
    }
    const int nv = k - 1;
    if (nv < 3 || nv > 6) return false;
    double cx = 0, cy = 0;
    for (int i = 0; i < nv; i++) { cx += h[i].first; cy += h[i].second; }
    cx /= nv;
    cy /= nv;
    for (int i = 0; i < 6; i++) {
        f.e[i][0] = 0.0f;
        f.e[i][1] = 0.0f;
        f.e[i][2] = 1.0f;
    }
    for (int i = 0; i < nv; i++) {
        const double ex = h[i + 1].first - h[i].first, ey = h[i + 1].second - h[i].second;
        const double len = sqrt(ex * ex + ey * ey);
        if (!(len > 0)) return false;
        const double nx = ey / len, ny = -ex / len;
        const double off = nx * (h[i].first - cx) + ny * (h[i].second - cy) + pad;
        f.e[i][0] = (float)nx;
        f.e[i][1] = (float)ny;
        f.e[i][2] = (float)(off * (1.0 + 1e-6) + 1e-6);
    }
    f.cx = cx;
    f.cy = cy;
    f.kind = 0;
    return true;
}

// Beam lists of the stage-1 elements (DevBeam).  Element i's list holds
// every stage-1 hit box that a ray starting in the sphere S about i's near
// box, in a direction within angle alpha of the axis a, can reach: for a
// box inside the sphere B (centre v from S's centre, radius rho together
// with S's radius) the directions from S to B lie within asin(rho / |v|) of
// v, so the box is kept when angle(v, a) <= alpha + asin(rho / |v|) or
// |v| <= rho.  The host test uses a slightly larger sphere and angle than
// the device's check, so rounding there cannot admit a ray the list does
// not cover.  Lists go nearest box first (an early hit prunes the rest).
static void compile_beams(CompiledScene &out, const double sun_dir[3]) {
    out.beam.clear();
    out.beam_items.clear();
    if (!out.use_grid || out.stages.empty()) return;
    const DevStage &s0 = out.stages[0];
    if (s0.nbvh == 0 || s0.nelem == 0) return;
    const DevSun &sun = out.sun;
    double th_sun;
    switch (sun.shape) {
    case 'p': case 'P': th_sun = sun.sigma / 1000.0; break;
    case 'g': case 'G': th_sun = 5.0 * sun.sigma / 1000.0; break;
    case 'd': case 'D': th_sun = sun.max_angle / 1000.0; break;
    default: return;
    }
    // error angle bound of a face distribution: the pillbox's radius, five
    // standard deviations of the Gaussian (a draw beyond it, about 4e-6 of
    // them, takes the BVH walk); a diffuse face has no bound
    auto err_bound = [](int dist, double mrad, double *th) {
        if (mrad == 0.0) { *th = 0.0; return true; }
        if (dist == 'p' || dist == 'P') { *th = mrad / 1000.0; return true; }
        if (dist == 'g' || dist == 'G') { *th = 5.0 * mrad / 1000.0; return true; }
        return false;
    };
    // element -> its near and far box in s0_boxes (-1: none)
    std::vector<int> box_of[2] = {std::vector<int>(s0.nelem, -1), std::vector<int>(s0.nelem, -1)};
    for (size_t q = 0; q < out.s0_boxes.size(); q++) box_of[out.s0_boxes[q].box][out.s0_boxes[q].j] = (int)q;
    const std::vector<int> &near = box_of[0];
    std::vector<int> stamp(out.s0_boxes.size(), -1);  // box already taken for element i
    auto centre = [](const DevBeamItem &b, double c[3]) {
        double r2 = 0.0;
        for (int k = 0; k < 3; k++) {
            c[k] = 0.5 * ((double)b.lo[k] + (double)b.hi[k]);
            const double h = 0.5 * ((double)b.hi[k] - (double)b.lo[k]);
            r2 += h * h;
        }
        return sqrt(r2);
    };
    out.beam.assign(s0.nelem, DevBeam{});
    int lists = 0;
    for (int i = 0; i < s0.nelem; i++) {
        DevBeam &bm = out.beam[i];
        bm.count = -1;
        const DevElem &d = out.elems[s0.elem0 + i];
        if (!d.enabled || near[i] < 0 || d.interaction != 2 || d.face_front < 0) continue;
        if (!(d.surf == SURF_FLAT || (d.surf == SURF_SPHERE && d.stype != 7))) continue;
        const DevFace &f = out.faces[d.face_front];
        double th_slope, th_spec;
        if (!err_bound(f.dist, f.slope, &th_slope) || !err_bound(f.dist, f.spec, &th_spec)) continue;
        // normal tilt over the aperture: asin(r c) for the sphere
        double th_curv = 0.0;
        if (d.surf == SURF_SPHERE) {
            const double *bx = &out.ebox[12 * (size_t)(s0.elem0 + i)];
            const double rx = std::max(fabs(bx[0]), fabs(bx[3])), ry = std::max(fabs(bx[1]), fabs(bx[4]));
            const double rc = sqrt(rx * rx + ry * ry) * fabs(d.vcx);
            if (!(rc < 1.0)) continue;
            th_curv = asin(rc);
        }
        const double alpha = th_sun + 2.0 * (th_curv + th_slope) + th_spec + 2e-3;
        if (!(alpha < 0.5)) continue;
        // axis: the sun direction reflected at the vertex normal (the
        // element's z axis in stage frame)
        const double n[3] = {d.R[6], d.R[7], d.R[8]};
        const double dn = sun_dir[0] * n[0] + sun_dir[1] * n[1] + sun_dir[2] * n[2];
        double a[3], an = 0.0;
        for (int k = 0; k < 3; k++) { a[k] = sun_dir[k] - 2.0 * dn * n[k]; an += a[k] * a[k]; }
        an = sqrt(an);
        for (int k = 0; k < 3; k++) a[k] /= an;
        double c[3];
        const double r = centre(out.s0_boxes[near[i]], c);
        const double r_host = r * (1.0 + 1e-6) + 1e-3, alpha_host = alpha + 1e-4;
        // whether a ray from S within the cone can reach the sphere (cq, rq);
        // dist: from S's centre.  angle(v, a) <= alpha + asin(rho / |v|) is
        // tested as cos(angle) >= cos(alpha + beta), beta = asin(rho / |v|)
        // (alpha + beta < pi here).
        const double ca = cos(alpha_host), sa = sin(alpha_host);
        auto reach = [&](const double cq[3], double rq, double *dist_out) {
            const double rho = rq + r_host;
            double dist = 0.0, va = 0.0;
            for (int k = 0; k < 3; k++) {
                const double v = cq[k] - c[k];
                dist += v * v;
                va += v * a[k];
            }
            dist = sqrt(dist);
            *dist_out = dist;
            if (dist <= rho) return true;
            const double sb = rho / dist, cb = sqrt(1.0 - sb * sb);
            return va / dist >= ca * cb - sa * sb;
        };
        // the stage's BVH, pruned by the same test on each node's bounding
        // sphere (a box inside a node lies inside the node's sphere, so a
        // node that fails holds no box that passes)
        std::vector<std::pair<double, int>> keep;
        bool too_long = false;
        for (int k = 0; k < s0.nbvh && !too_long;) {
            const DevBvhNode &nd = out.bvh[(size_t)s0.bvh0 + k];
            double cn[3], rn2 = 0.0, dist;
            for (int q = 0; q < 3; q++) {
                cn[q] = 0.5 * ((double)nd.lo[q] + (double)nd.hi[q]);
                const double h = 0.5 * ((double)nd.hi[q] - (double)nd.lo[q]);
                rn2 += h * h;
            }
            if (!reach(cn, sqrt(rn2), &dist)) {
                k = nd.skip;
            } else if (nd.leaf) {
                for (int t = nd.leaf >> 3, te = t + (nd.leaf & 7); t < te; t++) {
                    const int j = out.bvh_items[t];
                    for (int q = 0; q < 2; q++) {
                        const int bq = box_of[q][j];
                        if (bq < 0 || stamp[bq] == i) continue;  // j sits in a leaf once per box
                        stamp[bq] = i;
                        double cq[3];
                        const double rq = centre(out.s0_boxes[bq], cq);
                        if (reach(cq, rq, &dist)) keep.emplace_back(dist, bq);
                    }
                }
                k = nd.skip;
            } else {
                k = k + 1;
            }
            // a long list walks no faster than the BVH
            too_long = keep.size() > 32;
        }
        if (too_long) continue;
        std::sort(keep.begin(), keep.end());
        for (int k = 0; k < 3; k++) { bm.c[k] = c[k]; bm.a[k] = (float)a[k]; }
        bm.r2 = (float)(r * r);
        bm.cosa = (float)cos(alpha);
        bm.first = (int32_t)out.beam_items.size();
        bm.count = (int32_t)keep.size();
        for (auto &kq : keep) out.beam_items.push_back(out.s0_boxes[kq.second]);
        lists++;
    }
    if (lists == 0) {
        out.beam.clear();
        out.beam_items.clear();
    }
}

// Sun frame, sun-plane box and the stage-1 sun-plane grid
// (coretrace/suntoprimarystage.cpp:59-293).
bool compile_sun(SceneModel &m, CompiledScene &out, std::string &msg) {
    HStage &st = m.stages[0];
    HSun &sun = m.sun;
    double dx = 0.0 - sun.origin[0], dy = 0.0 - sun.origin[1], dz = 0.0 - sun.origin[2];
    double dtot = sqrt(dx * dx + dy * dy + dz * dz);
    if (dtot == 0.0) {
        msg = "error calculating sun position in primary stage, dtot = 0.0\n";
        return false;
    }
    double cg[3] = {dx / dtot, dy / dtot, dz / dtot};
    double dd[3] = {st.origin[0] - st.origin[0], st.origin[1] - st.origin[1], st.origin[2] - st.origin[2]};
    double pss[3], cs[3];
    mv(st.R, dd, pss);
    mv(st.R, cg, cs);
    double se[3] = {atan2(cs[0], cs[2]), asin(cs[1]), 0.0};
    double Rs[9];
    euler_matrix(se, Rs);
    DevSun &ds = out.sun;
    for (int k = 0; k < 3; k++) { ds.origin[k] = sun.origin[k]; ds.pos_sun_stage[k] = pss[k]; }
    for (int k = 0; k < 9; k++) ds.R[k] = Rs[k];
    double xs = 0.0, ys = 0.0;
    for (const HElement &e : st.elements) {
        if (!e.enabled) continue;
        double pl[3];
        mv(Rs, e.origin, pl);
        xs = xs + pl[0];
        ys = ys + pl[1];
    }
    double xcm = xs / st.elements.size(), ycm = ys / st.elements.size();
    (void)xcm; (void)ycm;
    double minx = 1.0e20, maxx = -1.0e20, miny = 1.0e20, maxy = -1.0e20;
    size_t n = 0;
    for (const HElement &e : st.elements) {
        if (!e.enabled) continue;
        double pl[3];
        mv(Rs, e.origin, pl);
        double xmn = pl[0], xmx = pl[0], ymn = pl[1], ymx = pl[1];
        const double A = e.p[0], B = e.p[1], C = e.p[2];
        double pad = 0.0;
        bool padded = true;
        switch (e.shape) {
        case 'c': case 'C': case 'h': case 'H': pad = A / 2.0; break;
        case 'r': case 'R': pad = sqrt(((A * A) + (B * B)) / 4.0); break;
        case 't': case 'T': pad = A / (2.0 * cos(30.0 * (kDegRad))); break;
        case 'a': case 'A': pad = B; break;
        case 'l': case 'L': {
            double xl = (fabs(B) >= fabs(A)) ? B : A;
            pad = sqrt((xl * xl) + 0.25 * C * C);
            break;
        }
        case 'i': case 'I': case 'q': case 'Q': {
            int nv = (e.shape == 'q' || e.shape == 'Q') ? 4 : 3;
            double rt = 0.0;
            for (int k = 0; k < nv; k++) {
                double r = sqrt((e.p[2 * k] * e.p[2 * k]) + (e.p[2 * k + 1] * e.p[2 * k + 1]));
                if (k == 0 || r > rt) rt = r;
            }
            pad = rt;
            break;
        }
        default: padded = false; break;
        }
        if (padded) { xmn = xmn - pad; ymn = ymn - pad; xmx = xmx + pad; ymx = ymx + pad; }
        if (xmn < minx) minx = xmn;
        if (xmx > maxx) maxx = xmx;
        if (ymn < miny) miny = ymn;
        if (ymx > maxy) maxy = ymx;
        n++;
    }
    if (n == 0) {
        msg = "error calculating sun position in primary stage because no elements were enabled";
        return false;
    }
    sun.minx = ds.minx = minx;
    sun.maxx = ds.maxx = maxx;
    sun.miny = ds.miny = miny;
    sun.maxy = ds.maxy = maxy;
    ds.shape = (unsigned char)sun.shape;
    ds.point_source = sun.point_source;
    ds.sigma = sun.sigma;
    ds.max_angle = sun.max_angle;
    ds.max_intensity = sun.max_intensity;
    ds.tab_n = (int)sun.angle.size();
    if (ds.tab_n > 0) {
        ds.tab = (int)out.aux.size();
        out.aux.insert(out.aux.end(), sun.angle.begin(), sun.angle.end());
        out.aux.insert(out.aux.end(), sun.intensity.begin(), sun.intensity.end());
        ds.tab_ascending = 1;
        for (int k = 1; k < ds.tab_n; k++)
            if (!(sun.angle[k] > sun.angle[k - 1])) ds.tab_ascending = 0;
        // Upper envelope of the interpolated profile / max intensity per
        // polar-angle bin over [0, max angle], so a uniform above it rejects
        // without the table walk.  Each bin is widened by 1e-3 of a bin on
        // both sides: the kernels find the bin from a float square root,
        // whose error is below 1e-4 of a bin.
        if (ds.tab_ascending && sun.max_angle > 0 && sun.max_intensity > 0) {
            const int NB = 256;
            const double ma = sun.max_angle, dth = ma / NB;
            auto prof = [&](double th) {  // the interpolation of coretrace/errors.cpp:167-178
                int i = 0;
                while (i < ds.tab_n - 1 && sun.angle[i] < th) i++;
                if (i == 0) return sun.intensity[0];
                return sun.intensity[i - 1] + (sun.intensity[i] - sun.intensity[i - 1]) * (th - sun.angle[i - 1]) /
                                                  (sun.angle[i] - sun.angle[i - 1]);
            };
            ds.env = (int)out.aux.size();
            ds.env_n = NB;
            ds.env_scale = NB / ma;
            for (int b = 0; b < NB; b++) {
                double lo = std::max(0.0, b * dth - 1e-3 * dth), hi = (b + 1) * dth + 1e-3 * dth;
                double m = std::max(prof(lo), prof(hi));
                for (int k = 0; k < ds.tab_n; k++)
                    if (sun.angle[k] > lo && sun.angle[k] < hi) m = std::max(m, sun.intensity[k]);
                double v = m / sun.max_intensity;
                out.aux.push_back((v * (1.0 + 1e-5) + 1e-7) * 4294967296.0);  // in units of 2^-32
            }
            // bins past the last one (theta up to sqrt(2) * max angle) fall
            // to the exact test
        }
        // Inverse-CDF segments of the same law (sunshape_icdf in
        // st_device.h): density I(theta) theta on [0, max angle], I the
        // interpolation of coretrace/errors.cpp:167-178 between mirrored
        // entries i-1 and i.  Built for ascending tables of non-negative
        // intensity; others keep the rejection loop.
        bool nonneg = true;
        for (double v : sun.intensity) nonneg = nonneg && v >= 0.0;
        if (ds.tab_ascending && nonneg && sun.max_angle > 0 && sun.max_intensity > 0) {
            std::vector<double> segs;
            double cum = 0.0;
            for (int i = 1; i < ds.tab_n; i++) {
                const double s0 = sun.angle[i - 1], s1 = sun.angle[i];
                const double lo = std::max(s0, 0.0), hi = std::min(s1, sun.max_angle);
                if (!(hi > lo)) continue;
                const double b = (sun.intensity[i] - sun.intensity[i - 1]) / (s1 - s0);
                const double a = sun.intensity[i - 1] + (sun.intensity[i] - sun.intensity[i - 1]) * (lo - s0) / (s1 - s0);
                const double w = hi - lo;
                const double p0 = a * lo, p1 = (a + b * lo) / 2.0, p2 = b / 3.0;
                segs.insert(segs.end(), {lo, w, a, b, p0, p1, p2, cum});
                cum = cum + w * (p0 + w * (p1 + w * p2));
            }
            if (!segs.empty() && cum > 0.0) {
                ds.icdf = (int)out.aux.size();
                ds.icdf_n = (int)(segs.size() / kIcdfStride);
                out.aux.insert(out.aux.end(), segs.begin(), segs.end());
                out.aux.push_back(cum);
            }
        }
    }

    // Sun-plane footprints of the stage-1 elements: each hit box projected
    // along the sun direction.  Its convex hull (DevFoot) screens candidates
    // and its bounding rectangle places the element in grid cells.  A
    // footprint wholly outside the sun-plane box is never reached, since sun
    // rays start inside that box.
    const DevStage &s0 = out.stages[0];
    struct Rect { int j; double x0, x1, y0, y1; };
    std::vector<Rect> rects;
    std::vector<double> sizes;
    out.foot.assign(2 * (size_t)std::max(1, s0.nelem), DevFoot{});
    out.grid_always.clear();
    for (DevFoot &f : out.foot) f.kind = 2;
    // every sun ray travels along the sun-frame z axis
    double sun_dir[3];
    {
        const double ez[3] = {0.0, 0.0, 1.0};
        mtv(Rs, ez, sun_dir);
    }
    for (int j = 0; j < s0.nelem; j++) {
        const DevElem &d = out.elems[s0.elem0 + j];
        if (!d.enabled) continue;
        const int nb = out.enbox[s0.elem0 + j];
        if (nb == 0) {
            out.foot[2 * j].kind = 1;
            out.grid_always.push_back(j);
            continue;
        }
        for (int b = 0; b < nb; b++) {
            const double *bx = &out.ebox[12 * (size_t)(s0.elem0 + j) + 6 * b];
            if (b == 1 && d.surf == SURF_SPHERE && d.vcx != 0.0) {
                // The far box of a sphere holds only exit points: a hit
                // there is the larger root t1 (the smaller one is taken only
                // at z <= zap, on the near branch), where the ray leaves the
                // sphere, so D . grad F = 2 (Kx x Dx + Ky y Dy + (z - 1/c) Dz)
                // >= 0.  A sun direction for which that is strongly negative
                // over the whole box never ends a sun ray there.
                double dl[3];
                mv(d.R, sun_dir, dl);
                const double zc = 1.0 / d.vcx, ky = d.stype == 7 ? 0.0 : 1.0;
                double g = -HUGE_VAL;
                for (int q = 0; q < 8; q++) {
                    const double x = (q & 1) ? bx[3] : bx[0], y = (q & 2) ? bx[4] : bx[1], z = (q & 4) ? bx[5] : bx[2];
                    g = std::max(g, x * dl[0] + ky * y * dl[1] + (z - zc) * dl[2]);
                }
                if (g < -0.01 * fabs(zc)) continue;
            }
            std::vector<std::pair<double, double>> pts;
            double x0 = HUGE_VAL, x1 = -HUGE_VAL, y0 = HUGE_VAL, y1 = -HUGE_VAL;
            for (int q = 0; q < 8; q++) {
                const double v[3] = {(q & 1) ? bx[3] : bx[0], (q & 2) ? bx[4] : bx[1], (q & 4) ? bx[5] : bx[2]};
                double w[3], c[3], u[3];
                mtv(d.R, v, w);
                for (int k = 0; k < 3; k++) c[k] = w[k] + d.org[k] - pss[k];
                mv(Rs, c, u);
                pts.emplace_back(u[0], u[1]);
                x0 = std::min(x0, u[0]); x1 = std::max(x1, u[0]);
                y0 = std::min(y0, u[1]); y1 = std::max(y1, u[1]);
            }
            // 1 mm, plus what rounding of sun-plane coordinates can reach
            const double pad = 1e-3 + 1e-9 * (fabs(x0) + fabs(x1) + fabs(y0) + fabs(y1));
            x0 -= pad; x1 += pad; y0 -= pad; y1 += pad;
            if (x1 < minx || x0 > maxx || y1 < miny || y0 > maxy) continue;
            if (!foot_hull(pts, pad, out.foot[2 * j + b])) out.foot[2 * j + b].kind = 1;
            if (b == 1) out.foot[2 * j].more = 1;
            rects.push_back({j, x0, x1, y0, y1});
            sizes.push_back(std::max(x1 - x0, y1 - y0));
        }
    }
    double med = 0.0;
    if (!sizes.empty()) {
        std::nth_element(sizes.begin(), sizes.begin() + sizes.size() / 2, sizes.end());
        med = sizes[sizes.size() / 2];
    }
    double w = maxx - minx, h = maxy - miny;
    double cell = std::max(med, std::max(w, h) / 4096.0);
    if (!(cell > 0)) cell = std::max(w, h) + 1.0;
    int nx = std::max(1, std::min(4096, (int)ceil(w / cell)));
    int ny = std::max(1, std::min(4096, (int)ceil(h / cell)));
    DevSunGrid &g = out.grid;
    g.x0 = minx;
    g.y0 = miny;
    g.nx = nx;
    g.ny = ny;
    g.inv_dx = nx / (w > 0 ? w : 1.0);
    g.inv_dy = ny / (h > 0 ? h : 1.0);
    std::vector<int32_t> cnt((size_t)nx * ny + 1, 0);
    auto cell_of = [](double v, int n) { return (int)std::max(0.0, std::min((double)(n - 1), floor(v))); };
    struct Span { int ax, bx, ay, by; };
    auto span = [&](const Rect &r) {
        return Span{cell_of((r.x0 - g.x0) * g.inv_dx, nx), cell_of((r.x1 - g.x0) * g.inv_dx, nx),
                    cell_of((r.y0 - g.y0) * g.inv_dy, ny), cell_of((r.y1 - g.y0) * g.inv_dy, ny)};
    };
    // each element enters a cell once: a second box skips the first's cells
    auto each_cell = [&](auto &&fn) {
        for (size_t k = 0; k < rects.size(); k++) {
            const Span s = span(rects[k]);
            const bool second = k > 0 && rects[k - 1].j == rects[k].j;
            const Span p = second ? span(rects[k - 1]) : Span{1, 0, 1, 0};
            for (int y = s.ay; y <= s.by; y++)
                for (int x = s.ax; x <= s.bx; x++) {
                    if (x >= p.ax && x <= p.bx && y >= p.ay && y <= p.by) continue;
                    fn((size_t)y * nx + x, rects[k].j);
                }
        }
    };
    each_cell([&](size_t c, int) { cnt[c + 1]++; });
    for (size_t c = 1; c < cnt.size(); c++) cnt[c] += cnt[c - 1];
    out.grid_items.assign(cnt.back(), 0);
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
    each_cell([&](size_t c, int j) { out.grid_items[fill[c]++] = j; });
    out.grid_cells.swap(cnt);
    out.use_grid = !sun.point_source;
    compile_beams(out, sun_dir);
    return true;
}

}  // namespace st
