// SolTrace .stinput reader and .rays writer over the public C API (see
// stinput.h for the reference functions each part follows).
#include "stinput.h"

#include <algorithm>
#include <cctype>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace st_io {

namespace {

// One line of at most 1023 characters with its "\n" and then "\r" removed
// (coretrace/strace/main.cpp:58-66).  End of file reads as an empty line.
struct Lines {
    FILE *fp;
    char buf[1024];
    const char *next() {
        if (!fgets(buf, sizeof buf, fp)) buf[0] = 0;
        size_t n = strlen(buf);
        if (n > 0 && buf[n - 1] == '\n') buf[--n] = 0;
        if (n > 0 && buf[n - 1] == '\r') buf[--n] = 0;
        return buf;
    }
};

// Fields between tabs, empty ones kept; text after a final tab is no field
// (the split() of coretrace/strace/main.cpp:18-52 with ret_empty set).
std::vector<std::string> tab_fields(const char *line) {
    std::vector<std::string> out;
    const char *p = line;
    while (*p) {
        const char *t = strchr(p, '\t');
        if (!t) {
            out.emplace_back(p);
            break;
        }
        out.emplace_back(p, (size_t)(t - p));
        p = t + 1;
    }
    return out;
}

void say(FILE *log, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void say(FILE *log, const char *fmt, ...) {
    if (!log) return;
    va_list ap;
    va_start(ap, fmt);
    vfprintf(log, fmt, ap);
    va_end(ap);
}

void print_messages(st_context_t cxt, FILE *log) {
    for (int j = 0; j < st_num_messages(cxt); j++) say(log, "error: %s\n", st_message(cxt, (st_uint_t)j));
}

// Two-column table rows "x y" (coretrace/strace/main.cpp:188-212).
void read_table(Lines &in, int n, std::vector<double> &x, std::vector<double> &y) {
    x.assign((size_t)std::max(n, 0), 0.0);
    y.assign((size_t)std::max(n, 0), 0.0);
    for (int i = 0; i < n; i++) sscanf(in.next(), "%lg %lg", &x[i], &y[i]);
}

bool read_sun(Lines &in, st_context_t cxt, FILE *log) {
    int ps = 0, count = 0, use_ldh = 0;
    char shape = 'g';
    double sigma = 0, halfwidth = 0;
    sscanf(in.next(), "SUN\tPTSRC\t%d\tSHAPE\t%c\tSIGMA\t%lg\tHALFWIDTH\t%lg", &ps, &shape, &sigma, &halfwidth);
    shape = (char)tolower((unsigned char)shape);
    st_sun(cxt, ps != 0 ? 1 : 0, shape, shape == 'g' ? sigma : halfwidth);

    double x = 0, y = 0, z = 0, lat = 0, day = 0, hour = 0;
    sscanf(in.next(), "XYZ\t%lg\t%lg\t%lg\tUSELDH\t%d\tLDH\t%lg\t%lg\t%lg", &x, &y, &z, &use_ldh, &lat, &day, &hour);
    if (use_ldh != 0) st_sun_position(cxt, lat, day, hour, &x, &y, &z);
    st_sun_xyz(cxt, x, y, z);
    say(log, "sun ps? %d cs: %c  %lg %lg %lg\n", ps != 0 ? 1 : 0, shape, x, y, z);

    sscanf(in.next(), "USER SHAPE DATA\t%d", &count);
    if (count > 0) {
        std::vector<double> angle((size_t)count), intensity((size_t)count);
        for (int i = 0; i < count; i++) sscanf(in.next(), "%lg\t%lg", &angle[i], &intensity[i]);
        st_sun_userdata(cxt, (st_uint_t)count, angle.data(), intensity.data());
    }
    return true;
}

// One face of an optical pair: an OPTICAL / OPTICAL.v2 line and its tables.
bool read_face(Lines &in, st_context_t cxt, int iopt, int fb, FILE *log) {
    const char *line = in.next();
    const std::vector<std::string> f = tab_fields(line);
    if (f.size() < 15) {
        say(log, "too few tokens for optical surface: %d\n\t>> %s\n", (int)f.size(), line);
        return false;
    }
    auto num = [&](size_t i) { return atof(f[i].c_str()); };
    auto inum = [&](size_t i) { return atoi(f[i].c_str()); };
    const char dist = f[1].empty() ? 'g' : f[1][0];
    double grating[4] = {num(11), num(12), num(13), num(14)};
    bool use_refl = false, use_trans = false;
    int n_refl = 0, n_trans = 0;
    if (f.size() >= 17) {
        use_refl = inum(15) > 0;
        n_refl = inum(16);
        if (f.size() >= 19) {
            use_trans = inum(17) > 0;
            n_trans = inum(18);
        }
    }
    std::vector<double> ra, rv, ta, tv;
    if (use_refl) read_table(in, n_refl, ra, rv);
    if (use_trans) read_table(in, n_trans, ta, tv);
    st_optic(cxt, (st_uint_t)iopt, fb, dist, inum(3), inum(2), inum(4), num(9), num(10), num(5), num(6), grating,
             num(7), num(8), use_refl ? 1 : 0, n_refl, use_refl ? ra.data() : nullptr, use_refl ? rv.data() : nullptr,
             use_trans ? 1 : 0, n_trans, use_trans ? ta.data() : nullptr, use_trans ? tv.data() : nullptr);
    return true;
}

bool read_optic(Lines &in, st_context_t cxt, FILE *log) {
    const char *line = in.next();
    if (strncmp(line, "OPTICAL PAIR", 12) != 0) return false;
    const int iopt = st_add_optic(cxt, line + 13);
    // a face with too few fields is reported and skipped, as strace does
    read_face(in, cxt, iopt, 1, log);
    read_face(in, cxt, iopt, 2, log);
    return true;
}

bool read_element(Lines &in, st_context_t cxt, int istage, FILE *log) {
    const st_uint_t s = (st_uint_t)istage;
    const st_uint_t e = (st_uint_t)st_add_element(cxt, s);
    const char *line = in.next();
    const std::vector<std::string> f = tab_fields(line);
    if (f.size() < 29) {
        say(log, "too few tokens for element: %d\n\t>> %s\n", (int)f.size(), line);
        return false;
    }
    auto num = [&](size_t i) { return atof(f[i].c_str()); };
    st_element_enabled(cxt, s, e, atoi(f[0].c_str()) ? 1 : 0);
    st_element_xyz(cxt, s, e, num(1), num(2), num(3));
    st_element_aim(cxt, s, e, num(4), num(5), num(6));
    st_element_zrot(cxt, s, e, num(7));
    if (!f[8].empty()) st_element_aperture(cxt, s, e, f[8][0]);
    double p[8];
    for (int i = 0; i < 8; i++) p[i] = num(9 + (size_t)i);
    st_element_aperture_params(cxt, s, e, p);
    if (!f[17].empty()) st_element_surface(cxt, s, e, f[17][0]);
    for (int i = 0; i < 8; i++) p[i] = num(18 + (size_t)i);
    if (!f[26].empty()) {
        if (st_element_surface_file(cxt, s, e, f[26].c_str()) < 0) {
            print_messages(cxt, log);
            return false;
        }
    } else {
        st_element_surface_params(cxt, s, e, p);
    }
    st_element_optic(cxt, s, e, f[27].c_str());
    st_element_interaction(cxt, s, e, atoi(f[28].c_str()));
    return true;
}

bool read_stage(Lines &in, st_context_t cxt, FILE *log) {
    int virt = 0, multi = 1, count = 0, through = 0;
    double x = 0, y = 0, z = 0, ax = 0, ay = 0, az = 0, zrot = 0;
    sscanf(in.next(),
           "STAGE\tXYZ\t%lg\t%lg\t%lg\tAIM\t%lg\t%lg\t%lg\tZROT\t%lg\tVIRTUAL\t%d\tMULTIHIT\t%d\tELEMENTS\t%d\t"
           "TRACETHROUGH\t%d",
           &x, &y, &z, &ax, &ay, &az, &zrot, &virt, &multi, &count, &through);
    const std::string name = in.next();
    const int istage = st_add_stage(cxt);
    const st_uint_t s = (st_uint_t)istage;
    st_stage_flags(cxt, s, virt, multi, through);
    st_stage_xyz(cxt, s, x, y, z);
    st_stage_aim(cxt, s, ax, ay, az);
    st_stage_zrot(cxt, s, zrot);
    say(log, "stage '%s': [%d] %lg %lg %lg   %lg %lg %lg   %lg   %d %d %d\n", name.c_str(), count, x, y, z, ax, ay, az,
        zrot, virt, multi, through);
    st_clear_elements(cxt, s);
    for (int i = 0; i < count; i++) {
        if (!read_element(in, cxt, istage, log)) {
            say(log, "error in element %d of %d in stage %d\n", i, count, istage);
            return false;
        }
    }
    return true;
}

}  // namespace

bool read_system(FILE *fp, st_context_t cxt, FILE *log) {
    if (!fp) return false;
    const int c = fgetc(fp);
    if (c != '#') {
        if (c != EOF) ungetc(c, fp);
        say(log, "input file must start with '#'\n");
        return false;
    }
    Lines in{fp, {0}};
    int vmaj = 0, vmin = 0, vmic = 0;
    sscanf(in.next(), " SOLTRACE VERSION %d.%d.%d INPUT FILE", &vmaj, &vmin, &vmic);
    say(log, "loading input file version %d.%d.%d\n", vmaj, vmin, vmic);
    if (!read_sun(in, cxt, log)) return false;

    int count = 0;
    sscanf(in.next(), "OPTICS LIST COUNT\t%d", &count);
    st_clear_optics(cxt);
    for (int i = 0; i < count; i++)
        if (!read_optic(in, cxt, log)) return false;

    count = 0;
    sscanf(in.next(), "STAGE LIST COUNT\t%d", &count);
    st_clear_stages(cxt);
    for (int i = 0; i < count; i++)
        if (!read_stage(in, cxt, log)) return false;
    return true;
}

bool write_rays(const char *path, st_context_t cxt) {
    FILE *out = fopen(path, "w");
    if (!out) return false;
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
    int nsun = 0;
    st_sun_stats(cxt, &xmin, &xmax, &ymin, &ymax, &nsun);
    const int n = std::max(0, st_num_intersections(cxt));
    std::vector<double> px((size_t)n), py((size_t)n), pz((size_t)n), cx((size_t)n), cy((size_t)n), cz((size_t)n);
    std::vector<int> elem((size_t)n), stage((size_t)n), ray((size_t)n);
    if (n > 0) {
        st_locations(cxt, px.data(), py.data(), pz.data());
        st_cosines(cxt, cx.data(), cy.data(), cz.data());
        st_elementmap(cxt, elem.data());
        st_stagemap(cxt, stage.data());
        st_raynumbers(cxt, ray.data());
    }
    // every value in a 15-wide column: reals as %.6e, maps and ray numbers as
    // integers (the iostream formats of coretrace/strace/main.cpp:440-470)
    fputs("#       SunXMin        SunXMax        SunYMin        SunYMax    SunRayCount         Length\n", out);
    const double head[6] = {xmin, xmax, ymin, ymax, (double)nsun, (double)n};
    for (double v : head) fprintf(out, "%15.6e", v);
    fputs("\n", out);
    fputs("#             X              Y              Z           Xcos           Ycos           Zcos       Stagemap"
          "     Elementmap      Raynumber\n",
          out);
    for (int i = 0; i < n; i++)
        fprintf(out, "%15.6e%15.6e%15.6e%15.6e%15.6e%15.6e%15d%15d%15d\n", px[i], py[i], pz[i], cx[i], cy[i], cz[i],
                stage[i], elem[i], ray[i]);
    const bool ok = fclose(out) == 0;
    return ok;
}

bool write_rays_binary(const char *path, st_context_t cxt) {
    FILE *out = fopen(path, "wb");
    if (!out) return false;
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
    int nsun = 0;
    st_sun_stats(cxt, &xmin, &xmax, &ymin, &ymax, &nsun);
    const int n = std::max(0, st_num_intersections(cxt));
    std::vector<double> px((size_t)n), py((size_t)n), pz((size_t)n), cx((size_t)n), cy((size_t)n), cz((size_t)n);
    std::vector<int> elem((size_t)n), stage((size_t)n), ray((size_t)n);
    if (n > 0) {
        st_locations(cxt, px.data(), py.data(), pz.data());
        st_cosines(cxt, cx.data(), cy.data(), cz.data());
        st_elementmap(cxt, elem.data());
        st_stagemap(cxt, stage.data());
        st_raynumbers(cxt, ray.data());
    }
    const double head[6] = {xmin, xmax, ymin, ymax, (double)nsun, (double)n};
    bool ok = fwrite(head, sizeof(double), 6, out) == 6;
    // records in blocks, each record nine doubles as the GUI writes them
    std::vector<double> buf;
    const int kBlock = 65536;
    for (int i0 = 0; ok && i0 < n; i0 += kBlock) {
        const int m = std::min(kBlock, n - i0);
        buf.resize((size_t)m * 9);
        for (int k = 0; k < m; k++) {
            const int i = i0 + k;
            double *r = &buf[(size_t)k * 9];
            r[0] = px[i]; r[1] = py[i]; r[2] = pz[i];
            r[3] = cx[i]; r[4] = cy[i]; r[5] = cz[i];
            r[6] = stage[i]; r[7] = elem[i]; r[8] = ray[i];
        }
        ok = fwrite(buf.data(), sizeof(double), buf.size(), out) == buf.size();
    }
    return fclose(out) == 0 && ok;
}

namespace {
int progress(st_uint_t, st_uint_t, st_uint_t, st_uint_t curstage, st_uint_t, void *data) {
    long *last = static_cast<long *>(data);
    if ((long)curstage != *last) {
        printf("Stage %2d:\n", (int)curstage);
        *last = (long)curstage;
    }
    return 1;
}
}  // namespace

int strace_main(int argc, char **argv, void (*prepare)(st_context_t cxt),
                void (*after)(st_context_t cxt, const char *file)) {
    if (argc < 7) {
        printf("strace: too few arguments.  usage:\n\t"
               "strace <system.stinput> <number of rays> <max rays> <seed> <sunshape? 0/1> <errors? 0/1> "
               "<pointfocus? 0/1>\n");
        return -1;
    }
    const char *file = argv[1];
    const int nrays = atoi(argv[2]), maxrays = atoi(argv[3]), seed = atoi(argv[4]);
    const int sunshape = atoi(argv[5]), errors = atoi(argv[6]);
    const int power_tower = argc > 7 ? atoi(argv[7]) : 0;

    FILE *fp = fopen(file, "r");
    if (!fp) {
        printf("failed to open system input file\n");
        return -1;
    }
    st_context_t cxt = st_create_context();
    printf("input file: %s\n", file);
    const bool ok = read_system(fp, cxt, stdout);
    fclose(fp);
    if (!ok) {
        printf("error in input file.\n");
        st_free_context(cxt);
        return -1;
    }
    if (prepare) prepare(cxt);
    st_sim_params(cxt, nrays, maxrays, power_tower);
    st_sim_errors(cxt, sunshape, errors);
    long last_stage = -1;
    const int code = st_sim_run(cxt, (unsigned int)seed, progress, &last_stage);
    if (code >= 0) {
        printf("trace finished with code %d\n", code);
        const std::string out = std::string(file) + ".rays";
        if (!write_rays(out.c_str(), cxt)) printf("failed to write %s\n", out.c_str());
        if (after) after(cxt, file);
    } else {
        print_messages(cxt, stdout);
    }
    st_free_context(cxt);
    return 0;
}

}  // namespace st_io
