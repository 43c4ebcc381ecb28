// Plain-array access to the reference build's st_sim_run_data, for tests
// only: the same two entry points the library and the oracle export
// (st_amd_sim_run_data, st_amd_saved_stage_data), implemented by calling the
// reference's own C++ function (coretrace/stapi.cpp:680-728) on vectors.
// Built by `make -C oracle ref` into oracle/_ref/ref_run_data.so, linked
// against oracle/_ref/coretrace_ref.so.
#include <map>
#include <mutex>
#include <utility>
#include <vector>

typedef unsigned long st_uint_t;
typedef void *st_context_t;
typedef int (*progress_cb)(st_uint_t, st_uint_t, st_uint_t, st_uint_t, st_uint_t, void *);
typedef std::vector<std::vector<double> > Rows;

// the reference's prototype (coretrace/stapi.cpp:680-685, C++ linkage)
int st_sim_run_data(st_context_t pcxt, unsigned int seed, Rows *data_s1, Rows *data_s2, bool save_st_data,
                    progress_cb callback, void *cbdata);

namespace {
std::mutex g_mu;
std::map<st_context_t, std::pair<std::vector<double>, std::vector<double> > > g_saved;

Rows rows_of(const double *v, st_uint_t n, size_t w) {
    Rows r;
    for (st_uint_t i = 0; i < n; i++) r.emplace_back(v + i * w, v + (i + 1) * w);
    return r;
}
}  // namespace

extern "C" __attribute__((visibility("default"))) int st_amd_sim_run_data(st_context_t p, unsigned int seed,
                                                                          const double *s1, st_uint_t n1,
                                                                          const double *s2, st_uint_t n2, int save,
                                                                          progress_cb cb, void *data) {
    Rows r1 = rows_of(s1, n1, 8), r2 = rows_of(s2, n2, 7);
    const int n = st_sim_run_data(p, seed, &r1, &r2, save != 0, cb, data);
    // the reference appends what it saves after the rows it was given
    std::vector<double> f1, f2;
    for (size_t i = n1; i < r1.size(); i++) f1.insert(f1.end(), r1[i].begin(), r1[i].end());
    for (size_t i = n2; i < r2.size(); i++) f2.insert(f2.end(), r2[i].begin(), r2[i].end());
    std::lock_guard<std::mutex> lk(g_mu);
    g_saved[p] = std::make_pair(f1, f2);
    return n;
}

extern "C" __attribute__((visibility("default"))) int st_amd_saved_stage_data(st_context_t p, int which, double *out) {
    if (which != 1 && which != 2) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    const std::vector<double> &v = which == 1 ? g_saved[p].first : g_saved[p].second;
    if (out)
        for (size_t i = 0; i < v.size(); i++) out[i] = v[i];
    return (int)(v.size() / (which == 1 ? 8 : 7));
}
