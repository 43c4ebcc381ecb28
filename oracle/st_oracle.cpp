// CPU oracle for SolTrace's coretrace trace loop (st_sim_run -> Trace).
//
// TEST INFRASTRUCTURE ONLY.  Built into oracle/_build/st_oracle.so by
// oracle/Makefile; loaded only by tests/, by the smoke check and by the
// CPU-baseline leg of bench.py, always as the checker.  The shipped library
// (soltrace_amd/coretrace_api.so) neither links nor loads it.
//
// It exports the reference's C API (coretrace/stapi.h:76-171) so the same
// scene-building code drives the oracle, the reference build in oracle/_ref
// and the HIP library, plus three oracle-only calls:
//   st_oracle_set_rng(ctx, mode)   0 = the reference's sequential libstdc++
//                                   stream, 1 = the per-ray Philox streams
//                                   the GPU kernels draw from;
//   st_oracle_ray_sun_index(ctx, out)  0-based sun-ray index of each ray;
//   st_oracle_flux(...)            the GUI flux map of one element;
//   st_oracle_flux_by_source(...)  that map split by each ray's stage-1
//                                   source element;
//   st_oracle_set_torus_start(ctx, on)  start torus solves where the HIP
//                                   library starts them for far origins.
// In mode 0 its records are checked bit for bit against oracle/_ref.
//
// Parity pins: RNG streams vs libstdc++ and Random123 vectors; full record
// streams vs oracle/_ref on the sample-like scenes of tests/; the trough
// intercept-factor known answers; closed forms (quadric roots, reflection
// law, Snell) in tests/test_oracle_math.py.
#include <chrono>
#include <ctime>

#include "st_oracle_optics.h"

namespace oracle {

typedef int (*progress_cb)(unsigned long, unsigned long, unsigned long, unsigned long, unsigned long, void *);

struct GlobalRay {
    double pos[3];
    double cos[3];
    uint32_t num;
    uint64_t sun_index;
};

// Closest-hit search over all elements of a stage
// (coretrace/raytrace.cpp:653-770, the brute-force branch).
struct HitState {
    bool hit = false;
    double last_path = 1e99;
    double pos_el[3] = {0, 0, 0}, cos_el[3] = {0, 0, 0}, df[3] = {0, 0, 0};
    double pos_st[3] = {0, 0, 0}, cos_st[3] = {0, 0, 0};
    int elem = 0;  // 1-based
    int back = 0;
};

// Test element j; keep it when it is the closest accepted hit so far
// (coretrace/raytrace.cpp:722-768).
static void test_element(const Element &e, size_t j, const double ps[3], const double cs[3], HitState &h) {
    double pe[3], ce[3];
    to_local(ps, cs, e.origin, e.R, pe, ce);
    pe[0] = pe[0] + 1.0e-5 * ce[0];
    pe[1] = pe[1] + 1.0e-5 * ce[1];
    pe[2] = pe[2] + 1.0e-5 * ce[2];
    double X[3], C[3], DF[3], path = 0.0;
    int err = 0, hit = 0, back = 0;
    element_intersection(e, pe, ce, X, C, DF, &path, &err, &hit, &back);
    if (!hit) return;
    if (path < h.last_path) {
        if (X[2] <= e.zap || e.surf == 'm' || e.surf == 'M' || e.surf == 'r' || e.surf == 'R') {
            h.hit = true;
            h.last_path = path;
            for (int k = 0; k < 3; k++) {
                h.pos_el[k] = X[k];
                h.cos_el[k] = C[k];
                h.df[k] = DF[k];
            }
            h.elem = (int)j + 1;
            to_ref(X, C, e.origin, e.RT, h.pos_st, h.cos_st);
            h.back = back;
        }
    }
}

static void closest_hit(const Stage &st, int si, const double ps[3], const double cs[3], HitState &h) {
    h.last_path = 1e99;
    h.hit = false;
    for (size_t j = 0; j < st.elements.size(); j++) {
        const Element &e = *st.elements[j];
        if (!e.enabled) continue;
        ST_OPS_TEST(si);
        test_element(e, j, ps, cs, h);
        ST_OPS_POP();
    }
}

// Optical efficiency lookup (coretrace/raytrace.cpp:885-946).
static bool test_value(const Element &e, const OpticFace *opt, const double cos_el[3], const double df[3],
                       double *tv) {
    const std::vector<TableEntry> *tab = nullptr;
    if (e.interaction == 1) {
        if (opt->use_trans_table) tab = &opt->trans_table;
        else { *tv = opt->transmissivity; return true; }
    } else if (e.interaction == 2) {
        if (opt->use_refl_table) tab = &opt->refl_table;
        else { *tv = opt->reflectivity; return true; }
    } else {
        return false;
    }
    int n = (int)tab->size();
    double u[3];
    u[0] = -df[0] / sqrt(dot3(df, df));
    u[1] = -df[1] / sqrt(dot3(df, df));
    u[2] = -df[2] / sqrt(dot3(df, df));
    double ia = acos(dot3(cos_el, u)) * 1000.;
    if (ia >= (*tab)[n - 1].angle) {
        *tv = (*tab)[n - 1].value;
    } else {
        int m = 0;
        while ((*tab)[m].angle < ia) m++;
        if (m == 0)
            *tv = (*tab)[m].value;
        else
            *tv = ((*tab)[m].value + (*tab)[m - 1].value) / 2.0;
    }
    return true;
}

// Per-ray state carried through the multi-hit loop of one stage.
struct RayState {
    double pg[3], cg[3];  // global
    double ps[3], cs[3];  // stage frame
    double cos_out_el[3] = {0, 0, 0};  // persists like the reference's CosRayOutElement
};

// The trace loop, coretrace/raytrace.cpp:137-1113, restated as structured
// code.  Returns false on error (message logged); *cancelled set when the
// callback asked to stop (then the partial records are kept).
static bool trace(System &sys, uint32_t seed, uint64_t N, uint64_t maxrays, bool inc_sun, bool inc_err,
                  progress_cb cb, void *cbdata, bool *cancelled) {
    *cancelled = false;
    sys.sunraycount = 0;
    sys.ray_sun_index.clear();
    if (N < 1) {
        sys.errlog("invalid number of rays: %d", (int)N);
        return false;
    }
    if (sys.stages.size() < 1) {
        sys.errlog("no stages defined.");
        return false;
    }
    Rng rng(sys.rng_mode, seed);
    double pos_sun_stage[3];
    if (!sun_to_primary_stage(sys, *sys.stages[0], pos_sun_stage)) return false;

    // stage 1 from saved rows when both row sets are given
    // (coretrace/raytrace.cpp:159-163)
    const bool load = sys.load_n1 > 0 && sys.load_n2 > 0;
    std::vector<GlobalRay> incoming;  // rays passed from the previous stage
    bool prev_has_rays = false;
    uint64_t last_prev = N;  // LastRayNumberInPreviousStage
    uint64_t traced_total = 0;
    uint64_t per_cb = 50;
    auto t0 = std::chrono::steady_clock::now();
    RayState rs;

    for (size_t si = 0; si < sys.stages.size(); si++) {
        if (si > 0 && !prev_has_rays) break;  // nothing reaches this or any later stage
        Stage &st = *sys.stages[si];
        std::vector<GlobalRay> outgoing;
        int64_t prev_idx = 0;  // PreviousStageDataArrayIndex
        size_t read = 0;
        uint64_t ray_number = 1;
        bool stage_done = false;
        if (si == 0 && load) {
            // coretrace/raytrace.cpp:544-578, generateray.cpp:165-219: the
            // records are the rows of s1, the rays entering stage 2 those of
            // s2; the fields convert as the reference's casts do
            for (uint64_t j = 0; j < sys.load_n1; j++) {
                const double *v = sys.load_s1 + 8 * j;
                Record r;
                for (int k = 0; k < 3; k++) { r.pos[k] = v[k]; r.cos[k] = v[k + 3]; }
                r.element = (int)(int64_t)v[6];  // double -> st_uint_t -> int
                r.stage = 1;
                r.raynum = (uint32_t)(uint64_t)(int64_t)v[7];
                st.recs.push_back(r);
            }
            for (uint64_t j = 0; j < sys.load_n2; j++) {
                const double *v = sys.load_s2 + 7 * j;
                GlobalRay g;
                for (int k = 0; k < 3; k++) { g.pos[k] = v[k]; g.cos[k] = v[k + 3]; }
                g.num = (uint32_t)(int)v[6];
                g.sun_index = (uint64_t)g.num - 1;  // the counter streams of a loaded ray
                outgoing.push_back(g);
            }
            prev_has_rays = true;
            prev_idx = (int64_t)sys.load_n2 - 1;
            sys.sunraycount = st.recs.back().raynum;
            stage_done = true;
        }

        while (!stage_done) {
            // ---- Label_StartRayLoop ----
            uint64_t mhc = 0;
            uint64_t sun_index = 0;
            if (si == 0) {
                ST_OPS_PHASE(15);  // the search of this sun ray, until it is settled
                sun_index = sys.sunraycount;
                rng.begin_ray(0, sun_index);
                double psun[3];
                generate_ray(rng, pos_sun_stage, st, sys.sun, rs.pg, rs.cg, psun);
                sys.sunraycount++;
                if (sys.sunraycount > maxrays) {
                    sys.errlog("generated sun rays reached maximum count: %d", (int)maxrays);
                    return false;
                }
            } else {
                ST_OPS_PHASE(4 + (int)si);
                const GlobalRay &g = incoming[read++];
                ray_number = g.num;
                sun_index = g.sun_index;
                for (int k = 0; k < 3; k++) {
                    rs.pg[k] = g.pos[k];
                    rs.cg[k] = g.cos[k];
                }
                rng.begin_ray((uint32_t)si, sun_index);
            }
            if (rng.mode() == 1) rs.cos_out_el[0] = rs.cos_out_el[1] = rs.cos_out_el[2] = 0.0;
            to_local(rs.pg, rs.cg, st.origin, st.R, rs.ps, rs.cs);

            if (cb != nullptr && traced_total++ % per_cb == 0) {
                if (traced_total > 1) {
                    double ms = 1e-3 * (double)std::chrono::duration_cast<std::chrono::microseconds>(
                        std::chrono::steady_clock::now() - t0).count();
                    double msper = ms / (double)traced_total;
                    per_cb = msper > 0 ? (uint64_t)(200.0 / msper) : 1000000;
                    if (per_cb < 5) per_cb = 5;
                }
                if (!cb(traced_total, ray_number, last_prev, si + 1, sys.stages.size(), cbdata)) {
                    *cancelled = true;
                    return true;
                }
            }

            // ---- Label_MultiHitLoop ----
            for (;;) {
                HitState h;
                closest_hit(st, (int)si, rs.ps, rs.cs, h);
                if (si == 0 && mhc == 0) ST_OPS_SETTLE(h.hit ? 1 : 0, 2);  // missed or hit; the rest is phase 2
                bool flag_miss = false;
                if (!h.hit) {
                    if (si == 0) {
                        if (mhc == 0) break;  // new sun ray, same ray number
                        GlobalRay g;
                        for (int k = 0; k < 3; k++) { g.pos[k] = rs.pg[k]; g.cos[k] = rs.cg[k]; }
                        g.num = (uint32_t)ray_number;
                        g.sun_index = sun_index;
                        outgoing.resize(prev_idx + 1);
                        outgoing[prev_idx] = g;
                        if (ray_number == N) { stage_done = true; break; }
                        prev_idx++;
                        prev_has_rays = true;
                        ray_number++;
                        break;
                    }
                    if (st.tracethrough || mhc > 0) {
                        GlobalRay g;
                        for (int k = 0; k < 3; k++) { g.pos[k] = rs.pg[k]; g.cos[k] = rs.cg[k]; }
                        g.num = (uint32_t)ray_number;
                        g.sun_index = sun_index;
                        outgoing.resize(prev_idx + 1);
                        outgoing[prev_idx] = g;
                        if (ray_number == last_prev) { stage_done = true; break; }
                        prev_idx++;
                        prev_has_rays = true;
                        if (mhc > 0) break;
                    }
                    flag_miss = true;  // Label_FlagMiss
                }
                Record r;
                if (flag_miss) {
                    for (int k = 0; k < 3; k++) { r.pos[k] = rs.ps[k]; r.cos[k] = rs.cs[k]; }
                    r.element = 0;
                } else {
                    for (int k = 0; k < 3; k++) { r.pos[k] = h.pos_st[k]; r.cos[k] = h.cos_st[k]; }
                    r.element = h.elem;
                }
                r.stage = (int)si + 1;
                r.raynum = (uint32_t)ray_number;
                st.recs.push_back(r);
                if (si == 0 && mhc == 0 && !flag_miss) sys.ray_sun_index.push_back(sun_index);

                if (flag_miss) {
                    if (ray_number == last_prev) {
                        if (!st.tracethrough) {
                            prev_has_rays = false;
                            if (prev_idx > 0) {
                                prev_has_rays = true;
                                prev_idx--;
                            }
                        }
                        stage_done = true;
                    }
                    break;
                }
                mhc++;

                const Element &el = *st.elements[h.elem - 1];
                double pos_out[3], cos_out[3];
                if (st.virt) {
                    for (int k = 0; k < 3; k++) { pos_out[k] = h.pos_el[k]; cos_out[k] = h.cos_el[k]; }
                } else {
                    const OpticFace *opt = h.back ? &el.optics->back : &el.optics->front;
                    double tv;
                    if (!test_value(el, opt, h.cos_el, h.df, &tv)) {
                        sys.errlog("Bad optical interaction type = %d (stage %d)", (int)si, el.interaction);
                        return false;
                    }
                    if (tv <= rng()) {
                        st.recs.back().element = -st.recs.back().element;
                        if (ray_number == last_prev) {
                            prev_has_rays = false;
                            if (prev_idx > 0) {
                                prev_idx--;
                                prev_has_rays = true;
                            }
                            stage_done = true;
                        } else if (si == 0) {
                            if (ray_number == N)
                                stage_done = true;
                            else
                                ray_number++;
                        }
                        break;
                    }
                    double cin[3] = {h.cos_el[0], h.cos_el[1], h.cos_el[2]};
                    double df[3] = {h.df[0], h.df[1], h.df[2]};
                    if (inc_sun && si == 0 && mhc == 1) {
                        double co[3];
                        errors(rng, cin, 1, sys.sun, el, opt, co, df);
                        for (int k = 0; k < 3; k++) cin[k] = co[k];
                    }
                    if (inc_err) {
                        double co[3];
                        surface_normal_errors(rng, df, opt, co);
                        for (int k = 0; k < 3; k++) df[k] = co[k];
                    }
                    for (int k = 0; k < 3; k++) pos_out[k] = h.pos_el[k];
                    interaction(rng, cin, df, el.interaction, opt, rs.cos_out_el);
                    if (inc_err) {
                        const double *base = (opt->dist == 'F' || opt->dist == 'f') ? df : rs.cos_out_el;
                        double b[3] = {base[0], base[1], base[2]};
                        double co[3];
                        errors(rng, b, 2, sys.sun, el, opt, co, df);
                        for (int k = 0; k < 3; k++) rs.cos_out_el[k] = co[k];
                    }
                    for (int k = 0; k < 3; k++) cos_out[k] = rs.cos_out_el[k];
                }
                if (st.virt)
                    for (int k = 0; k < 3; k++) rs.cos_out_el[k] = cos_out[k];
                to_ref(pos_out, cos_out, el.origin, el.RT, rs.ps, rs.cs);
                to_ref(rs.ps, rs.cs, st.origin, st.RT, rs.pg, rs.cg);
                if (!st.multihit) {
                    // Label_StageHitLogic with StageHit = false and at least one hit
                    GlobalRay g;
                    for (int k = 0; k < 3; k++) { g.pos[k] = rs.pg[k]; g.cos[k] = rs.cg[k]; }
                    g.num = (uint32_t)ray_number;
                    g.sun_index = sun_index;
                    outgoing.resize(prev_idx + 1);
                    outgoing[prev_idx] = g;
                    if (si == 0) {
                        if (ray_number == N) { stage_done = true; break; }
                        prev_idx++;
                        prev_has_rays = true;
                        ray_number++;
                    } else {
                        if (ray_number == last_prev) { stage_done = true; break; }
                        prev_idx++;
                        prev_has_rays = true;
                    }
                    break;
                }
            }
        }

        // ---- Label_EndStageLoop ----
        if (sys.save_data && si == 0) {
            // every stage-1 record (coretrace/raytrace.cpp:1044-1063)
            for (const Record &r : st.recs)
                sys.saved_s1.insert(sys.saved_s1.end(), {r.pos[0], r.pos[1], r.pos[2], r.cos[0], r.cos[1], r.cos[2],
                                                         (double)r.element, (double)r.raynum});
        }
        if (sys.save_data && si == 1) {
            // the reference's one ray array as stage 2 leaves it
            // (coretrace/raytrace.cpp:1065-1078): the first slots hold the
            // rays written out of stage 2, the rest still the rays read in
            for (size_t k = 0; k < read; k++) {
                const GlobalRay &g = k < outgoing.size() ? outgoing[k] : incoming[k];
                sys.saved_s2.insert(sys.saved_s2.end(),
                                    {g.pos[0], g.pos[1], g.pos[2], g.cos[0], g.cos[1], g.cos[2], (double)g.num});
            }
        }
        if (!prev_has_rays) {
            last_prev = 0;
            continue;
        }
        if (prev_idx < (int64_t)outgoing.size()) {
            last_prev = outgoing[prev_idx].num;
            if (last_prev == 0) {
                const uint32_t pp = prev_idx > 0 ? outgoing[prev_idx - 1].num : 0;
                sys.errlog("LastRayNumberInPreviousStage=0, stage %d, PrevIdx=%d, CurIdx=%d, pp=%d", (int)si + 1,
                           (int)prev_idx, (int)read, (int)pp);
                return false;
            }
        } else {
            sys.errlog("Invalid PreviousStageDataArrayIndex: %u, @ stage %d", (unsigned)prev_idx, (int)si + 1);
            return false;
        }
        outgoing.resize(prev_idx + 1);
        incoming.swap(outgoing);
    }
    return true;
}

// TRayData::Merge order (coretrace/types.cpp:311-355 as called from
// stapi.cpp:718-719): full 8192-record blocks first, partial blocks after.
static void merge_records(System &sys) {
    const size_t B = 8192;
    std::vector<Record> all;
    for (Stage *st : sys.stages) {
        std::vector<Record> &src = st->recs;
        size_t fa = (all.size() / B) * B, fs = (src.size() / B) * B;
        std::vector<Record> out;
        out.reserve(all.size() + src.size());
        out.insert(out.end(), all.begin(), all.begin() + fa);
        out.insert(out.end(), src.begin(), src.begin() + fs);
        out.insert(out.end(), all.begin() + fa, all.end());
        out.insert(out.end(), src.begin() + fs, src.end());
        all.swap(out);
        src.clear();
    }
    sys.all.swap(all);
}

}  // namespace oracle

using namespace oracle;

#define SYS(p, r)                                         \
    System *sys = reinterpret_cast<System *>(p);          \
    if (!sys) return r;
#define STG(i)                                            \
    if ((i) >= sys->stages.size()) return -1;             \
    Stage *s = sys->stages[i];
#define ELM(i)                                            \
    if ((i) >= s->elements.size()) return -1;             \
    Element *e = s->elements[i];

extern "C" {
typedef unsigned long st_uint_t;
typedef void *st_context_t;

st_context_t st_create_context() { return reinterpret_cast<st_context_t>(new System); }
int st_free_context(st_context_t p) { SYS(p, -1); delete sys; return 1; }
int st_num_messages(st_context_t p) { SYS(p, -1); return (int)sys->messages.size(); }
const char *st_message(st_context_t p, st_uint_t idx) {
    SYS(p, nullptr);
    return idx < sys->messages.size() ? sys->messages[idx].c_str() : nullptr;
}
void st_clear_messages(st_context_t p) { System *sys = reinterpret_cast<System *>(p); if (sys) sys->messages.clear(); }

int st_num_optics(st_context_t p) { SYS(p, -1); return (int)sys->optics.size(); }
int st_add_optic(st_context_t p, const char *name) {
    SYS(p, -1);
    OpticSet *o = new OpticSet;
    o->name = name;
    sys->optics.push_back(o);
    return (int)sys->optics.size() - 1;
}
int st_delete_optic(st_context_t p, st_uint_t idx) {
    SYS(p, -1);
    if (idx >= sys->optics.size()) return -1;
    delete sys->optics[idx];
    sys->optics.erase(sys->optics.begin() + idx);
    return 1;
}
int st_clear_optics(st_context_t p) {
    SYS(p, -1);
    for (OpticSet *o : sys->optics) delete o;
    sys->optics.clear();
    return 1;
}
int st_optic(st_context_t p, st_uint_t idx, int fb, char dist, int optnum, int apgr, int order, double rreal,
             double rimag, double ref, double tra, double gratingab12[3], double rmsslope, double rmsspec,
             int userefltable, int refl_npoints, double *refl_angles, double *refls, int usetranstable,
             int trans_npoints, double *trans_angles, double *transs) {
    SYS(p, -1);
    if (idx >= sys->optics.size()) return -1;
    OpticFace *o = (fb == 2) ? &sys->optics[idx]->back : &sys->optics[idx]->front;
    o->dist = dist;
    o->optnum = optnum;
    o->apgr = apgr;
    o->order = order;
    o->refr[0] = rreal;
    o->refr[1] = rimag;
    o->reflectivity = ref;
    o->transmissivity = tra;
    for (int i = 0; i < 4; i++) o->ab12[i] = gratingab12[i];
    o->slope = rmsslope;
    o->spec = rmsspec;
    if (userefltable && refl_npoints > 0 && refl_angles && refls) {
        o->use_refl_table = true;
        o->refl_table.resize(refl_npoints);
        for (int i = 0; i < refl_npoints; i++) o->refl_table[i] = {refl_angles[i], refls[i]};
    }
    if (usetranstable && trans_npoints > 0 && trans_angles && transs) {
        o->use_trans_table = true;
        o->trans_table.resize(trans_npoints);
        for (int i = 0; i < trans_npoints; i++) o->trans_table[i] = {trans_angles[i], transs[i]};
    }
    return 1;
}

int st_num_stages(st_context_t p) { SYS(p, -1); return (int)sys->stages.size(); }
int st_add_stage(st_context_t p) { SYS(p, -1); sys->stages.push_back(new Stage); return (int)sys->stages.size() - 1; }
int st_add_stages(st_context_t p, st_uint_t num) {
    SYS(p, -1);
    for (st_uint_t i = 0; i < num; i++) sys->stages.push_back(new Stage);
    return (int)num;
}
int st_delete_stage(st_context_t p, st_uint_t idx) {
    SYS(p, -1);
    if (idx >= sys->stages.size()) return -1;
    delete sys->stages[idx];
    sys->stages.erase(sys->stages.begin() + idx);
    return 1;
}
int st_clear_stages(st_context_t p) {
    SYS(p, -1);
    for (Stage *s : sys->stages) delete s;
    sys->stages.clear();
    return 1;
}
int st_stage_xyz(st_context_t p, st_uint_t i, double x, double y, double z) {
    SYS(p, -1); STG(i); s->origin[0] = x; s->origin[1] = y; s->origin[2] = z; return 1;
}
int st_stage_aim(st_context_t p, st_uint_t i, double x, double y, double z) {
    SYS(p, -1); STG(i); s->aim[0] = x; s->aim[1] = y; s->aim[2] = z; return 1;
}
int st_stage_zrot(st_context_t p, st_uint_t i, double z) { SYS(p, -1); STG(i); s->zrot = z; return 1; }
int st_stage_flags(st_context_t p, st_uint_t i, int virt, int multihit, int tracethrough) {
    SYS(p, -1); STG(i);
    s->virt = virt != 0; s->multihit = multihit != 0; s->tracethrough = tracethrough != 0;
    return 1;
}
int st_num_elements(st_context_t p, st_uint_t i) { SYS(p, -1); STG(i); return (int)s->elements.size(); }
int st_add_element(st_context_t p, st_uint_t i) {
    SYS(p, -1); STG(i); s->elements.push_back(new Element); return (int)s->elements.size() - 1;
}
int st_add_elements(st_context_t p, st_uint_t i, st_uint_t num) {
    SYS(p, -1);
    if (num < 1) return -1;
    STG(i);
    for (st_uint_t k = 0; k < num; k++) s->elements.push_back(new Element);
    return (int)num;
}
int st_delete_element(st_context_t p, st_uint_t i, st_uint_t idx) {
    SYS(p, -1); STG(i); ELM(idx);
    delete e;
    s->elements.erase(s->elements.begin() + idx);
    return 1;
}
int st_clear_elements(st_context_t p, st_uint_t i) {
    SYS(p, -1); STG(i);
    for (Element *e : s->elements) delete e;
    s->elements.clear();
    return 1;
}
int st_element_enabled(st_context_t p, st_uint_t i, st_uint_t j, int en) {
    SYS(p, -1); STG(i); ELM(j); e->enabled = en != 0; return 1;
}
int st_element_xyz(st_context_t p, st_uint_t i, st_uint_t j, double x, double y, double z) {
    SYS(p, -1); STG(i); ELM(j); e->origin[0] = x; e->origin[1] = y; e->origin[2] = z; return 1;
}
int st_element_aim(st_context_t p, st_uint_t i, st_uint_t j, double x, double y, double z) {
    SYS(p, -1); STG(i); ELM(j); e->aim[0] = x; e->aim[1] = y; e->aim[2] = z; return 1;
}
int st_element_zrot(st_context_t p, st_uint_t i, st_uint_t j, double z) { SYS(p, -1); STG(i); ELM(j); e->zrot = z; return 1; }
int st_element_aperture(st_context_t p, st_uint_t i, st_uint_t j, char ap) {
    SYS(p, -1); STG(i); ELM(j); e->shape = ap; return 1;
}
int st_element_aperture_params(st_context_t p, st_uint_t i, st_uint_t j, double params[8]) {
    SYS(p, -1); STG(i); ELM(j);
    for (int k = 0; k < 8; k++) e->p[k] = params[k];
    return 1;
}
int st_element_surface(st_context_t p, st_uint_t i, st_uint_t j, char surf) {
    SYS(p, -1); STG(i); ELM(j); e->surf = surf; return 1;
}
int st_element_surface_params(st_context_t p, st_uint_t i, st_uint_t j, double params[8]) {
    SYS(p, -1); STG(i); ELM(j);
    return translate_surface_params(*sys, *e, params) ? 1 : -1;
}
int st_element_surface_file(st_context_t p, st_uint_t i, st_uint_t j, const char *file) {
    SYS(p, -1); STG(i); ELM(j);
    return read_surface_file(file, *e, *sys) ? 1 : -1;
}
int st_element_interaction(st_context_t p, st_uint_t i, st_uint_t j, int type) {
    SYS(p, -1); STG(i); ELM(j); e->interaction = type; return 1;
}
int st_element_optic(st_context_t p, st_uint_t i, st_uint_t j, const char *name) {
    SYS(p, -1); STG(i); ELM(j); e->optic_name = name; return 1;
}

int st_sun(st_context_t p, int point_source, char shape, double sigma) {
    SYS(p, -1);
    sys->sun.point_source = point_source != 0;
    sys->sun.shape = shape;
    sys->sun.sigma = sigma;
    return 1;
}
int st_sun_xyz(st_context_t p, double x, double y, double z) {
    SYS(p, -1); sys->sun.origin[0] = x; sys->sun.origin[1] = y; sys->sun.origin[2] = z; return 1;
}
// coretrace/stapi.cpp:497-527
int st_sun_position(st_context_t p, double lat, double day, double hour, double *x, double *y, double *z) {
    (void)p;
    double decl = 180 / M_PI * asin(0.39795 * cos(0.98563 * M_PI / 180 * (day - 173)));
    double ha = 15 * (hour - 12);
    double el = 180 / M_PI *
                asin(sin(decl * M_PI / 180) * sin(lat * M_PI / 180) +
                     cos(decl * M_PI / 180) * cos(ha * M_PI / 180) * cos(lat * M_PI / 180));
    double az = 180 / M_PI *
                acos((sin(M_PI / 180 * decl) * cos(M_PI / 180 * lat) -
                      cos(M_PI / 180 * decl) * sin(M_PI / 180 * lat) * cos(M_PI / 180 * ha)) /
                         cos(M_PI / 180 * el) +
                     0.0000000001);
    if (sin(ha * M_PI / 180) > 0.0) az = 360 - az;
    *x = -sin(az * M_PI / 180) * cos(el * M_PI / 180);
    *y = sin(el * M_PI / 180);
    *z = cos(az * M_PI / 180) * cos(el * M_PI / 180);
    return 1;
}
// coretrace/stapi.cpp:529-560
int st_sun_userdata(st_context_t p, st_uint_t n, double angle[], double intensity[]) {
    SYS(p, -1);
    if (n <= 2) return -1;
    Sun &s = sys->sun;
    s.max_angle = 0;
    s.max_intensity = 0;
    s.angle.assign(2 * n - 1, 0.0);
    s.intensity.assign(2 * n - 1, 0.0);
    for (st_uint_t i = 0; i < n; i++) {
        s.angle[n + i - 1] = angle[i];
        s.intensity[n + i - 1] = intensity[i];
        if (angle[i] > s.max_angle) s.max_angle = angle[i];
        if (intensity[i] > s.max_intensity) s.max_intensity = intensity[i];
    }
    for (st_uint_t i = 0; i < n - 1; i++) {
        s.angle[i] = -angle[n - i - 1];
        s.intensity[i] = intensity[n - i - 1];
    }
    sun_icdf_build(s);
    return (int)n;
}

int st_num_intersections(st_context_t p) { SYS(p, -1); return (int)sys->all.size(); }
int st_locations(st_context_t p, double *x, double *y, double *z) {
    SYS(p, -1);
    for (size_t i = 0; i < sys->all.size(); i++) {
        x[i] = sys->all[i].pos[0]; y[i] = sys->all[i].pos[1]; z[i] = sys->all[i].pos[2];
    }
    return 1;
}
int st_cosines(st_context_t p, double *x, double *y, double *z) {
    SYS(p, -1);
    for (size_t i = 0; i < sys->all.size(); i++) {
        x[i] = sys->all[i].cos[0]; y[i] = sys->all[i].cos[1]; z[i] = sys->all[i].cos[2];
    }
    return 1;
}
int st_elementmap(st_context_t p, int *m) { SYS(p, -1); for (size_t i = 0; i < sys->all.size(); i++) m[i] = sys->all[i].element; return 1; }
int st_stagemap(st_context_t p, int *m) { SYS(p, -1); for (size_t i = 0; i < sys->all.size(); i++) m[i] = sys->all[i].stage; return 1; }
int st_raynumbers(st_context_t p, int *m) { SYS(p, -1); for (size_t i = 0; i < sys->all.size(); i++) m[i] = (int)sys->all[i].raynum; return 1; }
int st_sun_stats(st_context_t p, double *xmin, double *xmax, double *ymin, double *ymax, int *nsun) {
    SYS(p, -1);
    if (xmin) *xmin = sys->sun.minx;
    if (xmax) *xmax = sys->sun.maxx;
    if (ymin) *ymin = sys->sun.miny;
    if (ymax) *ymax = sys->sun.maxy;
    if (nsun) *nsun = (int)sys->sunraycount;
    return 1;
}
int st_sim_params(st_context_t p, int raycount, int maxcount, int dyn) {
    SYS(p, -1); sys->raycount = raycount; sys->raymax = maxcount; sys->dyn_group = dyn != 0; return 1;
}
int st_sim_errors(st_context_t p, int sun, int optics) {
    SYS(p, -1); sys->err_sunshape = sun != 0; sys->err_optical = optics != 0; return 1;
}
int st_sim_run(st_context_t p, unsigned int seed, progress_cb cb, void *data);
// coretrace/stapi.cpp:680-728, with plain arrays for the row vectors
int st_amd_sim_run_data(st_context_t p, unsigned int seed, const double *s1, st_uint_t n1, const double *s2,
                        st_uint_t n2, int save, progress_cb cb, void *data) {
    SYS(p, -1);
    if ((n1 && !s1) || (n2 && !s2)) return -1;
    sys->load_s1 = s1;
    sys->load_n1 = n1;
    sys->load_s2 = s2;
    sys->load_n2 = n2;
    sys->save_data = save != 0;
    sys->saved_s1.clear();
    sys->saved_s2.clear();
    const int r = st_sim_run(p, seed, cb, data);
    sys->load_s1 = sys->load_s2 = nullptr;
    sys->load_n1 = sys->load_n2 = 0;
    sys->save_data = false;
    return r;
}
int st_amd_saved_stage_data(st_context_t p, int which, double *out) {
    SYS(p, -1);
    if (which != 1 && which != 2) return -1;
    const std::vector<double> &v = which == 1 ? sys->saved_s1 : sys->saved_s2;
    if (out) std::copy(v.begin(), v.end(), out);
    return (int)(v.size() / (which == 1 ? 8 : 7));
}
// coretrace/stapi.cpp:680-734 (st_sim_run is st_sim_run_data without rows)
int st_sim_run(st_context_t p, unsigned int seed, progress_cb cb, void *data) {
    SYS(p, -1);
    sys->all.clear();
    for (Stage *s : sys->stages) s->recs.clear();
    if (!init_geometries(*sys)) return -1;
    for (Stage *s : sys->stages)
        for (Element *e : s->elements) {
            e->torus_start_r = 0.0;
            if (sys->torus_start && e->stype == 10) torus_start_sphere(*e);
        }
    bool cancelled = false;
    try {
        // the ray count is the number of loaded stage-2 rays when there are
        // any (coretrace/stapi.cpp:697-700)
        const int rayct = sys->load_n2 > 0 ? (int)sys->load_n2 : sys->raycount;
        if (!trace(*sys, seed, (uint64_t)(int64_t)rayct, (uint64_t)(int64_t)sys->raymax, sys->err_sunshape,
                   sys->err_optical, cb, data, &cancelled))
            return -1;
    } catch (const FeFailure &) {
        // the reference throws out of Trace here (coretrace/vshot.cpp:177,
        // raytrace.cpp:1108-1112)
        sys->errlog("trace error: Interpolation failed to find nearby nodes to ray hit in FEA surface.");
        return -1;
    }
    merge_records(*sys);
    return (int)sys->all.size();
}

void st_calc_euler_angles(double origin[3], double aimpoint[3], double zrot, double euler[3]) {
    euler[0] = euler[1] = euler[2] = 0;
    double e[3];
    if (euler_from_aim(origin, aimpoint, zrot, e))
        for (int i = 0; i < 3; i++) euler[i] = e[i];
}
// The exported transforms write their outputs in coretrace/mathproc.cpp:
// 115-173's order (no copy of the direction), so aliased arguments give the
// reference's result.
void st_transform_to_local(double posref[3], double cosref[3], double origin[3], double R[3][3], double posloc[3],
                           double cosloc[3]) {
    double d[3] = {posref[0] - origin[0], posref[1] - origin[1], posref[2] - origin[2]};
    mat_vec(R, d, posloc);
    mat_vec(R, cosref, cosloc);
}
void st_transform_to_reference(double posloc[3], double cosloc[3], double origin[3], double RT[3][3],
                               double posref[3], double cosref[3]) {
    double t[3];
    mat_vec(RT, posloc, t);
    mat_vec(RT, cosloc, cosref);
    for (int i = 0; i < 3; i++) posref[i] = t[i] + origin[i];
}
void st_matrix_vector_mult(double m[3][3], double v[3], double mxv[3]) { mat_vec(m, v, mxv); }
void st_calc_transform_matrices(double euler[3], double R[3][3], double RT[3][3]) { euler_matrices(euler, R, RT); }
void st_matrix_transpose(double in[3][3], double out[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i][j] = in[j][i];
}

// ---- oracle-only entry points ----
int st_oracle_set_rng(st_context_t p, int mode) { SYS(p, -1); sys->rng_mode = mode; return 1; }
int st_oracle_set_torus_start(st_context_t p, int on) { SYS(p, -1); sys->torus_start = on != 0; return 1; }
int st_oracle_ray_sun_index(st_context_t p, uint64_t *out) {
    SYS(p, -1);
    for (size_t i = 0; i < sys->ray_sun_index.size(); i++) out[i] = sys->ray_sun_index[i];
    return (int)sys->ray_sun_index.size();
}
// Draw checks for the RNG restatements.
void st_oracle_ref_uniforms(uint32_t seed, int n, double *out) {
    RefStream s(seed);
    for (int i = 0; i < n; i++) out[i] = s.uniform();
}
void st_oracle_ref_normals(uint32_t seed, int n, double *out) {
    RefStream s(seed);
    for (int i = 0; i < n; i++) out[i] = s.normal(0.0, 1.0);
}
void st_oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    Philox4 b = philox4x32_10(ctr, key);
    for (int i = 0; i < 4; i++) out[i] = b.v[i];
}
double st_oracle_counter_uniform(uint32_t seed, uint32_t stage, uint64_t id, uint64_t k) {
    return counter_uniform(seed, stage, id, k);
}

// GUI flux map of one flat or cylindrical element,
// app/src/project.cpp:1139-1436 (ElementStatistics::Compute / BinRaysXY).
// Bins the LAST record of each ray on (stage, element), absorbed-only when
// asked; counts[ix*ny+iy].  Ranges are taken as given unless autoscale.
// src_counts (may be NULL): the same binning split by source, the stage-1
// element of the ray's last stage-1 record, src_counts[(src-1)*nx*ny + bin]
// for every stage-1 element (the library's st_amd_flux_by_source maps).
static int oracle_flux(st_context_t p, int stage_idx, int elem_idx, int nx, int ny, int absorbed_only, int autoscale,
                       double *minx, double *maxx, double *miny, double *maxy, double *counts, double *stats,
                       double *src_counts) {
    SYS(p, -1);
    if (stage_idx < 0 || stage_idx >= (int)sys->stages.size()) return -1;
    Stage &st = *sys->stages[stage_idx];
    if (elem_idx < 0 || elem_idx >= (int)st.elements.size()) return -1;
    Element &e = *st.elements[elem_idx];
    char surf = (char)tolower(e.surf), aper = (char)tolower(e.shape);
    if (surf != 't' && surf != 'f' && !(aper == 'l' && surf == 's')) return -1;
    const std::vector<Record> &rd = sys->all;
    if (autoscale) {
        *minx = *miny = 1e199;
        *maxx = *maxy = -1e199;
        for (size_t i = 0; i < rd.size(); i++) {
            if (std::abs(rd[i].element) == elem_idx + 1 && rd[i].stage == stage_idx + 1) {
                double pe[3], ce[3];
                to_local(rd[i].pos, rd[i].cos, e.origin, e.R, pe, ce);
                if (pe[0] < *minx) *minx = pe[0];
                if (pe[0] > *maxx) *maxx = pe[0];
                if (pe[1] < *miny) *miny = pe[1];
                if (pe[1] > *maxy) *maxy = pe[1];
            }
        }
    }
    if (nx <= 1 || ny <= 1 || *maxx <= *minx || *maxy <= *miny) return -1;
    double gx = *maxx - *minx, gy = *maxy - *miny;
    double radius = 0.0;
    if (surf == 't' || surf == 's') {
        radius = 1.0 / (surf == 't' ? e.curv_rev : e.vcx);
        if (surf == 't') {
            *minx = -M_PI * radius;
            *maxx = M_PI * radius;
            gx = 2.0 * M_PI * radius;
        } else {
            double amin = asin(e.p[0] / radius), amax = asin(e.p[1] / radius);
            *minx = radius * amin;
            *maxx = radius * amax;
            gx = (amax - amin) * radius;
        }
    }
    double bx = gx / nx, by = gy / ny;
    for (int i = 0; i < nx * ny; i++) counts[i] = 0.0;
    std::vector<int> src_of;  // ray number -> source element (1-based), from the stage-1 records in order
    if (src_counts) {
        const size_t nsrc = sys->stages[0]->elements.size();
        for (size_t i = 0; i < nsrc * (size_t)nx * ny; i++) src_counts[i] = 0.0;
        for (const Record &r : rd)
            if (r.stage == 1) {
                if (r.raynum >= src_of.size()) src_of.resize((size_t)r.raynum + 1, 0);
                src_of[r.raynum] = std::abs(r.element);
            }
    }
    std::vector<size_t> idx;
    for (size_t i = 0; i < rd.size(); i++)
        if (std::abs(rd[i].element) == elem_idx + 1 && rd[i].stage == stage_idx + 1) idx.push_back(i);
    bool cyl = surf == 't', pcyl = (surf == 's' && aper == 'l');
    size_t binned = 0;
    for (size_t j = 0; j < idx.size(); j++) {
        if (!(j == idx.size() - 1 || rd[idx[j]].raynum != rd[idx[j + 1]].raynum)) continue;
        const Record &r = rd[idx[j]];
        if (absorbed_only && r.element >= 0) continue;
        double pe[3], ce[3];
        to_local(r.pos, r.cos, e.origin, e.R, pe, ce);
        double x = pe[0], y = pe[1], z = pe[2];
        if (cyl) {
            if (z <= radius)
                x = radius * asin(x / radius);
            else if (z > radius) {
                if (x < 0) x = -(M_PI * radius / 2.0 + radius * acos(fabs(x) / radius));
                if (x >= 0) x = M_PI * radius / 2.0 + radius * acos(x / radius);
            }
        } else if (pcyl) {
            x = radius * asin(x / radius);
        }
        int gxi = -1, gyi = -1;
        while ((*minx + (gxi + 1) * bx) < x) gxi++;
        while ((*miny + (gyi + 1) * by) < y) gyi++;
        if (gxi >= 0 && gxi < nx && gyi >= 0 && gyi < ny) {
            counts[gxi * ny + gyi] += 1;
            binned++;
            if (src_counts && r.raynum < src_of.size() && src_of[r.raynum] > 0)
                src_counts[(size_t)(src_of[r.raynum] - 1) * nx * ny + gxi * ny + gyi] += 1;
        }
    }
    if (stats) {
        stats[0] = bx;
        stats[1] = by;
        stats[2] = (double)binned;
    }
    return 1;
}
int st_oracle_flux(st_context_t p, int stage_idx, int elem_idx, int nx, int ny, int absorbed_only, int autoscale,
                   double *minx, double *maxx, double *miny, double *maxy, double *counts, double *stats) {
    return oracle_flux(p, stage_idx, elem_idx, nx, ny, absorbed_only, autoscale, minx, maxx, miny, maxy, counts,
                       stats, nullptr);
}
// Per-source maps of a flux element on stage 2 or later, fixed ranges.
int st_oracle_flux_by_source(st_context_t p, int stage_idx, int elem_idx, int nx, int ny, int absorbed_only,
                             double minx, double maxx, double miny, double maxy, double *counts,
                             double *src_counts) {
    if (stage_idx < 1) return -1;
    return oracle_flux(p, stage_idx, elem_idx, nx, ny, absorbed_only, 0, &minx, &maxx, &miny, &maxy, counts,
                       nullptr, src_counts);
}

// The GUI's element statistics in full (app/src/project.cpp:1139-1436):
// the binning above, then the centroid and z value of BinRaysXY
// (:1340-1436) and the flux statistics of Compute (:1248-1288), serially in
// record order.  Same signature and stats[15] layout as the library's
// st_amd_element_stats.
int st_oracle_element_stats(st_context_t p, int stage_idx, int elem_idx, int nx, int ny, int autoscale,
                            int absorbed_only, double dni, double bounds[4], unsigned int *grid, double *stats) {
    SYS(p, -1);
    if (stage_idx < 0 || stage_idx >= (int)sys->stages.size()) return -1;
    Stage &st = *sys->stages[stage_idx];
    if (elem_idx < 0 || elem_idx >= (int)st.elements.size()) return -1;
    Element &e = *st.elements[elem_idx];
    char surf = (char)tolower(e.surf), aper = (char)tolower(e.shape);
    if (surf != 't' && surf != 'f' && !(aper == 'l' && surf == 's')) return -1;
    const std::vector<Record> &rd = sys->all;
    double minx = bounds[0], maxx = bounds[1], miny = bounds[2], maxy = bounds[3];
    if (autoscale) {
        minx = miny = 1e199;
        maxx = maxy = -1e199;
        for (size_t i = 0; i < rd.size(); i++) {
            if (std::abs(rd[i].element) == elem_idx + 1 && rd[i].stage == stage_idx + 1) {
                double pe[3], ce[3];
                to_local(rd[i].pos, rd[i].cos, e.origin, e.R, pe, ce);
                if (pe[0] < minx) minx = pe[0];
                if (pe[0] > maxx) maxx = pe[0];
                if (pe[1] < miny) miny = pe[1];
                if (pe[1] > maxy) maxy = pe[1];
            }
        }
    }
    if (nx <= 1 || ny <= 1 || maxx <= minx || maxy <= miny) return -1;
    double gx = maxx - minx, gy = maxy - miny;
    double radius = 0.0;
    if (surf == 't' || surf == 's') {
        radius = 1.0 / (surf == 't' ? e.curv_rev : e.vcx);
        if (surf == 't') {
            minx = -M_PI * radius;
            maxx = M_PI * radius;
            gx = 2.0 * M_PI * radius;
        } else {
            double amin = asin(e.p[0] / radius), amax = asin(e.p[1] / radius);
            minx = radius * amin;
            maxx = radius * amax;
            gx = (amax - amin) * radius;
        }
    }
    const double bx = gx / nx, by = gy / ny;
    std::vector<double> fg((size_t)nx * ny, 0.0);
    std::vector<size_t> idx;
    for (size_t i = 0; i < rd.size(); i++)
        if (std::abs(rd[i].element) == elem_idx + 1 && rd[i].stage == stage_idx + 1) idx.push_back(i);
    const bool cyl = surf == 't', pcyl = (surf == 's' && aper == 'l');
    size_t binned = 0, npoints = 0;
    double cen[3] = {0, 0, 0}, zval = 0.0;
    for (size_t j = 0; j < idx.size(); j++) {
        if (!(j == idx.size() - 1 || rd[idx[j]].raynum != rd[idx[j + 1]].raynum)) continue;
        const Record &r = rd[idx[j]];
        if (absorbed_only && r.element >= 0) continue;
        double pe[3], ce[3];
        to_local(r.pos, r.cos, e.origin, e.R, pe, ce);
        double x = pe[0], y = pe[1], z = pe[2];
        cen[0] += x;
        cen[1] += y;
        cen[2] += z;
        npoints++;
        if (cyl) {
            zval = radius;
            if (z <= radius)
                x = radius * asin(x / radius);
            else if (z > radius) {
                if (x < 0) x = -(M_PI * radius / 2.0 + radius * acos(fabs(x) / radius));
                if (x >= 0) x = M_PI * radius / 2.0 + radius * acos(x / radius);
            }
        } else if (pcyl) {
            zval = radius;
            x = radius * asin(x / radius);
        } else {
            zval = z;
        }
        int gxi = -1, gyi = -1;
        while ((minx + (gxi + 1) * bx) < x) gxi++;
        while ((miny + (gyi + 1) * by) < y) gyi++;
        if (gxi >= 0 && gxi < nx && gyi >= 0 && gyi < ny) {
            fg[(size_t)gxi * ny + gyi] += 1;
            binned++;
        }
    }
    if (npoints > 0)
        for (int k = 0; k < 3; k++) cen[k] /= (double)npoints;
    const double ppr = dni * (sys->sun.maxx - sys->sun.minx) * (sys->sun.maxy - sys->sun.miny) / sys->sunraycount;
    const double zscale = ppr / (bx * by);
    double peak = 0, sum = 0, sum2 = 0, minf = 1e99;
    int npeak = -1;
    for (int r = 0; r < nx; r++)
        for (int c = 0; c < ny; c++) {
            double z = fg[(size_t)r * ny + c] * zscale;
            sum += z;
            sum2 += z * z;
            if (z > peak) {
                peak = z;
                npeak = (int)fg[(size_t)r * ny + c];
            }
            if (z < minf) minf = z;
        }
    const double nb = (double)nx * (double)ny;
    const double ave = sum / nb;
    const double sigma = sqrt((nb * sum2 - sum * sum) / (nb * nb));
    const double out[15] = {ppr, peak, minf, ave, sigma, sigma / ave, 100.0 / sqrt((double)npeak),
                            100.0 / sqrt((double)rd.size()), cen[0], cen[1], cen[2], (double)binned, zval, bx, by};
    for (int k = 0; k < 15; k++) stats[k] = out[k];
    bounds[0] = minx;
    bounds[1] = maxx;
    bounds[2] = miny;
    bounds[3] = maxy;
    if (grid)
        for (size_t k = 0; k < fg.size(); k++) grid[k] = (unsigned int)fg[k];
    return 1;
}
}  // extern "C"
