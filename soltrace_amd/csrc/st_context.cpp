// C-ABI of the HIP backend: the 55 entry points of the reference's
// coretrace/stapi.h (declared in include/stapi.h) plus the st_amd_*
// extensions of include/stapi_amd.h.
//
// A context owns its scene model, a HIP stream (default priority), the
// device tables of the last compile, and the device-resident records of the
// last run.  st_sim_run (coretrace/stapi.cpp:680-734) compiles the scene,
// traces every stage on the GPU and leaves the records in HBM; the result
// getters copy them out in the reference's TRayData::Merge order
// (coretrace/types.cpp:311-355).
#include "st_context.h"

#include <deque>
#include <thread>

namespace st {

// Scratch of st_amd_element_stats_noalloc, the allocation-free element
// statistics: bin counts (at most kNoAllocMaxBins u32), then the centroid
// accumulators and the autoscale min/max keys.  Each GPU holds its own copy
// for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress gives on that device.
static constexpr size_t kNoAllocMaxBins = (size_t)1 << 20;
static constexpr size_t kNoAllocAccOff = kNoAllocMaxBins * sizeof(uint32_t);  // 6 x u64 accumulators
static constexpr size_t kNoAllocMmOff = kNoAllocAccOff + 64;                 // 4 x u64 min/max keys
static __device__ __align__(256) unsigned char scratch[kNoAllocMmOff + 64];

static constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static void *scratch_dev[kMaxDevices] = {nullptr};

// The device address of `scratch` on GPU `dev`, looked up once per device
// with that device current.
void *scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[dev]) {
        int cur = 0;
        HC(hipGetDevice(&cur));
        if (cur != dev) HC(hipSetDevice(dev));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) HC(hipSetDevice(cur));
        HC(e);
        scratch_dev[dev] = ptr;
    }
    return scratch_dev[dev];
}

// The streams of every context on GPU `dev`, made once per device with that
// device current and kept until the process ends.  A process gets few
// hardware queues (four by default) and streams beyond them share queues,
// where one stream's work waits behind another's: contexts share one set so
// that two or more contexts of a process keep their streams on queues of
// their own.
static std::mutex streams_mu;
static DevStreams streams_dev[kMaxDevices];
const DevStreams &streams_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lk(streams_mu);
    DevStreams &d = streams_dev[dev];
    if (!d.trace) {
        int cur = 0;
        HC(hipGetDevice(&cur));
        if (cur != dev) HC(hipSetDevice(dev));
        DevStreams n{};
        hipError_t e = hipStreamCreateWithFlags(&n.trace, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&n.mask, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&n.comp, hipStreamNonBlocking);
        if (cur != dev) HC(hipSetDevice(cur));
        HC(e);
        d = n;
    }
    return d;
}

void Context::materialize() {
    if (cols_valid) return;
    // TRayData::Merge applied stage by stage: the records merged so far up
    // to their last full 8192-record block, then this stage's full blocks,
    // then the two partial blocks (coretrace/types.cpp:311-355).  Tracked as
    // pieces of stage chunks, then gathered on the device into columns.
    struct Run {
        int stage;
        size_t first, n;  // records [first, first + n) of the stage
    };
    const size_t B = 8192;
    std::vector<Run> all;
    size_t n_all = 0;
    auto take = [](const std::vector<Run> &v, size_t from, size_t to, std::vector<Run> &out) {
        size_t pos = 0;
        for (const Run &r : v) {
            const size_t a = std::max(pos, from), b = std::min(pos + r.n, to);
            if (a < b) out.push_back(Run{r.stage, r.first + (a - pos), b - a});
            pos += r.n;
        }
    };
    for (size_t s = 0; s < recs.size(); s++) {
        const size_t ns = recs[s].total;
        const std::vector<Run> src = {Run{(int)s, 0, ns}};
        const size_t fa = (n_all / B) * B, fs = (ns / B) * B;
        std::vector<Run> out;
        take(all, 0, fa, out);
        take(src, 0, fs, out);
        take(all, fa, n_all, out);
        take(src, fs, ns, out);
        all.swap(out);
        n_all += ns;
    }
    // runs -> pieces of the stages' chunks
    std::vector<MergePiece> pieces;
    uint64_t dst = 0;
    for (const Run &r : all) {
        const StageRecords &sr = recs[r.stage];
        size_t pos = 0;
        for (size_t k = 0; k < sr.chunks.size(); k++) {
            const size_t a = std::max(pos, r.first), b = std::min(pos + sr.counts[k], r.first + r.n);
            if (a < b) {
                pieces.push_back(MergePiece{sr.ptr(k) + (a - pos), dst, r.stage + 1});
                dst += b - a;
            }
            pos += sr.counts[k];
        }
    }
    if (dst != n_all) throw std::runtime_error("record merge lost records");
    const size_t n = n_all;
    char *base = (char *)d_cols.ensure(std::max<size_t>(n, 1) * (6 * sizeof(double) + 3 * sizeof(int32_t)));
    double *d = (double *)base;
    cols = RecordColumns{d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, (int32_t *)(d + 6 * n),
                         (int32_t *)(d + 6 * n) + n, (int32_t *)(d + 6 * n) + 2 * n};
    if (n) {
        if (!cus) HC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        const MergePiece *dp = upload(d_pieces, pieces, stream);
        launch_gather_records(dp, (int)pieces.size(), n, cols, cus, stream);
        HC(hipStreamSynchronize(stream));
    }
    cols_valid = true;
}

// Columns go through two pinned staging halves in at least four chunks:
// the device copies chunk k into one half while host threads copy chunk
// k - 1 from the other into the caller's (pageable) memory.  Columns of
// 256 KiB or less are copied directly.
void Context::copy_out(void *host, const void *dev, size_t bytes) {
    constexpr size_t kMaxHalf = (size_t)32 << 20, kMinHalf = (size_t)64 << 10;
    if (bytes == 0) return;
    if (bytes <= 4 * kMinHalf) {
        HC(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream));
        return;
    }
    const size_t kHalf = std::min(kMaxHalf, std::max(kMinHalf, (bytes / 4 + 4095) & ~(size_t)4095));
    if (!h_stage) HC(hipHostMalloc((void **)&h_stage, 2 * kMaxHalf, hipHostMallocDefault));
    for (hipEvent_t &e : ev_stage)
        if (!e) HC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const unsigned hw = std::thread::hardware_concurrency();
    auto host_copy = [&](char *dst, const char *src, size_t len) {
        const int nthr = len >= ((size_t)4 << 20) ? (int)std::max(1u, std::min(8u, hw ? hw : 1u)) : 1;
        std::vector<std::thread> th;
        const size_t part = (len + nthr - 1) / nthr;
        for (int t = 1; t < nthr; t++) {
            const size_t a = std::min(len, t * part), b = std::min(len, (t + 1) * part);
            if (a < b) th.emplace_back([=]() { memcpy(dst + a, src + a, b - a); });
        }
        memcpy(dst, src, std::min(len, part));
        for (auto &x : th) x.join();
    };
    const char *d = (const char *)dev;
    char *h = (char *)host;
    const size_t nch = (bytes + kHalf - 1) / kHalf;
    for (size_t k = 0; k <= nch; k++) {
        if (k < nch) {
            const size_t len = std::min(kHalf, bytes - k * kHalf);
            HC(hipMemcpyAsync(h_stage + (k & 1) * kHalf, d + k * kHalf, len, hipMemcpyDeviceToHost, stream));
            HC(hipEventRecord(ev_stage[k & 1], stream));
        }
        if (k >= 1) {
            const size_t j = k - 1, len = std::min(kHalf, bytes - j * kHalf);
            HC(hipEventSynchronize(ev_stage[j & 1]));
            host_copy(h + j * kHalf, h_stage + (j & 1) * kHalf, len);
        }
    }
}

// Compile the scene when a setter changed it (geometry set-up, cull
// structures), check the run's inputs, and upload the device tables: the
// reference's InitGeometries, then Trace's checks and SunToPrimaryStage
// (coretrace/stapi.cpp:699-712, raytrace.cpp:137-250), in its order.
bool Context::prepare() {
    std::string msg;
    const bool recompile = dirty || compiled_cull != cull;
    if (recompile && !compile_scene(model, cull, cs, msg)) {
        errlog("%s", msg.c_str());
        return false;
    }
    const int rayct = load_n2 > 0 ? (int)load_n2 : raycount;
    if ((int64_t)rayct < 1) {
        errlog("invalid number of rays: %d", rayct);
        return false;
    }
    if (model.stages.empty()) {
        errlog("no stages defined.");
        return false;
    }
    if ((int)model.stages.size() > kMaxStages) {
        errlog("at most %d stages are supported", (int)kMaxStages);
        return false;
    }
    hipStream_t s = stream;
    if (recompile) {
        if (!compile_sun(model, cs, msg)) {
            errlog("%s", msg.c_str());
            return false;
        }
        SceneView &v = scene_view;
        v = SceneView{};
        v.elems = upload(d_elems, cs.elems, s);
        v.faces = upload(d_faces, cs.faces, s);
        v.aux = upload(d_aux, cs.aux, s);
        v.stages = upload(d_stages, cs.stages, s);
        v.bvh = upload(d_bvh, cs.bvh, s);
        v.bvh_items = upload(d_bvh_items, cs.bvh_items, s);
        v.always_items = upload(d_always_items, cs.always_items, s);
        v.grid_cells = upload(d_grid_cells, cs.grid_cells, s);
        v.grid_items = upload(d_grid_items, cs.grid_items, s);
        v.grid_always = upload(d_grid_always, cs.grid_always, s);
        v.n_always = (int32_t)cs.grid_always.size();
        v.foot = upload(d_foot, cs.foot, s);
        if (!cs.beam.empty()) {
            v.beam = upload(d_beam, cs.beam, s);
            v.beam_items = upload(d_beam_items, cs.beam_items, s);
        }
        v.grid = cs.grid;
        v.use_grid = cs.use_grid ? 1 : 0;
        v.nstages = (int)cs.stages.size();
        v.sun = cs.sun;
        HC(hipStreamSynchronize(s));
        dirty = false;
        compiled_cull = cull;
    }
    return true;
}

int Context::run(uint32_t seed, st_callback_t cb, void *cbdata) {
    try {
        return precision == 32 ? run_t<float>(seed, cb, cbdata) : run_t<double>(seed, cb, cbdata);
    } catch (const std::exception &e) {
        errlog("trace error: %s", e.what());
        return -1;
    }
}

template <typename R>
int Context::run_t(uint32_t seed, st_callback_t cb, void *cbdata) {
    auto t_start = std::chrono::steady_clock::now();
    clear_results();
    timing = Timing();
    std::string msg;
    // The scene is compiled (geometry set-up, cull structures, device tables)
    // only when a setter changed it since the last successful compile.
    if (!prepare()) return -1;
    // rays to trace: the number of loaded stage-2 rays when there are any
    // (coretrace/stapi.cpp:697-700)
    const int rayct = load_n2 > 0 ? (int)load_n2 : raycount;
    const bool load = load_n1 > 0 && load_n2 > 0;  // coretrace/raytrace.cpp:159-163
    saved_s1.clear();
    saved_s2.clear();
    const uint64_t N = (uint64_t)(int64_t)rayct;
    const uint64_t maxrays = (uint64_t)(int64_t)raymax;
    hipStream_t s = stream;
    const SceneView sv = scene_view;

    // total[4p .. 4p+3]: launch-buffer set p's scan totals (records, rays)
    // and its launches' error flags, read together; total[8..9]: the least
    // sun-ray index whose finite-element interpolation failed in a hit-mask
    // launch (all ones: none); total[10]: error flags of the sunshape
    // sampler, read once stage 1 is done
    uint32_t *total = (uint32_t *)d_total.ensure(12 * sizeof(uint32_t));
    uint32_t *err = total + 10;
    unsigned long long *fe_fail = (unsigned long long *)(total + 8);
    launch_fill_u32(total, 8, 0u, s);
    launch_fill_u32(err, 1, 0u, s);
    launch_fill_u32(total + 8, 2, 0xFFFFFFFFu, s);
    uint32_t *ovf_counts = (uint32_t *)d_ovf_count.ensure(2 * sizeof(uint32_t));
    // Per-launch scratch sized for this run: no launch takes more than N
    // rays (stage-1 hits or rays entering a later stage), so a context of a
    // small scene holds little device memory, however many contexts exist.
    // The second set is sized when a second launch first needs it.
    const uint32_t launch_cap = (uint32_t)std::min<uint64_t>(kMaxChunk, std::max<uint64_t>(N, 1));
    const uint32_t ovf_cap = std::max<uint32_t>(launch_cap / 2, std::min<uint32_t>(kMaxChunk / 2, 65536u));
    auto size_set = [&](int p) {
        d_ovf_rec[p].ensure((size_t)ovf_cap * sizeof(DevRecord));
        d_ovf_meta[p].ensure((size_t)ovf_cap * 2 * sizeof(uint32_t));
        d_rec_slots[p].ensure((size_t)launch_cap * kSlots * sizeof(DevRecord));
        d_rec_count[p].ensure((size_t)launch_cap * sizeof(uint32_t));
        d_rec_offs[p].ensure((size_t)launch_cap * sizeof(uint32_t));
        d_out_slots[p].ensure((size_t)launch_cap * sizeof(DevRay));
        d_out_flag[p].ensure((size_t)launch_cap * sizeof(uint32_t));
        d_out_offs[p].ensure((size_t)launch_cap * sizeof(uint32_t));
        d_scan_tmp[p].ensure(scan_tmp_words((size_t)1 << 26) * sizeof(uint32_t));
        d_scan_tmp2[p].ensure(scan_tmp_words((size_t)1 << 26) * sizeof(uint32_t));
    };
    size_set(0);
    bool set_sized[2] = {true, false};

    unsigned long long *ctrs = nullptr;
    if (count_events) {
        ctrs = (unsigned long long *)d_counters.ensure(((1 + kMaxStages) * 5 + 2) * sizeof(unsigned long long));
        launch_fill_u32((uint32_t *)ctrs, ((1 + kMaxStages) * 5 + 2) * 2, 0u, s);
    }
    FluxSpec fs = flux;
    if (fs.active) {
        d_flux.ensure((size_t)fs.nx * fs.ny * sizeof(uint32_t));
        launch_fill_u32(d_flux.as<uint32_t>(), (size_t)fs.nx * fs.ny, 0u, s);
    }
    // per-source maps: one per stage-1 element, for a flux element on a later stage
    const bool by_src = fs.active && flux_by_source && fs.stage >= 1;
    const size_t src_words = by_src ? model.stages[0].elements.size() * (size_t)fs.nx * fs.ny : 0;
    if (by_src) {
        d_flux_src.ensure(src_words * sizeof(uint32_t));
        launch_fill_u32(d_flux_src.as<uint32_t>(), src_words, 0u, s);
    }

    // kernel times: events recorded around the launches, read once the run
    // has finished (no synchronisation per launch)
    marks.clear();
    size_t ev_used = 0;
    auto timed_on = [&](hipStream_t st, double &acc, auto &&fn) {
        while (ev_used + 2 > ev_pool.size()) {
            hipEvent_t e;
            HC(hipEventCreate(&e));
            ev_pool.push_back(e);
        }
        HC(hipEventRecord(ev_pool[ev_used], st));
        fn();
        HC(hipEventRecord(ev_pool[ev_used + 1], st));
        marks.emplace_back(&acc, ev_used);
        ev_used += 2;
    };
    auto timed = [&](double &acc, auto &&fn) { timed_on(s, acc, fn); };
    auto read_u32 = [&](const uint32_t *p) {
        uint32_t v = 0;
        HC(hipMemcpyAsync(&v, p, sizeof v, hipMemcpyDeviceToHost, s));
        HC(hipStreamSynchronize(s));
        return v;
    };
    auto take_buffer = [&](size_t bytes) -> std::shared_ptr<DevBuf> {
        std::shared_ptr<DevBuf> b;
        for (size_t i = 0; i < rec_pool.size(); i++)
            if (rec_pool[i]->cap >= bytes) {
                b = rec_pool[i];
                rec_pool.erase(rec_pool.begin() + (long)i);
                break;
            }
        if (!b) {
            if (!rec_pool.empty()) {
                b = rec_pool.back();
                rec_pool.pop_back();
            } else {
                b = std::make_shared<DevBuf>();
            }
        }
        b->ensure(bytes);
        return b;
    };
    auto check_err = [&](uint32_t e) -> bool {
        if (e & 16u) {
            // the reference throws out of Trace (coretrace/vshot.cpp:177-178,
            // raytrace.cpp:1108-1112)
            errlog("trace error: Interpolation failed to find nearby nodes to ray hit in FEA surface.");
            return false;
        }
        if (e & 1u) {
            errlog("Bad optical interaction type (stage)");
            return false;
        }
        if (e & 2u) {
            errlog("internal error: sun-ray hit test disagreed between launches");
            return false;
        }
        if (e & 4u) {
            errlog("internal error: record overflow buffer full");
            return false;
        }
        if (e & 8u) {
            errlog("sunshape table sampling found no accepted angle in 2^26 tries");
            return false;
        }
        return true;
    };

    recs.assign(model.stages.size(), StageRecords());
    const size_t nst = model.stages.size();
    uint64_t traced_total = 0;
    bool cancelled = false;

    // One trace launch over n rays (chunk), then record and ray compaction.
    // Appends the passed rays to next_rays at *n_next.
    const bool presample = err_sun && (sv.sun.shape == 'd' || sv.sun.shape == 'D') && !sv.sun.point_source;
    // Tabulated sunshape: the polar angles of a list of stage-0 rays, drawn
    // cooperatively ahead of their trace.
    auto launch_presample = [&](hipStream_t st, const uint64_t *sun_idx, int n, SunPre *out) {
        SunPreArgs pa{};
        pa.sun = sv.sun;
        pa.aux = sv.aux;
        pa.seed = seed;
        pa.k0 = 3;  // two sun-plane draws and the reflectivity draw come first
        pa.sun_index = sun_idx;
        pa.n = n;
        pa.out = out;
        pa.err_flags = err;
        pa.counters = ctrs ? ctrs + 5 * (1 + kMaxStages) : nullptr;
        timed_on(st, timing.sunpre_ms, [&]() { launch_sunshape_d<R>(pa, st); });
    };
    // One trace launch over n rays (chunk), then record and ray compaction,
    // in two halves so the host can queue other work between them: the
    // launch half ends with an asynchronous copy of the totals into pinned
    // memory; the finish half waits for it, checks the error flags, takes a
    // record buffer and compacts on the compaction stream.  Passed rays are
    // appended to next_rays at *n_next.  Launch k works in buffer set k mod
    // 2 and first waits for the compaction of launch k - 2 (same set), so
    // with two launches in flight a launch traces while the one before it is
    // compacted.
    struct Pending {
        TraceArgs a;
        int si, n, set;
        DevRay *next_rays;
        uint64_t *n_next;
        std::chrono::steady_clock::time_point t0;
        bool last;  // the stage's last launch: compacted on the trace stream, nothing to overlap
    };
    hipStream_t cs_ = comp_stream;
    uint64_t launch_seq = 0;
    bool comp_queued[2] = {false, false};
    auto trace_launch = [&](int si, int n, const uint64_t *sun_idx, uint32_t raynum_base, const DevRay *in,
                            uint32_t last_raynum, DevRay *next_rays, uint64_t *n_next,
                            const SunPre *pre, const uint32_t *order = nullptr) -> Pending {
        const int set = (int)(launch_seq++ & 1);
        if (!set_sized[set]) {
            size_set(set);
            set_sized[set] = true;
        }
        if (comp_queued[set]) HC(hipStreamWaitEvent(s, ev_comp[set], 0));
        uint32_t *tot_set = total + 4 * set;
        uint32_t *ovf_count = ovf_counts + set;
        TraceArgs a{};
        a.sv = sv;
        a.si = si;
        a.seed = seed;
        a.inc_sun = err_sun;
        a.inc_err = err_opt;
        a.n_rays = n;
        a.ray_base = 0;
        a.slot_stride = (uint32_t)n;
        a.slots = kSlots;
        a.sun_index = sun_idx;
        a.raynum_base = raynum_base;
        a.in_rays = in;
        a.last_raynum = last_raynum;
        a.rec_slots = d_rec_slots[set].as<DevRecord>();
        a.rec_count = d_rec_count[set].as<uint32_t>();
        a.ovf_rec = d_ovf_rec[set].as<DevRecord>();
        a.ovf_meta = d_ovf_meta[set].as<uint32_t>();
        a.ovf_count = ovf_count;
        a.ovf_cap = ovf_cap;
        a.out_slots = next_rays ? d_out_slots[set].as<DevRay>() : nullptr;
        a.out_flag = next_rays ? d_out_flag[set].as<uint32_t>() : nullptr;
        a.err_flags = tot_set + 3;
        a.flux = fs;
        a.flux_bins = fs.active ? d_flux.as<uint32_t>() : nullptr;
        a.flux_src = by_src ? d_flux_src.as<uint32_t>() : nullptr;
        a.flux_nbins = (uint32_t)(fs.nx * fs.ny);
        a.counters = ctrs ? ctrs + 5 * (1 + si) : nullptr;
        a.generic = cs.stages[si].generic;
        a.sun_pre = pre;
        a.order = order;
        if (si == 0 && presample && pre == nullptr) {
            SunPre *out = (SunPre *)d_sunpre.ensure((size_t)launch_cap * sizeof(SunPre));
            launch_presample(s, sun_idx, n, out);
            a.sun_pre = out;
        }
        timing.trace_launches[si]++;
        launch_fill_u32(ovf_count, 1, 0u, s);
        if (!h_args) HC(hipHostMalloc((void **)&h_args, kArgsRing * sizeof(TraceArgs), hipHostMallocDefault));
        TraceArgs *dev_a = (TraceArgs *)d_args.ensure(kArgsRing * sizeof(TraceArgs)) + args_next;
        h_args[args_next] = a;
        HC(hipMemcpyAsync(dev_a, &h_args[args_next], sizeof(TraceArgs), hipMemcpyHostToDevice, s));
        args_next = (args_next + 1) % kArgsRing;
        timed(timing.trace_ms[si], [&]() { launch_trace_stage<R>(a, dev_a, s); });
        timing.rays_in[si] += (uint64_t)n;
        launch_tile_scan(a.rec_count, d_rec_offs[set].as<uint32_t>(), (size_t)n, d_scan_tmp[set].as<uint32_t>(),
                         tot_set, s);
        if (next_rays)
            launch_tile_scan(a.out_flag, d_out_offs[set].as<uint32_t>(), (size_t)n, d_scan_tmp2[set].as<uint32_t>(),
                             tot_set + 1, s);
        else
            launch_fill_u32(tot_set + 1, 1, 0u, s);
        HC(hipMemcpyAsync(h_tot + 4 * set, tot_set, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HC(hipEventRecord(ev_tot[set], s));
        return Pending{a, si, n, set, next_rays, n_next, std::chrono::steady_clock::now(), false};
    };
    auto trace_finish = [&](Pending &p) -> bool {
        const int set = p.set;
        HC(hipEventSynchronize(ev_tot[set]));
        const uint32_t *ht = h_tot + 4 * set;
        uint32_t tot[4] = {ht[0], ht[1], ht[2], ht[3]};
        if (!check_err(tot[3])) return false;
        const TraceArgs &a = p.a;
        const int n = p.n, si = p.si;
        std::shared_ptr<DevBuf> chunk = take_buffer((size_t)std::max<uint32_t>(1, tot[0]) * sizeof(DevRecord));
        const hipStream_t cst = p.last ? s : cs_;
        if (!p.last) HC(hipStreamWaitEvent(cs_, ev_tot[set], 0));
        launch_compact_records(a.rec_slots, a.rec_count, d_rec_offs[set].as<uint32_t>(), d_scan_tmp[set].as<uint32_t>(),
                               n, kSlots, chunk->as<DevRecord>(), cst);
        launch_compact_overflow(a.ovf_rec, a.ovf_meta, a.ovf_count, ovf_cap, d_rec_offs[set].as<uint32_t>(),
                                d_scan_tmp[set].as<uint32_t>(), chunk->as<DevRecord>(), cst);
        if (tot[1] && p.next_rays)
            launch_compact_rays(a.out_slots, a.out_flag, d_out_offs[set].as<uint32_t>(),
                                d_scan_tmp2[set].as<uint32_t>(), n, p.next_rays + *p.n_next, cst);
        if (!p.last) HC(hipEventRecord(ev_comp[set], cs_));
        comp_queued[set] = !p.last;
        recs[si].add(chunk, 0, tot[0], true);
        timing.records[si] += tot[0];
        if (p.n_next) *p.n_next += tot[1];
        timing.other_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - p.t0).count();
        traced_total += (uint64_t)n;
        return true;
    };
    // Launches whose totals the host has not read yet: with a progress
    // callback each launch is finished before the next (the callback sees
    // every launch's rays done), else two stay in flight.
    std::deque<Pending> inflight;
    const size_t depth = cb ? 1 : 2;
    auto finish_to = [&](size_t keep) -> bool {
        while (inflight.size() > keep) {
            Pending p = inflight.front();
            inflight.pop_front();
            if (!trace_finish(p)) {
                inflight.clear();
                return false;
            }
        }
        return true;
    };
    // every queued compaction done before stream s goes on
    auto join_comp = [&]() {
        HC(hipEventRecord(ev_join, cs_));
        HC(hipStreamWaitEvent(s, ev_join, 0));
    };

    // ------------------------------------------------------ stage 0 -------
    // Batches of sun rays, pipelined over two streams.  The mask stream
    // draws batch k+1's sun-ray mask, compacts its hit list and samples its
    // tabulated sunshape angles while the trace stream traces batch k; the
    // host reads batch k+1's hit count after queueing batch k's trace and
    // before waiting for batch k's totals.  Events hand each hit list over.
    // The mask side has its own scan workspace and totals; the sunshape
    // angles alternate between two buffers.
    DevRay *next = (DevRay *)d_rays_a.ensure((size_t)N * sizeof(DevRay));
    uint64_t n_next = 0;
    if (!load) {
        uint64_t *sun_index_all = (uint64_t *)d_hits.ensure((size_t)N * sizeof(uint64_t));
        const uint64_t kMaxBatch = 1u << 25;
        const size_t max_words = (size_t)kMaxBatch / 64;
        uint64_t *mask = (uint64_t *)d_mask.ensure(max_words * sizeof(uint64_t));
        uint32_t *popc = (uint32_t *)d_popc.ensure(max_words * sizeof(uint32_t));
        uint32_t *offs = (uint32_t *)d_offs.ensure(max_words * sizeof(uint32_t));
        uint32_t *mscan = (uint32_t *)d_mask_scan_tmp.ensure(scan_tmp_words(max_words) * sizeof(uint32_t));
        uint32_t *mtotal = (uint32_t *)d_mask_total.ensure(4 * sizeof(uint32_t));
        hipStream_t sm = mask_stream;
        HC(hipEventRecord(ev_ready, s));  // counters and flags are cleared
        HC(hipStreamWaitEvent(sm, ev_ready, 0));
        uint64_t hits = 0, s0 = 0, batch = 0;
        int parity = 0;
        // many stage-1 elements under a BVH: order each launch's rays by
        // sun-plane bucket (the walks of a wave's reflected rays then agree)
        const bool sort_rays = sv.use_grid && !sv.sun.point_source && cs.stages[0].nbvh > 0;
        double rate = 0.0;  // hits per sun ray observed so far
        auto plan = [&](uint64_t want) -> uint64_t {
            uint64_t b = (rate > 0) ? (uint64_t)(want / rate * 1.05) + 4096
                                    : std::min<uint64_t>(1u << 22, want * 4 + 4096);
            b = std::min<uint64_t>(b, kMaxBatch);
            b = std::max<uint64_t>(b, 4096);
            if (s0 + b > maxrays) b = maxrays - s0;
            return b;
        };
        auto launch_mask_batch = [&]() {
            const size_t nwords = (batch + 63) / 64;
            MaskArgs ma{};
            ma.sv = sv;
            ma.seed = seed;
            ma.s0 = s0;
            ma.n = (uint32_t)batch;
            ma.mask = mask;
            ma.fe_fail = fe_fail;
            ma.counters = ctrs;
            ma.generic = cs.stages[0].generic;
            // pairs pack the element index in 24 bits
            ma.compact = sv.use_grid && !sv.sun.point_source && cs.stages[0].nelem < (1 << 24);
            timing.mask_launches++;
            timed_on(sm, timing.mask_ms, [&]() { launch_sun_mask<R>(ma, sm); });
            timing.sun_rays += batch;
            launch_mask_popc(mask, popc, nwords, sm);
            launch_tile_scan(popc, offs, nwords, mscan, mtotal, sm);
        };
        struct Ready {
            uint64_t *hit_list;
            uint32_t take;
            uint64_t raynum_base;
            SunPre *pre;
            const uint32_t *order;  // lane -> ray of the batch (nullptr: identity)
            hipEvent_t ev;
            int par;             // parity of its sunshape-angle and order buffers
            bool more;           // another batch was launched behind this one
            uint64_t sun_count;  // sun rays drawn up to this batch's last hit (or its end)
        };
        bool maxed = false;
        bool traced_rec[2] = {false, false};
        // Read the pending mask batch's hit count, compact its hit list (and
        // sample its sunshape angles), and launch the next mask batch if hits
        // are still missing.
        auto prepare = [&]() -> Ready {
            Ready r{};
            const uint64_t want = N - hits;
            const size_t nwords = (batch + 63) / 64;
            uint32_t h = 0;
            HC(hipMemcpyAsync(&h, mtotal, sizeof h, hipMemcpyDeviceToHost, sm));
            HC(hipStreamSynchronize(sm));
            // the batch two back, which used this parity's buffers, is traced
            if (traced_rec[parity]) HC(hipStreamWaitEvent(sm, ev_traced[parity], 0));
            r.par = parity;
            r.take = (uint32_t)std::min<uint64_t>(h, want);
            r.hit_list = sun_index_all + hits;
            r.raynum_base = hits;
            if (r.take) launch_mask_compact(mask, offs, mscan, nwords, s0, r.take, r.hit_list, sm);
            r.pre = nullptr;
            if (presample && r.take) {
                r.pre = (SunPre *)d_sunpre_batch[parity].ensure((size_t)r.take * sizeof(SunPre));
                launch_presample(sm, r.hit_list, (int)r.take, r.pre);
            }
            // the batch's rays in sun-plane bucket order, for a single launch
            r.order = nullptr;
            if (sort_rays && r.take && r.take <= launch_cap) {
                OrderWork w;
                w.key_rank = (uint32_t *)d_keyrank.ensure((size_t)r.take * sizeof(uint32_t));
                w.table = (uint32_t *)d_hist.ensure(order_table_words((int)r.take) * sizeof(uint32_t));
                w.colsum = (uint32_t *)d_hoffs.ensure(2 * kOrderBuckets * sizeof(uint32_t));
                w.cbase = w.colsum + kOrderBuckets;
                w.ctile = (uint32_t *)d_htmp.ensure(scan_tmp_words(kOrderBuckets) * sizeof(uint32_t));
                w.ctot = (uint32_t *)d_htot.ensure(sizeof(uint32_t));
                uint32_t *ord = (uint32_t *)d_order[parity].ensure((size_t)r.take * sizeof(uint32_t));
                launch_sun_order<R>(sv, seed, r.hit_list, (int)r.take, w, ord, sm);
                r.order = ord;
            }
            if (r.take == want) {
                uint64_t last = 0;
                HC(hipMemcpyAsync(&last, r.hit_list + r.take - 1, sizeof last, hipMemcpyDeviceToHost, sm));
                HC(hipStreamSynchronize(sm));
                r.sun_count = last + 1;
            } else {
                r.sun_count = s0 + batch;
            }
            r.ev = ev_hits[parity];
            HC(hipEventRecord(r.ev, sm));
            parity ^= 1;
            hits += r.take;
            s0 += batch;
            rate = (double)hits / (double)s0;
            r.more = false;
            if (hits < N) {
                if (s0 >= maxrays) {
                    maxed = true;
                } else {
                    batch = plan(N - hits);
                    launch_mask_batch();
                    r.more = true;
                }
            }
            return r;
        };
        batch = plan(N);
        if (batch == 0) {
            errlog("generated sun rays reached maximum count: %d", raymax);
            return -1;
        }
        launch_mask_batch();
        Ready cur_b = prepare();
        for (;;) {
            HC(hipStreamWaitEvent(s, cur_b.ev, 0));
            bool ok = true;
            Ready next_b{};
            bool have_next = false;
            for (uint32_t c = 0; c < cur_b.take && ok; c += launch_cap) {
                const int n = (int)std::min<uint32_t>(launch_cap, cur_b.take - c);
                // a single stage hands no rays on
                inflight.push_back(trace_launch(0, n, cur_b.hit_list + c, (uint32_t)(cur_b.raynum_base + c), nullptr,
                                                0, nst > 1 ? next : nullptr, &n_next,
                                                cur_b.pre != nullptr ? cur_b.pre + c : nullptr,
                                                c == 0 && (uint32_t)n == cur_b.take ? cur_b.order : nullptr));
                inflight.back().last = !cur_b.more && c + (uint32_t)n >= cur_b.take;
                if (c == 0 && cur_b.more) {
                    next_b = prepare();  // overlaps the trace just queued
                    have_next = true;
                }
                ok = finish_to(depth - 1);
            }
            HC(hipEventRecord(ev_traced[cur_b.par], s));
            traced_rec[cur_b.par] = true;
            if (!ok) {
                HC(hipStreamSynchronize(sm));
                HC(hipStreamSynchronize(cs_));
                return -1;
            }
            n_hits_done = (uint64_t)cur_b.raynum_base + cur_b.take;
            sun_ray_count = cur_b.sun_count;
            if (cb && !cb((st_uint_t)traced_total, (st_uint_t)n_hits_done, (st_uint_t)N, 1, (st_uint_t)nst, cbdata)) {
                cancelled = true;
                break;
            }
            if (!have_next && cur_b.more) {  // an empty batch: nothing was traced
                next_b = prepare();
                have_next = true;
            }
            if (!have_next) break;
            cur_b = next_b;
        }
        if (!finish_to(0)) {
            HC(hipStreamSynchronize(sm));
            HC(hipStreamSynchronize(cs_));
            return -1;
        }
        join_comp();
        HC(hipStreamSynchronize(sm));
        if (cs.stages[0].generic == 2) {
            // hit-mask launches run ahead of the trace: a failed interpolation
            // counts when the reference reaches its sun ray, i.e. before the
            // last sun ray used
            unsigned long long f = 0;
            HC(hipMemcpyAsync(&f, fe_fail, sizeof f, hipMemcpyDeviceToHost, sm));
            HC(hipStreamSynchronize(sm));
            if (f < (unsigned long long)sun_ray_count) {
                check_err(16u);
                return -1;
            }
        }
        if (maxed && !cancelled) {
            errlog("generated sun rays reached maximum count: %d", raymax);
            return -1;
        }
    } else {
        // stage 1 from saved data (coretrace/raytrace.cpp:544-578,
        // generateray.cpp:165-219): its records are the rows of s1, stage 2
        // starts from the rays of s2; the row fields convert as the
        // reference's casts do (element: double -> st_uint_t -> int)
        std::vector<DevRecord> r0(load_n1);
        for (uint64_t j = 0; j < load_n1; j++) {
            const double *v = load_s1 + 8 * j;
            DevRecord &r = r0[j];
            for (int k = 0; k < 3; k++) { r.pos[k] = v[k]; r.cos[k] = v[k + 3]; }
            r.element = (int32_t)(int64_t)v[6];
            r.raynum = (uint32_t)(uint64_t)(int64_t)v[7];
        }
        std::vector<DevRay> in(load_n2);
        for (uint64_t j = 0; j < load_n2; j++) {
            const double *v = load_s2 + 7 * j;
            DevRay &r = in[j];
            for (int k = 0; k < 3; k++) { r.pos[k] = v[k]; r.cos[k] = v[k + 3]; }
            r.raynum = (uint32_t)(int32_t)v[6];
            r.src = 0;
            r.sun_index = (uint64_t)r.raynum - 1;  // the draws of a loaded ray
        }
        std::shared_ptr<DevBuf> chunk = take_buffer(r0.size() * sizeof(DevRecord));
        HC(hipMemcpyAsync(chunk->p, r0.data(), r0.size() * sizeof(DevRecord), hipMemcpyHostToDevice, s));
        HC(hipMemcpyAsync(next, in.data(), in.size() * sizeof(DevRay), hipMemcpyHostToDevice, s));
        HC(hipStreamSynchronize(s));
        recs[0].add(chunk, 0, r0.size(), true);
        if (fs.active && fs.stage == 0) {
            // a traced stage is binned into the flux map at record time;
            // loaded records are binned here, each ray's last record on the
            // element (the rows of a ray are contiguous, as saved)
            DevBuf acc;
            acc.ensure(6 * sizeof(unsigned long long));
            launch_fill_u32(acc.as<uint32_t>(), 12, 0u, s);
            launch_stats_bin(chunk->as<DevRecord>(), (uint32_t)r0.size(), fs.elem, fs, 0, d_flux.as<uint32_t>(),
                             acc.as<unsigned long long>(), s);
            HC(hipStreamSynchronize(s));
        }
        n_next = load_n2;
        sun_ray_count = r0.back().raynum;
        n_hits_done = N;
    }
    if (save_data && !cancelled) {
        // every stage-1 record in trace order (coretrace/raytrace.cpp:1044-1063)
        std::vector<DevRecord> r0(recs[0].total);
        size_t off = 0;
        for (size_t c = 0; c < recs[0].chunks.size(); c++) {
            if (recs[0].counts[c])
                HC(hipMemcpyAsync(r0.data() + off, recs[0].ptr(c), recs[0].counts[c] * sizeof(DevRecord),
                                  hipMemcpyDeviceToHost, s));
            off += recs[0].counts[c];
        }
        HC(hipStreamSynchronize(s));
        saved_s1.reserve(r0.size() * 8);
        for (const DevRecord &r : r0) {
            saved_s1.insert(saved_s1.end(), {r.pos[0], r.pos[1], r.pos[2], r.cos[0], r.cos[1], r.cos[2],
                                             (double)r.element, (double)r.raynum});
        }
    }

    // stage-0 pass-on rule: the last ray's pass counts only when an earlier
    // ray was passed too (coretrace/raytrace.cpp:786-797, 1081-1085)
    auto last_raynum_of = [&](DevRay *rays, uint64_t n) -> uint32_t {
        DevRay r;
        HC(hipMemcpyAsync(&r, rays + n - 1, sizeof r, hipMemcpyDeviceToHost, s));
        HC(hipStreamSynchronize(s));
        return r.raynum;
    };
    DevRay *cur = next;
    uint64_t n_cur = n_next;
    bool prev_has = false;
    uint32_t last_prev = 0;
    if (load) {
        prev_has = true;
        last_prev = last_raynum_of(cur, n_cur);
        if (last_prev == 0) {  // coretrace/raytrace.cpp:1087-1096
            const uint32_t pp = n_cur > 1 ? last_raynum_of(cur, n_cur - 1) : 0;
            errlog("LastRayNumberInPreviousStage=0, stage %d, PrevIdx=%d, CurIdx=%d, pp=%d", 1, (int)(n_cur - 1), 0,
                   (int)pp);
            return -1;
        }
    } else if (!cancelled && n_cur > 0) {
        // the last ray number handed on matters to the pass-on rule only for
        // a single ray, to a trace-through stage 2 and to the callback
        // (reading it costs a round trip while the GPU waits)
        if (n_cur == 1 || cb != nullptr || (nst > 1 && cs.stages[1].tracethrough))
            last_prev = last_raynum_of(cur, n_cur);
        uint64_t below = (last_prev == N) ? n_cur - 1 : n_cur;
        prev_has = below > 0;
    }

    // ------------------------------------------------- later stages -------
    DevBuf *bufs[2] = {&d_rays_a, &d_rays_b};
    int which = 0;
    for (size_t si = 1; si < nst && prev_has && !cancelled; si++) {
        // the last stage hands no rays on, unless its rays are saved
        const bool hand_on = si + 1 < nst || (si == 1 && save_data);
        DevRay *nxt = hand_on ? (DevRay *)bufs[which ^ 1]->ensure(std::max<uint64_t>(n_cur, 1) * sizeof(DevRay))
                              : nullptr;
        uint64_t n_out = 0;
        for (uint64_t c = 0; c < n_cur; c += launch_cap) {
            int n = (int)std::min<uint64_t>(launch_cap, n_cur - c);
            inflight.push_back(trace_launch((int)si, n, nullptr, 0, cur + c, last_prev, nxt, &n_out, nullptr));
            inflight.back().last = c + (uint64_t)n >= n_cur;
            if (!finish_to(depth - 1)) {
                HC(hipStreamSynchronize(cs_));
                return -1;
            }
            if (cb && !cb((st_uint_t)traced_total, (st_uint_t)(c + n), (st_uint_t)last_prev, (st_uint_t)si + 1,
                          (st_uint_t)nst, cbdata)) {
                cancelled = true;
                break;
            }
        }
        if (!finish_to(0)) {
            HC(hipStreamSynchronize(cs_));
            return -1;
        }
        join_comp();
        if (si == 1 && save_data && !cancelled) {
            // the reference's stage-2 ray array as stage 2 leaves it
            // (coretrace/raytrace.cpp:1065-1078): the array is reused in
            // place, so its first n_out rows are the rays leaving stage 2 and
            // the rest the rays that entered it
            std::vector<DevRay> rows(n_cur);
            if (n_out)
                HC(hipMemcpyAsync(rows.data(), nxt, n_out * sizeof(DevRay), hipMemcpyDeviceToHost, s));
            if (n_cur > n_out)
                HC(hipMemcpyAsync(rows.data() + n_out, cur + n_out, (n_cur - n_out) * sizeof(DevRay),
                                  hipMemcpyDeviceToHost, s));
            HC(hipStreamSynchronize(s));
            saved_s2.reserve(rows.size() * 7);
            for (const DevRay &r : rows)
                saved_s2.insert(saved_s2.end(),
                                {r.pos[0], r.pos[1], r.pos[2], r.cos[0], r.cos[1], r.cos[2], (double)r.raynum});
        }
        which ^= 1;
        cur = nxt;
        n_cur = n_out;
        prev_has = n_cur > 0;
        if (prev_has && (cb != nullptr || (si + 1 < nst && cs.stages[si + 1].tracethrough)))
            last_prev = last_raynum_of(cur, n_cur);
    }
    // the run's last copies behind one synchronisation: the sunshape
    // sampler's error flags, the flux map, the event counters
    uint32_t sampler_err = 0;
    if (presample) HC(hipMemcpyAsync(&sampler_err, err, sizeof sampler_err, hipMemcpyDeviceToHost, s));
    if (fs.active) {
        flux_host.assign((size_t)fs.nx * fs.ny, 0u);
        HC(hipMemcpyAsync(flux_host.data(), d_flux.p, flux_host.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          s));
        flux_src_words = src_words;  // per-source maps stay in HBM until asked for
    }
    if (ctrs) {
        HC(hipMemcpyAsync(&timing.ev[0][0], ctrs, sizeof(timing.ev), hipMemcpyDeviceToHost, s));
        HC(hipMemcpyAsync(timing.ev_sun, ctrs + 5 * (1 + kMaxStages), sizeof(timing.ev_sun), hipMemcpyDeviceToHost,
                          s));
    }
    HC(hipStreamSynchronize(s));
    if (!check_err(sampler_err)) return -1;
    timing.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return (int)total_records();
}

}  // namespace st

using namespace st;

#define CTX(p, r)                       \
    Context *c = ctx_of(p);             \
    if (!c) return r;
// setters also mark the compiled scene stale
#define CTXW(p, r)                      \
    CTX(p, r)                           \
    c->dirty = true;
#define STAGE(i)                                        \
    if ((i) >= c->model.stages.size()) return -1;      \
    HStage &stg = c->model.stages[i];
#define ELEM(i)                                         \
    if ((i) >= stg.elements.size()) return -1;         \
    HElement &el = stg.elements[i];

// ------------------------------------------------------------- contexts ---
STCORE_API st_context_t st_create_context() {
    try {
        return reinterpret_cast<st_context_t>(new Context);
    } catch (const std::exception &e) {
        fprintf(stderr, "st_create_context: %s\n", e.what());
        return nullptr;
    }
}
STCORE_API int st_free_context(st_context_t p) {
    CTX(p, -1);
    delete c;
    return 1;
}
STCORE_API int st_num_messages(st_context_t p) {
    CTX(p, -1);
    return (int)c->messages.size();
}
STCORE_API const char *st_message(st_context_t p, st_uint_t idx) {
    CTX(p, nullptr);
    return idx < c->messages.size() ? c->messages[idx].c_str() : nullptr;
}
STCORE_API void st_clear_messages(st_context_t p) {
    Context *c = ctx_of(p);
    if (c) c->messages.clear();
}

// --------------------------------------------------------------- optics ---
STCORE_API int st_num_optics(st_context_t p) {
    CTX(p, -1);
    return (int)c->model.optics.size();
}
STCORE_API int st_add_optic(st_context_t p, const char *name) {
    CTXW(p, -1);
    c->model.optics.emplace_back();
    c->model.optics.back().name = name ? name : "";
    return (int)c->model.optics.size() - 1;
}
STCORE_API int st_delete_optic(st_context_t p, st_uint_t idx) {
    CTXW(p, -1);
    if (idx >= c->model.optics.size()) return -1;
    c->model.optics.erase(c->model.optics.begin() + idx);
    return 1;
}
STCORE_API int st_clear_optics(st_context_t p) {
    CTXW(p, -1);
    c->model.optics.clear();
    return 1;
}
STCORE_API int st_optic(st_context_t p, st_uint_t idx, int fb, char dist, int optnum, int apgr, int order,
                        double rreal, double rimag, double ref, double tra, double gratingab12[3], double rmsslope,
                        double rmsspec, int userefltable, int refl_npoints, double *refl_angles, double *refls,
                        int usetranstable, int trans_npoints, double *trans_angles, double *transs) {
    CTXW(p, -1);
    if (idx >= c->model.optics.size()) return -1;
    HFace &f = (fb == 2) ? c->model.optics[idx].back : c->model.optics[idx].front;
    f.dist = dist;
    f.optnum = optnum;
    f.apgr = apgr;
    f.order = order;
    f.refr[0] = rreal;
    f.refr[1] = rimag;
    f.reflectivity = ref;
    f.transmissivity = tra;
    // the reference copies four grating coefficients (coretrace/stapi.cpp:190-191)
    for (int i = 0; i < 4; i++) f.ab12[i] = gratingab12 ? gratingab12[i] : 0.0;
    f.slope = rmsslope;
    f.spec = rmsspec;
    if (userefltable && refl_npoints > 0 && refl_angles && refls) {
        f.use_refl = true;
        f.refl_ang.assign(refl_angles, refl_angles + refl_npoints);
        f.refl_val.assign(refls, refls + refl_npoints);
    }
    if (usetranstable && trans_npoints > 0 && trans_angles && transs) {
        f.use_trans = true;
        f.trans_ang.assign(trans_angles, trans_angles + trans_npoints);
        f.trans_val.assign(transs, transs + trans_npoints);
    }
    return 1;
}

// --------------------------------------------------------------- stages ---
STCORE_API int st_num_stages(st_context_t p) {
    CTX(p, -1);
    return (int)c->model.stages.size();
}
STCORE_API int st_add_stage(st_context_t p) {
    CTXW(p, -1);
    c->model.stages.emplace_back();
    return (int)c->model.stages.size() - 1;
}
STCORE_API int st_add_stages(st_context_t p, st_uint_t num) {
    CTXW(p, -1);
    for (st_uint_t i = 0; i < num; i++) c->model.stages.emplace_back();
    return (int)num;
}
STCORE_API int st_delete_stage(st_context_t p, st_uint_t idx) {
    CTXW(p, -1);
    if (idx >= c->model.stages.size()) return -1;
    c->model.stages.erase(c->model.stages.begin() + idx);
    return 1;
}
STCORE_API int st_clear_stages(st_context_t p) {
    CTXW(p, -1);
    c->model.stages.clear();
    return 1;
}
STCORE_API int st_stage_xyz(st_context_t p, st_uint_t i, double x, double y, double z) {
    CTXW(p, -1);
    STAGE(i);
    stg.origin[0] = x; stg.origin[1] = y; stg.origin[2] = z;
    return 1;
}
STCORE_API int st_stage_aim(st_context_t p, st_uint_t i, double x, double y, double z) {
    CTXW(p, -1);
    STAGE(i);
    stg.aim[0] = x; stg.aim[1] = y; stg.aim[2] = z;
    return 1;
}
STCORE_API int st_stage_zrot(st_context_t p, st_uint_t i, double z) {
    CTXW(p, -1);
    STAGE(i);
    stg.zrot = z;
    return 1;
}
STCORE_API int st_stage_flags(st_context_t p, st_uint_t i, int virt, int multihit, int tracethrough) {
    CTXW(p, -1);
    STAGE(i);
    stg.virt = virt != 0;
    stg.multihit = multihit != 0;
    stg.tracethrough = tracethrough != 0;
    return 1;
}

// ------------------------------------------------------------- elements ---
STCORE_API int st_num_elements(st_context_t p, st_uint_t i) {
    CTX(p, -1);
    STAGE(i);
    return (int)stg.elements.size();
}
STCORE_API int st_add_element(st_context_t p, st_uint_t i) {
    CTXW(p, -1);
    STAGE(i);
    stg.elements.emplace_back();
    return (int)stg.elements.size() - 1;
}
STCORE_API int st_add_elements(st_context_t p, st_uint_t i, st_uint_t num) {
    CTXW(p, -1);
    if (num < 1) return -1;
    STAGE(i);
    stg.elements.resize(stg.elements.size() + num);
    return (int)num;
}
STCORE_API int st_delete_element(st_context_t p, st_uint_t i, st_uint_t idx) {
    CTXW(p, -1);
    STAGE(i);
    if (idx >= stg.elements.size()) return -1;
    stg.elements.erase(stg.elements.begin() + idx);
    return 1;
}
STCORE_API int st_clear_elements(st_context_t p, st_uint_t i) {
    CTXW(p, -1);
    STAGE(i);
    stg.elements.clear();
    return 1;
}
STCORE_API int st_element_enabled(st_context_t p, st_uint_t i, st_uint_t j, int en) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    el.enabled = en != 0;
    return 1;
}
STCORE_API int st_element_xyz(st_context_t p, st_uint_t i, st_uint_t j, double x, double y, double z) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    el.origin[0] = x; el.origin[1] = y; el.origin[2] = z;
    return 1;
}
STCORE_API int st_element_aim(st_context_t p, st_uint_t i, st_uint_t j, double x, double y, double z) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    el.aim[0] = x; el.aim[1] = y; el.aim[2] = z;
    return 1;
}
STCORE_API int st_element_zrot(st_context_t p, st_uint_t i, st_uint_t j, double z) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    el.zrot = z;
    return 1;
}
STCORE_API int st_element_aperture(st_context_t p, st_uint_t i, st_uint_t j, char ap) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    el.shape = ap;
    return 1;
}
STCORE_API int st_element_aperture_params(st_context_t p, st_uint_t i, st_uint_t j, double params[8]) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    for (int k = 0; k < 8; k++) el.p[k] = params[k];
    return 1;
}
STCORE_API int st_element_surface(st_context_t p, st_uint_t i, st_uint_t j, char surf) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    el.surf = surf;
    return 1;
}
STCORE_API int st_element_surface_params(st_context_t p, st_uint_t i, st_uint_t j, double params[8]) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    std::string msg;
    if (!translate_surface_params(el, params, msg)) {
        c->errlog("%s", msg.c_str());
        return -1;
    }
    return 1;
}
STCORE_API int st_element_surface_file(st_context_t p, st_uint_t i, st_uint_t j, const char *file) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    std::string msg;
    if (!read_surface_file(el, file, msg)) {
        c->errlog("%s", msg.c_str());
        return -1;
    }
    return 1;
}
STCORE_API int st_element_interaction(st_context_t p, st_uint_t i, st_uint_t j, int type) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    el.interaction = type;
    return 1;
}
STCORE_API int st_element_optic(st_context_t p, st_uint_t i, st_uint_t j, const char *name) {
    CTXW(p, -1);
    STAGE(i);
    ELEM(j);
    el.optic_name = name ? name : "";
    return 1;
}

// ------------------------------------------------------------------ sun ---
STCORE_API int st_sun(st_context_t p, int point_source, char shape, double sigma) {
    CTXW(p, -1);
    c->model.sun.point_source = point_source != 0;
    c->model.sun.shape = shape;
    c->model.sun.sigma = sigma;
    return 1;
}
STCORE_API int st_sun_xyz(st_context_t p, double x, double y, double z) {
    CTXW(p, -1);
    c->model.sun.origin[0] = x; c->model.sun.origin[1] = y; c->model.sun.origin[2] = z;
    return 1;
}
// coretrace/stapi.cpp:497-527
STCORE_API int st_sun_position(st_context_t p, double lat, double day, double hour, double *x, double *y,
                               double *z) {
    (void)p;
    const double d2r = M_PI / 180;
    double decl = 180 / M_PI * asin(0.39795 * cos(0.98563 * M_PI / 180 * (day - 173)));
    double ha = 15 * (hour - 12);
    double el = 180 / M_PI * asin(sin(decl * d2r) * sin(lat * d2r) + cos(decl * d2r) * cos(ha * d2r) * cos(lat * d2r));
    double az = 180 / M_PI *
                acos((sin(d2r * decl) * cos(d2r * lat) - cos(d2r * decl) * sin(d2r * lat) * cos(d2r * ha)) /
                         cos(d2r * el) +
                     0.0000000001);
    if (sin(ha * d2r) > 0.0) az = 360 - az;
    *x = -sin(az * d2r) * cos(el * d2r);
    *y = sin(el * d2r);
    *z = cos(az * d2r) * cos(el * d2r);
    return 1;
}
// coretrace/stapi.cpp:529-560 (table mirrored about zero)
STCORE_API int st_sun_userdata(st_context_t p, st_uint_t n, double angle[], double intensity[]) {
    CTXW(p, -1);
    if (n <= 2) return -1;
    HSun &s = c->model.sun;
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
    return (int)n;
}

// -------------------------------------------------------------- results ---
STCORE_API int st_num_intersections(st_context_t p) {
    CTX(p, -1);
    return (int)c->total_records();
}
#define MATERIALIZE()                                      \
    try {                                                  \
        c->materialize();                                  \
    } catch (const std::exception &e) {                    \
        c->errlog("result copy failed: %s", e.what());     \
        return -2;                                         \
    }
// The copies of one getter, then a wait for them: the caller's arrays are
// complete when the getter returns.
#define GETTER_COPY(...)                                   \
    try {                                                  \
        __VA_ARGS__;                                       \
        HC(hipStreamSynchronize(c->stream));               \
        return 1;                                          \
    } catch (const std::exception &e) {                    \
        c->errlog("result copy failed: %s", e.what());     \
        return -2;                                         \
    }
STCORE_API int st_locations(st_context_t p, double *x, double *y, double *z) {
    CTX(p, -1);
    MATERIALIZE();
    const size_t b = c->total_records() * sizeof(double);
    GETTER_COPY(c->copy_out(x, c->cols.x, b); c->copy_out(y, c->cols.y, b); c->copy_out(z, c->cols.z, b));
}
STCORE_API int st_cosines(st_context_t p, double *x, double *y, double *z) {
    CTX(p, -1);
    MATERIALIZE();
    const size_t b = c->total_records() * sizeof(double);
    GETTER_COPY(c->copy_out(x, c->cols.cx, b); c->copy_out(y, c->cols.cy, b); c->copy_out(z, c->cols.cz, b));
}
STCORE_API int st_elementmap(st_context_t p, int *m) {
    CTX(p, -1);
    MATERIALIZE();
    GETTER_COPY(c->copy_out(m, c->cols.element, c->total_records() * sizeof(int)));
}
STCORE_API int st_stagemap(st_context_t p, int *m) {
    CTX(p, -1);
    MATERIALIZE();
    GETTER_COPY(c->copy_out(m, c->cols.stage, c->total_records() * sizeof(int)));
}
STCORE_API int st_raynumbers(st_context_t p, int *m) {
    CTX(p, -1);
    MATERIALIZE();
    GETTER_COPY(c->copy_out(m, c->cols.raynum, c->total_records() * sizeof(int)));
}
STCORE_API int st_sun_stats(st_context_t p, double *xmin, double *xmax, double *ymin, double *ymax, int *nsun) {
    CTX(p, -1);
    if (xmin) *xmin = c->model.sun.minx;
    if (xmax) *xmax = c->model.sun.maxx;
    if (ymin) *ymin = c->model.sun.miny;
    if (ymax) *ymax = c->model.sun.maxy;
    if (nsun) *nsun = (int)c->sun_ray_count;
    return 1;
}

// ----------------------------------------------------------- simulation ---
STCORE_API int st_sim_params(st_context_t p, int raycount, int maxcount, int dyn) {
    CTX(p, -1);
    c->raycount = raycount;
    c->raymax = maxcount;
    c->dyn_group = dyn != 0;
    return 1;
}
STCORE_API int st_sim_errors(st_context_t p, int sun, int optics) {
    CTX(p, -1);
    c->err_sun = sun != 0;
    c->err_opt = optics != 0;
    return 1;
}
STCORE_API int st_sim_run(st_context_t p, unsigned int seed, st_callback_t cb, void *data) {
    CTX(p, -1);
    return c->run(seed, cb, data);
}

// coretrace/stapi.cpp:680-728 with plain arrays
STCORE_API int st_amd_sim_run_data(st_context_t p, unsigned int seed, const double *s1, st_uint_t n1, const double *s2,
                                   st_uint_t n2, int save_stage_data, st_callback_t cb, void *data) {
    CTX(p, -1);
    if ((n1 && !s1) || (n2 && !s2)) return -1;
    c->load_s1 = s1;
    c->load_n1 = n1;
    c->load_s2 = s2;
    c->load_n2 = n2;
    c->save_data = save_stage_data != 0;
    const int r = c->run(seed, cb, data);
    c->load_s1 = c->load_s2 = nullptr;
    c->load_n1 = c->load_n2 = 0;
    c->save_data = false;
    return r;
}
STCORE_API int st_amd_saved_stage_data(st_context_t p, int which, double *out) {
    CTX(p, -1);
    if (which != 1 && which != 2) return -1;
    const std::vector<double> &v = which == 1 ? c->saved_s1 : c->saved_s2;
    if (out) std::copy(v.begin(), v.end(), out);
    return (int)(v.size() / (which == 1 ? 8 : 7));
}

// The reference's own entry point, with C++ linkage as there
// (coretrace/stapi.cpp:680-685; its stapi.h declaration is commented out).
// Rows are appended to the vectors, as the reference appends them.  A loaded
// row of the wrong length keeps the previous row's values, as the reference's
// loaders leave their outputs untouched then (generateray.cpp:165-219).
__attribute__((visibility("default"))) int st_sim_run_data(st_context_t p, unsigned int seed,
                                                           std::vector<std::vector<double>> *data_s1,
                                                           std::vector<std::vector<double>> *data_s2,
                                                           bool save_st_data, st_callback_t cb, void *cbdata) {
    auto flat = [](const std::vector<std::vector<double>> *rows, size_t w) {
        std::vector<double> f;
        if (!rows) return f;
        std::vector<double> prev(w, 0.0);
        for (const auto &r : *rows) {
            if (r.size() == w) prev.assign(r.begin(), r.end());
            f.insert(f.end(), prev.begin(), prev.end());
        }
        return f;
    };
    const std::vector<double> f1 = flat(data_s1, 8), f2 = flat(data_s2, 7);
    const int n = st_amd_sim_run_data(p, seed, f1.data(), f1.size() / 8, f2.data(), f2.size() / 7,
                                      save_st_data ? 1 : 0, cb, cbdata);
    Context *c = ctx_of(p);
    if (n >= 0 && c && save_st_data) {
        if (data_s1)
            for (size_t i = 0; i < c->saved_s1.size(); i += 8)
                data_s1->emplace_back(c->saved_s1.begin() + i, c->saved_s1.begin() + i + 8);
        if (data_s2)
            for (size_t i = 0; i < c->saved_s2.size(); i += 7)
                data_s2->emplace_back(c->saved_s2.begin() + i, c->saved_s2.begin() + i + 7);
    }
    return n;
}

// ------------------------------------------------------------ utilities ---
STCORE_API void st_calc_euler_angles(double origin[3], double aimpoint[3], double zrot, double euler[3]) {
    euler[0] = euler[1] = euler[2] = 0;
    double e[3];
    if (euler_from_aim(origin, aimpoint, zrot, e))
        for (int i = 0; i < 3; i++) euler[i] = e[i];
}
// The utilities write their outputs in the order coretrace/mathproc.cpp:58-173
// does, so a caller passing one array as both input and output gets the
// reference's result.
static void matrix_vector_mult(const double m[3][3], const double *v, double *mxv) {
    mxv[0] = m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2];
    mxv[1] = m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2];
    mxv[2] = m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2];
}
STCORE_API void st_transform_to_local(double posref[3], double cosref[3], double origin[3], double R[3][3],
                                      double posloc[3], double cosloc[3]) {
    double d[3];
    for (int i = 0; i < 3; i++) d[i] = posref[i] - origin[i];
    matrix_vector_mult(R, d, posloc);
    matrix_vector_mult(R, cosref, cosloc);
}
STCORE_API void st_transform_to_reference(double posloc[3], double cosloc[3], double origin[3], double RT[3][3],
                                          double posref[3], double cosref[3]) {
    double d[3];
    matrix_vector_mult(RT, posloc, d);
    matrix_vector_mult(RT, cosloc, cosref);
    for (int i = 0; i < 3; i++) posref[i] = d[i] + origin[i];
}
STCORE_API void st_matrix_vector_mult(double m[3][3], double v[3], double mxv[3]) { matrix_vector_mult(m, v, mxv); }
STCORE_API void st_calc_transform_matrices(double euler[3], double R[3][3], double RT[3][3]) {
    double m[9];
    euler_matrix(euler, m);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) R[i][j] = m[3 * i + j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) RT[i][j] = R[j][i];
}
STCORE_API void st_matrix_transpose(double in[3][3], double out[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i][j] = in[j][i];
}

// --------------------------------------------------- st_amd extensions ----
STCORE_API int st_amd_set_precision(st_context_t p, int bits) {
    CTX(p, -1);
    if (bits != 32 && bits != 64) return -1;
    c->precision = bits;
    return 1;
}
STCORE_API int st_amd_set_cull(st_context_t p, int on) {
    CTXW(p, -1);
    c->cull = on != 0;
    return 1;
}
STCORE_API int st_amd_flux_request(st_context_t p, int stage, int elem, int nx, int ny, int absorbed_only,
                                   double xmin, double xmax, double ymin, double ymax) {
    CTX(p, -1);
    if (stage < 0) {
        c->flux = FluxSpec{};
        return 1;
    }
    if ((size_t)stage >= c->model.stages.size() || elem < 0 ||
        (size_t)elem >= c->model.stages[stage].elements.size() || nx <= 1 || ny <= 1)
        return -1;
    const HElement &e = c->model.stages[stage].elements[elem];
    char sf = (char)tolower(e.surf), ap = (char)tolower(e.shape);
    if (sf != 't' && sf != 'f' && !(ap == 'l' && sf == 's')) return -1;
    FluxSpec f{};
    f.active = 1;
    f.stage = stage;
    f.elem = elem + 1;
    f.nx = nx;
    f.ny = ny;
    f.absorbed_only = absorbed_only != 0;
    double gx = xmax - xmin, gy = ymax - ymin;
    // cylinder and partial-cylinder unrolling (app/src/project.cpp:1204-1222)
    if (sf == 't' || sf == 's') {
        double Rr = 1.0 / (sf == 't' ? e.curv_rev : e.vcx);
        f.radius = Rr;
        if (sf == 't') {
            f.kind = 1;
            xmin = -M_PI * Rr;
            gx = 2.0 * M_PI * Rr;
        } else {
            f.kind = 2;
            double amin = asin(e.p[0] / Rr), amax = asin(e.p[1] / Rr);
            xmin = Rr * amin;
            gx = (amax - amin) * Rr;
        }
    }
    if (!(gx > 0) || !(gy > 0)) return -1;
    f.xmin = xmin;
    f.ymin = ymin;
    f.bx = gx / nx;
    f.by = gy / ny;
    double eu[3];
    if (!euler_from_aim(e.origin, e.aim, e.zrot, eu)) return -1;
    euler_matrix(eu, f.R);
    for (int k = 0; k < 3; k++) f.org[k] = e.origin[k];
    c->flux = f;
    return 1;
}
// Order key of a double (st_stats.hip) and back.
static double from_order_key(unsigned long long k) {
    const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    memcpy(&v, &u, sizeof v);
    return v;
}
static unsigned long long host_order_key(double v) {
    unsigned long long u;
    memcpy(&u, &v, sizeof u);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// The GUI's element statistics (app/src/project.cpp:1139-1436) over the last
// run's records, binned on the GPU where the records are.  d_bins (nx*ny
// u32), d_acc (6 u64) and d_mm (4 u64) are device scratch; h_bins receives
// the counts (nx*ny, host).  Makes no allocation of its own.
static int element_stats_impl(Context *c, int stage, int elem, int nx, int ny, int autoscale, int absorbed_only,
                              double dni, double bounds[4], double *stats, uint32_t *d_bins,
                              unsigned long long *d_acc, unsigned long long *d_mm, unsigned int *h_bins) {
    const HElement &e = c->model.stages[stage].elements[elem];
    const char sf = (char)tolower(e.surf);
    FluxSpec f{};
    double eu[3];
    if (!euler_from_aim(e.origin, e.aim, e.zrot, eu)) return -1;
    euler_matrix(eu, f.R);
    for (int k = 0; k < 3; k++) f.org[k] = e.origin[k];
    const StageRecords &sr = c->recs[stage];
    hipStream_t s = c->stream;
    double minx = bounds[0], maxx = bounds[1], miny = bounds[2], maxy = bounds[3];
    if (autoscale) {
        const unsigned long long init[4] = {host_order_key(1e199), host_order_key(-1e199), host_order_key(1e199),
                                            host_order_key(-1e199)};
        HC(hipMemcpyAsync(d_mm, init, sizeof init, hipMemcpyHostToDevice, s));
        for (size_t k = 0; k < sr.chunks.size(); k++)
            launch_stats_bounds(sr.ptr(k), (uint32_t)sr.counts[k], elem + 1, f, d_mm, s);
        unsigned long long out[4];
        HC(hipMemcpyAsync(out, d_mm, sizeof out, hipMemcpyDeviceToHost, s));
        HC(hipStreamSynchronize(s));
        minx = from_order_key(out[0]);
        maxx = from_order_key(out[1]);
        miny = from_order_key(out[2]);
        maxy = from_order_key(out[3]);
    }
    if (maxx <= minx || maxy <= miny) return -1;
    double gx = maxx - minx, gy = maxy - miny;
    if (sf == 't' || sf == 's') {
        const double Rr = 1.0 / (sf == 't' ? e.curv_rev : e.vcx);  // SurfaceParams[0]
        f.radius = Rr;
        if (sf == 't') {
            f.kind = 1;
            minx = -M_PI * Rr;
            maxx = M_PI * Rr;
            gx = 2.0 * M_PI * Rr;
        } else {
            f.kind = 2;
            const double amin = asin(e.p[0] / Rr), amax = asin(e.p[1] / Rr);
            minx = Rr * amin;
            maxx = Rr * amax;
            gx = (amax - amin) * Rr;
        }
    }
    f.active = 1;
    f.nx = nx;
    f.ny = ny;
    f.absorbed_only = absorbed_only != 0;
    f.xmin = minx;
    f.ymin = miny;
    f.bx = gx / nx;
    f.by = gy / ny;
    const size_t nb = (size_t)nx * ny;
    launch_fill_u32(d_bins, nb, 0u, s);
    launch_fill_u32((uint32_t *)d_acc, 12, 0u, s);
    uint64_t base = 0;
    for (size_t k = 0; k < sr.chunks.size(); k++) {
        launch_stats_bin(sr.ptr(k), (uint32_t)sr.counts[k], elem + 1, f, base, d_bins, d_acc, s);
        base += sr.counts[k];
    }
    unsigned long long a[6];
    HC(hipMemcpyAsync(h_bins, d_bins, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HC(hipMemcpyAsync(a, d_acc, sizeof a, hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    // z value: the cylinder radius, or the z of the last ray considered
    double zval = 0.0;
    if (f.kind != 0) {
        zval = f.radius;
    } else if (a[5] > 0) {
        uint64_t idx = a[5] - 1;
        size_t k = 0;
        while (k < sr.chunks.size() && idx >= sr.counts[k]) idx -= sr.counts[k++];
        DevRecord r;
        HC(hipMemcpyAsync(&r, sr.ptr(k) + idx, sizeof r, hipMemcpyDeviceToHost, s));
        HC(hipStreamSynchronize(s));
        // the third row of st_transform_to_local (coretrace/mathproc.cpp)
        const double d[3] = {r.pos[0] - f.org[0], r.pos[1] - f.org[1], r.pos[2] - f.org[2]};
        zval = f.R[6] * d[0] + f.R[7] * d[1] + f.R[8] * d[2];
    }
    // statistics in the GUI's order (app/src/project.cpp:1248-1288)
    const double ppr = dni * (c->model.sun.maxx - c->model.sun.minx) * (c->model.sun.maxy - c->model.sun.miny) /
                       (double)c->sun_ray_count;
    const double zscale = ppr / (f.bx * f.by);
    double peak = 0, sum = 0, sum2 = 0, minf = 1e99;
    int npeak = -1;
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            const double z = (double)h_bins[(size_t)ix * ny + iy] * zscale;
            sum += z;
            sum2 += z * z;
            if (z > peak) {
                peak = z;
                npeak = (int)h_bins[(size_t)ix * ny + iy];
            }
            if (z < minf) minf = z;
        }
    const double nbins = (double)nx * (double)ny;
    const double ave = sum / nbins;
    const double sigma = sqrt((nbins * sum2 - sum * sum) / (nbins * nbins));
    const double considered = (double)a[3];
    auto centroid = [&](int q) { return considered > 0 ? (double)(int64_t)a[q] * 1e-6 / considered : 0.0; };
    const double out[15] = {ppr, peak, minf, ave, sigma, sigma / ave, 100.0 / sqrt((double)npeak),
                            100.0 / sqrt((double)c->total_records()), centroid(0), centroid(1), centroid(2),
                            (double)a[4], zval, f.bx, f.by};
    for (int q = 0; q < 15; q++) stats[q] = out[q];
    bounds[0] = minx;
    bounds[1] = maxx;
    bounds[2] = miny;
    bounds[3] = maxy;
    return 1;
}

// Arguments every element-statistics entry point accepts: an element the GUI
// can bin (flat, cylinder, partial cylinder), a run to bin, bins > 1.
static bool element_stats_args_ok(Context *c, int stage, int elem, int nx, int ny, const double *bounds,
                                  const double *stats) {
    if (stage < 0 || (size_t)stage >= c->model.stages.size() || elem < 0 ||
        (size_t)elem >= c->model.stages[stage].elements.size() || !bounds || !stats || nx <= 1 || ny <= 1)
        return false;
    const HElement &e = c->model.stages[stage].elements[elem];
    const char sf = (char)tolower(e.surf), ap = (char)tolower(e.shape);
    if (sf != 't' && sf != 'f' && !(ap == 'l' && sf == 's')) return false;
    return (size_t)stage < c->recs.size();  // a run to bin
}

STCORE_API int st_amd_element_stats(st_context_t p, int stage, int elem, int nx, int ny, int autoscale,
                                    int absorbed_only, double dni, double bounds[4], unsigned int *grid,
                                    double *stats) {
    CTX(p, -1);
    try {
        if (!element_stats_args_ok(c, stage, elem, nx, ny, bounds, stats)) return -1;
        const size_t nb = (size_t)nx * ny;
        DevBuf bins, acc, mm;
        bins.ensure(nb * sizeof(uint32_t));
        acc.ensure(6 * sizeof(unsigned long long));
        mm.ensure(4 * sizeof(unsigned long long));
        std::vector<uint32_t> h(nb);
        const int r = element_stats_impl(c, stage, elem, nx, ny, autoscale, absorbed_only, dni, bounds, stats,
                                         bins.as<uint32_t>(), acc.as<unsigned long long>(),
                                         mm.as<unsigned long long>(), h.data());
        if (r == 1 && grid)
            for (size_t q = 0; q < nb; q++) grid[q] = h[q];
        return r;
    } catch (const std::exception &ex) {
        c->errlog("element statistics: %s", ex.what());
        return -1;
    }
}

// st_amd_element_stats without a device allocation: the bins and
// accumulators live in this GPU's `scratch`, the counts go straight to the
// caller's grid.  Calls overlapping on one GPU share that scratch and race.
STCORE_API int st_amd_element_stats_noalloc(st_context_t p, int stage, int elem, int nx, int ny, int autoscale,
                                            int absorbed_only, double dni, double bounds[4], unsigned int *grid,
                                            double *stats) {
    CTX(p, -1);
    try {
        if (!grid || !element_stats_args_ok(c, stage, elem, nx, ny, bounds, stats)) return -1;
        if ((size_t)nx * (size_t)ny > kNoAllocMaxBins) {
            c->errlog("element statistics: %d x %d bins exceed the %zu the allocation-free entry point holds", nx,
                      ny, kNoAllocMaxBins);
            return -1;
        }
        unsigned char *sp = (unsigned char *)c->scratch_ptr;
        return element_stats_impl(c, stage, elem, nx, ny, autoscale, absorbed_only, dni, bounds, stats,
                                  (uint32_t *)sp, (unsigned long long *)(sp + kNoAllocAccOff),
                                  (unsigned long long *)(sp + kNoAllocMmOff), grid);
    } catch (const std::exception &ex) {
        c->errlog("element statistics: %s", ex.what());
        return -1;
    }
}

STCORE_API int st_amd_flux_by_source(st_context_t p, int on) {
    CTX(p, -1);
    c->flux_by_source = on != 0;
    return 1;
}
STCORE_API long long st_amd_flux_source_result(st_context_t p, unsigned int *counts, long long n) {
    CTX(p, -1);
    if (!c->flux.active || c->flux_src_words == 0) return -1;
    const long long words = (long long)c->flux_src_words;
    if (counts != nullptr) {
        if (n < words) return -1;
        try {
            HC(hipMemcpyAsync(counts, c->d_flux_src.p, (size_t)words * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              c->stream));
            HC(hipStreamSynchronize(c->stream));
        } catch (const std::exception &e) {
            c->errlog("per-source map copy failed: %s", e.what());
            return -1;
        }
    }
    return words;
}
STCORE_API long long st_amd_flux_source_result_device(st_context_t p, unsigned int *counts, long long n,
                                                      void *stream) {
    CTX(p, -1);
    if (!c->flux.active || c->flux_src_words == 0) return -1;
    const long long words = (long long)c->flux_src_words;
    if (counts != nullptr) {
        if (n < words) return -1;
        try {
            HC(hipMemcpyAsync(counts, c->d_flux_src.p, (size_t)words * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                              reinterpret_cast<hipStream_t>(stream)));
        } catch (const std::exception &e) {
            c->errlog("per-source map copy failed: %s", e.what());
            return -1;
        }
    }
    return words;
}
STCORE_API int st_amd_flux_result(st_context_t p, unsigned int *counts, double *bin_dims) {
    CTX(p, -1);
    if (!c->flux.active) return -1;
    for (size_t i = 0; i < c->flux_host.size(); i++) counts[i] = c->flux_host[i];
    if (bin_dims) {
        bin_dims[0] = c->flux.xmin;
        bin_dims[1] = c->flux.ymin;
        bin_dims[2] = c->flux.bx;
        bin_dims[3] = c->flux.by;
    }
    return (int)c->flux_host.size();
}
STCORE_API int st_amd_ray_sun_index(st_context_t p, unsigned long long *out, int n) {
    CTX(p, -1);
    if (n < 0) return -1;
    size_t m = std::min<size_t>((size_t)n, c->n_hits_done);
    if (m == 0) return 0;
    try {
        HC(hipMemcpyAsync(out, c->d_hits.p, m * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
        HC(hipStreamSynchronize(c->stream));
    } catch (const std::exception &e) {
        c->errlog("%s", e.what());
        return -1;
    }
    return (int)m;
}
STCORE_API int st_amd_count_events(st_context_t p, int on) {
    CTX(p, -1);
    c->count_events = on != 0;
    return 1;
}
// Launch counts and event counters of the last run: [0] sun-mask launches,
// [1..16] trace launches per stage, then 5 counters for the sun mask and 5
// per stage (rays, element tests, BVH nodes, uniforms drawn, records).
STCORE_API int st_amd_events(st_context_t p, double *out, int n) {
    CTX(p, -1);
    std::vector<double> v;
    v.push_back((double)c->timing.mask_launches);
    for (int i = 0; i < kMaxStages; i++) v.push_back((double)c->timing.trace_launches[i]);
    for (int k = 0; k < 1 + kMaxStages; k++)
        for (int j = 0; j < 5; j++) v.push_back((double)c->timing.ev[k][j]);
    v.push_back((double)c->timing.ev_sun[0]);
    v.push_back((double)c->timing.ev_sun[1]);
    int m = std::min<int>(n, (int)v.size());
    for (int i = 0; i < m; i++) out[i] = v[i];
    return (int)v.size();
}
// Timing of the last run: [0] mask ms, [1..kMaxStages] trace ms per stage,
// [17] other ms, [18] total ms, [19] sun rays, [20..] rays per stage,
// [36..] records per stage, [52] sampler ms.
STCORE_API int st_amd_timing(st_context_t p, double *out, int n) {
    CTX(p, -1);
    c->settle_timing();
    std::vector<double> v;
    v.push_back(c->timing.mask_ms);
    for (int i = 0; i < kMaxStages; i++) v.push_back(c->timing.trace_ms[i]);
    v.push_back(c->timing.other_ms);
    v.push_back(c->timing.total_ms);
    v.push_back((double)c->timing.sun_rays);
    for (int i = 0; i < kMaxStages; i++) v.push_back((double)c->timing.rays_in[i]);
    for (int i = 0; i < kMaxStages; i++) v.push_back((double)c->timing.records[i]);
    v.push_back(c->timing.sunpre_ms);
    int m = std::min<int>(n, (int)v.size());
    for (int i = 0; i < m; i++) out[i] = v[i];
    return (int)v.size();
}
