// Batched runs: st_amd_sim_run_batch traces several contexts' scenes
// together, each as its own st_sim_run would (coretrace/stapi.cpp:680-728,
// raytrace.cpp:137-1113), with every kernel launch covering all of them.
//
// Per stage-1 round, for every scene still short of its ray hits:
//   1. one persistent sun-mask launch over every scene's next batch of sun
//      rays (per-wave tiles drawn from a device queue), the masks laid end to
//      end;
//   2. one popcount and one scan over all masks; the hit count of each scene
//      (a small gather) decides how many of its hits it takes;
//   3. one compaction writing each scene's hits, in sun-index order, to its
//      hit list;
//   4. per chunk of at most 4M rays across the scenes: one tabulated-
//      sunshape launch, one persistent trace launch (per kernel variant), one
//      record scan + compaction and one ray scan + compaction, with each
//      scene's records and passed rays contiguous and in ray order.
// Later stages trace every scene's passed rays the same way.  Each scene's
// records, ray numbers and sun-ray count are therefore those of its own run:
// the per-ray work and its random streams do not depend on what else shares
// a launch.
//
// The scenes are split over two lanes, each its own stream and host thread,
// so that one lane's masks overlap the other's traces.
//
// The trace launches' shape (rays per tile, waves per SIMD) is chosen per
// scene class (precision, kernel variant, first or later stage) by timing
// the candidates on the first launch of that class; the choice is kept for
// the process (st_amd_batch_tuning).
#include "st_context.h"

#include <map>
#include <thread>

namespace st {

namespace {

constexpr uint64_t kRoundSunMax = 1ull << 27;  // sun rays per stage-1 round, all scenes
constexpr uint32_t kChunkRays = 1u << 22;      // rays per trace launch, all scenes

struct Scene {
    Context *c = nullptr;
    uint32_t seed = 0;
    int nst = 0;
    uint64_t N = 0, maxrays = 0;
    bool presample = false;
    int result = 0;        // -1 once the scene failed
    // stage 1
    uint64_t hits = 0, s0 = 0, sun_count = 0;
    double rate = 0.0;
    bool done = false, maxed = false;
    // rays entering the current stage, and those it hands on
    DevRay *cur = nullptr, *nxt = nullptr;
    uint64_t n_cur = 0, n_nxt = 0;
    bool prev_has = false;
    uint32_t last_prev = 0;
};

// Segment of a trace chunk: rays [first, first + n) of scene k's stage input.
struct Seg {
    int k;
    uint64_t first, n;
    uint32_t base;  // place in the chunk's ray space
};

}  // namespace

struct BatchWork {
    DevBuf mask, popc, offs, scan_tmp, totals, pos, pre_out;
    DevBuf mask_items, compact_items, pre_items, trace_items, copy_items, last_items, prefix, queue;
    DevBuf rec_slots, rec_count, rec_offs, rec_tmp, ovf_rec, ovf_meta, ovf_count;
    DevBuf out_slots, out_flag, out_offs, out_tmp, rays_round, sunpre;
    DevBuf err, fe_fail, counters;
    DevBuf bucket_items, keyrank, hist, hoffs, htmp, htot, order;  // stage-0 ray order by sun-plane bucket
    std::vector<std::shared_ptr<DevBuf>> arenas;  // record buffers; reused once no context holds them
    uint32_t *h = nullptr;                        // pinned staging for small reads
    size_t h_words = 0;
    // pinned staging for the launches' small uploads (item lists, prefixes):
    // filled front to back, reset whenever the lane's stream is idle
    char *pin = nullptr;
    size_t pin_cap = 0, pin_off = 0;
    int cus = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    BatchTiming timing;
    ~BatchWork() {
        if (h) (void)hipHostFree(h);
        if (pin) (void)hipHostFree(pin);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    uint32_t *host(size_t words) {
        if (words > h_words) {
            if (h) HC(hipHostFree(h));
            h = nullptr;
            HC(hipHostMalloc((void **)&h, words * sizeof(uint32_t), hipHostMallocDefault));
            h_words = words;
        }
        return h;
    }
    std::shared_ptr<DevBuf> arena(size_t bytes) {
        for (auto &a : arenas)
            if (a.use_count() == 1 && a->cap >= bytes) return a;
        for (auto &a : arenas)
            if (a.use_count() == 1) {
                a->ensure(bytes);
                return a;
            }
        arenas.push_back(std::make_shared<DevBuf>());
        arenas.back()->ensure(bytes);
        return arenas.back();
    }
};

namespace {

// Upload a small host array through the lane's pinned staging (one DMA
// copy, no pageable staging), in stream order.
template <typename T>
T *put(BatchWork &W, DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    constexpr size_t kPin = (size_t)4 << 20;
    b.ensure(std::max<size_t>(1, v.size()) * sizeof(T));
    if (v.empty()) return b.as<T>();
    const size_t bytes = v.size() * sizeof(T), need = (bytes + 255) & ~(size_t)255;
    if (need > kPin) return upload(b, v, s);
    if (!W.pin) {
        HC(hipHostMalloc((void **)&W.pin, kPin, hipHostMallocDefault));
        W.pin_cap = kPin;
        W.pin_off = 0;
    }
    if (W.pin_off + need > W.pin_cap) {  // earlier copies must finish before their staging is reused
        HC(hipStreamSynchronize(s));
        W.pin_off = 0;
    }
    memcpy(W.pin + W.pin_off, v.data(), bytes);
    HC(hipMemcpyAsync(b.p, W.pin + W.pin_off, bytes, hipMemcpyHostToDevice, s));
    W.pin_off += need;
    return b.as<T>();
}
// The lane's stream is idle: its staging can be reused from the start.
void idle(BatchWork &W, hipStream_t s) {
    HC(hipStreamSynchronize(s));
    W.pin_off = 0;
}

// Trace-launch shapes chosen per class: key = bits * 100 + generic * 10 +
// (later stage ? 1 : 0).
std::mutex tune_mu;
std::map<int, TraceShape> tuned;

int trace_grid(const BatchWork &W, const TraceShape &sh, int generic) {
    return W.cus * (generic == 2 ? 1 : sh.waves);
}

}  // namespace

int trace_batch_waves(int bits, int generic, int *out, int max) {
    int v[3], n = 0;
    if (generic == 1) { v[n++] = 2; v[n++] = 1; }
    else if (generic) v[n++] = 1;
    else if (bits == 64) { v[n++] = 2; v[n++] = 1; }
    else { v[n++] = 4; v[n++] = 3; v[n++] = 2; }
    for (int i = 0; i < n && i < max; i++) out[i] = v[i];
    return n;
}

template <typename R>
static bool run_batch(std::vector<Scene> &sc, BatchWork &W, bool count, hipStream_t s) {
    constexpr int bits = sizeof(R) == 8 ? 64 : 32;
    const int n = (int)sc.size();
    W.timing = BatchTiming();
    auto t_start = std::chrono::steady_clock::now();
    for (hipEvent_t &e : W.ev)
        if (!e) HC(hipEventCreate(&e));
    auto elapsed = [&]() {
        float ms = 0;
        HC(hipEventElapsedTime(&ms, W.ev[0], W.ev[1]));
        return (double)ms;
    };

    // per-scene error words, finite-element failure indices, counters
    uint32_t *err = (uint32_t *)W.err.ensure((size_t)n * sizeof(uint32_t));
    unsigned long long *fe_fail = (unsigned long long *)W.fe_fail.ensure((size_t)n * sizeof(unsigned long long));
    launch_fill_u32(err, (size_t)n, 0u, s);
    launch_fill_u32((uint32_t *)fe_fail, 2 * (size_t)n, 0xFFFFFFFFu, s);
    const size_t ctr_words = (size_t)(1 + kMaxStages) * 5 + 2;
    unsigned long long *ctr = nullptr;
    if (count) {
        ctr = (unsigned long long *)W.counters.ensure((size_t)n * ctr_words * sizeof(unsigned long long));
        launch_fill_u32((uint32_t *)ctr, (size_t)n * ctr_words * 2, 0u, s);
    }
    uint32_t *queue = (uint32_t *)W.queue.ensure(64);
    uint32_t *ovf_count = (uint32_t *)W.ovf_count.ensure(sizeof(uint32_t));
    const uint32_t ovf_cap = kChunkRays / 2;
    W.ovf_rec.ensure((size_t)ovf_cap * sizeof(DevRecord));
    W.ovf_meta.ensure((size_t)ovf_cap * 2 * sizeof(uint32_t));
    W.rec_slots.ensure((size_t)kChunkRays * kSlots * sizeof(DevRecord));
    W.rec_count.ensure((size_t)kChunkRays * sizeof(uint32_t));
    W.rec_offs.ensure((size_t)kChunkRays * sizeof(uint32_t));
    W.rec_tmp.ensure(scan_tmp_words(kChunkRays) * sizeof(uint32_t));
    W.out_slots.ensure((size_t)kChunkRays * sizeof(DevRay));
    W.out_flag.ensure((size_t)kChunkRays * sizeof(uint32_t));
    W.out_offs.ensure((size_t)kChunkRays * sizeof(uint32_t));
    W.out_tmp.ensure(scan_tmp_words(kChunkRays) * sizeof(uint32_t));
    W.rays_round.ensure((size_t)kChunkRays * sizeof(DevRay));
    W.sunpre.ensure((size_t)kChunkRays * sizeof(SunPre));
    uint32_t *totals = (uint32_t *)W.totals.ensure(64);

    // prefix positions -> exclusive prefix values, read to the host
    auto prefix_read = [&](const uint32_t *offs, const uint32_t *tmp, size_t len, const uint32_t *total,
                           const std::vector<uint32_t> &pos) -> std::vector<uint32_t> {
        const uint32_t *dpos = put(W, W.pos, pos, s);
        uint32_t *dout = (uint32_t *)W.pre_out.ensure(pos.size() * sizeof(uint32_t));
        launch_prefix_at(offs, tmp, len, total, dpos, (int)pos.size(), dout, s);
        uint32_t *h = W.host(pos.size());
        HC(hipMemcpyAsync(h, dout, pos.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        idle(W, s);
        return std::vector<uint32_t>(h, h + pos.size());
    };
    auto check_errors = [&]() {
        uint32_t *h = W.host((size_t)n);
        HC(hipMemcpyAsync(h, err, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        idle(W, s);
        for (int k = 0; k < n; k++) {
            if (!h[k] || sc[k].result < 0) continue;
            Context *c = sc[k].c;
            const uint32_t e = h[k];
            if (e & 16u) c->errlog("trace error: Interpolation failed to find nearby nodes to ray hit in FEA surface.");
            else if (e & 1u) c->errlog("Bad optical interaction type (stage)");
            else if (e & 2u) c->errlog("internal error: sun-ray hit test disagreed between launches");
            else if (e & 4u) c->errlog("internal error: record overflow buffer full");
            else c->errlog("sunshape table sampling found no accepted angle in 2^26 tries");
            sc[k].result = -1;
        }
    };

    // One stage over chunks of the scenes' input rays.  Stage 0 reads the
    // round's new hits (sun indices at hit_list + first), later stages the
    // rays of Scene::cur.
    auto trace_stage = [&](int si, std::vector<Seg> segs_all) {
        size_t i0 = 0;
        while (i0 < segs_all.size()) {
            // a chunk: whole or split segments up to kChunkRays rays
            std::vector<Seg> segs;
            uint32_t tot = 0;
            size_t i = i0;
            while (i < segs_all.size() && tot < kChunkRays) {
                Seg g = segs_all[i];
                const uint64_t room = kChunkRays - tot;
                if (g.n > room) {  // split: the rest opens the next chunk
                    Seg head = g;
                    head.n = room;
                    head.base = tot;
                    segs.push_back(head);
                    tot += (uint32_t)room;
                    segs_all[i].first += room;
                    segs_all[i].n -= room;
                    break;
                }
                g.base = tot;
                segs.push_back(g);
                tot += (uint32_t)g.n;
                i++;
            }
            i0 = i;
            if (tot == 0) continue;
            // tabulated-sunshape angles of the chunk's stage-1 rays
            SunPre *pre = W.sunpre.as<SunPre>();
            if (si == 0) {
                std::vector<SunPreArgs> items;
                std::vector<uint32_t> pfx;
                uint32_t acc = 0;
                for (const Seg &g : segs) {
                    const Scene &S = sc[g.k];
                    if (!S.presample) continue;
                    SunPreArgs pa{};
                    pa.sun = S.c->scene_view.sun;
                    pa.aux = S.c->scene_view.aux;
                    pa.seed = S.seed;
                    pa.k0 = 3;
                    pa.sun_index = S.c->d_hits.as<uint64_t>() + g.first;
                    pa.n = (int32_t)g.n;
                    pa.out = pre + g.base;
                    pa.err_flags = err + g.k;
                    pa.counters = ctr ? ctr + (size_t)g.k * ctr_words + 5 * (1 + kMaxStages) : nullptr;
                    items.push_back(pa);
                    pfx.push_back(acc);
                    acc += (uint32_t)g.n;
                }
                if (acc) {
                    pfx.push_back(acc);
                    BatchQueue q{put(W, W.prefix, pfx, s), (int32_t)items.size(), 1u, queue};
                    HC(hipEventRecord(W.ev[0], s));
                    launch_sunshape_batch<R>(put(W, W.pre_items, items, s), q, acc, s);
                    HC(hipEventRecord(W.ev[1], s));
                    HC(hipEventSynchronize(W.ev[1]));
                    W.timing.sunshape_ms += elapsed();
                    W.timing.sunshape_launches++;
                }
            }
            // stage-0 rays assigned to lanes by sun-plane bucket within each
            // scene (keys: segment * 64^2 + bucket), when a scene's stage 0
            // has a BVH for its reflected rays to walk
            const uint32_t *order = nullptr;
            if (si == 0) {
                bool any = false;
                for (const Seg &g : segs) {
                    const Context *c = sc[g.k].c;
                    any = any || (c->scene_view.use_grid && !c->scene_view.sun.point_source &&
                                  c->cs.stages[0].nbvh > 0);
                }
                if (any) {
                    constexpr uint32_t per = 1u << (2 * kBatchBucketLog2);
                    std::vector<BucketItem> bi;
                    std::vector<uint32_t> bp;
                    for (size_t j = 0; j < segs.size(); j++) {
                        const Seg &g = segs[j];
                        const Scene &S = sc[g.k];
                        BucketItem it{};
                        it.sun = S.c->scene_view.sun;
                        it.seed = S.seed;
                        it.key0 = (uint32_t)j * per;
                        it.sun_index = S.c->d_hits.as<uint64_t>() + g.first;
                        bi.push_back(it);
                        bp.push_back(g.base);
                    }
                    bp.push_back(tot);
                    const uint32_t nkeys = (uint32_t)segs.size() * per;
                    uint32_t *kr = (uint32_t *)W.keyrank.ensure((size_t)tot * 2 * sizeof(uint32_t));
                    uint32_t *hist = (uint32_t *)W.hist.ensure((size_t)nkeys * sizeof(uint32_t));
                    uint32_t *hoffs = (uint32_t *)W.hoffs.ensure((size_t)nkeys * sizeof(uint32_t));
                    uint32_t *htmp = (uint32_t *)W.htmp.ensure(scan_tmp_words(nkeys) * sizeof(uint32_t));
                    uint32_t *htot = (uint32_t *)W.htot.ensure(sizeof(uint32_t));
                    uint32_t *ord = (uint32_t *)W.order.ensure((size_t)tot * sizeof(uint32_t));
                    BatchQueue bq{put(W, W.prefix, bp, s), (int32_t)bi.size(), 1u, queue};
                    launch_sun_buckets_batch<R>(put(W, W.bucket_items, bi, s), bq, tot, kr, hist, nkeys, s);
                    launch_tile_scan(hist, hoffs, nkeys, htmp, htot, s);
                    launch_order_scatter(kr, (int)tot, hoffs, htmp, ord, s);
                    order = ord;
                }
            }
            // the trace launches, one per kernel variant present
            std::vector<TraceArgs> all(segs.size());
            for (size_t j = 0; j < segs.size(); j++) {
                const Seg &g = segs[j];
                const Scene &S = sc[g.k];
                Context *c = S.c;
                TraceArgs a{};
                a.sv = c->scene_view;
                a.si = si;
                a.seed = S.seed;
                a.inc_sun = c->err_sun;
                a.inc_err = c->err_opt;
                a.n_rays = (int32_t)g.n;
                a.ray_base = g.base;
                a.slot_stride = tot;
                a.slots = kSlots;
                if (si == 0) {
                    a.sun_index = c->d_hits.as<uint64_t>() + g.first;
                    a.raynum_base = (uint32_t)g.first;
                    a.sun_pre = S.presample ? pre + g.base : nullptr;
                } else {
                    a.in_rays = S.cur + g.first;
                    a.last_raynum = S.last_prev;
                }
                a.rec_slots = W.rec_slots.as<DevRecord>();
                a.rec_count = W.rec_count.as<uint32_t>();
                a.ovf_rec = W.ovf_rec.as<DevRecord>();
                a.ovf_meta = W.ovf_meta.as<uint32_t>();
                a.ovf_count = ovf_count;
                a.ovf_cap = ovf_cap;
                a.out_slots = W.out_slots.as<DevRay>();
                a.out_flag = W.out_flag.as<uint32_t>();
                a.err_flags = err + g.k;
                a.flux = FluxSpec{};
                a.flux_bins = nullptr;
                a.counters = ctr ? ctr + (size_t)g.k * ctr_words + 5 * (1 + si) : nullptr;
                a.generic = c->cs.stages[si].generic;
                a.order = order;
                all[j] = a;
            }
            launch_fill_u32(ovf_count, 1, 0u, s);
            for (int gv = 0; gv <= 2; gv++) {
                std::vector<TraceArgs> items;
                for (const TraceArgs &a : all)
                    if (a.generic == gv) items.push_back(a);
                if (items.empty()) continue;
                const int key = bits * 100 + gv * 10 + (si > 0 ? 1 : 0);
                TraceShape shape{};
                bool have = false;
                {
                    std::lock_guard<std::mutex> lk(tune_mu);
                    auto it = tuned.find(key);
                    if (it != tuned.end()) {
                        shape = it->second;
                        have = true;
                    }
                }
                const TraceArgs *ditems = put(W, W.trace_items, items, s);
                auto launch = [&](const TraceShape &sh) {
                    std::vector<uint32_t> pfx(items.size() + 1, 0u);
                    for (size_t j = 0; j < items.size(); j++)
                        pfx[j + 1] = pfx[j] + ((uint32_t)items[j].n_rays + sh.tile - 1) / sh.tile;
                    BatchQueue q{put(W, W.prefix, pfx, s), (int32_t)items.size(), (uint32_t)sh.tile, queue};
                    launch_fill_u32(queue, 1, 0u, s);
                    launch_fill_u32(ovf_count, 1, 0u, s);
                    HC(hipEventRecord(W.ev[0], s));
                    launch_trace_batch<R>(ditems, q, gv, sh, trace_grid(W, sh, gv), s);
                    HC(hipEventRecord(W.ev[1], s));
                    HC(hipEventSynchronize(W.ev[1]));
                    return elapsed();
                };
                if (!have) {
                    // autotune: every candidate on this launch (each rewrites
                    // the same outputs from the same inputs); keep the fastest
                    int waves[3];
                    const int nw = trace_batch_waves(bits, gv, waves, 3);
                    double best = 1e30;
                    for (int w = 0; w < nw; w++)
                        for (int tile : {64, 128, 256, 512}) {
                            TraceShape sh{tile, waves[w]};
                            const double ms = launch(sh);
                            if (ms < best) {
                                best = ms;
                                shape = sh;
                            }
                        }
                    std::lock_guard<std::mutex> lk(tune_mu);
                    tuned[key] = shape;
                }
                W.timing.trace_ms[si] += launch(shape);
                W.timing.trace_launches[si]++;
            }
            W.timing.rays_in[si] += tot;
            // records: one scan over the chunk's ray space, one compaction
            launch_tile_scan(W.rec_count.as<uint32_t>(), W.rec_offs.as<uint32_t>(), tot, W.rec_tmp.as<uint32_t>(),
                             totals, s);
            launch_tile_scan(W.out_flag.as<uint32_t>(), W.out_offs.as<uint32_t>(), tot, W.out_tmp.as<uint32_t>(),
                             totals + 1, s);
            std::vector<uint32_t> pos;
            for (const Seg &g : segs) pos.push_back(g.base);
            pos.push_back(tot);
            const std::vector<uint32_t> rp = prefix_read(W.rec_offs.as<uint32_t>(), W.rec_tmp.as<uint32_t>(), tot,
                                                         totals, pos);
            const std::vector<uint32_t> op = prefix_read(W.out_offs.as<uint32_t>(), W.out_tmp.as<uint32_t>(), tot,
                                                         totals + 1, pos);
            check_errors();
            const uint32_t nrec = rp.back();
            std::shared_ptr<DevBuf> arena = W.arena((size_t)std::max<uint32_t>(1, nrec) * sizeof(DevRecord));
            launch_compact_records(W.rec_slots.as<DevRecord>(), W.rec_count.as<uint32_t>(), W.rec_offs.as<uint32_t>(),
                                   W.rec_tmp.as<uint32_t>(), (int)tot, kSlots, arena->as<DevRecord>(), s);
            launch_compact_overflow(W.ovf_rec.as<DevRecord>(), W.ovf_meta.as<uint32_t>(), ovf_count, ovf_cap,
                                    W.rec_offs.as<uint32_t>(), W.rec_tmp.as<uint32_t>(), arena->as<DevRecord>(), s);
            for (size_t j = 0; j < segs.size(); j++) {
                Scene &S = sc[segs[j].k];
                const uint32_t cnt = rp[j + 1] - rp[j];
                S.c->recs[si].add(arena, rp[j], cnt, false);
                S.c->timing.records[si] += cnt;
            }
            W.timing.records[si] += nrec;
            // passed rays: compacted for the chunk, then appended to each
            // scene's next-stage rays
            const uint32_t npass = op.back();
            if (npass) {
                launch_compact_rays(W.out_slots.as<DevRay>(), W.out_flag.as<uint32_t>(), W.out_offs.as<uint32_t>(),
                                    W.out_tmp.as<uint32_t>(), (int)tot, W.rays_round.as<DevRay>(), s);
                std::vector<RayCopyItem> items;
                std::vector<uint32_t> pfx;
                uint32_t acc = 0;
                for (size_t j = 0; j < segs.size(); j++) {
                    Scene &S = sc[segs[j].k];
                    const uint32_t cnt = op[j + 1] - op[j];
                    if (!cnt) continue;
                    items.push_back(RayCopyItem{W.rays_round.as<DevRay>() + op[j], S.nxt + S.n_nxt});
                    pfx.push_back(acc);
                    acc += cnt;
                    S.n_nxt += cnt;
                }
                pfx.push_back(acc);
                BatchQueue q{put(W, W.prefix, pfx, s), (int32_t)items.size(), 1u, queue};
                launch_copy_rays_batch(put(W, W.copy_items, items, s), q, acc, s);
            }
            idle(W, s);
        }
    };

    // ------------------------------------------------------ stage 1 -------
    for (;;) {
        // plan each unfinished scene's batch of sun rays, as a single run plans
        std::vector<int> act;
        std::vector<uint64_t> B;
        uint64_t budget = kRoundSunMax;
        for (int k = 0; k < n; k++) {
            Scene &S = sc[k];
            if (S.done || S.result < 0) continue;
            const uint64_t want = S.N - S.hits;
            uint64_t b = (S.rate > 0) ? (uint64_t)(want / S.rate * 1.05) + 4096 : std::min<uint64_t>(1u << 22, want * 4 + 4096);
            b = std::min<uint64_t>(b, 1u << 25);
            b = std::max<uint64_t>(b, 4096);
            if (S.s0 + b > S.maxrays) b = S.maxrays - S.s0;
            if (b == 0) {
                S.maxed = true;
                S.done = true;
                continue;
            }
            if (b > budget && !act.empty()) break;  // the rest waits for the next round
            b = std::min(b, std::max<uint64_t>(budget, 4096));
            act.push_back(k);
            B.push_back(b);
            budget = budget > b ? budget - b : 0;
        }
        if (act.empty()) break;
        const int m = (int)act.size();
        std::vector<uint32_t> wpfx(m + 1, 0u);  // mask words
        for (int j = 0; j < m; j++) wpfx[j + 1] = wpfx[j] + (uint32_t)((B[j] + 63) / 64);
        const size_t nwords = wpfx[m];
        uint64_t *mask = (uint64_t *)W.mask.ensure(nwords * sizeof(uint64_t));
        uint32_t *popc = (uint32_t *)W.popc.ensure(nwords * sizeof(uint32_t));
        uint32_t *offs = (uint32_t *)W.offs.ensure(nwords * sizeof(uint32_t));
        uint32_t *stmp = (uint32_t *)W.scan_tmp.ensure(scan_tmp_words(nwords) * sizeof(uint32_t));
        // sun masks, one persistent launch per kernel variant
        std::vector<MaskArgs> mitems(m);
        for (int j = 0; j < m; j++) {
            const Scene &S = sc[act[j]];
            MaskArgs ma{};
            ma.sv = S.c->scene_view;
            ma.seed = S.seed;
            ma.s0 = S.s0;
            ma.n = (uint32_t)B[j];
            ma.mask = mask + wpfx[j];
            ma.fe_fail = fe_fail + act[j];
            ma.counters = ctr ? ctr + (size_t)act[j] * ctr_words : nullptr;
            ma.generic = S.c->cs.stages[0].generic;
            ma.compact = ma.sv.use_grid && !ma.sv.sun.point_source && S.c->cs.stages[0].nelem < (1 << 24);
            mitems[j] = ma;
        }
        for (int gv = 0; gv <= 2; gv++) {
            std::vector<MaskArgs> items;
            std::vector<uint32_t> tp(1, 0u);
            for (const MaskArgs &ma : mitems)
                if (ma.generic == gv) {
                    items.push_back(ma);
                    tp.push_back(tp.back() + (ma.n + 64 * kMaskRays - 1) / (64 * kMaskRays));
                }
            if (items.empty()) continue;
            BatchQueue q{put(W, W.prefix, tp, s), (int32_t)items.size(), 0u, queue};
            launch_fill_u32(queue, 1, 0u, s);
            HC(hipEventRecord(W.ev[0], s));
            launch_mask_batch<R>(put(W, W.mask_items, items, s), q, gv, W.cus * 8, s);
            HC(hipEventRecord(W.ev[1], s));
            HC(hipEventSynchronize(W.ev[1]));
            W.timing.mask_ms += elapsed();
            W.timing.mask_launches++;
        }
        for (int j = 0; j < m; j++) W.timing.sun_rays += B[j];
        launch_mask_popc(mask, popc, nwords, s);
        launch_tile_scan(popc, offs, nwords, stmp, totals + 2, s);
        const std::vector<uint32_t> hp = prefix_read(offs, stmp, nwords, totals + 2, wpfx);
        // each scene's hits in sun-index order, as many as it still needs
        std::vector<MaskCompactItem> citems(m);
        std::vector<uint64_t> take(m);
        for (int j = 0; j < m; j++) {
            Scene &S = sc[act[j]];
            const uint64_t h = hp[j + 1] - hp[j];
            take[j] = std::min<uint64_t>(h, S.N - S.hits);
            citems[j] = MaskCompactItem{S.s0, S.c->d_hits.as<uint64_t>() + S.hits, (uint32_t)take[j]};
        }
        BatchQueue cq{put(W, W.prefix, wpfx, s), (int32_t)m, 0u, queue};
        launch_mask_compact_batch(mask, offs, stmp, nwords, put(W, W.compact_items, citems, s), cq, s);
        idle(W, s);
        // sun-ray counts: up to the last hit taken when a scene is complete
        std::vector<Seg> segs;
        for (int j = 0; j < m; j++) {
            Scene &S = sc[act[j]];
            if (take[j] && S.hits + take[j] == S.N) {
                uint64_t last = 0;
                HC(hipMemcpyAsync(&last, S.c->d_hits.as<uint64_t>() + S.hits + take[j] - 1, sizeof last,
                                  hipMemcpyDeviceToHost, s));
                idle(W, s);
                S.sun_count = last + 1;
            } else {
                S.sun_count = S.s0 + B[j];
            }
            if (take[j]) segs.push_back(Seg{act[j], S.hits, take[j], 0});
            S.hits += take[j];
            S.s0 += B[j];
            S.rate = (double)S.hits / (double)S.s0;
            if (S.hits >= S.N) S.done = true;
        }
        trace_stage(0, segs);
    }
    // finite-element failures of the masks count when the scene's run
    // reaches that sun ray
    {
        std::vector<unsigned long long> f(n);
        HC(hipMemcpyAsync(f.data(), fe_fail, (size_t)n * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        idle(W, s);
        for (int k = 0; k < n; k++)
            if (sc[k].result >= 0 && sc[k].c->cs.stages[0].generic == 2 && f[k] < sc[k].sun_count) {
                sc[k].c->errlog("trace error: Interpolation failed to find nearby nodes to ray hit in FEA surface.");
                sc[k].result = -1;
            }
    }
    for (Scene &S : sc)
        if (S.result >= 0 && S.maxed) {
            S.c->errlog("generated sun rays reached maximum count: %d", S.c->raymax);
            S.result = -1;
        }

    // stage-1 pass-on rule (coretrace/raytrace.cpp:786-797, 1081-1085) and
    // each scene's last ray number, read together
    auto read_last = [&]() {
        std::vector<LastRayItem> items(n);
        for (int k = 0; k < n; k++) items[k] = LastRayItem{sc[k].nxt, sc[k].n_nxt};
        uint32_t *d = (uint32_t *)W.pre_out.ensure((size_t)n * sizeof(uint32_t));
        launch_last_raynum(put(W, W.last_items, items, s), n, d, s);
        uint32_t *h = W.host((size_t)n);
        HC(hipMemcpyAsync(h, d, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        idle(W, s);
        return std::vector<uint32_t>(h, h + n);
    };
    {
        const std::vector<uint32_t> last = read_last();
        for (int k = 0; k < n; k++) {
            Scene &S = sc[k];
            S.cur = S.nxt;
            S.n_cur = S.n_nxt;
            S.prev_has = false;
            if (S.result < 0 || S.nst < 2 || S.n_cur == 0) continue;
            S.last_prev = last[k];
            const uint64_t below = (S.last_prev == S.N) ? S.n_cur - 1 : S.n_cur;
            S.prev_has = below > 0;
        }
    }
    // ------------------------------------------------- later stages -------
    for (int si = 1; si < kMaxStages; si++) {
        std::vector<Seg> segs;
        for (int k = 0; k < n; k++) {
            Scene &S = sc[k];
            if (S.result < 0 || si >= S.nst || !S.prev_has) continue;
            // the other ray buffer receives what this stage hands on
            S.nxt = (S.cur == S.c->d_rays_a.as<DevRay>()) ? (DevRay *)S.c->d_rays_b.ensure(std::max<uint64_t>(S.n_cur, 1) * sizeof(DevRay))
                                                           : (DevRay *)S.c->d_rays_a.ensure(std::max<uint64_t>(S.n_cur, 1) * sizeof(DevRay));
            S.n_nxt = 0;
            segs.push_back(Seg{k, 0, S.n_cur, 0});
        }
        if (segs.empty()) break;
        trace_stage(si, segs);
        const std::vector<uint32_t> last = read_last();
        for (const Seg &g : segs) {
            Scene &S = sc[g.k];
            S.cur = S.nxt;
            S.n_cur = S.n_nxt;
            S.prev_has = S.n_cur > 0 && si + 1 < S.nst;
            if (S.prev_has) S.last_prev = last[g.k];
        }
    }
    // event counters back to each context
    if (ctr) {
        std::vector<unsigned long long> h((size_t)n * ctr_words);
        HC(hipMemcpyAsync(h.data(), ctr, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        idle(W, s);
        for (int k = 0; k < n; k++) {
            Timing &t = sc[k].c->timing;
            memcpy(&t.ev[0][0], h.data() + (size_t)k * ctr_words, sizeof(t.ev));
            memcpy(t.ev_sun, h.data() + (size_t)k * ctr_words + 5 * (1 + kMaxStages), sizeof(t.ev_sun));
        }
    }
    for (Scene &S : sc) {
        if (S.result < 0) continue;
        S.c->sun_ray_count = S.sun_count;
        S.c->n_hits_done = S.hits;
        S.result = (int)S.c->total_records();
    }
    W.timing.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return true;
}

}  // namespace st

using namespace st;

STCORE_API int st_amd_sim_run_batch(st_context_t *pcxts, int n, const unsigned int *seeds, int *results) {
    if (!pcxts || !seeds || n < 1) return -1;
    std::vector<Scene> sc(n);
    Context *c0 = ctx_of(pcxts[0]);
    if (!c0) return -1;
    for (int k = 0; k < n; k++) {
        Context *c = ctx_of(pcxts[k]);
        if (!c) return -1;
        for (int j = 0; j < k; j++)
            if (sc[j].c == c) return -1;  // a context twice
        sc[k].c = c;
        sc[k].seed = seeds[k];
        if (c->device != c0->device || c->precision != c0->precision || c->flux.active || c->load_n2 > 0 ||
            c->save_data) {
            c->errlog("batched run: contexts must share the device and precision, without flux maps or stage data");
            return -1;
        }
    }
    try {
        for (Scene &S : sc) {
            Context *c = S.c;
            c->clear_results();
            c->timing = Timing();
            S.result = 0;
            if (!c->prepare()) {
                S.result = -1;
                continue;
            }
            S.nst = (int)c->model.stages.size();
            S.N = (uint64_t)(int64_t)c->raycount;
            S.maxrays = (uint64_t)(int64_t)c->raymax;
            S.presample = c->err_sun && (c->scene_view.sun.shape == 'd' || c->scene_view.sun.shape == 'D') &&
                          !c->scene_view.sun.point_source && c->scene_view.sun.icdf_n > 0;
            c->recs.assign(c->model.stages.size(), StageRecords());
            c->d_hits.ensure((size_t)S.N * sizeof(uint64_t));
            S.nxt = (DevRay *)c->d_rays_a.ensure((size_t)S.N * sizeof(DevRay));
            S.n_nxt = 0;
        }
        // Two lanes, each a run_batch over half the scenes on its own stream
        // from its own host thread: one lane's sun masks overlap the other
        // lane's traces, as a single run overlaps its next batch's mask with
        // its current trace.  Scenes go to the lane with less estimated work
        // (ray hits, times 8 for scenes needing the generic kernel).
        std::vector<int> live;
        for (int k = 0; k < n; k++)
            if (sc[k].result >= 0) live.push_back(k);
        const int lanes = live.size() >= 2 ? 2 : 1;
        hipStream_t streams[2] = {c0->stream, c0->mask_stream};
        for (int l = 0; l < lanes; l++)
            if (!c0->batch[l]) {
                c0->batch[l] = std::make_shared<BatchWork>();
                HC(hipDeviceGetAttribute(&c0->batch[l]->cus, hipDeviceAttributeMultiprocessorCount, c0->device));
            }
        std::vector<std::pair<double, int>> order;
        for (int k : live)
            order.emplace_back(-(double)sc[k].N * (sc[k].c->cs.stages[0].generic ? 8.0 : 1.0), k);
        std::sort(order.begin(), order.end());
        std::vector<Scene> part[2];
        std::vector<int> where[2];
        double load[2] = {0.0, 0.0};
        for (const auto &o : order) {
            const int l = (lanes == 2 && load[1] < load[0]) ? 1 : 0;
            load[l] -= o.first;
            part[l].push_back(sc[o.second]);
            where[l].push_back(o.second);
        }
        const bool count = c0->count_events;
        const bool fp32 = c0->precision == 32;
        std::string lane_err[2];
        auto lane = [&](int l) {
            try {
                HC(hipSetDevice(c0->device));
                if (fp32) run_batch<float>(part[l], *c0->batch[l], count, streams[l]);
                else run_batch<double>(part[l], *c0->batch[l], count, streams[l]);
            } catch (const std::exception &e) {
                lane_err[l] = e.what();
            }
        };
        if (lanes == 1 && c0->batch[1]) c0->batch[1]->timing = BatchTiming();
        std::thread other;
        if (lanes == 2) other = std::thread(lane, 1);
        if (!part[0].empty()) lane(0);
        if (other.joinable()) other.join();
        for (int l = 0; l < lanes; l++) {
            if (!lane_err[l].empty()) throw std::runtime_error(lane_err[l]);
            for (size_t j = 0; j < part[l].size(); j++) sc[where[l][j]].result = part[l][j].result;
        }
        for (int k = 0; k < n; k++) {
            if (results) results[k] = sc[k].result;
            if (sc[k].result < 0) sc[k].c->clear_results();
        }
        for (const Scene &S : sc)
            if (S.result < 0) return -1;
        return 1;
    } catch (const std::exception &e) {
        c0->errlog("batched run: %s", e.what());
        return -1;
    }
}

STCORE_API int st_amd_batch_timing(st_context_t p, double *out, int nmax) {
    Context *c = ctx_of(p);
    if (!c || !c->batch[0]) return -1;
    // both lanes' kernel times and counts added; the wall time is the longer
    BatchTiming t = c->batch[0]->timing;
    if (c->batch[1]) {
        const BatchTiming &u = c->batch[1]->timing;
        t.mask_ms += u.mask_ms;
        t.sunshape_ms += u.sunshape_ms;
        t.total_ms = std::max(t.total_ms, u.total_ms);
        t.sun_rays += u.sun_rays;
        t.mask_launches += u.mask_launches;
        t.sunshape_launches += u.sunshape_launches;
        for (int i = 0; i < kMaxStages; i++) {
            t.trace_ms[i] += u.trace_ms[i];
            t.trace_launches[i] += u.trace_launches[i];
            t.rays_in[i] += u.rays_in[i];
            t.records[i] += u.records[i];
        }
    }
    std::vector<double> v = {t.mask_ms, t.sunshape_ms, t.total_ms, (double)t.sun_rays, (double)t.mask_launches,
                             (double)t.sunshape_launches};
    for (int i = 0; i < kMaxStages; i++) v.push_back(t.trace_ms[i]);
    for (int i = 0; i < kMaxStages; i++) v.push_back((double)t.trace_launches[i]);
    for (int i = 0; i < kMaxStages; i++) v.push_back((double)t.rays_in[i]);
    for (int i = 0; i < kMaxStages; i++) v.push_back((double)t.records[i]);
    const int m = std::min<int>(nmax, (int)v.size());
    for (int i = 0; i < m; i++) out[i] = v[i];
    return (int)v.size();
}

STCORE_API int st_amd_batch_tuning(int *out, int nmax) {
    std::lock_guard<std::mutex> lk(tune_mu);
    int m = 0;
    for (const auto &kv : tuned) {
        if (m + 3 <= nmax) {
            out[m] = kv.first;
            out[m + 1] = kv.second.tile;
            out[m + 2] = kv.second.waves;
        }
        m += 3;
    }
    return (int)tuned.size();
}
