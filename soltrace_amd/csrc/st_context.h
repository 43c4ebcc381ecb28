// Host side of the HIP backend shared by the C API (st_context.cpp) and the
// batched multi-scene runs (st_batch.cpp): device buffers, the per-context
// scene, run state and results.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "st_kernels.h"
#include "st_scene.h"

#define STCORE_API extern "C" __attribute__((visibility("default")))

typedef unsigned long st_uint_t;
typedef void *st_context_t;
typedef int (*st_callback_t)(st_uint_t, st_uint_t, st_uint_t, st_uint_t, st_uint_t, void *);

namespace st {

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        char b[256];
        snprintf(b, sizeof b, "HIP error in %s: %s", what, hipGetErrorString(e));
        throw std::runtime_error(b);
    }
}
#define HC(x) hip_check((x), #x)

// Device buffer that grows on demand and keeps its capacity.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void *ensure(size_t bytes) {
        if (bytes <= cap && p) return p;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes, 256);
        HC(hipMalloc(&p, want));
        cap = want;
        return p;
    }
    template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

template <typename T>
inline T *upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    b.ensure(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) HC(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return b.as<T>();
}

// The device address of the allocation-free scratch array on GPU `dev`
// (st_context.cpp).
void *scratch_on(int dev);

// The trace, sun-mask and compaction streams shared by the contexts on GPU
// `dev` (st_context.cpp).
struct DevStreams {
    hipStream_t trace, mask, comp;
};
const DevStreams &streams_on(int dev);

// A stage's records of one run: chunks of device records, each a range of a
// buffer.  Buffers a context allocated return to its pool when the results
// are cleared; a batched run's buffers are shared by its contexts and are
// not pooled.
struct StageRecords {
    std::vector<std::shared_ptr<DevBuf>> chunks;
    std::vector<size_t> offsets;  // first record of each chunk in its buffer
    std::vector<size_t> counts;
    std::vector<char> pooled;
    size_t total = 0;
    void add(const std::shared_ptr<DevBuf> &b, size_t off, size_t n, bool pool) {
        chunks.push_back(b);
        offsets.push_back(off);
        counts.push_back(n);
        pooled.push_back(pool ? 1 : 0);
        total += n;
    }
    DevRecord *ptr(size_t k) const { return chunks[k]->as<DevRecord>() + offsets[k]; }
};

struct Timing {
    double mask_ms = 0, trace_ms[kMaxStages] = {0}, other_ms = 0, total_ms = 0, sunpre_ms = 0;
    uint64_t sun_rays = 0, rays_in[kMaxStages] = {0}, records[kMaxStages] = {0};
    uint64_t mask_launches = 0, trace_launches[kMaxStages] = {0};
    // event counters per kernel: [0] sun mask, [1 + stage] trace
    unsigned long long ev[1 + kMaxStages][5] = {{0}};
    unsigned long long ev_sun[2] = {0, 0};  // sampler: tries evaluated, tries needed
};

// Kernel times and counts of the last batched run (st_amd_batch_timing).
struct BatchTiming {
    double mask_ms = 0, sunshape_ms = 0, total_ms = 0, trace_ms[kMaxStages] = {0};
    uint64_t sun_rays = 0, mask_launches = 0, sunshape_launches = 0, trace_launches[kMaxStages] = {0};
    uint64_t rays_in[kMaxStages] = {0}, records[kMaxStages] = {0};
};

struct BatchWork;  // workspace of batched runs (st_batch.cpp)

struct Context {
    SceneModel model;
    int raycount = 1000, raymax = 100000;
    bool dyn_group = false, err_sun = true, err_opt = true;
    int precision = 64;
    bool cull = true;
    bool count_events = false;
    DevBuf d_counters;
    std::vector<std::string> messages;
    // the device's shared streams (streams_on): trace; the next batch's sun
    // mask, beside the current batch's trace; record and ray compaction of a
    // trace launch, beside the next launch
    hipStream_t stream = nullptr;
    hipStream_t mask_stream = nullptr;
    hipStream_t comp_stream = nullptr;
    hipEvent_t ev_hits[2] = {nullptr, nullptr}, ev_ready = nullptr, ev_join = nullptr;
    // per launch-buffer set: its totals reached the host; its compaction is done
    hipEvent_t ev_tot[2] = {nullptr, nullptr}, ev_comp[2] = {nullptr, nullptr};
    // per sun-ray batch parity: the batch's trace launches are done (its
    // sunshape angles and ray order may be overwritten)
    hipEvent_t ev_traced[2] = {nullptr, nullptr};
    uint32_t *h_tot = nullptr;  // pinned: per-launch totals and error flags (4 words per set)
    int device = 0;

    CompiledScene cs;
    bool dirty = true;            // a setter ran since the last compile
    bool compiled_cull = true;
    SceneView scene_view{};
    DevBuf d_elems, d_faces, d_aux, d_stages, d_bvh, d_bvh_items, d_always_items, d_grid_cells, d_grid_items, d_grid_always, d_foot, d_beam, d_beam_items;
    // run scratch
    DevBuf d_mask_scan_tmp, d_mask_total;
    DevBuf d_mask, d_popc, d_offs, d_total, d_hits, d_ovf_count, d_flux;
    // per-launch buffers, two sets: launch k uses set k mod 2, so a launch
    // runs while the one before it is compacted
    DevBuf d_scan_tmp[2], d_scan_tmp2[2], d_rec_slots[2], d_rec_count[2], d_rec_offs[2], d_ovf_rec[2], d_ovf_meta[2],
        d_out_slots[2], d_out_flag[2], d_out_offs[2];
    DevBuf d_rays_a, d_rays_b, d_sunpre, d_sunpre_batch[2];
    // stage-0 ray order by sun-plane bucket (OrderWork: key_rank, table,
    // column totals and starts, tile starts, total; the order of each parity)
    DevBuf d_keyrank, d_hist, d_hoffs, d_htmp, d_htot, d_order[2];
    // trace-launch arguments in device memory: a ring of slots, each filled
    // from its pinned host twin by a copy queued before its launch (at most
    // two launches are queued ahead of the host, so a slot is free again
    // long before the ring comes round)
    static constexpr int kArgsRing = 8;
    TraceArgs *h_args = nullptr;
    DevBuf d_args;
    int args_next = 0;

    // results
    std::vector<StageRecords> recs;
    // record buffers of earlier runs, reused so a run allocates no device
    // memory once the buffers have grown to the scene's size
    std::vector<std::shared_ptr<DevBuf>> rec_pool;
    std::vector<hipEvent_t> ev_pool;  // kernel timing events, reused
    // the last run's kernel intervals (a time field, the first of two pool
    // events), turned into times when the timing is asked for
    std::vector<std::pair<double *, size_t>> marks;
    void settle_timing() {
        for (const auto &m : marks) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ev_pool[m.second], ev_pool[m.second + 1]) == hipSuccess) *m.first += ms;
        }
        marks.clear();
    }
    uint64_t sun_ray_count = 0;
    // the records in merged order as the getters' columns, in HBM
    bool cols_valid = false;
    DevBuf d_cols, d_pieces;
    RecordColumns cols{};
    int cus = 0;
    char *h_stage = nullptr;  // pinned staging for the getters' copies (2 x 32 MiB)
    hipEvent_t ev_stage[2] = {nullptr, nullptr};
    uint64_t n_hits_done = 0;

    FluxSpec flux{};
    std::vector<uint32_t> flux_host;
    // per-source maps: one flux map per stage-1 element (st_amd_flux_by_source)
    bool flux_by_source = false;
    DevBuf d_flux_src;
    size_t flux_src_words = 0;  // words of d_flux_src the last run filled (0: none)
    Timing timing;

    // stage-data save/resume (coretrace/stapi.cpp:680-728): rows to load in
    // the next run, and the rows the last run saved
    const double *load_s1 = nullptr, *load_s2 = nullptr;
    uint64_t load_n1 = 0, load_n2 = 0;
    bool save_data = false;
    std::vector<double> saved_s1, saved_s2;  // 8 and 7 doubles a row

    void *scratch_ptr = nullptr;  // this context's GPU's copy of `scratch`
    std::shared_ptr<BatchWork> batch[2];  // the two lanes of a batched run, kept by its first context

    Context() {
        HC(hipGetDevice(&device));
        scratch_ptr = scratch_on(device);
        const DevStreams &ds = streams_on(device);
        stream = ds.trace;
        mask_stream = ds.mask;
        comp_stream = ds.comp;
        for (auto &e : ev_hits) HC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (auto &e : ev_tot) HC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (auto &e : ev_comp) HC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (auto &e : ev_traced) HC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HC(hipEventCreateWithFlags(&ev_ready, hipEventDisableTiming));
        HC(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        HC(hipHostMalloc((void **)&h_tot, 16 * sizeof(uint32_t), hipHostMallocDefault));
    }
    ~Context() {
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        for (auto e : ev_hits)
            if (e) (void)hipEventDestroy(e);
        if (ev_ready) (void)hipEventDestroy(ev_ready);
        if (ev_join) (void)hipEventDestroy(ev_join);
        for (auto e : ev_tot)
            if (e) (void)hipEventDestroy(e);
        for (auto e : ev_comp)
            if (e) (void)hipEventDestroy(e);
        for (auto e : ev_traced)
            if (e) (void)hipEventDestroy(e);
        if (h_tot) (void)hipHostFree(h_tot);
        if (h_stage) (void)hipHostFree(h_stage);
        if (h_args) (void)hipHostFree(h_args);
        for (hipEvent_t e : ev_stage)
            if (e) (void)hipEventDestroy(e);
    }
    void errlog(const char *fmt, ...) {
        char b[513];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(b, 512, fmt, ap);
        va_end(ap);
        messages.push_back(b);
    }
    void clear_results() {
        for (auto &st : recs)
            for (size_t k = 0; k < st.chunks.size(); k++)
                if (st.pooled[k]) rec_pool.push_back(st.chunks[k]);
        recs.clear();
        cols_valid = false;
        flux_src_words = 0;
        n_hits_done = 0;
        sun_ray_count = 0;
    }
    bool prepare();
    int run(uint32_t seed, st_callback_t cb, void *cbdata);
    template <typename R> int run_t(uint32_t seed, st_callback_t cb, void *cbdata);
    void materialize();
    // copy `bytes` of the merged columns at dev into the caller's host memory
    void copy_out(void *host, const void *dev, size_t bytes);
    size_t total_records() const {
        size_t n = 0;
        for (const auto &s : recs) n += s.total;
        return n;
    }
};

inline Context *ctx_of(st_context_t p) { return reinterpret_cast<Context *>(p); }

constexpr uint32_t kMaxChunk = 1u << 22;   // rays per trace launch
constexpr int kSlots = 4;                  // record slots per ray per launch

}  // namespace st
