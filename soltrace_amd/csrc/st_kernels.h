// Host-visible launch interface of the trace kernels (st_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "st_types.h"

namespace st {

// Device pointers of one compiled scene.
struct SceneView {
    const DevElem *elems;
    const DevFace *faces;
    const double *aux;
    const DevStage *stages;
    const DevBvhNode *bvh;
    const int32_t *bvh_items;
    const int32_t *always_items;  // a stage's elements no box bounds (DevStage::always0)
    const int32_t *grid_cells;   // ncell + 1 starts
    const int32_t *grid_items;
    const int32_t *grid_always;  // stage-1 elements every sun ray is tested against
    int32_t n_always;
    const DevFoot *foot;         // per stage-1 element
    const DevBeam *beam;         // per stage-1 element, or nullptr
    const DevBeamItem *beam_items;
    DevSunGrid grid;
    int32_t use_grid;
    int32_t nstages;
    DevSun sun;
};

// Sunshape polar angle of one stage-1 ray drawn ahead of the trace.
struct SunPre {
    double theta2;     // accepted sample scaled to rad^2
    uint32_t kend;     // stream index of the phi draw that follows
    uint32_t ok;
};

struct SunPreArgs {
    DevSun sun;
    const double *aux;
    uint32_t seed;
    uint32_t k0;                 // first draw of the rejection loop
    const uint64_t *sun_index;
    int32_t n;
    SunPre *out;
    uint32_t *err_flags;
    // event counters (nullptr: off): tries evaluated, tries the serial
    // loop would have made
    unsigned long long *counters;
};

// Arguments of one stage-trace launch.
struct TraceArgs {
    const SunPre *sun_pre;       // per ray, or nullptr
    SceneView sv;
    int32_t si;                  // 0-based stage
    uint32_t seed;
    int32_t inc_sun, inc_err;
    int32_t n_rays;              // rays in this launch
    int32_t slots;               // record slots per ray
    // stage 0 input
    const uint64_t *sun_index;   // per ray: 0-based sun-ray index
    uint32_t raynum_base;        // ray t has ray number raynum_base + t + 1
    // later-stage input
    const DevRay *in_rays;
    uint32_t last_raynum;        // LastRayNumberInPreviousStage
    // outputs, indexed by the ray's place in the launch's ray space,
    // ray_base + t (a batched launch lays several scenes' rays end to end)
    uint32_t ray_base;
    uint32_t slot_stride;        // rays in the launch's ray space
    DevRecord *rec_slots;        // slots * slot_stride, slot-major: record j of ray t at [j * slot_stride + ray_base + t]
    uint32_t *rec_count;         // per ray of the ray space
    DevRecord *ovf_rec;          // overflow records (seq >= slots)
    uint32_t *ovf_meta;          // (ray in the ray space, seq) pairs
    uint32_t *ovf_count;
    uint32_t ovf_cap;
    DevRay *out_slots;           // per ray of the ray space (nullptr: no later stage takes the rays)
    uint32_t *out_flag;          // per ray of the ray space (nullptr with out_slots)
    uint32_t *err_flags;         // bit 0: bad interaction, 1: phase mismatch, 2: overflow full
                                 // 3: sunshape sampling, 4: finite-element interpolation
    FluxSpec flux;
    uint32_t *flux_bins;
    // per-source maps (nullptr: off): stage 0 tags the rays it hands on with
    // the element that sent them; the flux stage also bins each ray into the
    // map of its source, flux_src[(src - 1) * flux_nbins + bin]
    uint32_t *flux_src;
    uint32_t flux_nbins;
    // event counters (nullptr: off): rays, element tests, BVH nodes,
    // uniforms drawn, records
    unsigned long long *counters;
    int32_t generic;             // kernel variant: DevStage::generic of stage si
    // lane -> ray, in the launch's ray space: the lane of place ray_base + i
    // traces the ray at place order[ray_base + i] (nullptr: ray i).  Every
    // output is indexed by the ray, so the order changes no result; stage-0
    // launches order their rays by sun-plane bucket, so a wave's rays leave
    // neighbouring mirrors in similar directions and walk the BVH alike.
    const uint32_t *order;
};

// Arguments of the sun-ray hit-mask launch.
struct MaskArgs {
    SceneView sv;
    uint32_t seed;
    uint64_t s0;                 // first sun-ray index
    uint32_t n;                  // sun rays in this launch
    uint64_t *mask;              // one word per 64 sun rays
    unsigned long long *fe_fail; // least sun-ray index whose finite-element interpolation failed
    unsigned long long *counters;  // as TraceArgs::counters (nullptr: off)
    int32_t generic;             // kernel variant: DevStage::generic of stage 0
    int32_t compact;             // screen grid candidates into per-wave pair queues
};

// Sun rays each lane of the grid sun-mask kernel draws (a wave covers
// 64 * kMaskRays consecutive sun rays).
constexpr int kMaskRays = 4;

// ----------------------------------------------------- batched scenes ----
// A batched launch runs one kernel over the work of several scenes.  The
// work is cut into tiles; items[k] holds scene k's launch arguments and
// prefix[k] its first tile (prefix[n] is the total).  The launch is
// persistent: each wave takes the next tile from a device counter until none
// is left, so scenes of uneven size share the whole chip.
struct BatchQueue {
    const uint32_t *prefix;      // n + 1 tile offsets
    int32_t n;                   // items
    uint32_t tile;               // trace: rays per tile (a multiple of 64)
    uint32_t *counter;           // tiles taken; zero before the launch
};

// Trace-kernel launch shape the batch autotuner chooses per scene class.
struct TraceShape {
    int32_t tile;                // rays per tile
    int32_t waves;               // waves per SIMD the kernel is built for
};

// Items of the batched sun-mask compaction: the hit sun indices of scene k's
// mask words, in order, up to `take` of them.
struct MaskCompactItem {
    uint64_t s0;                 // sun index of the scene's first mask bit
    uint64_t *dst;               // where the scene's next hit goes
    uint32_t take;
};

template <typename R> void launch_sun_mask(const MaskArgs &a, hipStream_t s);
// Batched launches (items in device memory, queue counter cleared by the
// caller).  grid: blocks of the persistent launch.
template <typename R> void launch_mask_batch(const MaskArgs *items, const BatchQueue &q, int generic, int grid,
                                             hipStream_t s);
template <typename R> void launch_sunshape_batch(const SunPreArgs *items, const BatchQueue &q, uint32_t total,
                                                 hipStream_t s);
template <typename R> void launch_trace_batch(const TraceArgs *items, const BatchQueue &q, int generic,
                                              const TraceShape &shape, int grid, hipStream_t s);
// The waves-per-SIMD builds launch_trace_batch has for a precision and
// kernel variant (generic), best first by the single-scene measurements.
int trace_batch_waves(int bits, int generic, int *out, int max);
template <typename R> void launch_sunshape_d(const SunPreArgs &a, hipStream_t s);
// a: the launch's arguments on the host (they pick the kernel and grid);
// dev_a: the same bytes in device memory, which the kernel reads (held by
// pointer, its many uniform fields need not stay in scalar registers).
template <typename R> void launch_trace_stage(const TraceArgs &a, const TraceArgs *dev_a, hipStream_t s);

// Tile scan of n uint32 values: out[i] is the exclusive prefix sum of in[i]
// within its tile of kScanTileElems values, tile_off[k] the exclusive prefix
// of the tile sums, *total (device) the sum.  The exclusive prefix of i is
// out[i] + tile_off[i / kScanTileElems]: consumers add the tile offset, which
// saves a pass over out.  tile_off must hold scan_tmp_words(n) words.
constexpr uint32_t kScanTileElems = 4096;
size_t scan_tmp_words(size_t n);
void launch_tile_scan(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tile_off, uint32_t *total,
                      hipStream_t s);

// Fill n words with v (a kernel, so it may sit inside a captured graph).
void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
// out[i] = exclusive prefix at position pos[i] of a tile scan over n values
// (the scan's total when pos[i] >= n).
void launch_prefix_at(const uint32_t *offsets, const uint32_t *tile_off, size_t n, const uint32_t *total,
                      const uint32_t *pos, int m, uint32_t *out, hipStream_t s);
// Batched mask compaction: word w of scene k (q.prefix in words) writes its
// hits' sun indices to items[k].dst, ranked within the scene.
void launch_mask_compact_batch(const uint64_t *mask, const uint32_t *offsets, const uint32_t *tile_off,
                               size_t nwords, const MaskCompactItem *items, const BatchQueue &q, hipStream_t s);
// Copy ray ranges: ray i of item k (q.prefix in rays) from src[k] to dst[k].
struct RayCopyItem {
    const DevRay *src;
    DevRay *dst;
};
void launch_copy_rays_batch(const RayCopyItem *items, const BatchQueue &q, uint32_t total, hipStream_t s);
// out[k] = ray number of the last of items[k]'s n rays (0 when n is 0).
struct LastRayItem {
    const DevRay *rays;
    uint64_t n;
};
void launch_last_raynum(const LastRayItem *items, int m, uint32_t *out, hipStream_t s);
// Mask words -> per-word popcounts.
void launch_mask_popc(const uint64_t *mask, uint32_t *counts, size_t nwords, hipStream_t s);
// Hit list: the first `limit` set bits of the mask, as sun-ray indices.
// (offsets, tile_off: a tile scan of the counts, as launch_tile_scan leaves it)
void launch_mask_compact(const uint64_t *mask, const uint32_t *offsets, const uint32_t *tile_off, size_t nwords,
                         uint64_t s0, uint32_t limit, uint64_t *hits, hipStream_t s);
// Record and ray compaction after a stage launch.
void launch_compact_records(const DevRecord *slots, const uint32_t *count, const uint32_t *offsets,
                            const uint32_t *tile_off, int n_rays, int nslots, DevRecord *out, hipStream_t s);
void launch_compact_overflow(const DevRecord *ovf, const uint32_t *meta, const uint32_t *n_ovf_dev, uint32_t cap,
                             const uint32_t *offsets, const uint32_t *tile_off, DevRecord *out, hipStream_t s);
// Flux statistics over resident records (st_stats.hip).  mm: min x, max x,
// min y, max y as order keys; acc: centroid sums (um), considered, binned,
// 1 + last considered index.
void launch_stats_bounds(const DevRecord *rec, uint32_t n, int32_t elem, const FluxSpec &fl, unsigned long long *mm,
                         hipStream_t s);
void launch_stats_bin(const DevRecord *rec, uint32_t n, int32_t elem, const FluxSpec &fl, uint64_t base,
                      uint32_t *bins, unsigned long long *acc, hipStream_t s);
void launch_compact_rays(const DevRay *slots, const uint32_t *flag, const uint32_t *offsets, const uint32_t *tile_off,
                         int n_rays, DevRay *out, hipStream_t s);
// Ray order of a stage-0 launch by sun-plane bucket, a counting sort with
// no global atomics (launch_sun_order, per precision file).  Blocks of
// kOrderRays rays rank their rays' buckets (a 2^kOrderLog2-square Morton
// grid over the sun box) with LDS atomics and store the block's bucket
// counts as a row of `table`; a column scan turns each column into the
// blocks' starts within the bucket and the bucket's total, a tile scan of
// the totals gives the buckets' starts, and order[bucket start + block
// start + rank] = ray.  The order within a bucket varies between runs; the
// results do not.
constexpr uint32_t kOrderLog2 = 7;
constexpr uint32_t kOrderBuckets = 1u << (2 * kOrderLog2);  // 16384 (a 64 KiB LDS histogram)
constexpr uint32_t kOrderRays = 8192;                        // rays per block (ranks fit 13 bits)
struct OrderWork {
    uint32_t *key_rank;  // n words: key | rank << 2 kOrderLog2
    uint32_t *table;     // order_table_words(n)
    uint32_t *colsum, *cbase, *ctile, *ctot;  // kOrderBuckets, kOrderBuckets, scan_tmp_words(kOrderBuckets), 1
};
inline size_t order_table_words(int n) { return (size_t)((n + kOrderRays - 1) / kOrderRays) * kOrderBuckets; }
template <typename R>
void launch_sun_order(const SceneView &sv, uint32_t seed, const uint64_t *sun_index, int n, const OrderWork &w,
                      uint32_t *order, hipStream_t s);
// the generic passes of launch_sun_order (st_compact.hip)
void launch_order_colscan(uint32_t *table, uint32_t nblocks, uint32_t *colsum, hipStream_t s);
void launch_order_place(const uint32_t *key_rank, int n, const uint32_t *table, const uint32_t *cbase,
                        const uint32_t *ctile, uint32_t *order, hipStream_t s);
// Bucket order over the segments of a batched launch: ray i of item k
// (q.prefix in rays) gets key items[k].key0 + its bucket on a
// 2^side_log2-square grid, counted into hist (nkeys words, zeroed here) with
// its rank from a global atomic; then launch_order_scatter.
struct BucketItem {
    DevSun sun;
    uint32_t seed;
    uint32_t key0;
    const uint64_t *sun_index;
};
constexpr uint32_t kBatchBucketLog2 = 6;  // 64 x 64 buckets per scene
template <typename R>
void launch_sun_buckets_batch(const BucketItem *items, const BatchQueue &q, uint32_t total, uint32_t *key_rank,
                              uint32_t *hist, uint32_t nkeys, hipStream_t s);
void launch_order_scatter(const uint32_t *key_rank, int n, const uint32_t *offsets, const uint32_t *tile_off,
                          uint32_t *order, hipStream_t s);
// The getters' columns of a run's records in merged order: records
// [dst, dst + n) of the output come from src, of stage `stage` (1-based);
// pieces are sorted by dst and cover [0, total).
struct MergePiece {
    const DevRecord *src;
    uint64_t dst;
    int32_t stage;
};
struct RecordColumns {
    double *x, *y, *z, *cx, *cy, *cz;
    int32_t *element, *stage, *raynum;
};
void launch_gather_records(const MergePiece *pieces, int npieces, uint64_t n, const RecordColumns &out, int cus,
                           hipStream_t s);

}  // namespace st
