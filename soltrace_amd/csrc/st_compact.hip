// Scans and compaction kernels shared by both precisions.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "st_kernels.h"

namespace st {

// ------------------------------------------------------------- scans ------
static constexpr int kScanItems = 16;
static constexpr int kScanBlock = 256;
static constexpr int kScanTile = kScanBlock * kScanItems;  // 4096
static_assert(kScanTile == (int)kScanTileElems, "tile size");

// Block-wide exclusive scan of one value per thread (LDS, 64-wide waves).
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wsum[kScanBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) {
        if (w < wid) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

// A thread's kScanItems values are 64 contiguous bytes: whole threads move
// them as four 16-byte loads and stores (the arrays are 16-byte aligned),
// the tail thread element by element.
__device__ __forceinline__ void load_items(const uint32_t *in, size_t base, size_t n, uint32_t v[kScanItems]) {
    if (base + kScanItems <= n) {
        const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            const uint4 q = p[k];
            v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) v[i] = (base + i < n) ? in[base + i] : 0u;
    }
}
__device__ __forceinline__ void store_items(uint32_t *out, size_t base, size_t n, const uint32_t v[kScanItems]) {
    if (base + kScanItems <= n) {
        uint4 *p = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) p[k] = make_uint4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i)
            if (base + i < n) out[base + i] = v[i];
    }
}

__global__ __launch_bounds__(kScanBlock) void scan_tiles(const uint32_t *in, uint32_t *out, size_t n,
                                                         uint32_t *tile_sums) {
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    load_items(in, base, n, v);
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) s += v[i];
    uint32_t tot;
    uint32_t off = block_excl_scan(s, &tot);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const uint32_t x = v[i];
        v[i] = off;
        off += x;
    }
    store_items(out, base, n, v);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void scan_sums(uint32_t *sums, size_t ntiles, uint32_t *total) {
    // single block; ntiles <= kScanTile
    const size_t base = (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = (base + i < ntiles) ? sums[base + i] : 0u;
        s += v[i];
    }
    uint32_t tot;
    uint32_t off = block_excl_scan(s, &tot);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < ntiles) sums[base + i] = off;
        off += v[i];
    }
    if (threadIdx.x == 0) *total = tot;
}


__global__ void fill_u32_kernel(uint32_t *p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

size_t scan_tmp_words(size_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_tile_scan(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tile_off, uint32_t *total,
                      hipStream_t s) {
    const size_t ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles == 0) {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, total, (size_t)1, 0u);
        return;
    }
    hipLaunchKernelGGL(scan_tiles, dim3((unsigned)ntiles), dim3(kScanBlock), 0, s, in, out, n, tile_off);
    hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanBlock), 0, s, tile_off, ntiles, total);
}

// -------------------------------------------------------- compaction ------
__global__ void mask_popc_kernel(const uint64_t *mask, uint32_t *counts, size_t nwords) {
    size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nwords) counts[w] = (uint32_t)__popcll(mask[w]);
}
void launch_mask_popc(const uint64_t *mask, uint32_t *counts, size_t nwords, hipStream_t s) {
    hipLaunchKernelGGL(mask_popc_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, mask, counts,
                       nwords);
}

__global__ void mask_compact_kernel(const uint64_t *mask, const uint32_t *offsets, const uint32_t *tile_off,
                                    size_t nwords, uint64_t s0, uint32_t limit, uint64_t *hits) {
    size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    uint64_t m = mask[w];
    uint32_t o = offsets[w] + tile_off[w / kScanTile];
    while (m != 0 && o < limit) {
        int b = __ffsll((unsigned long long)m) - 1;
        hits[o++] = s0 + w * 64 + (uint64_t)b;
        m &= m - 1;
    }
}
void launch_mask_compact(const uint64_t *mask, const uint32_t *offsets, const uint32_t *tile_off, size_t nwords,
                         uint64_t s0, uint32_t limit, uint64_t *hits, hipStream_t s) {
    hipLaunchKernelGGL(mask_compact_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, mask, offsets,
                       tile_off, nwords, s0, limit, hits);
}

// Slots are slot-major (record j of ray t at slots[j * n_rays + t]), so the
// first records of a wave's rays, most of the data, are read contiguously.
// A wave whose 64 rays have one record each copies them as one block: its
// 64 records are contiguous in the slots and in the output, and the lanes
// move the block in 8-byte words, each load and store instruction covering
// 512 contiguous bytes (a record per lane would touch 56 lines per
// instruction).  Other waves copy per lane.
static_assert(sizeof(DevRecord) == 56, "record layout");
__global__ __launch_bounds__(256) void compact_records_kernel(const DevRecord *slots, const uint32_t *count,
                                                              const uint32_t *offsets, const uint32_t *tile_off,
                                                              int n_rays, int nslots, DevRecord *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const uint32_t c = t < n_rays ? count[t] : 0u;
    const uint32_t o = t < n_rays ? offsets[t] + tile_off[t / kScanTile] : 0u;
    if (__ballot(c == 1u) == ~0ull) {
        // every lane is in range with one record: out[o0 + l] = slots[t0 + l]
        const int t0 = t - lane;
        const uint32_t o0 = (uint32_t)__shfl((int)o, 0);
        const uint2 *src = reinterpret_cast<const uint2 *>(slots + t0);
        uint2 *dst = reinterpret_cast<uint2 *>(out + o0);
        constexpr int kWords = 64 * (int)sizeof(DevRecord) / 8;  // 448
#pragma unroll
        for (int k = 0; k < kWords / 64; ++k) dst[k * 64 + lane] = src[k * 64 + lane];
        return;
    }
    if (t >= n_rays) return;
    const uint32_t m = c < (uint32_t)nslots ? c : (uint32_t)nslots;
    for (uint32_t j = 0; j < m; ++j) out[o + j] = slots[(size_t)j * n_rays + t];
}
void launch_compact_records(const DevRecord *slots, const uint32_t *count, const uint32_t *offsets,
                            const uint32_t *tile_off, int n_rays, int nslots, DevRecord *out, hipStream_t s) {
    hipLaunchKernelGGL(compact_records_kernel, dim3((n_rays + 255) / 256), dim3(256), 0, s, slots, count, offsets,
                       tile_off, n_rays, nslots, out);
}

// Overflow records are rare: a small grid walks however many there are.
__global__ void compact_overflow_kernel(const DevRecord *ovf, const uint32_t *meta, const uint32_t *n_ovf_dev,
                                        uint32_t cap, const uint32_t *offsets, const uint32_t *tile_off,
                                        DevRecord *out) {
    uint32_t n = *n_ovf_dev;
    if (n > cap) n = cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t t = meta[2 * i], seq = meta[2 * i + 1];
        out[offsets[t] + tile_off[t / kScanTile] + seq] = ovf[i];
    }
}
void launch_compact_overflow(const DevRecord *ovf, const uint32_t *meta, const uint32_t *n_ovf_dev, uint32_t cap,
                             const uint32_t *offsets, const uint32_t *tile_off, DevRecord *out, hipStream_t s) {
    const unsigned blocks = std::min<unsigned>(256u, (cap + 255) / 256);
    hipLaunchKernelGGL(compact_overflow_kernel, dim3(std::max(1u, blocks)), dim3(256), 0, s, ovf, meta, n_ovf_dev,
                       cap, offsets, tile_off, out);
}

// A wave's passed rays land contiguously in the output (their offsets are
// consecutive), so the wave copies them as one block of 16-byte words: word
// q of the block is word q mod 4 of the (q / 4)-th passed ray, whose lane
// the wave keeps in LDS by rank.
static_assert(sizeof(DevRay) == 64, "ray layout");
__global__ __launch_bounds__(256) void compact_rays_kernel(const DevRay *slots, const uint32_t *flag,
                                                           const uint32_t *offsets, const uint32_t *tile_off,
                                                           int n_rays, DevRay *out) {
    __shared__ int src_of[256 / 64][64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool f = t < n_rays && flag[t] != 0u;
    const uint64_t m = __ballot(f);
    if (m == 0) return;
    const int first = __ffsll((unsigned long long)m) - 1;
    const uint32_t o = f ? offsets[t] + tile_off[t / kScanTile] : 0u;
    const uint32_t o0 = (uint32_t)__shfl((int)o, first);
    if (f) src_of[wv][__popcll(m & ((1ull << lane) - 1ull))] = t;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int words = __popcll(m) * 4;
    const uint4 *src = reinterpret_cast<const uint4 *>(slots);
    uint4 *dst = reinterpret_cast<uint4 *>(out + o0);
    for (int q = lane; q < words; q += 64) dst[q] = src[(size_t)src_of[wv][q >> 2] * 4 + (q & 3)];
}
void launch_compact_rays(const DevRay *slots, const uint32_t *flag, const uint32_t *offsets, const uint32_t *tile_off,
                         int n_rays, DevRay *out, hipStream_t s) {
    hipLaunchKernelGGL(compact_rays_kernel, dim3((n_rays + 255) / 256), dim3(256), 0, s, slots, flag, offsets,
                       tile_off, n_rays, out);
}

// ---------------------------------------------------- batched scenes ----
__device__ __forceinline__ int batch_item_of(const BatchQueue &q, uint32_t x) {
    int lo = 0, hi = q.n - 1;  // the last item whose first tile is <= x
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (q.prefix[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void prefix_at_kernel(const uint32_t *offsets, const uint32_t *tile_off, size_t n, const uint32_t *total,
                                 const uint32_t *pos, int m, uint32_t *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const size_t p = pos[i];
    out[i] = p < n ? offsets[p] + tile_off[p / kScanTile] : *total;
}
void launch_prefix_at(const uint32_t *offsets, const uint32_t *tile_off, size_t n, const uint32_t *total,
                      const uint32_t *pos, int m, uint32_t *out, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(prefix_at_kernel, dim3((m + 255) / 256), dim3(256), 0, s, offsets, tile_off, n, total, pos, m,
                       out);
}

__global__ void mask_compact_batch_kernel(const uint64_t *mask, const uint32_t *offsets, const uint32_t *tile_off,
                                          size_t nwords, const MaskCompactItem *items, BatchQueue q) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    const int k = batch_item_of(q, (uint32_t)w);
    const uint32_t w0 = q.prefix[k];
    const MaskCompactItem it = items[k];
    uint64_t m = mask[w];
    // rank of the word's first hit within its scene
    uint32_t o = (offsets[w] + tile_off[w / kScanTile]) - (offsets[w0] + tile_off[w0 / kScanTile]);
    while (m != 0 && o < it.take) {
        const int b = __ffsll((unsigned long long)m) - 1;
        it.dst[o++] = it.s0 + (uint64_t)(w - w0) * 64 + (uint64_t)b;
        m &= m - 1;
    }
}
void launch_mask_compact_batch(const uint64_t *mask, const uint32_t *offsets, const uint32_t *tile_off,
                               size_t nwords, const MaskCompactItem *items, const BatchQueue &q, hipStream_t s) {
    if (nwords == 0) return;
    hipLaunchKernelGGL(mask_compact_batch_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, mask,
                       offsets, tile_off, nwords, items, q);
}

__global__ void copy_rays_batch_kernel(const RayCopyItem *items, BatchQueue q, uint32_t total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int k = batch_item_of(q, i);
    const RayCopyItem it = items[k];
    const uint32_t j = i - q.prefix[k];
    it.dst[j] = it.src[j];
}
void launch_copy_rays_batch(const RayCopyItem *items, const BatchQueue &q, uint32_t total, hipStream_t s) {
    if (total == 0) return;
    hipLaunchKernelGGL(copy_rays_batch_kernel, dim3((total + 255) / 256), dim3(256), 0, s, items, q, total);
}

__global__ void last_raynum_kernel(const LastRayItem *items, int m, uint32_t *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const LastRayItem it = items[k];
    out[k] = it.n ? it.rays[it.n - 1].raynum : 0u;
}
void launch_last_raynum(const LastRayItem *items, int m, uint32_t *out, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(last_raynum_kernel, dim3((m + 255) / 256), dim3(256), 0, s, items, m, out);
}



// ----------------------------------------------------- ray order ---------
__global__ void order_scatter_kernel(const uint32_t *key_rank, int n, const uint32_t *offsets,
                                     const uint32_t *tile_off, uint32_t *order) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = key_rank[2 * i], rank = key_rank[2 * i + 1];
    order[offsets[key] + tile_off[key / kScanTile] + rank] = (uint32_t)i;
}
void launch_order_scatter(const uint32_t *key_rank, int n, const uint32_t *offsets, const uint32_t *tile_off,
                          uint32_t *order, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(order_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, s, key_rank, n, offsets, tile_off,
                       order);
}

// Ray order (launch_sun_order): column b of table (nblocks rows of
// kOrderBuckets counts) -> each row's exclusive prefix within the column,
// and colsum[b] = the column's total.  Four lanes share a column, each
// scanning a quarter of the rows; a shuffle joins the quarters.
__global__ __launch_bounds__(256) void order_colscan_kernel(uint32_t *table, uint32_t nblocks, uint32_t *colsum) {
    const uint32_t b = blockIdx.x * 64 + (threadIdx.x >> 2), q = threadIdx.x & 3;
    const uint32_t seg = (nblocks + 3) / 4, r0 = min(nblocks, q * seg), r1 = min(nblocks, r0 + seg);
    uint32_t sum = 0;
    if (b < kOrderBuckets)
        for (uint32_t r = r0; r < r1; r++) sum += table[(size_t)r * kOrderBuckets + b];
    uint32_t incl = sum;
    uint32_t v = (uint32_t)__shfl_up((int)incl, 1, 4);
    if (q >= 1) incl += v;
    v = (uint32_t)__shfl_up((int)incl, 2, 4);
    if (q >= 2) incl += v;
    const uint32_t total = (uint32_t)__shfl((int)incl, 3, 4);
    uint32_t run = incl - sum;
    if (b < kOrderBuckets) {
        for (uint32_t r = r0; r < r1; r++) {
            uint32_t *p = table + (size_t)r * kOrderBuckets + b;
            const uint32_t c = *p;
            *p = run;
            run += c;
        }
        if (q == 3) colsum[b] = total;
    }
}
void launch_order_colscan(uint32_t *table, uint32_t nblocks, uint32_t *colsum, hipStream_t s) {
    hipLaunchKernelGGL(order_colscan_kernel, dim3(kOrderBuckets / 64), dim3(256), 0, s, table, nblocks, colsum);
}

// Ray order, last pass: ray i of block i / kOrderRays goes to its bucket's
// start, plus its block's start within the bucket, plus its rank.  A block
// of the pass covers one block of the count pass and reads that block's
// row of starts into LDS first.
__global__ __launch_bounds__(1024) void order_place_kernel(const uint32_t *key_rank, int n, const uint32_t *table,
                                                           const uint32_t *cbase, const uint32_t *ctile,
                                                           uint32_t *order) {
    __shared__ uint32_t start[kOrderBuckets];
    const uint32_t *row = table + (size_t)blockIdx.x * kOrderBuckets;
    for (uint32_t b = threadIdx.x; b < kOrderBuckets; b += 1024)
        start[b] = row[b] + cbase[b] + ctile[b / kScanTile];
    __syncthreads();
    const int base = (int)blockIdx.x * (int)kOrderRays;
    for (int k = 0; k < (int)(kOrderRays / 1024); k++) {
        const int i = base + k * 1024 + (int)threadIdx.x;
        if (i < n) {
            const uint32_t kr = key_rank[i];
            const uint32_t key = kr & (kOrderBuckets - 1u), rank = kr >> (2 * kOrderLog2);
            order[start[key] + rank] = (uint32_t)i;
        }
    }
}
void launch_order_place(const uint32_t *key_rank, int n, const uint32_t *table, const uint32_t *cbase,
                        const uint32_t *ctile, uint32_t *order, hipStream_t s) {
    if (n <= 0) return;
    const uint32_t nblocks = (uint32_t)((n + kOrderRays - 1) / kOrderRays);
    hipLaunchKernelGGL(order_place_kernel, dim3(nblocks), dim3(1024), 0, s, key_rank, n, table, cbase, ctile, order);
}

// ------------------------------------------------ merged record columns ---
// The getters' view of a run's records (coretrace/stapi.cpp:564-659 over
// TRayData::Merge, types.cpp:311-355): output record i comes from the piece
// holding i, found by a binary search over the pieces' first output
// positions; the record is split into the columns the getters return.
__global__ __launch_bounds__(256) void gather_records_kernel(const MergePiece *pieces, int npieces, uint64_t n,
                                                             RecordColumns out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int lo = 0, hi = npieces - 1;  // the last piece starting at or below i
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pieces[mid].dst <= i) lo = mid;
            else hi = mid - 1;
        }
        const MergePiece p = pieces[lo];
        const DevRecord r = p.src[i - p.dst];
        out.x[i] = r.pos[0];
        out.y[i] = r.pos[1];
        out.z[i] = r.pos[2];
        out.cx[i] = r.cos[0];
        out.cy[i] = r.cos[1];
        out.cz[i] = r.cos[2];
        out.element[i] = r.element;
        out.stage[i] = p.stage;
        out.raynum[i] = (int32_t)r.raynum;
    }
}

void launch_gather_records(const MergePiece *pieces, int npieces, uint64_t n, const RecordColumns &out,
                           int cus, hipStream_t s) {
    if (n == 0 || npieces == 0) return;
    const uint64_t want = (n + 255) / 256;
    const int blocks = (int)std::min<uint64_t>(want, (uint64_t)cus * 16);
    hipLaunchKernelGGL(gather_records_kernel, dim3(blocks), dim3(256), 0, s, pieces, npieces, n, out);
}

}  // namespace st
