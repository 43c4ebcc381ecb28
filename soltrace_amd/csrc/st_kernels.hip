// Trace kernels for CDNA4 (gfx950).
//
// Stage 0 runs in two launches per batch of sun rays:
//   1. sun_mask   -- one lane per sun-ray index s: draw the two sun-plane
//                    uniforms, find whether any stage-1 element accepts the
//                    ray, ballot the answer into one 64-bit word per wave.
//   2. trace_stage -- one lane per hitting sun ray, in sun-index order (ray
//                    numbers are ranks among hitting sun rays, as the
//                    reference numbers them, coretrace/raytrace.cpp:780-797).
// Later stages run trace_stage over the rays the previous stage passed on,
// in ray-number order.  Records go to per-lane slots and are compacted by an
// exclusive scan, so the record stream is deterministic and in the
// reference's per-stage order (ray number, then hit order).
#include <hip/hip_runtime.h>

#include "st_device.h"
#include "st_kernels.h"

namespace st {

static constexpr int kBlock = 256;

// ------------------------------------------------------ sun-ray source ----
// coretrace/generateray.cpp:58-161 (the sun-plane draws are full 53-bit)
template <typename R>
STD void gen_sun_ray(Rng<R> &rng, const DevSun &sun, const DevStage &st, double pg[3], double cg[3],
                     double psun[2]) {
    if (sun.point_source) {
        pg[0] = sun.origin[0]; pg[1] = sun.origin[1]; pg[2] = sun.origin[2];
        for (int i = 0; i < 3; i++) {
            double sgn = (rng.uniform_d() <= 0.5) ? -1.0 : 1.0;
            cg[i] = sgn * rng.uniform_d();
        }
        double m = ::sqrt(cg[0] * cg[0] + cg[1] * cg[1] + cg[2] * cg[2]);
        cg[0] = cg[0] / m; cg[1] = cg[1] / m; cg[2] = cg[2] / m;
        psun[0] = psun[1] = 0.0;
        return;
    }
    double x = sun.minx + (sun.maxx - sun.minx) * rng.uniform_d();
    double y = sun.miny + (sun.maxy - sun.miny) * rng.uniform_d();
    psun[0] = x; psun[1] = y;
    const double ps[3] = {x, y, -10000.0}, cs[3] = {0.0, 0.0, 1.0};
    double pst[3], cst[3];
    to_ref(ps, cs, sun.pos_sun_stage, sun.R, pst, cst);
    to_ref(pst, cst, st.org, st.R, pg, cg);
}

// ------------------------------------------------------ closest hit -------
template <typename R>
struct Best {
    double path;
    int j;               // element index within the stage
    ElemHit<R> h;
    double ce[3];        // element-frame incoming direction
};

// Test element j; keep it when it is the closest accepted hit so far
// (coretrace/raytrace.cpp:740-768; ties go to the lower element index, the
// order a brute-force scan meets them).
template <typename R>
STD void consider(const SceneView &sv, const DevStage &st, int j, const double ps[3], const double cs[3],
                  Best<R> &b) {
    const DevElem &e = sv.elems[st.elem0 + j];
    if (!e.enabled) return;
    double pe[3], ce[3];
    to_local(ps, cs, e.org, e.R, pe, ce);
    pe[0] = pe[0] + 1.0e-5 * ce[0];
    pe[1] = pe[1] + 1.0e-5 * ce[1];
    pe[2] = pe[2] + 1.0e-5 * ce[2];
    ElemHit<R> h;
    if (!element_hit<R>(e, sv.aux, pe, ce, h)) return;
    if (h.path < b.path || (h.path == b.path && j < b.j)) {
        if (!e.zap_filter || (double)h.X[2] <= e.zap) {
            b.path = h.path;
            b.j = j;
            b.h = h;
            b.ce[0] = ce[0]; b.ce[1] = ce[1]; b.ce[2] = ce[2];
        }
    }
}

// Candidates: the sun-plane grid cell for a first hit of a sun ray, the
// stage BVH when present, every element otherwise.
template <typename R>
STD void closest_hit(const SceneView &sv, const DevStage &st, const double ps[3], const double cs[3],
                     const double *psun, Best<R> &b) {
    b.path = 1e99;
    b.j = 0x7fffffff;
    if (psun != nullptr) {
        const DevSunGrid &g = sv.grid;
        int cx = (int)((psun[0] - g.x0) * g.inv_dx);
        int cy = (int)((psun[1] - g.y0) * g.inv_dy);
        cx = cx < 0 ? 0 : (cx >= g.nx ? g.nx - 1 : cx);
        cy = cy < 0 ? 0 : (cy >= g.ny ? g.ny - 1 : cy);
        int c = cy * g.nx + cx;
        int a = sv.grid_cells[c], z = sv.grid_cells[c + 1];
        for (int k = a; k < z; ++k) consider<R>(sv, st, sv.grid_items[k], ps, cs, b);
        return;
    }
    if (st.nbvh > 0) {
        // stack traversal; boxes are padded outward so a float slab test
        // never drops a box the exact ray touches
        const DevBvhNode *nodes = sv.bvh + st.bvh0;
        float o[3] = {(float)ps[0], (float)ps[1], (float)ps[2]};
        float inv[3];
        for (int i = 0; i < 3; i++) {
            float d = (float)cs[i];
            inv[i] = (d == 0.0f) ? 3.0e38f : 1.0f / d;
        }
        int stack[48];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const DevBvhNode nd = nodes[stack[--sp]];
            float t0 = 0.0f, t1 = 3.0e38f;
            for (int i = 0; i < 3; i++) {
                float ta = (nd.lo[i] - o[i]) * inv[i];
                float tb = (nd.hi[i] - o[i]) * inv[i];
                float lo = fminf(ta, tb), hi = fmaxf(ta, tb);
                t0 = fmaxf(t0, lo);
                t1 = fminf(t1, hi);
            }
            if (t0 > t1) continue;
            if ((double)t0 > b.path * 1.000001 + 1.0e-3) continue;
            if (nd.left < 0) {
                int first = -nd.left - 1;
                for (int k = 0; k < nd.right; ++k) consider<R>(sv, st, sv.bvh_items[first + k], ps, cs, b);
            } else if (sp < 46) {
                stack[sp++] = nd.right;
                stack[sp++] = nd.left;
            }
        }
        return;
    }
    for (int j = 0; j < st.nelem; ++j) consider<R>(sv, st, j, ps, cs, b);
}

// Any accepted hit at all (the first-hit test of a sun ray).
template <typename R>
STD bool any_hit(const SceneView &sv, const DevStage &st, const double ps[3], const double cs[3],
                 const double *psun) {
    Best<R> b;
    closest_hit<R>(sv, st, ps, cs, psun, b);
    return b.path < 1e99;
}

// ---------------------------------------------------------- sun mask ------
template <typename R>
__global__ __launch_bounds__(kBlock) void sun_mask_kernel(MaskArgs a) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    bool hit = false;
    if (t < a.n) {
        const DevStage st = a.sv.stages[0];
        Rng<R> rng;
        rng.begin(a.seed, 0, a.s0 + t);
        double pg[3], cg[3], psun[2], ps[3], cs[3];
        gen_sun_ray<R>(rng, a.sv.sun, st, pg, cg, psun);
        to_local(pg, cg, st.org, st.R, ps, cs);
        hit = any_hit<R>(a.sv, st, ps, cs, (a.sv.use_grid && !a.sv.sun.point_source) ? psun : nullptr);
    }
    const uint64_t word = __ballot(hit);
    if ((threadIdx.x & 63) == 0) {
        const uint32_t w = t >> 6;
        if ((uint64_t)w * 64 < a.n) a.mask[w] = word;
    }
}

// ------------------------------------------------------- stage trace ------
STD void grid_bin(double x, double lo, double b, int n, int *g) {
    // GridIncrement of app/src/project.cpp:1398-1405: the first k >= 0 with
    // lo + k*b >= x, minus one.
    double est = (x - lo) / b;
    int k;
    if (!(est > 0.0)) k = 0;
    else if (est > (double)(n + 1)) k = n + 2;
    else k = (int)::ceil(est);
    while (k > 0 && !((lo + (k - 1) * b) < x)) k--;
    while (k <= n + 1 && (lo + k * b) < x) k++;
    *g = k - 1;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void trace_stage_kernel(TraceArgs a) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= a.n_rays) return;
    const SceneView &sv = a.sv;
    const DevStage st = sv.stages[a.si];
    const bool sun_stage = (a.si == 0);

    double pg[3], cg[3], psun[2];
    uint32_t raynum;
    uint64_t sun_index;
    Rng<R> rng;
    if (sun_stage) {
        sun_index = a.sun_index[t];
        raynum = a.raynum_base + (uint32_t)t + 1u;
        rng.begin(a.seed, 0, sun_index);
        gen_sun_ray<R>(rng, sv.sun, st, pg, cg, psun);
    } else {
        const DevRay r = a.in_rays[t];
        raynum = r.raynum;
        sun_index = r.sun_index;
        for (int i = 0; i < 3; i++) { pg[i] = r.pos[i]; cg[i] = r.cos[i]; }
        rng.begin(a.seed, (uint32_t)a.si, raynum);
    }
    double ps[3], cs[3];
    to_local(pg, cg, st.org, st.R, ps, cs);

    R cos_out_el[3] = {R(0), R(0), R(0)};
    uint32_t nrec = 0;
    int mhc = 0;
    bool passed = false;
    DevRecord pend;
    bool has_pend = false;
    // flux: last record of this ray on the flux element
    const bool flux_here = a.flux.active && a.flux.stage == a.si;
    double fpos[3] = {0, 0, 0}, fcos[3] = {0, 0, 0};
    bool fhave = false, fabs_ = false;

    auto flush = [&]() {
        if (!has_pend) return;
        if (nrec < (uint32_t)a.slots) {
            a.rec_slots[(size_t)t * a.slots + nrec] = pend;
        } else {
            uint32_t o = atomicAdd(a.ovf_count, 1u);
            if (o < a.ovf_cap) {
                a.ovf_rec[o] = pend;
                a.ovf_meta[2 * o] = (uint32_t)t;
                a.ovf_meta[2 * o + 1] = nrec;
            } else {
                atomicOr(a.err_flags, 4u);
            }
        }
        nrec++;
        has_pend = false;
    };

    for (;;) {
        Best<R> b;
        const bool first_sun = sun_stage && mhc == 0 && sv.use_grid && !sv.sun.point_source;
        closest_hit<R>(sv, st, ps, cs, first_sun ? psun : nullptr, b);
        if (!(b.path < 1e99)) {
            if (sun_stage) {
                if (mhc == 0) atomicOr(a.err_flags, 2u);
                else passed = true;
                break;
            }
            if (st.tracethrough || mhc > 0) {
                passed = true;
                if (raynum == a.last_raynum || mhc > 0) break;
            }
            flush();
            for (int i = 0; i < 3; i++) { pend.pos[i] = ps[i]; pend.cos[i] = cs[i]; }
            pend.element = 0;
            pend.raynum = raynum;
            has_pend = true;
            break;
        }
        const DevElem &e = sv.elems[st.elem0 + b.j];
        double xd[3] = {(double)b.h.X[0], (double)b.h.X[1], (double)b.h.X[2]};
        flush();
        to_ref(xd, b.ce, e.org, e.R, pend.pos, pend.cos);
        pend.element = b.j + 1;
        pend.raynum = raynum;
        has_pend = true;
        mhc++;
        if (flux_here && b.j + 1 == a.flux.elem) {
            for (int i = 0; i < 3; i++) { fpos[i] = pend.pos[i]; fcos[i] = pend.cos[i]; }
            fhave = true;
            fabs_ = false;
        }

        R pos_out[3] = {b.h.X[0], b.h.X[1], b.h.X[2]};
        R cos_out[3];
        const R cin0[3] = {(R)b.ce[0], (R)b.ce[1], (R)b.ce[2]};
        if (st.virt) {
            for (int i = 0; i < 3; i++) { cos_out[i] = cin0[i]; cos_out_el[i] = cin0[i]; }
        } else {
            if (e.interaction != 1 && e.interaction != 2) {
                atomicOr(a.err_flags, 1u);
                break;
            }
            const DevFace &f = sv.faces[b.h.back ? e.face_back : e.face_front];
            R tv = test_value<R>(e.interaction, f, sv.aux, cin0, b.h.DF);
            if (tv <= rng.uniform()) {
                pend.element = -pend.element;
                if (fhave && b.j + 1 == a.flux.elem) fabs_ = true;
                break;
            }
            R cin[3] = {cin0[0], cin0[1], cin0[2]};
            R df[3] = {b.h.DF[0], b.h.DF[1], b.h.DF[2]};
            if (a.inc_sun && sun_stage && mhc == 1) {
                R co[3];
                errors<R>(rng, cin, 1, sv.sun, sv.aux, e.interaction, f, co, df);
                cin[0] = co[0]; cin[1] = co[1]; cin[2] = co[2];
            }
            if (a.inc_err) {
                R co[3];
                normal_errors<R>(rng, df, f, co);
                df[0] = co[0]; df[1] = co[1]; df[2] = co[2];
            }
            interact<R>(rng, cin, df, e.interaction, f, cos_out_el);
            if (a.inc_err) {
                R base[3];
                const bool diffuse = (f.dist == 'f' || f.dist == 'F');
                for (int i = 0; i < 3; i++) base[i] = diffuse ? df[i] : cos_out_el[i];
                R co[3];
                errors<R>(rng, base, 2, sv.sun, sv.aux, e.interaction, f, co, df);
                cos_out_el[0] = co[0]; cos_out_el[1] = co[1]; cos_out_el[2] = co[2];
            }
            for (int i = 0; i < 3; i++) cos_out[i] = cos_out_el[i];
        }
        double po[3] = {(double)pos_out[0], (double)pos_out[1], (double)pos_out[2]};
        double co[3] = {(double)cos_out[0], (double)cos_out[1], (double)cos_out[2]};
        to_ref(po, co, e.org, e.R, ps, cs);
        to_ref(ps, cs, st.org, st.R, pg, cg);
        if (!st.multihit) {
            passed = true;
            break;
        }
    }
    flush();
    a.rec_count[t] = nrec;
    a.out_flag[t] = passed ? 1u : 0u;
    if (passed) {
        DevRay r;
        for (int i = 0; i < 3; i++) { r.pos[i] = pg[i]; r.cos[i] = cg[i]; }
        r.raynum = raynum;
        r.pad = 0;
        r.sun_index = sun_index;
        a.out_slots[t] = r;
    }
    if (flux_here && fhave && (!a.flux.absorbed_only || fabs_)) {
        // app/src/project.cpp:1358-1415: stage -> element frame, unroll the
        // cylinder, linear-scan bin search, integer counts
        double pe[3], ce[3];
        to_local(fpos, fcos, a.flux.org, a.flux.R, pe, ce);
        double x = pe[0], y = pe[1], z = pe[2];
        const double Rr = a.flux.radius;
        if (a.flux.kind == 1) {
            if (z <= Rr) x = Rr * ::asin(x / Rr);
            else if (z > Rr) {
                if (x < 0) x = -(M_PI * Rr / 2.0 + Rr * ::acos(::fabs(x) / Rr));
                if (x >= 0) x = M_PI * Rr / 2.0 + Rr * ::acos(x / Rr);
            }
        } else if (a.flux.kind == 2) {
            x = Rr * ::asin(x / Rr);
        }
        int gx, gy;
        grid_bin(x, a.flux.xmin, a.flux.bx, a.flux.nx, &gx);
        grid_bin(y, a.flux.ymin, a.flux.by, a.flux.ny, &gy);
        if (gx >= 0 && gx < a.flux.nx && gy >= 0 && gy < a.flux.ny)
            atomicAdd(&a.flux_bins[gx * a.flux.ny + gy], 1u);
    }
}

template <typename R>
void launch_sun_mask(const MaskArgs &a, hipStream_t s) {
    const uint32_t blocks = (a.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(sun_mask_kernel<R>, dim3(blocks), dim3(kBlock), 0, s, a);
}
template <typename R>
void launch_trace_stage(const TraceArgs &a, hipStream_t s) {
    const int blocks = (a.n_rays + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(trace_stage_kernel<R>, dim3(blocks), dim3(kBlock), 0, s, a);
}
template void launch_sun_mask<double>(const MaskArgs &, hipStream_t);
template void launch_sun_mask<float>(const MaskArgs &, hipStream_t);
template void launch_trace_stage<double>(const TraceArgs &, hipStream_t);
template void launch_trace_stage<float>(const TraceArgs &, hipStream_t);

// ------------------------------------------------------------- scans ------
static constexpr int kScanItems = 16;
static constexpr int kScanBlock = 256;
static constexpr int kScanTile = kScanBlock * kScanItems;  // 4096

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

__global__ __launch_bounds__(kScanBlock) void scan_tiles(const uint32_t *in, uint32_t *out, size_t n,
                                                         uint32_t *tile_sums) {
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        s += v[i];
    }
    uint32_t tot;
    uint32_t off = block_excl_scan(s, &tot);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[base + i] = off;
        off += v[i];
    }
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

__global__ __launch_bounds__(kScanBlock) void scan_add(uint32_t *out, size_t n, const uint32_t *tile_off) {
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    const uint32_t add = tile_off[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i)
        if (base + i < n) out[base + i] += add;
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

void launch_exclusive_scan(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp, uint32_t *total,
                           hipStream_t s) {
    const size_t ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles == 0) {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, total, (size_t)1, 0u);
        return;
    }
    hipLaunchKernelGGL(scan_tiles, dim3((unsigned)ntiles), dim3(kScanBlock), 0, s, in, out, n, tmp);
    hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanBlock), 0, s, tmp, ntiles, total);
    hipLaunchKernelGGL(scan_add, dim3((unsigned)ntiles), dim3(kScanBlock), 0, s, out, n, tmp);
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

__global__ void mask_compact_kernel(const uint64_t *mask, const uint32_t *offsets, size_t nwords, uint64_t s0,
                                    uint32_t limit, uint64_t *hits) {
    size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    uint64_t m = mask[w];
    uint32_t o = offsets[w];
    while (m != 0 && o < limit) {
        int b = __ffsll((unsigned long long)m) - 1;
        hits[o++] = s0 + w * 64 + (uint64_t)b;
        m &= m - 1;
    }
}
void launch_mask_compact(const uint64_t *mask, const uint32_t *offsets, size_t nwords, uint64_t s0, uint32_t limit,
                         uint64_t *hits, hipStream_t s) {
    hipLaunchKernelGGL(mask_compact_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, mask, offsets,
                       nwords, s0, limit, hits);
}

__global__ void compact_records_kernel(const DevRecord *slots, const uint32_t *count, const uint32_t *offsets,
                                       int n_rays, int nslots, DevRecord *out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_rays) return;
    uint32_t c = count[t];
    uint32_t m = c < (uint32_t)nslots ? c : (uint32_t)nslots;
    uint32_t o = offsets[t];
    for (uint32_t j = 0; j < m; ++j) out[o + j] = slots[(size_t)t * nslots + j];
}
void launch_compact_records(const DevRecord *slots, const uint32_t *count, const uint32_t *offsets, int n_rays,
                            int nslots, DevRecord *out, hipStream_t s) {
    hipLaunchKernelGGL(compact_records_kernel, dim3((n_rays + 255) / 256), dim3(256), 0, s, slots, count, offsets,
                       n_rays, nslots, out);
}

__global__ void compact_overflow_kernel(const DevRecord *ovf, const uint32_t *meta, const uint32_t *n_ovf_dev,
                                        uint32_t cap, const uint32_t *offsets, DevRecord *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = *n_ovf_dev;
    if (n > cap) n = cap;
    if (i >= n) return;
    uint32_t t = meta[2 * i], seq = meta[2 * i + 1];
    out[offsets[t] + seq] = ovf[i];
}
void launch_compact_overflow(const DevRecord *ovf, const uint32_t *meta, const uint32_t *n_ovf_dev, uint32_t cap,
                             const uint32_t *offsets, DevRecord *out, hipStream_t s) {
    hipLaunchKernelGGL(compact_overflow_kernel, dim3((cap + 255) / 256), dim3(256), 0, s, ovf, meta, n_ovf_dev, cap,
                       offsets, out);
}

__global__ void compact_rays_kernel(const DevRay *slots, const uint32_t *flag, const uint32_t *offsets, int n_rays,
                                    DevRay *out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_rays) return;
    if (flag[t]) out[offsets[t]] = slots[t];
}
void launch_compact_rays(const DevRay *slots, const uint32_t *flag, const uint32_t *offsets, int n_rays, DevRay *out,
                         hipStream_t s) {
    hipLaunchKernelGGL(compact_rays_kernel, dim3((n_rays + 255) / 256), dim3(256), 0, s, slots, flag, offsets, n_rays,
                       out);
}

}  // namespace st
