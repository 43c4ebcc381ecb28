// Trace kernels for CDNA4 (gfx950): templates, instantiated for FP64 in
// st_trace_f64.hip and for FP32 in st_trace_f32.hip.
//
// Stage 0 runs in two launches per batch of sun rays:
//   1. sun_mask   -- per sun-ray index s: draw the two sun-plane uniforms,
//                    find whether any stage-1 element accepts the ray, and
//                    set bit s of the hit mask (one 64-bit word per 64 rays).
//                    Over the sun-plane grid, each lane draws several rays
//                    and screens their candidates, and the element tests run
//                    on per-wave queues of the pairs that pass.
//   2. trace_stage -- one lane per hitting sun ray, in sun-index order (ray
//                    numbers are ranks among hitting sun rays, as the
//                    reference numbers them, coretrace/raytrace.cpp:780-797).
// Later stages run trace_stage over the rays the previous stage passed on,
// in ray-number order.  Records go to per-lane slots and are compacted by an
// exclusive scan, so the record stream is deterministic and in the
// reference's per-stage order (ray number, then hit order).
#pragma once
#include <hip/hip_runtime.h>

#include "st_device.h"
#include "st_kernels.h"

namespace st {

static constexpr int kBlock = 256;

// Event counters (a diagnostic mode): the sum of v over the wave's active
// lanes, added by one lane, so a launch makes one atomic per wave and
// counter rather than one per lane on the same address.
STD void wave_add(unsigned long long *p, unsigned long long v) {
    const uint64_t act = __ballot(true);
    unsigned long long sum = 0;
    for (uint64_t m = act; m != 0; m &= m - 1) {
        const int l = __ffsll((unsigned long long)m) - 1;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
        sum += ((unsigned long long)hi << 32) | lo;
    }
    if ((int)(threadIdx.x & 63) == __ffsll((unsigned long long)act) - 1) atomicAdd(p, sum);
}
static constexpr int32_t kNoRecord = INT32_MIN;  // a record lost to a full overflow buffer

// ------------------------------------------------------ sun-ray source ----
// coretrace/generateray.cpp:58-161 (the sun-plane draws are full 53-bit):
// the point on the sun plane, then the ray it starts in the global frame.
// x = MinX + (MaxX - MinX) U (coretrace/generateray.cpp:143-144) with the
// uniform's 2^-32 folded into the width: the product is the same exact
// number, so it rounds the same.
STD void sun_plane_point(Rng &rng, const DevSun &sun, double psun[2]) {
    const double wx = (sun.maxx - sun.minx) * 2.3283064365386963e-10;
    const double wy = (sun.maxy - sun.miny) * 2.3283064365386963e-10;
    psun[0] = sun.minx + (double)rng.next_word() * wx;
    psun[1] = sun.miny + (double)rng.next_word() * wy;
}

STD void sun_ray_from_plane(const DevSun &sun, const DevStage &st, const double psun[2], double pg[3],
                            double cg[3]) {
    const double ps[3] = {psun[0], psun[1], -10000.0}, cs[3] = {0.0, 0.0, 1.0};
    double pst[3], cst[3];
    to_ref(ps, cs, sun.pos_sun_stage, sun.R, pst, cst);
    to_ref(pst, cst, st.org, st.R, pg, cg);
}

STD void gen_sun_ray(Rng &rng, const DevSun &sun, const DevStage &st, double pg[3], double cg[3],
                     double psun[2]) {
    if (sun.point_source) {
        pg[0] = sun.origin[0]; pg[1] = sun.origin[1]; pg[2] = sun.origin[2];
        for (int i = 0; i < 3; i++) {
            double sgn = (rng.uniform() <= 0.5) ? -1.0 : 1.0;
            cg[i] = sgn * rng.uniform();
        }
        double m = ::sqrt(cg[0] * cg[0] + cg[1] * cg[1] + cg[2] * cg[2]);
        cg[0] = cg[0] / m; cg[1] = cg[1] / m; cg[2] = cg[2] / m;
        psun[0] = psun[1] = 0.0;
        return;
    }
    sun_plane_point(rng, sun, psun);
    sun_ray_from_plane(sun, st, psun, pg, cg);
}

// The sun-plane grid cell holding (sx, sy), and its candidate range.
STD int grid_cell(const SceneView &sv, double sx, double sy) {
    const DevSunGrid &g = sv.grid;
    int cx = (int)((sx - g.x0) * g.inv_dx);
    int cy = (int)((sy - g.y0) * g.inv_dy);
    cx = cx < 0 ? 0 : (cx >= g.nx ? g.nx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= g.ny ? g.ny - 1 : cy);
    return cy * g.nx + cx;
}
STD void grid_cell_range(const SceneView &sv, double sx, double sy, int *k, int *kend) {
    const int c = grid_cell(sv, sx, sy);
    *k = sv.grid_cells[c];
    *kend = sv.grid_cells[c + 1];
}

// ------------------------------------------------------ closest hit -------
// Only the winner's path and index are kept during the search; its hit
// point and normal are recomputed once afterwards, which keeps the search
// loop's live state small.
struct Best {
    double path;
    int j;               // element index within the stage
    uint32_t tests;      // element tests (event counter)
    uint32_t nodes;      // BVH nodes visited (event counter)
    uint32_t fefail;     // a finite-element interpolation failed on the way
};

// The ray in element j's frame, nudged 1e-5 m along itself
// (coretrace/raytrace.cpp:725-731).
STD void element_ray(const DevElem &e, const double ps[3], const double cs[3], double pe[3], double ce[3]) {
    to_local(ps, cs, e.org, e.R, pe, ce);
    pe[0] = pe[0] + 1.0e-5 * ce[0];
    pe[1] = pe[1] + 1.0e-5 * ce[1];
    pe[2] = pe[2] + 1.0e-5 * ce[2];
}

// Test element j; keep it when it is the closest accepted hit so far
// (coretrace/raytrace.cpp:740-768; ties go to the lower element index, the
// order a brute-force scan meets them).
template <typename R, int G>
STD void consider(const SceneView &sv, const DevStage &st, int j, const double ps[3], const double cs[3], Best &b) {
    const DevElem &e = sv.elems[st.elem0 + j];
    if (!e.enabled) return;
    b.tests++;
    double pe[3], ce[3];
    element_ray(e, ps, cs, pe, ce);
    ElemHit<R> h;
    h.back = 0;
    if (!element_hit<R, G>(e, sv.aux, pe, ce, h)) {
        if (G == 2 && h.back == 2) b.fefail = 1;
        return;
    }
    if (h.path < b.path || (h.path == b.path && j < b.j)) {
        if (!e.zap_filter || (double)h.X[2] <= e.zap) {
            b.path = h.path;
            b.j = j;
        }
    }
}

// Whether sun-plane point (sx, sy) lies inside element j's footprints.  The
// first record is read whole before any test, so its loads issue together.
STD bool foot_contains(const DevFoot *foot, int j, double sx, double sy) {
    const DevFoot f = foot[2 * j];
    const float px = (float)(sx - f.cx), py = (float)(sy - f.cy);
    bool in = true;
#pragma unroll
    for (int i = 0; i < 6; i++) in = in && (f.e[i][0] * px + f.e[i][1] * py <= f.e[i][2]);
    in = f.kind == 0 ? in : f.kind == 1;
    if (in || !f.more) return in;
    const DevFoot g = foot[2 * j + 1];
    const float qx = (float)(sx - g.cx), qy = (float)(sy - g.cy);
    in = true;
#pragma unroll
    for (int i = 0; i < 6; i++) in = in && (g.e[i][0] * qx + g.e[i][1] * qy <= g.e[i][2]);
    return g.kind == 0 ? in : g.kind == 1;
}

// Whether a ray leaving stage-1 element `prev` at ps along cs is covered by
// that element's beam list (DevBeam).
STD bool in_beam(const SceneView &sv, int prev, const double ps[3], const double cs[3]) {
    if (prev < 0 || sv.beam == nullptr) return false;
    const DevBeam &bm = sv.beam[prev];
    if (bm.count < 0) return false;
    const float dx = (float)(ps[0] - bm.c[0]), dy = (float)(ps[1] - bm.c[1]), dz = (float)(ps[2] - bm.c[2]);
    const float dot = (float)cs[0] * bm.a[0] + (float)cs[1] * bm.a[1] + (float)cs[2] * bm.a[2];
    return dx * dx + dy * dy + dz * dz <= bm.r2 && dot >= bm.cosa;
}

// Candidates: sun-plane grid cell `cell` for a first hit of a sun ray (cell
// >= 0), screened by the elements' footprints at (sx, sy) when `screen`; the
// beam list of stage-1 element `prev`, the element the ray leaves, when the
// ray is within its beam; the stage BVH when present; every element
// otherwise.  One loop feeds them all to a single consider() so the element
// code is instantiated once.
template <typename R, int G>
STD void closest_hit(const SceneView &sv, const DevStage &st, const double ps[3], const double cs[3], int cell,
                     bool screen, double sx, double sy, Best &b, int prev = -1) {
    b.path = 1e99;
    b.j = 0x7fffffff;
    b.fefail = 0;
    const int32_t *items = nullptr;  // nullptr: the item is the element index
    const DevBeamItem *boxes = nullptr;  // beam list entries (a box test each), else nullptr
    int k = 0, kend = 0;
    bool bvh = false, always = false, beam = false;
    if (cell >= 0) {
        // the cell's candidates, then the elements no footprint bounds
        items = sv.grid_items;
        k = sv.grid_cells[cell];
        kend = sv.grid_cells[cell + 1];
        always = sv.n_always > 0;
    } else if (st.nbvh > 0) {
        // the stage's unbounded elements (torus, finite element), then the
        // beam list or the walk
        beam = in_beam(sv, prev, ps, cs);
        bvh = !beam;
        items = sv.always_items;
        k = st.always0;
        kend = st.always0 + st.nalways;
    } else {
        kend = st.nelem;
    }
    // stackless depth-first BVH walk: descend on a box hit, else jump to the
    // node's skip index (no per-lane stack, so no scratch memory); boxes are
    // padded outward so a float slab test never drops a box the ray touches
    const DevBvhNode *nodes = sv.bvh + st.bvh0;
    float o[3], inv[3];
    for (int q = 0; q < 3; q++) {
        o[q] = (float)ps[q];
        const float d = (float)cs[q];
        inv[q] = (d == 0.0f) ? 3.0e38f : 1.0f / d;
    }
    int i = 0;
    for (;;) {
        if (k < kend) {
            int j;
            if (boxes != nullptr) {
                const DevBeamItem it = boxes[k];
                ++k;
                b.nodes++;
                float t0 = 0.0f, t1 = 3.0e38f;
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const float ta = (it.lo[q] - o[q]) * inv[q];
                    const float tb = (it.hi[q] - o[q]) * inv[q];
                    t0 = fmaxf(t0, fminf(ta, tb));
                    t1 = fminf(t1, fmaxf(ta, tb));
                }
                if (!(t0 <= t1) || (double)t0 > b.path * 1.000001 + 1.0e-3) continue;
                j = it.j;
            } else {
                j = items != nullptr ? items[k] : k;
                ++k;
                if (screen && !foot_contains(sv.foot, j, sx, sy)) continue;
            }
            consider<R, G>(sv, st, j, ps, cs, b);
            continue;
        }
        if (beam) {
            beam = false;
            const DevBeam &bm = sv.beam[prev];
            boxes = sv.beam_items;
            k = bm.first;
            kend = bm.first + bm.count;
            continue;
        }
        if (always) {
            always = false;
            items = sv.grid_always;
            k = 0;
            kend = sv.n_always;
            continue;
        }
        if (!bvh) break;
        items = sv.bvh_items;
        bool leaf = false;
        while (i < st.nbvh) {
            const DevBvhNode nd = nodes[i];
            b.nodes++;
            float t0 = 0.0f, t1 = 3.0e38f;
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const float ta = (nd.lo[q] - o[q]) * inv[q];
                const float tb = (nd.hi[q] - o[q]) * inv[q];
                t0 = fmaxf(t0, fminf(ta, tb));
                t1 = fminf(t1, fmaxf(ta, tb));
            }
            const bool enter = (t0 <= t1) && !((double)t0 > b.path * 1.000001 + 1.0e-3);
            if (!enter) {
                i = nd.skip;
            } else if (nd.leaf) {
                k = nd.leaf >> 3;
                kend = k + (nd.leaf & 7);
                i = nd.skip;
                leaf = true;
                break;
            } else {
                i = i + 1;
            }
        }
        if (!leaf) break;
    }
}

// Any accepted hit at all (the first-hit test of a sun ray).
template <typename R, int G>
STD bool any_hit(const SceneView &sv, const DevStage &st, const double ps[3], const double cs[3], bool grid,
                 double sx, double sy, uint32_t *tests, uint32_t *fefail) {
    Best b;
    b.tests = 0;
    b.nodes = 0;
    closest_hit<R, G>(sv, st, ps, cs, grid ? grid_cell(sv, sx, sy) : -1, grid, sx, sy, b);
    *tests += b.tests;
    *fefail = b.fefail;
    return b.path < 1e99;
}

// ---------------------------------------------------------- sun mask ------
// Sun ray t of a mask launch tested by one lane (no sun-plane grid); the
// wave's 64 rays share one mask word.
template <typename R, int G>
__device__ __forceinline__ void mask_lane(const MaskArgs &a, const uint32_t t) {
    bool hit = false;
    if (t < a.n) {
        const DevStage &st = a.sv.stages[0];
        Rng rng;
        rng.begin(a.seed, 0, a.s0 + t);
        double pg[3], cg[3], psun[2], ps[3], cs[3];
        gen_sun_ray(rng, a.sv.sun, st, pg, cg, psun);
        to_local(pg, cg, st.org, st.R, ps, cs);
        uint32_t tests = 0, fefail = 0;
        hit = any_hit<R, G>(a.sv, st, ps, cs, a.sv.use_grid && !a.sv.sun.point_source, psun[0], psun[1], &tests,
                            &fefail);
        if (G == 2 && fefail) atomicMin(a.fe_fail, (unsigned long long)(a.s0 + t));
        if (a.counters) {
            wave_add(&a.counters[0], 1ull);
            wave_add(&a.counters[1], (unsigned long long)tests);
            wave_add(&a.counters[3], (unsigned long long)rng.k);
        }
    }
    const uint64_t word = __ballot(hit);
    if ((threadIdx.x & 63) == 0) {
        const uint32_t w = t >> 6;
        if ((uint64_t)w * 64 < a.n) a.mask[w] = word;
    }
}

template <typename R, int G>
__global__ __launch_bounds__(kBlock) void sun_mask_kernel(MaskArgs a) {
    mask_lane<R, G>(a, blockIdx.x * kBlock + threadIdx.x);
}

// Sun mask over the sun-plane grid, compacted per wave.  About one sun ray in
// ten hits something, and those rays are spread over every wave, so a
// lane-per-ray candidate loop runs the full element test in nearly every
// wave for a few lanes.  Here each lane draws kMaskRays sun points; the
// wave's (ray, grid item) candidates, 0.6 per sun ray on the tower field but
// unevenly spread, are flattened and screened by their footprints 64 at a
// time; the pairs that pass go to a per-wave queue in LDS, and the element
// tests run on the queue, one pair per lane.  A ray hits when any of its
// pairs does, which is what any_hit() answers for it.
constexpr int kMaskQueue = 128;      // ring of pairs per wave (holds up to 127)
// a pair packs its ray slot (r * 64 + lane) in 8 bits above a 24-bit element index
static_assert(kMaskRays * 64 <= 256, "ray slot must fit in 8 bits");

STD uint64_t lanes_below() { return (1ull << (threadIdx.x & 63)) - 1ull; }

// Wave `wave` of a sun-mask launch: sun rays wave * 64 * kMaskRays onwards.
// Each wave works in its own share of the block's LDS, so waves need not
// move together.
template <typename R, int G>
__device__ __forceinline__ void mask_grid_wave(const MaskArgs &a, const uint64_t wave) {
    constexpr int W = kBlock / 64;
    __shared__ uint32_t ring[W][kMaskQueue];
    __shared__ double2 pts[W][kMaskRays * 64];
    __shared__ uint32_t hitw[W][kMaskRays * 2];
    __shared__ uint32_t cstart[W][kMaskRays * 64];  // candidate-list starts, lane-major (lane * kMaskRays + r)
    __shared__ int32_t kbase[W][kMaskRays * 64];    // first grid item of each ray slot (r * 64 + lane)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t ray0 = wave * (64 * kMaskRays);
    const SceneView &sv = a.sv;
    const DevStage &st = sv.stages[0];
    if (lane < kMaskRays * 2) hitw[wv][lane] = 0u;
    int kb[kMaskRays], ke[kMaskRays];
    uint32_t rays = 0, draws = 0;
#pragma unroll
    for (int r = 0; r < kMaskRays; r++) {
        const uint64_t i = ray0 + (uint64_t)(r * 64 + lane);
        double p[2] = {0.0, 0.0};
        kb[r] = ke[r] = 0;
        if (i < a.n) {
            Rng rng;
            rng.begin(a.seed, 0, a.s0 + i);
            sun_plane_point(rng, sv.sun, p);
            grid_cell_range(sv, p[0], p[1], &kb[r], &ke[r]);
            rays++;
            draws += rng.k;
        }
        pts[wv][r * 64 + lane] = make_double2(p[0], p[1]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t head = 0, count = 0, tests = 0;
    // test the m pairs at the ring's head, one per lane
    auto drain = [&](uint32_t m) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if ((uint32_t)lane < m) {
            const uint32_t e = ring[wv][(head + lane) & (kMaskQueue - 1)];
            const int slot = (int)(e >> 24), j = (int)(e & 0xFFFFFFu);
            const double2 q = pts[wv][slot];
            const double psun[2] = {q.x, q.y};
            double pg[3], cg[3], ps[3], cs[3];
            sun_ray_from_plane(sv.sun, st, psun, pg, cg);
            to_local(pg, cg, st.org, st.R, ps, cs);
            Best b;
            b.path = 1e99;
            b.j = 0x7fffffff;
            b.tests = 0;
            b.nodes = 0;
            b.fefail = 0;
            consider<R, G>(sv, st, j, ps, cs, b);
            if (b.path < 1e99) atomicOr(&hitw[wv][(slot >> 6) * 2 + ((slot >> 5) & 1)], 1u << (slot & 31));
            if (G == 2 && b.fefail) atomicMin(a.fe_fail, (unsigned long long)(a.s0 + ray0 + (uint64_t)slot));
        }
        head += m;
        count -= m;
        tests += m;
    };
    // The candidates of the wave's rays, flattened: lane l's rays own
    // positions cstart[l * kMaskRays + r] onwards; each pass hands the next
    // 64 positions to the 64 lanes, so no lane idles while another walks a
    // long cell list.
    uint32_t mine = 0;
#pragma unroll
    for (int r = 0; r < kMaskRays; r++) mine += (uint32_t)(ke[r] - kb[r]);
    uint32_t incl = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, d, 64);
        if (lane >= d) incl += y;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 63);
    {
        uint32_t c = incl - mine;
#pragma unroll
        for (int r = 0; r < kMaskRays; r++) {
            cstart[wv][lane * kMaskRays + r] = c;
            kbase[wv][r * 64 + lane] = kb[r];
            c += (uint32_t)(ke[r] - kb[r]);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t f = base + (uint32_t)lane;
        bool pass = false;
        int j = 0, slot = 0;
        if (f < total) {
            // the last list start at or below f owns position f
            int q = 0;
#pragma unroll
            for (int step = kMaskRays * 32; step > 0; step >>= 1)
                if (cstart[wv][q + step] <= f) q += step;
            slot = (q % kMaskRays) * 64 + q / kMaskRays;
            j = sv.grid_items[kbase[wv][slot] + (int)(f - cstart[wv][q])];
            const double2 p = pts[wv][slot];
            pass = foot_contains(sv.foot, j, p.x, p.y);
        }
        const uint64_t bits = __ballot(pass);
        if (pass) {
            const uint32_t pos = (head + count + (uint32_t)__popcll(bits & lanes_below())) & (kMaskQueue - 1);
            ring[wv][pos] = ((uint32_t)slot << 24) | (uint32_t)j;
        }
        count += (uint32_t)__popcll(bits);
        if (count >= 64) drain(64);
    }
    if (count > 0) drain(count);
    // The elements no footprint bounds (torus, finite element), after the
    // others: only for rays no element has accepted yet, as the mask asks
    // whether any element accepts a ray.  Finite-element stages test every
    // ray, for the least failing sun-ray index.
    if (sv.n_always > 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        bool open[kMaskRays];
#pragma unroll
        for (int r = 0; r < kMaskRays; r++) {
            const int slot = r * 64 + lane;
            const bool hit = (hitw[wv][r * 2 + (lane >> 5)] >> (lane & 31)) & 1u;
            open[r] = ray0 + (uint64_t)slot < a.n && (G == 2 || !hit);
        }
        for (int q = 0; q < sv.n_always; q++) {
            const int j = sv.grid_always[q];
#pragma unroll
            for (int r = 0; r < kMaskRays; r++) {
                const uint64_t bits = __ballot(open[r]);
                if (open[r]) {
                    const uint32_t pos = (head + count + (uint32_t)__popcll(bits & lanes_below())) & (kMaskQueue - 1);
                    ring[wv][pos] = ((uint32_t)(r * 64 + lane) << 24) | (uint32_t)j;
                }
                count += (uint32_t)__popcll(bits);
                if (count >= 64) drain(64);
            }
        }
        if (count > 0) drain(count);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < kMaskRays) {
        const uint64_t w = ray0 / 64 + (uint64_t)lane;
        if (w * 64 < a.n) a.mask[w] = (uint64_t)hitw[wv][lane * 2] | ((uint64_t)hitw[wv][lane * 2 + 1] << 32);
    }
    if (a.counters) {
        wave_add(&a.counters[0], (unsigned long long)rays);
        wave_add(&a.counters[3], (unsigned long long)draws);
        if (lane == 0) atomicAdd(&a.counters[1], (unsigned long long)tests);
    }
}

template <typename R, int G>
__global__ __launch_bounds__(kBlock) void sun_mask_grid_kernel(MaskArgs a) {
    mask_grid_wave<R, G>(a, (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
}

// ------------------------------------------- cooperative sunshape draw -----
// The 'd' sunshape accepts about one rejection iteration in a hundred for a
// Buie aureole, and a lane-per-ray loop runs a whole wave until its slowest
// lane is done.  Here the 64 lanes of a wave share the wave's unresolved
// rays: each round, lane l takes ray (l mod m) of the m unresolved ones and
// groups next_g + 2(l/m) and next_g + 2(l/m) + 1 of that ray's loop (eight
// tries: the round's shuffles and ballots are paid once per eight); each
// ray keeps the smallest accepted iteration.  The result is exactly what
// the serial loop of coretrace/errors.cpp:160-185 returns on the same
// stream.
// Position of set bit number n (from 0) of v, which has more than n set
// bits: six halving steps on population counts.
STD int nth_set_bit(uint64_t v, int n) {
    int pos = 0;
    int c = __popc((uint32_t)v);
    if (n >= c) { n -= c; v >>= 32; pos += 32; }
    c = __popc((uint32_t)v & 0xFFFFu);
    if (n >= c) { n -= c; v >>= 16; pos += 16; }
    c = __popc((uint32_t)v & 0xFFu);
    if (n >= c) { n -= c; v >>= 8; pos += 8; }
    c = __popc((uint32_t)v & 0xFu);
    if (n >= c) { n -= c; v >>= 4; pos += 4; }
    c = __popc((uint32_t)v & 0x3u);
    if (n >= c) { n -= c; v >>= 2; pos += 2; }
    if (n >= (int)(v & 1u)) pos += 1;
    return pos;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void sunshape_d_kernel(SunPreArgs a) {
    const int lane = threadIdx.x & 63;
    const int myray = blockIdx.x * kBlock + (threadIdx.x & ~63) + lane;
    const bool valid = myray < a.n;
    const uint64_t sid = valid ? a.sun_index[myray] : 0;
    const uint32_t sid_lo = (uint32_t)sid, sid_hi = (uint32_t)(sid >> 32);
    uint64_t U = __ballot(valid);
    uint32_t next_g = 0;  // next group of four tries of this lane's ray
    uint32_t ev_eval = 0, ev_need = 0;
    for (int round = 0; U != 0; ++round) {
        const int m = __popcll(U);
        const int rank = lane % m, rep = lane / m;
        const int rl = nth_set_bit(U, rank);
        const uint32_t g = (uint32_t)__shfl((int)next_g, rl) + 2u * (uint32_t)rep;
        const uint64_t s = ((uint64_t)(uint32_t)__shfl((int)sid_hi, rl) << 32) | (uint32_t)__shfl((int)sid_lo, rl);
        double t2 = 0.0;
        int off = sunshape_d_group(a.seed, s, a.k0, g, a.sun, a.aux, &t2);
        ev_eval += 4u;
        if (off < 0) {
            off = sunshape_d_group(a.seed, s, a.k0, g + 1u, a.sun, a.aux, &t2);
            if (off >= 0) off += 4;
            ev_eval += 4u;
        }
        const bool acc = off >= 0;
        const uint64_t accm = __ballot(acc);
        // lanes that served this lane's ray: rank r_own, r_own + m, ...
        const bool owner = (U >> lane) & 1ull;
        const int r_own = __popcll(U & ((1ull << lane) - 1ull));
        uint64_t pat = 1ull;  // bits 0, m, 2m, ... below 64
        for (int sh = m; sh < 64; sh <<= 1) pat |= pat << sh;
        const uint64_t mine = owner ? (pat << r_own) : 0ull;
        const uint64_t hit = accm & mine;
        const int first = hit ? (__ffsll((unsigned long long)hit) - 1) : 0;
        const double t2_first = __shfl(t2, first);
        const int off_first = __shfl(off, first);
        bool resolved = false;
        if (owner) {
            if (hit) {
                const uint32_t gg = next_g + 2u * (uint32_t)((first - r_own) / m);
                const uint32_t jj = 4u * gg + (uint32_t)off_first - sunshape_d_shift(a.k0);
                SunPre p;
                p.theta2 = t2_first / 1000000.0;
                p.kend = a.k0 + 3u * (jj + 1u);
                p.ok = 1u;
                a.out[myray] = p;
                resolved = true;
                ev_need += jj + 1u;
            } else {
                next_g += 2u * (uint32_t)__popcll(mine);
                if (next_g > (1u << 24)) {  // no acceptance in 2^26 tries: give up on this ray
                    SunPre p;
                    p.theta2 = 0.0;
                    p.kend = 0;
                    p.ok = 0u;
                    a.out[myray] = p;
                    atomicOr(a.err_flags, 8u);
                    resolved = true;
                }
            }
        }
        U &= ~__ballot(resolved);
    }
    if (a.counters) {
        wave_add(&a.counters[0], (unsigned long long)ev_eval);
        wave_add(&a.counters[1], (unsigned long long)ev_need);
    }
}

// The 'd' polar angle by inversion (sunshape_icdf): one lane per ray, two
// draws at k0 and k0+1 of the ray's stage-0 stream.  One instantiation per
// precision file, so each takes that file's contraction setting.
template <typename R>
__device__ __forceinline__ void sunshape_icdf_one(const SunPreArgs &a, const int i) {
    Rng rng;
    rng.begin(a.seed, 0, a.sun_index[i]);
    rng.seek(a.k0);
    const uint32_t w0 = rng.next_word(), w1 = rng.next_word();
    SunPre p;
    p.theta2 = sunshape_icdf(w0, w1, a.sun, a.aux) / 1000000.0;
    p.kend = rng.k;
    p.ok = 1u;
    a.out[i] = p;
    if (a.counters) {
        const uint64_t m = __ballot(true);
        if ((threadIdx.x & 63) == (uint32_t)(__ffsll((unsigned long long)m) - 1)) {
            atomicAdd(&a.counters[0], (unsigned long long)__popcll(m));
            atomicAdd(&a.counters[1], (unsigned long long)__popcll(m));
        }
    }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void sunshape_icdf_kernel(SunPreArgs a) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    sunshape_icdf_one<R>(a, i);
}

template <typename R>
void launch_sunshape_d(const SunPreArgs &a, hipStream_t s) {
    const int blocks = (a.n + kBlock - 1) / kBlock;
    if (a.sun.icdf_n > 0) hipLaunchKernelGGL(sunshape_icdf_kernel<R>, dim3(blocks), dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL(sunshape_d_kernel<R>, dim3(blocks), dim3(kBlock), 0, s, a);
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

// Flux-map bin of a record on the flux element, -1 outside the grid
// (app/src/project.cpp:1358-1415: stage -> element frame, cylinder
// unrolled, linear-scan bin search).
STD int flux_bin(const FluxSpec &fl, const double pos[3], const double cos[3]) {
    double pe[3], ce[3];
    to_local(pos, cos, fl.org, fl.R, pe, ce);
    double x = pe[0], y = pe[1], z = pe[2];
    const double Rr = fl.radius;
    if (fl.kind == 1) {
        if (z <= Rr) x = Rr * ::asin(x / Rr);
        else if (z > Rr) {
            if (x < 0) x = -(M_PI * Rr / 2.0 + Rr * ::acos(::fabs(x) / Rr));
            if (x >= 0) x = M_PI * Rr / 2.0 + Rr * ::acos(x / Rr);
        }
    } else if (fl.kind == 2) {
        x = Rr * ::asin(x / Rr);
    }
    int gx, gy;
    grid_bin(x, fl.xmin, fl.bx, fl.nx, &gx);
    grid_bin(y, fl.ymin, fl.by, fl.ny, &gy);
    return (gx >= 0 && gx < fl.nx && gy >= 0 && gy < fl.ny) ? gx * fl.ny + gy : -1;
}

// Ray t of a stage-trace launch: the closest-hit loop, records, optics and
// multi-hit of coretrace/raytrace.cpp:653-1040 for one ray.
template <typename R, int G>
__device__ __forceinline__ void trace_one(const TraceArgs &a, const int t) {
    const SceneView &sv = a.sv;
    const DevStage &st = sv.stages[a.si];  // read where used: a copy would hold ~33 scalar registers
    const bool sun_stage = (a.si == 0);
    constexpr bool FAST = sizeof(R) == 4;

    // Loop-carried state is kept small for occupancy: the sun-ray index is
    // read again at the end, the first hit's grid cell replaces the
    // sun-plane point, and event counts live in LDS.
    double ps[3], cs[3];  // ray in the stage frame
    uint32_t raynum;
    int cell = -1;        // sun-plane grid cell of a sun ray's first hit
    Rng rng;
    {
        double pg[3], cg[3];
        if (sun_stage) {
            const uint64_t sun_index = a.sun_index[t];
            raynum = a.raynum_base + (uint32_t)t + 1u;
            rng.begin(a.seed, 0, sun_index);
            double psun[2];
            gen_sun_ray(rng, sv.sun, st, pg, cg, psun);
            if (sv.use_grid && !sv.sun.point_source) cell = grid_cell(sv, psun[0], psun[1]);
        } else {
            const DevRay r = a.in_rays[t];
            raynum = r.raynum;
            for (int i = 0; i < 3; i++) { pg[i] = r.pos[i]; cg[i] = r.cos[i]; }
            rng.begin(a.seed, (uint32_t)a.si, r.sun_index);
        }
        to_local(pg, cg, st.org, st.R, ps, cs);
    }
    __shared__ uint32_t ev_lds[3][kBlock];  // element tests, BVH nodes, draws made by the sampler
    if (a.counters) ev_lds[0][threadIdx.x] = ev_lds[1][threadIdx.x] = ev_lds[2][threadIdx.x] = 0u;

    double cos_out_el[3] = {0.0, 0.0, 0.0};  // CosRayOutElement, kept across hits
    uint32_t nrec = 0;
    int mhc = 0;
    bool passed = false;
    // records go out as they are made; an absorbed hit later rewrites the
    // element field of its own record (slot index, or -1 - overflow index)
    int32_t last_loc = 0;  // < 2^31 slots and overflow entries per launch
    auto emit = [&](const double pos[3], const double cos[3], int element) {
        DevRecord r;
        for (int i = 0; i < 3; i++) { r.pos[i] = pos[i]; r.cos[i] = cos[i]; }
        r.element = element;
        r.raynum = raynum;
        if (nrec < (uint32_t)a.slots) {
            // slot-major: a wave's first records are contiguous
            last_loc = (int32_t)(nrec * a.slot_stride + a.ray_base + (uint32_t)t);
            a.rec_slots[last_loc] = r;
        } else {
            uint32_t o = atomicAdd(a.ovf_count, 1u);
            if (o < a.ovf_cap) {
                a.ovf_rec[o] = r;
                a.ovf_meta[2 * o] = a.ray_base + (uint32_t)t;
                a.ovf_meta[2 * o + 1] = nrec;
                last_loc = -1 - (int32_t)o;
            } else {
                atomicOr(a.err_flags, 4u);
                last_loc = kNoRecord;
            }
        }
        nrec++;
    };
    // flux: bin of this ray's last record on the flux element
    const bool flux_here = a.flux.active && a.flux.stage == a.si;
    int fbin = -1;
    bool fabs_ = false;
    int prev = -1;  // the element the ray leaves

    for (;;) {
        Best b;
        b.tests = 0;
        b.nodes = 0;
        // rays here are known to hit, so footprints would screen out little;
        // a ray leaving a stage-1 element may take its beam list
        closest_hit<R, G>(sv, st, ps, cs, mhc == 0 ? cell : -1, false, 0.0, 0.0, b, sun_stage ? prev : -1);
        if (G == 2 && b.fefail) atomicOr(a.err_flags, 16u);
        if (a.counters) {
            ev_lds[0][threadIdx.x] += b.tests;
            ev_lds[1][threadIdx.x] += b.nodes;
        }
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
            emit(ps, cs, 0);
            break;
        }
        const DevElem &e = sv.elems[st.elem0 + b.j];
        prev = b.j;
        double pe[3], ce[3], xd[3], df[3];
        int back;
        element_ray(e, ps, cs, pe, ce);
        if constexpr (FAST) {
            if (!hit_finish_d<G>(e, sv.aux, pe, ce, b.path, xd, df, &back)) atomicOr(a.err_flags, 16u);
        } else {
            // the winner again: the same arithmetic gives the same point
            ElemHit<R> h;
            element_hit<R, G>(e, sv.aux, pe, ce, h);
            for (int i = 0; i < 3; i++) { xd[i] = h.X[i]; df[i] = h.DF[i]; }
            back = h.back;
        }
        {
            double rp[3], rc[3];
            to_ref(xd, ce, e.org, e.R, rp, rc);
            emit(rp, rc, b.j + 1);
            if (flux_here && b.j + 1 == a.flux.elem) {
                fbin = flux_bin(a.flux, rp, rc);
                fabs_ = false;
            }
        }
        mhc++;

        // directions from here on are double in both precisions
        double cos_out[3];
        if (st.virt) {
            for (int i = 0; i < 3; i++) { cos_out[i] = ce[i]; cos_out_el[i] = ce[i]; }
        } else {
            if (e.interaction != 1 && e.interaction != 2) {
                atomicOr(a.err_flags, 1u);
                break;
            }
            const DevFace &f = sv.faces[back ? e.face_back : e.face_front];
            const double tv = test_value(e.interaction, f, sv.aux, ce, df);
            if (tv <= rng.uniform()) {
                if (last_loc >= 0) a.rec_slots[last_loc].element = -(b.j + 1);
                else if (last_loc != kNoRecord) a.ovf_rec[-1 - last_loc].element = -(b.j + 1);
                if (flux_here && b.j + 1 == a.flux.elem) fabs_ = true;
                break;
            }
            // sunshape error (first hit of a sun ray), slope error, reflection
            // or refraction, specularity error -- one perturb() call site
            double cin[3] = {ce[0], ce[1], ce[2]};
            const bool sun_err = a.inc_sun && sun_stage && mhc == 1;
            double pre = 0.0;
            const bool use_pre = sun_err && a.sun_pre != nullptr && rng.k == 3u && a.sun_pre[t].ok;
            if (use_pre) {
                pre = a.sun_pre[t].theta2;
                if (a.counters) ev_lds[2][threadIdx.x] = a.sun_pre[t].kend - rng.k;  // drawn by the sampler kernel
                rng.seek(a.sun_pre[t].kend);
            }
            const bool diffuse = (f.dist == 'f' || f.dist == 'F');
            for (int step = 0; step < 3; ++step) {
                if (step == 2) interact(rng, cin, df, e.interaction, f, cos_out_el);
                if (step == 0 ? !sun_err : !a.inc_err) continue;
                double bs[3], co[3];
                for (int i = 0; i < 3; i++)
                    bs[i] = step == 0 ? cin[i] : ((step == 1 || diffuse) ? df[i] : cos_out_el[i]);
                const int dist = step == 0 ? sv.sun.shape : f.dist;
                const double delop = (step == 0 ? sv.sun.sigma : (step == 1 ? f.slope : f.spec)) / 1000.0;
                perturb<FAST>(rng, bs, dist, delop, step == 1, sv.sun, sv.aux, step == 0 && use_pre, pre,
                              step == 2 && e.interaction == 2, df, co);
                for (int i = 0; i < 3; i++) {
                    if (step == 0) cin[i] = co[i];
                    else if (step == 1) df[i] = co[i];
                    else cos_out_el[i] = co[i];
                }
            }
            for (int i = 0; i < 3; i++) cos_out[i] = cos_out_el[i];
        }
        to_ref(xd, cos_out, e.org, e.R, ps, cs);
        if (!st.multihit) {
            passed = true;
            break;
        }
    }
    const uint32_t g = a.ray_base + (uint32_t)t;  // place in the launch's ray space
    a.rec_count[g] = nrec;
    if (a.out_flag != nullptr) a.out_flag[g] = passed ? 1u : 0u;
    if (a.counters) {
        wave_add(&a.counters[0], 1ull);
        wave_add(&a.counters[1], (unsigned long long)ev_lds[0][threadIdx.x]);
        wave_add(&a.counters[2], (unsigned long long)ev_lds[1][threadIdx.x]);
        wave_add(&a.counters[3], (unsigned long long)(rng.k - ev_lds[2][threadIdx.x]));
        wave_add(&a.counters[4], (unsigned long long)nrec);
    }
    if (passed && a.out_slots != nullptr) {
        DevRay r;
        if (mhc == 0) {
            r = a.in_rays[t];  // passed through untouched: hand on the ray as it came
        } else {
            to_ref(ps, cs, st.org, st.R, r.pos, r.cos);
            r.raynum = raynum;
            // source tag: the stage-1 element that sent the ray on (the
            // element of its last stage-1 record, a reflection), carried
            // unchanged through later stages
            r.src = 0;
            if (a.flux_src != nullptr) {
                if (!sun_stage) r.src = a.in_rays[t].src;
                else if (last_loc >= 0) r.src = (uint32_t)a.rec_slots[last_loc].element;
                else if (last_loc != kNoRecord) r.src = (uint32_t)a.ovf_rec[-1 - last_loc].element;
            }
            r.sun_index = sun_stage ? a.sun_index[t] : a.in_rays[t].sun_index;
        }
        a.out_slots[g] = r;
    }
    if (fbin >= 0 && (!a.flux.absorbed_only || fabs_)) {
        atomicAdd(&a.flux_bins[fbin], 1u);
        if (a.flux_src != nullptr && !sun_stage) {
            const uint32_t src = a.in_rays[t].src;
            if (src > 0u) atomicAdd(&a.flux_src[(size_t)(src - 1u) * a.flux_nbins + (uint32_t)fbin], 1u);
        }
    }
}

// Waves per SIMD (WV): the generic variant runs 2 (its registers spill,
// but a second wave hides more latency than the spills cost; the one with
// finite elements keeps the whole register file); the basic FP32 variant
// runs 3 (168 registers, 20 B of spills per lane: with beam lists the rays
// of the sun stage wait little on the BVH, and 4 waves in 128 registers
// spill more than the fourth wave hides); the basic FP64 variant 2 (3 is
// slower).
template <typename R, int G, int WV>
__global__ __launch_bounds__(kBlock, WV) void trace_stage_kernel(const TraceArgs *__restrict__ ap) {
    const TraceArgs &a = *ap;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n_rays) return;
    const int t = a.order != nullptr ? (int)(a.order[a.ray_base + (uint32_t)i] - a.ray_base) : i;
    trace_one<R, G>(a, t);
}

// Sun-plane bucket of stage-0 ray i (the point of its sun-ray index's two
// sun-plane draws on a kSunBucketSide-square grid over the sun box, in Morton
// order so neighbouring buckets stay close), counted into hist: key_rank
// holds (bucket, rank within the bucket) for each ray.
STD uint32_t sun_bucket(const DevSun &sun, uint32_t seed, uint64_t sun_index, uint32_t side_log2) {
    Rng rng;
    rng.begin(seed, 0, sun_index);
    double p[2];
    sun_plane_point(rng, sun, p);
    const double side = (double)(1u << side_log2);
    const double fx = (p[0] - sun.minx) / (sun.maxx - sun.minx), fy = (p[1] - sun.miny) / (sun.maxy - sun.miny);
    const uint32_t bx = (uint32_t)fmin(fmax(fx * side, 0.0), side - 1.0);
    const uint32_t by = (uint32_t)fmin(fmax(fy * side, 0.0), side - 1.0);
    uint32_t key = 0;
    for (uint32_t b = 0; b < side_log2; b++)
        key |= (((bx >> b) & 1u) << (2 * b)) | (((by >> b) & 1u) << (2 * b + 1));
    return key;
}

// Ray order, first pass: block b ranks rays [b kOrderRays, (b + 1)
// kOrderRays) within their buckets in LDS and writes its bucket counts as
// row b of table.
template <typename R>
__global__ __launch_bounds__(1024) void sun_order_count_kernel(DevSun sun, uint32_t seed, const uint64_t *sun_index,
                                                               int n, uint32_t *key_rank, uint32_t *table) {
    __shared__ uint32_t hist[kOrderBuckets];
    for (uint32_t b = threadIdx.x; b < kOrderBuckets; b += 1024) hist[b] = 0u;
    __syncthreads();
    const int base = (int)blockIdx.x * (int)kOrderRays;
    for (int k = 0; k < (int)(kOrderRays / 1024); k++) {
        const int i = base + k * 1024 + (int)threadIdx.x;
        if (i < n) {
            const uint32_t key = sun_bucket(sun, seed, sun_index[i], kOrderLog2);
            const uint32_t rank = atomicAdd(&hist[key], 1u);
            key_rank[i] = key | (rank << (2 * kOrderLog2));
        }
    }
    __syncthreads();
    uint32_t *row = table + (size_t)blockIdx.x * kOrderBuckets;
    for (uint32_t b = threadIdx.x; b < kOrderBuckets; b += 1024) row[b] = hist[b];
}

template <typename R>
void launch_sun_order(const SceneView &sv, uint32_t seed, const uint64_t *sun_index, int n, const OrderWork &w,
                      uint32_t *order, hipStream_t s) {
    static_assert(kOrderRays <= (1u << (32 - 2 * kOrderLog2)), "rank bits");
    if (n <= 0) return;
    const uint32_t nblocks = (uint32_t)((n + kOrderRays - 1) / kOrderRays);
    hipLaunchKernelGGL(sun_order_count_kernel<R>, dim3(nblocks), dim3(1024), 0, s, sv.sun, seed, sun_index, n,
                       w.key_rank, w.table);
    launch_order_colscan(w.table, nblocks, w.colsum, s);
    launch_tile_scan(w.colsum, w.cbase, kOrderBuckets, w.ctile, w.ctot, s);
    launch_order_place(w.key_rank, n, w.table, w.cbase, w.ctile, order, s);
}

// ------------------------------------------------------ batched scenes ---
// The item holding tile x: the last one whose first tile is <= x.
STD int batch_item(const BatchQueue &q, uint32_t x) {
    int lo = 0, hi = q.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (q.prefix[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void sun_bucket_batch_kernel(const BucketItem *items, BatchQueue q,
                                                                  uint32_t total, uint32_t *key_rank,
                                                                  uint32_t *hist) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const int k = batch_item(q, i);
    const BucketItem &it = items[k];
    const uint32_t key = it.key0 + sun_bucket(it.sun, it.seed, it.sun_index[i - q.prefix[k]], kBatchBucketLog2);
    key_rank[2 * i] = key;
    key_rank[2 * i + 1] = atomicAdd(&hist[key], 1u);
}

template <typename R>
void launch_sun_buckets_batch(const BucketItem *items, const BatchQueue &q, uint32_t total, uint32_t *key_rank,
                              uint32_t *hist, uint32_t nkeys, hipStream_t s) {
    launch_fill_u32(hist, nkeys, 0u, s);
    if (total == 0) return;
    hipLaunchKernelGGL(sun_bucket_batch_kernel<R>, dim3((total + kBlock - 1) / kBlock), dim3(kBlock), 0, s, items,
                       q, total, key_rank, hist);
}
// The wave's next tile (wave-uniform).
STD uint32_t wave_take(uint32_t *counter) {
    uint32_t v = 0;
    if ((threadIdx.x & 63) == 0) v = atomicAdd(counter, 1u);
    return (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)v, 0));
}

// Sun masks of several scenes: a tile is one wave's 64 * kMaskRays sun
// rays of one scene, and a wave takes kMaskChunk tiles in a row per atomic
// on the queue counter (one atomic per 256 sun rays made the counter's
// address the bottleneck of a launch).
constexpr uint32_t kMaskChunk = 8;
// Waves per SIMD held to those of the single-scene mask kernels (4 basic, 2
// generic): the queue loop would otherwise cost registers and occupancy.
template <typename R, int G>
__global__ __launch_bounds__(kBlock, G ? 2 : 4) void mask_batch_kernel(const MaskArgs *__restrict__ items,
                                                                       BatchQueue q) {
    const uint32_t total = q.prefix[q.n];
    for (;;) {
        const uint32_t t0 = wave_take(q.counter) * kMaskChunk;
        if (t0 >= total) break;
        const uint32_t t1 = min(t0 + kMaskChunk, total);
        for (uint32_t tile = t0; tile < t1; tile++) {
            const int k = __builtin_amdgcn_readfirstlane(batch_item(q, tile));
            const MaskArgs &a = items[k];
            const uint32_t local = tile - q.prefix[k];
            if (a.compact) {
                mask_grid_wave<R, G>(a, local);
            } else {
                for (int r = 0; r < kMaskRays; r++)
                    mask_lane<R, G>(a, (local * kMaskRays + (uint32_t)r) * 64u + (threadIdx.x & 63));
            }
        }
    }
}

// Tabulated-sunshape angles of several scenes' hitting rays: one lane per
// ray (q.prefix in rays).
template <typename R>
__global__ __launch_bounds__(kBlock) void sunshape_batch_kernel(const SunPreArgs *items, BatchQueue q,
                                                                uint32_t total) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const int k = batch_item(q, i);
    sunshape_icdf_one<R>(items[k], (int)(i - q.prefix[k]));
}

// A stage of several scenes: a tile is q.tile rays of one scene, traced by
// one wave (lane l takes rays l, l + 64, ...).
template <typename R, int G, int WV>
__global__ __launch_bounds__(kBlock, WV) void trace_batch_kernel(const TraceArgs *__restrict__ items,
                                                                         BatchQueue q) {
    const uint32_t total = q.prefix[q.n];
    const uint32_t lane = threadIdx.x & 63;
    for (;;) {
        const uint32_t tile = wave_take(q.counter);
        if (tile >= total) break;
        const int k = __builtin_amdgcn_readfirstlane(batch_item(q, tile));
        const TraceArgs &a = items[k];
        const uint32_t first = (tile - q.prefix[k]) * q.tile;
        for (uint32_t r = lane; r < q.tile; r += 64) {
            const uint32_t t = first + r;
            if (t < (uint32_t)a.n_rays)
                trace_one<R, G>(a, a.order != nullptr ? (int)(a.order[a.ray_base + t] - a.ray_base) : (int)t);
        }
    }
}

template <typename R>
void launch_mask_batch(const MaskArgs *items, const BatchQueue &q, int generic, int grid, hipStream_t s) {
    if (generic == 2) hipLaunchKernelGGL((mask_batch_kernel<R, 2>), dim3(grid), dim3(kBlock), 0, s, items, q);
    else if (generic) hipLaunchKernelGGL((mask_batch_kernel<R, 1>), dim3(grid), dim3(kBlock), 0, s, items, q);
    else hipLaunchKernelGGL((mask_batch_kernel<R, 0>), dim3(grid), dim3(kBlock), 0, s, items, q);
}
template <typename R>
void launch_sunshape_batch(const SunPreArgs *items, const BatchQueue &q, uint32_t total, hipStream_t s) {
    if (total == 0) return;
    hipLaunchKernelGGL(sunshape_batch_kernel<R>, dim3((total + kBlock - 1) / kBlock), dim3(kBlock), 0, s, items, q,
                       total);
}
template <typename R>
void launch_trace_batch(const TraceArgs *items, const BatchQueue &q, int generic, const TraceShape &shape, int grid,
                        hipStream_t s) {
    if (generic == 2)
        hipLaunchKernelGGL((trace_batch_kernel<R, 2, 1>), dim3(grid), dim3(kBlock), 0, s, items, q);
    else if (generic && shape.waves >= 2)
        hipLaunchKernelGGL((trace_batch_kernel<R, 1, 2>), dim3(grid), dim3(kBlock), 0, s, items, q);
    else if (generic)
        hipLaunchKernelGGL((trace_batch_kernel<R, 1, 1>), dim3(grid), dim3(kBlock), 0, s, items, q);
    else if constexpr (sizeof(R) == 8) {
        if (shape.waves >= 2) hipLaunchKernelGGL((trace_batch_kernel<R, 0, 2>), dim3(grid), dim3(kBlock), 0, s, items, q);
        else hipLaunchKernelGGL((trace_batch_kernel<R, 0, 1>), dim3(grid), dim3(kBlock), 0, s, items, q);
    } else {
        if (shape.waves >= 4) hipLaunchKernelGGL((trace_batch_kernel<R, 0, 4>), dim3(grid), dim3(kBlock), 0, s, items, q);
        else if (shape.waves == 3)
            hipLaunchKernelGGL((trace_batch_kernel<R, 0, 3>), dim3(grid), dim3(kBlock), 0, s, items, q);
        else hipLaunchKernelGGL((trace_batch_kernel<R, 0, 2>), dim3(grid), dim3(kBlock), 0, s, items, q);
    }
}

template <typename R>
void launch_sun_mask(const MaskArgs &a, hipStream_t s) {
    if (a.compact) {
        const uint32_t per_block = kBlock * kMaskRays;
        const uint32_t blocks = (a.n + per_block - 1) / per_block;
        if (a.generic == 2) hipLaunchKernelGGL((sun_mask_grid_kernel<R, 2>), dim3(blocks), dim3(kBlock), 0, s, a);
        else if (a.generic) hipLaunchKernelGGL((sun_mask_grid_kernel<R, 1>), dim3(blocks), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((sun_mask_grid_kernel<R, 0>), dim3(blocks), dim3(kBlock), 0, s, a);
        return;
    }
    const uint32_t blocks = (a.n + kBlock - 1) / kBlock;
    if (a.generic == 2) hipLaunchKernelGGL((sun_mask_kernel<R, 2>), dim3(blocks), dim3(kBlock), 0, s, a);
    else if (a.generic) hipLaunchKernelGGL((sun_mask_kernel<R, 1>), dim3(blocks), dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL((sun_mask_kernel<R, 0>), dim3(blocks), dim3(kBlock), 0, s, a);
}
template <typename R>
void launch_trace_stage(const TraceArgs &a, const TraceArgs *dev_a, hipStream_t s) {
    const int blocks = (a.n_rays + kBlock - 1) / kBlock;
    if (a.generic == 2) hipLaunchKernelGGL((trace_stage_kernel<R, 2, 1>), dim3(blocks), dim3(kBlock), 0, s, dev_a);
    else if (a.generic) hipLaunchKernelGGL((trace_stage_kernel<R, 1, 2>), dim3(blocks), dim3(kBlock), 0, s, dev_a);
    else if constexpr (sizeof(R) == 8)
        hipLaunchKernelGGL((trace_stage_kernel<R, 0, 2>), dim3(blocks), dim3(kBlock), 0, s, dev_a);
    else hipLaunchKernelGGL((trace_stage_kernel<R, 0, 3>), dim3(blocks), dim3(kBlock), 0, s, dev_a);
}

}  // namespace st
