// Trace kernels for CDNA4 (gfx950): templates, instantiated for FP64 in
// st_trace_f64.hip and for FP32 in st_trace_f32.hip.
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
#pragma once
#include <hip/hip_runtime.h>

#include "st_device.h"
#include "st_kernels.h"

namespace st {

static constexpr int kBlock = 256;

// ------------------------------------------------------ sun-ray source ----
// coretrace/generateray.cpp:58-161 (the sun-plane draws are full 53-bit)
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
    double x = sun.minx + (sun.maxx - sun.minx) * rng.uniform();
    double y = sun.miny + (sun.maxy - sun.miny) * rng.uniform();
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
    uint32_t tests;      // element tests (event counter)
    uint32_t nodes;      // BVH nodes visited (event counter)
};

// Test element j; keep it when it is the closest accepted hit so far
// (coretrace/raytrace.cpp:740-768; ties go to the lower element index, the
// order a brute-force scan meets them).
template <typename R>
STD void consider(const SceneView &sv, const DevStage &st, int j, const double ps[3], const double cs[3],
                  Best<R> &b) {
    const DevElem &e = sv.elems[st.elem0 + j];
    if (!e.enabled) return;
    b.tests++;
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
        // stackless depth-first walk: descend on a box hit, else jump to the
        // node's skip index (no per-lane stack, so no scratch memory)
        int i = 0;
        while (i < st.nbvh) {
            const DevBvhNode nd = nodes[i];
            b.nodes++;
            float t0 = 0.0f, t1 = 3.0e38f;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                float ta = (nd.lo[k] - o[k]) * inv[k];
                float tb = (nd.hi[k] - o[k]) * inv[k];
                t0 = fmaxf(t0, fminf(ta, tb));
                t1 = fminf(t1, fmaxf(ta, tb));
            }
            const bool enter = (t0 <= t1) && !((double)t0 > b.path * 1.000001 + 1.0e-3);
            if (!enter) {
                i = nd.skip;
                continue;
            }
            if (nd.leaf) {
                const int first = nd.leaf >> 3, cnt = nd.leaf & 7;
                for (int k = 0; k < cnt; ++k) consider<R>(sv, st, sv.bvh_items[first + k], ps, cs, b);
                i = nd.skip;
            } else {
                i = i + 1;
            }
        }
        return;
    }
    for (int j = 0; j < st.nelem; ++j) consider<R>(sv, st, j, ps, cs, b);
}

// Any accepted hit at all (the first-hit test of a sun ray).
template <typename R>
STD bool any_hit(const SceneView &sv, const DevStage &st, const double ps[3], const double cs[3],
                 const double *psun, uint32_t *tests) {
    Best<R> b;
    b.tests = 0;
    b.nodes = 0;
    closest_hit<R>(sv, st, ps, cs, psun, b);
    *tests += b.tests;
    return b.path < 1e99;
}

// ---------------------------------------------------------- sun mask ------
template <typename R>
__global__ __launch_bounds__(kBlock) void sun_mask_kernel(MaskArgs a) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    bool hit = false;
    if (t < a.n) {
        const DevStage st = a.sv.stages[0];
        Rng rng;
        rng.begin(a.seed, 0, a.s0 + t);
        double pg[3], cg[3], psun[2], ps[3], cs[3];
        gen_sun_ray(rng, a.sv.sun, st, pg, cg, psun);
        to_local(pg, cg, st.org, st.R, ps, cs);
        uint32_t tests = 0;
        hit = any_hit<R>(a.sv, st, ps, cs, (a.sv.use_grid && !a.sv.sun.point_source) ? psun : nullptr, &tests);
        if (a.counters) {
            atomicAdd(&a.counters[0], 1ull);
            atomicAdd(&a.counters[1], (unsigned long long)tests);
            atomicAdd(&a.counters[3], (unsigned long long)rng.k);
        }
    }
    const uint64_t word = __ballot(hit);
    if ((threadIdx.x & 63) == 0) {
        const uint32_t w = t >> 6;
        if ((uint64_t)w * 64 < a.n) a.mask[w] = word;
    }
}

// ------------------------------------------- cooperative sunshape draw -----
// The 'd' sunshape accepts about one rejection iteration in a hundred for a
// Buie aureole, and a lane-per-ray loop runs a whole wave until its slowest
// lane is done.  Here the 64 lanes of a wave share the wave's unresolved
// rays: each round, lane l takes ray (l mod m) of the m unresolved ones and
// iteration next_j + l/m of that ray's loop; each ray keeps the smallest
// accepted iteration.  The result is exactly what the serial loop of
// coretrace/errors.cpp:160-185 returns on the same stream.
STD int nth_set_bit(uint64_t v, int n) {
    for (int i = 0; i < n; ++i) v &= v - 1;
    return __ffsll((unsigned long long)v) - 1;
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
    for (int round = 0; U != 0; ++round) {
        const int m = __popcll(U);
        const int rank = lane % m, rep = lane / m;
        const int rl = nth_set_bit(U, rank);
        const uint32_t g = (uint32_t)__shfl((int)next_g, rl) + (uint32_t)rep;
        const uint64_t s = ((uint64_t)(uint32_t)__shfl((int)sid_hi, rl) << 32) | (uint32_t)__shfl((int)sid_lo, rl);
        double t2 = 0.0;
        const int off = sunshape_d_group(a.seed, s, a.k0, g, a.sun, a.aux, &t2);
        const bool acc = off >= 0;
        const uint64_t accm = __ballot(acc);
        // lanes that served this lane's ray: rank r_own, r_own + m, ...
        const bool owner = (U >> lane) & 1ull;
        const int r_own = __popcll(U & ((1ull << lane) - 1ull));
        uint64_t mine = 0;
        if (owner)
            for (int q = r_own; q < 64; q += m) mine |= 1ull << q;
        const uint64_t hit = accm & mine;
        const int first = hit ? (__ffsll((unsigned long long)hit) - 1) : 0;
        const double t2_first = __shfl(t2, first);
        const int off_first = __shfl(off, first);
        bool resolved = false;
        if (owner) {
            if (hit) {
                const uint32_t gg = next_g + (uint32_t)((first - r_own) / m);
                const uint32_t jj = 4u * gg + (uint32_t)off_first;
                SunPre p;
                p.theta2 = t2_first / 1000000.0;
                p.kend = a.k0 + 3u * (jj + 1u);
                p.ok = 1u;
                a.out[myray] = p;
                resolved = true;
            } else {
                next_g += (uint32_t)__popcll(mine);
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
}

template <typename R>
void launch_sunshape_d(const SunPreArgs &a, hipStream_t s) {
    const int blocks = (a.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(sunshape_d_kernel<R>, dim3(blocks), dim3(kBlock), 0, s, a);
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
    Rng rng;
    if (sun_stage) {
        sun_index = a.sun_index[t];
        raynum = a.raynum_base + (uint32_t)t + 1u;
        rng.begin(a.seed, 0, sun_index);
        gen_sun_ray(rng, sv.sun, st, pg, cg, psun);
    } else {
        const DevRay r = a.in_rays[t];
        raynum = r.raynum;
        sun_index = r.sun_index;
        for (int i = 0; i < 3; i++) { pg[i] = r.pos[i]; cg[i] = r.cos[i]; }
        rng.begin(a.seed, (uint32_t)a.si, sun_index);
    }
    double ps[3], cs[3];
    to_local(pg, cg, st.org, st.R, ps, cs);

    constexpr bool FAST = sizeof(R) == 4;
    double cos_out_el[3] = {0.0, 0.0, 0.0};
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

    uint32_t ev_tests = 0, ev_nodes = 0;
    for (;;) {
        Best<R> b;
        b.tests = 0;
        b.nodes = 0;
        const bool first_sun = sun_stage && mhc == 0 && sv.use_grid && !sv.sun.point_source;
        closest_hit<R>(sv, st, ps, cs, first_sun ? psun : nullptr, b);
        ev_tests += b.tests;
        ev_nodes += b.nodes;
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
        double xd[3], df[3];
        int back;
        if constexpr (FAST) {
            double pe[3], ce[3];
            to_local(ps, cs, e.org, e.R, pe, ce);
            pe[0] = pe[0] + 1.0e-5 * ce[0];
            pe[1] = pe[1] + 1.0e-5 * ce[1];
            pe[2] = pe[2] + 1.0e-5 * ce[2];
            hit_finish_d(e, sv.aux, pe, ce, b.path, xd, df, &back);
        } else {
            for (int i = 0; i < 3; i++) { xd[i] = b.h.X[i]; df[i] = b.h.DF[i]; }
            back = b.h.back;
        }
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

        // directions from here on are double in both precisions
        double cos_out[3];
        const double *cin0 = b.ce;
        if (st.virt) {
            for (int i = 0; i < 3; i++) { cos_out[i] = cin0[i]; cos_out_el[i] = cin0[i]; }
        } else {
            if (e.interaction != 1 && e.interaction != 2) {
                atomicOr(a.err_flags, 1u);
                break;
            }
            const DevFace &f = sv.faces[back ? e.face_back : e.face_front];
            const double tv = test_value(e.interaction, f, sv.aux, cin0, df);
            if (tv <= rng.uniform()) {
                pend.element = -pend.element;
                if (fhave && b.j + 1 == a.flux.elem) fabs_ = true;
                break;
            }
            double cin[3] = {cin0[0], cin0[1], cin0[2]};
            if (a.inc_sun && sun_stage && mhc == 1) {
                double co[3];
                if (a.sun_pre != nullptr && rng.k == 3u && a.sun_pre[t].ok) {
                    rng.seek(a.sun_pre[t].kend);
                    sun_error_pre<FAST>(rng, cin, a.sun_pre[t].theta2, co);
                } else {
                    errors<FAST>(rng, cin, 1, sv.sun, sv.aux, e.interaction, f, co, df);
                }
                cin[0] = co[0]; cin[1] = co[1]; cin[2] = co[2];
            }
            if (a.inc_err) {
                double co[3];
                normal_errors<FAST>(rng, df, f, co);
                df[0] = co[0]; df[1] = co[1]; df[2] = co[2];
            }
            interact(rng, cin, df, e.interaction, f, cos_out_el);
            if (a.inc_err) {
                double base[3];
                const bool diffuse = (f.dist == 'f' || f.dist == 'F');
                for (int i = 0; i < 3; i++) base[i] = diffuse ? df[i] : cos_out_el[i];
                double co[3];
                errors<FAST>(rng, base, 2, sv.sun, sv.aux, e.interaction, f, co, df);
                cos_out_el[0] = co[0]; cos_out_el[1] = co[1]; cos_out_el[2] = co[2];
            }
            for (int i = 0; i < 3; i++) cos_out[i] = cos_out_el[i];
        }
        double po[3] = {xd[0], xd[1], xd[2]};
        to_ref(po, cos_out, e.org, e.R, ps, cs);
        to_ref(ps, cs, st.org, st.R, pg, cg);
        if (!st.multihit) {
            passed = true;
            break;
        }
    }
    flush();
    a.rec_count[t] = nrec;
    a.out_flag[t] = passed ? 1u : 0u;
    if (a.counters) {
        atomicAdd(&a.counters[0], 1ull);
        atomicAdd(&a.counters[1], (unsigned long long)ev_tests);
        atomicAdd(&a.counters[2], (unsigned long long)ev_nodes);
        atomicAdd(&a.counters[3], (unsigned long long)rng.k);
        atomicAdd(&a.counters[4], (unsigned long long)nrec);
    }
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

}  // namespace st
