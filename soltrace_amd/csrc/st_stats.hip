// Flux statistics of one element over a run's records, which stay in HBM
// (the GUI's ElementStatistics::Compute / BinRaysXY,
// app/src/project.cpp:1139-1436).  Two passes over each stage's record
// chunks, one lane per record:
//   bounds -- element-frame x/y range of every record on the element
//             (autoscale), as order-preserving integer min/max;
//   bin    -- each ray's last record on the element (absorbed ones only when
//             asked): element-frame point, cylinder unrolled, linear-scan bin,
//             centroid sums in fixed point, so every sum is independent of
//             the order lanes finish in.
// Compiled without FMA contraction: the stage -> element transform rounds as
// the reference's st_transform_to_local does, so every bin matches.
#include <hip/hip_runtime.h>

#include "st_device.h"
#include "st_kernels.h"

namespace st {

static constexpr int kStatsBlock = 256;

// Monotonic map of a double onto uint64 (a < b  <=>  key(a) < key(b)).
__device__ __forceinline__ unsigned long long order_key(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__global__ __launch_bounds__(kStatsBlock) void stats_bounds_kernel(const DevRecord *rec, uint32_t n, int32_t elem,
                                                                   FluxSpec fl, unsigned long long *mm) {
    const uint32_t i = blockIdx.x * kStatsBlock + threadIdx.x;
    if (i >= n) return;
    const DevRecord r = rec[i];
    if (abs(r.element) != elem) return;
    double pe[3], ce[3];
    to_local(r.pos, r.cos, fl.org, fl.R, pe, ce);
    atomicMin(&mm[0], order_key(pe[0]));
    atomicMax(&mm[1], order_key(pe[0]));
    atomicMin(&mm[2], order_key(pe[1]));
    atomicMax(&mm[3], order_key(pe[1]));
}

// GridIncrement of app/src/project.cpp:1398-1405: the first k >= 0 with
// lo + k*b >= x, minus one (the same search as the trace kernel's flux map).
__device__ __forceinline__ int stats_bin(double x, double lo, double b, int n) {
    const double est = (x - lo) / b;
    int k;
    if (!(est > 0.0)) k = 0;
    else if (est > (double)(n + 1)) k = n + 2;
    else k = (int)::ceil(est);
    while (k > 0 && !((lo + (k - 1) * b) < x)) k--;
    while (k <= n + 1 && (lo + k * b) < x) k++;
    return k - 1;
}

// acc: [0..2] centroid sums (micrometres, two's complement), [3] rays
// considered, [4] rays binned, [5] 1 + the largest considered record index
// (offset by `base`, the chunk's first index in stage order).
__global__ __launch_bounds__(kStatsBlock) void stats_bin_kernel(const DevRecord *rec, uint32_t n, int32_t elem,
                                                                FluxSpec fl, uint64_t base, uint32_t *bins,
                                                                unsigned long long *acc) {
    const uint32_t i = blockIdx.x * kStatsBlock + threadIdx.x;
    if (i >= n) return;
    const DevRecord r = rec[i];
    if (abs(r.element) != elem) return;
    // a ray's records are contiguous within a chunk: the record counts only
    // if no later record of the same ray lands on the element
    for (uint32_t j = i + 1; j < n; j++) {
        const DevRecord q = rec[j];
        if (q.raynum != r.raynum) break;
        if (abs(q.element) == elem) return;
    }
    if (fl.absorbed_only && r.element >= 0) return;
    double pe[3], ce[3];
    to_local(r.pos, r.cos, fl.org, fl.R, pe, ce);
    double x = pe[0];
    const double y = pe[1], z = pe[2];
    atomicAdd(&acc[0], (unsigned long long)llrint(x * 1e6));
    atomicAdd(&acc[1], (unsigned long long)llrint(y * 1e6));
    atomicAdd(&acc[2], (unsigned long long)llrint(z * 1e6));
    atomicAdd(&acc[3], 1ull);
    atomicMax(&acc[5], (unsigned long long)(base + i + 1));
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
    const int gx = stats_bin(x, fl.xmin, fl.bx, fl.nx), gy = stats_bin(y, fl.ymin, fl.by, fl.ny);
    if (gx >= 0 && gx < fl.nx && gy >= 0 && gy < fl.ny) {
        atomicAdd(&bins[gx * fl.ny + gy], 1u);
        atomicAdd(&acc[4], 1ull);
    }
}

void launch_stats_bounds(const DevRecord *rec, uint32_t n, int32_t elem, const FluxSpec &fl, unsigned long long *mm,
                         hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(stats_bounds_kernel, dim3((n + kStatsBlock - 1) / kStatsBlock), dim3(kStatsBlock), 0, s, rec,
                       n, elem, fl, mm);
}

void launch_stats_bin(const DevRecord *rec, uint32_t n, int32_t elem, const FluxSpec &fl, uint64_t base,
                      uint32_t *bins, unsigned long long *acc, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(stats_bin_kernel, dim3((n + kStatsBlock - 1) / kStatsBlock), dim3(kStatsBlock), 0, s, rec, n,
                       elem, fl, base, bins, acc);
}

}  // namespace st
