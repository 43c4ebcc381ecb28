// Compiled with -ffp-contract=off so the FP64 path keeps the reference's
// separate multiply and add roundings (coretrace is built without FMA).
#include "st_trace_impl.h"

namespace st {
template void launch_sun_mask<double>(const MaskArgs &, hipStream_t);
template void launch_sunshape_d<double>(const SunPreArgs &, hipStream_t);
template void launch_trace_stage<double>(const TraceArgs &, const TraceArgs *, hipStream_t);
template void launch_sun_buckets_batch<double>(const BucketItem *, const BatchQueue &, uint32_t, uint32_t *, uint32_t *,
                                                uint32_t, hipStream_t);
template void launch_sun_order<double>(const SceneView &, uint32_t, const uint64_t *, int, const OrderWork &, uint32_t *,
                                        hipStream_t);
template void launch_mask_batch<double>(const MaskArgs *, const BatchQueue &, int, int, hipStream_t);
template void launch_sunshape_batch<double>(const SunPreArgs *, const BatchQueue &, uint32_t, hipStream_t);
template void launch_trace_batch<double>(const TraceArgs *, const BatchQueue &, int, const TraceShape &, int,
                                      hipStream_t);
}  // namespace st
