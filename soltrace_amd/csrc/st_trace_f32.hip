// FP32-bulk instantiation of the trace kernels.
#include "st_trace_impl.h"

namespace st {
template void launch_sun_mask<float>(const MaskArgs &, hipStream_t);
template void launch_sunshape_d<float>(const SunPreArgs &, hipStream_t);
template void launch_trace_stage<float>(const TraceArgs &, const TraceArgs *, hipStream_t);
template void launch_sun_buckets_batch<float>(const BucketItem *, const BatchQueue &, uint32_t, uint32_t *, uint32_t *,
                                                uint32_t, hipStream_t);
template void launch_sun_order<float>(const SceneView &, uint32_t, const uint64_t *, int, const OrderWork &, uint32_t *,
                                        hipStream_t);
template void launch_mask_batch<float>(const MaskArgs *, const BatchQueue &, int, int, hipStream_t);
template void launch_sunshape_batch<float>(const SunPreArgs *, const BatchQueue &, uint32_t, hipStream_t);
template void launch_trace_batch<float>(const TraceArgs *, const BatchQueue &, int, const TraceShape &, int,
                                      hipStream_t);
}  // namespace st
