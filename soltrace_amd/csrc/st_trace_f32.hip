// FP32-bulk instantiation of the trace kernels.
#include "st_trace_impl.h"

namespace st {
template void launch_sun_mask<float>(const MaskArgs &, hipStream_t);
template void launch_sunshape_d<float>(const SunPreArgs &, hipStream_t);
template void launch_trace_stage<float>(const TraceArgs &, hipStream_t);
}  // namespace st
