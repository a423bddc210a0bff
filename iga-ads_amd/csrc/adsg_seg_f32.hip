// Segmented line sweeps, FP32 instantiation.
#include "adsg_seg.inc"

namespace adsg {
template hipError_t launch_line_sweep_seg<float, float>(float*, const line_geometry&,
                                                        const seg_factors<float>&, hipStream_t,
                                                        const rhs_fixup<float>*, const seg_apply*, bool*);
template hipError_t launch_sweep_seg_batch<float>(const seg_item<float>*, int, int, hipStream_t);
}  // namespace adsg
