// Segmented line sweeps, FP64 instantiation.
#include "adsg_seg.inc"

namespace adsg {
template hipError_t launch_line_sweep_seg<double, double>(double*, const line_geometry&,
                                                          const seg_factors<double>&, hipStream_t,
                                                          const rhs_fixup<double>*, const seg_apply*, bool*);
template hipError_t launch_sweep_seg_batch<double>(const seg_item<double>*, int, int, hipStream_t);
}  // namespace adsg
