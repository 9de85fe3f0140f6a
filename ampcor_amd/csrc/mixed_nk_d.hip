// mixed_nk_d.hip -- instantiations of the MFMA contraction for NK = 14 .. 17 (K-steps of
// 16 target columns per chip-row pair), and of the packed contraction for NK = 17.
#define AMPCOR_MIXED_INSTANTIATE
#include "packed_correlate.h"

namespace ampcor_hip {
template hipError_t launch_correlate_mixed_nk<14>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_mixed_nk<15>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_mixed_nk<16>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_mixed_nk<17>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_packed_nk<17>(hipStream_t, const MixedShape &,
                                                   const uint16_t *, const uint16_t *, int, int,
                                                   int, bool, float *, const uint16_t *,
                                                   const uint16_t *);
}  // namespace ampcor_hip
