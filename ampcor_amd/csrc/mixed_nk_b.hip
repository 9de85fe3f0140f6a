// mixed_nk_b.hip -- instantiations of the MFMA contractions (tiled, streaming, packed)
// for NK = 6 .. 9 (K-steps of
// 16 target columns per chip-row pair).
#define AMPCOR_MIXED_INSTANTIATE
#include "mixed_correlate.h"
#include "packed_correlate.h"

namespace ampcor_hip {
template hipError_t launch_correlate_mixed_nk<6>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_mixed_nk<7>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_mixed_nk<8>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_mixed_nk<9>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_packed_nk<6>(hipStream_t, const MixedShape &,
                                                  const uint16_t *, const uint16_t *, int, int, int,
                                                  bool, float *, const uint16_t *,
                                                  const uint16_t *);
template hipError_t launch_correlate_packed_nk<7>(hipStream_t, const MixedShape &,
                                                  const uint16_t *, const uint16_t *, int, int, int,
                                                  bool, float *, const uint16_t *,
                                                  const uint16_t *);
template hipError_t launch_correlate_packed_nk<8>(hipStream_t, const MixedShape &,
                                                  const uint16_t *, const uint16_t *, int, int, int,
                                                  bool, float *, const uint16_t *,
                                                  const uint16_t *);
template hipError_t launch_correlate_packed_nk<9>(hipStream_t, const MixedShape &,
                                                  const uint16_t *, const uint16_t *, int, int, int,
                                                  bool, float *, const uint16_t *,
                                                  const uint16_t *);
}  // namespace ampcor_hip
