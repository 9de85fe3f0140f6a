// mixed_nk_c.hip -- instantiations of the MFMA contractions (tiled, streaming, packed)
// for NK = 10 .. 13 (K-steps of
// 16 target columns per chip-row pair).
#define AMPCOR_MIXED_INSTANTIATE
#include "mixed_correlate.h"
#include "packed_correlate.h"

namespace ampcor_hip {
template hipError_t launch_correlate_mixed_nk<10>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_mixed_nk<11>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_mixed_nk<12>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_mixed_nk<13>(hipStream_t, const MixedShape &, const uint16_t *,
                                                const uint16_t *, int, int, int, bool, float *,
                                                const uint16_t *, const uint16_t *);
template hipError_t launch_correlate_packed_nk<10>(hipStream_t, const MixedShape &,
                                                  const uint16_t *, const uint16_t *, int, int, int,
                                                  bool, float *, const uint16_t *,
                                                  const uint16_t *);
template hipError_t launch_correlate_packed_nk<11>(hipStream_t, const MixedShape &,
                                                  const uint16_t *, const uint16_t *, int, int, int,
                                                  bool, float *, const uint16_t *,
                                                  const uint16_t *);
}  // namespace ampcor_hip
