// k_momentum_zm_f32.hip - FP32 instantiation of the fused z-marching
// momentum kernel (k_momentum_zm.inl) for the precision form.  Products
// feeding a sum inside one expression may fuse into FMAs, and quotients by
// the grid spacings and by 3 are products with reciprocals (one rounding
// each): the precision form is checked against the FP64 oracle by its error
// bars, not bit for bit, and these forms save VALU issue slots.
#pragma clang fp contract(on)
#define ELMM_ZM_RECIP 1

#include "k_momentum_zm.inl"

namespace elmm {

template void launch_momentum_zm<float>(const Geo &, const float *, const float *, const float *, const float *,
                                        const float *, float *const[3], float *const[3], const uint8_t *const[3],
                                        StageConst, int, hipStream_t);

}  // namespace elmm
