// k_momentum_zm_f32.hip - FP32 instantiation of the fused momentum kernel
// (k_momentum_zm.inl) for the precision form.  Built with
// -ffp-contract=fast: the kernel issues one VALU instruction per 4 cycles
// per SIMD, the wave64 peak, so fusing its multiply-add pairs shortens it;
// the FP64 form stays uncontracted.
#include "k_momentum_zm.inl"

namespace elmm {

template void launch_momentum_zm<float>(const Geo &, const float *, const float *, const float *, const float *, const float *,
                                        const float *, const float *, float *const[3], float *const[3], const uint8_t *const[3],
                                        StageConst, int, int, hipStream_t);

}  // namespace elmm
