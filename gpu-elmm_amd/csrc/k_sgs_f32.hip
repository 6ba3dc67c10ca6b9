// k_sgs_f32.hip - FP32 instantiation of the subgrid pass (k_sgs.inl) for the
// precision form: quotients by the spacings as products with reciprocals and
// FMA contraction inside expressions (checked against the FP64 oracle by the
// precision form's error bars).
#pragma clang fp contract(on)
#define ELMM_SGS_RECIP 1
#include "k_sgs.inl"

namespace elmm {
template void launch_sgs<float>(const Geo &, const float *, const float *, const float *, float *, double, double,
                                double, SgsWidth, int, double *, int, hipStream_t);
}  // namespace elmm
