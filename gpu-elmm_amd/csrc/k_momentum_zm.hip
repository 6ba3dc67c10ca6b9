// k_momentum_zm.hip - FP64 instantiation of the fused momentum kernel
// (k_momentum_zm.inl); built with -ffp-contract=off, so its arithmetic is
// the reference's operation by operation.
#include "k_momentum_zm.inl"

namespace elmm {

size_t momentum_zm_lds_bytes(int elem) { return (size_t)NSLOT * Tile<32, 8>::PS * elem; }

bool momentum_zm_supported(const Geo &g) { return !g.variable_z; }

template void launch_momentum_zm<double>(const Geo &, const double *, const double *, const double *, const double *, const double *,
                                         const double *, const double *, double *const[3], double *const[3],
                                         const uint8_t *const[3], StageConst, int, int, hipStream_t);
}  // namespace elmm
