// k_sgs.hip - FP64 instantiation of the subgrid pass (k_sgs.inl).
#define ELMM_SGS_RECIP 0
#include "k_sgs.inl"

namespace elmm {
template void launch_sgs<double>(const Geo &, const double *, const double *, const double *, double *, double,
                                 double, double, SgsWidth, int, double *, int, hipStream_t);
}  // namespace elmm
