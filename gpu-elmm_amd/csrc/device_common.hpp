// device_common.hpp - HBM layout and shared device helpers of the gfx950
// LES kernels.
//
// Every 3-D field lives in one layout: index space -2..n+3 in each direction
// (n = Prnx/Prny/Prnz), x fastest, with a front pad so that interior element
// i=1 of every row starts on a 128-byte boundary for FP64 (64 bytes for FP32)
// and the row pitch a multiple of 16 elements.  A wavefront of 64 lanes
// sweeping i=1..64 of a row therefore issues aligned, fully coalesced loads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace elmm {

constexpr int kFrontPad = 13;  // elements before i=-2 in a row
constexpr int kRowAlign = 16;  // row pitch multiple (elements)

struct Geo {
  int nx, ny, nz;     // Prnx, Prny, Prnz
  int n[3][3];        // component extents: n[c][d]
  long sy, sz;        // strides in elements
  long base;          // element offset of (0,0,0)
  long ntot;          // elements per field allocation
  double dxmin, dymin, dzmin;
  // RN(1/dxmin) etc. in both field types, or 0 where the reciprocal is not
  // close enough for div_rn (then the kernels divide)
  double rdx, rdy, rdz;
  float rdxf, rdyf, rdzf;
  int variable_z;
  int sgs_cellwise;   // one thread per cell SGS kernel instead of the z-march
  int sgs_zchunk;     // levels per block of the SGS z-march (0: default)
  int bt[6];          // velocity boundary types
  int sbt[6];         // scalar boundary types
  double side_scal[6];
  // 1-D coordinate tables (device pointers already offset so index i works
  // directly for the reference's index i)
  const double *dxPr, *dyPr, *dzPr, *dzW, *zPr, *zW;
};

__host__ __device__ inline long gix(const Geo &g, int i, int j, int k) {
  return g.base + (long)i + (long)j * g.sy + (long)k * g.sz;
}

// x / d correctly rounded for a divisor d whose reciprocal r = RN(1/d) is
// known: q = RN(x r) is within one ulp of x/d when |r d - 1| <= 2^-(p+1)
// (checked on the host, see div_recip), the remainder x - q d is then exact
// in one fma, and RN(q + rem r) is the correctly rounded quotient
// (P. Markstein, IBM J. Res. Dev. 34, 1990).  For rem = 0 that sum is q
// itself except for the sign of a zero quotient, which q carries (sign(q) =
// sign(x/d), and a non-zero result has the sign of q): copysign restores it
// without a branch.  div_rn_nz needs r != 0; div_rn divides when r = 0.
template <typename T>
__device__ inline T div_rn_nz(T x, T d, T r) {
  const T q = x * r;
  const T rem = fma(-q, d, x);
  return copysign(fma(rem, r, q), q);
}
template <typename T>
__device__ inline T div_rn(T x, T d, T r) {
  if (r == (T)0) return x / d;
  return div_rn_nz(x, d, r);
}

// RN(1/d) if it satisfies div_rn's condition, else 0 (host).
inline double div_recip(double d) {
  const double r = 1.0 / d;
  const double e = __builtin_fma(r, d, -1.0);  // exact
  return __builtin_fabs(e) <= 0x1p-54 ? r : 0.0;
}
inline float div_recipf(float d) {
  const float r = 1.0f / d;
  const double e = (double)r * (double)d - 1.0;  // exact in FP64
  return __builtin_fabs(e) <= 0x1p-25 ? r : 0.0f;
}

template <typename T>
__device__ inline T recip3() {
  return sizeof(T) == 4 ? (T)(1.0f / 3.0f) : (T)(1.0 / 3.0);
}

// Host operations that HIP refuses in every thread while any stream is
// being captured into a graph (allocation, free, synchronous copies, event
// and attribute changes) hold this lock shared; a graph capture holds it
// exclusively.  Re-entrant within a thread, and a no-op in the capturing
// thread itself (which makes no such call during its capture).
class HostOpLock {
 public:
  HostOpLock();
  ~HostOpLock();
  HostOpLock(const HostOpLock &) = delete;
  HostOpLock &operator=(const HostOpLock &) = delete;

 private:
  bool own_;
};

// Raises a kernel's dynamic-LDS limit above 64 KiB once per kernel and size
// (host; later launches, including those recorded into a graph, make no
// attribute call).
bool lds_limit_raise(const void *fn, size_t bytes);

// Per-cell code bytes used by the bulk kernels.
enum : uint8_t {
  CODE_ADV_MASK = 3,   // 0: 4th order, 1: 2nd order (Xtype<0), 2: none (Xtype>0)
  CODE_CORRECTED = 4,  // cell handled by the ordered-correction kernel
  CODE_SOLID = 8       // Prtype>0 (scalar code only)
};

// Stage constants shared by the momentum and scalar kernels.
struct StageConst {
  int stage;          // 0,1,2
  double dt;
  double rho_dt;      // RK_rho(stage)*dt   (scalar RK uses the product)
  double beta_dt;     // RK_beta(stage)*dt
  double rho, beta;   // momentum RK uses (Ustar*rho)*dt, (Ustar*beta)*dt
  double coef;        // 2*RK_alpha(stage)*dt
  double A, Ax, Ay, Az;     // -coef, -coef/dxmin, ...
  double prgx, prgy;  // uniform pressure gradient (0 = off)
  // buoyancy (BuoyancyForce, 2nd order): W += bA2 (T(k+1) + T(k)) - bATref;
  // moist: the virtual temperature T (1 + 0.61 q) in the reference's
  // two-pass order (odd k first)
  int buoy, moist;
  double bA2, bATref;
};

// Ordered per-cell corrections (built on the host in the reference's scatter
// order).  Cells are listed once; terms[ptr[c]..ptr[c+1]) apply in order.
struct TermList {
  int n;                  // number of cells
  const int32_t *cell;    // 3 per cell: i,j,k
  const int32_t *ptr;     // n+1
  const int32_t *term;    // 4 per term, meaning depends on the family
};

}  // namespace elmm
