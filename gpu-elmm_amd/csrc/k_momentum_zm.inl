// k_momentum_zm.inl - the fused z-marching momentum kernel and its launcher,
// included by k_momentum_zm.hip (FP64 instantiation, every multiply and add
// rounded separately) and k_momentum_zm_f32.hip (FP32 instantiation of the
// precision form, built with multiply-add contraction).
//
// Fused momentum right-hand side of all three velocity
// components on uniform grids, marching in z.
//
// Per cell and component the arithmetic is that of CD4divU/V/W
// (src/advection.f90:334-834), MomentumDiffusion_nobranch_2ord
// (src/momentum_diffusion.f90:169-288), the RK3 combination of Convection
// (src/dynamics.f90:266-295, 450-478), PressureGrad (src/dynamics.f90:
// 156-194) and OtherTerms' U2 = U2 + U (src/tsteps.f90:350-355), term by
// term in the same order as the per-component kernels, so the FP64
// instantiation gives identical results.
//
// Data movement (gfx950): a block owns a TX x TY tile of columns and walks
// k upward.  In-plane neighbours come from LDS rings of planes with a halo of
// 3: U and V at levels k-1..k+2, W and the viscosity at k-1..k+1.  Column
// neighbours (k-3..k+3 of U, V, W) and the x/y-interpolated W of the cross
// terms (levels k-2..k+1) live in registers.  The next level's planes are
// fetched into registers before the current level is computed, then written
// to the ring slot that the current level no longer needs.  Each field is
// thus read from HBM about once (plus the tile halo) instead of once per
// stencil tap.
#include <cstdlib>
#include <string>

#include "kernels.hpp"

namespace elmm {

namespace {

constexpr int HALO = 3;
// ring slots: U 0-3, V 4-7, W 8-11, viscosity 12-15
constexpr int SU = 0, SV = 4, SW = 8, SN = 12, NSLOT = 16;

// tile geometry of a TX x TY block
template <int TX, int TY>
struct Tile {
  static constexpr int PW = TX + 2 * HALO, PH = TY + 2 * HALO, PS = PW * PH;
  static constexpr int NTH = TX * TY;
  static constexpr int PLOAD = (PS + NTH - 1) / NTH;
};

template <typename T>
__device__ inline T m4(T a, T b, T c, T d) {
  return ((-a + 9 * b) + 9 * c) + -d;
}

template <typename T>
struct MomZArgs {
  const T *U, *V, *W, *Visc, *Pr;
  const T *Temp;   // temperature (buoyancy), stage start
  const T *Moist;  // moisture (moist buoyancy), stage start
  T *Us[3];
  T *U2[3];
  const uint8_t *code[3];
};

// d of one direction, 4th order, from the UV1/UV3 products
template <typename T>
__device__ inline T d4(T uv1b, T uv1a, T uv3b, T uv3a) {
  T d = 9 * (uv1b - uv1a);
  if constexpr (sizeof(T) == 4)
    return d - (uv3b - uv3a) * recip3<T>();
  else
    return d - div_rn_nz<T>(uv3b - uv3a, (T)3, recip3<T>());
}

template <typename T, typename FF>
__device__ inline T self4(FF F) {
  T Fi, uv1b, uv1a, uv3b, uv3a;
  Fi = m4<T>(F(-1), F(0), F(1), F(2));
  uv1b = Fi * (F(0) + F(1));
  Fi = m4<T>(F(-2), F(-1), F(0), F(1));
  uv1a = Fi * (F(-1) + F(0));
  Fi = m4<T>(F(0), F(1), F(2), F(3));
  uv3b = Fi * (F(0) + F(3));
  Fi = m4<T>(F(-3), F(-2), F(-1), F(0));
  uv3a = Fi * (F(-3) + F(0));
  return d4<T>(uv1b, uv1a, uv3b, uv3a);
}

template <typename T, typename FF, typename AF>
__device__ inline T cross4(FF F, AF AI) {
  const T uv1b = AI(0) * (F(0) + F(1));
  const T uv1a = AI(-1) * (F(-1) + F(0));
  const T uv3b = AI(1) * (F(0) + F(3));
  const T uv3a = AI(-2) * (F(-3) + F(0));
  return d4<T>(uv1b, uv1a, uv3b, uv3a);
}

template <typename T, typename FF>
__device__ inline T self2(FF F) {
  return 64 * ((F(1) + F(0)) * (F(1) + F(0)) - (F(0) + F(-1)) * (F(0) + F(-1)));
}

template <typename T, typename FF>
__device__ inline T cross2(FF F, T ac, T acm) {
  return 64 * ((F(1) + F(0)) * ac - (F(0) + F(-1)) * acm);
}

template <typename T>
__device__ inline void finish_cell(const StageConst &sc, int C, T w, T f0, T dpr, T dprz_var, bool varz, T us_prev,
                                   T *__restrict__ Us, T *__restrict__ U2, long p, bool corr) {
  T u2 = sc.stage > 0 ? us_prev * (T)sc.rho * (T)sc.dt : (T)0;
  u2 = u2 + w * (T)sc.beta * (T)sc.dt;
  if (C == 0 && sc.prgx != 0) u2 = u2 + (T)sc.A * (T)sc.prgx;
  if (C == 1 && sc.prgy != 0) u2 = u2 + (T)sc.A * (T)sc.prgy;
  if (C == 0) u2 = u2 + (T)sc.Ax * dpr;
  if (C == 1) u2 = u2 + (T)sc.Ay * dpr;
  if (C == 2) {
    if (varz)
      u2 = u2 - (T)sc.coef * dprz_var;
    else
      u2 = u2 + (T)sc.Az * dpr;
  }
  u2 = u2 + f0;
  // a cell with ordered corrections hands over to the correction kernel: the
  // bulk value in Ustar, the previous Ustar in U2
  Us[p] = w;
  U2[p] = corr ? us_prev : u2;
}

// Quotients of the advective terms.  FP64 (the reference's arithmetic): each
// direction's difference divided by the spacing, correctly rounded (with
// FAST the reciprocals satisfy div_rn's condition, checked on the host, so
// there is no fallback branch), summed, then scaled by -1/256.  FP32 (the
// precision form): one product per direction with -1/(256 d) rounded once,
// and the 1/4 of the viscosity averages folded into 1/d^2.
//
// Ring slots hold level L at (L - kbeg) mod 4 for every field and the z loop
// is unrolled by 4, so every slot, and with it every LDS address, is a
// constant offset from the thread's own base.
// (launch bounds, in waves per SIMD: 4-byte fields 4 for 512-thread blocks
// (128 VGPRs, two blocks per CU) and 3 for 256-thread blocks (three blocks,
// as many as their LDS allows); 8-byte fields 2)
// BUOY: 0 no buoyancy, 1 the temperature's, 2 with the virtual temperature
// (moisture); a compile-time choice, so the case without buoyancy keeps the
// registers of the temperature and moisture columns
template <typename T, int TX, int TY, bool FAST, int BUOY>
__global__ void __launch_bounds__(TX *TY, sizeof(T) == 4 ? (TX * TY == 512 ? 4 : 3) : 2)
    k_momentum_zm(Geo g, MomZArgs<T> a, StageConst sc, int kchunk) {
  constexpr bool F32 = sizeof(T) == 4;
  constexpr int PW = Tile<TX, TY>::PW, PS = Tile<TX, TY>::PS, NTH = Tile<TX, TY>::NTH;
  constexpr int PLOAD = Tile<TX, TY>::PLOAD;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int i0 = 1 + blockIdx.x * TX, j0 = 1 + blockIdx.y * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int kbeg = 1 + blockIdx.z * kchunk;
  const int kend = min(kbeg + kchunk - 1, g.nz);
  const long sz = g.sz;
  const int lx = tx + HALO, ly = ty + HALO;

  // plane element offsets at level 0 (clamped into the allocation) and LDS
  // positions.  A thread past the end of the plane loads the plane's last
  // element and stores it where its owner does (the same value), so plane
  // loads and stores carry no branches.
  long poff[PLOAD];
  int pidx[PLOAD];
#pragma unroll
  for (int r = 0; r < PLOAD; r++) {
    const int ec = min(tid + r * NTH, PS - 1);
    pidx[r] = ec;
    const int ex = ec % PW, ey = ec / PW;
    const int gi = min(i0 - HALO + ex, g.nx + 3), gj = min(j0 - HALO + ey, g.ny + 3);
    poff[r] = gix(g, gi, gj, 0);
  }
  const int ic = min(i, g.nx), jc = min(j, g.ny);
  const long pc = gix(g, ic, jc, 0);

  // slot of the level kbeg + rel
  auto slot = [](int base, int rel) { return base + ((rel + 8) & 3); };
  auto at = [&](int s, int dx, int dy) -> T & { return lds[s * PS + (ly + dy) * PW + lx + dx]; };
  auto load_plane = [&](const T *__restrict__ F, int L, int s) {
#pragma unroll
    for (int r = 0; r < PLOAD; r++) lds[s * PS + pidx[r]] = F[poff[r] + (long)L * sz];
  };

  // ---- prologue: rings for k = kbeg
  T cu[7], cv[7], cw[7], wx[4], wy[4];
#pragma unroll
  for (int q = 0; q < 7; q++) {
    const long o = pc + (long)(kbeg - 3 + q) * sz;
    cu[q] = a.U[o];
    cv[q] = a.V[o];
    cw[q] = a.W[o];
  }
#pragma unroll
  for (int rel = -1; rel <= 2; rel++) {
    load_plane(a.U, kbeg + rel, slot(SU, rel));
    load_plane(a.V, kbeg + rel, slot(SV, rel));
  }
#pragma unroll
  for (int rel = -1; rel <= 1; rel++) {
    load_plane(a.W, kbeg + rel, slot(SW, rel));
    load_plane(a.Visc, kbeg + rel, slot(SN, rel));
  }
  {
    const long o = pc + (long)(kbeg - 2) * sz;
    wx[0] = m4<T>(a.W[o - 1], a.W[o], a.W[o + 1], a.W[o + 2]);
    wy[0] = m4<T>(a.W[o - g.sy], a.W[o], a.W[o + g.sy], a.W[o + 2 * g.sy]);
  }
  __syncthreads();
#pragma unroll
  for (int q = 1; q < 4; q++) {
    const int s = slot(SW, q - 2);
    wx[q] = m4<T>(at(s, -1, 0), at(s, 0, 0), at(s, 1, 0), at(s, 2, 0));
    wy[q] = m4<T>(at(s, 0, -1), at(s, 0, 0), at(s, 0, 1), at(s, 0, 2));
  }

  const T q4 = (T)0.25, m256 = (T)(-1.0 / 256);
  const T rx = (T)1.0 / ((T)g.dxmin * (T)g.dxmin), ry = (T)1.0 / ((T)g.dymin * (T)g.dymin),
          rz = (T)1.0 / ((T)g.dzmin * (T)g.dzmin);
  // factors of the 4-term viscosity averages (FP32: the 1/4 folded in)
  const T rxq = F32 ? rx * q4 : rx, ryq = F32 ? ry * q4 : ry, rzq = F32 ? rz * q4 : rz;
  auto qa = [&](T s4) -> T {
    if constexpr (F32)
      return s4;
    else
      return q4 * s4;
  };
  const T ix = (T)g.dxmin, iy = (T)g.dymin, iz = (T)g.dzmin;
  const T rix = F32 ? (T)g.rdxf : (T)g.rdx, riy = F32 ? (T)g.rdyf : (T)g.rdy, riz = F32 ? (T)g.rdzf : (T)g.rdz;
  const T cx = (T)(-1.0 / (256.0 * g.dxmin)), cy = (T)(-1.0 / (256.0 * g.dymin)), cz = (T)(-1.0 / (256.0 * g.dzmin));
  auto dv = [](T x, T d, T r) {
    if constexpr (FAST)
      return div_rn_nz<T>(x, d, r);
    else
      return div_rn<T>(x, d, r);
  };
  // advective term from the three directions' differences
  auto adv = [&](T dX, T dY, T dZ) -> T {
    if constexpr (F32) {
      return (dX * cx + dY * cy) + dZ * cz;
    } else {
      T acc = 0;
      acc = acc + dv(dX, ix, rix);
      acc = acc + dv(dY, iy, riy);
      acc = acc + dv(dZ, iz, riz);
      return acc * m256;
    }
  };
  const T wnone = (T)0 * m256;  // the advective term of a cell without one
  const bool in_xy = i <= g.nx && j <= g.ny;
  // per-cell operands of the level being computed (Pr and its +x, +y, +z
  // neighbours, the three codes, the three previous Ustar values), fetched
  // one level ahead like the planes
  T cpr = a.Pr[pc + (long)kbeg * sz], cprx = a.Pr[pc + (long)kbeg * sz + 1],
    cpry = a.Pr[pc + (long)kbeg * sz + g.sy], cprz = a.Pr[pc + (long)(kbeg + 1) * sz];
  // temperature at k and k+1 of the column (buoyancy only)
  T ct0 = 0, ct1 = 0;
  if constexpr (BUOY >= 1) {
    ct0 = a.Temp[pc + (long)kbeg * sz];
    ct1 = a.Temp[pc + (long)(kbeg + 1) * sz];
  }
  // moisture at k and k+1 (moist buoyancy)
  T cq0 = 0, cq1 = 0;
  if constexpr (BUOY == 2) {
    cq0 = a.Moist[pc + (long)kbeg * sz];
    cq1 = a.Moist[pc + (long)(kbeg + 1) * sz];
  }
  uint8_t ccd[3];
  T cus[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    ccd[c] = a.code[c][pc + (long)kbeg * sz];
    cus[c] = sc.stage > 0 ? a.Us[c][pc + (long)kbeg * sz] : (T)0;
  }
  const int kcol = g.nz + 3;  // top level of the allocation

  for (int k0 = kbeg; k0 <= kend; k0 += 4) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int k = k0 + u;
      if (k > kend) break;
      // prefetch of the next level's planes and column values (after the
      // chunk's last level they are loaded and not used; levels stay inside
      // the allocation)
      T pu[PLOAD], pv[PLOAD], pw[PLOAD], pn[PLOAD];
#pragma unroll
      for (int r = 0; r < PLOAD; r++) {
        pu[r] = a.U[poff[r] + (long)(k + 3) * sz];
        pv[r] = a.V[poff[r] + (long)(k + 3) * sz];
        pw[r] = a.W[poff[r] + (long)(k + 2) * sz];
        pn[r] = a.Visc[poff[r] + (long)(k + 2) * sz];
      }
      T ou, ov, ow;
      {
        const long o = pc + (long)min(k + 4, kcol) * sz;
        ou = a.U[o];
        ov = a.V[o];
        ow = a.W[o];
      }
      T nt = 0;
      if constexpr (BUOY >= 1) nt = a.Temp[pc + (long)(k + 2) * sz];
      T nq = 0;
      if constexpr (BUOY == 2) nq = a.Moist[pc + (long)(k + 2) * sz];
      const long q = pc + (long)(k + 1) * sz;
      const T nprx = a.Pr[q + 1], npry = a.Pr[q + g.sy], nprz = a.Pr[q + sz];
      uint8_t ncd[3];
      T nus[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        ncd[c] = a.code[c][q];
        nus[c] = sc.stage > 0 ? a.Us[c][q] : (T)0;
      }

      if (in_xy) {
        const long p = gix(g, i, j, k);
        const int u0 = slot(SU, u), um = slot(SU, u - 1), u1 = slot(SU, u + 1), u2s = slot(SU, u + 2);
        const int v0 = slot(SV, u), vm = slot(SV, u - 1), v1 = slot(SV, u + 1), v2s = slot(SV, u + 2);
        const int w0 = slot(SW, u), wm = slot(SW, u - 1);
        const int n0 = slot(SN, u), nm = slot(SN, u - 1), n1 = slot(SN, u + 1);
        const T pr0 = cpr;

        // ---------------- U
        if (i <= g.n[0][0] && j <= g.n[0][1] && k <= g.n[0][2]) {
          const uint8_t cd = ccd[0];
          const bool corr = (cd & CODE_CORRECTED) != 0;
          const int cls = cd & CODE_ADV_MASK;
          T w = wnone;
          if (cls != 2) {
            auto Fx = [&](int o) { return at(u0, o, 0); };
            auto Fy = [&](int o) { return at(u0, 0, o); };
            auto Fz = [&](int o) { return cu[3 + o]; };
            T dX, dY, dZ;
            if (cls == 0) {
              dX = self4<T>(Fx);
              dY = cross4<T>(Fy, [&](int o) { return m4<T>(at(v0, -1, o), at(v0, 0, o), at(v0, 1, o), at(v0, 2, o)); });
              dZ = cross4<T>(Fz, [&](int o) { return wx[2 + o]; });
            } else {
              dX = self2<T>(Fx);
              dY = cross2<T>(Fy, at(v0, 1, 0) + at(v0, 0, 0), at(v0, 1, -1) + at(v0, 0, -1));
              dZ = cross2<T>(Fz, at(w0, 1, 0) + at(w0, 0, 0), at(wm, 1, 0) + at(wm, 0, 0));
            }
            w = adv(dX, dY, dZ);
          }
          const T f0 = at(u0, 0, 0);
          w = w + at(n0, 1, 0) * (at(u0, 1, 0) - f0) * rx;
          w = w - at(n0, 0, 0) * (f0 - at(u0, -1, 0)) * rx;
          w = w + qa(at(n0, 1, 1) + at(n0, 1, 0) + at(n0, 0, 1) + at(n0, 0, 0)) * (at(u0, 0, 1) - f0) * ryq;
          w = w - qa(at(n0, 1, 0) + at(n0, 1, -1) + at(n0, 0, 0) + at(n0, 0, -1)) * (f0 - at(u0, 0, -1)) * ryq;
          w = w + qa(at(n1, 1, 0) + at(n0, 1, 0) + at(n1, 0, 0) + at(n0, 0, 0)) * (cu[4] - cu[3]) * rzq;
          w = w - qa(at(n0, 1, 0) + at(nm, 1, 0) + at(n0, 0, 0) + at(nm, 0, 0)) * (cu[3] - cu[2]) * rzq;
          finish_cell<T>(sc, 0, w, f0, cprx - pr0, (T)0, false, cus[0], a.Us[0], a.U2[0], p, corr);
        }
        // ---------------- V
        if (i <= g.n[1][0] && j <= g.n[1][1] && k <= g.n[1][2]) {
          const uint8_t cd = ccd[1];
          const bool corr = (cd & CODE_CORRECTED) != 0;
          const int cls = cd & CODE_ADV_MASK;
          T w = wnone;
          if (cls != 2) {
            auto Fx = [&](int o) { return at(v0, o, 0); };
            auto Fy = [&](int o) { return at(v0, 0, o); };
            auto Fz = [&](int o) { return cv[3 + o]; };
            T dX, dY, dZ;
            if (cls == 0) {
              dX = cross4<T>(Fx, [&](int o) { return m4<T>(at(u0, o, -1), at(u0, o, 0), at(u0, o, 1), at(u0, o, 2)); });
              dY = self4<T>(Fy);
              dZ = cross4<T>(Fz, [&](int o) { return wy[2 + o]; });
            } else {
              dX = cross2<T>(Fx, at(u0, 0, 1) + at(u0, 0, 0), at(u0, -1, 1) + at(u0, -1, 0));
              dY = self2<T>(Fy);
              dZ = cross2<T>(Fz, at(w0, 0, 1) + at(w0, 0, 0), at(wm, 0, 1) + at(wm, 0, 0));
            }
            w = adv(dX, dY, dZ);
          }
          const T f0 = at(v0, 0, 0);
          w = w + qa(at(n0, 1, 1) + at(n0, 1, 0) + at(n0, 0, 1) + at(n0, 0, 0)) * (at(v0, 1, 0) - f0) * rxq;
          w = w - qa(at(n0, 0, 1) + at(n0, 0, 0) + at(n0, -1, 1) + at(n0, -1, 0)) * (f0 - at(v0, -1, 0)) * rxq;
          w = w + at(n0, 0, 1) * (at(v0, 0, 1) - f0) * ry;
          w = w - at(n0, 0, 0) * (f0 - at(v0, 0, -1)) * ry;
          w = w + qa(at(n1, 0, 1) + at(n0, 0, 1) + at(n1, 0, 0) + at(n0, 0, 0)) * (cv[4] - cv[3]) * rzq;
          w = w - qa(at(n0, 0, 1) + at(nm, 0, 1) + at(n0, 0, 0) + at(nm, 0, 0)) * (cv[3] - cv[2]) * rzq;
          finish_cell<T>(sc, 1, w, f0, cpry - pr0, (T)0, false, cus[1], a.Us[1], a.U2[1], p, corr);
        }
        // ---------------- W
        if (i <= g.n[2][0] && j <= g.n[2][1] && k <= g.n[2][2]) {
          const uint8_t cd = ccd[2];
          const bool corr = (cd & CODE_CORRECTED) != 0;
          const int cls = cd & CODE_ADV_MASK;
          T w = wnone;
          if (cls != 2) {
            auto Fx = [&](int o) { return at(w0, o, 0); };
            auto Fy = [&](int o) { return at(w0, 0, o); };
            auto Fz = [&](int o) { return cw[3 + o]; };
            T dX, dY, dZ;
            if (cls == 0) {
              dX = cross4<T>(Fx, [&](int o) { return m4<T>(at(um, o, 0), at(u0, o, 0), at(u1, o, 0), at(u2s, o, 0)); });
              dY = cross4<T>(Fy, [&](int o) { return m4<T>(at(vm, 0, o), at(v0, 0, o), at(v1, 0, o), at(v2s, 0, o)); });
              dZ = self4<T>(Fz);
            } else {
              dX = cross2<T>(Fx, at(u1, 0, 0) + at(u0, 0, 0), at(u1, -1, 0) + at(u0, -1, 0));
              dY = cross2<T>(Fy, at(v1, 0, 0) + at(v0, 0, 0), at(v1, 0, -1) + at(v0, 0, -1));
              dZ = self2<T>(Fz);
            }
            w = adv(dX, dY, dZ);
          }
          if constexpr (BUOY >= 1) {
            if constexpr (BUOY == 2) {
              // W(k) gets Tv(k) in the pass over odd k and Tv(k+1) in the
              // pass over even k (and the reverse for even k)
              const T tv0 = ct0 * ((T)1 + (T)0.61 * cq0), tv1 = ct1 * ((T)1 + (T)0.61 * cq1);
              if (k & 1)
                w = ((w + (T)sc.bA2 * tv0) - (T)sc.bATref) + (T)sc.bA2 * tv1;
              else
                w = ((w + (T)sc.bA2 * tv1) + (T)sc.bA2 * tv0) - (T)sc.bATref;
            } else {
              w = w + (T)sc.bA2 * (ct1 + ct0) - (T)sc.bATref;
            }
          }
          const T f0 = cw[3];
          w = w + qa(at(n1, 1, 0) + at(n0, 1, 0) + at(n1, 0, 0) + at(n0, 0, 0)) * (at(w0, 1, 0) - f0) * rxq;
          w = w - qa(at(n1, 0, 0) + at(n0, 0, 0) + at(n1, -1, 0) + at(n0, -1, 0)) * (f0 - at(w0, -1, 0)) * rxq;
          w = w + qa(at(n1, 0, 1) + at(n1, 0, 0) + at(n0, 0, 1) + at(n0, 0, 0)) * (at(w0, 0, 1) - f0) * ryq;
          w = w - qa(at(n1, 0, 0) + at(n1, 0, -1) + at(n0, 0, 0) + at(n0, 0, -1)) * (f0 - at(w0, 0, -1)) * ryq;
          w = w + at(n1, 0, 0) * (cw[4] - f0) * rz;
          w = w - at(n0, 0, 0) * (f0 - cw[2]) * rz;
          finish_cell<T>(sc, 2, w, f0, cprz - pr0, (T)0, false, cus[2], a.Us[2], a.U2[2], p, corr);
        }
      }

      // U and V of level k+3 replace level k-1, read above; W and the
      // viscosity of level k+2 replace level k-2
      __syncthreads();
#pragma unroll
      for (int r = 0; r < PLOAD; r++) {
        lds[slot(SU, u + 3) * PS + pidx[r]] = pu[r];
        lds[slot(SV, u + 3) * PS + pidx[r]] = pv[r];
        lds[slot(SW, u + 2) * PS + pidx[r]] = pw[r];
        lds[slot(SN, u + 2) * PS + pidx[r]] = pn[r];
      }
      __syncthreads();
#pragma unroll
      for (int q2 = 0; q2 < 6; q2++) {
        cu[q2] = cu[q2 + 1];
        cv[q2] = cv[q2 + 1];
        cw[q2] = cw[q2 + 1];
      }
      cu[6] = ou;
      cv[6] = ov;
      cw[6] = ow;
      ct0 = ct1;
      ct1 = nt;
      cq0 = cq1;
      cq1 = nq;
      cpr = cprz;
      cprx = nprx;
      cpry = npry;
      cprz = nprz;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        ccd[c] = ncd[c];
        cus[c] = nus[c];
      }
#pragma unroll
      for (int q2 = 0; q2 < 3; q2++) {
        wx[q2] = wx[q2 + 1];
        wy[q2] = wy[q2 + 1];
      }
      const int s = slot(SW, u + 2);
      wx[3] = m4<T>(at(s, -1, 0), at(s, 0, 0), at(s, 1, 0), at(s, 2, 0));
      wy[3] = m4<T>(at(s, 0, -1), at(s, 0, 0), at(s, 0, 1), at(s, 0, 2));
    }
  }
}

}  // namespace



namespace {
template <typename T, int TX, int TY, bool FAST, int BUOY>
void launch_tile_b(const Geo &g, const MomZArgs<T> &a, StageConst sc, int kchunk, hipStream_t st) {
  const size_t lds = (size_t)NSLOT * Tile<TX, TY>::PS * sizeof(T);
  if (lds > 65536) (void)lds_limit_raise((const void *)k_momentum_zm<T, TX, TY, FAST, BUOY>, lds);
  dim3 b(TX * TY), gr((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, (g.nz + kchunk - 1) / kchunk);
  hipLaunchKernelGGL((k_momentum_zm<T, TX, TY, FAST, BUOY>), gr, b, lds, st, g, a, sc, kchunk);
}
template <typename T, int TX, int TY, bool FAST>
void launch_tile_f(const Geo &g, const MomZArgs<T> &a, StageConst sc, int kchunk, hipStream_t st) {
  if (!sc.buoy)
    launch_tile_b<T, TX, TY, FAST, 0>(g, a, sc, kchunk, st);
  else if (!sc.moist)
    launch_tile_b<T, TX, TY, FAST, 1>(g, a, sc, kchunk, st);
  else
    launch_tile_b<T, TX, TY, FAST, 2>(g, a, sc, kchunk, st);
}
template <typename T, int TX, int TY>
void launch_tile(const Geo &g, const MomZArgs<T> &a, StageConst sc, int kchunk, hipStream_t st) {
  // FP32 multiplies by reciprocals whatever the spacing
  if constexpr (sizeof(T) == 4)
    launch_tile_f<T, TX, TY, true>(g, a, sc, kchunk, st);
  else if (g.rdx != 0 && g.rdy != 0 && g.rdz != 0)
    launch_tile_f<T, TX, TY, true>(g, a, sc, kchunk, st);
  else
    launch_tile_f<T, TX, TY, false>(g, a, sc, kchunk, st);
}
}  // namespace

// Tile ids: kMomTile64x4 (default; measured 5 % faster than 32x8 in FP32 on
// 512x512x128, equal in FP64), kMomTile32x8, kMomTile64x8 (4-byte fields
// only: its FP64 rings take 125 KiB of LDS, one block per CU; an FP64 launch
// asking for it runs 32x8).  The engine's autotuner times them per grid.
template <typename T>
void launch_momentum_zm(const Geo &g, const T *U, const T *V, const T *W, const T *Visc, const T *Pr, const T *Temp,
                        const T *Moist, T *const Us[3], T *const U2[3], const uint8_t *const code[3], StageConst sc,
                        int kchunk, int tile, hipStream_t st) {
  MomZArgs<T> a;
  a.Temp = Temp;
  a.Moist = Moist;
  a.U = U;
  a.V = V;
  a.W = W;
  a.Visc = Visc;
  a.Pr = Pr;
  for (int c = 0; c < 3; c++) {
    a.Us[c] = Us[c];
    a.U2[c] = U2[c];
    a.code[c] = code[c];
  }
  if (kchunk <= 0) kchunk = g.nz;
  switch (tile) {
    case kMomTile64x4: launch_tile<T, 64, 4>(g, a, sc, kchunk, st); break;
    case kMomTile64x8:
      if constexpr (sizeof(T) == 4) {
        launch_tile<T, 64, 8>(g, a, sc, kchunk, st);
        break;
      }
      launch_tile<T, 32, 8>(g, a, sc, kchunk, st);
      break;
    default: launch_tile<T, 32, 8>(g, a, sc, kchunk, st);
  }
}

}  // namespace elmm
