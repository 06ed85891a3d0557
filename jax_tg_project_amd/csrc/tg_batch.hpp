// Batched ensemble: many independent hydro instances of different grid
// sizes advance together, one launch per pass for the whole ensemble.
//
// A batched launch is a 1D grid whose blocks are the concatenation of every
// instance's own grid for that pass.  Each block finds its instance with a
// binary search in a per-pass prefix table, reads the instance descriptor
// (size, viscosity, buffer pointers) and runs the size-templated pass body
// (the same code the single-instance kernels run) on its local block index.
// The block size is the largest any compiled size needs; smaller bodies
// leave the extra threads idle at the barriers.
#pragma once
#include "tg_passes.hpp"

namespace tg {

// Grid sizes compiled into the batched kernels (tile variant 0 for each).
#define TG_ENS_SIZES_LIST 16, 32, 48, 64, 96, 128
constexpr int kEnsSizes[] = {TG_ENS_SIZES_LIST};
constexpr int kNumEnsSizes = sizeof(kEnsSizes) / sizeof(kEnsSizes[0]);

enum EnsPass { EP_INV_X = 0, EP_INV_Y = 1, EP_Z = 2, EP_FWD_Y = 3, EP_FWD_X = 4, EP_COUNT = 5 };

template <typename R>
struct EnsInst {
  int n;
  R nu, ks;
  cpx<R>* u0;   // compact state (3 components)
  cpx<R>* u1;   // stage input
  cpx<R>* acc;  // r1 + 2 r2 + 2 r3
  cpx<R>* A;    // workspace (6 N K KZ)
  cpx<R>* B;    // workspace (6 N N KZ)
  const cpx<R>* tw;
};

inline __host__ __device__ constexpr int cdiv_c(int a, int b) { return (a + b - 1) / b; }

// Blocks of one instance's grid for each pass.
template <typename R, int N>
struct EnsGrid {
  using G = Geo<N>;
  using XY = CfgXY<R, N, 0>;
  using XF = CfgXF<R, N, 0>;
  static constexpr int GX_XY = cdiv_c(G::KZ, XY::L);
  static constexpr int LZ = CfgZM<R, N, ZM_HYDRO, 0>::LZ;
  static constexpr int GX_XF = cdiv_c(G::KZ, XF::LINES / 3);
  static constexpr int blocks(int pass) {
    return pass == EP_INV_X ? GX_XY * G::K
         : pass == EP_INV_Y ? GX_XY * N * 6
         : pass == EP_Z ? (N / LZ) * N
         : pass == EP_FWD_Y ? GX_XY * N * 3
         : GX_XF * G::K;
  }
};

template <typename R, int PASS, int... Ns>
struct EnsMax;
template <typename R, int PASS, int N0, int... Ns>
struct EnsMax<R, PASS, N0, Ns...> {
  static constexpr int nt0 = PASS == EP_Z ? CfgZM<R, N0, ZM_HYDRO, 0>::NT
                            : PASS == EP_FWD_X ? CfgXF<R, N0, 0>::NT : CfgXY<R, N0, 0>::NT;
  static constexpr int sm0 = PASS == EP_INV_X ? XInvBody<R, N0, XM_HYDRO, 0>::SMEM
                            : PASS == EP_Z ? ZBody<R, N0, ZM_HYDRO, 0>::SMEM
                            : PASS == EP_FWD_X ? XFBody<R, N0, FM_HYDRO_STAGE, 0>::SMEM
                            : YBody<R, N0, 1, 0>::SMEM;
  static constexpr int NT = nt0 > EnsMax<R, PASS, Ns...>::NT ? nt0 : EnsMax<R, PASS, Ns...>::NT;
  static constexpr int SMEM = sm0 > EnsMax<R, PASS, Ns...>::SMEM ? sm0 : EnsMax<R, PASS, Ns...>::SMEM;
};
template <typename R, int PASS>
struct EnsMax<R, PASS> {
  static constexpr int NT = 64;
  static constexpr int SMEM = 1;
};

template <typename R, int PASS>
using EnsShape = EnsMax<R, PASS, TG_ENS_SIZES_LIST>;

template <typename R, int N, int PASS>
__device__ __forceinline__ void ens_run(cpx<R>* sm, int lb, const EnsInst<R>& I, R dt, int stage) {
  using Gd = EnsGrid<R, N>;
  if constexpr (PASS == EP_INV_X) {
    const cpx<R>* us = stage == 0 ? I.u0 : I.u1;
    XInvBody<R, N, XM_HYDRO, 0>::run(sm, lb % Gd::GX_XY, lb / Gd::GX_XY, us, nullptr, I.A, I.tw, I.ks);
  } else if constexpr (PASS == EP_INV_Y || PASS == EP_FWD_Y) {
    const int bx = lb % Gd::GX_XY, rest = lb / Gd::GX_XY;
    if constexpr (PASS == EP_INV_Y) YBody<R, N, +1, 0>::run(sm, bx, rest % N, rest / N, I.A, I.B, I.tw);
    else YBody<R, N, -1, 0>::run(sm, bx, rest % N, rest / N, I.A, I.B, I.tw);
  } else if constexpr (PASS == EP_Z) {
    constexpr int gx = N / Gd::LZ;
    ZBody<R, N, ZM_HYDRO, 0>::run(sm, lb % gx, lb / gx, I.B, I.A, nullptr, nullptr, I.tw, R(0));
  } else {
    StageArgs<R> a{};
    a.u0 = I.u0;
    a.us = stage == 0 ? I.u0 : I.u1;
    a.unext = stage == 3 ? I.u0 : I.u1;
    a.uacc = I.acc;
    a.dt = dt;
    a.nu = I.nu;
    a.ks = I.ks;
    a.stage = stage;
    XFBody<R, N, FM_HYDRO_STAGE, 0>::run(sm, lb % Gd::GX_XF, lb / Gd::GX_XF, I.B, a, I.tw);
  }
}

template <typename R, int PASS, int N0, int... Ns>
__device__ __forceinline__ void ens_dispatch(int n, cpx<R>* sm, int lb, const EnsInst<R>& I, R dt, int stage) {
  if (n == N0) {
    ens_run<R, N0, PASS>(sm, lb, I, dt, stage);
    return;
  }
  if constexpr (sizeof...(Ns) > 0) ens_dispatch<R, PASS, Ns...>(n, sm, lb, I, dt, stage);
}

template <typename R, int PASS>
__global__ __launch_bounds__((EnsShape<R, PASS>::NT)) void kb_pass(const EnsInst<R>* __restrict__ inst,
                                                                  const int* __restrict__ prefix, int ninst,
                                                                  R dt, int stage) {
  __shared__ cpx<R> sm[EnsShape<R, PASS>::SMEM];
  const int b = blockIdx.x;
  int lo = 0, hi = ninst;  // prefix[lo] <= b < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid;
  }
  const EnsInst<R> I = inst[lo];
  ens_dispatch<R, PASS, TG_ENS_SIZES_LIST>(I.n, sm, b - prefix[lo], I, dt, stage);
}

// Host-side block counts (same formulas as EnsGrid).
template <typename R>
inline int ens_blocks(int n, int pass) {
  switch (n) {
    case 16: return EnsGrid<R, 16>::blocks(pass);
    case 32: return EnsGrid<R, 32>::blocks(pass);
    case 48: return EnsGrid<R, 48>::blocks(pass);
    case 64: return EnsGrid<R, 64>::blocks(pass);
    case 96: return EnsGrid<R, 96>::blocks(pass);
    case 128: return EnsGrid<R, 128>::blocks(pass);
  }
  return -1;
}

}  // namespace tg
