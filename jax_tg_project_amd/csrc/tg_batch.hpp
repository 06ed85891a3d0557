// Batched ensemble: many independent hydro instances of different grid
// sizes advance together, one launch per pass and grid size.
//
// Instances are grouped by grid size.  For every pass, one launch per size
// class covers all instances of that size: a 1D grid whose blocks are the
// concatenation of each instance's own grid.  Each block finds its instance
// with a binary search in the class's prefix table, reads the instance
// descriptor (viscosity, buffer pointers) and runs the size-templated pass
// body (the same code the single-instance kernels run) on its local block
// index, with that body's exact block shape.
#pragma once
#include "tg_common.hpp"
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
  static constexpr int ZE = ZwDefault<N>::E, ZWPB = ZwDefault<N>::WPB;
  static constexpr int zlines() {
    if constexpr (ZE > 0) return CfgZW<R, N, ZE, ZWPB>::LINES;
    else return LZ;
  }
  static constexpr int ZBLOCKS = N * N / zlines();
  static constexpr int blocks(int pass) {
    return pass == EP_INV_X ? GX_XY * G::K
         : pass == EP_INV_Y ? GX_XY * N * 6
         : pass == EP_Z ? ZBLOCKS
         : pass == EP_FWD_Y ? GX_XY * N * 3
         : GX_XF * G::K;
  }
};

template <typename R, int N, int PASS>
__device__ __forceinline__ void ens_run(cpx<R>* sm, int lb, const EnsInst<R>& I, R dt, int stage) {
  using Gd = EnsGrid<R, N>;
  if constexpr (PASS == EP_INV_X) {
    const cpx<R>* us = stage == 0 ? I.u0 : I.u1;
    XInvBody<R, N, XM_HYDRO5, 0>::run(sm, lb % Gd::GX_XY, lb / Gd::GX_XY, us, nullptr, I.A, I.tw, I.ks);
  } else if constexpr (PASS == EP_INV_Y) {
    const int bx = lb % Gd::GX_XY, rest = lb / Gd::GX_XY;
    YCurlBody<R, N, 0>::run(sm, bx, rest % N, rest / N, 0, I.A, I.B, I.tw, I.ks);
  } else if constexpr (PASS == EP_FWD_Y) {
    const int bx = lb % Gd::GX_XY, rest = lb / Gd::GX_XY;
    YBody<R, N, -1, 0>::run(sm, bx, rest % N, rest / N, I.A, I.B, I.tw);
  } else if constexpr (PASS == EP_Z) {
    if constexpr (Gd::ZE > 0) {
      ZWBody<R, N, Gd::ZE, Gd::ZWPB>::run(sm, lb, I.B, I.A, I.tw);
    } else {
      constexpr int gx = N / Gd::LZ;
      ZBody<R, N, ZM_HYDRO, 0>::run(sm, lb % gx, lb / gx, I.B, I.A, nullptr, nullptr, I.tw, R(0));
    }
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

template <typename R, int N, bool ZW = (ZwDefault<N>::E > 0)>
struct EnsZ {  // LDS-resident z body
  static constexpr int NT = CfgZM<R, N, ZM_HYDRO, 0>::NT;
  static constexpr int SMEM = ZBody<R, N, ZM_HYDRO, 0>::SMEM;
  static constexpr int WAVES = 1;
};
template <typename R, int N>
struct EnsZ<R, N, true> {  // register-resident z body
  static constexpr int NT = 64 * ZwDefault<N>::WPB;
  static constexpr int SMEM = ZWBody<R, N, ZwDefault<N>::E, ZwDefault<N>::WPB>::SMEM;
  static constexpr int WAVES = zw_waves_per_eu<R, ZwDefault<N>::E>();
};

template <typename R, int N, int PASS>
struct EnsBlock {
  static constexpr int NT = PASS == EP_Z ? EnsZ<R, N>::NT
                          : PASS == EP_FWD_X ? CfgXF<R, N, 0>::NT : CfgXY<R, N, 0>::NT;
  static constexpr int SMEM = PASS == EP_INV_X ? XInvBody<R, N, XM_HYDRO5, 0>::SMEM
                            : PASS == EP_INV_Y ? YCurlBody<R, N, 0>::SMEM
                            : PASS == EP_Z ? EnsZ<R, N>::SMEM
                            : PASS == EP_FWD_X ? XFBody<R, N, FM_HYDRO_STAGE, 0>::SMEM
                            : YBody<R, N, 1, 0>::SMEM;
  static constexpr int WAVES = PASS == EP_Z ? EnsZ<R, N>::WAVES : 1;
};

// One launch covers every instance of grid size N (a size class): the
// descriptors of the class are contiguous, `prefix` holds their block
// offsets for this pass.
template <typename R, int N, int PASS>
__global__ __launch_bounds__((EnsBlock<R, N, PASS>::NT))
__attribute__((amdgpu_waves_per_eu(EnsBlock<R, N, PASS>::WAVES))) void kb_pass(const EnsInst<R>* __restrict__ inst,
                                                                     const int* __restrict__ prefix, int ninst,
                                                                     R dt, int stage) {
  __shared__ cpx<R> sm[EnsBlock<R, N, PASS>::SMEM];
  const int b = blockIdx.x;
  int lo = 0, hi = ninst;  // prefix[lo] <= b < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid;
  }
  ens_run<R, N, PASS>(sm, b - prefix[lo], inst[lo], dt, stage);
}

inline constexpr bool is_ens_size(int n) {
  for (int s : kEnsSizes)
    if (s == n) return true;
  return false;
}

// Batched entry points of one (precision, N); n == 0 when N has no batched
// kernels.  `stage` launches the five passes of RK4 stage `st` for the `cnt`
// instances whose descriptors start at `inst`; dpre / hpre hold, on the
// device and on the host, EP_COUNT prefix tables of cnt + 1 block offsets.
template <typename R>
struct EnsTable {
  int n;
  int blocks[EP_COUNT];
  int (*stage)(const EnsInst<R>* inst, const int* dpre, const int* hpre, int cnt, int st, double dt,
               hipStream_t s);
};

template <typename R, int N, int PASS>
int ens_launch(const EnsInst<R>* inst, const int* dpre, const int* hpre, int cnt, int st, double dt, hipStream_t s) {
  const int nblk = hpre[PASS * (cnt + 1) + cnt];
  hipLaunchKernelGGL((kb_pass<R, N, PASS>), dim3(nblk), dim3(EnsBlock<R, N, PASS>::NT), 0, s, inst,
                     dpre + PASS * (cnt + 1), cnt, R(dt), st);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? TG_OK : fail(TG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
}

template <typename R, int N>
int ens_stage(const EnsInst<R>* inst, const int* dpre, const int* hpre, int cnt, int st, double dt, hipStream_t s) {
  TG_RET(ens_launch<R, N, EP_INV_X>(inst, dpre, hpre, cnt, st, dt, s));
  TG_RET(ens_launch<R, N, EP_INV_Y>(inst, dpre, hpre, cnt, st, dt, s));
  TG_RET(ens_launch<R, N, EP_Z>(inst, dpre, hpre, cnt, st, dt, s));
  TG_RET(ens_launch<R, N, EP_FWD_Y>(inst, dpre, hpre, cnt, st, dt, s));
  return ens_launch<R, N, EP_FWD_X>(inst, dpre, hpre, cnt, st, dt, s);
}

template <typename R, int N>
EnsTable<R> make_ens_table() {
  EnsTable<R> t{};
  if constexpr (is_ens_size(N)) {
    t.n = N;
    for (int p = 0; p < EP_COUNT; ++p) t.blocks[p] = EnsGrid<R, N>::blocks(p);
    t.stage = &ens_stage<R, N>;
  }
  return t;
}

}  // namespace tg
