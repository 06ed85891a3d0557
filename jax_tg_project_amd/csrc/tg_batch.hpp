// Batched ensemble: many independent hydro instances of different grid
// sizes advance together, one launch per pass and grid size, on the dense
// layout of tg_ens.hpp.
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
#include "tg_ens.hpp"

#include <algorithm>

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
  cpx<R>* s0;   // state (3 components, dense rows, plane pitch PA): copied in and out per call
  cpx<R>* u1;   // stage input
  cpx<R>* acc;  // r1 + 2 r2 + 2 r3
  cpx<R>* A;    // inverse-x / z / ... workspace (see EnsCfg)
  cpx<R>* B;
  const cpx<R>* tw;
};

// The descriptor as the kernels read it: the same layout, with every pointer
// in the global address space.  A pointer loaded from memory through the
// generic pointers of EnsInst is generic to the compiler, which then accesses
// it with flat instructions; those count in both the vector-memory and the
// LDS wait counters, so every LDS wait of a transform would also wait for the
// wave's outstanding loads and stores.  Descriptors only point into device
// memory.
#define TG_GLOBAL_AS __attribute__((address_space(1)))
template <typename R>
struct EnsInstG {
  int n;
  R nu, ks;
  TG_GLOBAL_AS cpx<R>* s0;
  TG_GLOBAL_AS cpx<R>* u1;
  TG_GLOBAL_AS cpx<R>* acc;
  TG_GLOBAL_AS cpx<R>* A;
  TG_GLOBAL_AS cpx<R>* B;
  TG_GLOBAL_AS const cpx<R>* tw;
};
static_assert(sizeof(EnsInstG<float>) == sizeof(EnsInst<float>) && sizeof(EnsInstG<double>) == sizeof(EnsInst<double>),
              "one descriptor layout");

// Descriptor i, read as its global-pointer view (a byte copy: the two
// structs share the layout but not the type).
template <typename R>
__device__ __forceinline__ EnsInstG<R> load_inst(const EnsInst<R>* inst, int i) {
  EnsInstG<R> g;
  __builtin_memcpy(&g, inst + i, sizeof(g));
  return g;
}

inline __host__ __device__ constexpr int cdiv_c(int a, int b) { return (a + b - 1) / b; }

// Blocks of one instance's grid for each pass.
template <typename R, int N>
struct EnsGrid {
  using C = EnsCfg<R, N>;
  static constexpr int LZ = CfgZM<R, N, ZM_HYDRO, 0>::LZ;
  static constexpr int ZE = ZwDefault<N>::E, ZWPB = ZwDefault<N>::WPB;
  static constexpr int zlines() {
    if constexpr (ZE > 0) return CfgZW<R, N, ZE, ZWPB>::LINES;
    else return LZ;
  }
  static constexpr int GROUPS = N / C::PPB;  // plane groups of one field
  // The six output fields of a plane group go to one XCD (workgroups are
  // dealt to the 8 XCDs round robin), so the components they share come
  // from that XCD's L2; possible when every instance of the class has a
  // multiple of 48 inverse-y blocks, which keeps the block offsets of the
  // instances multiples of 8.
  static constexpr bool XCD_GROUPED = (GROUPS * 6) % 48 == 0;
  static constexpr int blocks(int pass) {
    return pass == EP_INV_X ? C::TX
         : pass == EP_INV_Y ? GROUPS * 6
         : pass == EP_Z ? N * N / zlines()
         : pass == EP_FWD_Y ? GROUPS * 3
         : C::TX;
  }
};

template <typename R, int N, int PASS>
__device__ __forceinline__ void ens_run(cpx<R>* sm, int lb, const EnsInstG<R>& I, R dt, int stage) {
  using Gd = EnsGrid<R, N>;
  constexpr int KZ = Geo<N>::KZ;
  // explicit casts: the host pass has no implicit global -> generic conversion
  cpx<R>* const s0 = (cpx<R>*)I.s0;
  cpx<R>* const u1 = (cpx<R>*)I.u1;
  cpx<R>* const A = (cpx<R>*)I.A;
  cpx<R>* const B = (cpx<R>*)I.B;
  const cpx<R>* const tw = (const cpx<R>*)I.tw;
  const cpx<R>* us = stage == 0 ? s0 : u1;
  if constexpr (PASS == EP_INV_X) {
    EnsInvX<R, N>::run(sm, lb, us, A, tw, I.ks);
  } else if constexpr (PASS == EP_INV_Y) {
    int g, f;
    if constexpr (Gd::XCD_GROUPED) {
      const int xcd = lb & 7, j = lb >> 3;
      g = (j / 6) * 8 + xcd;
      f = j % 6;
    } else {
      g = lb % Gd::GROUPS;
      f = lb / Gd::GROUPS;
    }
    EnsInvY<R, N>::run(sm, g, f, A, B, tw, I.ks);
  } else if constexpr (PASS == EP_FWD_Y) {
    EnsFwdY<R, N>::run(sm, lb % Gd::GROUPS, lb / Gd::GROUPS, A, B, tw);
  } else if constexpr (PASS == EP_Z) {
    if constexpr (Gd::ZE > 0) {
      ZWBody<R, N, Gd::ZE, Gd::ZWPB, KZ>::run(sm, lb, B, A, tw);
    } else {
      constexpr int gx = N / Gd::LZ;
      ZBody<R, N, ZM_HYDRO, 0, KZ>::run(sm, lb % gx, lb / gx, B, A, nullptr, nullptr, tw, R(0));
    }
  } else {
    EnsStage<R> a;
    a.us = us;
    a.u0 = s0;
    a.unext = stage == 3 ? s0 : u1;
    a.acc = (cpx<R>*)I.acc;
    a.dt = dt;
    a.nu = I.nu;
    a.ks = I.ks;
    a.stage = stage;
    EnsFwdX<R, N>::run(sm, lb, B, a, tw);
  }
}

template <typename R, int N, bool ZW = (ZwDefault<N>::E > 0)>
struct EnsZ {  // LDS-resident z body
  static constexpr int NT = CfgZM<R, N, ZM_HYDRO, 0>::NT;
  static constexpr int SMEM = ZBody<R, N, ZM_HYDRO, 0, Geo<N>::KZ>::SMEM;
  static constexpr int WAVES = 1;
};
template <typename R, int N>
struct EnsZ<R, N, true> {  // register-resident z body
  static constexpr int NT = 64 * EnsGrid<R, N>::ZWPB;
  static constexpr int SMEM = ZWBody<R, N, ZwDefault<N>::E, EnsGrid<R, N>::ZWPB, Geo<N>::KZ>::SMEM;
  static constexpr int WAVES = zw_waves_per_eu<R, ZwDefault<N>::E>();
};

template <typename R, int N, int PASS>
struct EnsBlock {
  using C = EnsCfg<R, N>;
  static constexpr int NT = PASS == EP_Z ? EnsZ<R, N>::NT
                          : (PASS == EP_INV_Y || PASS == EP_FWD_Y) ? C::NTY : C::NTX;
  static constexpr int SMEM = PASS == EP_INV_X ? EnsInvX<R, N>::SMEM
                            : PASS == EP_INV_Y ? EnsInvY<R, N>::SMEM
                            : PASS == EP_Z ? EnsZ<R, N>::SMEM
                            : PASS == EP_FWD_X ? EnsFwdX<R, N>::SMEM
                            : EnsFwdY<R, N>::SMEM;
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
  ens_run<R, N, PASS>(sm, b - prefix[lo], load_inst(inst, lo), dt, stage);
}

// Sizes whose batched x bodies (256 threads) and z bodies (one wavefront)
// share their block shapes, so one launch per pass can cover instances of
// all of them: a block finds its instance as in kb_pass and branches (block-
// uniformly) to that instance's size body.  rev_n > 0 (the launch's block
// count) runs the blocks in reverse order.  The y bodies differ in threads
// (PPB planes x KZ x N/E); the mixed y launch takes the largest, and the
// threads a smaller body does not use stay inactive through its transforms
// while still reaching every barrier.
#define TG_MIX_SIZES_LIST 96, 64, 48, 32
constexpr int kMixSizes[] = {TG_MIX_SIZES_LIST};

inline constexpr bool is_mix_size(int n) {
  for (int s : kMixSizes)
    if (s == n) return true;
  return false;
}

template <typename R, int PASS>
struct MixBlock {
  template <int... Ns>
  static constexpr int max_nt() { int m = 0; ((m = EnsBlock<R, Ns, PASS>::NT > m ? EnsBlock<R, Ns, PASS>::NT : m), ...); return m; }
  template <int... Ns>
  static constexpr int min_nt() { int m = 1 << 20; ((m = EnsBlock<R, Ns, PASS>::NT < m ? EnsBlock<R, Ns, PASS>::NT : m), ...); return m; }
  template <int... Ns>
  static constexpr int max_smem() { int m = 0; ((m = EnsBlock<R, Ns, PASS>::SMEM > m ? EnsBlock<R, Ns, PASS>::SMEM : m), ...); return m; }
  template <int... Ns>
  static constexpr int min_waves() { int m = 16; ((m = EnsBlock<R, Ns, PASS>::WAVES < m ? EnsBlock<R, Ns, PASS>::WAVES : m), ...); return m; }
  static constexpr int NT = max_nt<TG_MIX_SIZES_LIST>();
  static constexpr int SMEM = max_smem<TG_MIX_SIZES_LIST>();
  static constexpr int WAVES = min_waves<TG_MIX_SIZES_LIST>();
  // the x and z bodies index lanes by their own block shape: it must match
  static_assert(PASS == EP_INV_Y || PASS == EP_FWD_Y || min_nt<TG_MIX_SIZES_LIST>() == NT,
                "mixed x / z launches need one block shape");
};

template <typename R, int PASS>
__global__ __launch_bounds__((MixBlock<R, PASS>::NT))
__attribute__((amdgpu_waves_per_eu(MixBlock<R, PASS>::WAVES))) void kb_mixed(const EnsInst<R>* __restrict__ inst,
                                                                      const int* __restrict__ prefix, int ninst,
                                                                      R dt, int stage, int rev_n) {
  __shared__ cpx<R> sm[MixBlock<R, PASS>::SMEM];
  const int b = rev_n ? rev_n - 1 - int(blockIdx.x) : int(blockIdx.x);
  int lo = 0, hi = ninst;  // prefix[lo] <= b < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid;
  }
  const int lb = b - prefix[lo];
  const EnsInstG<R> I = load_inst(inst, lo);
  switch (I.n) {
    case 96: ens_run<R, 96, PASS>(sm, lb, I, dt, stage); break;
    case 64: ens_run<R, 64, PASS>(sm, lb, I, dt, stage); break;
    case 48: ens_run<R, 48, PASS>(sm, lb, I, dt, stage); break;
    default: ens_run<R, 32, PASS>(sm, lb, I, dt, stage); break;
  }
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
  int pa;                    // plane pitch of the state arrays (complex elements)
  long long a_elems, b_elems;  // workspace arrays A, B (complex elements)
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

// The stage of the mixed launches: `cnt` descriptors from `inst` (every one
// of a kMixSizes size), dpre / hpre as for EnsTable::stage.
template <typename R>
struct EnsMixTable {
  int (*stage)(const EnsInst<R>* inst, const int* dpre, const int* hpre, int cnt, int st, double dt,
               hipStream_t s);
};

template <typename R, int PASS>
int mix_launch(const EnsInst<R>* inst, const int* dpre, const int* hpre, int cnt, int st, double dt, hipStream_t s) {
  const int nblk = hpre[PASS * (cnt + 1) + cnt];
  // Every other launch of a step (20 per step, so also across steps) runs its
  // blocks in reverse order: a pass then first reads what the previous pass
  // wrote last, which is still in the last-level cache when the previous
  // pass's output exceeds it.
  const bool rev = (st * EP_COUNT + PASS) & 1;
  hipLaunchKernelGGL((kb_mixed<R, PASS>), dim3(nblk), dim3(MixBlock<R, PASS>::NT), 0, s, inst,
                     dpre + PASS * (cnt + 1), cnt, R(dt), st, rev ? nblk : 0);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? TG_OK : fail(TG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
}

template <typename R>
int mix_stage(const EnsInst<R>* inst, const int* dpre, const int* hpre, int cnt, int st, double dt, hipStream_t s) {
  TG_RET(mix_launch<R, EP_INV_X>(inst, dpre, hpre, cnt, st, dt, s));
  TG_RET(mix_launch<R, EP_INV_Y>(inst, dpre, hpre, cnt, st, dt, s));
  TG_RET(mix_launch<R, EP_Z>(inst, dpre, hpre, cnt, st, dt, s));
  TG_RET(mix_launch<R, EP_FWD_Y>(inst, dpre, hpre, cnt, st, dt, s));
  return mix_launch<R, EP_FWD_X>(inst, dpre, hpre, cnt, st, dt, s);
}

template <typename R>
EnsMixTable<R> make_mix_table() {
  EnsMixTable<R> t{};
  t.stage = &mix_stage<R>;
  return t;
}

template <typename R, int N>
EnsTable<R> make_ens_table() {
  EnsTable<R> t{};
  if constexpr (is_ens_size(N)) {
    using C = EnsCfg<R, N>;
    t.n = N;
    for (int p = 0; p < EP_COUNT; ++p) t.blocks[p] = EnsGrid<R, N>::blocks(p);
    t.pa = C::PA;
    t.a_elems = std::max(5LL * N * C::PA, 3LL * N * C::PB);
    t.b_elems = std::max(6LL * N * C::PB, 3LL * N * C::PA);
    t.stage = &ens_stage<R, N>;
  }
  return t;
}

}  // namespace tg
