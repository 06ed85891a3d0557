// libtgdns.so: host side of the C ABI declared in include/tgdns.h.
// Plans own twiddle tables; callers own every other buffer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/tgdns.h"
#include "tg_passes.hpp"
#include "tg_batch.hpp"

using namespace tg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define TG_HIP(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(TG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

#define TG_LAUNCHED()                                                                     \
  do {                                                                                    \
    hipError_t e_ = hipGetLastError();                                                    \
    if (e_ != hipSuccess)                                                                 \
      return fail(TG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_));   \
  } while (0)

constexpr int kSizes[] = {4, 8, 16, 32, 48, 64, 96, 128, 256};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

struct tg_plan {
  int n, prec, mhd;
  int M, K, KZ;
  double ks;          // 2 pi / L
  void* tw;           // N complex twiddles exp(-2 pi i j / N) in plan precision
  size_t cb;          // bytes per complex element
  size_t compact;     // complex elements per compact component
  // workspace layout (byte offsets)
  size_t offA, offB, offU0, offU1, offAcc, offPart, offMax, offOut, wsBytes;
  int nDiagBlk, nMax;
  int var[TG_NPASS] = {0, 0, 0, 0, 0};  // tile variant per pass
  // optional per-pass timing: pairs of events around each pass launch
  bool timing = false;
  std::vector<hipEvent_t> evpool;
  std::vector<int> evpass;   // pass id of each recorded pair
  size_t evused = 0;         // pairs recorded since the last read
  // adjoint: private stream (default priority) on which the reverse sweep
  // recomputes the next segment while the current one is reversed
  hipStream_t aux = nullptr;
  hipEvent_t adjev[3] = {nullptr, nullptr, nullptr};
};

namespace {

template <typename R>
struct Ptrs {
  cpx<R>* A;
  cpx<R>* B;
  cpx<R>* U0;    // padded step state u (and b right after it for MHD)
  cpx<R>* U1;    // stage input u (and b right after it for MHD)
  cpx<R>* Acc;   // accumulator u (and b)
  double* part;
  R* maxu;
  R* maxb;
  double* out;
};

template <typename R>
Ptrs<R> carve(const tg_plan* p, void* ws) {
  char* w = static_cast<char*>(ws);
  Ptrs<R> q;
  q.A = reinterpret_cast<cpx<R>*>(w + p->offA);
  q.B = reinterpret_cast<cpx<R>*>(w + p->offB);
  q.U0 = reinterpret_cast<cpx<R>*>(w + p->offU0);
  q.U1 = reinterpret_cast<cpx<R>*>(w + p->offU1);
  q.Acc = reinterpret_cast<cpx<R>*>(w + p->offAcc);
  q.part = reinterpret_cast<double*>(w + p->offPart);
  q.maxu = reinterpret_cast<R*>(w + p->offMax);
  q.maxb = q.maxu + p->nMax;
  q.out = reinterpret_cast<double*>(w + p->offOut);
  return q;
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------- launchers
template <int V>
using VC = std::integral_constant<int, V>;

// Call f(VC<v>) for the run-time tile variant v (0 when N has one variant).
template <int N, class F>
int with_variant(int v, F&& f) {
  if constexpr (num_variants<N>() == 1) {
    return f(VC<0>{});
  } else {
    switch (v) {
      case 1: return f(VC<1>{});
      case 2: return f(VC<2>{});
      default: return f(VC<0>{});
    }
  }
}

// Grids whose adjoint z pass is the register-resident kernel; their reverse
// sweep also keeps each stage's inverse-y output (u, w) for the VJPs.
template <int N>
constexpr bool adj_reg_z() { return N >= 32 && (N & (N - 1)) == 0; }
inline bool adj_reg_z_rt(int n) { return n >= 32 && (n & (n - 1)) == 0; }

template <typename R, int N>
struct Launch {
  using G = Geo<N>;
  static int inv_x(int mode, int var, const cpx<R>* su, const cpx<R>* sb, cpx<R>* out, const cpx<R>* tw, R ks,
                   hipStream_t s, R al = R(1), int cpz = Geo<N>::KZ) {
    return with_variant<N>(var, [&](auto vc) {
      constexpr int V = decltype(vc)::value;
      using C = CfgXY<R, N, V>;
      dim3 grid(cdiv(G::KZ, C::L), G::K), blk(C::NT);
      switch (mode) {
        case XM_HYDRO: hipLaunchKernelGGL((k_inv_x<R, N, XM_HYDRO, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_MHD: hipLaunchKernelGGL((k_inv_x<R, N, XM_MHD, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_COPY3: hipLaunchKernelGGL((k_inv_x<R, N, XM_COPY3, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_CURL3: hipLaunchKernelGGL((k_inv_x<R, N, XM_CURL3, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_ADJ: hipLaunchKernelGGL((k_inv_x<R, N, XM_ADJ, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_ADJQ: hipLaunchKernelGGL((k_inv_x<R, N, XM_ADJQ, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_HYDRO5: hipLaunchKernelGGL((k_inv_x<R, N, XM_HYDRO5, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_MHD5: hipLaunchKernelGGL((k_inv_x<R, N, XM_MHD5, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        default: hipLaunchKernelGGL((k_inv_x<R, N, XM_DIV, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
      }
      TG_LAUNCHED();
      return TG_OK;
    });
  }
  static int y(int dir, int var, int nf, const cpx<R>* in, cpx<R>* out, const cpx<R>* tw, hipStream_t s) {
    return with_variant<N>(var, [&](auto vc) {
      constexpr int V = decltype(vc)::value;
      using C = CfgXY<R, N, V>;
      dim3 grid(cdiv(G::KZ, C::L), N, nf), blk(C::NT);
      if (dir > 0) hipLaunchKernelGGL((k_y<R, N, +1, V>), grid, blk, 0, s, in, out, tw);
      else hipLaunchKernelGGL((k_y<R, N, -1, V>), grid, blk, 0, s, in, out, tw);
      TG_LAUNCHED();
      return TG_OK;
    });
  }
  // Inverse y of `nvec` vectors handed over as five fields each (HYDRO5 /
  // MHD5 x pass), writing (v, i k x v) per vector.
  static int y_curl(int var, int nvec, const cpx<R>* in, cpx<R>* out, const cpx<R>* tw, R ks, hipStream_t s) {
    return with_variant<N>(var, [&](auto vc) {
      constexpr int V = decltype(vc)::value;
      using C = CfgXY<R, N, V>;
      const int units = cdiv(G::KZ, C::L) * N * nvec;
      hipLaunchKernelGGL((k_y_curl<R, N, V>), dim3(48 * cdiv(units, 8)), dim3(C::NT), 0, s, in, out, tw, ks, units);
      TG_LAUNCHED();
      return TG_OK;
    });
  }
  template <int MODE, int V>
  static int z_(const cpx<R>* in, cpx<R>* out, const R* pin, R* pout, const cpx<R>* tw, R lor, hipStream_t s) {
    using Z = CfgZM<R, N, MODE, V>;
    dim3 grid(N / Z::LZ, N), blk(Z::NT);
    hipLaunchKernelGGL((k_z<R, N, MODE, V>), grid, blk, 0, s, in, out, pin, pout, tw, lor);
    TG_LAUNCHED();
    return TG_OK;
  }
  // Register-resident hydro z pass: E points per lane, WPB wavefronts per block.
  template <int E, int WPB>
  static int zw(const cpx<R>* in, cpx<R>* out, const cpx<R>* tw, hipStream_t s) {
    using C = CfgZW<R, N, E, WPB>;
    static_assert((N * N) % C::LINES == 0, "lines per block must divide N^2");
    hipLaunchKernelGGL((k_zw<R, N, E, WPB>), dim3(N * N / C::LINES), dim3(C::NT), 0, s, in, out, tw);
    TG_LAUNCHED();
    return TG_OK;
  }
  // RHS modes take the tuned variant; edge modes (physical I/O, max |div|)
  // always run variant 0.  The hydro z pass is the register-resident kernel:
  // N >= 128 variants 0 / 1 hold 16 / 8 points per lane (variant 2 is the
  // LDS-resident kernel); smaller grids hold 8 (12 for N = 48, 96) points
  // per lane, one wavefront per block.
  static int z(int mode, int var, const cpx<R>* in, cpx<R>* out, const R* pin, R* pout, const cpx<R>* tw, R lor,
               hipStream_t s) {
    switch (mode) {
      case ZM_HYDRO:
        if constexpr (N >= 128) {
          if (var == 0) return zw<16, 4>(in, out, tw, s);
          if (var == 1) return zw<8, 4>(in, out, tw, s);
          return z_<ZM_HYDRO, 2>(in, out, pin, pout, tw, lor, s);
        } else if constexpr (ZwDefault<N>::E > 0) {
          return zw<ZwDefault<N>::E, ZwDefault<N>::WPB>(in, out, tw, s);
        } else {
          return z_<ZM_HYDRO, 0>(in, out, pin, pout, tw, lor, s);
        }
      case ZM_MHD:
        // variant 0 of the power-of-two grids >= 32 is the register-resident kernel
        if constexpr (adj_reg_z<N>()) {
          if (var == 0) {
            constexpr int WPB = ZwDefault<N>::WPB;
            using C = CfgZW<R, N, 8, WPB>;
            hipLaunchKernelGGL((k_zw_mhd<R, N, WPB>), dim3(N * N / C::LINES), dim3(C::NT), 0, s, in, out, tw, lor);
            TG_LAUNCHED();
            return TG_OK;
          }
        }
        return with_variant<N>(var, [&](auto vc) {
          return z_<ZM_MHD, decltype(vc)::value>(in, out, pin, pout, tw, lor, s);
        });
      case ZM_PHYS_OUT3: return z_<ZM_PHYS_OUT3, 0>(in, out, pin, pout, tw, lor, s);
      case ZM_MAXABS: return z_<ZM_MAXABS, 0>(in, out, pin, pout, tw, lor, s);
      case ZM_PHYS_OUT1: return z_<ZM_PHYS_OUT1, 0>(in, out, pin, pout, tw, lor, s);
      case ZM_ADJ:
        if constexpr (adj_reg_z<N>()) {
          return z_adj(in, in + size_t(6) * N * N * kzp<R, N>(), out, tw, s);  // q = fields 6-8 of in
        } else {
          return z_<ZM_ADJ, 0>(in, out, pin, pout, tw, lor, s);
        }
      default: return z_<ZM_PHYS_IN3, 0>(in, out, pin, pout, tw, lor, s);
    }
  }
  // Register-resident adjoint z pass: u, w from fields 0-5 of `in`, q from
  // fields 0-2 of `qin`.
  static int z_adj(const cpx<R>* in, const cpx<R>* qin, cpx<R>* out, const cpx<R>* tw, hipStream_t s) {
    if constexpr (adj_reg_z<N>()) {
      constexpr int WPB = ZwDefault<N>::WPB;
      using C = CfgZW<R, N, 8, WPB>;
      hipLaunchKernelGGL((k_zw_adj<R, N, WPB>), dim3(N * N / C::LINES), dim3(C::NT), 0, s, in, qin, out, tw);
      TG_LAUNCHED();
      return TG_OK;
    } else {
      return fail(TG_ERR_UNSUPPORTED, "register adjoint z pass: N must be a power of two >= 32");
    }
  }
  static int nmax() { return (N / CfgZM<R, N, ZM_MAXABS, 0>::LZ) * N; }
  // RK4 stage variant 1 of N >= 128 (hydro and MHD) is the register-epilogue
  // kernel; so is every adjoint launch on those grids.
  static int fwd_x(int mode, int var, const cpx<R>* in, const StageArgs<R>& a, const cpx<R>* tw, hipStream_t s) {
    if constexpr (N >= 128) {
      if (mode == FM_ADJ) {
        using C = CfgXR<R, N>;
        hipLaunchKernelGGL((k_fwd_xr_adj<R, N>), dim3(cdiv(G::KZ, C::L), G::K), dim3(C::NT), 0, s, in, a, tw);
        TG_LAUNCHED();
        return TG_OK;
      }
      if ((mode == FM_HYDRO_STAGE || mode == FM_MHD_STAGE) && var == 1) {
        using C = CfgXR<R, N>;
        const dim3 grid(cdiv(G::KZ, C::L), G::K), blk(C::NT);
        if (mode == FM_HYDRO_STAGE) hipLaunchKernelGGL((k_fwd_xr<R, N, false>), grid, blk, 0, s, in, a, tw);
        else hipLaunchKernelGGL((k_fwd_xr<R, N, true>), grid, blk, 0, s, in, a, tw);
        TG_LAUNCHED();
        return TG_OK;
      }
    }
    return with_variant<N>(var, [&](auto vc) {
      constexpr int V = decltype(vc)::value;
      using C = CfgXF<R, N, V>;
      const int ff = (mode == FM_MHD_STAGE || mode == FM_MHD_RHS || mode == FM_ADJ) ? 6 : 3;
      const int L = C::LINES / ff;
      dim3 grid(cdiv(G::KZ, L), G::K), blk(C::NT);
      switch (mode) {
        case FM_HYDRO_STAGE: hipLaunchKernelGGL((k_fwd_x<R, N, FM_HYDRO_STAGE, V>), grid, blk, 0, s, in, a, tw); break;
        case FM_MHD_STAGE: hipLaunchKernelGGL((k_fwd_x<R, N, FM_MHD_STAGE, V>), grid, blk, 0, s, in, a, tw); break;
        case FM_STORE3: hipLaunchKernelGGL((k_fwd_x<R, N, FM_STORE3, V>), grid, blk, 0, s, in, a, tw); break;
        case FM_HYDRO_RHS: hipLaunchKernelGGL((k_fwd_x<R, N, FM_HYDRO_RHS, V>), grid, blk, 0, s, in, a, tw); break;
        case FM_ADJ: hipLaunchKernelGGL((k_fwd_x<R, N, FM_ADJ, V>), grid, blk, 0, s, in, a, tw); break;
        default: hipLaunchKernelGGL((k_fwd_x<R, N, FM_MHD_RHS, V>), grid, blk, 0, s, in, a, tw); break;
      }
      TG_LAUNCHED();
      return TG_OK;
    });
  }
  // blocks of the adjoint forward-x pass (slots of the nu-gradient partials)
  static int adj_x_blocks() {
    if constexpr (N >= 128) return cdiv(G::KZ, CfgXR<R, N>::L) * G::K;
    else return cdiv(G::KZ, CfgXF<R, N, 0>::LINES / 6) * G::K;
  }
  static int diag(const cpx<R>* su, const cpx<R>* sb, double* part, int nblk, double ks, hipStream_t s) {
    hipLaunchKernelGGL((k_diag<R, N>), dim3(nblk), dim3(DIAG_NT), 0, s, su, sb, part, ks);
    TG_LAUNCHED();
    return TG_OK;
  }
};

#define TG_RET(...)          \
  do {                       \
    int r_ = (__VA_ARGS__);  \
    if (r_) return r_;       \
  } while (0)

// Pass ids for the timing hook (TG_PASS_* in include/tgdns.h).
int timed_begin(const tg_plan* cp, int pass, hipStream_t s) {
  tg_plan* p = const_cast<tg_plan*>(cp);
  if (!p->timing) return TG_OK;
  if (2 * (p->evused + 1) > p->evpool.size()) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      TG_HIP(hipEventCreate(&e));
      p->evpool.push_back(e);
    }
  }
  p->evpass.resize(p->evused + 1);
  p->evpass[p->evused] = pass;
  TG_HIP(hipEventRecord(p->evpool[2 * p->evused], s));
  return TG_OK;
}

int timed_end(const tg_plan* cp, hipStream_t s) {
  tg_plan* p = const_cast<tg_plan*>(cp);
  if (!p->timing) return TG_OK;
  TG_HIP(hipEventRecord(p->evpool[2 * p->evused + 1], s));
  ++p->evused;
  return TG_OK;
}

#define TG_TIMED(pass, ...)                  \
  do {                                       \
    TG_RET(timed_begin(p, (pass), s));       \
    TG_RET(__VA_ARGS__);                     \
    TG_RET(timed_end(p, s));                 \
  } while (0)

// One right-hand side: stage input (us, bs) -> z-pass input in A/B -> the
// forward-x epilogue described by `mode` / `a`.
template <typename R, int N>
// bslot (hydro only): where the inverse-y output (u, w) goes and the z pass
// reads it, instead of the workspace B (the adjoint keeps it per stage).
int rhs_passes(const tg_plan* p, const Ptrs<R>& w, const cpx<R>* us, const cpx<R>* bs, int fmode,
               const StageArgs<R>& a, R lor, hipStream_t s, cpx<R>* bslot = nullptr) {
  const int cpz = a.cpz ? a.cpz : Geo<N>::KZ;
  using Lc = Launch<R, N>;
  const cpx<R>* tw = static_cast<const cpx<R>*>(p->tw);
  const bool mhd = p->mhd != 0;
  const int fo = mhd ? 6 : 3;
  const int* v = p->var;
  TG_TIMED(TG_PASS_INV_X, Lc::inv_x(mhd ? XM_MHD5 : XM_HYDRO5, v[TG_PASS_INV_X], us, bs, w.A, tw, R(p->ks), s, R(1),
                                    cpz));
  cpx<R>* bm = bslot ? bslot : w.B;
  TG_TIMED(TG_PASS_INV_Y, Lc::y_curl(v[TG_PASS_INV_Y], mhd ? 2 : 1, w.A, bm, tw, R(p->ks), s));
  TG_TIMED(TG_PASS_Z, Lc::z(mhd ? ZM_MHD : ZM_HYDRO, v[TG_PASS_Z], bm, w.A, nullptr, nullptr, tw, lor, s));
  TG_TIMED(TG_PASS_FWD_Y, Lc::y(-1, v[TG_PASS_FWD_Y], fo, w.A, w.B, tw, s));
  TG_TIMED(TG_PASS_FWD_X, Lc::fwd_x(fmode, v[TG_PASS_FWD_X], w.B, a, tw, s));
  return TG_OK;
}

template <typename R, int N>
struct Ops {
  static int forward(const tg_plan* p, const R* phys, cpx<R>* spec, void* ws, hipStream_t s) {
    using Lc = Launch<R, N>;
    Ptrs<R> w = carve<R>(p, ws);
    const cpx<R>* tw = static_cast<const cpx<R>*>(p->tw);
    TG_RET(Lc::z(ZM_PHYS_IN3, 0, nullptr, w.A, phys, nullptr, tw, R(0), s));
    TG_RET(Lc::y(-1, 0, 3, w.A, w.B, tw, s));
    StageArgs<R> a{};
    a.unext = spec;
    a.ks = R(p->ks);
    return Lc::fwd_x(FM_STORE3, 0, w.B, a, tw, s);
  }
  static int inverse(const tg_plan* p, const cpx<R>* spec, int op, R* phys, void* ws, hipStream_t s) {
    using Lc = Launch<R, N>;
    Ptrs<R> w = carve<R>(p, ws);
    const cpx<R>* tw = static_cast<const cpx<R>*>(p->tw);
    const int xm = op == TG_OP_CURL ? XM_CURL3 : op == TG_OP_DIV ? XM_DIV : XM_COPY3;
    const int nf = op == TG_OP_DIV ? 1 : 3;
    TG_RET(Lc::inv_x(xm, 0, spec, nullptr, w.A, tw, R(p->ks), s));
    TG_RET(Lc::y(+1, 0, nf, w.A, w.B, tw, s));
    return Lc::z(op == TG_OP_DIV ? ZM_PHYS_OUT1 : ZM_PHYS_OUT3, 0, w.B, nullptr, nullptr, phys, tw, R(0), s);
  }
  static int rhs(const tg_plan* p, const cpx<R>* u, const cpx<R>* b, double nu, double eta, double lor,
                 cpx<R>* ru, cpx<R>* rb, void* ws, hipStream_t s) {
    Ptrs<R> w = carve<R>(p, ws);
    StageArgs<R> a{};
    a.us = u; a.bs = b; a.unext = ru; a.bnext = rb;
    a.nu = R(nu); a.eta = R(eta); a.ks = R(p->ks);
    return rhs_passes<R, N>(p, w, u, b, p->mhd ? FM_MHD_RHS : FM_HYDRO_RHS, a, R(lor), s);
  }
  // nsteps RK4 steps.  The loop runs on line-padded internal copies of the
  // state (row pitch kzp instead of the dense KZ of the ABI layout), copied
  // in once and out once per call.
  static int rk4(const tg_plan* p, cpx<R>* u, cpx<R>* b, int nsteps, double dt, double nu, double eta,
                 double lor, void* ws, hipStream_t s) {
    if (nsteps == 0) return TG_OK;
    Ptrs<R> w = carve<R>(p, ws);
    constexpr int KZP = kzp<R, N>();
    constexpr int KZ = Geo<N>::KZ;
    const size_t rows = size_t(3) * p->K * p->K;
    const size_t vec = rows * KZP;  // padded vector (3 components)
    cpx<R>* u0 = w.U0;
    cpx<R>* b0 = w.U0 + vec;
    const size_t e = sizeof(cpx<R>);
    TG_HIP(hipMemcpy2DAsync(u0, KZP * e, u, KZ * e, KZ * e, rows, hipMemcpyDeviceToDevice, s));
    if (p->mhd) TG_HIP(hipMemcpy2DAsync(b0, KZP * e, b, KZ * e, KZ * e, rows, hipMemcpyDeviceToDevice, s));
    TG_RET(rk4_padded(p, u0, b0, nsteps, dt, nu, eta, lor, ws, s));
    TG_HIP(hipMemcpy2DAsync(u, KZ * e, u0, KZP * e, KZ * e, rows, hipMemcpyDeviceToDevice, s));
    if (p->mhd) TG_HIP(hipMemcpy2DAsync(b, KZ * e, b0, KZP * e, KZ * e, rows, hipMemcpyDeviceToDevice, s));
    return TG_OK;
  }
  // nsteps RK4 steps on line-padded state vectors u0 (b0) in place; the
  // stage input and accumulator live in the workspace.
  static int rk4_padded(const tg_plan* p, cpx<R>* u0, cpx<R>* b0, int nsteps, double dt, double nu, double eta,
                        double lor, void* ws, hipStream_t s) {
    Ptrs<R> w = carve<R>(p, ws);
    constexpr int KZP = kzp<R, N>();
    const size_t vec = size_t(3) * p->K * p->K * KZP;
    cpx<R>* u1 = w.U1;
    cpx<R>* b1 = w.U1 + vec;
    cpx<R>* ua = w.Acc;
    cpx<R>* ba = w.Acc + vec;
    const int fm = p->mhd ? FM_MHD_STAGE : FM_HYDRO_STAGE;
    for (int step = 0; step < nsteps; ++step) {
      for (int st = 0; st < 4; ++st) {
        StageArgs<R> a{};
        a.u0 = u0; a.b0 = b0;
        a.us = st == 0 ? u0 : u1;
        a.bs = st == 0 ? b0 : b1;
        a.unext = st == 3 ? u0 : u1;
        a.bnext = st == 3 ? b0 : b1;
        a.uacc = ua; a.bacc = ba;
        a.dt = R(dt); a.nu = R(nu); a.eta = R(eta); a.ks = R(p->ks);
        a.stage = st;
        a.cpz = KZP;
        TG_RET(rhs_passes<R, N>(p, w, a.us, a.bs, fm, a, R(lor), s));
      }
    }
    return TG_OK;
  }
  // One vector-Jacobian product of the hydro RHS at x for the cotangent a
  // (scaled by a.alpha), with the epilogue bookkeeping described in `a`.
  // uw: the stage's kept inverse-y output (u, w) of x, if any: then only the
  // cotangent's projection q is transformed.
  static int vjp_passes(const tg_plan* p, const Ptrs<R>& w, const cpx<R>* x, const StageArgs<R>& a,
                        hipStream_t s, const cpx<R>* uw = nullptr) {
    using Lc = Launch<R, N>;
    const cpx<R>* tw = static_cast<const cpx<R>*>(p->tw);
    const int cpz = a.cpz ? a.cpz : Geo<N>::KZ;  // row pitch of x, a and the epilogue vectors
    if (uw) {
      TG_TIMED(TG_PASS_INV_X, Lc::inv_x(XM_ADJQ, p->var[TG_PASS_INV_X], a.a, nullptr, w.A, tw, R(p->ks), s, a.alpha,
                                        cpz));
      TG_TIMED(TG_PASS_INV_Y, Lc::y(+1, p->var[TG_PASS_INV_Y], 3, w.A, w.B, tw, s));
      TG_TIMED(TG_PASS_Z, Lc::z_adj(uw, w.B, w.A, tw, s));
      TG_TIMED(TG_PASS_FWD_Y, Lc::y(-1, p->var[TG_PASS_FWD_Y], 6, w.A, w.B, tw, s));
      TG_TIMED(TG_PASS_FWD_X, Lc::fwd_x(FM_ADJ, 0, w.B, a, tw, s));
      return TG_OK;
    }
    // the 8-points-per-thread tile (variant 1, where compiled) keeps the six
    // state and cotangent components in registers
    TG_TIMED(TG_PASS_INV_X, Lc::inv_x(XM_ADJ, num_variants<N>() > 1 ? 1 : 0, x, a.a, w.A, tw, R(p->ks), s, a.alpha,
                                      cpz));
    TG_TIMED(TG_PASS_INV_Y, Lc::y(+1, p->var[TG_PASS_INV_Y], 9, w.A, w.B, tw, s));
    TG_TIMED(TG_PASS_Z, Lc::z(ZM_ADJ, 0, w.B, w.A, nullptr, nullptr, tw, R(0), s));
    TG_TIMED(TG_PASS_FWD_Y, Lc::y(-1, p->var[TG_PASS_FWD_Y], 6, w.A, w.B, tw, s));
    TG_TIMED(TG_PASS_FWD_X, Lc::fwd_x(FM_ADJ, 0, w.B, a, tw, s));
    return TG_OK;
  }
  static int vjp(const tg_plan* p, const cpx<R>* x, const cpx<R>* av, double nu, cpx<R>* out, void* ws,
                 hipStream_t s) {
    Ptrs<R> w = carve<R>(p, ws);
    StageArgs<R> a{};
    a.nu = R(nu); a.ks = R(p->ks);
    a.a = av; a.alpha = R(1); a.lacc = out; a.first = 2; a.last = 1;
    return vjp_passes(p, w, x, a, s);
  }
  // Adjoint workspace: cur | lamp | lacc | abuf | nck checkpoints | 4 seg
  // stage states | nu-gradient partials, all vectors with line-padded rows
  // (pitch kzp, as in the RK4 loop), and with more than one segment a
  // second set for the recompute that overlaps the reverse sweep: cur2 |
  // 4 seg stage states | RHS workspace.
  // On grids with the register adjoint z pass, each stage buffer has a
  // matching set of 4 seg slots for the stages' inverse-y output (u, w).
  struct AdjLayout {
    cpx<R>* cur; cpx<R>* lamp; cpx<R>* lacc; cpx<R>* abuf; cpx<R>* ck; cpx<R>* st; double* nupart;
    cpx<R>* cur2; cpx<R>* st2; void* ws2;
    cpx<R>* slots; cpx<R>* slots2;
    size_t vstride, sstride; int nck, nblk;
  };
  // one padded 3-component vector
  static size_t vec_bytes(const tg_plan* p) { return align256(size_t(3) * p->K * p->K * kzp<R, N>() * sizeof(cpx<R>)); }
  static size_t slot_bytes(const tg_plan* p) { return align256(size_t(6) * N * N * kzp<R, N>() * sizeof(cpx<R>)); }
  static size_t nupart_bytes(const tg_plan* p) { return align256(size_t(p->K) * p->KZ * 8 * sizeof(double)); }
  static AdjLayout adj_layout(const tg_plan* p, int nsteps, int seg, void* aws) {
    AdjLayout L;
    const size_t vb = vec_bytes(p);
    L.nck = (nsteps + seg - 1) / seg;
    char* base = static_cast<char*>(aws);
    L.cur = reinterpret_cast<cpx<R>*>(base);
    L.lamp = reinterpret_cast<cpx<R>*>(base + vb);
    L.lacc = reinterpret_cast<cpx<R>*>(base + 2 * vb);
    L.abuf = reinterpret_cast<cpx<R>*>(base + 3 * vb);
    L.ck = reinterpret_cast<cpx<R>*>(base + 4 * vb);
    L.st = reinterpret_cast<cpx<R>*>(base + (4 + L.nck) * vb);
    L.nupart = reinterpret_cast<double*>(base + (4 + L.nck + 4 * size_t(seg)) * vb);
    char* second = base + (4 + L.nck + 4 * size_t(seg)) * vb + nupart_bytes(p);
    L.cur2 = L.nck > 1 ? reinterpret_cast<cpx<R>*>(second) : nullptr;
    L.st2 = L.nck > 1 ? reinterpret_cast<cpx<R>*>(second + vb) : nullptr;
    L.ws2 = L.nck > 1 ? static_cast<void*>(second + (1 + 4 * size_t(seg)) * vb) : nullptr;
    char* third = L.nck > 1 ? second + (1 + 4 * size_t(seg)) * vb + align256(p->wsBytes) : second;
    const size_t sb = slot_bytes(p);
    L.slots = adj_reg_z<N>() ? reinterpret_cast<cpx<R>*>(third) : nullptr;
    L.slots2 = (adj_reg_z<N>() && L.nck > 1) ? reinterpret_cast<cpx<R>*>(third + 4 * size_t(seg) * sb) : nullptr;
    L.sstride = sb / sizeof(cpx<R>);
    L.vstride = vb / sizeof(cpx<R>);
    L.nblk = Launch<R, N>::adj_x_blocks();
    return L;
  }
  // `len` RK4 steps from cur keeping the four stage states of step j in
  // st[4j .. 4j+3] (stage 1 = the step's input state).
  static int steps_keeping_stages(const tg_plan* p, cpx<R>* cur, cpx<R>* st, size_t vstride, int len, double dt,
                                  double nu, void* ws, hipStream_t s, cpx<R>* slots = nullptr,
                                  size_t sstride = 0) {
    Ptrs<R> w = carve<R>(p, ws);
    const size_t bytes = size_t(3) * p->K * p->K * kzp<R, N>() * sizeof(cpx<R>);
    for (int j = 0; j < len; ++j) {
      cpx<R>* S = st + size_t(4 * j) * vstride;
      TG_HIP(hipMemcpyAsync(S, cur, bytes, hipMemcpyDeviceToDevice, s));
      for (int stg = 0; stg < 4; ++stg) {
        StageArgs<R> a{};
        a.u0 = cur;
        a.us = S + size_t(stg) * vstride;
        a.unext = stg == 3 ? cur : S + size_t(stg + 1) * vstride;
        a.uacc = w.Acc;
        a.dt = R(dt); a.nu = R(nu); a.ks = R(p->ks); a.stage = stg;
        a.cpz = kzp<R, N>();
        cpx<R>* slot = slots ? slots + size_t(4 * j + stg) * sstride : nullptr;
        TG_RET(rhs_passes<R, N>(p, w, a.us, nullptr, FM_HYDRO_STAGE, a, R(0), s, slot));
      }
    }
    return TG_OK;
  }
  // Forward sweep: nsteps RK4 steps from x0 with a checkpoint every `seg`
  // steps, J = E(T) (objective 0) or Z(T) (1), and the terminal cotangent.
  static int adj_forward(const tg_plan* p, const cpx<R>* x0, int nsteps, double dt, double nu, int objective,
                         int seg, double* J, void* ws, void* aws, hipStream_t s) {
    AdjLayout L = adj_layout(p, nsteps, seg, aws);
    constexpr int KZP = kzp<R, N>(), KZ = Geo<N>::KZ;
    const size_t rows = size_t(3) * p->K * p->K, e = sizeof(cpx<R>);
    const size_t bytes = rows * KZP * e;
    TG_HIP(hipMemcpy2DAsync(L.cur, KZP * e, x0, KZ * e, KZ * e, rows, hipMemcpyDeviceToDevice, s));
    for (int c = 0; c < L.nck; ++c) {
      TG_HIP(hipMemcpyAsync(L.ck + c * L.vstride, L.cur, bytes, hipMemcpyDeviceToDevice, s));
      const int len = std::min(seg, nsteps - c * seg);
      if (c < L.nck - 1) TG_RET(rk4_padded(p, L.cur, nullptr, len, dt, nu, 0.0, -1.0, ws, s));
      else TG_RET(steps_keeping_stages(p, L.cur, L.st, L.vstride, len, dt, nu, ws, s, L.slots,
                                       L.sstride));  // the reverse sweep starts here
    }
    // diagnostics read the dense layout: a copy of the final state in abuf
    TG_HIP(hipMemcpy2DAsync(L.abuf, KZ * e, L.cur, KZP * e, KZ * e, rows, hipMemcpyDeviceToDevice, s));
    double d[NDIAG + 2];
    TG_RET(diagnostics(p, L.abuf, nullptr, 0, d, ws, s));
    *J = objective == 0 ? d[0] : d[1];
    hipLaunchKernelGGL((k_obj_grad<R, N>), dim3(cdiv(int(rows / 3 * KZP), 256)), dim3(256), 0, s, L.cur, L.lamp,
                       objective, R(p->ks), KZP);
    TG_LAUNCHED();
    return TG_OK;
  }
  // Reverse sweep: per segment (last first) recompute its steps from the
  // checkpoint keeping the four stage states of each step, then four
  // vector-Jacobian products per step in reverse.  grad = dJ/d(stored
  // entries of x0), dJ/dnu to the host.
  static int adj_backward(const tg_plan* p, int nsteps, double dt, double nu, int seg, cpx<R>* grad,
                          double* dJdnu, void* ws, void* aws, hipStream_t s) {
    Ptrs<R> w = carve<R>(p, ws);
    AdjLayout L = adj_layout(p, nsteps, seg, aws);
    constexpr int KZP = kzp<R, N>();
    const size_t bytes = size_t(3) * p->K * p->K * KZP * sizeof(cpx<R>);
    TG_HIP(hipMemsetAsync(L.nupart, 0, L.nblk * sizeof(double), s));
    cpx<R>* lamp = L.lamp;
    cpx<R>* lacc = L.lacc;
    const R h = R(dt);
    const R cl[4] = {h / R(6), h / R(3), h / R(3), h / R(6)};  // weight of lamp in a_s (s = 1..4)
    const R cm[4] = {R(0), h / R(2), h / R(2), h};              // weight of mu_{s+1} in a_s
    // The last segment's stages are kept by the forward sweep (buffer 0).
    // Segment c - 1 is recomputed on the private stream into the other
    // buffer while segment c is reversed on `s`; events order the buffer
    // reuse both ways.
    hipStream_t s2 = nullptr;
    hipEvent_t ev_in = nullptr, ev_rec = nullptr, ev_rev = nullptr;
    if (L.nck > 1) {
      tg_plan* mp = const_cast<tg_plan*>(p);
      if (!mp->aux) TG_HIP(hipStreamCreateWithFlags(&mp->aux, hipStreamNonBlocking));
      for (hipEvent_t& e : mp->adjev)
        if (!e) TG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      s2 = mp->aux;
      ev_in = mp->adjev[0]; ev_rec = mp->adjev[1]; ev_rev = mp->adjev[2];
      TG_HIP(hipEventRecord(ev_in, s));  // checkpoints written, nupart cleared
      TG_HIP(hipStreamWaitEvent(s2, ev_in, 0));
    }
    for (int c = L.nck - 1; c >= 0; --c) {
      const int len = std::min(seg, nsteps - c * seg);
      const bool even = (L.nck - 1 - c) % 2 == 0;
      cpx<R>* stc = even ? L.st : L.st2;
      cpx<R>* slc = even ? L.slots : L.slots2;
      if (c < L.nck - 1) TG_HIP(hipStreamWaitEvent(s, ev_rec, 0));  // segment c recomputed
      if (c > 0) {
        // recompute segment c - 1 into the buffer segment c + 1 used
        const bool evn = (L.nck - c) % 2 == 0;
        cpx<R>* stn = evn ? L.st : L.st2;
        cpx<R>* sln = evn ? L.slots : L.slots2;
        const int lenn = std::min(seg, nsteps - (c - 1) * seg);
        TG_HIP(hipMemcpyAsync(L.cur2, L.ck + (c - 1) * L.vstride, bytes, hipMemcpyDeviceToDevice, s2));
        TG_RET(steps_keeping_stages(p, L.cur2, stn, L.vstride, lenn, dt, nu, L.ws2, s2, sln, L.sstride));
        TG_HIP(hipEventRecord(ev_rec, s2));
      }
      for (int j = len - 1; j >= 0; --j) {
        cpx<R>* S = stc + size_t(4 * j) * L.vstride;
        for (int sidx = 3; sidx >= 0; --sidx) {
          StageArgs<R> a{};
          a.nu = R(nu); a.ks = R(p->ks); a.cpz = KZP;
          a.lamp = lamp; a.lacc = lacc; a.xs = S + size_t(sidx) * L.vstride;
          a.nupart = L.nupart;
          if (sidx == 3) { a.a = lamp; a.alpha = cl[3]; a.first = 1; }
          else { a.a = L.abuf; a.alpha = R(1); a.first = 0; }
          a.last = sidx == 0;
          if (sidx > 0) { a.anext = L.abuf; a.clam = cl[sidx - 1]; a.cmu = cm[sidx]; }
          const cpx<R>* uw = slc ? slc + size_t(4 * j + sidx) * L.sstride : nullptr;
          TG_RET(vjp_passes(p, w, a.xs, a, s, uw));
        }
        std::swap(lamp, lacc);
      }
      if (c > 1) {
        // the recompute of segment c - 2 reuses this buffer: after this reverse
        TG_HIP(hipEventRecord(ev_rev, s));
        TG_HIP(hipStreamWaitEvent(s2, ev_rev, 0));
      }
    }
    hipLaunchKernelGGL((k_adj_final<R, N>), dim3(cdiv(int(p->compact), 256)), dim3(256), 0, s, lamp, grad, KZP);
    TG_LAUNCHED();
    std::vector<double> part(L.nblk);
    TG_HIP(hipMemcpyAsync(part.data(), L.nupart, L.nblk * sizeof(double), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    double sum = 0.0, comp = 0.0;
    for (double v : part) {
      const double t = sum + v;
      comp += std::fabs(sum) >= std::fabs(v) ? (sum - t) + v : (v - t) + sum;
      sum = t;
    }
    const double n3 = double(p->n) * p->n * p->n;
    *dJdnu = -(sum + comp) / n3;
    return TG_OK;
  }
  static int diagnostics(const tg_plan* p, const cpx<R>* u, const cpx<R>* b, int divmax, double* out,
                         void* ws, hipStream_t s) {
    using Lc = Launch<R, N>;
    Ptrs<R> w = carve<R>(p, ws);
    const cpx<R>* tw = static_cast<const cpx<R>*>(p->tw);
    TG_RET(Lc::diag(u, b, w.part, p->nDiagBlk, p->ks, s));
    int nm = 0;
    if (divmax) {
      nm = p->nMax;
      TG_RET(Lc::inv_x(XM_DIV, 0, u, nullptr, w.A, tw, R(p->ks), s));
      TG_RET(Lc::y(+1, 0, 1, w.A, w.B, tw, s));
      TG_RET(Lc::z(ZM_MAXABS, 0, w.B, nullptr, nullptr, w.maxu, tw, R(0), s));
      if (b) {
        TG_RET(Lc::inv_x(XM_DIV, 0, b, nullptr, w.A, tw, R(p->ks), s));
        TG_RET(Lc::y(+1, 0, 1, w.A, w.B, tw, s));
        TG_RET(Lc::z(ZM_MAXABS, 0, w.B, nullptr, nullptr, w.maxb, tw, R(0), s));
      }
    }
    hipLaunchKernelGGL((k_diag_final<R>), dim3(1), dim3(64), 0, s, w.part, p->nDiagBlk,
                       divmax ? w.maxu : nullptr, (divmax && b) ? w.maxb : nullptr, nm, w.out);
    TG_LAUNCHED();
    double host[NDIAG + 2];
    TG_HIP(hipMemcpyAsync(host, w.out, sizeof(host), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    const double n3 = double(p->n) * p->n * p->n;
    const double n6 = n3 * n3;
    out[0] = 0.5 * host[0] / n6;
    out[1] = 0.5 * host[1] / n6;
    out[2] = host[2] / n6;
    out[3] = 0.5 * host[3] / n6;
    out[4] = host[4] / n6;
    out[5] = host[5] / n6;
    out[6] = host[6] / n6;
    out[7] = host[7] / n6;
    out[8] = host[8];
    out[9] = host[9];
    return TG_OK;
  }
};

// Runtime N -> template dispatch.
template <typename R, template <typename, int> class Fn>
struct Dispatch {
  template <typename Call>
  static int run(int n, Call&& c) {
    switch (n) {
      case 4: return c(Fn<R, 4>{});
      case 8: return c(Fn<R, 8>{});
      case 16: return c(Fn<R, 16>{});
      case 32: return c(Fn<R, 32>{});
      case 48: return c(Fn<R, 48>{});
      case 64: return c(Fn<R, 64>{});
      case 96: return c(Fn<R, 96>{});
      case 128: return c(Fn<R, 128>{});
      case 256: return c(Fn<R, 256>{});
      default: return fail(TG_ERR_UNSUPPORTED, "grid size N=" + std::to_string(n) + " has no compiled plan");
    }
  }
};

template <typename Call>
int with_plan(const tg_plan* p, Call&& c) {
  if (p->prec == 64) return Dispatch<double, Ops>::run(p->n, [&](auto ops) { return c(ops, (double*)nullptr); });
  return Dispatch<float, Ops>::run(p->n, [&](auto ops) { return c(ops, (float*)nullptr); });
}

int check_plan(const tg_plan* p) {
  if (!p) return fail(TG_ERR_ARG, "plan is NULL");
  return TG_OK;
}

template <int N>
int nmax_for() { return Launch<float, N>::nmax(); }

// Padded row pitch of the intermediate arrays for a run-time N.
template <typename R>
int kzp_rt(int n) {
  switch (n) {
    case 4: return kzp<R, 4>();
    case 8: return kzp<R, 8>();
    case 16: return kzp<R, 16>();
    case 32: return kzp<R, 32>();
    case 48: return kzp<R, 48>();
    case 64: return kzp<R, 64>();
    case 96: return kzp<R, 96>();
    case 128: return kzp<R, 128>();
    case 256: return kzp<R, 256>();
  }
  return 0;
}

int nmax_rt(int n) {
  switch (n) {
    case 4: return nmax_for<4>();
    case 8: return nmax_for<8>();
    case 16: return nmax_for<16>();
    case 32: return nmax_for<32>();
    case 48: return nmax_for<48>();
    case 64: return nmax_for<64>();
    case 96: return nmax_for<96>();
    case 128: return nmax_for<128>();
    case 256: return nmax_for<256>();
  }
  return 0;
}

}  // namespace

extern "C" {

int tg_version(void) { return 1; }

int tg_supported_sizes(int* out, int cap) {
  const int n = int(sizeof(kSizes) / sizeof(kSizes[0]));
  for (int i = 0; i < n && i < cap; ++i) out[i] = kSizes[i];
  return n;
}

int tg_plan_create(int n, int precision_bits, int mhd, double domain_length, tg_plan** out) {
  if (!out) return fail(TG_ERR_ARG, "out is NULL");
  *out = nullptr;
  bool ok = false;
  for (int s : kSizes) ok |= (s == n);
  if (!ok) return fail(TG_ERR_UNSUPPORTED, "grid size N=" + std::to_string(n) + " has no compiled plan");
  if (precision_bits != 64 && precision_bits != 32)
    return fail(TG_ERR_ARG, "precision_bits must be 64 or 32");
  if (!(domain_length > 0.0)) return fail(TG_ERR_ARG, "domain_length must be positive");
  tg_plan* p = new tg_plan{};
  p->n = n;
  p->prec = precision_bits;
  p->mhd = mhd ? 1 : 0;
  p->M = n / 3;
  p->K = 2 * p->M + 1;
  p->KZ = p->M + 1;
  p->ks = 2.0 * M_PI / domain_length;
  if (domain_length == 2.0 * M_PI) p->ks = 1.0;
  p->cb = precision_bits == 64 ? 16 : 8;
  p->compact = size_t(p->K) * p->K * p->KZ;
  const size_t N = n;
  const size_t fi = p->mhd ? 12 : 9;  // inverse fields: hydro RHS 6, adjoint 9, MHD 12
  const size_t fo = 6;                 // z-pass outputs: hydro 3, adjoint and MHD 6
  // A holds the inverse-x output (fi x N K KZ) and later the z output (fo x N N KZ)
  const size_t kzpad = precision_bits == 64 ? size_t(kzp_rt<double>(n)) : size_t(kzp_rt<float>(n));
  const size_t sA = std::max(fi * N * p->K, fo * N * N) * kzpad * p->cb;
  const size_t sB = fi * N * N * kzpad * p->cb;
  // padded state vectors (row pitch kzpad) for the RK4 loop
  const size_t sU = (p->mhd ? 2 : 1) * 3 * size_t(p->K) * p->K * kzpad * p->cb;
  p->nDiagBlk = 512;
  p->nMax = nmax_rt(n);
  size_t o = 0;
  p->offA = o; o += align256(sA);
  p->offB = o; o += align256(sB);
  p->offU0 = o; o += align256(sU);
  p->offU1 = o; o += align256(sU);
  p->offAcc = o; o += align256(sU);
  p->offPart = o; o += align256(size_t(p->nDiagBlk) * NDIAG * sizeof(double));
  p->offMax = o; o += align256(2 * size_t(p->nMax) * (p->cb / 2));
  p->offOut = o; o += align256((NDIAG + 2) * sizeof(double));
  p->wsBytes = o;
  // twiddles exp(-2 pi i j / N), evaluated in long double on the host
  std::vector<char> host(N * p->cb);
  for (size_t j = 0; j < N; ++j) {
    const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)N;
    const long double c = cosl(ang), sn = sinl(ang);
    if (p->prec == 64) {
      double v[2] = {double(c), double(sn)};
      memcpy(host.data() + j * 16, v, 16);
    } else {
      float v[2] = {float(c), float(sn)};
      memcpy(host.data() + j * 8, v, 8);
    }
  }
  hipError_t e = hipMalloc(&p->tw, host.size());
  if (e != hipSuccess) {
    delete p;
    return fail(TG_ERR_HIP, std::string("hipMalloc(twiddles): ") + hipGetErrorString(e));
  }
  e = hipMemcpy(p->tw, host.data(), host.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(p->tw);
    delete p;
    return fail(TG_ERR_HIP, std::string("hipMemcpy(twiddles): ") + hipGetErrorString(e));
  }
  *out = p;
  return TG_OK;
}

int tg_plan_destroy(tg_plan* p) {
  if (!p) return TG_OK;
  for (hipEvent_t e : p->evpool) (void)hipEventDestroy(e);
  for (hipEvent_t e : p->adjev)
    if (e) (void)hipEventDestroy(e);
  if (p->aux) (void)hipStreamDestroy(p->aux);
  if (p->tw) (void)hipFree(p->tw);
  delete p;
  return TG_OK;
}

int tg_plan_query(const tg_plan* p, long long* info) {
  TG_RET(check_plan(p));
  if (!info) return fail(TG_ERR_ARG, "info is NULL");
  info[TG_INFO_N] = p->n;
  info[TG_INFO_K] = p->K;
  info[TG_INFO_KZ] = p->KZ;
  info[TG_INFO_COMPACT_ELEMS] = (long long)p->compact;
  info[TG_INFO_WS_BYTES] = (long long)p->wsBytes;
  info[TG_INFO_PRECISION] = p->prec;
  info[TG_INFO_MHD] = p->mhd;
  info[7] = 0;
  return TG_OK;
}

int tg_forward(tg_plan* p, const void* phys, void* spec, void* ws, void* stream) {
  TG_RET(check_plan(p));
  if (!phys || !spec || !ws) return fail(TG_ERR_ARG, "tg_forward: NULL buffer");
  return with_plan(p, [&](auto ops, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return decltype(ops)::forward(p, static_cast<const R*>(phys), static_cast<cpx<R>*>(spec), ws,
                                  static_cast<hipStream_t>(stream));
  });
}

int tg_inverse(tg_plan* p, const void* spec, int op, void* phys, void* ws, void* stream) {
  TG_RET(check_plan(p));
  if (!phys || !spec || !ws) return fail(TG_ERR_ARG, "tg_inverse: NULL buffer");
  if (op != TG_OP_FIELD && op != TG_OP_CURL && op != TG_OP_DIV)
    return fail(TG_ERR_ARG, "tg_inverse: op must be TG_OP_FIELD, TG_OP_CURL or TG_OP_DIV");
  return with_plan(p, [&](auto ops, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return decltype(ops)::inverse(p, static_cast<const cpx<R>*>(spec), op, static_cast<R*>(phys), ws,
                                  static_cast<hipStream_t>(stream));
  });
}

int tg_rhs(tg_plan* p, const void* u_hat, const void* b_hat, double nu, double eta, double lorentz_sign,
           void* ru_hat, void* rb_hat, void* ws, void* stream) {
  TG_RET(check_plan(p));
  if (!u_hat || !ru_hat || !ws) return fail(TG_ERR_ARG, "tg_rhs: NULL buffer");
  if (p->mhd && (!b_hat || !rb_hat)) return fail(TG_ERR_ARG, "tg_rhs: MHD plan needs b_hat and rb_hat");
  return with_plan(p, [&](auto ops, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return decltype(ops)::rhs(p, static_cast<const cpx<R>*>(u_hat), static_cast<const cpx<R>*>(b_hat), nu, eta,
                              lorentz_sign, static_cast<cpx<R>*>(ru_hat), static_cast<cpx<R>*>(rb_hat), ws,
                              static_cast<hipStream_t>(stream));
  });
}

int tg_rk4_steps(tg_plan* p, void* u_hat, void* b_hat, int nsteps, double dt, double nu, double eta,
                 double lorentz_sign, void* ws, void* stream) {
  TG_RET(check_plan(p));
  if (!u_hat || !ws) return fail(TG_ERR_ARG, "tg_rk4_steps: NULL buffer");
  if (p->mhd && !b_hat) return fail(TG_ERR_ARG, "tg_rk4_steps: MHD plan needs b_hat");
  if (nsteps < 0) return fail(TG_ERR_ARG, "tg_rk4_steps: nsteps must be >= 0");
  return with_plan(p, [&](auto ops, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return decltype(ops)::rk4(p, static_cast<cpx<R>*>(u_hat), static_cast<cpx<R>*>(b_hat), nsteps, dt, nu, eta,
                              lorentz_sign, ws, static_cast<hipStream_t>(stream));
  });
}

int tg_diagnostics(tg_plan* p, const void* u_hat, const void* b_hat, int with_divmax, double* out, void* ws,
                   void* stream) {
  TG_RET(check_plan(p));
  if (!u_hat || !out || !ws) return fail(TG_ERR_ARG, "tg_diagnostics: NULL buffer");
  return with_plan(p, [&](auto ops, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return decltype(ops)::diagnostics(p, static_cast<const cpx<R>*>(u_hat), static_cast<const cpx<R>*>(b_hat),
                                      with_divmax, out, ws, static_cast<hipStream_t>(stream));
  });
}

int tg_vjp(tg_plan* p, const void* x_hat, const void* a_hat, double nu, void* out_hat, void* ws, void* stream) {
  TG_RET(check_plan(p));
  if (p->mhd) return fail(TG_ERR_UNSUPPORTED, "tg_vjp: hydro plans only");
  if (!x_hat || !a_hat || !out_hat || !ws) return fail(TG_ERR_ARG, "tg_vjp: NULL buffer");
  return with_plan(p, [&](auto ops, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return decltype(ops)::vjp(p, static_cast<const cpx<R>*>(x_hat), static_cast<const cpx<R>*>(a_hat), nu,
                              static_cast<cpx<R>*>(out_hat), ws, static_cast<hipStream_t>(stream));
  });
}

int tg_adjoint_workspace(tg_plan* p, int nsteps, int seg, long long* bytes) {
  TG_RET(check_plan(p));
  if (nsteps < 1 || seg < 1 || !bytes) return fail(TG_ERR_ARG, "tg_adjoint_workspace: need nsteps, seg >= 1");
  const size_t kzpad = p->prec == 64 ? size_t(kzp_rt<double>(p->n)) : size_t(kzp_rt<float>(p->n));
  const size_t vb = align256(3 * size_t(p->K) * p->K * kzpad * p->cb);  // line-padded vector
  const size_t nck = (nsteps + seg - 1) / seg;
  size_t b = (4 + nck + 4 * size_t(seg)) * vb + align256(size_t(p->K) * p->KZ * 8 * sizeof(double));
  if (nck > 1) b += (1 + 4 * size_t(seg)) * vb + align256(p->wsBytes);  // overlapped recompute
  if (adj_reg_z_rt(p->n)) {  // kept inverse-y output (u, w) per stage, per stage buffer
    const size_t sb = align256(6 * size_t(p->n) * p->n * kzpad * p->cb);
    b += (nck > 1 ? 2 : 1) * 4 * size_t(seg) * sb;
  }
  *bytes = (long long)b;
  return TG_OK;
}

int tg_adjoint_forward(tg_plan* p, const void* x0_hat, int nsteps, double dt, double nu, int objective, int seg,
                       double* J, void* ws, void* adj_ws, void* stream) {
  TG_RET(check_plan(p));
  if (p->mhd) return fail(TG_ERR_UNSUPPORTED, "adjoint: hydro plans only");
  if (!x0_hat || !J || !ws || !adj_ws) return fail(TG_ERR_ARG, "tg_adjoint_forward: NULL buffer");
  if (nsteps < 1 || seg < 1) return fail(TG_ERR_ARG, "adjoint: need nsteps, seg >= 1");
  if (objective != 0 && objective != 1) return fail(TG_ERR_ARG, "objective must be 0 (energy) or 1 (enstrophy)");
  return with_plan(p, [&](auto ops, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return decltype(ops)::adj_forward(p, static_cast<const cpx<R>*>(x0_hat), nsteps, dt, nu, objective, seg, J,
                                      ws, adj_ws, static_cast<hipStream_t>(stream));
  });
}

int tg_adjoint_backward(tg_plan* p, int nsteps, double dt, double nu, int seg, void* grad_hat, double* dJdnu,
                        void* ws, void* adj_ws, void* stream) {
  TG_RET(check_plan(p));
  if (p->mhd) return fail(TG_ERR_UNSUPPORTED, "adjoint: hydro plans only");
  if (!grad_hat || !dJdnu || !ws || !adj_ws) return fail(TG_ERR_ARG, "tg_adjoint_backward: NULL buffer");
  if (nsteps < 1 || seg < 1) return fail(TG_ERR_ARG, "adjoint: need nsteps, seg >= 1");
  return with_plan(p, [&](auto ops, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return decltype(ops)::adj_backward(p, nsteps, dt, nu, seg, static_cast<cpx<R>*>(grad_hat), dJdnu, ws, adj_ws,
                                       static_cast<hipStream_t>(stream));
  });
}

int tg_plan_variants(tg_plan* p, const int* set, int* get, int* count) {
  TG_RET(check_plan(p));
  const int nv = p->n >= 128 ? NVAR : 1;
  if (count) *count = nv;
  if (set) {
    for (int i = 0; i < TG_NPASS; ++i)
      if (set[i] < 0 || set[i] >= nv) return fail(TG_ERR_ARG, "tg_plan_variants: variant out of range");
    for (int i = 0; i < TG_NPASS; ++i) p->var[i] = set[i];
  }
  if (get)
    for (int i = 0; i < TG_NPASS; ++i) get[i] = p->var[i];
  return TG_OK;
}

int tg_timing_enable(tg_plan* p, int on) {
  TG_RET(check_plan(p));
  p->timing = on != 0;
  p->evused = 0;
  return TG_OK;
}

int tg_timing_read(tg_plan* p, double* ms, long long* count) {
  TG_RET(check_plan(p));
  if (!ms || !count) return fail(TG_ERR_ARG, "tg_timing_read: NULL output");
  for (int i = 0; i < TG_NPASS; ++i) { ms[i] = 0.0; count[i] = 0; }
  if (p->evused) TG_HIP(hipEventSynchronize(p->evpool[2 * p->evused - 1]));
  for (size_t i = 0; i < p->evused; ++i) {
    float t = 0.f;
    TG_HIP(hipEventElapsedTime(&t, p->evpool[2 * i], p->evpool[2 * i + 1]));
    ms[p->evpass[i]] += t;
    count[p->evpass[i]] += 1;
  }
  p->evused = 0;
  return TG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- ensembles
struct tg_batch {
  int ninst = 0, prec = 64;
  double ks = 1.0;
  std::vector<int> n;
  std::vector<double> nu;
  std::vector<long long> state_off;  // complex elements, ninst + 1 entries
  std::vector<size_t> ws_off;        // bytes, ninst + 1 entries
  std::vector<int> order;            // descriptor slot -> instance (grouped by size)
  std::vector<int> cls_n, cls_first, cls_count;  // size classes over descriptor slots
  std::vector<size_t> cls_prefix_off;            // offset of each class's EP_COUNT x (count+1) table
  std::vector<int> prefix;
  std::vector<void*> tw_by_size;     // twiddles per entry of kEnsSizes
  void* d_inst = nullptr;            // EnsInst<R>[ninst]
  int* d_prefix = nullptr;
  size_t cb = 16;
  bool bound = false;
  hipStream_t cap = nullptr;         // private capture stream
  std::vector<hipStream_t> branch;   // capture streams of size classes 1.. (graph branches)
  std::vector<hipEvent_t> fork;      // per size class: fork and join events
  hipGraphExec_t exec = nullptr;
  double graph_dt = 0.0;
};

namespace {

int ens_size_index(int n) {
  for (int i = 0; i < kNumEnsSizes; ++i)
    if (kEnsSizes[i] == n) return i;
  return -1;
}

template <typename R, int N>
int batch_launch_class(const tg_batch* b, int c, int st, double dt, hipStream_t s) {
  const int first = b->cls_first[c], cnt = b->cls_count[c];
  const auto* inst = static_cast<const EnsInst<R>*>(b->d_inst) + first;
  for (int pass = 0; pass < EP_COUNT; ++pass) {
    const size_t po = b->cls_prefix_off[c] + size_t(pass) * (cnt + 1);
    const int nblk = b->prefix[po + cnt];
    const int* pre = b->d_prefix + po;
    switch (pass) {
      case EP_INV_X:
        hipLaunchKernelGGL((kb_pass<R, N, EP_INV_X>), dim3(nblk), dim3(EnsBlock<R, N, EP_INV_X>::NT), 0, s, inst,
                           pre, cnt, R(dt), st);
        break;
      case EP_INV_Y:
        hipLaunchKernelGGL((kb_pass<R, N, EP_INV_Y>), dim3(nblk), dim3(EnsBlock<R, N, EP_INV_Y>::NT), 0, s, inst,
                           pre, cnt, R(dt), st);
        break;
      case EP_Z:
        hipLaunchKernelGGL((kb_pass<R, N, EP_Z>), dim3(nblk), dim3(EnsBlock<R, N, EP_Z>::NT), 0, s, inst, pre, cnt,
                           R(dt), st);
        break;
      case EP_FWD_Y:
        hipLaunchKernelGGL((kb_pass<R, N, EP_FWD_Y>), dim3(nblk), dim3(EnsBlock<R, N, EP_FWD_Y>::NT), 0, s, inst,
                           pre, cnt, R(dt), st);
        break;
      default:
        hipLaunchKernelGGL((kb_pass<R, N, EP_FWD_X>), dim3(nblk), dim3(EnsBlock<R, N, EP_FWD_X>::NT), 0, s, inst,
                           pre, cnt, R(dt), st);
        break;
    }
    TG_LAUNCHED();
  }
  return TG_OK;
}

// The four stages of one RK4 step of size class c on stream s.
template <typename R>
int batch_class_step(tg_batch* b, int c, double dt, hipStream_t s) {
  for (int st = 0; st < 4; ++st) {
    int rc = TG_OK;
    switch (b->cls_n[c]) {
      case 16: rc = batch_launch_class<R, 16>(b, c, st, dt, s); break;
      case 32: rc = batch_launch_class<R, 32>(b, c, st, dt, s); break;
      case 48: rc = batch_launch_class<R, 48>(b, c, st, dt, s); break;
      case 64: rc = batch_launch_class<R, 64>(b, c, st, dt, s); break;
      case 96: rc = batch_launch_class<R, 96>(b, c, st, dt, s); break;
      case 128: rc = batch_launch_class<R, 128>(b, c, st, dt, s); break;
      default: rc = fail(TG_ERR_UNSUPPORTED, "ensemble size class");
    }
    if (rc) return rc;
  }
  return TG_OK;
}

// One RK4 step of every instance.  Size classes are independent: with
// `branches` (the graph capture) class c >= 1 runs on its own stream forked
// from and joined back to s, so the captured graph holds one parallel
// branch per class; without, the classes run one after the other on s.
template <typename R>
int batch_launch_step(tg_batch* b, double dt, hipStream_t s, bool branches) {
  const int ncls = int(b->cls_n.size());
  if (!branches || ncls == 1) {
    for (int c = 0; c < ncls; ++c) TG_RET(batch_class_step<R>(b, c, dt, s));
    return TG_OK;
  }
  TG_HIP(hipEventRecord(b->fork[0], s));
  for (int c = 1; c < ncls; ++c) {
    TG_HIP(hipStreamWaitEvent(b->branch[c - 1], b->fork[0], 0));
    TG_RET(batch_class_step<R>(b, c, dt, b->branch[c - 1]));
    TG_HIP(hipEventRecord(b->fork[c], b->branch[c - 1]));
  }
  TG_RET(batch_class_step<R>(b, 0, dt, s));
  for (int c = 1; c < ncls; ++c) TG_HIP(hipStreamWaitEvent(s, b->fork[c], 0));
  return TG_OK;
}

template <typename R>
int batch_bind(tg_batch* b, void* state, void* ws) {
  std::vector<EnsInst<R>> h(b->ninst);
  char* w = static_cast<char*>(ws);
  cpx<R>* st = static_cast<cpx<R>*>(state);
  for (int slot = 0; slot < b->ninst; ++slot) {
    const int i = b->order[slot];
    const int N = b->n[i], M = N / 3, K = 2 * M + 1, KZ = M + 1;
    const size_t vec = size_t(3) * K * K * KZ * sizeof(cpx<R>);
    const size_t sA = size_t(6) * N * K * kzp_rt<R>(N) * sizeof(cpx<R>);
    char* base = w + b->ws_off[i];
    EnsInst<R>& I = h[slot];
    I.n = N;
    I.nu = R(b->nu[i]);
    I.ks = R(b->ks);
    I.u0 = st + b->state_off[i];
    I.u1 = reinterpret_cast<cpx<R>*>(base);
    I.acc = reinterpret_cast<cpx<R>*>(base + align256(vec));
    I.A = reinterpret_cast<cpx<R>*>(base + 2 * align256(vec));
    I.B = reinterpret_cast<cpx<R>*>(base + 2 * align256(vec) + align256(sA));
    I.tw = static_cast<const cpx<R>*>(b->tw_by_size[ens_size_index(N)]);
  }
  TG_HIP(hipMemcpy(b->d_inst, h.data(), h.size() * sizeof(EnsInst<R>), hipMemcpyHostToDevice));
  b->bound = true;
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_batch_create(int ninst, const int* n, const double* nu, int precision_bits, double domain_length,
                    tg_batch** out) {
  if (!out) return fail(TG_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (ninst <= 0 || !n || !nu) return fail(TG_ERR_ARG, "tg_batch_create: need ninst > 0, n and nu");
  if (precision_bits != 64 && precision_bits != 32) return fail(TG_ERR_ARG, "precision_bits must be 64 or 32");
  if (!(domain_length > 0.0)) return fail(TG_ERR_ARG, "domain_length must be positive");
  for (int i = 0; i < ninst; ++i) {
    if (ens_size_index(n[i]) < 0)
      return fail(TG_ERR_UNSUPPORTED, "ensemble grid size N=" + std::to_string(n[i]) + " has no batched plan");
    if (!(nu[i] >= 0.0)) return fail(TG_ERR_ARG, "viscosity must be nonnegative");
  }
  tg_batch* b = new tg_batch{};
  b->ninst = ninst;
  b->prec = precision_bits;
  b->cb = precision_bits == 64 ? 16 : 8;
  b->ks = domain_length == 2.0 * M_PI ? 1.0 : 2.0 * M_PI / domain_length;
  b->n.assign(n, n + ninst);
  b->nu.assign(nu, nu + ninst);
  b->state_off.assign(ninst + 1, 0);
  b->ws_off.assign(ninst + 1, 0);
  for (int i = 0; i < ninst; ++i) {
    const size_t N = n[i], M = N / 3, K = 2 * M + 1, KZ = M + 1;
    const size_t vec = 3 * K * K * KZ * b->cb;
    b->state_off[i + 1] = b->state_off[i] + (long long)(3 * K * K * KZ);
    const size_t kzpad = precision_bits == 64 ? size_t(kzp_rt<double>(n[i])) : size_t(kzp_rt<float>(n[i]));
    b->ws_off[i + 1] = b->ws_off[i] + 2 * align256(vec) + align256(6 * N * K * kzpad * b->cb) +
                       align256(6 * N * N * kzpad * b->cb);
  }
  for (int k = 0; k < kNumEnsSizes; ++k) {
    int cnt = 0;
    for (int i = 0; i < ninst; ++i)
      if (n[i] == kEnsSizes[k]) {
        if (!cnt) {
          b->cls_n.push_back(kEnsSizes[k]);
          b->cls_first.push_back(int(b->order.size()));
        }
        b->order.push_back(i);
        ++cnt;
      }
    if (cnt) b->cls_count.push_back(cnt);
  }
  for (size_t c = 0; c < b->cls_n.size(); ++c) {
    const int cnt = b->cls_count[c];
    b->cls_prefix_off.push_back(b->prefix.size());
    for (int pass = 0; pass < EP_COUNT; ++pass) {
      const int per = precision_bits == 64 ? ens_blocks<double>(b->cls_n[c], pass)
                                           : ens_blocks<float>(b->cls_n[c], pass);
      for (int j = 0; j <= cnt; ++j) b->prefix.push_back(j * per);
    }
  }
  auto cleanup = [&](hipError_t e, const char* what) {
    for (void* t : b->tw_by_size)
      if (t) (void)hipFree(t);
    if (b->d_inst) (void)hipFree(b->d_inst);
    if (b->d_prefix) (void)hipFree(b->d_prefix);
    delete b;
    return fail(TG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  b->tw_by_size.assign(kNumEnsSizes, nullptr);
  for (int k = 0; k < kNumEnsSizes; ++k) {
    const size_t N = kEnsSizes[k];
    std::vector<char> host(N * b->cb);
    for (size_t j = 0; j < N; ++j) {
      const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)N;
      if (b->prec == 64) {
        double v[2] = {double(cosl(ang)), double(sinl(ang))};
        memcpy(host.data() + j * 16, v, 16);
      } else {
        float v[2] = {float(cosl(ang)), float(sinl(ang))};
        memcpy(host.data() + j * 8, v, 8);
      }
    }
    hipError_t e = hipMalloc(&b->tw_by_size[k], host.size());
    if (e != hipSuccess) return cleanup(e, "hipMalloc(twiddles)");
    e = hipMemcpy(b->tw_by_size[k], host.data(), host.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(e, "hipMemcpy(twiddles)");
  }
  const size_t instBytes = size_t(ninst) * (b->prec == 64 ? sizeof(EnsInst<double>) : sizeof(EnsInst<float>));
  hipError_t e = hipMalloc(&b->d_inst, instBytes);
  if (e != hipSuccess) return cleanup(e, "hipMalloc(descriptors)");
  e = hipMalloc(reinterpret_cast<void**>(&b->d_prefix), b->prefix.size() * sizeof(int));
  if (e != hipSuccess) return cleanup(e, "hipMalloc(prefix)");
  e = hipMemcpy(b->d_prefix, b->prefix.data(), b->prefix.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) return cleanup(e, "hipMemcpy(prefix)");
  *out = b;
  return TG_OK;
}

int tg_batch_destroy(tg_batch* b) {
  if (!b) return TG_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  for (hipStream_t t : b->branch) (void)hipStreamDestroy(t);
  for (hipEvent_t e : b->fork) (void)hipEventDestroy(e);
  for (void* t : b->tw_by_size)
    if (t) (void)hipFree(t);
  if (b->d_inst) (void)hipFree(b->d_inst);
  if (b->d_prefix) (void)hipFree(b->d_prefix);
  delete b;
  return TG_OK;
}

int tg_batch_query(const tg_batch* b, long long* state_offsets, long long* ws_bytes) {
  if (!b) return fail(TG_ERR_ARG, "batch is NULL");
  if (state_offsets)
    for (int i = 0; i <= b->ninst; ++i) state_offsets[i] = b->state_off[i];
  if (ws_bytes) *ws_bytes = (long long)b->ws_off[b->ninst];
  return TG_OK;
}

int tg_batch_bind(tg_batch* b, void* state, void* ws) {
  if (!b) return fail(TG_ERR_ARG, "batch is NULL");
  if (!state || !ws) return fail(TG_ERR_ARG, "tg_batch_bind: NULL buffer");
  return b->prec == 64 ? batch_bind<double>(b, state, ws) : batch_bind<float>(b, state, ws);
}

int tg_batch_rk4_steps(tg_batch* b, int nsteps, double dt, int use_graph, void* stream) {
  if (!b) return fail(TG_ERR_ARG, "batch is NULL");
  if (!b->bound) return fail(TG_ERR_ARG, "tg_batch_rk4_steps: call tg_batch_bind first");
  if (nsteps < 0) return fail(TG_ERR_ARG, "nsteps must be >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto step = [&](hipStream_t st, bool branches) {
    return b->prec == 64 ? batch_launch_step<double>(b, dt, st, branches)
                         : batch_launch_step<float>(b, dt, st, branches);
  };
  if (!use_graph) {
    for (int i = 0; i < nsteps; ++i) TG_RET(step(s, false));
    return TG_OK;
  }
  if (!b->exec || b->graph_dt != dt) {
    if (b->exec) {
      TG_HIP(hipGraphExecDestroy(b->exec));
      b->exec = nullptr;
    }
    if (!b->cap) TG_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    const size_t ncls = b->cls_n.size();
    while (b->branch.size() + 1 < ncls) {
      hipStream_t t = nullptr;
      TG_HIP(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
      b->branch.push_back(t);
    }
    while (b->fork.size() < ncls) {
      hipEvent_t e = nullptr;
      TG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      b->fork.push_back(e);
    }
    hipGraph_t g = nullptr;
    TG_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    const int rc = step(b->cap, true);
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(TG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiateWithFlags(&b->exec, g, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(TG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    b->graph_dt = dt;
  }
  for (int i = 0; i < nsteps; ++i) TG_HIP(hipGraphLaunch(b->exec, s));
  return TG_OK;
}

int tg_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return TG_ERR_ARG;
  snprintf(buf, n, "%s", g_err.c_str());
  return TG_OK;
}

}  // extern "C"
