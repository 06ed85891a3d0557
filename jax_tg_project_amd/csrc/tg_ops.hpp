// Launchers and host-side pass sequences of one (precision, grid size):
// the RHS passes, RK4 loop, transforms, diagnostics and adjoint sweeps.  Each
// (precision, N) is instantiated in its own translation unit (tg_inst.hip)
// and exported as an OpsTable; the C ABI (tgdns.hip) dispatches through it.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

#include "tg_common.hpp"
#include "tg_passes.hpp"

namespace tg {

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
      dim3 grid(cdiv(G::KZ, C::L), G::K, mode == XM_MHD5 ? 2 : 1), blk(C::NT);
      switch (mode) {
        case XM_HYDRO: hipLaunchKernelGGL((k_inv_x<R, N, XM_HYDRO, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_MHD: hipLaunchKernelGGL((k_inv_x<R, N, XM_MHD, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_COPY3: hipLaunchKernelGGL((k_inv_x<R, N, XM_COPY3, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_CURL3: hipLaunchKernelGGL((k_inv_x<R, N, XM_CURL3, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_ADJ: hipLaunchKernelGGL((k_inv_x<R, N, XM_ADJ, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_ADJQ: hipLaunchKernelGGL((k_inv_x<R, N, XM_ADJQ, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_HYDRO5: hipLaunchKernelGGL((k_inv_x<R, N, XM_HYDRO5, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_MHD5: hipLaunchKernelGGL((k_inv_x<R, N, XM_MHD5, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
        case XM_GRAD3: hipLaunchKernelGGL((k_inv_x<R, N, XM_GRAD3, V>), grid, blk, 0, s, su, sb, out, tw, ks, al, cpz); break;
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
  // Forward y of the adjoint z pass's kz-folded fields into the five fields
  // of the register forward x (k_y_adj).
  static int y_adj(int var, const cpx<R>* in, cpx<R>* out, const cpx<R>* tw, R ks, hipStream_t s) {
    return with_variant<N>(var, [&](auto vc) {
      constexpr int V = decltype(vc)::value;
      using C = CfgXY<R, N, V>;
      hipLaunchKernelGGL((k_y_adj<R, N, V>), dim3(cdiv(G::KZ, C::L), N, 4), dim3(C::NT), 0, s, in, out, tw, ks);
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
  // N >= 128 variants 0 / 1 hold ZwDefault E / E_ALT points per lane
  // (variant 2 is the LDS-resident kernel); smaller grids hold ZwDefault E
  // points per lane, one wavefront per block, or (E = 0) run the LDS kernel.
  static int z(int mode, int var, const cpx<R>* in, cpx<R>* out, const R* pin, R* pout, const cpx<R>* tw, R lor,
               hipStream_t s) {
    switch (mode) {
      case ZM_HYDRO:
        if constexpr (N >= 128) {
          if (var == 0) return zw<ZwDefault<N>::E, 4>(in, out, tw, s);
          if (var == 1) return zw<ZwDefault<N>::E_ALT, 4>(in, out, tw, s);
          return z_<ZM_HYDRO, 2>(in, out, pin, pout, tw, lor, s);
        } else if constexpr (ZwDefault<N>::E > 0) {
          return zw<ZwDefault<N>::E, ZwDefault<N>::WPB>(in, out, tw, s);
        } else {
          return z_<ZM_HYDRO, 0>(in, out, pin, pout, tw, lor, s);
        }
      case ZM_MHD:
        // variant 0 of the power-of-two grids >= 32 is the register-resident
        // kernel at 8 points per lane; variant 1 of those up to 256 the same
        // at 4 points per lane (N / 4 lanes per line fit a wavefront)
        if constexpr (adj_reg_z<N>()) {
          if (var == 0) return zw_mhd<8>(in, out, tw, lor, s);
          if constexpr (N <= 256 && num_variants<N>() > 1) {
            if (var == 1) return zw_mhd<4>(in, out, tw, lor, s);
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
  template <int E>
  static int zw_mhd(const cpx<R>* in, cpx<R>* out, const cpx<R>* tw, R lor, hipStream_t s) {
    constexpr int WPB = ZwDefault<N>::WPB;
    using C = CfgZW<R, N, E, WPB>;
    static_assert((N * N) % C::LINES == 0, "lines per block must divide N^2");
    hipLaunchKernelGGL((k_zw_mhd<R, N, E, WPB>), dim3(N * N / C::LINES), dim3(C::NT), 0, s, in, out, tw, lor);
    TG_LAUNCHED();
    return TG_OK;
  }
  // Register-resident adjoint z pass: u, w from fields 0-5 of `in`, q from
  // fields 0-2 of `qin`.
  // fold: store the kz-folded fields k_y_adj reads (see ZWAdjBody).
  static int z_adj(const cpx<R>* in, const cpx<R>* qin, cpx<R>* out, const cpx<R>* tw, hipStream_t s,
                   bool fold = false, R ks = R(1)) {
    if constexpr (adj_reg_z<N>()) {
      constexpr int WPB = ZwDefault<N>::WPB;
      using C = CfgZW<R, N, adj_ze<R>(), WPB>;
      if (fold)
        hipLaunchKernelGGL((k_zw_adj<R, N, WPB, true>), dim3(N * N / C::LINES), dim3(C::NT), 0, s, in, qin, out, tw,
                           ks);
      else
        hipLaunchKernelGGL((k_zw_adj<R, N, WPB, false>), dim3(N * N / C::LINES), dim3(C::NT), 0, s, in, qin, out,
                           tw, ks);
      TG_LAUNCHED();
      return TG_OK;
    } else {
      return fail(TG_ERR_UNSUPPORTED, "register adjoint z pass: N must be a power of two >= 32");
    }
  }
  static int nmax() { return (N / CfgZM<R, N, ZM_MAXABS, 0>::LZ) * N; }
  // RK4 stage variant 1 of the power-of-two grids N >= 128 (hydro and MHD)
  // is the register-epilogue kernel; so is every adjoint launch on those grids.
  static int fwd_x(int mode, int var, const cpx<R>* in, const StageArgs<R>& a, const cpx<R>* tw, hipStream_t s) {
    if constexpr (use_xr<N>()) {
      if (mode == FM_ADJ || mode == FM_ADJ5) {
        using C = CfgXR<R, N, true>;
        if (mode == FM_ADJ5)
          hipLaunchKernelGGL((k_fwd_xr_adj<R, N, true>), dim3(cdiv(G::KZ, C::L), G::K), dim3(C::NT), 0, s, in, a, tw);
        else
          hipLaunchKernelGGL((k_fwd_xr_adj<R, N, false>), dim3(cdiv(G::KZ, C::L), G::K), dim3(C::NT), 0, s, in, a,
                             tw);
        TG_LAUNCHED();
        return TG_OK;
      }
      if ((mode == FM_HYDRO_STAGE || mode == FM_MHD_STAGE) && var == 1) {
        using C = CfgXR<R, N>;
        const dim3 grid(cdiv(G::KZ, C::L), G::K, mode == FM_MHD_STAGE ? 2 : 1), blk(C::NT);
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
    if constexpr (use_xr<N>()) return cdiv(G::KZ, CfgXR<R, N, true>::L) * G::K;
    else return cdiv(G::KZ, CfgXF<R, N, 0>::LINES / 6) * G::K;
  }
  static int diag(const cpx<R>* su, const cpx<R>* sb, double* part, int nblk, double ks, hipStream_t s) {
    hipLaunchKernelGGL((k_diag<R>), dim3(nblk), dim3(DIAG_NT), 0, s, su, sb, part, ks, G::M);
    TG_LAUNCHED();
    return TG_OK;
  }
};


inline int timed_begin(const tg_plan* cp, int pass, hipStream_t s) {
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

inline int timed_end(const tg_plan* cp, hipStream_t s) {
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

template <typename R, int N>
// bslot (hydro only): where the inverse-y output (u, w) goes and the z pass
// reads it, instead of the workspace B (the adjoint keeps it per stage).
int rhs_passes(const tg_plan* p, const Ptrs<R>& w, const cpx<R>* us, const cpx<R>* bs, int fmode,
               const StageArgs<R>& a, R lor, hipStream_t s, cpx<R>* bslot = nullptr, const int* vars = nullptr) {
  const int cpz = a.cpz ? a.cpz : Geo<N>::KZ;
  using Lc = Launch<R, N>;
  const cpx<R>* tw = static_cast<const cpx<R>*>(p->tw);
  const bool mhd = p->mhd != 0;
  const int fo = mhd ? 6 : 3;
  const int* v = vars ? vars : p->var;
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
    const int xm = op == TG_OP_CURL ? XM_CURL3 : op == TG_OP_DIV ? XM_DIV : op == TG_OP_GRAD ? XM_GRAD3 : XM_COPY3;
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
  // cotangent's projection q is transformed.  q_in_A: the previous stage
  // VJP's forward x already left q's inverse-x output in w.A (a.qout of that
  // launch), so the first pass is skipped.
  static int vjp_passes(const tg_plan* p, const Ptrs<R>& w, const cpx<R>* x, const StageArgs<R>& a,
                        hipStream_t s, const cpx<R>* uw = nullptr, bool q_in_A = false) {
    using Lc = Launch<R, N>;
    const cpx<R>* tw = static_cast<const cpx<R>*>(p->tw);
    const int cpz = a.cpz ? a.cpz : Geo<N>::KZ;  // row pitch of x, a and the epilogue vectors
    if (uw) {
      if (!q_in_A)
        TG_TIMED(TG_PASS_INV_X, Lc::inv_x(XM_ADJQ, p->var[TG_PASS_INV_X], a.a, nullptr, w.A, tw, R(p->ks), s,
                                          a.alpha, cpz));
      TG_TIMED(TG_PASS_INV_Y, Lc::y(+1, p->var[TG_PASS_INV_Y], 3, w.A, w.B, tw, s));
      if constexpr (use_xr<N>()) {
        // i k x p2 folded along the way: kz in the z pass, ky in k_y_adj, kx
        // in the register forward x, which then transforms five fields
        TG_TIMED(TG_PASS_Z, Lc::z_adj(uw, w.B, w.A, tw, s, true, R(p->ks)));
        TG_TIMED(TG_PASS_FWD_Y, Lc::y_adj(p->var[TG_PASS_FWD_Y], w.A, w.B, tw, R(p->ks), s));
        TG_TIMED(TG_PASS_FWD_X, Lc::fwd_x(FM_ADJ5, 0, w.B, a, tw, s));
      } else {
        TG_TIMED(TG_PASS_Z, Lc::z_adj(uw, w.B, w.A, tw, s));
        TG_TIMED(TG_PASS_FWD_Y, Lc::y(-1, p->var[TG_PASS_FWD_Y], 6, w.A, w.B, tw, s));
        TG_TIMED(TG_PASS_FWD_X, Lc::fwd_x(FM_ADJ, 0, w.B, a, tw, s));
      }
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
  // The forward sweep keeps the last kKeptSegments segments: each one past
  // the second has a stage buffer of its own (4 seg stage states, then its
  // slots) after those, when there are that many segments.
  static constexpr int kKeptSegments = kAdjKeptSegments;
  struct AdjLayout {
    cpx<R>* cur; cpx<R>* lamp; cpx<R>* lacc; cpx<R>* abuf; cpx<R>* ck; cpx<R>* st; double* nupart;
    cpx<R>* cur2; cpx<R>* st2; void* ws2;
    cpx<R>* slots; cpx<R>* slots2;
    cpx<R>* stk[kKeptSegments]; cpx<R>* slk[kKeptSegments];  // stage buffer k: states, slots
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
    char* extra = third + (adj_reg_z<N>() ? size_t(L.nck > 1 ? 2 : 1) * 4 * size_t(seg) * sb : 0);
    L.stk[0] = L.st;
    L.slk[0] = L.slots;
    L.stk[1] = L.st2;
    L.slk[1] = L.slots2;
    for (int k = 2; k < kKeptSegments; ++k) {
      const bool has = L.nck > k;
      L.stk[k] = has ? reinterpret_cast<cpx<R>*>(extra) : nullptr;
      L.slk[k] = (has && adj_reg_z<N>()) ? reinterpret_cast<cpx<R>*>(extra + 4 * size_t(seg) * vb) : nullptr;
      if (has) extra += 4 * size_t(seg) * (vb + (adj_reg_z<N>() ? sb : 0));
    }
    L.sstride = sb / sizeof(cpx<R>);
    L.vstride = vb / sizeof(cpx<R>);
    L.nblk = Launch<R, N>::adj_x_blocks();
    return L;
  }
  // Segments of the checkpoint schedule: the first one takes the remainder
  // of nsteps over seg, every later one seg steps, so the two segments the
  // forward sweep keeps (the last two) are whole.
  static int seg_len(int nsteps, int seg, int c) {
    const int r = nsteps - ((nsteps + seg - 1) / seg - 1) * seg;
    return c == 0 ? r : seg;
  }
  // `len` RK4 steps from cur keeping the four stage states of step j in
  // st[4j .. 4j+3] (stage 1 = the step's input state); cur ends as the
  // state after the last step.
  static int steps_keeping_stages(const tg_plan* p, cpx<R>* cur, cpx<R>* st, size_t vstride, int len, double dt,
                                  double nu, void* ws, hipStream_t s, cpx<R>* slots = nullptr,
                                  size_t sstride = 0) {
    const size_t bytes = size_t(3) * p->K * p->K * kzp<R, N>() * sizeof(cpx<R>);
    if (len > 0) TG_HIP(hipMemcpyAsync(st, cur, bytes, hipMemcpyDeviceToDevice, s));
    for (int j = 0; j < len; ++j) TG_RET(keep_step(p, cur, st, vstride, j, len, dt, nu, ws, s, slots, sstride));
    return TG_OK;
  }
  // Step j of steps_keeping_stages (after its copy st[0] <- cur): it starts
  // from st[4j]; its last stage writes the next step's input slot directly
  // (cur only after the last step).
  // vars: tile variants of its passes (nullptr: the plan's).
  static int keep_step(const tg_plan* p, cpx<R>* cur, cpx<R>* st, size_t vstride, int j, int len, double dt,
                       double nu, void* ws, hipStream_t s, cpx<R>* slots, size_t sstride,
                       const int* vars = nullptr) {
    Ptrs<R> w = carve<R>(p, ws);
    cpx<R>* S = st + size_t(4 * j) * vstride;
    for (int stg = 0; stg < 4; ++stg) {
      StageArgs<R> a{};
      a.u0 = S;
      a.us = S + size_t(stg) * vstride;
      a.unext = stg < 3 ? S + size_t(stg + 1) * vstride : (j + 1 < len ? S + size_t(4) * vstride : cur);
      a.uacc = w.Acc;
      a.dt = R(dt); a.nu = R(nu); a.ks = R(p->ks); a.stage = stg;
      a.cpz = kzp<R, N>();
      cpx<R>* slot = slots ? slots + size_t(4 * j + stg) * sstride : nullptr;
      TG_RET(rhs_passes<R, N>(p, w, a.us, nullptr, FM_HYDRO_STAGE, a, R(0), s, slot, vars));
    }
    return TG_OK;
  }
  // Stage buffer (states, slots) of segment c in the reverse sweep: the
  // i-th last segment (i = nck - 1 - c) in buffer i while i < kKeptSegments,
  // then the first and the last kept buffer in turn (segment c - 1 is
  // recomputed into the buffer whose segment was reversed before segment c,
  // while segment c is reversed).
  static std::pair<cpx<R>*, cpx<R>*> stage_buffer(const AdjLayout& L, int c) {
    const int i = L.nck - 1 - c;
    if (i < kKeptSegments) return {L.stk[i], L.slk[i]};
    const int k = (i - (kKeptSegments - 1)) % 2 == 0 ? kKeptSegments - 1 : 0;
    return {L.stk[k], L.slk[k]};
  }
  // Forward sweep: nsteps RK4 steps from x0 with a checkpoint every `seg`
  // steps, J = E(T) (objective 0) or Z(T) (1), and the terminal cotangent.
  // The stage states (and kept inverse-y outputs) of the last kKeptSegments
  // segments go to their stage buffers (stage_buffer).
  static int adj_forward(const tg_plan* p, const cpx<R>* x0, int nsteps, double dt, double nu, int objective,
                         int seg, double* J, void* ws, void* aws, hipStream_t s) {
    AdjLayout L = adj_layout(p, nsteps, seg, aws);
    constexpr int KZP = kzp<R, N>(), KZ = Geo<N>::KZ;
    const size_t rows = size_t(3) * p->K * p->K, e = sizeof(cpx<R>);
    const size_t bytes = rows * KZP * e;
    TG_HIP(hipMemcpy2DAsync(L.cur, KZP * e, x0, KZ * e, KZ * e, rows, hipMemcpyDeviceToDevice, s));
    for (int c = 0; c < L.nck; ++c) {
      TG_HIP(hipMemcpyAsync(L.ck + c * L.vstride, L.cur, bytes, hipMemcpyDeviceToDevice, s));
      const int len = seg_len(nsteps, seg, c);
      if (c < L.nck - kKeptSegments) {
        TG_RET(rk4_padded(p, L.cur, nullptr, len, dt, nu, 0.0, -1.0, ws, s));
      } else {  // kept: the reverse sweep never recomputes it
        const auto b = stage_buffer(L, c);
        TG_RET(steps_keeping_stages(p, L.cur, b.first, L.vstride, len, dt, nu, ws, s, b.second, L.sstride));
      }
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
    // The stages of the last kKeptSegments segments are kept by the forward
    // sweep (stage_buffer).  Segment c - 1 is recomputed on the private stream
    // into its buffer while segment c is reversed on `s`; events order the
    // buffer reuse both ways.
    hipStream_t s2 = nullptr;
    tg_plan* mp = const_cast<tg_plan*>(p);
    auto ev_rec = [&](int c) { return mp->adjev[1 + 2 * size_t(c)]; };  // recompute of segment c done
    auto ev_rev = [&](int c) { return mp->adjev[2 + 2 * size_t(c)]; };  // reverse of segment c done
    if (L.nck > 1) {
      if (!mp->aux) TG_HIP(hipStreamCreateWithFlags(&mp->aux, hipStreamNonBlocking));
      while (mp->adjev.size() < 1 + 2 * size_t(L.nck)) {
        hipEvent_t e = nullptr;
        TG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        mp->adjev.push_back(e);
      }
      s2 = mp->aux;
      TG_HIP(hipEventRecord(mp->adjev[0], s));  // checkpoints written, nupart cleared
      TG_HIP(hipStreamWaitEvent(s2, mp->adjev[0], 0));
    }
    // the recompute's tile variants: chosen for running beside the reverse
    // sweep's kernels (smaller register tiles share the CUs better)
    const int* rv = p->rvar[0] >= 0 ? p->rvar : p->var;
    // On grids with the register forward x, every stage VJP but the last
    // also does the next one's inverse x of q (StageArgs::qout), so only the
    // first stage VJP of the sweep runs that pass on its own.
    bool q_in_A = false;
    for (int c = L.nck - 1; c >= 0; --c) {
      const int len = seg_len(nsteps, seg, c);
      const auto bc = stage_buffer(L, c);
      cpx<R>* stc = bc.first;
      cpx<R>* slc = bc.second;
      // segment c recomputed (the last ones were kept by the forward sweep)
      if (c < L.nck - kKeptSegments) TG_HIP(hipStreamWaitEvent(s, ev_rec(c), 0));
      // recompute segment c - 1 into its buffer (free: its last user was
      // reversed earlier); its steps are enqueued one per reversed step of
      // segment c, so the host feeds both streams together instead of
      // queueing the whole recompute first
      const bool rec = c > 0 && c - 1 < L.nck - kKeptSegments;
      const auto bn = rec ? stage_buffer(L, c - 1) : std::make_pair(static_cast<cpx<R>*>(nullptr),
                                                                     static_cast<cpx<R>*>(nullptr));
      cpx<R>* stn = bn.first;
      cpx<R>* sln = bn.second;
      const int lenn = rec ? seg_len(nsteps, seg, c - 1) : 0;
      int jr = 0;
      if (rec) {
        TG_HIP(hipMemcpyAsync(L.cur2, L.ck + (c - 1) * L.vstride, bytes, hipMemcpyDeviceToDevice, s2));
        TG_HIP(hipMemcpyAsync(stn, L.cur2, bytes, hipMemcpyDeviceToDevice, s2));
      }
      for (int j = len - 1; j >= 0; --j) {
        for (; jr < lenn && jr * len <= (len - 1 - j) * lenn; ++jr)
          TG_RET(keep_step(p, L.cur2, stn, L.vstride, jr, lenn, dt, nu, L.ws2, s2, sln, L.sstride, rv));
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
          const bool fuse = use_xr<N>() && uw && !(c == 0 && j == 0 && sidx == 0);
          if (fuse) {
            a.qout = w.A;
            a.qalpha = sidx > 0 ? R(1) : cl[3];  // the next VJP's cotangent: abuf, or lamp of the step before
          }
          TG_RET(vjp_passes(p, w, a.xs, a, s, uw, q_in_A));
          q_in_A = fuse;
        }
        std::swap(lamp, lacc);
      }
      for (; jr < lenn; ++jr)
        TG_RET(keep_step(p, L.cur2, stn, L.vstride, jr, lenn, dt, nu, L.ws2, s2, sln, L.sstride, rv));
      if (rec) TG_HIP(hipEventRecord(ev_rec(c - 1), s2));
      if (c > 1) {
        // a later recompute reuses this buffer: after this reverse
        TG_HIP(hipEventRecord(ev_rev(c), s));
        TG_HIP(hipStreamWaitEvent(s2, ev_rev(c), 0));
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
  // Scratch bytes diagnostics_dev needs (partials, sums, maxima and, with
  // divmax, one field of each inverse-transform intermediate).
  static size_t diag_scratch_bytes(const tg_plan* p, int divmax) {
    size_t b = align256(size_t(p->nDiagBlk) * NDIAG * sizeof(double)) + align256((NDIAG + 2) * sizeof(double)) +
               align256(2 * size_t(p->nMax) * sizeof(R));
    if (divmax)
      b += align256(size_t(N) * p->K * kzp<R, N>() * sizeof(cpx<R>)) +
           align256(size_t(N) * N * kzp<R, N>() * sizeof(cpx<R>));
    return b;
  }
  // tg_diagnostics with every intermediate in `scratch` and the normalised
  // results written to the device array out_dev: no allocation, no host copy,
  // no synchronisation.
  static int diagnostics_dev(const tg_plan* p, const cpx<R>* u, const cpx<R>* b, int divmax, double* out_dev,
                             void* scratch, hipStream_t s) {
    using Lc = Launch<R, N>;
    const cpx<R>* tw = static_cast<const cpx<R>*>(p->tw);
    char* c = static_cast<char*>(scratch);
    double* part = reinterpret_cast<double*>(c);
    c += align256(size_t(p->nDiagBlk) * NDIAG * sizeof(double));
    double* sums = reinterpret_cast<double*>(c);
    c += align256((NDIAG + 2) * sizeof(double));
    R* maxu = reinterpret_cast<R*>(c);
    R* maxb = maxu + p->nMax;
    c += align256(2 * size_t(p->nMax) * sizeof(R));
    cpx<R>* A = reinterpret_cast<cpx<R>*>(c);
    c += align256(size_t(N) * p->K * kzp<R, N>() * sizeof(cpx<R>));
    cpx<R>* B = reinterpret_cast<cpx<R>*>(c);
    TG_RET(Lc::diag(u, b, part, p->nDiagBlk, p->ks, s));
    if (divmax) {
      TG_RET(Lc::inv_x(XM_DIV, 0, u, nullptr, A, tw, R(p->ks), s));
      TG_RET(Lc::y(+1, 0, 1, A, B, tw, s));
      TG_RET(Lc::z(ZM_MAXABS, 0, B, nullptr, nullptr, maxu, tw, R(0), s));
      if (b) {
        TG_RET(Lc::inv_x(XM_DIV, 0, b, nullptr, A, tw, R(p->ks), s));
        TG_RET(Lc::y(+1, 0, 1, A, B, tw, s));
        TG_RET(Lc::z(ZM_MAXABS, 0, B, nullptr, nullptr, maxb, tw, R(0), s));
      }
    }
    hipLaunchKernelGGL((k_diag_final<R>), dim3(1), dim3(64), 0, s, part, p->nDiagBlk, divmax ? maxu : nullptr,
                       (divmax && b) ? maxb : nullptr, divmax ? p->nMax : 0, sums);
    TG_LAUNCHED();
    hipLaunchKernelGGL(k_diag_normalize<0>, dim3(1), dim3(64), 0, s, sums, p->n, out_dev);
    TG_LAUNCHED();
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
    out[8] = host[NDIAG];      // max |div u|
    out[9] = host[NDIAG + 1];  // max |div b|
    out[10] = host[8];         // non-finite stored values
    return TG_OK;
  }
};


// Entry points of one (precision, N), filled by make_table in the
// translation unit that instantiates them.
template <typename R>
struct OpsTable {
  int n, kzp, nmax;
  int (*forward)(const tg_plan*, const R*, cpx<R>*, void*, hipStream_t);
  int (*inverse)(const tg_plan*, const cpx<R>*, int, R*, void*, hipStream_t);
  int (*rhs)(const tg_plan*, const cpx<R>*, const cpx<R>*, double, double, double, cpx<R>*, cpx<R>*, void*,
             hipStream_t);
  int (*rk4)(const tg_plan*, cpx<R>*, cpx<R>*, int, double, double, double, double, void*, hipStream_t);
  int (*vjp)(const tg_plan*, const cpx<R>*, const cpx<R>*, double, cpx<R>*, void*, hipStream_t);
  int (*adj_forward)(const tg_plan*, const cpx<R>*, int, double, double, int, int, double*, void*, void*,
                     hipStream_t);
  int (*adj_backward)(const tg_plan*, int, double, double, int, cpx<R>*, double*, void*, void*, hipStream_t);
  int (*diagnostics)(const tg_plan*, const cpx<R>*, const cpx<R>*, int, double*, void*, hipStream_t);
  size_t (*diag_scratch_bytes)(const tg_plan*, int);
  int (*diagnostics_dev)(const tg_plan*, const cpx<R>*, const cpx<R>*, int, double*, void*, hipStream_t);
};

template <typename R, int N>
OpsTable<R> make_table() {
  using O = Ops<R, N>;
  OpsTable<R> t;
  t.n = N;
  t.kzp = kzp<R, N>();
  t.nmax = Launch<R, N>::nmax();
  t.forward = &O::forward;
  t.inverse = &O::inverse;
  t.rhs = &O::rhs;
  t.rk4 = &O::rk4;
  t.vjp = &O::vjp;
  t.adj_forward = &O::adj_forward;
  t.adj_backward = &O::adj_backward;
  t.diagnostics = &O::diagnostics;
  t.diag_scratch_bytes = &O::diag_scratch_bytes;
  t.diagnostics_dev = &O::diagnostics_dev;
  return t;
}

}  // namespace tg
