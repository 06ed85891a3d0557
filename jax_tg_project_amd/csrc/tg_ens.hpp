// Pass bodies of the batched ensemble on a dense layout.
//
// The grids of an ensemble are small (N <= 128), so their kz rows are short
// (KZ = N/3 + 1: 11 to 43 columns) and padding each row to whole 128-byte
// lines, as the single-grid passes do, fetches up to 1.9x the data.  The
// ensemble instead keeps every row dense (pitch KZ) and pads only whole
// planes to a line boundary:
//
//   state, stage input, accumulator   S[c][ix][PA]   row iy at iy * KZ
//   inverse-x out, forward-y out      A[f][x][PA]    row iy at iy * KZ
//   inverse-y out, z out              B[f][x][PB]    row y  at y  * KZ
//
// with PA = K * KZ rounded up to a line and PB = N * KZ (already a whole
// number of lines).  The x passes see an x-plane as ONE flat index
// col = iy * KZ + iz and transform tiles of LX consecutive flat columns
// (whole lines, line-aligned); the y passes give a block PPB whole planes
// (every kz column of each); the z pass reads dense lines (pitch KZ).  Every
// access is then a full, aligned line except the last tile of a plane.
#pragma once
#include "tg_passes.hpp"

namespace tg {

template <typename R, int N>
struct EnsCfg {
  using G = Geo<N>;
  static constexpr int PL = 128 / int(2 * sizeof(R));  // complex elements per 128-byte line
  static constexpr int K = G::K, KZ = G::KZ;
  static constexpr int PA = (K * KZ + PL - 1) / PL * PL;
  static constexpr int PB = N * KZ;
  static_assert(PB % PL == 0, "N * KZ is a whole number of lines for N a multiple of 16");
  static constexpr int E = base_e<N>();  // points per thread: 12 (48, 96), 16 (128), else 8
  static constexpr int TPL = N / E;      // threads per line
  // x passes: LX flat columns (whole lines) per block, 256 threads
  static constexpr int LX = 256 / TPL;
  static_assert(LX % PL == 0 && LX * TPL == 256, "x tiles are whole lines");
  static constexpr int NTX = 256;
  static constexpr int TX = (K * KZ + LX - 1) / LX;  // x tiles per plane
  // y passes: PPB whole planes per block, the largest divisor of N that
  // keeps the block within 256 threads (at least one plane)
  static constexpr int ppb() {
    int best = 1;
    for (int p = 1; p <= N; ++p)
      if (N % p == 0 && p * KZ * TPL <= 256) best = p;
    return best;
  }
  static constexpr int PPB = ppb();
  static constexpr int LY = PPB * KZ;  // lines (plane, kz) per y block
  static constexpr int NTY = LY * TPL;
};

// ---------------------------------------------------------------- inverse x
// Flat tile `bt` of an x-plane: the five fields (u, i kx u_y, i kx u_z) of
// the stage input s (components at c * K * PA) along x, zero-padded to N,
// inverse-transformed (1/N^3 applied) into A[f][x][PA].
template <typename R, int N>
struct EnsInvX {
  using G = Geo<N>;
  using C = EnsCfg<R, N>;
  static constexpr int SMEM = N * C::LX + N;
  __device__ __forceinline__ static void run(cpx<R>* sm, int bt, const cpx<R>* __restrict__ s,
                                             cpx<R>* __restrict__ out, const cpx<R>* __restrict__ tw, R ks) {
    constexpr int L = C::LX, E = C::E, TPL = C::TPL, PA = C::PA;
    cpx<R>* buf = sm;
    cpx<R>* tws = sm + N * L;
    const size_t cs = size_t(G::K) * PA;
    const R scale = R(1) / (R(N) * R(N) * R(N));
    stage_twiddles<R, N>(tws, tw);
    const int tid = threadIdx.x, l = tid % L, t = tid / L;
    const bool act = t < TPL;
    const int colr = bt * L + l;
    const bool colok = colr < G::K * G::KZ;
    const int col = colok ? colr : G::K * G::KZ - 1;
    using Pl = typename PlanFor<N, E>::type;
    constexpr int RAD0 = FirstRadix<Pl>::value;
    constexpr int NB0 = E / RAD0, SIN0 = N / RAD0;
    cpx<R> sv[3][E];  // the three components at the thread's first-stage points
#pragma unroll
    for (int j = 0; j < NB0; ++j) {
#pragma unroll
      for (int m = 0; m < RAD0; ++m) {
        const int n = (act ? t : 0) * NB0 + j + m * SIN0;
        const int ix = G::kept(n) ? G::compact_of(n) : G::M;
        const size_t pt = size_t(ix) * PA + col;
#pragma unroll
        for (int c = 0; c < 3; ++c) sv[c][j * RAD0 + m] = s[c * cs + pt];
      }
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      if (f) __syncthreads();  // the previous field's last LDS reads are done
      auto ld = [&](int n, int, int slot) -> cpx<R> {
        const bool ok = colok && G::kept(n);
        const R kx = ks * R(G::kint(ok ? G::compact_of(n) : 0));
        const cpx<R> v = cscale(xfield5<R>(f, sv[0][slot], sv[1][slot], sv[2][slot], kx), scale);
        return ok ? v : cpx<R>{R(0), R(0)};
      };
      auto st = [&](int n, int, cpx<R> v, int) {
        if (colok) out[(size_t(f) * N + n) * PA + colr] = v;
      };
      fft_io<R, N, E, +1, 1, true>(buf, L, L, tws, ld, st);
    }
  }
};

// ---------------------------------------------------------------- inverse y
// Output field F of the planes x0 .. x0 + PPB - 1: from the five fields
// (v, D_y, D_z) of A (K rows each, pitch KZ) it forms v_F (F < 3) or the
// curl component i k x v (F >= 3, completed from D_y = i kx v_y and
// D_z = i kx v_z while ky and kz are spectral; see YCurlBody), zero-pads
// y to N, inverse-transforms into B[F][x][PB].
template <typename R, int N>
struct EnsInvY {
  using G = Geo<N>;
  using C = EnsCfg<R, N>;
  static constexpr int SMEM = N * C::LY + N;
  template <int F>
  __device__ __forceinline__ static void field(cpx<R>* sm, int x0, const cpx<R>* __restrict__ a,
                                               cpx<R>* __restrict__ b, const cpx<R>* __restrict__ tws, R ks) {
    constexpr int LY = C::LY, E = C::E, PA = C::PA, PB = C::PB, KZ = G::KZ;
    constexpr size_t AF = size_t(N) * PA, BF = size_t(N) * PB;
    auto ld = [&](int n, int l, int) -> cpx<R> {
      const int p = l / KZ, iz = l - p * KZ;
      const bool ok = G::kept(n);
      const int iy = ok ? G::compact_of(n) : 0;
      const size_t pt = size_t(x0 + p) * PA + size_t(iy) * KZ + iz;
      const R ky = ks * R(G::kint(iy)), kz = ks * R(iz);
      cpx<R> v;
      if constexpr (F < 3) {
        v = a[F * AF + pt];
      } else if constexpr (F == 3) {  // w_x = i (ky v_z - kz v_y)
        const cpx<R> y = a[AF + pt], z = a[2 * AF + pt];
        v = cmuli(cpx<R>{ky * z.x - kz * y.x, ky * z.y - kz * y.y});
      } else {  // w_y = i kz v_x - D_z,  w_z = D_y - i ky v_x
        const cpx<R> vx = a[pt];
        const cpx<R> d = a[(F == 4 ? 4 : 3) * AF + pt];
        const R k = F == 4 ? kz : ky;
        const cpx<R> t = cmuli(cpx<R>{k * vx.x, k * vx.y});
        v = F == 4 ? cpx<R>{t.x - d.x, t.y - d.y} : cpx<R>{d.x - t.x, d.y - t.y};
      }
      return ok ? v : cpx<R>{R(0), R(0)};
    };
    auto st = [&](int n, int l, cpx<R> v, int) {
      const int p = l / KZ, iz = l - p * KZ;
      b[F * BF + size_t(x0 + p) * PB + size_t(n) * KZ + iz] = v;
    };
    fft_io<R, N, E, +1, 1, true>(sm, LY, LY, tws, ld, st);
  }
  // block (g, F): plane group g, output field F
  __device__ __forceinline__ static void run(cpx<R>* sm, int g, int f, const cpx<R>* __restrict__ a,
                                             cpx<R>* __restrict__ b, const cpx<R>* __restrict__ tw, R ks) {
    cpx<R>* tws = sm + N * C::LY;
    stage_twiddles<R, N>(tws, tw);
    const int x0 = g * C::PPB;
    switch (f) {
      case 0: field<0>(sm, x0, a, b, tws, ks); break;
      case 1: field<1>(sm, x0, a, b, tws, ks); break;
      case 2: field<2>(sm, x0, a, b, tws, ks); break;
      case 3: field<3>(sm, x0, a, b, tws, ks); break;
      case 4: field<4>(sm, x0, a, b, tws, ks); break;
      default: field<5>(sm, x0, a, b, tws, ks); break;
    }
  }
};

// ---------------------------------------------------------------- forward y
// Field f of the planes x0 .. x0 + PPB - 1: forward transform along y of the
// z-pass output (rows y, pitch KZ, plane pitch PB), keeping the K masked
// rows, into A[f][x][PA].
template <typename R, int N>
struct EnsFwdY {
  using G = Geo<N>;
  using C = EnsCfg<R, N>;
  static constexpr int SMEM = N * C::LY + N;
  __device__ __forceinline__ static void run(cpx<R>* sm, int g, int f, const cpx<R>* __restrict__ in,
                                             cpx<R>* __restrict__ out, const cpx<R>* __restrict__ tw) {
    constexpr int LY = C::LY, E = C::E, PA = C::PA, PB = C::PB, KZ = G::KZ;
    cpx<R>* tws = sm + N * LY;
    stage_twiddles<R, N>(tws, tw);
    const int x0 = g * C::PPB;
    const cpx<R>* src = in + size_t(f) * N * PB;
    cpx<R>* dst = out + size_t(f) * N * PA;
    auto ld = [&](int n, int l, int) -> cpx<R> {
      const int p = l / KZ, iz = l - p * KZ;
      return src[size_t(x0 + p) * PB + size_t(n) * KZ + iz];
    };
    auto st = [&](int n, int l, cpx<R> v, int) {
      if (!G::kept(n)) return;
      const int p = l / KZ, iz = l - p * KZ;
      dst[size_t(x0 + p) * PA + size_t(G::compact_of(n)) * KZ + iz] = v;
    };
    fft_io<R, N, E, -1, 1, true>(sm, LY, LY, tws, ld, st);
  }
};

// ---------------------------------------------------------------- forward x + stage
// Flat tile `bt`: forward transform along x of the three fields of the
// forward-y output, keeping the K masked rows; each thread keeps its
// column's last-stage outputs in registers and applies the stage epilogue
// (Leray projection, -nu k^2 u_s, RK4 update) at its own (kx, ky, kz) points
// with every load of a point issued before its stores.
template <typename R>
struct EnsStage {
  const cpx<R>* us;  // stage input
  const cpx<R>* u0;  // state at the start of the step
  cpx<R>* unext;     // next stage input (u0 itself at stage 3)
  cpx<R>* acc;       // r1 + 2 r2 + 2 r3
  R dt, nu, ks;
  int stage;
};

template <typename R, int N>
struct EnsFwdX {
  using G = Geo<N>;
  using C = EnsCfg<R, N>;
  static constexpr int SMEM = N * C::LX + N;
  __device__ __forceinline__ static void run(cpx<R>* sm, int bt, const cpx<R>* __restrict__ in,
                                             const EnsStage<R>& a, const cpx<R>* __restrict__ tw) {
    constexpr int L = C::LX, E = C::E, PA = C::PA;
    using Pl = typename PlanFor<N, E>::type;
    constexpr int RL = LastRadix<Pl>::value, NBL = E / RL, SL = N / RL;
    cpx<R>* buf = sm;
    cpx<R>* tws = sm + N * L;
    stage_twiddles<R, N>(tws, tw);
    const int tid = threadIdx.x, l = tid % L, t = tid / L;
    const int colr = bt * L + l;
    const bool colok = colr < G::K * G::KZ;
    const int col = colok ? colr : G::K * G::KZ - 1;
    cpx<R> o[3][E];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      if (f) __syncthreads();  // the previous field's last LDS reads are done
      auto ld = [&](int n, int, int) -> cpx<R> { return in[(size_t(f) * N + n) * PA + col]; };
      auto st = [&](int, int, cpx<R> v, int slot) { o[f][slot] = v; };
      fft_io<R, N, E, -1, 1, true>(buf, L, L, tws, ld, st);
    }
    if (!colok || t >= C::TPL) return;
    const size_t cs = size_t(G::K) * PA;
    const int iy = col / G::KZ, iz = col - iy * G::KZ;
    const R ky = a.ks * R(G::kint(iy)), kz = a.ks * R(iz);
#pragma unroll
    for (int j = 0; j < NBL; ++j) {
#pragma unroll
      for (int m = 0; m < RL; ++m) {
        const int n = t * NBL + j + m * SL;
        if (!G::kept(n)) continue;
        const int ix = G::compact_of(n);
        const size_t pt = size_t(ix) * PA + col;
        cpx<R> v[3] = {o[0][j * RL + m], o[1][j * RL + m], o[2][j * RL + m]};
        const R kx = a.ks * R(G::kint(ix));
        const R k2 = kx * kx + ky * ky + kz * kz;
        const R k2s = (k2 == R(0)) ? R(1) : k2;
        cpx<R> su[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) su[c] = a.us[c * cs + pt];
        const RK4In<R> in0 = rk4_load(a.stage, pt, cs, a.u0, a.acc);
        project(kx, ky, kz, k2s, v);
        cpx<R> ru[3];
        const R nk = a.nu * k2;
#pragma unroll
        for (int c = 0; c < 3; ++c) ru[c] = {v[c].x - nk * su[c].x, v[c].y - nk * su[c].y};
        rk4_store(a.stage, a.dt, pt, cs, ru, in0, a.unext, a.acc);
      }
    }
  }
};

}  // namespace tg
