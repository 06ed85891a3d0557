// The five fused passes of one pseudo-spectral right-hand side on the
// pruned (2/3-truncated) half spectrum, plus the diagnostics reductions.
//
// Index conventions (all arrays C order, z fastest):
//   N    grid points per axis,  M = N/3 (integer), K = 2M+1, KZ = M+1
//   compact spectral field  S[c][ix][iy][iz]    ix,iy in [0,K), iz in [0,KZ)
//       compact index i <-> signed wavenumber  i <= M ? i : i - K
//       and full FFT index                     i <= M ? i : i + N - K
//   after the x pass        A[f][x][iy][iz]     x in [0,N)
//   after the y pass        B[f][x][y][iz]      y in [0,N)
//   physical field          P[f][x][y][z]       z in [0,N)
// Only modes inside the inclusive cube mask |k_i| <= N/3 are ever stored, so
// the 2/3 rule is applied by construction and every pass reads/writes only
// the kept part of its axis (zero padding lives in registers and LDS).
#pragma once
#include "tg_fft.hpp"

namespace tg {

template <int N>
struct Geo {
  static constexpr int M = N / 3;
  static constexpr int K = 2 * M + 1;
  static constexpr int KZ = M + 1;
  __device__ static constexpr bool kept(int n) { return n <= M || n >= N - M; }
  __device__ static constexpr int compact_of(int n) { return n <= M ? n : n - (N - K); }
  __device__ static constexpr int full_of(int i) { return i <= M ? i : i + (N - K); }
  __device__ static constexpr int kint(int i) { return i <= M ? i : i - K; }
};

// Row pitch (complex elements) of the intermediate arrays A and B: KZ
// rounded up to a whole number of 128-byte lines, so every tile of 8 (FP64)
// or 16 (FP32) columns starts on a cache-line boundary.  The compact state
// keeps the dense pitch KZ.
template <typename R, int N>
constexpr int kzp() {
  constexpr int per_line = 128 / int(2 * sizeof(R));
  return (Geo<N>::KZ + per_line - 1) / per_line * per_line;
}

// The intermediate arrays A and B of grids of 256^3 and up (0.2-0.6 GB
// each in FP64) exceed the 256 MB last-level cache, so the passes stream
// them with nontemporal loads and stores; the compact state (re-read by the
// next pass while still cached) and smaller grids use ordinary accesses.
template <int N>
constexpr bool nt_mid() { return N >= 256; }
// The exceptions are outputs that the next pass reads at once: the x pass's
// output (5 fields, read by the inverse y), the hydro z pass's (3 fields,
// read by the forward y) and, in a stage VJP, the fused inverse x of the
// next cotangent (3 fields).  While such an output is not much larger
// than the last-level cache it is stored with ordinary stores, so most of it
// is still cached when the next pass reads it.  At 256^3 (FP64: 308 and 277
// MB; FP32: 168 and 151 MB) this took FP64 from 402 to 422 steps/s and FP32
// from 745 to 765.  Cached stores of the larger inverse-y and forward-y
// outputs, or of every array, were measured slower, and so were cached
// stores of the VJP's y-pass outputs (the recompute stream shares the cache).
constexpr bool nt_store_for(long long bytes) { return bytes > (320LL << 20); }

// Launch shapes.  Grids of N >= 128 compile NVAR tile variants per pass;
// the plan picks one per pass at run time (see the autotuner); smaller grids
// compile variant 0 only.  Every shape is a function of N: the points per
// lane E divide N with a radix plan (N = 2^a or 3 * 2^a), a block's LDS tile
// stays within the 160 KiB of a CU and its thread count within 1024.
constexpr int NVAR = 3;
template <int N>
constexpr int num_variants() { return N >= 128 ? NVAR : 1; }

template <typename R>
constexpr bool lds_fits(long elems) { return elems * long(2 * sizeof(R)) <= 160L * 1024; }

// points per lane of the line passes: 12 when 3 divides N, else 16 (N >= 128),
// 8, or 4 (N = 4)
template <int N>
constexpr int base_e() { return N % 3 == 0 ? 12 : (N >= 128 ? 16 : (N == 4 ? 4 : 8)); }
// the fewer-points alternative of variant 1 (8, or 12 when 3 divides N)
template <int N>
constexpr int alt_e() { return N % 3 == 0 ? 12 : 8; }

// X/Y passes: L columns of kz per block, E points per thread.
//   V0: 16 columns x base_e points   V1: one 128-byte line of columns (8
//   FP64, 16 FP32) x alt_e points   V2: 32 columns x base_e points
// Grids above 256 would double the block: V0 halves L while the block
// exceeds 256 threads, V1 while it exceeds 512, and V2 takes 16 columns x
// 2 base_e points when 32 columns exceed LDS or 512 threads.
template <int N, int E>
constexpr int xy_cap_l(int l, int cap) {
  while (l > 8 && l * N / E > cap) l /= 2;
  return l;
}
template <typename R, int N>
constexpr bool xy_wide() { return lds_fits<R>(long(N) * 33) && 32 * N / base_e<N>() <= 512; }
// At N = 128 variant 2 is instead one line of columns x 4 points: a 128^3
// grid has only K * KZ / 8 = 457 FP64 column lines, so 8 or 16 points per
// thread leave about one wavefront per SIMD; 4 points double the wavefronts.
template <int N>
constexpr bool xy_narrow() { return N <= 128 && N % 3 != 0; }
// At N = 256 in FP64 variant 2 is one 128-byte line of columns x 16 points
// (128 threads): one LDS exchange per transform instead of the two of
// variant 1's 8 points; the y passes run 3% faster with it (32 columns x 16
// points, the generic variant 2, spills).
template <typename R, int N>
constexpr bool xy_mid() { return sizeof(R) == 8 && N == 256; }
template <typename R, int N, int V>
struct CfgXY {
  static constexpr int E = V == 1 ? alt_e<N>()
                         : V == 2 ? (xy_narrow<N>() ? 4
                                     : xy_mid<R, N>() ? base_e<N>()
                                     : !xy_wide<R, N>() ? 2 * base_e<N>() : base_e<N>())
                         : base_e<N>();
  static constexpr int L = V == 1 ? xy_cap_l<N, E>(sizeof(R) == 4 ? 16 : 8, 512)
                         : V == 2 ? (xy_narrow<N>() ? (sizeof(R) == 4 ? 16 : 8)
                                     : xy_mid<R, N>() ? 8
                                     : xy_wide<R, N>() ? 32 : 16)
                         : xy_cap_l<N, E>(16, 256);
  static constexpr int NT = L * N / E;
  static_assert(NT <= 1024 && lds_fits<R>(long(N) * (L + 1)), "x/y tile exceeds a workgroup");
};
// Z pass: NB = 3 x (hydro lines per block); E points per thread for the
// inverse, E/2 for the forward (half as many sequences).
//   V0: NB = 12 (N >= 256 or N not a multiple of 8) or 24    V1: twice V0
//   where LDS holds it    V2: half of V0
template <int N>
constexpr int z_nb0() { return (N >= 256 || N % 8 != 0) ? 12 : 24; }
template <int N>
constexpr int z_e() { return N % 3 == 0 ? (N >= 48 ? 24 : 12) : (N >= 128 ? 16 : (N == 4 ? 4 : 8)); }
template <typename R, int N, int V>
struct CfgZ {
  static constexpr int NB0 = z_nb0<N>();
  // the widest block (NB sequences of N points) plus pad and twiddles
  static constexpr int NB1 = lds_fits<R>(long(N) * (2 * NB0 + 3) + 1) ? 2 * NB0 : NB0;
  static constexpr int NB = V == 1 ? NB1 : (V == 2 ? NB0 / 2 : NB0);
  static constexpr int E = z_e<N>();
};
// Forward-x stage pass: LINES lines per block (3 fields x LINES/3 kz, or 6 x LINES/6).
//   V0: 24 lines, E = base_e   V1: 24 lines, E = alt_e   V2: 48 lines (FP32)
//   / 12 lines (FP64), E = base_e; each halved (down to 6) until the N x
//   LINES tile fits LDS and the block 1024 threads (512 above N = 256)
template <typename R, int N, int E>
constexpr int xf_lines(int want) {
  constexpr int cap = N > 256 ? 512 : 1024;
  int l = want;
  while (l > 6 && (!lds_fits<R>(long(N) * (l + 1)) || l * N / E > cap)) l /= 2;
  return l;
}
template <typename R, int N, int V>
struct CfgXF {
  static constexpr int E = (V == 1) ? alt_e<N>() : base_e<N>();
  static constexpr int LINES = xf_lines<R, N, E>(V == 2 ? (sizeof(R) == 4 ? 48 : 12) : 24);
  static constexpr int NT = LINES * N / E;
  static_assert(NT <= 1024 && LINES % 6 == 0 && lds_fits<R>(long(N) * (LINES + 1)), "x tile exceeds a workgroup");
};

enum XMode { XM_HYDRO = 0, XM_MHD = 1, XM_COPY3 = 2, XM_DIV = 3, XM_CURL3 = 4, XM_ADJ = 5, XM_ADJQ = 6,
             XM_HYDRO5 = 7, XM_MHD5 = 8, XM_GRAD3 = 9 };
enum ZMode { ZM_HYDRO = 0, ZM_MHD = 1, ZM_PHYS_OUT3 = 2, ZM_MAXABS = 3, ZM_PHYS_IN3 = 4, ZM_PHYS_OUT1 = 5, ZM_ADJ = 6 };
// FM_ADJ5: FM_ADJ from the five fields of k_y_adj (register forward x only)
enum FMode { FM_HYDRO_STAGE = 0, FM_MHD_STAGE = 1, FM_STORE3 = 2, FM_HYDRO_RHS = 3, FM_MHD_RHS = 4, FM_ADJ = 5,
             FM_ADJ5 = 6 };

__host__ __device__ constexpr int xmode_fields(int m) {
  return m == XM_HYDRO ? 6 : m == XM_MHD ? 12 : m == XM_ADJ ? 9 : m == XM_HYDRO5 ? 5 : m == XM_MHD5 ? 10
       : (m == XM_COPY3 || m == XM_CURL3 || m == XM_ADJQ || m == XM_GRAD3) ? 3 : 1;
}

// ----------------------------------------------------------------- pass 1
// Inverse transform along x of the fields derived from the compact state:
// HYDRO -> (u, omega) = (S, i k x S); MHD -> (u, omega, b, j); COPY3 -> S;
// CURL3 -> i k x S; DIV -> i k . S; ADJ -> (u, omega, P(alpha a)) of the
// state and the cotangent; ADJQ -> P(alpha S) of a cotangent S alone;
// HYDRO5 -> (u, i kx u_y, i kx u_z), the five fields from which the y pass
// completes omega (k_y_curl); MHD5 -> the same for u and for b; GRAD3 ->
// i k S_0, the gradient of the scalar field held in component 0.  The 1/N^3
// inverse normalisation is applied here.
template <typename R>
__device__ __forceinline__ cpx<R> xfield5(int f, cpx<R> a, cpx<R> b, cpx<R> c, R kx) {
  if (f == 0) return a;
  if (f == 1) return b;
  if (f == 2) return c;
  return cmuli(f == 3 ? cpx<R>{kx * b.x, kx * b.y} : cpx<R>{kx * c.x, kx * c.y});
}

template <typename R, int N, int MODE>
__device__ __forceinline__ cpx<R> xfield(int f, const cpx<R>* __restrict__ su,
                                         const cpx<R>* __restrict__ sb, size_t pt, size_t cs,
                                         R kx, R ky, R kz, R alpha = R(1)) {
  if constexpr (MODE == XM_HYDRO5 || MODE == XM_MHD5) {
    const cpx<R>* s = f >= 5 ? sb : su;
    const int g = f >= 5 ? f - 5 : f;
    if (g < 3) return s[g * cs + pt];
    return xfield5<R>(g, cpx<R>{}, s[cs + pt], s[2 * cs + pt], kx);
  }
  if constexpr (MODE == XM_GRAD3) {
    const cpx<R> a = su[pt];
    const R k = f == 0 ? kx : (f == 1 ? ky : kz);
    return cmuli(cpx<R>{k * a.x, k * a.y});
  }
  if constexpr (MODE == XM_ADJQ) {  // P(alpha * a) of the cotangent su
    sb = su;
    f += 6;
  }
  if constexpr (MODE == XM_ADJ || MODE == XM_ADJQ) {
    // fields 0-5: u, curl u of the state su; 6-8: P(alpha * a) of the
    // cotangent sb (Leray projection, k2 == 0 -> 1)
    if (f >= 6) {
      const cpx<R> a = sb[pt], b = sb[cs + pt], c = sb[2 * cs + pt];
      const R k2 = kx * kx + ky * ky + kz * kz;
      const R k2s = (k2 == R(0)) ? R(1) : k2;
      const cpx<R> kd = {kx * a.x + ky * b.x + kz * c.x, kx * a.y + ky * b.y + kz * c.y};
      const R kc = f == 6 ? kx : (f == 7 ? ky : kz);
      const cpx<R> v = f == 6 ? a : (f == 7 ? b : c);
      return {alpha * (v.x - kc * kd.x / k2s), alpha * (v.y - kc * kd.y / k2s)};
    }
  }
  const cpx<R>* s = su;
  if constexpr (MODE == XM_MHD) {
    if (f >= 6) { s = sb; f -= 6; }
  }
  if constexpr (MODE == XM_CURL3) f += 3;
  if (f < 3 && MODE != XM_DIV) return s[f * cs + pt];  // (ADJ f < 6 falls through to the curl)
  const cpx<R> a = s[pt], b = s[cs + pt], c = s[2 * cs + pt];
  if constexpr (MODE == XM_DIV) {
    cpx<R> d = {kx * a.x + ky * b.x + kz * c.x, kx * a.y + ky * b.y + kz * c.y};
    return cmuli(d);
  }
  cpx<R> r;
  if (f == 3) r = {ky * c.x - kz * b.x, ky * c.y - kz * b.y};
  else if (f == 4) r = {kz * a.x - kx * c.x, kz * a.y - kx * c.y};
  else r = {kx * b.x - ky * a.x, kx * b.y - ky * a.y};
  return cmuli(r);
}

// xfield on register-held components: sv[c][slot] holds component c (0-2
// the state, 3-5 the second field: b for MHD, the cotangent for ADJ).
template <typename R, int N, int MODE, int NS, int E>
__device__ __forceinline__ cpx<R> xfield_reg(int f, const cpx<R> (&sv)[NS][E], int slot, R kx, R ky, R kz,
                                             R alpha) {
  if constexpr (MODE == XM_GRAD3) {
    const cpx<R> a = sv[0][slot];
    const R k = f == 0 ? kx : (f == 1 ? ky : kz);
    return cmuli(cpx<R>{k * a.x, k * a.y});
  } else {
  int o = 0;
  if constexpr (MODE == XM_HYDRO5 || MODE == XM_MHD5) {
    if (f >= 5) { o = 3; f -= 5; }
    return xfield5<R>(f, sv[o][slot], sv[o + 1][slot], sv[o + 2][slot], kx);
  }
  if constexpr (MODE == XM_ADJ || MODE == XM_ADJQ) {
    if (MODE == XM_ADJQ || f >= 6) {
      constexpr int q0 = MODE == XM_ADJQ ? 0 : 3;  // components of the cotangent
      if constexpr (MODE == XM_ADJQ) f += 6;
      const cpx<R> a = sv[q0][slot], b = sv[q0 + 1][slot], c = sv[q0 + 2][slot];
      const R k2 = kx * kx + ky * ky + kz * kz;
      const R k2s = (k2 == R(0)) ? R(1) : k2;
      const cpx<R> kd = {kx * a.x + ky * b.x + kz * c.x, kx * a.y + ky * b.y + kz * c.y};
      const R kc = f == 6 ? kx : (f == 7 ? ky : kz);
      const cpx<R> v = f == 6 ? a : (f == 7 ? b : c);
      return {alpha * (v.x - kc * kd.x / k2s), alpha * (v.y - kc * kd.y / k2s)};
    }
  }
  if constexpr (MODE == XM_MHD) {
    if (f >= 6) { o = 3; f -= 6; }
  }
  if constexpr (MODE == XM_CURL3) f += 3;
  const cpx<R> a = sv[o][slot], b = sv[o + 1][slot], c = sv[o + 2][slot];
  if constexpr (MODE == XM_DIV) {
    return cmuli(cpx<R>{kx * a.x + ky * b.x + kz * c.x, kx * a.y + ky * b.y + kz * c.y});
  } else {
    if (f == 0) return a;
    if (f == 1) return b;
    if (f == 2) return c;
    cpx<R> r;
    if (f == 3) r = {ky * c.x - kz * b.x, ky * c.y - kz * b.y};
    else if (f == 4) r = {kz * a.x - kx * c.x, kz * a.y - kx * c.y};
    else r = {kx * b.x - ky * a.x, kx * b.y - ky * a.y};
    return cmuli(r);
  }
  }
}

template <typename R, int N, int MODE, int V>
struct XInvBody {
  using G = Geo<N>;
  using C = CfgXY<R, N, V>;
  static constexpr int SMEM = N * C::L + N;  // line tile + twiddles
  // Block (bx, by) = (kz tile, iy).
  // cpz: row pitch (complex elements) of the compact state arrays su, sb.
  __device__ __forceinline__ static void run(cpx<R>* sm, int bx, int by, const cpx<R>* __restrict__ su,
                             const cpx<R>* __restrict__ sb, cpx<R>* __restrict__ out,
                             const cpx<R>* __restrict__ tw, R ks, R alpha = R(1), int cpz = Geo<N>::KZ) {
    constexpr int L = C::L, E = C::E;
    constexpr int F = xmode_fields(MODE);
    // the launch's whole output (both vectors of the MHD5 split)
    constexpr bool NT_OUT = nt_store_for(
        (long long)(MODE == XM_MHD5 ? 2 * F : F) * N * G::K * kzp<R, N>() * (long long)sizeof(cpx<R>));
    cpx<R>* buf = sm;
    cpx<R>* tws = sm + N * L;
    const int iz0 = bx * L, iy = by;
    const size_t cs = size_t(G::K) * G::K * cpz;
    const R scale = R(1) / (R(N) * R(N) * R(N));
    const R ky = ks * R(G::kint(iy));
    stage_twiddles<R, N>(tws, tw);
    // The first FFT stage of every field reads the same E points per thread:
    // load the state (and cotangent) components at those points once.
    constexpr int NS = (MODE == XM_GRAD3) ? 1
                     : (MODE == XM_HYDRO || MODE == XM_CURL3 || MODE == XM_DIV || MODE == XM_COPY3) ? 3
                     : (MODE == XM_MHD || MODE == XM_MHD5 || MODE == XM_ADJ) ? 6 : 3;
    constexpr int TPL = N / E;
    const int tid = threadIdx.x;
    const int l = tid % L, t = tid / L;
    const bool act = t < TPL;
    const int izr = iz0 + l;
    const bool colok = izr < G::KZ;
    const int iz = colok ? izr : G::KZ - 1;  // clamp inside the row: same cache line
    // Registers only while the preload stays small; otherwise every field
    // reads its components from global memory (cache-served re-reads).
    constexpr bool PRE = NS * E <= 48;
    cpx<R> sv[PRE ? NS : 1][PRE ? E : 1];
    if constexpr (PRE) {
      using Pl = typename PlanFor<N, E>::type;
      constexpr int RAD0 = FirstRadix<Pl>::value;
      constexpr int NB0 = E / RAD0, SIN0 = N / RAD0;
#pragma unroll
      for (int j = 0; j < NB0; ++j) {
#pragma unroll
        for (int m = 0; m < RAD0; ++m) {
          const int n = (act ? t : 0) * NB0 + j + m * SIN0;
          const bool ok = G::kept(n);
          const int ix = ok ? G::compact_of(n) : G::M;
          const size_t pt = (size_t(ix) * G::K + iy) * cpz + iz;
#pragma unroll
          for (int c = 0; c < NS; ++c) {
            const cpx<R>* src = c < 3 ? su : sb;
            sv[c][j * RAD0 + m] = src[(c % 3) * cs + pt];
          }
          if constexpr (MODE == XM_ADJQ) {
            // q = alpha P(a) once per point (the three fields then only read it)
            const R kx = ks * R(G::kint(ix)), kz = ks * R(iz);
            const R k2 = kx * kx + ky * ky + kz * kz;
            const R r = R(1) / ((k2 == R(0)) ? R(1) : k2);
            const cpx<R> a0 = sv[0][j * RAD0 + m], a1 = sv[1][j * RAD0 + m], a2 = sv[2][j * RAD0 + m];
            const cpx<R> kd = {(kx * a0.x + ky * a1.x + kz * a2.x) * r, (kx * a0.y + ky * a1.y + kz * a2.y) * r};
            sv[0][j * RAD0 + m] = {alpha * (a0.x - kx * kd.x), alpha * (a0.y - kx * kd.y)};
            sv[1][j * RAD0 + m] = {alpha * (a1.x - ky * kd.x), alpha * (a1.y - ky * kd.y)};
            sv[2][j * RAD0 + m] = {alpha * (a2.x - kz * kd.x), alpha * (a2.y - kz * kd.y)};
          }
        }
      }
    }
    // Unrolled: each field's formula is resolved at compile time (a run-time
    // field index costs a branch chain per point and keeps every candidate's
    // operands live).
#pragma unroll
    for (int f = 0; f < F; ++f) {
      if (f) __syncthreads();  // the previous field's last LDS reads are done
      auto ld = [&](int n, int ll, int slot) -> cpx<R> {
        const bool ok = colok && G::kept(n);
        const int ix = ok ? G::compact_of(n) : 0;
        const R kx = ks * R(G::kint(ix)), kz = ks * R(iz);
        cpx<R> v;
        if constexpr (PRE && MODE == XM_ADJQ) {
          v = sv[f][slot];  // projected at preload
        } else if constexpr (PRE) {
          v = xfield_reg<R, N, MODE, NS, E>(f, sv, slot, kx, ky, kz, alpha);
        } else {
          const size_t pt = (size_t(ix) * G::K + iy) * cpz + iz;
          v = xfield<R, N, MODE>(f, su, sb, pt, cs, kx, ky, kz, alpha);
        }
        v = cscale(v, scale);
        return ok ? v : cpx<R>{R(0), R(0)};
      };
      auto st = [&](int n, int ll, cpx<R> v, int) {
        const int izw = iz0 + ll;
        if (izw < G::KZ) gst<NT_OUT>(&out[((size_t(f) * N + n) * G::K + iy) * kzp<R, N>() + izw], v);
      };
      fft_io<R, N, E, +1, 1, true>(buf, L, L, tws, ld, st);
    }
  }
};

// FP32 variant 1 (8 points per thread) fits four wavefronts per SIMD (not
// the adjoint mode, whose nine fields need more registers).
template <typename R, int MODE, int V>
constexpr int inv_x_waves_per_eu() { return (V == 1 && sizeof(R) == 4 && MODE != XM_ADJ) ? 4 : 1; }

template <typename R, int N, int MODE, int V>
__global__ __launch_bounds__((CfgXY<R, N, V>::NT)) __attribute__((amdgpu_waves_per_eu(inv_x_waves_per_eu<R, MODE, V>())))
void k_inv_x(const cpx<R>* __restrict__ su,
                                                              const cpx<R>* __restrict__ sb,
                                                              cpx<R>* __restrict__ out,
                                                              const cpx<R>* __restrict__ tw, R ks,
                                                              R alpha, int cpz) {
  if constexpr (MODE == XM_MHD5) {
    // the five fields of u (blockIdx.z = 0) and those of b (fields 5-9,
    // blockIdx.z = 1) are independent: two blocks per tile instead of one
    using B = XInvBody<R, N, XM_HYDRO5, V>;
    __shared__ cpx<R> sm[B::SMEM];
    const bool mag = blockIdx.z != 0;
    B::run(sm, blockIdx.x, blockIdx.y, mag ? sb : su, nullptr,
           out + (mag ? size_t(5) * N * Geo<N>::K * kzp<R, N>() : 0), tw, ks, alpha, cpz);
  } else {
    using B = XInvBody<R, N, MODE, V>;
    __shared__ cpx<R> sm[B::SMEM];
    B::run(sm, blockIdx.x, blockIdx.y, su, sb, out, tw, ks, alpha, cpz);
  }
}

// ----------------------------------------------------------------- pass 2/4
// Transform along y for F fields.  DIR=+1: A[f][x][iy][iz] (K rows, zero
// padded) -> B[f][x][y][iz].  DIR=-1: B -> A keeping the K masked rows.
template <typename R, int N, int DIR, int V>
struct YBody {
  using G = Geo<N>;
  using C = CfgXY<R, N, V>;
  static constexpr int SMEM = N * C::L + N;
  // Block (bx, by, bz) = (kz tile, x, field).
  __device__ __forceinline__ static void run(cpx<R>* sm, int bx, int by, int bz, const cpx<R>* __restrict__ in,
                             cpx<R>* __restrict__ out, const cpx<R>* __restrict__ tw) {
    constexpr int L = C::L, E = C::E;
    cpx<R>* buf = sm;
    cpx<R>* tws = sm + N * L;
    const int iz0 = bx * L;
    const size_t plane = size_t(bz) * N + by;
    stage_twiddles<R, N>(tws, tw);
    auto ld = [&](int n, int ll, int) -> cpx<R> {
      const int izr = iz0 + ll;
      bool ok = izr < G::KZ;
      const int iz = ok ? izr : G::KZ - 1;
      cpx<R> v;
      if constexpr (DIR > 0) {
        ok = ok && G::kept(n);
        v = gld<nt_mid<N>()>(&in[(plane * G::K + (ok ? G::compact_of(n) : 0)) * kzp<R, N>() + iz]);
      } else {
        v = gld<nt_mid<N>()>(&in[(plane * N + n) * kzp<R, N>() + iz]);
      }
      return ok ? v : cpx<R>{R(0), R(0)};
    };
    auto st = [&](int n, int ll, cpx<R> v, int) {
      const int iz = iz0 + ll;
      if (iz >= G::KZ) return;
      if constexpr (DIR > 0) {
        gst<nt_mid<N>()>(&out[(plane * N + n) * kzp<R, N>() + iz], v);
      } else {
        if (G::kept(n)) gst<nt_mid<N>()>(&out[(plane * G::K + G::compact_of(n)) * kzp<R, N>() + iz], v);
      }
    };
    fft_io<R, N, E, DIR, 1, true>(buf, L, L, tws, ld, st);
  }
};

template <typename R, int N, int DIR, int V>
__global__ __launch_bounds__((CfgXY<R, N, V>::NT)) void k_y(const cpx<R>* __restrict__ in,
                                                          cpx<R>* __restrict__ out,
                                                          const cpx<R>* __restrict__ tw) {
  using B = YBody<R, N, DIR, V>;
  __shared__ cpx<R> sm[B::SMEM];
  B::run(sm, blockIdx.x, blockIdx.y, blockIdx.z, in, out, tw);
}

// Forward y of the six fields the adjoint z pass leaves with the kz terms of
// i k x p2 folded (A0 = p1_x - i kz p2_y, A1 = p1_y + i kz p2_x, A2 = p1_z,
// A3..A5 = p2): folds the ky terms and keeps the K masked rows of five
// fields, X = A0 + i ky p2_z, Y = A1, Z = A2 - i ky p2_x, p2_y, p2_z (the
// register forward x adds the kx terms).  Block type bz: 0 -> X and p2_z
// from two transforms, 1 -> Y, 2 -> Z from two transforms, 3 -> p2_y.
template <typename R, int N, int V>
struct YAdjBody {
  using G = Geo<N>;
  using C = CfgXY<R, N, V>;
  static constexpr int SMEM = N * C::L + N;
  __device__ __forceinline__ static void run(cpx<R>* sm, int bx, int x, int bz, const cpx<R>* __restrict__ in,
                                             cpx<R>* __restrict__ out, const cpx<R>* __restrict__ tw, R ks) {
    constexpr int L = C::L, E = C::E;
    cpx<R>* buf = sm;
    cpx<R>* tws = sm + N * L;
    const int iz0 = bx * L;
    stage_twiddles<R, N>(tws, tw);
    auto ldf = [&](int f) {
      return [&, f](int n, int ll, int) -> cpx<R> {
        const int izr = iz0 + ll;
        const bool ok = izr < G::KZ;
        const cpx<R> v = gld<nt_mid<N>()>(&in[((size_t(f) * N + x) * N + n) * kzp<R, N>() + (ok ? izr : G::KZ - 1)]);
        return ok ? v : cpx<R>{R(0), R(0)};
      };
    };
    auto put = [&](int f, int n, int ll, cpx<R> v) {
      const int iz = iz0 + ll;
      if (iz < G::KZ && G::kept(n))
        gst<nt_mid<N>()>(&out[((size_t(f) * N + x) * G::K + G::compact_of(n)) * kzp<R, N>() + iz], v);
    };
    auto kyi = [&](int n, cpx<R> v) -> cpx<R> {  // i ky v at output row n
      const R ky = ks * R(G::kint(G::kept(n) ? G::compact_of(n) : 0));
      return cmuli(cpx<R>{ky * v.x, ky * v.y});
    };
    if (bz == 1 || bz == 3) {
      const int f = bz == 1 ? 1 : 4;
      fft_io<R, N, E, -1, 1, true>(buf, L, L, tws, ldf(f),
                                   [&](int n, int ll, cpx<R> v, int) { put(bz == 1 ? 1 : 3, n, ll, v); });
      return;
    }
    cpx<R> h[E];  // the first transform (p2_z for bz 0, p2_x for bz 2) by slot
    fft_io<R, N, E, -1, 1, true>(buf, L, L, tws, ldf(bz == 0 ? 5 : 3), [&](int n, int ll, cpx<R> v, int slot) {
      h[slot] = v;
      if (bz == 0) put(4, n, ll, v);
    });
    __syncthreads();  // the first transform's last LDS reads are done
    fft_io<R, N, E, -1, 1, true>(buf, L, L, tws, ldf(bz == 0 ? 0 : 2), [&](int n, int ll, cpx<R> v, int slot) {
      put(bz == 0 ? 0 : 2, n, ll, bz == 0 ? cadd(v, kyi(n, h[slot])) : csub(v, kyi(n, h[slot])));
    });
  }
};

template <typename R, int N, int V>
__global__ __launch_bounds__((CfgXY<R, N, V>::NT)) void k_y_adj(const cpx<R>* __restrict__ in,
                                                              cpx<R>* __restrict__ out,
                                                              const cpx<R>* __restrict__ tw, R ks) {
  using B = YAdjBody<R, N, V>;
  __shared__ cpx<R> sm[B::SMEM];
  B::run(sm, blockIdx.x, blockIdx.y, blockIdx.z, in, out, tw, ks);
}

// Inverse y transform that completes the curl.  The x pass (HYDRO5 / MHD5)
// hands over five fields per vector v: (v_x, v_y, v_z, D_y, D_z) with
// D_c = i kx v_c after the x transform.  While ky and kz are still spectral
// the remaining derivatives are one multiply each, so this pass writes the
// six fields (v, i k x v):
//   w_x = i (ky v_z - kz v_y),   w_y = i kz v_x - D_z,   w_z = D_y - i ky v_x.
// Compared with transforming six fields along x, the x pass writes and this
// pass reads one field in six less.  One block transforms one output field
// of one (x, kz tile); the components of v are read by up to three of the
// six blocks of that tile, with ordinary loads so that the copies the other
// blocks fetched serve them from L2 (the launcher places the six blocks on
// one XCD, one after another), and D_y, D_z once, streamed.
template <typename R, int N, int V>
struct YCurlBody {
  using G = Geo<N>;
  using C = CfgXY<R, N, V>;
  static constexpr int SMEM = N * C::L + N;
  static constexpr size_t AF = size_t(N) * G::K * kzp<R, N>();  // field stride of A
  static constexpr size_t BF = size_t(N) * N * kzp<R, N>();     // field stride of B
  // Output field F of one (x, kz tile).  F is a template argument so the
  // first stage's loads carry no per-point branch: all of them issue before
  // the curl arithmetic waits on any.
  template <int F>
  __device__ __forceinline__ static void field(cpx<R>* sm, int bx, const cpx<R>* __restrict__ a,
                                               cpx<R>* __restrict__ b, const cpx<R>* __restrict__ tws, R ks) {
    constexpr int L = C::L, E = C::E, KZP = kzp<R, N>();
    constexpr bool NT = nt_mid<N>();
    const int iz0 = bx * L;
    auto ld = [&](int n, int ll, int) -> cpx<R> {
      const int izr = iz0 + ll;
      const bool colok = izr < G::KZ;
      const bool ok = colok && G::kept(n);
      const int iy = ok ? G::compact_of(n) : 0;
      const size_t pt = size_t(iy) * KZP + (colok ? izr : G::KZ - 1);
      const R ky = ks * R(G::kint(iy)), kz = ks * R(izr);
      cpx<R> v;
      if constexpr (F < 3) {
        v = gld<false>(&a[F * AF + pt]);
      } else if constexpr (F == 3) {  // w_x = i (ky v_z - kz v_y)
        const cpx<R> y = gld<false>(&a[AF + pt]), z = gld<false>(&a[2 * AF + pt]);
        v = cmuli(cpx<R>{ky * z.x - kz * y.x, ky * z.y - kz * y.y});
      } else {  // w_y = i kz v_x - D_z,  w_z = D_y - i ky v_x
        const cpx<R> vx = gld<false>(&a[pt]);
        const cpx<R> d = gld<NT>(&a[(F == 4 ? 4 : 3) * AF + pt]);
        const R k = F == 4 ? kz : ky;
        const cpx<R> t = cmuli(cpx<R>{k * vx.x, k * vx.y});
        v = F == 4 ? cpx<R>{t.x - d.x, t.y - d.y} : cpx<R>{d.x - t.x, d.y - t.y};
      }
      return ok ? v : cpx<R>{R(0), R(0)};
    };
    auto st = [&](int n, int ll, cpx<R> v, int) {
      const int iz = iz0 + ll;
      if (iz < G::KZ) gst<NT>(&b[F * BF + size_t(n) * KZP + iz], v);
    };
    fft_io<R, N, E, +1, 1, true>(sm, L, L, tws, ld, st);
  }

  // (bx, x, f, vec): kz tile, x plane, output field 0-5, vector.  A[5 vec ..]
  // -> B[6 vec + f].
  __device__ __forceinline__ static void run(cpx<R>* sm, int bx, int x, int f, int vec,
                                             const cpx<R>* __restrict__ in, cpx<R>* __restrict__ out,
                                             const cpx<R>* __restrict__ tw, R ks) {
    constexpr int KZP = kzp<R, N>();
    const cpx<R>* a = in + size_t(5 * vec) * AF + size_t(x) * G::K * KZP;
    cpx<R>* b = out + size_t(6 * vec) * BF + size_t(x) * N * KZP;
    cpx<R>* tws = sm + N * C::L;
    stage_twiddles<R, N>(tws, tw);
    switch (f) {
      case 0: field<0>(sm, bx, a, b, tws, ks); break;
      case 1: field<1>(sm, bx, a, b, tws, ks); break;
      case 2: field<2>(sm, bx, a, b, tws, ks); break;
      case 3: field<3>(sm, bx, a, b, tws, ks); break;
      case 4: field<4>(sm, bx, a, b, tws, ks); break;
      default: field<5>(sm, bx, a, b, tws, ks); break;
    }
  }
};

// Linear block b of a 1-D grid of 48 * ceil(U / 8) blocks, U = tiles x N x
// vectors: workgroups go to the 8 XCDs round robin (b mod 8), so the six
// fields of tile u take the consecutive slots 6 j .. 6 j + 5 of XCD u mod 8
// (j = u / 8).  Blocks past the last tile return at once.
template <typename R, int N, int V>
__global__ __launch_bounds__((CfgXY<R, N, V>::NT)) void k_y_curl(const cpx<R>* __restrict__ in,
                                                               cpx<R>* __restrict__ out,
                                                               const cpx<R>* __restrict__ tw, R ks, int units) {
  using B = YCurlBody<R, N, V>;
  constexpr int T = (Geo<N>::KZ + CfgXY<R, N, V>::L - 1) / CfgXY<R, N, V>::L;
  __shared__ cpx<R> sm[B::SMEM];
  const int b = blockIdx.x, xcd = b & 7, j = b >> 3;
  const int u = (j / 6) * 8 + xcd;
  if (u >= units) return;
  B::run(sm, u % T, (u / T) % N, j % 6, u / (T * N), in, out, tw, ks);
}

// ----------------------------------------------------------------- pass 3
// Along z: complex-to-real of the inverse fields, the pointwise nonlinear
// product, real-to-complex of the results, all for a group of lines
// (x, y0..y0+LZ-1) held in LDS.  Two real fields share one complex sequence
// (a + i b); Hermitian extension supplies the negative kz half.
template <int MODE>
struct ZShape;
template <> struct ZShape<ZM_HYDRO> { static constexpr int SPL = 3, FIN = 6, FOUT = 3; };
template <> struct ZShape<ZM_MHD> { static constexpr int SPL = 6, FIN = 12, FOUT = 6; };
template <> struct ZShape<ZM_PHYS_OUT3> { static constexpr int SPL = 2, FIN = 3, FOUT = 0; };
template <> struct ZShape<ZM_MAXABS> { static constexpr int SPL = 1, FIN = 1, FOUT = 0; };
template <> struct ZShape<ZM_PHYS_IN3> { static constexpr int SPL = 3, FIN = 0, FOUT = 3; };
template <> struct ZShape<ZM_PHYS_OUT1> { static constexpr int SPL = 1, FIN = 1, FOUT = 0; };
// ADJ: 9 inverse fields (u, curl u, q) per line packed as 9 sequences per
// pair of lines, 6 forward fields per line (SPL counts per 2 lines / 2).
template <> struct ZShape<ZM_ADJ> { static constexpr int SPL = 0, FIN = 9, FOUT = 6; };

// Per-mode block shape: LZ lines, NSEQ = LZ * SPL inverse-side sequences.
template <typename R, int N, int MODE, int V>
struct CfgZM {
  static constexpr int LZ = (MODE == ZM_MHD || MODE == ZM_ADJ) ? CfgZ<R, N, V>::NB / 6 : CfgZ<R, N, V>::NB / 3;
  static constexpr int NSEQ = (MODE == ZM_ADJ) ? 9 * LZ / 2 : LZ * ZShape<MODE>::SPL;
  static constexpr int E = CfgZ<R, N, V>::E;
  static constexpr int NT = NSEQ * N / E;
  static constexpr int LP = NSEQ | 1;  // odd pitch: conflict-free column access
};

template <typename R>
__device__ __forceinline__ void cross3(const R* a, const R* b, R* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// PITCH: row pitch (complex elements) of the (x, y, kz) arrays it reads and
// writes (kzp for the single-grid plans, KZ for the dense ensemble layout).
template <typename R, int N, int MODE, int V, int PITCH = kzp<R, N>()>
struct ZBody {
  using G = Geo<N>;
  using Z = CfgZM<R, N, MODE, V>;
  using S = ZShape<MODE>;
  static constexpr int PSI = pad_shift<N, Z::E>();
  static constexpr int PSF = (MODE == ZM_ADJ) ? PSI : (S::FOUT > 0) ? pad_shift<N, Z::E / 2>() : PSI;
  static constexpr int PSMIN = PSI < PSF ? PSI : PSF;
  static constexpr int BUF = N * Z::LP + (N >> PSMIN) + 1;
  static constexpr int SMEM = BUF + N;
  // Block (bx, by) = (group of LZ y lines, x).  Threads at or beyond Z::NT
  // (a batched launch with a larger block) only join the barriers.
  __device__ __forceinline__ static void run(cpx<R>* sm, int bx, int by, const cpx<R>* __restrict__ in,
                             cpx<R>* __restrict__ out, const R* __restrict__ phys_in,
                             R* __restrict__ phys_out, const cpx<R>* __restrict__ tw, R lorentz) {
  constexpr int NSEQ = Z::NSEQ, LP = Z::LP, NT = Z::NT, E = Z::E;
  constexpr int LZ = Z::LZ;  // lines per block
  constexpr int KZ = G::KZ;
  static_assert(N % LZ == 0, "lines per block must divide N");
  // padded LDS layouts of the inverse output and of the forward input
  cpx<R>* buf = sm;
  cpx<R>* tws = sm + BUF;
  auto bi = [&](int n, int l) -> cpx<R>& { return buf[lidx<PSI>(n, l, LP)]; };
  stage_twiddles<R, N>(tws, tw);
  auto bf = [&](int n, int l) -> cpx<R>& { return buf[lidx<PSF>(n, l, LP)]; };
  const int tid = threadIdx.x;
  const bool live = tid < NT;
  const int x = by, y0 = bx * LZ;
  constexpr int KZP = PITCH;  // pitch of the A/B rows
  const size_t NNKZ = size_t(N) * N * KZP;
  const size_t line0 = size_t(x) * N + y0;  // (x, y0) line index

  if constexpr (MODE == ZM_ADJ) {
    // ---- inverse of u (f 0-2), curl u (3-5), q (6-8) for line pairs: per
    // pair g, sequences j = 0..3 (line 2g), 4..7 (line 2g+1) pack (u0,u1),
    // (u2,w0), (w1,w2), (q0,q1); j = 8 packs (q2 of 2g, q2 of 2g+1).
    auto ld = [&](int k, int s, int) -> cpx<R> {
      const int g = s / 9, j = s % 9;
      int la, lb, fa, fb;
      if (j < 8) {
        la = lb = 2 * g + (j >> 2);
        fa = 2 * (j & 3);
        fb = fa + 1;
      } else {
        la = 2 * g; lb = 2 * g + 1; fa = fb = 8;
      }
      const bool neg = k > N / 2;
      const int kzr = neg ? N - k : k;
      const bool ok = kzr < KZ;
      const int kz = ok ? kzr : 0;
      cpx<R> a = in[fa * NNKZ + (line0 + la) * KZP + kz];
      cpx<R> b = in[fb * NNKZ + (line0 + lb) * KZP + kz];
      if (k == 0) { a.y = R(0); b.y = R(0); }
      if (neg) { a.y = -a.y; b.y = -b.y; }
      const cpx<R> z = {a.x - b.y, a.y + b.x};
      return ok ? z : cpx<R>{R(0), R(0)};
    };
    fft_io<R, N, E, +1, 1, false, true, PSI>(buf, LP, NSEQ, tws, ld, NoIO{});
    // ---- products per line: p1 = w x q, p2 = q x u -> 3 forward sequences
    // per line: (p1x,p1y), (p1z,p2x), (p2y,p2z)
    constexpr int ITEMS = (LZ / 2) * N;
    constexpr int IPT = (ITEMS + NT - 1) / NT;
    R res[IPT][12];
#pragma unroll
    for (int it = 0; it < IPT; ++it) {
      const int item = tid + it * NT;
      if (!live || item >= ITEMS) break;
      const int z = item % N, g = item / N;
      const cpx<R> q2 = bi(z, 9 * g + 8);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const cpx<R> s0 = bi(z, 9 * g + 4 * h), s1 = bi(z, 9 * g + 4 * h + 1);
        const cpx<R> s2 = bi(z, 9 * g + 4 * h + 2), s3 = bi(z, 9 * g + 4 * h + 3);
        const R u[3] = {s0.x, s0.y, s1.x};
        const R w[3] = {s1.y, s2.x, s2.y};
        const R q[3] = {s3.x, s3.y, h ? q2.y : q2.x};
        cross3(w, q, &res[it][6 * h]);
        cross3(q, u, &res[it][6 * h + 3]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IPT; ++it) {
      const int item = tid + it * NT;
      if (!live || item >= ITEMS) break;
      const int z = item % N, g = item / N;
      const R* r = res[it];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int line = 2 * g + h;
        bf(z, 3 * line + 0) = {r[6 * h + 0], r[6 * h + 1]};
        bf(z, 3 * line + 1) = {r[6 * h + 2], r[6 * h + 3]};
        bf(z, 3 * line + 2) = {r[6 * h + 4], r[6 * h + 5]};
      }
    }
    __syncthreads();
    fft_io<R, N, E, -1, 0, false, true, PSF>(buf, LP, 3 * LZ, tws, NoIO{}, NoIO{});
    for (int idx = tid; live && idx < 3 * LZ * KZ; idx += NT) {
      const int kz = idx % KZ, q = idx / KZ;
      const cpx<R> P = bf(kz, q);
      const cpx<R> Qc = bf((N - kz) % N, q);
      const cpx<R> re = {R(0.5) * (P.x + Qc.x), R(0.5) * (P.y - Qc.y)};
      const cpx<R> im = {R(0.5) * (P.y + Qc.y), R(0.5) * (Qc.x - P.x)};
      const int line = q / 3, p = q % 3;
      out[(2 * p) * NNKZ + (line0 + line) * KZP + kz] = re;
      out[(2 * p + 1) * NNKZ + (line0 + line) * KZP + kz] = im;
    }
    return;
  }
  if constexpr (MODE != ZM_PHYS_IN3 && MODE != ZM_ADJ) {
    // ---- inverse: sequence s = line*SPL + p packs fields (fa, fb) as a + i b;
    // the first FFT stage reads the Hermitian-extended coefficients directly.
    auto ld = [&](int k, int s, int) -> cpx<R> {
      const int line = s / S::SPL, p = s % S::SPL;
      int fa, fb;
      if constexpr (MODE == ZM_HYDRO) { fa = p; fb = p + 3; }
      else if constexpr (MODE == ZM_MHD) { fa = (p < 3) ? p : p + 3; fb = fa + 3; }
      else if constexpr (MODE == ZM_PHYS_OUT3) { fa = 2 * p; fb = (2 * p + 1 < 3) ? 2 * p + 1 : -1; }
      else { fa = 0; fb = -1; }
      const bool neg = k > N / 2;
      const int kzr = neg ? N - k : k;
      const bool ok = kzr < KZ;
      const int kz = ok ? kzr : 0;  // branch-free: clamped load, then select
      const size_t off = (line0 + line) * KZP + kz;
      cpx<R> a = in[fa * NNKZ + off];
      cpx<R> b = in[(fb >= 0 ? fb : fa) * NNKZ + off];
      if (fb < 0) b = {R(0), R(0)};
      if (k == 0) { a.y = R(0); b.y = R(0); }
      if (neg) { a.y = -a.y; b.y = -b.y; }
      const cpx<R> z = {a.x - b.y, a.y + b.x};
      return ok ? z : cpx<R>{R(0), R(0)};
    };
    fft_io<R, N, E, +1, 1, false, true, PSI>(buf, LP, NSEQ, tws, ld, NoIO{});
  }

  if constexpr (MODE == ZM_PHYS_OUT3) {
    for (int idx = tid; live && idx < LZ * 3 * N; idx += NT) {
      const int z = idx % N, f = (idx / N) % 3, line = idx / (3 * N);
      const cpx<R> v = bi(z, line * S::SPL + f / 2);
      phys_out[(size_t(f) * N * N + line0 + line) * N + z] = (f & 1) ? v.y : v.x;
    }
    return;
  } else if constexpr (MODE == ZM_PHYS_OUT1) {
    for (int idx = tid; live && idx < LZ * N; idx += NT) {
      const int z = idx % N, line = idx / N;
      phys_out[(line0 + line) * N + z] = bi(z, line).x;
    }
    return;
  } else if constexpr (MODE == ZM_MAXABS) {
    __shared__ R red[NT];
    R m = R(0);
    for (int idx = tid; live && idx < LZ * N; idx += NT) {
      const int z = idx % N, line = idx / N;
      m = fmax(m, fabs(bi(z, line).x));
    }
    if (live) red[tid] = m;
    __syncthreads();
    if (tid == 0) {
      R mm = R(0);
      for (int i = 0; i < NT; ++i) mm = fmax(mm, red[i]);
      phys_out[size_t(by) * (N / LZ) + bx] = mm;
    }
    return;
  } else {
    // ---- products -> forward sequences.  Group g: HYDRO = 2 lines, 6 real
    // outputs (n of line a, n of line b); MHD = 1 line, 6 real outputs
    // (m = u x w + lorentz * j x b, e = u x b).  Each group fills 3 sequences
    // q = 3g + {0,1,2} with pairs (r0,r1), (r2,r3), (r4,r5).
    constexpr int NG = (MODE == ZM_MHD) ? LZ : LZ / 2;
    constexpr int ITEMS = NG * N;
    constexpr int IPT = (ITEMS + NT - 1) / NT;
    R res[IPT][6];
#pragma unroll
    for (int it = 0; it < IPT; ++it) {
      const int item = tid + it * NT;
      if (!live || item >= ITEMS) break;
      const int z = item % N, g = item / N;
      if constexpr (MODE == ZM_HYDRO) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int line = 2 * g + h;
          R u[3], w[3];
#pragma unroll
          for (int p = 0; p < 3; ++p) {
            const cpx<R> v = bi(z, line * 3 + p);
            u[p] = v.x; w[p] = v.y;
          }
          cross3(u, w, &res[it][3 * h]);
        }
      } else if constexpr (MODE == ZM_MHD) {
        R u[3], w[3], b[3], j[3];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          const cpx<R> v = bi(z, g * 6 + p);
          const cpx<R> q = bi(z, g * 6 + 3 + p);
          u[p] = v.x; w[p] = v.y; b[p] = q.x; j[p] = q.y;
        }
        R uw[3], jb[3];
        cross3(u, w, uw);
        cross3(j, b, jb);
#pragma unroll
        for (int p = 0; p < 3; ++p) res[it][p] = uw[p] + lorentz * jb[p];
        cross3(u, b, &res[it][3]);
      } else {  // PHYS_IN3: lines (2g, 2g+1), fields u,v,w read from physical
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const size_t line = line0 + 2 * g + h;
#pragma unroll
          for (int p = 0; p < 3; ++p) res[it][3 * h + p] = phys_in[(p * size_t(N) * N + line) * N + z];
        }
      }
    }
    if constexpr (MODE != ZM_PHYS_IN3) __syncthreads();
    // Pack: HYDRO/PHYS_IN: (nx_a + i ny_a), (nx_b + i ny_b), (nz_a + i nz_b);
    //       MHD:           (m_x + i m_y), (m_z + i e_x), (e_y + i e_z).
#pragma unroll
    for (int it = 0; it < IPT; ++it) {
      const int item = tid + it * NT;
      if (!live || item >= ITEMS) break;
      const int z = item % N, g = item / N;
      const R* r = res[it];
      if constexpr (MODE == ZM_MHD) {
        bf(z, 3 * g + 0) = {r[0], r[1]};
        bf(z, 3 * g + 1) = {r[2], r[3]};
        bf(z, 3 * g + 2) = {r[4], r[5]};
      } else {
        bf(z, 3 * g + 0) = {r[0], r[1]};
        bf(z, 3 * g + 1) = {r[3], r[4]};
        bf(z, 3 * g + 2) = {r[2], r[5]};
      }
    }
    __syncthreads();
    fft_io<R, N, E / 2, -1, 0, false, true, PSF>(buf, LP, NSEQ / 2, tws, NoIO{}, NoIO{});
    // ---- unpack the kept kz of each pair and store
    constexpr int NQ = NSEQ / 2;
    for (int idx = tid; live && idx < NQ * KZ; idx += NT) {
      const int kz = idx % KZ, q = idx / KZ;
      const cpx<R> P = bf(kz, q);
      const cpx<R> Qc = bf((N - kz) % N, q);  // conj taken below
      const cpx<R> re = {R(0.5) * (P.x + Qc.x), R(0.5) * (P.y - Qc.y)};
      const cpx<R> im = {R(0.5) * (P.y + Qc.y), R(0.5) * (Qc.x - P.x)};
      const int g = q / 3, p = q % 3;
      int line_a, fa, line_b, fb;
      if constexpr (MODE == ZM_MHD) {
        line_a = line_b = g;
        fa = 2 * p; fb = 2 * p + 1;
      } else {
        if (p < 2) { line_a = line_b = 2 * g + p; fa = 0; fb = 1; }
        else { line_a = 2 * g; line_b = 2 * g + 1; fa = 2; fb = 2; }
      }
      out[fa * NNKZ + (line0 + line_a) * KZP + kz] = re;
      out[fb * NNKZ + (line0 + line_b) * KZP + kz] = im;
    }
  }
  }
};

template <typename R, int N, int MODE, int V>
__global__ __launch_bounds__((CfgZM<R, N, MODE, V>::NT)) void k_z(const cpx<R>* __restrict__ in,
                                                                 cpx<R>* __restrict__ out,
                                                                 const R* __restrict__ phys_in,
                                                                 R* __restrict__ phys_out,
                                                                 const cpx<R>* __restrict__ tw, R lorentz) {
  using B = ZBody<R, N, MODE, V>;
  __shared__ cpx<R> sm[B::SMEM];
  B::run(sm, blockIdx.x, blockIdx.y, in, out, phys_in, phys_out, tw, lorentz);
}

// Hydro z pass with the lines in registers.  TPL = N / E lanes own one
// (x, y) line and hold all three packed sequences u_p + i w_p of it (E
// points each), so the product u x w is formed in registers right after the
// inverse transforms, and the forward transforms start from there (their
// plan is the inverse plan reversed, whose first-stage order is the inverse's
// output order).  LDS serves only as a per-line exchange buffer between
// stages, private to one wavefront: no workgroup barrier after the twiddle
// table is staged.  Forward sequences per line: n_x + i n_y (split by
// Hermitian symmetry on the way out) and n_z + 0 i.
template <int N, int E>
struct ZWPlan {
  using type = typename PlanFor<N, E>::type;
};
template <> struct ZWPlan<128, 8> { using type = Radices<8, 4, 4>; };

// LDS layout of the register z bodies' line buffers: one pad slot every
// 2^PS points (pidx) and a line pitch of LPL points.  The default keeps
// consecutive lines 32 banks apart.  In FP32, 256 bytes of lanes (32 lanes:
// the share of one LDS cycle at 64 banks of 4 bytes) hold 4 to 16 lines on
// the grids below 256, and the default layout puts those lines on the same
// banks: 2 to 4 times the conflict-free cycles, and 5.4 conflict cycles
// per LDS instruction measured in the ensemble's z launch.  Those shapes
// take the PS and LPL that a model of the banks gives the fewest cycles
// over every put and get of the line transforms and the Hermitian split
// (the smallest LPL on ties).  In FP64 the model's choices ran slower; the
// 128^3 shape at 8 points per lane (the MHD and adjoint z passes) takes the
// best of a few measured layouts (unpadded pitch: adjoint -0.7%), the
// others keep the default.
template <typename R, int N, int E>
struct ZwLds {
  static constexpr int PS = ilog2(E);
  static constexpr int USED = N + (N >> PS);
  static constexpr int LPL = sizeof(R) == 8 ? (USED + 7) / 16 * 16 + 8 : (USED + 15) / 32 * 32 + 16;
};
#define TG_ZW_LDS(R_, N_, E_, PS_, LPL_)                 \
  template <>                                           \
  struct ZwLds<R_, N_, E_> {                            \
    static constexpr int PS = PS_, LPL = LPL_;          \
    static constexpr int USED = N_ + (N_ >> PS_);       \
    static_assert(LPL_ >= USED, "line pitch below the padded line"); \
  };
TG_ZW_LDS(float, 32, 8, 3, 36)
TG_ZW_LDS(float, 32, 16, 4, 34)
TG_ZW_LDS(float, 48, 12, 2, 60)
TG_ZW_LDS(float, 64, 8, 3, 72)
TG_ZW_LDS(float, 64, 16, 3, 76)
TG_ZW_LDS(float, 96, 12, 1, 152)
TG_ZW_LDS(float, 128, 16, 4, 136)
TG_ZW_LDS(double, 128, 8, 3, 144)
#undef TG_ZW_LDS

template <typename R, int N, int E_, int WPB>
struct CfgZW {
  static constexpr int E = E_;
  static constexpr int TPL = N / E;          // lanes per line
  static constexpr int LW = 64 / TPL;        // lines per wavefront
  static_assert(N % E == 0 && 64 % TPL == 0, "a wavefront holds whole lines");
  static constexpr int LINES = WPB * LW;     // lines per block
  static constexpr int NT = 64 * WPB;
  static constexpr int PS = ZwLds<R, N, E>::PS;
  static constexpr int USED = ZwLds<R, N, E>::USED;
  static constexpr int LPL = ZwLds<R, N, E>::LPL;
  using IP = typename ZWPlan<N, E>::type;
  using FP = ReversePlan<IP>;
  static constexpr bool SAME = std::is_same<IP, FP>::value;
  static constexpr int TWI = PlanTw<N, IP>::SIZE;
  static constexpr int TWF = SAME ? 0 : PlanTw<N, FP>::SIZE;
  static constexpr int SMEM = TWI + TWF + LINES * LPL;
};

// every full index k in [lo, hi] lies in the truncated band (|kz| > N/3)
template <int N, int KZ>
__device__ __forceinline__ constexpr bool band_zero(int lo, int hi) { return lo >= KZ && hi <= N - KZ; }

// Load the Hermitian-extended packed sequence a + i b of fields FA, FB (FB <
// 0: b = 0) of one line into registers in the first-stage order of plan IP:
// slot j*R1 + m holds point k = t*NB1 + j + m*S1.
template <typename R, int N, class C, int FA, int FB>
__device__ __forceinline__ void zw_load(cpx<R>* v, const cpx<R>* __restrict__ src, size_t NNKZ, int t, int lane) {
  using IP = typename C::IP;
  constexpr int E = C::E, TPL = C::TPL, KZ = Geo<N>::KZ;
  constexpr int R1 = FirstRadix<IP>::value, NB1 = E / R1, S1 = N / R1;
  auto fetch = [&](int kz, cpx<R>& a, cpx<R>& b) {
    a = gld<nt_mid<N>()>(&src[FA * NNKZ + kz]);
    if constexpr (FB >= 0) b = gld<nt_mid<N>()>(&src[FB * NNKZ + kz]);
    else b = {R(0), R(0)};
  };
  if constexpr (NB1 == 1 && (TPL & (TPL - 1)) == 0) {
    // Slots m < R1/2 hold k = t + m*TPL <= N/2 - 1, read from memory; slot
    // m >= R1/2 holds N - k' of the lower slot R1-1-m of lane TPL - t (lane
    // 0: its own slot R1 - m; k = N/2 is truncated), so every coefficient
    // is read once and the mirror comes from the partner lane.
    const int partner = (lane & ~(TPL - 1)) | ((TPL - t) & (TPL - 1));
    cpx<R> w[R1 / 2];  // conj(a) + i conj(b): the value at N - k
#pragma unroll
    for (int m = 0; m < R1 / 2; ++m) {
      cpx<R> z = {R(0), R(0)}, zm = {R(0), R(0)};
      if (!band_zero<N, KZ>(m * S1, m * S1 + S1 - 1)) {
        const int k = t + m * S1;
        const bool ok = k < KZ;
        cpx<R> a, b;
        fetch(ok ? k : KZ - 1, a, b);
        if (k == 0) { a.y = R(0); b.y = R(0); }
        if (ok) {
          z = {a.x - b.y, a.y + b.x};
          zm = {a.x + b.y, b.x - a.y};
        }
      }
      v[m] = z;
      w[m] = zm;
    }
#pragma unroll
    for (int m = R1 / 2; m < R1; ++m) {
      cpx<R> z = {R(0), R(0)};
      if (!band_zero<N, KZ>(m * S1, m * S1 + S1 - 1)) {
        const cpx<R> f = lane_get(w[R1 - 1 - m], partner);
        const cpx<R> own = (m == R1 / 2) ? cpx<R>{R(0), R(0)} : w[(R1 - m) % R1];
        z = t == 0 ? own : f;
      }
      v[m] = z;
    }
  } else {
#pragma unroll
    for (int j = 0; j < NB1; ++j) {
#pragma unroll
      for (int m = 0; m < R1; ++m) {
        cpx<R> z = {R(0), R(0)};
        if (!band_zero<N, KZ>(j + m * S1, j + m * S1 + S1 - NB1)) {
          const int k = t * NB1 + j + m * S1;
          const bool neg = k > N / 2;
          const int kzr = neg ? N - k : k;
          const bool ok = kzr < KZ;
          cpx<R> a, b;
          fetch(ok ? kzr : KZ - 1, a, b);
          if (k == 0) { a.y = R(0); b.y = R(0); }
          if (neg) { a.y = -a.y; b.y = -b.y; }
          if (ok) z = {a.x - b.y, a.y + b.x};
        }
        v[j * R1 + m] = z;
      }
    }
  }
}

// zw_load in two phases for lines whose first stage holds one point per
// radix slot (NB1 == 1, TPL a power of two): zw_fetch issues the loads of
// the slots m < R1/2 into ra/rb, zw_combine forms the packed sequence and
// its mirror half exactly as zw_load does.  A caller fetches several
// sequences before combining any, so their loads are in flight together
// instead of one sequence per memory round trip.
template <typename R, int N, class C, int FA, int FB>
__device__ __forceinline__ void zw_fetch(cpx<R>* ra, cpx<R>* rb, const cpx<R>* __restrict__ src, size_t NNKZ,
                                         int t) {
  using IP = typename C::IP;
  constexpr int E = C::E, TPL = C::TPL, KZ = Geo<N>::KZ;
  constexpr int R1 = FirstRadix<IP>::value, NB1 = E / R1, S1 = N / R1;
  static_assert(NB1 == 1 && (TPL & (TPL - 1)) == 0, "two-phase load needs one point per radix slot");
#pragma unroll
  for (int m = 0; m < R1 / 2; ++m) {
    ra[m] = rb[m] = {R(0), R(0)};
    if (!band_zero<N, KZ>(m * S1, m * S1 + S1 - 1)) {
      const int k = t + m * S1;
      const int kk = k < KZ ? k : KZ - 1;
      ra[m] = gld<nt_mid<N>()>(&src[FA * NNKZ + kk]);
      if constexpr (FB >= 0) rb[m] = gld<nt_mid<N>()>(&src[FB * NNKZ + kk]);
    }
  }
}

template <typename R, int N, class C>
__device__ __forceinline__ void zw_combine(cpx<R>* v, const cpx<R>* ra, const cpx<R>* rb, int t, int lane) {
  using IP = typename C::IP;
  constexpr int TPL = C::TPL, KZ = Geo<N>::KZ;
  constexpr int R1 = FirstRadix<IP>::value, S1 = N / R1;
  const int partner = (lane & ~(TPL - 1)) | ((TPL - t) & (TPL - 1));
  cpx<R> w[R1 / 2];  // conj(a) + i conj(b): the value at N - k
#pragma unroll
  for (int m = 0; m < R1 / 2; ++m) {
    cpx<R> z = {R(0), R(0)}, zm = {R(0), R(0)};
    if (!band_zero<N, KZ>(m * S1, m * S1 + S1 - 1)) {
      const int k = t + m * S1;
      cpx<R> a = ra[m], b = rb[m];
      if (k == 0) { a.y = R(0); b.y = R(0); }
      if (k < KZ) {
        z = {a.x - b.y, a.y + b.x};
        zm = {a.x + b.y, b.x - a.y};
      }
    }
    v[m] = z;
    w[m] = zm;
  }
#pragma unroll
  for (int m = R1 / 2; m < R1; ++m) {
    cpx<R> z = {R(0), R(0)};
    if (!band_zero<N, KZ>(m * S1, m * S1 + S1 - 1)) {
      const cpx<R> f = lane_get(w[R1 - 1 - m], partner);
      const cpx<R> own = (m == R1 / 2) ? cpx<R>{R(0), R(0)} : w[(R1 - m) % R1];
      z = t == 0 ? own : f;
    }
    v[m] = z;
  }
}

// Store the kept kz of the two real fields packed in the forward transform
// x (output order of plan FP) into fields FA (real part) and FB: they are
// split by Hermitian symmetry from the pair (k, N - k), through the line
// buffer.  Ends with the buffer free again.
// Split the forward transform x of the two real fields a + i b packed in it
// (output order of plan FP) into their kept kz coefficients, through the
// line buffer: fn(i, kz, a_kz, b_kz) for the lane's kz = t + i * TPL < KZ.
// Ends with the buffer free again.
template <typename R, int N, class C, class Fn>
__device__ __forceinline__ void zw_split_pair(const cpx<R>* x, cpx<R>* lb, int t, const Fn& fn) {
  using IP = typename C::IP;
  constexpr int E = C::E, TPL = C::TPL, PS = C::PS, KZ = Geo<N>::KZ;
  constexpr int R1 = FirstRadix<IP>::value, NB1 = E / R1, S1 = N / R1;
#pragma unroll
  for (int j = 0; j < NB1; ++j)
#pragma unroll
    for (int m = 0; m < R1; ++m) lb[pidx<PS>(t * NB1 + j + m * S1)] = x[j * R1 + m];
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < (KZ + TPL - 1) / TPL; ++i) {
    const int kz = t + i * TPL;
    if (kz < KZ) {
      const cpx<R> P = lb[pidx<PS>(kz)];
      const cpx<R> Qc = lb[pidx<PS>(kz == 0 ? 0 : N - kz)];
      fn(i, kz, cpx<R>{R(0.5) * (P.x + Qc.x), R(0.5) * (P.y - Qc.y)},
         cpx<R>{R(0.5) * (P.y + Qc.y), R(0.5) * (Qc.x - P.x)});
    }
  }
  wave_lds_sync();
}

// Store the kept kz of the two real fields packed in x into fields FA (real
// part) and FB of dst.
template <typename R, int N, class C, int FA, int FB, bool NT = nt_mid<N>()>
__device__ __forceinline__ void zw_store_pair(const cpx<R>* x, cpx<R>* lb, cpx<R>* __restrict__ dst, size_t NNKZ,
                                              int t) {
  zw_split_pair<R, N, C>(x, lb, t, [&](int, int kz, cpx<R> a, cpx<R> b) {
    gst<NT>(&dst[FA * NNKZ + kz], a);
    gst<NT>(&dst[FB * NNKZ + kz], b);
  });
}

template <typename R, int N, int E_, int WPB, int PITCH = kzp<R, N>()>
struct ZWBody {
  using G = Geo<N>;
  using C = CfgZW<R, N, E_, WPB>;
  using IP = typename C::IP;
  using FP = typename C::FP;
  static constexpr int SMEM = C::SMEM;
  __device__ __forceinline__ static void run(cpx<R>* sm, int blk, const cpx<R>* __restrict__ in,
                                             cpx<R>* __restrict__ out, const cpx<R>* __restrict__ tw) {
    constexpr int E = C::E, TPL = C::TPL, LW = C::LW, PS = C::PS, KZ = G::KZ;
    constexpr int KZP = PITCH;
    constexpr int R1 = FirstRadix<IP>::value, NB1 = E / R1, S1 = N / R1;
    const size_t NNKZ = size_t(N) * N * KZP;
    constexpr bool NT_OUT = nt_store_for(3LL * N * N * KZP * (long long)sizeof(cpx<R>));
    cpx<R>* tsi = sm;
    cpx<R>* tsf = C::SAME ? sm : sm + C::TWI;
    PlanTw<N, IP>::fill(tsi, tw, threadIdx.x, C::NT);
    if constexpr (!C::SAME) PlanTw<N, FP>::fill(tsf, tw, threadIdx.x, C::NT);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane % TPL, lw = lane / TPL;
    cpx<R>* lb = sm + C::TWI + C::TWF + (wave * LW + lw) * C::LPL;
    const size_t line = size_t(blk) * C::LINES + wave * LW + lw;
    const cpx<R>* src = in + line * KZP;
    cpx<R>* dst = out + line * KZP;

    // ---- the packed sequences u_p + i w_p
    cpx<R> v[3][E];
    if constexpr (NB1 == 1 && (TPL & (TPL - 1)) == 0) {
      cpx<R> ra[3][R1 / 2], rb[3][R1 / 2];  // every load of the three sequences first
      zw_fetch<R, N, C, 0, 3>(ra[0], rb[0], src, NNKZ, t);
      zw_fetch<R, N, C, 1, 4>(ra[1], rb[1], src, NNKZ, t);
      zw_fetch<R, N, C, 2, 5>(ra[2], rb[2], src, NNKZ, t);
#pragma unroll
      for (int p = 0; p < 3; ++p) zw_combine<R, N, C>(v[p], ra[p], rb[p], t, lane);
    } else {
      zw_load<R, N, C, 0, 3>(v[0], src, NNKZ, t, lane);
      zw_load<R, N, C, 1, 4>(v[1], src, NNKZ, t, lane);
      zw_load<R, N, C, 2, 5>(v[2], src, NNKZ, t, lane);
    }
    __syncthreads();  // twiddle rows staged
#pragma unroll
    for (int p = 0; p < 3; ++p) reg_fft<R, N, E, +1, PS>(v[p], t, lb, tsi, IP{});

    // ---- n = u x w at the lane's E points, packed (n_x + i n_y), (n_z + 0 i)
    cpx<R> a[E], c[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const R u[3] = {v[0][i].x, v[1][i].x, v[2][i].x};
      const R w[3] = {v[0][i].y, v[1][i].y, v[2][i].y};
      R n[3];
      cross3(u, w, n);
      a[i] = {n[0], n[1]};
      c[i] = {n[2], R(0)};
    }
    reg_fft<R, N, E, -1, PS>(a, t, lb, tsf, FP{});
    reg_fft<R, N, E, -1, PS>(c, t, lb, tsf, FP{});

    // ---- n_z: real input, kept coefficients straight from registers
#pragma unroll
    for (int j = 0; j < NB1; ++j)
#pragma unroll
      for (int m = 0; m < R1; ++m) {
        if (j + m * S1 < KZ) {
          const int k = t * NB1 + j + m * S1;
          if (k < KZ) gst<NT_OUT>(&dst[2 * NNKZ + k], c[j * R1 + m]);
        }
      }
    // ---- n_x, n_y from the pair (k, N - k) of the packed transform
    zw_store_pair<R, N, C, 0, 1, NT_OUT>(a, lb, dst, NNKZ, t);
  }
};

// Adjoint z pass with the lines in registers: per line the nine inverse
// fields u, w = curl u (fields 0-5 of `in`) and q (fields 0-2 of `qin`) as
// five packed sequences (u_x + i u_y), (u_z + i w_x), (w_y + i w_z), (q_x +
// i q_y), (q_z + 0 i); products p1 = w x q and p2 = q x u; three forward
// sequences (p1_x + i p1_y), (p1_z + i p2_x), (p2_y + i p2_z) stored as
// fields 0-5.  FOLD: fields 0 and 1 take the kz terms of i k x p2 as they
// are stored, p1_x - i kz p2_y and p1_y + i kz p2_x (k_y_adj folds the ky
// terms, the register forward x the kx terms).
template <typename R, int N, int E_, int WPB, bool FOLD = false>
struct ZWAdjBody {
  using G = Geo<N>;
  using C = CfgZW<R, N, E_, WPB>;
  using IP = typename C::IP;
  using FP = typename C::FP;
  static constexpr int SMEM = C::SMEM;
  __device__ __forceinline__ static void run(cpx<R>* sm, int blk, const cpx<R>* __restrict__ in,
                                             const cpx<R>* __restrict__ qin, cpx<R>* __restrict__ out,
                                             const cpx<R>* __restrict__ tw, R ks = R(1)) {
    constexpr int E = C::E, TPL = C::TPL, LW = C::LW, PS = C::PS;
    constexpr int KZP = kzp<R, N>();
    const size_t NNKZ = size_t(N) * N * KZP;
    cpx<R>* tsi = sm;
    cpx<R>* tsf = C::SAME ? sm : sm + C::TWI;
    PlanTw<N, IP>::fill(tsi, tw, threadIdx.x, C::NT);
    if constexpr (!C::SAME) PlanTw<N, FP>::fill(tsf, tw, threadIdx.x, C::NT);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane % TPL, lw = lane / TPL;
    cpx<R>* lb = sm + C::TWI + C::TWF + (wave * LW + lw) * C::LPL;
    const size_t line = size_t(blk) * C::LINES + wave * LW + lw;
    const cpx<R>* src = in + line * KZP;
    const cpx<R>* qsrc = qin + line * KZP;
    cpx<R>* dst = out + line * KZP;

    cpx<R> v[5][E];
    {
      constexpr int H = FirstRadix<IP>::value / 2;
      cpx<R> ra[5][H], rb[5][H];  // every load of the five sequences first
      zw_fetch<R, N, C, 0, 1>(ra[0], rb[0], src, NNKZ, t);
      zw_fetch<R, N, C, 2, 3>(ra[1], rb[1], src, NNKZ, t);
      zw_fetch<R, N, C, 4, 5>(ra[2], rb[2], src, NNKZ, t);
      zw_fetch<R, N, C, 0, 1>(ra[3], rb[3], qsrc, NNKZ, t);
      zw_fetch<R, N, C, 2, -1>(ra[4], rb[4], qsrc, NNKZ, t);
#pragma unroll
      for (int p = 0; p < 5; ++p) zw_combine<R, N, C>(v[p], ra[p], rb[p], t, lane);
    }
    __syncthreads();  // twiddle rows staged
#pragma unroll
    for (int p = 0; p < 5; ++p) reg_fft<R, N, E, +1, PS>(v[p], t, lb, tsi, IP{});
    cpx<R> f[3][E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const R u[3] = {v[0][i].x, v[0][i].y, v[1][i].x};
      const R w[3] = {v[1][i].y, v[2][i].x, v[2][i].y};
      const R q[3] = {v[3][i].x, v[3][i].y, v[4][i].x};
      R p1[3], p2[3];
      cross3(w, q, p1);
      cross3(q, u, p2);
      f[0][i] = {p1[0], p1[1]};
      f[1][i] = {p1[2], p2[0]};
      f[2][i] = {p2[1], p2[2]};
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) reg_fft<R, N, E, -1, PS>(f[p], t, lb, tsf, FP{});
    if constexpr (FOLD) {
      constexpr int NKZ = (G::KZ + TPL - 1) / TPL;
      cpx<R> px[NKZ], py[NKZ];  // p2_x, p2_y at the lane's kz
      zw_split_pair<R, N, C>(f[2], lb, t, [&](int i, int kz, cpx<R> a, cpx<R> b) {
        py[i] = a;
        gst<nt_mid<N>()>(&dst[4 * NNKZ + kz], a);
        gst<nt_mid<N>()>(&dst[5 * NNKZ + kz], b);
      });
      zw_split_pair<R, N, C>(f[1], lb, t, [&](int i, int kz, cpx<R> a, cpx<R> b) {
        px[i] = b;
        gst<nt_mid<N>()>(&dst[2 * NNKZ + kz], a);
        gst<nt_mid<N>()>(&dst[3 * NNKZ + kz], b);
      });
      zw_split_pair<R, N, C>(f[0], lb, t, [&](int i, int kz, cpx<R> a, cpx<R> b) {
        const R k = ks * R(kz);
        gst<nt_mid<N>()>(&dst[kz], cpx<R>{a.x + k * py[i].y, a.y - k * py[i].x});      // p1_x - i kz p2_y
        gst<nt_mid<N>()>(&dst[NNKZ + kz], cpx<R>{b.x - k * px[i].y, b.y + k * px[i].x});  // p1_y + i kz p2_x
      });
    } else {
      zw_store_pair<R, N, C, 0, 1>(f[0], lb, dst, NNKZ, t);
      zw_store_pair<R, N, C, 2, 3>(f[1], lb, dst, NNKZ, t);
      zw_store_pair<R, N, C, 4, 5>(f[2], lb, dst, NNKZ, t);
    }
  }
};

// MHD z pass with the lines in registers: per line the twelve inverse fields
// u (0-2), w = curl u (3-5), b (6-8), j = curl b (9-11) as six packed
// sequences u_p + i w_p and b_p + i j_p; products m = u x w + lorentz j x b
// and e = u x b; three forward sequences (m_x + i m_y), (m_z + i e_x),
// (e_y + i e_z) stored as fields 0-5.
template <typename R, int N, int E_, int WPB>
struct ZWMhdBody {
  using G = Geo<N>;
  using C = CfgZW<R, N, E_, WPB>;
  using IP = typename C::IP;
  using FP = typename C::FP;
  static constexpr int SMEM = C::SMEM;
  __device__ __forceinline__ static void run(cpx<R>* sm, int blk, const cpx<R>* __restrict__ in,
                                             cpx<R>* __restrict__ out, const cpx<R>* __restrict__ tw, R lorentz) {
    constexpr int E = C::E, TPL = C::TPL, LW = C::LW, PS = C::PS;
    constexpr int KZP = kzp<R, N>();
    const size_t NNKZ = size_t(N) * N * KZP;
    cpx<R>* tsi = sm;
    cpx<R>* tsf = C::SAME ? sm : sm + C::TWI;
    PlanTw<N, IP>::fill(tsi, tw, threadIdx.x, C::NT);
    if constexpr (!C::SAME) PlanTw<N, FP>::fill(tsf, tw, threadIdx.x, C::NT);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane % TPL, lw = lane / TPL;
    cpx<R>* lb = sm + C::TWI + C::TWF + (wave * LW + lw) * C::LPL;
    const size_t line = size_t(blk) * C::LINES + wave * LW + lw;
    const cpx<R>* src = in + line * KZP;
    cpx<R>* dst = out + line * KZP;

    cpx<R> v[6][E];
    {
      constexpr int H = FirstRadix<IP>::value / 2;
      cpx<R> ra[6][H], rb[6][H];  // every load of the six sequences first
      zw_fetch<R, N, C, 0, 3>(ra[0], rb[0], src, NNKZ, t);
      zw_fetch<R, N, C, 1, 4>(ra[1], rb[1], src, NNKZ, t);
      zw_fetch<R, N, C, 2, 5>(ra[2], rb[2], src, NNKZ, t);
      zw_fetch<R, N, C, 6, 9>(ra[3], rb[3], src, NNKZ, t);
      zw_fetch<R, N, C, 7, 10>(ra[4], rb[4], src, NNKZ, t);
      zw_fetch<R, N, C, 8, 11>(ra[5], rb[5], src, NNKZ, t);
#pragma unroll
      for (int p = 0; p < 6; ++p) zw_combine<R, N, C>(v[p], ra[p], rb[p], t, lane);
    }
    __syncthreads();  // twiddle rows staged
#pragma unroll
    for (int p = 0; p < 6; ++p) reg_fft<R, N, E, +1, PS>(v[p], t, lb, tsi, IP{});
    cpx<R> f[3][E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const R u[3] = {v[0][i].x, v[1][i].x, v[2][i].x};
      const R w[3] = {v[0][i].y, v[1][i].y, v[2][i].y};
      const R b[3] = {v[3][i].x, v[4][i].x, v[5][i].x};
      const R j[3] = {v[3][i].y, v[4][i].y, v[5][i].y};
      R uw[3], jb[3], e[3];
      cross3(u, w, uw);
      cross3(j, b, jb);
      cross3(u, b, e);
      const R m[3] = {uw[0] + lorentz * jb[0], uw[1] + lorentz * jb[1], uw[2] + lorentz * jb[2]};
      f[0][i] = {m[0], m[1]};
      f[1][i] = {m[2], e[0]};
      f[2][i] = {e[1], e[2]};
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) reg_fft<R, N, E, -1, PS>(f[p], t, lb, tsf, FP{});
    zw_store_pair<R, N, C, 0, 1>(f[0], lb, dst, NNKZ, t);
    zw_store_pair<R, N, C, 2, 3>(f[1], lb, dst, NNKZ, t);
    zw_store_pair<R, N, C, 4, 5>(f[2], lb, dst, NNKZ, t);
  }
};

// Default register z-pass shape of a grid size: points per lane and
// wavefronts per block (E = 0: the grid uses the LDS-resident kernel).  A
// wavefront holds whole lines, so N / E divides 64.
template <int N>
struct ZwDefault {
  static constexpr int E = N % 3 == 0 ? (N >= 48 ? 12 : 0) : (N >= 128 ? 16 : (N >= 32 ? 8 : 0));
  static constexpr int WPB = N >= 128 ? 4 : 1;
  // the fewer-points register variant of N >= 128
  static constexpr int E_ALT = N % 3 == 0 ? 6 : 8;
};

// Occupancy target (wavefronts per SIMD, i.e. a register budget of 512 /
// target per lane): FP64 with 16 points per lane keeps 192 registers of
// line data alone; in FP32 the 12- and 16-point plans spill under a
// 128-register budget.
template <typename R, int E>
constexpr int zw_waves_per_eu() {
  return sizeof(R) == 8 ? (E >= 12 ? 2 : 3) : (E >= 12 ? 3 : 4);
}

template <typename R, int N, int E, int WPB>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(zw_waves_per_eu<R, E>())))
void k_zw(const cpx<R>* __restrict__ in, cpx<R>* __restrict__ out, const cpx<R>* __restrict__ tw) {
  using B = ZWBody<R, N, E, WPB>;
  __shared__ cpx<R> sm[B::SMEM];
  B::run(sm, blockIdx.x, in, out, tw);
}

// Register-resident MHD z pass: E points per lane (six sequences per line),
// 8 (variant 0) or 4 (variant 1: half the line registers, more waves).
template <typename R, int E>
constexpr int zw_mhd_waves_per_eu() { return E <= 4 ? (sizeof(R) == 8 ? 3 : 4) : (sizeof(R) == 8 ? 2 : 3); }

template <typename R, int N, int E, int WPB>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(zw_mhd_waves_per_eu<R, E>())))
void k_zw_mhd(const cpx<R>* __restrict__ in, cpx<R>* __restrict__ out, const cpx<R>* __restrict__ tw, R lorentz) {
  using B = ZWMhdBody<R, N, E, WPB>;
  __shared__ cpx<R> sm[B::SMEM];
  B::run(sm, blockIdx.x, in, out, tw, lorentz);
}

// Register-resident adjoint z pass (five sequences per line): adj_ze points
// per lane, eight in FP64 (160 registers of line data, two wavefronts per
// SIMD; four points gave four wavefronts and one more exchange stage and ran
// slower) and sixteen in FP32 (one exchange stage fewer than eight).
template <typename R>
constexpr int adj_ze() { return sizeof(R) == 8 ? 8 : 16; }

template <typename R, int N, int WPB, bool FOLD>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(2)))
void k_zw_adj(const cpx<R>* __restrict__ in, const cpx<R>* __restrict__ qin, cpx<R>* __restrict__ out,
              const cpx<R>* __restrict__ tw, R ks) {
  using B = ZWAdjBody<R, N, adj_ze<R>(), WPB, FOLD>;
  __shared__ cpx<R> sm[B::SMEM];
  B::run(sm, blockIdx.x, in, qin, out, tw, ks);
}

// ----------------------------------------------------------------- pass 5
// Forward transform along x of FF fields (3 or 6) for one (iy, kz tile),
// keeping the K masked rows, followed by the spectral epilogue: Leray
// projection, viscous/resistive term and the RK4 stage update (or a plain
// store of the transform / of the right-hand side).
template <typename R>
struct StageArgs {
  const cpx<R>* u0;   // state at the start of the step (u) / (b)
  const cpx<R>* b0;
  const cpx<R>* us;   // stage input
  const cpx<R>* bs;
  cpx<R>* unext;      // next stage input (stages 0-2) ; u0 itself at stage 3
  cpx<R>* bnext;
  cpx<R>* uacc;       // r1 + 2 r2 + 2 r3 accumulator
  cpx<R>* bacc;
  R dt, nu, eta, ks;
  int stage;
  int cpz;  // row pitch of the compact arrays (0 = dense KZ)
  // adjoint (FM_ADJ): mu = P-less VJP of the RHS at the stage state xs for
  // the cotangent a_s = alpha * a;  lacc <- (lamp | lacc | 0) + mu by
  // `first` (1 | 0 | 2);  anext <- clam * lamp + cmu * mu unless `last`;
  // nupart[block] += sum w k^2 Re(conj(a_s) xs)
  const cpx<R>* a;
  const cpx<R>* lamp;
  cpx<R>* lacc;
  const cpx<R>* xs;
  cpx<R>* anext;
  R alpha, clam, cmu;
  int first, last;
  double* nupart;
  // register forward-x only: when qout is set, the pass also leaves the
  // next stage VJP's first pass done: q = P(qalpha a_next) (a_next = the
  // lacc it stores when `last`, else anext) inverse-transformed along x
  // into qout, laid out as k_inv_x<XM_ADJQ> writes it
  cpx<R>* qout;
  R qalpha;
};

// Sum of v over the block, returned to thread 0 in a fixed order (the
// result is reproducible): a butterfly over each wavefront, then thread 0
// adds the wavefront sums; blocks that are not whole wavefronts add every
// thread's value in turn.  red: one double per thread of LDS, free on entry.
// Every thread must call it.
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int nt = int(blockDim.x);
  const bool waves = nt % 64 == 0;
  if (waves) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  } else {
    red[threadIdx.x] = v;
  }
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (waves ? nt / 64 : nt); ++i) s += red[i];
  return s;
}

// Adjoint stage bookkeeping of one compact point (the FM_ADJ contract of
// StageArgs) given w = the forward-x output of the product's VJP through u
// plus the curl of the field that adds to it (so mu = w - nu k^2 a_s).
// adj_point_load issues every load of the point; adj_point_store forms and
// stores its results and returns the point's term of the nu gradient.  A
// caller issues a point's loads before any of its stores.
template <typename R>
struct AdjIn {
  cpx<R> ar[3], lp[3], l0[3], xv[3];
};

template <typename R>
__device__ __forceinline__ AdjIn<R> adj_point_load(const StageArgs<R>& a, size_t pt, size_t cs) {
  AdjIn<R> in;
  const bool uselp = a.first == 1 || !a.last;
#pragma unroll
  for (int c = 0; c < 3; ++c) in.ar[c] = a.a[c * cs + pt];
#pragma unroll
  for (int c = 0; c < 3; ++c) in.lp[c] = uselp ? a.lamp[c * cs + pt] : cpx<R>{R(0), R(0)};
#pragma unroll
  for (int c = 0; c < 3; ++c) in.l0[c] = a.first == 0 ? a.lacc[c * cs + pt] : cpx<R>{R(0), R(0)};
#pragma unroll
  for (int c = 0; c < 3; ++c) in.xv[c] = a.xs ? a.xs[c * cs + pt] : cpx<R>{R(0), R(0)};
  return in;
}

// an (if not null) receives the next stage's cotangent at the point: the
// lacc it stores when `last`, else anext.
template <typename R>
__device__ __forceinline__ double adj_point_store(const StageArgs<R>& a, size_t pt, size_t cs, R k2, int iz,
                                                  const cpx<R> (&w)[3], const AdjIn<R>& in,
                                                  cpx<R>* an = nullptr) {
  const R nk = a.nu * k2;
  const double wk = (iz == 0 ? 1.0 : 2.0) * double(k2);
  double nus = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t o = c * cs + pt;
    const cpx<R> as = {a.alpha * in.ar[c].x, a.alpha * in.ar[c].y};
    const cpx<R> mu = {w[c].x - nk * as.x, w[c].y - nk * as.y};
    cpx<R> acc = mu;
    if (a.first == 1) acc = {in.lp[c].x + mu.x, in.lp[c].y + mu.y};
    else if (a.first == 0) acc = {in.l0[c].x + mu.x, in.l0[c].y + mu.y};
    if (a.xs) nus += wk * (double(as.x) * double(in.xv[c].x) + double(as.y) * double(in.xv[c].y));
    a.lacc[o] = acc;
    const cpx<R> nx = {a.clam * in.lp[c].x + a.cmu * mu.x, a.clam * in.lp[c].y + a.cmu * mu.y};
    if (!a.last) a.anext[o] = nx;
    if (an) an[c] = a.last ? acc : nx;
  }
  return nus;
}

template <typename R>
__device__ __forceinline__ double adj_point_w(const StageArgs<R>& a, size_t pt, size_t cs, R k2, int iz,
                                              const cpx<R> (&w)[3]) {
  return adj_point_store(a, pt, cs, k2, iz, w, adj_point_load(a, pt, cs));
}

// The same from the six forward-x outputs: v[0..2] the product's VJP
// through u, v[3..5] the field whose curl adds to it.
template <typename R>
__device__ __forceinline__ double adj_point(const StageArgs<R>& a, size_t pt, size_t cs, R kx, R ky, R kz,
                                            int iz, const cpx<R> (&v)[6]) {
  cpx<R> w[3];
  w[0] = cadd(v[0], cmuli(cpx<R>{ky * v[5].x - kz * v[4].x, ky * v[5].y - kz * v[4].y}));
  w[1] = cadd(v[1], cmuli(cpx<R>{kz * v[3].x - kx * v[5].x, kz * v[3].y - kx * v[5].y}));
  w[2] = cadd(v[2], cmuli(cpx<R>{kx * v[4].x - ky * v[3].x, kx * v[4].y - ky * v[3].y}));
  return adj_point_w(a, pt, cs, kx * kx + ky * ky + kz * kz, iz, w);
}

// Leray projection n - k (k.n) / k2 (k2s: k2 with 0 -> 1).  FP64 divides
// each k_c (k.n) by k2 as the reference writes it; FP32, whose passes are
// bound by instruction issue, multiplies by one reciprocal instead (a last-
// place difference, far inside its error budget).
template <typename R>
__device__ __forceinline__ void project(R kx, R ky, R kz, R k2s, cpx<R>* n) {
  const cpx<R> kd = {kx * n[0].x + ky * n[1].x + kz * n[2].x, kx * n[0].y + ky * n[1].y + kz * n[2].y};
  if constexpr (sizeof(R) == 4) {
    const R q = R(1) / k2s;
    n[0] = {n[0].x - kx * kd.x * q, n[0].y - kx * kd.y * q};
    n[1] = {n[1].x - ky * kd.x * q, n[1].y - ky * kd.y * q};
    n[2] = {n[2].x - kz * kd.x * q, n[2].y - kz * kd.y * q};
  } else {
    n[0] = {n[0].x - kx * kd.x / k2s, n[0].y - kx * kd.y / k2s};
    n[1] = {n[1].x - ky * kd.x / k2s, n[1].y - ky * kd.y / k2s};
    n[2] = {n[2].x - kz * kd.x / k2s, n[2].y - kz * kd.y / k2s};
  }
}

// RK4 bookkeeping shared by the u and b halves of the state, split into the
// loads of a point (x0, and the accumulator after stage 0) and the update
// and stores: a caller issues every load of its point before any store, so
// the point costs one memory round trip (loads that follow a store in
// program order also wait for that store).
template <typename R>
struct RK4In {
  cpx<R> x0[3], acc[3];
};

template <typename R>
__device__ __forceinline__ RK4In<R> rk4_load(int stage, size_t pt, size_t cs, const cpx<R>* x0, const cpx<R>* acc) {
  RK4In<R> in;
#pragma unroll
  for (int c = 0; c < 3; ++c) in.x0[c] = x0[c * cs + pt];
#pragma unroll
  for (int c = 0; c < 3; ++c) in.acc[c] = stage > 0 ? acc[c * cs + pt] : cpx<R>{R(0), R(0)};
  return in;
}

template <typename R>
__device__ __forceinline__ void rk4_store(int stage, R dt, size_t pt, size_t cs, const cpx<R>* r, const RK4In<R>& in,
                                          cpx<R>* xnext, cpx<R>* acc) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t o = c * cs + pt;
    const cpx<R> a0 = in.x0[c];
    if (stage == 0) {
      acc[o] = r[c];
      const R h = R(0.5) * dt;
      xnext[o] = {a0.x + h * r[c].x, a0.y + h * r[c].y};
    } else if (stage < 3) {
      const cpx<R> s = in.acc[c];
      acc[o] = {s.x + R(2) * r[c].x, s.y + R(2) * r[c].y};
      const R h = (stage == 1) ? R(0.5) * dt : dt;
      xnext[o] = {a0.x + h * r[c].x, a0.y + h * r[c].y};
    } else {
      const cpx<R> s = in.acc[c];
      const R h = dt / R(6);
      xnext[o] = {a0.x + h * (s.x + r[c].x), a0.y + h * (s.y + r[c].y)};
    }
  }
}

template <typename R>
__device__ __forceinline__ void rk4_update(int stage, R dt, size_t pt, size_t cs, const cpx<R>* r,
                                           const cpx<R>* x0, cpx<R>* xnext, cpx<R>* acc) {
  rk4_store(stage, dt, pt, cs, r, rk4_load(stage, pt, cs, x0, acc), xnext, acc);
}

template <typename R, int N, int MODE, int V>
struct XFBody {
  using G = Geo<N>;
  static constexpr int LINES = CfgXF<R, N, V>::LINES;
  static constexpr int SMEM = N * LINES + N;
  // Block (bx, by) = (kz tile, iy).
  __device__ __forceinline__ static void run(cpx<R>* sm, int bx, int by, const cpx<R>* __restrict__ in, const StageArgs<R>& a,
                             const cpx<R>* __restrict__ tw) {
  constexpr int FF = (MODE == FM_MHD_STAGE || MODE == FM_MHD_RHS || MODE == FM_ADJ) ? 6 : 3;
  constexpr int NT = CfgXF<R, N, V>::NT, E = CfgXF<R, N, V>::E;
  double nusum = 0.0;
  constexpr int L = LINES / FF;
  constexpr int LP = LINES;
  cpx<R>* buf = sm;
  cpx<R>* tws = sm + N * LP;
  const int iz0 = bx * L, iy = by;
  const int tid = threadIdx.x;
  stage_twiddles<R, N>(tws, tw);
  auto ld = [&](int n, int ln, int) -> cpx<R> {
    const int f = ln / L, izr = iz0 + ln % L;
    const bool ok = izr < G::KZ;
    const cpx<R> v = gld<nt_mid<N>()>(&in[((size_t(f) * N + n) * G::K + iy) * kzp<R, N>() + (ok ? izr : G::KZ - 1)]);
    return ok ? v : cpx<R>{R(0), R(0)};
  };
  fft_io<R, N, E, -1, 1, false>(buf, LP, LINES, tws, ld, NoIO{});
  const int cpz = a.cpz ? a.cpz : G::KZ;
  const size_t cs = size_t(G::K) * G::K * cpz;
  const R ky = a.ks * R(G::kint(iy));
  for (int idx = tid; tid < NT && idx < G::K * L; idx += NT) {
    const int ll = idx % L, ix = idx / L, iz = iz0 + ll;
    if (iz >= G::KZ) continue;
    const int n = G::full_of(ix);
    const size_t pt = (size_t(ix) * G::K + iy) * cpz + iz;
    cpx<R> v[FF];
#pragma unroll
    for (int f = 0; f < FF; ++f) v[f] = buf[n * LP + f * L + ll];
    if constexpr (MODE == FM_STORE3) {
#pragma unroll
      for (int c = 0; c < 3; ++c) a.unext[c * cs + pt] = v[c];
      continue;
    } else if constexpr (MODE == FM_ADJ) {
      const R kx = a.ks * R(G::kint(ix)), kz = a.ks * R(iz);
      nusum += adj_point(a, pt, cs, kx, ky, kz, iz, v);
      continue;
    } else {
      const R kx = a.ks * R(G::kint(ix)), kz = a.ks * R(iz);
      const R k2 = kx * kx + ky * ky + kz * kz;
      const R k2s = (k2 == R(0)) ? R(1) : k2;
      // every load of the point first (see rk4_load)
      cpx<R> su[3], sbv[3];
      RK4In<R> inu{}, inb{};
#pragma unroll
      for (int c = 0; c < 3; ++c) su[c] = a.us[c * cs + pt];
      if constexpr (MODE == FM_HYDRO_STAGE || MODE == FM_MHD_STAGE) inu = rk4_load(a.stage, pt, cs, a.u0, a.uacc);
      if constexpr (FF == 6) {
#pragma unroll
        for (int c = 0; c < 3; ++c) sbv[c] = a.bs[c * cs + pt];
        if constexpr (MODE == FM_MHD_STAGE) inb = rk4_load(a.stage, pt, cs, a.b0, a.bacc);
      }
      // momentum: r_u = P(u x w [+ lorentz j x b]) - nu k2 u_s
      project(kx, ky, kz, k2s, v);
      cpx<R> ru[3];
      const R nk = a.nu * k2;
#pragma unroll
      for (int c = 0; c < 3; ++c) ru[c] = {v[c].x - nk * su[c].x, v[c].y - nk * su[c].y};
      if constexpr (FF == 6) {
        // induction: r_b = P(i k x e) - eta k2 b_s
        const cpx<R> e0 = v[3], e1 = v[4], e2 = v[5];
        cpx<R> cb[3];
        cb[0] = cmuli(cpx<R>{ky * e2.x - kz * e1.x, ky * e2.y - kz * e1.y});
        cb[1] = cmuli(cpx<R>{kz * e0.x - kx * e2.x, kz * e0.y - kx * e2.y});
        cb[2] = cmuli(cpx<R>{kx * e1.x - ky * e0.x, kx * e1.y - ky * e0.y});
        project(kx, ky, kz, k2s, cb);
        const R ek = a.eta * k2;
        cpx<R> rb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) rb[c] = {cb[c].x - ek * sbv[c].x, cb[c].y - ek * sbv[c].y};
        if constexpr (MODE == FM_MHD_RHS) {
#pragma unroll
          for (int c = 0; c < 3; ++c) { a.unext[c * cs + pt] = ru[c]; a.bnext[c * cs + pt] = rb[c]; }
        } else {
          rk4_store(a.stage, a.dt, pt, cs, ru, inu, a.unext, a.uacc);
          rk4_store(a.stage, a.dt, pt, cs, rb, inb, a.bnext, a.bacc);
        }
      } else {
        if constexpr (MODE == FM_HYDRO_RHS) {
#pragma unroll
          for (int c = 0; c < 3; ++c) a.unext[c * cs + pt] = ru[c];
        } else {
          rk4_store(a.stage, a.dt, pt, cs, ru, inu, a.unext, a.uacc);
        }
      }
    }
  }
  if constexpr (MODE == FM_ADJ) {
    if (a.nupart) {
      __syncthreads();  // every thread is done reading buf
      const double t = block_sum(nusum, reinterpret_cast<double*>(buf));
      if (threadIdx.x == 0) {
        a.nupart[size_t(by) * gridDim.x + bx] += t;
      }
    }
  }
  }
};

template <typename R, int N, int MODE, int V>
__global__ __launch_bounds__((CfgXF<R, N, V>::NT)) void k_fwd_x(const cpx<R>* __restrict__ in,
                                                              StageArgs<R> a,
                                                              const cpx<R>* __restrict__ tw) {
  using B = XFBody<R, N, MODE, V>;
  __shared__ cpx<R> sm[B::SMEM];
  B::run(sm, blockIdx.x, blockIdx.y, in, a, tw);
}

// Forward x + hydro stage epilogue with the transformed lines kept in
// registers: the block transforms the three fields of its L kz columns one
// after the other through one N x L LDS tile; the last stage of each leaves
// the thread's E outputs of its column in registers, so the projection,
// viscous term and RK4 update run on the thread's own (kx, kz) points with
// no LDS round trip.  LDS holds one field instead of three, so several
// blocks share a CU and overlap their load, transform and update phases.
template <typename R, int N, bool ADJ = false>
struct CfgXR {
  // power-of-two grids of N >= 128 only (use_xr); 4 points per thread at
  // N = 128 (twice the wavefronts of 8 on the small grid), else 8; the FP32
  // RK4 stage at N = 256 takes 16 (one LDS exchange instead of two, three
  // wavefronts per SIMD without spills; the adjoint's five fields and its
  // epilogue spill at 16)
  static constexpr int E = N <= 128 ? 4 : (sizeof(R) == 4 && N == 256 && !ADJ ? 16 : 8);
  // one 128-byte line of kz per row, halved in FP32 above 512 threads
  static constexpr int L = (sizeof(R) == 8 || 16 * N / E <= 512) ? 128 / int(2 * sizeof(R)) : 8;
  static constexpr int NT = L * N / E;
};

// Grids whose RK4 stage variant 1 and adjoint forward-x pass run the
// register-epilogue kernel: 8 points per lane need N a power of two.
template <int N>
constexpr bool use_xr() { return N >= 128 && (N & (N - 1)) == 0; }

// Epilogue over the E register slots of a thread.  AHEAD: the loads of
// slot s + 1 are issued before the stores of slot s (loads and stores share
// one in-order counter, so loads issued after a store also wait for that
// store); otherwise each slot loads, then stores, holding one slot's
// operands (fewer registers).
template <int E, class In, bool AHEAD = true, class Ld, class Fin>
__device__ __forceinline__ void pipelined_points(const Ld& load, const Fin& finish) {
  if constexpr (!AHEAD) {
#pragma unroll
    for (int sl = 0; sl < E; ++sl) finish(sl, load(sl));
  } else {
    In cur = load(0);
#pragma unroll
    for (int sl = 0; sl < E; ++sl) {
      In nxt;
      if (sl + 1 < E) nxt = load(sl + 1);
      finish(sl, cur);
      if (sl + 1 < E) cur = nxt;
    }
  }
}

// Epilogues of the register forward-x pass: hydro or MHD RK4 stage, or the
// adjoint bookkeeping of one stage VJP (the FM_ADJ epilogue of k_fwd_x).
// XR_ADJ5: the adjoint from the five fields of k_y_adj (only the kx terms
// of i k x p2 left to fold).
enum XRKind { XR_HYDRO = 0, XR_MHD = 1, XR_ADJ = 2, XR_ADJ5 = 3 };

template <typename R, int N, int KIND = XR_HYDRO>
struct XRBody {
  using G = Geo<N>;
  using C = CfgXR<R, N, KIND == XR_ADJ || KIND == XR_ADJ5>;
  static constexpr bool MHD = KIND == XR_MHD;
  static constexpr int SMEM = N * C::L + N;
  // fields per block: n (hydro); m or e (MHD: blockIdx.z = 0 updates u from
  // m, 1 updates b from e); adjoint 6
  static constexpr int FF = KIND == XR_ADJ ? 6 : 3;
  __device__ __forceinline__ static void run(cpx<R>* sm, int bx, int by, const cpx<R>* __restrict__ in,
                                             const StageArgs<R>& a, const cpx<R>* __restrict__ tw) {
    constexpr int L = C::L, E = C::E;
    using Pl = typename PlanFor<N, E>::type;
    constexpr int RL = LastRadix<Pl>::value, NBL = E / RL, SL = N / RL;
    cpx<R>* buf = sm;
    cpx<R>* tws = sm + N * L;
    const int iz0 = bx * L, iy = by;
    const int tid = threadIdx.x, l = tid % L, t = tid / L;
    const int izr = iz0 + l;
    const bool colok = izr < G::KZ;
    const int iz = colok ? izr : G::KZ - 1;
    stage_twiddles<R, N>(tws, tw);
    // First-stage points of field f of `in` (slot j * RAD0 + m holds point
    // t * NB0 + j + m * SIN0).  Each field's loads are issued while the
    // previous field is transformed (the barriers between fields wait on LDS
    // only), so a block pays one load latency instead of one per field.
    constexpr int RAD0 = FirstRadix<Pl>::value, NB0 = E / RAD0, SIN0 = N / RAD0;
    auto fetch = [&](int f, cpx<R>(&d)[E]) {
#pragma unroll
      for (int j = 0; j < NB0; ++j)
#pragma unroll
        for (int m = 0; m < RAD0; ++m) {
          const int n = t * NB0 + j + m * SIN0;
          const cpx<R> v = gld<nt_mid<N>()>(&in[((size_t(f) * N + n) * G::K + iy) * kzp<R, N>() + iz]);
          d[j * RAD0 + m] = colok ? v : cpx<R>{R(0), R(0)};
        }
    };
    if constexpr (KIND == XR_ADJ || KIND == XR_ADJ5) {
      // w = v + i k x p, folded as each p component leaves its transform: only
      // the three w fields and one transformed field are live at a time
      constexpr int NF = KIND == XR_ADJ5 ? 5 : 6;
      const R ky = a.ks * R(G::kint(iy)), kz = a.ks * R(iz);
      cpx<R> w[3][E];
      cpx<R> cur[E], nxt[E];
      // FP64 at 8 points per lane keeps no registers for the next field
      constexpr bool PREF = !(sizeof(R) == 8 && E >= 8);
      fetch(0, cur);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        if constexpr (PREF) {
          if (f + 1 < NF) fetch(f + 1, nxt);
        } else {
          if (f) fetch(f, cur);
        }
        if (f) __syncthreads();  // the previous field's last LDS reads are done
        auto ld = [&](int, int, int slot) -> cpx<R> { return cur[slot]; };
        // (i k x p)_x = i (ky p_z - kz p_y), _y = i (kz p_x - kx p_z), _z = i (kx p_y - ky p_x)
        auto st = [&](int n, int, cpx<R> p, int slot) {
          if (f < 3) {
            w[f][slot] = p;
            return;
          }
          const R kx = a.ks * R(G::kint(G::kept(n) ? G::compact_of(n) : 0));
          if constexpr (KIND == XR_ADJ5) {
            // fields 3, 4: p2_y -> w_z += i kx p2_y, p2_z -> w_y -= i kx p2_z
            if (f == 3) w[2][slot] = cadd(w[2][slot], cmuli(cpx<R>{kx * p.x, kx * p.y}));
            else w[1][slot] = csub(w[1][slot], cmuli(cpx<R>{kx * p.x, kx * p.y}));
            return;
          }
          if (f == 3) {
            w[1][slot] = cadd(w[1][slot], cmuli(cpx<R>{kz * p.x, kz * p.y}));
            w[2][slot] = csub(w[2][slot], cmuli(cpx<R>{ky * p.x, ky * p.y}));
          } else if (f == 4) {
            w[0][slot] = csub(w[0][slot], cmuli(cpx<R>{kz * p.x, kz * p.y}));
            w[2][slot] = cadd(w[2][slot], cmuli(cpx<R>{kx * p.x, kx * p.y}));
          } else {
            w[0][slot] = cadd(w[0][slot], cmuli(cpx<R>{ky * p.x, ky * p.y}));
            w[1][slot] = csub(w[1][slot], cmuli(cpx<R>{kx * p.x, kx * p.y}));
          }
        };
        fft_io<R, N, E, -1, 1, true>(buf, L, L, tws, ld, st);
        if constexpr (PREF) {
#pragma unroll
          for (int i = 0; i < E; ++i) cur[i] = nxt[i];
        }
      }
      cpx<R> qn[3][E];
      adjoint_epilogue(buf, w, colok, t, iy, iz, a, qn);
      if (a.qout) {
        // the next stage VJP's inverse x of q: the reversed plan's first stage
        // reads the points in the slots the epilogue left them in
        using IP = ReversePlan<Pl>;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          __syncthreads();  // buf is free (previous transform, nu reduction)
          auto ld = [&](int, int, int slot) -> cpx<R> { return qn[f][slot]; };
          auto st = [&](int n, int ll, cpx<R> v, int) {
            const int izw = iz0 + ll;
            if (izw < G::KZ)
              gst<nt_store_for(3LL * N * G::K * kzp<R, N>() * (long long)sizeof(cpx<R>))>(
                  &a.qout[((size_t(f) * N + n) * G::K + iy) * kzp<R, N>() + izw], v);
          };
          fft_io_plan<R, N, E, +1, 1, true, IP>(buf, L, L, tws, ld, st);
        }
      }
      return;
    }
    const bool mag = MHD && blockIdx.z != 0;
    const int f0 = mag ? 3 : 0;
    cpx<R> o[FF][E];
    cpx<R> cur[E], nxt[E];
    fetch(f0, cur);
#pragma unroll
    for (int f = 0; f < FF; ++f) {
      if (f + 1 < FF) fetch(f0 + f + 1, nxt);
      if (f) __syncthreads();  // the previous field's last LDS reads are done
      auto ld = [&](int, int, int slot) -> cpx<R> { return cur[slot]; };
      auto st = [&](int, int, cpx<R> v, int slot) { o[f][slot] = v; };
      fft_io<R, N, E, -1, 1, true>(buf, L, L, tws, ld, st);
#pragma unroll
      for (int i = 0; i < E; ++i) cur[i] = nxt[i];
    }
    if (!colok) return;
    const int cpz = a.cpz ? a.cpz : G::KZ;
    const size_t cs = size_t(G::K) * G::K * cpz;
    const R ky = a.ks * R(G::kint(iy)), kz = a.ks * R(iz);
    // The u half (momentum from n or m) or the b half (induction from e) of
    // every kept output point: the loads of the stage input and the RK4
    // operands, then the update and its stores.
    const cpx<R>* xsrc = mag ? a.bs : a.us;
    const cpx<R>* x0 = mag ? a.b0 : a.u0;
    const cpx<R>* xacc = mag ? a.bacc : a.uacc;
    struct PtIn {
      cpx<R> s[3];
      RK4In<R> in;
    };
    auto load = [&](int sl) -> PtIn {
      PtIn p;
      const int n = t * NBL + sl / RL + (sl % RL) * SL;
      if (G::kept(n)) {
        const size_t pt = (size_t(G::compact_of(n)) * G::K + iy) * cpz + iz;
#pragma unroll
        for (int c = 0; c < 3; ++c) p.s[c] = xsrc[c * cs + pt];
        p.in = rk4_load(a.stage, pt, cs, x0, xacc);
      }
      return p;
    };
    auto finish = [&](int sl, const PtIn& p) {
      const int n = t * NBL + sl / RL + (sl % RL) * SL;
      if (!G::kept(n)) return;
      const int ix = G::compact_of(n);
      const size_t pt = (size_t(ix) * G::K + iy) * cpz + iz;
      cpx<R> v[3] = {o[0][sl], o[1][sl], o[2][sl]};
      const R kx = a.ks * R(G::kint(ix));
      const R k2 = kx * kx + ky * ky + kz * kz;
      const R k2s = (k2 == R(0)) ? R(1) : k2;
      if (!mag) {
        // momentum: r_u = P(n or m) - nu k2 u_s
        project(kx, ky, kz, k2s, v);
        cpx<R> ru[3];
        const R nk = a.nu * k2;
#pragma unroll
        for (int c = 0; c < 3; ++c) ru[c] = {v[c].x - nk * p.s[c].x, v[c].y - nk * p.s[c].y};
        rk4_store(a.stage, a.dt, pt, cs, ru, p.in, a.unext, a.uacc);
      }
      if constexpr (MHD) {
        if (!mag) return;
        // induction: r_b = P(i k x e) - eta k2 b_s
        const cpx<R> e0 = v[0], e1 = v[1], e2 = v[2];
        cpx<R> cb[3];
        cb[0] = cmuli(cpx<R>{ky * e2.x - kz * e1.x, ky * e2.y - kz * e1.y});
        cb[1] = cmuli(cpx<R>{kz * e0.x - kx * e2.x, kz * e0.y - kx * e2.y});
        cb[2] = cmuli(cpx<R>{kx * e1.x - ky * e0.x, kx * e1.y - ky * e0.y});
        project(kx, ky, kz, k2s, cb);
        const R ek = a.eta * k2;
        cpx<R> rb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) rb[c] = {cb[c].x - ek * p.s[c].x, cb[c].y - ek * p.s[c].y};
        rk4_store(a.stage, a.dt, pt, cs, rb, p.in, a.bnext, a.bacc);
      }
    };
    // at 8 points per thread the next point's operands cost more in
    // registers (FP64: spills under its three-wavefront budget; FP32: a
    // wavefront per SIMD) than the overlap gains (measured at 256^3)
    pipelined_points<E, PtIn, (E < 8)>(load, finish);
  }

  // Stage-VJP bookkeeping on the register-held lines (w = the VJP of the
  // product through u plus the curl of the field that adds to it), as the
  // FM_ADJ epilogue of XFBody states it; the block's share of the nu
  // gradient goes to a.nupart[block].  Every thread reaches the barriers.
  // qn: the next stage's projected cotangent q = P(qalpha a_next) / N^3 at
  // each slot (zero off the kept kx), filled when a.qout is set.
  __device__ __forceinline__ static void adjoint_epilogue(cpx<R>* buf, const cpx<R> (&w)[3][C::E], bool colok,
                                                          int t, int iy, int iz, const StageArgs<R>& a,
                                                          cpx<R> (&qn)[3][C::E]) {
    constexpr int E = C::E;
    using Pl = typename PlanFor<N, E>::type;
    constexpr int RL = LastRadix<Pl>::value, NBL = E / RL, SL = N / RL;
    double nusum = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int sl = 0; sl < E; ++sl) qn[c][sl] = {R(0), R(0)};
    if (colok) {
      const int cpz = a.cpz ? a.cpz : G::KZ;
      const size_t cs = size_t(G::K) * G::K * cpz;
      const R ky = a.ks * R(G::kint(iy)), kz = a.ks * R(iz);
      const R qs = a.qalpha / (R(N) * R(N) * R(N));
      auto load = [&](int sl) -> AdjIn<R> {
        AdjIn<R> p;
        const int n = t * NBL + sl / RL + (sl % RL) * SL;
        if (G::kept(n)) p = adj_point_load(a, (size_t(G::compact_of(n)) * G::K + iy) * cpz + iz, cs);
        return p;
      };
      auto finish = [&](int sl, const AdjIn<R>& p) {
        const int n = t * NBL + sl / RL + (sl % RL) * SL;
        if (!G::kept(n)) return;
        const int ix = G::compact_of(n);
        const R kx = a.ks * R(G::kint(ix));
        const cpx<R> wv[3] = {w[0][sl], w[1][sl], w[2][sl]};
        const R k2 = kx * kx + ky * ky + kz * kz;
        cpx<R> an[3];
        nusum += adj_point_store(a, (size_t(ix) * G::K + iy) * cpz + iz, cs, k2, iz, wv, p, an);
        if (a.qout) {
          // q = P(qalpha a_next) / N^3, formed as k_inv_x<XM_ADJQ> forms it
          const R r = R(1) / ((k2 == R(0)) ? R(1) : k2);
          const cpx<R> kd = {(kx * an[0].x + ky * an[1].x + kz * an[2].x) * r,
                             (kx * an[0].y + ky * an[1].y + kz * an[2].y) * r};
          const R kc[3] = {kx, ky, kz};
#pragma unroll
          for (int c = 0; c < 3; ++c) qn[c][sl] = {qs * (an[c].x - kc[c] * kd.x), qs * (an[c].y - kc[c] * kd.y)};
        }
      };
      pipelined_points<E, AdjIn<R>, (E < 8)>(load, finish);
    }
    if (a.nupart) {
      __syncthreads();  // every thread is done reading buf
      const double s = block_sum(nusum, reinterpret_cast<double*>(buf));
      if (threadIdx.x == 0) a.nupart[size_t(blockIdx.y) * gridDim.x + blockIdx.x] += s;
    }
  }
};

// Register budget: three waves per SIMD for the FP64 hydro kernel and the
// FP64 MHD halves, three fields each (their few spills cost less than the
// lost occupancy, measured for hydro); the others keep the compiler's
// allocation (capping them spills more than it gains).
template <typename R, int KIND>
constexpr int xr_waves_per_eu() {
  return (KIND == XR_ADJ || KIND == XR_ADJ5) ? 2 : (sizeof(R) == 8 ? 3 : 1);
}

template <typename R, int N, bool MHD>
__global__ __launch_bounds__((CfgXR<R, N>::NT))
__attribute__((amdgpu_waves_per_eu(xr_waves_per_eu<R, MHD ? XR_MHD : XR_HYDRO>()))) void k_fwd_xr(const cpx<R>* __restrict__ in, StageArgs<R> a,
                                                             const cpx<R>* __restrict__ tw) {
  using B = XRBody<R, N, MHD ? XR_MHD : XR_HYDRO>;
  __shared__ cpx<R> sm[B::SMEM];
  B::run(sm, blockIdx.x, blockIdx.y, in, a, tw);
}

// Forward x of one stage VJP with its adjoint epilogue on register-held
// lines (grids of N >= 128; one 128-byte line of kz columns per block).
template <typename R, int N, bool FOLDED>
__global__ __launch_bounds__((CfgXR<R, N, true>::NT)) __attribute__((amdgpu_waves_per_eu(xr_waves_per_eu<R, XR_ADJ>())))
void k_fwd_xr_adj(const cpx<R>* __restrict__ in, StageArgs<R> a,
                                                                 const cpx<R>* __restrict__ tw) {
  using B = XRBody<R, N, FOLDED ? XR_ADJ5 : XR_ADJ>;
  __shared__ cpx<R> sm[B::SMEM];
  B::run(sm, blockIdx.x, blockIdx.y, in, a, tw);
}

// ----------------------------------------------------------------- adjoint helpers
// Cotangent of the final state for J = E (objective 0) or Z (1), as the
// unnormalised spectrum of the physical gradient: dE/du = u / N^3,
// dZ/du = curl curl u / N^3 = (k^2 u - k (k.u)) / N^3.  x and lam have row
// pitch cpz >= KZ (the padding columns of lam are zeroed).
template <typename R, int N>
__global__ __launch_bounds__(256) void k_obj_grad(const cpx<R>* __restrict__ x, cpx<R>* __restrict__ lam,
                                                  int objective, R ks, int cpz) {
  using G = Geo<N>;
  const size_t cs = size_t(G::K) * G::K * cpz;
  const size_t pt = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (pt >= cs) return;
  const int iz = int(pt % cpz), iy = int((pt / cpz) % G::K), ix = int(pt / (size_t(cpz) * G::K));
  if (iz >= G::KZ) {
#pragma unroll
    for (int c = 0; c < 3; ++c) lam[c * cs + pt] = {R(0), R(0)};
    return;
  }
  const R sc = R(1) / (R(N) * R(N) * R(N));
  cpx<R> v[3] = {x[pt], x[cs + pt], x[2 * cs + pt]};
  if (objective == 1) {
    const R kx = ks * R(G::kint(ix)), ky = ks * R(G::kint(iy)), kz = ks * R(iz);
    const R k2 = kx * kx + ky * ky + kz * kz;
    const cpx<R> kd = {kx * v[0].x + ky * v[1].x + kz * v[2].x, kx * v[0].y + ky * v[1].y + kz * v[2].y};
    const R kk[3] = {kx, ky, kz};
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = {k2 * v[c].x - kk[c] * kd.x, k2 * v[c].y - kk[c] * kd.y};
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) lam[c * cs + pt] = cscale(v[c], sc);
}

// Physical-cotangent spectrum (row pitch cpz) -> gradient w.r.t. the stored
// compact entries (dense pitch KZ; d/dRe + i d/dIm): weight (1 at kz = 0,
// 2 otherwise) / N^3.
template <typename R, int N>
__global__ __launch_bounds__(256) void k_adj_final(const cpx<R>* __restrict__ lam, cpx<R>* __restrict__ g, int cpz) {
  using G = Geo<N>;
  const size_t cs = size_t(G::K) * G::K * G::KZ, csl = size_t(G::K) * G::K * cpz;
  const size_t pt = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (pt >= cs) return;
  const size_t row = pt / G::KZ;
  const int iz = int(pt % G::KZ);
  const size_t src = row * cpz + iz;
  const R w = (iz == 0 ? R(1) : R(2)) / (R(N) * R(N) * R(N));
#pragma unroll
  for (int c = 0; c < 3; ++c) g[c * cs + pt] = cscale(lam[c * csl + src], w);
}

// ----------------------------------------------------------------- diagnostics
// Sums over the compact half spectrum with Hermitian weight (1 at kz=0, 2
// otherwise), accumulated in FP64 with Neumaier compensation.  Slots:
//  0 sum|u|^2  1 sum|k x u|^2  2 sum k^2|u|^2  3 sum|b|^2  4 sum Re(u.conj b)
//  5 sum|k.u|^2  6 sum|k.b|^2  7 sum|k x b|^2
//  8 number of stored real numbers (re and im parts of u, b) that are NaN
//    or infinite: the failure flag of a diverging run
constexpr int NDIAG = 9;
constexpr int DIAG_NT = 256;

struct Neu {
  double s = 0.0, c = 0.0;
  __device__ void add(double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  __device__ double val() const { return s + c; }
};

// The grid enters through M = N/3 only (K = 2M+1, KZ = M+1), so one
// instantiation per precision serves every grid size.
template <typename R>
__global__ __launch_bounds__(DIAG_NT) void k_diag(const cpx<R>* __restrict__ su,
                                                  const cpx<R>* __restrict__ sb,
                                                  double* __restrict__ partial, double ks, int M) {
  const int K = 2 * M + 1, KZ = M + 1;
  const size_t cs = size_t(K) * K * KZ;
  auto kint = [&](int i) { return i <= M ? i : i - K; };
  Neu acc[NDIAG];
  for (size_t pt = blockIdx.x * size_t(DIAG_NT) + threadIdx.x; pt < cs; pt += size_t(gridDim.x) * DIAG_NT) {
    const int iz = int(pt % KZ);
    const int iy = int((pt / KZ) % K);
    const int ix = int(pt / (size_t(KZ) * K));
    const double w = iz == 0 ? 1.0 : 2.0;
    const double kx = ks * kint(ix), ky = ks * kint(iy), kz = ks * iz;
    const double k2 = kx * kx + ky * ky + kz * kz;
    double ur[3], ui[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) { ur[c] = double(su[c * cs + pt].x); ui[c] = double(su[c * cs + pt].y); }
    int bad = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) bad += int(!isfinite(ur[c])) + int(!isfinite(ui[c]));
    const double uu = ur[0] * ur[0] + ui[0] * ui[0] + ur[1] * ur[1] + ui[1] * ui[1] + ur[2] * ur[2] + ui[2] * ui[2];
    auto curl2 = [&](const double* r, const double* i) {
      const double ar = ky * r[2] - kz * r[1], ai = ky * i[2] - kz * i[1];
      const double br = kz * r[0] - kx * r[2], bi = kz * i[0] - kx * i[2];
      const double cr = kx * r[1] - ky * r[0], ci = kx * i[1] - ky * i[0];
      return ar * ar + ai * ai + br * br + bi * bi + cr * cr + ci * ci;
    };
    auto div2 = [&](const double* r, const double* i) {
      const double dr = kx * r[0] + ky * r[1] + kz * r[2], di = kx * i[0] + ky * i[1] + kz * i[2];
      return dr * dr + di * di;
    };
    acc[0].add(w * uu);
    acc[1].add(w * curl2(ur, ui));
    acc[2].add(w * k2 * uu);
    acc[5].add(w * div2(ur, ui));
    if (sb) {
      double br[3], bi[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) { br[c] = double(sb[c * cs + pt].x); bi[c] = double(sb[c * cs + pt].y); }
      acc[3].add(w * (br[0] * br[0] + bi[0] * bi[0] + br[1] * br[1] + bi[1] * bi[1] + br[2] * br[2] + bi[2] * bi[2]));
      acc[4].add(w * (ur[0] * br[0] + ui[0] * bi[0] + ur[1] * br[1] + ui[1] * bi[1] + ur[2] * br[2] + ui[2] * bi[2]));
      acc[6].add(w * div2(br, bi));
      acc[7].add(w * curl2(br, bi));
#pragma unroll
      for (int c = 0; c < 3; ++c) bad += int(!isfinite(br[c])) + int(!isfinite(bi[c]));
    }
    if (bad) acc[8].add(double(bad));
  }
  __shared__ double red[DIAG_NT];
  for (int q = 0; q < NDIAG; ++q) {
    red[threadIdx.x] = acc[q].val();
    __syncthreads();
    if (threadIdx.x == 0) {
      Neu s;
      for (int i = 0; i < DIAG_NT; ++i) s.add(red[i]);
      partial[size_t(blockIdx.x) * NDIAG + q] = s.val();
    }
    __syncthreads();
  }
}

// Final reduction: out[0..NDIAG) = compensated sums of the partials,
// out[NDIAG] = max of `nmax` maxima (u divergence), out[NDIAG+1] = (b).
template <typename R>
__global__ __launch_bounds__(64) void k_diag_final(const double* __restrict__ partial, int nblk,
                                                   const R* __restrict__ maxu, const R* __restrict__ maxb,
                                                   int nmax, double* __restrict__ out) {
  const int q = threadIdx.x;
  if (q < NDIAG) {
    Neu s;
    for (int b = 0; b < nblk; ++b) s.add(partial[size_t(b) * NDIAG + q]);
    out[q] = s.val();
  } else if (q == NDIAG || q == NDIAG + 1) {
    const R* m = (q == NDIAG) ? maxu : maxb;
    double mm = 0.0;
    if (m)
      for (int i = 0; i < nmax; ++i) mm = fmax(mm, double(m[i]));
    out[q] = mm;
  }
}

// Normalised diagnostics on the device (the tg_diagnostics out[] layout):
// sums / N^6 (with 1/2 for the energies and the enstrophy), the divergence
// maxima and the non-finite count.
template <int UNUSED = 0>
__global__ __launch_bounds__(64) void k_diag_normalize(const double* __restrict__ sums, int n,
                                                       double* __restrict__ out) {
  const int q = threadIdx.x;
  const double n3 = double(n) * n * n;
  const double n6 = n3 * n3;
  if (q == 0 || q == 1 || q == 3) out[q] = 0.5 * sums[q] / n6;
  else if (q == 2 || (q >= 4 && q <= 7)) out[q] = sums[q] / n6;
  else if (q == 8) out[8] = sums[NDIAG];
  else if (q == 9) out[9] = sums[NDIAG + 1];
  else if (q == 10) out[10] = sums[8];
}

}  // namespace tg
