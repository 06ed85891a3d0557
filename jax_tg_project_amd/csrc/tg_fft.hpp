// Stockham FFT engine for gfx950, in two forms: batches of lines held in LDS
// and shared by the wavefronts of a block (the x and y passes), and lines
// held in the registers of the lanes of one wavefront with a wavefront-
// private LDS exchange buffer (the z passes, "register" section below).  One
// radix-R butterfly group per thread per stage, twiddles from a host-built
// table.
//
// Layout convention: element n of line l lives at buf[n * LP + l].  A thread
// owns line l = tid % L and slice t = tid / L of that line; in every stage it
// loads E/R butterflies of R points at stride N/R, applies twiddles, does the
// radix-R DFT in registers and writes the results at stride P (the product of
// the radices already applied).  After the last stage the line is in natural
// order (autosort), so no bit-reversal pass exists.
//
// DIR = -1 is the forward transform  X[k] = sum_n x[n] exp(-2 pi i n k / N),
// DIR = +1 the unnormalised inverse  x[n] = sum_k X[k] exp(+2 pi i n k / N).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace tg {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> cadd(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> csub(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) { return {a.x * s, a.y * s}; }
// multiply by i
template <typename R>
__device__ __forceinline__ cpx<R> cmuli(cpx<R> a) { return {-a.y, a.x}; }

// Global loads and stores with an optional nontemporal hint (NT), for
// arrays that a pass streams through once and that are too large to stay in
// the device's last-level cache until the next pass reads them (measured on
// gfx950: a plain device copy moves 4.7 TB/s, 6.3 TB/s with both hints).
template <typename R>
using vec2_t = R __attribute__((ext_vector_type(2)));
template <bool NT, typename R>
__device__ __forceinline__ cpx<R> gld(const cpx<R>* p) {
  if constexpr (NT) {
    const vec2_t<R> v = __builtin_nontemporal_load(reinterpret_cast<const vec2_t<R>*>(p));
    return {v.x, v.y};
  } else {
    return *p;
  }
}
template <bool NT, typename R>
__device__ __forceinline__ void gst(cpx<R>* p, cpx<R> c) {
  if constexpr (NT) {
    vec2_t<R> v;
    v.x = c.x;
    v.y = c.y;
    __builtin_nontemporal_store(v, reinterpret_cast<vec2_t<R>*>(p));
  } else {
    *p = c;
  }
}

// multiply by exp(DIR * 2 pi i / 4) = DIR * i
template <int DIR, typename R>
__device__ __forceinline__ cpx<R> rot4(cpx<R> a) {
  if constexpr (DIR < 0) return {a.y, -a.x};
  else return {-a.y, a.x};
}

// ---------------------------------------------------------------- radix DFTs
template <int DIR, typename R>
__device__ __forceinline__ void dft2(cpx<R>* v) {
  cpx<R> a = v[0];
  v[0] = cadd(a, v[1]);
  v[1] = csub(a, v[1]);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft3(cpx<R>* v) {
  // exp(DIR 2 pi i / 3) = -1/2 + DIR i sqrt(3)/2
  const R h = R(0.86602540378443864676372317075294);  // sqrt(3)/2
  cpx<R> s = cadd(v[1], v[2]);
  cpx<R> d = csub(v[1], v[2]);
  cpx<R> m = {v[0].x - R(0.5) * s.x, v[0].y - R(0.5) * s.y};
  cpx<R> r = {-R(DIR) * h * d.y, R(DIR) * h * d.x};  // DIR i h d
  v[0] = cadd(v[0], s);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft4s(cpx<R>& x0, cpx<R>& x1, cpx<R>& x2, cpx<R>& x3) {
  cpx<R> t0 = cadd(x0, x2), t1 = csub(x0, x2);
  cpx<R> t2 = cadd(x1, x3), t3 = rot4<DIR>(csub(x1, x3));
  x0 = cadd(t0, t2);
  x2 = csub(t0, t2);
  x1 = cadd(t1, t3);
  x3 = csub(t1, t3);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft4(cpx<R>* v) { dft4s<DIR>(v[0], v[1], v[2], v[3]); }

template <int DIR, typename R>
__device__ __forceinline__ cpx<R> mul_w8(cpx<R> a) {  // * exp(DIR 2 pi i/8)
  const R c = R(0.70710678118654752440084436210485);
  return {c * (a.x - R(DIR) * a.y), c * (a.y + R(DIR) * a.x)};
}
template <int DIR, typename R>
__device__ __forceinline__ cpx<R> mul_w8_3(cpx<R> a) {  // * exp(DIR 3 * 2 pi i/8)
  const R c = R(0.70710678118654752440084436210485);
  return {c * (-a.x - R(DIR) * a.y), c * (R(DIR) * a.x - a.y)};
}

template <int DIR, typename R>
__device__ __forceinline__ void dft8(cpx<R>* v) {
  cpx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cpx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4s<DIR>(e0, e1, e2, e3);
  dft4s<DIR>(o0, o1, o2, o3);
  o1 = mul_w8<DIR>(o1);
  o2 = rot4<DIR>(o2);
  o3 = mul_w8_3<DIR>(o3);
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft16(cpx<R>* v) {
  // n = 4 n1 + n2, k = k1 + 4 k2: inner DFT4 over n1, twiddle w16^(n2 k1),
  // outer DFT4 over n2.
  const R c1 = R(0.92387953251128675612818318939679);  // cos(pi/8)
  const R s1 = R(0.38268343236508977172845998403040);  // sin(pi/8)
  for (int n2 = 0; n2 < 4; ++n2) dft4s<DIR>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
  // now v[4*k1 + n2] holds A[n2][k1]
  // w16^1 = (c1, DIR s1), w16^2 = w8, w16^3 = (s1, DIR c1), w16^4 = DIR i,
  // w16^6 = w8^3, w16^9 = (-c1, -DIR s1)
  auto w1 = [&](cpx<R> a) -> cpx<R> {
    return {a.x * c1 - a.y * R(DIR) * s1, a.x * R(DIR) * s1 + a.y * c1};
  };
  auto w3 = [&](cpx<R> a) -> cpx<R> {
    return {a.x * s1 - a.y * R(DIR) * c1, a.x * R(DIR) * c1 + a.y * s1};
  };
  auto w9 = [&](cpx<R> a) -> cpx<R> {
    return {-a.x * c1 + a.y * R(DIR) * s1, -a.x * R(DIR) * s1 - a.y * c1};
  };
  // k1 = 1: n2 = 1,2,3 -> w^1, w^2, w^3
  v[5] = w1(v[5]); v[6] = mul_w8<DIR>(v[6]); v[7] = w3(v[7]);
  // k1 = 2: w^2, w^4, w^6
  v[9] = mul_w8<DIR>(v[9]); v[10] = rot4<DIR>(v[10]); v[11] = mul_w8_3<DIR>(v[11]);
  // k1 = 3: w^3, w^6, w^9
  v[13] = w3(v[13]); v[14] = mul_w8_3<DIR>(v[14]); v[15] = w9(v[15]);
  cpx<R> o[16];
  for (int k1 = 0; k1 < 4; ++k1) {
    cpx<R> a0 = v[4 * k1 + 0], a1 = v[4 * k1 + 1], a2 = v[4 * k1 + 2], a3 = v[4 * k1 + 3];
    dft4s<DIR>(a0, a1, a2, a3);
    o[k1 + 0] = a0; o[k1 + 4] = a1; o[k1 + 8] = a2; o[k1 + 12] = a3;
  }
  for (int i = 0; i < 16; ++i) v[i] = o[i];
}

template <int DIR, typename R>
__device__ __forceinline__ void dft12(cpx<R>* v) {
  // n = 4 n1 + n2, k = k1 + 3 k2: DFT3 over n1, twiddle w12^(n2 k1), DFT4
  // over n2.
  const R h = R(0.86602540378443864676372317075294);  // sqrt(3)/2
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    cpx<R> t[3] = {v[n2], v[4 + n2], v[8 + n2]};
    dft3<DIR>(t);
    v[n2] = t[0]; v[4 + n2] = t[1]; v[8 + n2] = t[2];
  }
  // now v[4*k1 + n2] holds A[n2][k1]; w12^m = (cos, DIR sin) of m*30 degrees
  auto rot = [&](cpx<R> a, R c, R sn) -> cpx<R> {
    return {a.x * c - a.y * R(DIR) * sn, a.x * R(DIR) * sn + a.y * c};
  };
  v[5] = rot(v[5], h, R(0.5));        // w^1
  v[6] = rot(v[6], R(0.5), h);        // w^2
  v[7] = rot4<DIR>(v[7]);             // w^3
  v[9] = rot(v[9], R(0.5), h);        // w^2
  v[10] = rot(v[10], R(-0.5), h);     // w^4
  v[11] = {-v[11].x, -v[11].y};       // w^6
  cpx<R> o[12];
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    cpx<R> a0 = v[4 * k1 + 0], a1 = v[4 * k1 + 1], a2 = v[4 * k1 + 2], a3 = v[4 * k1 + 3];
    dft4s<DIR>(a0, a1, a2, a3);
    o[k1 + 0] = a0; o[k1 + 3] = a1; o[k1 + 6] = a2; o[k1 + 9] = a3;
  }
#pragma unroll
  for (int i = 0; i < 12; ++i) v[i] = o[i];
}

template <int RAD, int DIR, typename R>
__device__ __forceinline__ void dft(cpx<R>* v) {
  if constexpr (RAD == 2) dft2<DIR>(v);
  else if constexpr (RAD == 3) dft3<DIR>(v);
  else if constexpr (RAD == 4) dft4<DIR>(v);
  else if constexpr (RAD == 8) dft8<DIR>(v);
  else if constexpr (RAD == 12) dft12<DIR>(v);
  else if constexpr (RAD == 16) dft16<DIR>(v);
  else static_assert(RAD == 2, "unsupported radix");
}

// ---------------------------------------------------------------- plans
// Plan<N, E>::run applies the stages for a transform of length N where every
// thread holds E points; each radix divides E and the radices multiply to N.
template <int... Rs>
struct Radices {};

// The default plan is greedy: at each stage the largest radix of
// {16, 12, 8, 4, 3, 2} that divides both E and what is left of N.  A size
// with no such factorisation (a prime factor above 3, or a factor 3 that E
// lacks) fails the static_assert.
struct RadixList {
  int r[24];
  int n;
};
__host__ __device__ constexpr RadixList greedy_radices(int N, int E) {
  RadixList a{};
  const int cand[6] = {16, 12, 8, 4, 3, 2};
  int rem = N;
  while (rem > 1) {
    int pick = 0;
    for (int c : cand)
      if (E % c == 0 && rem % c == 0) {
        pick = c;
        break;
      }
    if (pick == 0 || a.n == 24) {
      a.n = -1;
      return a;
    }
    a.r[a.n++] = pick;
    rem /= pick;
  }
  return a;
}

template <int N, int E>
struct GreedyPlan {
  static constexpr RadixList A = greedy_radices(N, E);
  static_assert(A.n > 0, "no radix plan: N must factor into radices of E drawn from {16, 12, 8, 4, 3, 2}");
  template <size_t... I>
  static Radices<A.r[I]...> conv(std::index_sequence<I...>);
  using type = decltype(conv(std::make_index_sequence<size_t(A.n > 0 ? A.n : 0)>{}));
};

template <int N, int E>
struct PlanFor {
  using type = typename GreedyPlan<N, E>::type;
};
// The 24-point plans of N = 48 and 96 (their LDS z pass) keep the radix
// order they were tuned with instead of the greedy 12 x 4 / 12 x 8.
template <> struct PlanFor<48, 24> { using type = Radices<8, 3, 2>; };
template <> struct PlanFor<96, 24> { using type = Radices<8, 4, 3>; };

// ---------------------------------------------------------------- global I/O
// Stage with optional direct I/O.  SRC: 0 = LDS, 1 = functor (no LDS read, so
// no barrier is needed before this stage writes LDS), 2 = functor that reads
// LDS itself (barrier before writing).  DST_GLOBAL: results go to the store
// functor instead of LDS (then no trailing barrier).
// LDS index of element n of line l with one pad element every 2^PS rows
// (PS = 0: no padding).  The pad keeps a radix-R first stage that writes at
// stride R*LP from landing every lane of a wave on the same bank.
template <int PS>
__device__ __forceinline__ int lidx(int n, int l, int LP) {
  if constexpr (PS > 0) return n * LP + (n >> PS) + l;
  else return n * LP + l;
}

template <typename R, int N, int E, int DIR, int P, int RAD, int SRC, bool DST_GLOBAL, int PS, class Ld, class St>
__device__ __forceinline__ void io_stage(cpx<R>* buf, int LP, int l, int t, bool active,
                                         const cpx<R>* __restrict__ tw, const Ld& ld, const St& st) {
  static_assert(E % RAD == 0, "radix must divide points per thread");
  constexpr int NB = E / RAD;
  constexpr int SIN = N / RAD;
  cpx<R> v[NB][RAD];
  if (active) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int b = t * NB + j;
#pragma unroll
      for (int m = 0; m < RAD; ++m) {
        const int n = b + m * SIN;
        if constexpr (SRC == 0) v[j][m] = buf[lidx<PS>(n, l, LP)];
        else v[j][m] = ld(n, l, j * RAD + m);  // slot j*RAD+m is a compile-time constant
      }
    }
  }
  if constexpr (SRC != 1 && !DST_GLOBAL) __syncthreads();
  if (active) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int b = t * NB + j;
      const int k = b % P;
      if constexpr (P > 1) {
#pragma unroll
        for (int m = 1; m < RAD; ++m) {
          cpx<R> w = tw[m * k * (N / (P * RAD))];
          if constexpr (DIR > 0) w.y = -w.y;
          v[j][m] = cmul(v[j][m], w);
        }
      }
      dft<RAD, DIR>(v[j]);
      const int base = (b - k) * RAD + k;
#pragma unroll
      for (int m = 0; m < RAD; ++m) {
        if constexpr (DST_GLOBAL) st(base + m * P, l, v[j][m], j * RAD + m);  // slot: compile-time
        else buf[lidx<PS>(base + m * P, l, LP)] = v[j][m];
      }
    }
  }
  if constexpr (!DST_GLOBAL) __syncthreads();
}

template <typename R, int N, int E, int DIR, int SRC, bool DST_GLOBAL, int PS, int P, class Ld, class St,
          int R0, int... Rest>
__device__ __forceinline__ void io_stages(cpx<R>* buf, int LP, int l, int t, bool active,
                                          const cpx<R>* __restrict__ tw, const Ld& ld, const St& st) {
  constexpr bool last = sizeof...(Rest) == 0;
  io_stage<R, N, E, DIR, P, R0, SRC, last && DST_GLOBAL, PS>(buf, LP, l, t, active, tw, ld, st);
  if constexpr (!last)
    io_stages<R, N, E, DIR, 0, DST_GLOBAL, PS, P * R0, Ld, St, Rest...>(buf, LP, l, t, active, tw, ld, st);
}

template <typename R, int N, int E, int DIR, int SRC, bool DST_GLOBAL, int PS, class Ld, class St, int... Rs>
__device__ __forceinline__ void io_apply(cpx<R>* buf, int LP, int l, int t, bool active,
                                         const cpx<R>* __restrict__ tw, const Ld& ld, const St& st,
                                         Radices<Rs...>) {
  io_stages<R, N, E, DIR, SRC, DST_GLOBAL, PS, 1, Ld, St, Rs...>(buf, LP, l, t, active, tw, ld, st);
}

template <class Plan>
struct FirstRadix;
template <int R0, int... Rest>
struct FirstRadix<Radices<R0, Rest...>> {
  static constexpr int value = R0;
};
template <class Plan>
struct LastRadix;
template <int R0>
struct LastRadix<Radices<R0>> {
  static constexpr int value = R0;
};
template <int R0, int R1, int... Rest>
struct LastRadix<Radices<R0, R1, Rest...>> : LastRadix<Radices<R1, Rest...>> {};
__host__ __device__ constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }
// Pad shift for a slice-fastest transform of length N with E points per thread.
template <int N, int E>
constexpr int pad_shift() { return ilog2(FirstRadix<typename PlanFor<N, E>::type>::value); }

// Transform `nlines` lines whose first stage reads through `ld(n, l, slot)`
// (slot = the thread's register slot of that point, 0..E-1, fixed for a
// given thread and n) and
// (if DST_GLOBAL) whose last stage writes through `st(n, l, v, slot)` in
// natural order; intermediate stages exchange through LDS (buf[n*LP + l]).  The
// caller guarantees the LDS buffer is free on entry (barrier after any
// earlier use).  With DST_GLOBAL the call ends without a barrier.
struct NoIO {
  template <typename... A>
  __device__ void operator()(A...) const {}
};

// Copy the N-entry twiddle table to LDS.  No barrier here: a first stage
// never reads twiddles (P = 1), and the barrier that ends it publishes the
// table to every later stage.
template <typename R, int N>
__device__ __forceinline__ void stage_twiddles(cpx<R>* tws, const cpx<R>* __restrict__ tw) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) tws[i] = tw[i];
}

// TFAST selects the thread mapping: false = consecutive threads take
// consecutive lines (coalesced when lines are adjacent columns in memory),
// true = consecutive threads take consecutive slices of one line (coalesced
// when a line is contiguous in memory).
template <typename R, int N, int E, int DIR, int SRC, bool DST_GLOBAL, bool TFAST = false, int PS = 0, class Ld,
          class St>
__device__ __forceinline__ void fft_io(cpx<R>* buf, int LP, int nlines, const cpx<R>* __restrict__ tw,
                                       const Ld& ld, const St& st) {
  const int tid = threadIdx.x;
  constexpr int TPL = N / E;
  const int l = TFAST ? tid / TPL : tid % nlines;
  const int t = TFAST ? tid % TPL : tid / nlines;
  const bool active = TFAST ? (l < nlines) : (t < TPL);
  io_apply<R, N, E, DIR, SRC, DST_GLOBAL, PS>(buf, LP, l, t, active, tw, ld, st, typename PlanFor<N, E>::type{});
}

// The same with an explicit radix plan (e.g. the reverse of PlanFor, whose
// first stage reads the points in the order the forward plan's last stage
// leaves them in registers).
template <typename R, int N, int E, int DIR, int SRC, bool DST_GLOBAL, class Plan, bool TFAST = false, int PS = 0,
          class Ld, class St>
__device__ __forceinline__ void fft_io_plan(cpx<R>* buf, int LP, int nlines, const cpx<R>* __restrict__ tw,
                                            const Ld& ld, const St& st) {
  const int tid = threadIdx.x;
  constexpr int TPL = N / E;
  const int l = TFAST ? tid / TPL : tid % nlines;
  const int t = TFAST ? tid % TPL : tid / nlines;
  const bool active = TFAST ? (l < nlines) : (t < TPL);
  io_apply<R, N, E, DIR, SRC, DST_GLOBAL, PS>(buf, LP, l, t, active, tw, ld, st, Plan{});
}

// ---------------------------------------------------------------- registers
// Register-resident Stockham transform for lines owned by lanes of a single
// wavefront.  The TPL = N / E lanes of a line keep its points in registers;
// between stages they trade them through a line buffer in LDS that only
// their own wavefront touches, so a wavefront-scope fence orders the
// exchange (LDS instructions of one wavefront execute in issue order) and no
// workgroup barrier is needed.
//
// Register order: before a stage of radix RAD, slot j*RAD + m of lane t
// holds point t*(E/RAD) + j + m*(N/RAD).  After the last stage (radix RL)
// slot j*RL + m holds output t*(E/RL) + j + m*(N/RL): exactly the input
// order of a first stage of radix RL, so a plan and its reverse chain in
// registers without an exchange.
template <class Done, class Todo>
struct RevImpl;
template <int... Done>
struct RevImpl<Radices<Done...>, Radices<>> {
  using type = Radices<Done...>;
};
template <int... Done, int H, int... T>
struct RevImpl<Radices<Done...>, Radices<H, T...>> {
  using type = typename RevImpl<Radices<H, Done...>, Radices<T...>>::type;
};
template <class Plan>
using ReversePlan = typename RevImpl<Radices<>, Plan>::type;

// Value held by lane `src` of the same wavefront (every lane must execute
// the call).
__device__ __forceinline__ float lane_get(float v, int src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src << 2, __builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double lane_get(double v, int src) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = unsigned(__builtin_amdgcn_ds_bpermute(src << 2, int(unsigned(b))));
  const unsigned hi = unsigned(__builtin_amdgcn_ds_bpermute(src << 2, int(unsigned(b >> 32))));
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
template <typename R>
__device__ __forceinline__ cpx<R> lane_get(cpx<R> v, int src) {
  return {lane_get(v.x, src), lane_get(v.y, src)};
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Line buffer position of point n: one pad slot every 2^PS points keeps a
// stage that writes E consecutive points per lane off a single bank.
template <int PS>
__device__ __forceinline__ int pidx(int n) { return n + (n >> PS); }

// Per-stage twiddle rows in LDS: a stage (P, RAD) with P > 1 owns
// (RAD - 1) * P entries, row m - 1 holding w^(m k N / (P RAD)) for k in
// [0, P), so the lanes of a stage (consecutive k) read consecutive entries.
// The rows of the stages follow each other in plan order.
template <int N, int P, class Plan>
struct StageTw;
template <int N, int P>
struct StageTw<N, P, Radices<>> {
  static constexpr int SIZE = 0;
  template <typename R>
  __device__ __forceinline__ static void fill(cpx<R>*, const cpx<R>* __restrict__, int, int) {}
};
template <int N, int P, int R0, int... Rest>
struct StageTw<N, P, Radices<R0, Rest...>> {
  static constexpr int OWN = P > 1 ? (R0 - 1) * P : 0;
  using Next = StageTw<N, P * R0, Radices<Rest...>>;
  static constexpr int SIZE = OWN + Next::SIZE;
  template <typename R>
  __device__ __forceinline__ static void fill(cpx<R>* ts, const cpx<R>* __restrict__ tw, int tid, int nt) {
    for (int i = tid; i < OWN; i += nt) {
      const int m = i / P + 1, k = i % P;
      ts[i] = tw[m * k * (N / (P * R0))];
    }
    Next::fill(ts + OWN, tw, tid, nt);
  }
};
template <int N, class Plan>
using PlanTw = StageTw<N, 1, Plan>;

template <typename R, int N, int E, int DIR, int P, int RAD>
__device__ __forceinline__ void rs_butterflies(cpx<R>* v, int t, const cpx<R>* ts) {
  constexpr int NB = E / RAD;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if constexpr (P > 1) {
      const int k = (t * NB + j) % P;
#pragma unroll
      for (int m = 1; m < RAD; ++m) {
        cpx<R> w = ts[(m - 1) * P + k];
        if constexpr (DIR > 0) w.y = -w.y;
        v[j * RAD + m] = cmul(v[j * RAD + m], w);
      }
    }
    dft<RAD, DIR>(&v[j * RAD]);
  }
}

template <typename R, int N, int E, int P, int RAD, int PS>
__device__ __forceinline__ void rs_put(const cpx<R>* v, int t, cpx<R>* lb) {
  constexpr int NB = E / RAD;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int b = t * NB + j, k = b % P, base = (b - k) * RAD + k;
#pragma unroll
    for (int m = 0; m < RAD; ++m) lb[pidx<PS>(base + m * P)] = v[j * RAD + m];
  }
}

template <typename R, int N, int E, int RAD, int PS>
__device__ __forceinline__ void rs_get(cpx<R>* v, int t, const cpx<R>* lb) {
  constexpr int NB = E / RAD;
#pragma unroll
  for (int j = 0; j < NB; ++j)
#pragma unroll
    for (int m = 0; m < RAD; ++m) v[j * RAD + m] = lb[pidx<PS>(t * NB + j + m * (N / RAD))];
}

// ts: the twiddle rows of this plan (PlanTw layout).
template <typename R, int N, int E, int DIR, int PS, int P, int R0, int... Rest>
__device__ __forceinline__ void rs_stages(cpx<R>* v, int t, cpx<R>* lb, const cpx<R>* ts) {
  rs_butterflies<R, N, E, DIR, P, R0>(v, t, ts);
  if constexpr (sizeof...(Rest) > 0) {
    constexpr int R1 = FirstRadix<Radices<Rest...>>::value;
    rs_put<R, N, E, P, R0, PS>(v, t, lb);
    wave_lds_sync();
    rs_get<R, N, E, R1, PS>(v, t, lb);
    wave_lds_sync();
    rs_stages<R, N, E, DIR, PS, P * R0, Rest...>(v, t, lb, ts + (P > 1 ? (R0 - 1) * P : 0));
  }
}

template <typename R, int N, int E, int DIR, int PS, int... Rs>
__device__ __forceinline__ void reg_fft(cpx<R>* v, int t, cpx<R>* lb, const cpx<R>* ts, Radices<Rs...>) {
  rs_stages<R, N, E, DIR, PS, 1, Rs...>(v, t, lb, ts);
}

}  // namespace tg
