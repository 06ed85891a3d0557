// Grid sizes without compiled kernels: every N >= 4 runs here, on direct
// (O(N) per output) discrete Fourier sums along each axis with the grid size
// a run-time value, so the radix plans and launch shapes of the compiled
// sizes are not needed.  The state keeps the compact layout of every plan
// (S[c][ix][iy][kz], K = 2M+1, KZ = M+1, M = N/3), so checkpoints,
// diagnostics and the host API are the same; the hot passes are not fused.
//
// The nonlinear terms follow the reference's formulation exactly, also
// where 3 divides N and the inclusive 2/3 mask keeps an aliased shell:
//   hydro  rhs = -P(F[u . grad u] mask) - nu k^2 u          (jax/taylor_green_jax.py:158-186)
//   MHD    rhs_u = -P(F[u . grad u]) + s P(F[j x b]) - nu k^2 u,
//          rhs_b = P(i k x F[u x b]) - eta k^2 b             (cuda/taylor_green_mhd/operators.py:57-107)
// with every derivative taken of the masked spectrum; the RK4 update is
// the classical four-stage scheme (jax/taylor_green_jax.py:189-196).
//
// Per transform of one field (K ~ 2N/3, KZ ~ N/3): x pass N K KZ outputs of
// K terms, y pass N^2 KZ outputs of K terms, z pass N^3 outputs of KZ terms;
// about 0.7 N^4 complex multiply-adds, so this path is for the sizes the
// compiled list lacks, not for speed.
#include "tg_ops.hpp"

namespace tg {
namespace gen {

constexpr int NT = 256;

__host__ __device__ inline int kint(int i, int M) { return i <= M ? i : i - (2 * M + 1); }

// Phase index of exp(sgn 2 pi i k n / N) in the twiddle table
// tw[j] = exp(-2 pi i j / N): j = (-sgn k n) mod N.
__device__ inline int phase(long long k, long long n, int N, int sgn) {
  long long j = (-sgn * k * n) % N;
  return int(j < 0 ? j + N : j);
}

// Sum over the kept compact indices i (signed k = kint(i)) of
// v(i) exp(+2 pi i k n / N), with the phase stepped by n per term.
template <typename R, class F>
__device__ inline cpx<R> kept_sum_inv(F v, int n, int N, int M, const cpx<R>* __restrict__ tw) {
  const int K = 2 * M + 1;
  const int step = (N - n % N) % N;  // index step of exp(+2 pi i n / N)
  cpx<R> acc = {R(0), R(0)};
  int j = 0;
  for (int i = 0; i <= M; ++i) {
    const cpx<R> w = tw[j];
    const cpx<R> x = v(i);
    acc.x += x.x * w.x - x.y * w.y;
    acc.y += x.x * w.y + x.y * w.x;
    j += step;
    if (j >= N) j -= N;
  }
  j = phase(-M, n, N, +1);
  for (int i = M + 1; i < K; ++i) {
    const cpx<R> w = tw[j];
    const cpx<R> x = v(i);
    acc.x += x.x * w.x - x.y * w.y;
    acc.y += x.x * w.y + x.y * w.x;
    j += step;
    if (j >= N) j -= N;
  }
  return acc;
}

// Sum over n in [0, N) of v(n) exp(-2 pi i k n / N).
template <typename R, class F>
__device__ inline cpx<R> full_sum_fwd(F v, int k, int N, const cpx<R>* __restrict__ tw) {
  const int step = ((k % N) + N) % N;
  cpx<R> acc = {R(0), R(0)};
  int j = 0;
  for (int n = 0; n < N; ++n) {
    const cpx<R> w = tw[j];
    const cpx<R> x = v(n);
    acc.x += x.x * w.x - x.y * w.y;
    acc.y += x.x * w.y + x.y * w.x;
    j += step;
    if (j >= N) j -= N;
  }
  return acc;
}

// Spectral operand of an inverse transform, formed from the compact state S
// (component stride cs) at (ix, iy, iz):
//   OP_FIELD  S_c            OP_DERIV  i k_d S_c
//   OP_CURL   i (k x S)_c    OP_DIV    i k . S
enum { OP_FIELD = 0, OP_DERIV = 1, OP_CURL = 2, OP_DIV = 3 };

template <typename R>
__device__ inline cpx<R> operand(const cpx<R>* __restrict__ S, size_t cs, size_t pt, int op, int c, int d, R kx,
                                 R ky, R kz) {
  auto imul = [](R k, cpx<R> v) { return cpx<R>{-k * v.y, k * v.x}; };
  const R kd[3] = {kx, ky, kz};
  if (op == OP_FIELD) return S[c * cs + pt];
  if (op == OP_DERIV) return imul(kd[d], S[c * cs + pt]);
  if (op == OP_DIV) {
    const cpx<R> a = S[pt], b = S[cs + pt], e = S[2 * cs + pt];
    return imul(R(1), cpx<R>{kx * a.x + ky * b.x + kz * e.x, kx * a.y + ky * b.y + kz * e.y});
  }
  const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
  const cpx<R> s1 = S[c1 * cs + pt], s2 = S[c2 * cs + pt];
  return imul(R(1), cpx<R>{kd[c1] * s2.x - kd[c2] * s1.x, kd[c1] * s2.y - kd[c2] * s1.y});
}

// x pass of the inverse: A[x][iy][kz] = sum_ix op(S)[ix][iy][kz] e^{+i kx x}.
template <typename R>
__global__ __launch_bounds__(NT) void g_inv_x(const cpx<R>* __restrict__ S, int op, int c, int d,
                                              cpx<R>* __restrict__ A, const cpx<R>* __restrict__ tw, int N, int M,
                                              R ks) {
  const int K = 2 * M + 1, KZ = M + 1;
  const size_t total = size_t(N) * K * KZ, cs = size_t(K) * K * KZ;
  const size_t o = size_t(blockIdx.x) * NT + threadIdx.x;
  if (o >= total) return;
  const int iz = int(o % KZ), iy = int((o / KZ) % K), x = int(o / (size_t(KZ) * K));
  const R ky = ks * R(kint(iy, M)), kz = ks * R(iz);
  auto v = [&](int ix) {
    const size_t pt = (size_t(ix) * K + iy) * KZ + iz;
    return operand<R>(S, cs, pt, op, c, d, ks * R(kint(ix, M)), ky, kz);
  };
  A[o] = kept_sum_inv<R>(v, x, N, M, tw);
}

// y pass of the inverse: B[x][y][kz] = sum_iy A[x][iy][kz] e^{+i ky y}.
template <typename R>
__global__ __launch_bounds__(NT) void g_inv_y(const cpx<R>* __restrict__ A, cpx<R>* __restrict__ B,
                                              const cpx<R>* __restrict__ tw, int N, int M) {
  const int K = 2 * M + 1, KZ = M + 1;
  const size_t total = size_t(N) * N * KZ;
  const size_t o = size_t(blockIdx.x) * NT + threadIdx.x;
  if (o >= total) return;
  const int iz = int(o % KZ), y = int((o / KZ) % N), x = int(o / (size_t(KZ) * N));
  const cpx<R>* row = A + size_t(x) * K * KZ + iz;
  B[o] = kept_sum_inv<R>([&](int iy) { return row[size_t(iy) * KZ]; }, y, N, M, tw);
}

// z pass of the inverse (Hermitian in kz): P[x][y][z] = scale (Re B0 +
// 2 Re sum_{kz >= 1} B[kz] e^{+i kz z}).
template <typename R>
__global__ __launch_bounds__(NT) void g_inv_z(const cpx<R>* __restrict__ B, R* __restrict__ P,
                                              const cpx<R>* __restrict__ tw, int N, int M, R scale) {
  const int KZ = M + 1;
  const size_t total = size_t(N) * N * N;
  const size_t o = size_t(blockIdx.x) * NT + threadIdx.x;
  if (o >= total) return;
  const int z = int(o % N);
  const cpx<R>* row = B + (o / N) * KZ;
  const int step = (N - z % N) % N;
  R acc = R(0);
  int j = step;
  for (int kz = 1; kz <= M; ++kz) {
    const cpx<R> w = tw[j], b = row[kz];
    acc += b.x * w.x - b.y * w.y;
    j += step;
    if (j >= N) j -= N;
  }
  P[o] = scale * (row[0].x + R(2) * acc);
}

// z pass of the forward: B[x][y][kz] = sum_z P[x][y][z] e^{-i kz z}.
template <typename R>
__global__ __launch_bounds__(NT) void g_fwd_z(const R* __restrict__ P, cpx<R>* __restrict__ B,
                                              const cpx<R>* __restrict__ tw, int N, int M) {
  const int KZ = M + 1;
  const size_t total = size_t(N) * N * KZ;
  const size_t o = size_t(blockIdx.x) * NT + threadIdx.x;
  if (o >= total) return;
  const int kz = int(o % KZ);
  const R* line = P + (o / KZ) * N;
  B[o] = full_sum_fwd<R>([&](int z) { return cpx<R>{line[z], R(0)}; }, kz, N, tw);
}

// y pass of the forward: A[x][iy][kz] = sum_y B[x][y][kz] e^{-i ky y}.
template <typename R>
__global__ __launch_bounds__(NT) void g_fwd_y(const cpx<R>* __restrict__ B, cpx<R>* __restrict__ A,
                                              const cpx<R>* __restrict__ tw, int N, int M) {
  const int K = 2 * M + 1, KZ = M + 1;
  const size_t total = size_t(N) * K * KZ;
  const size_t o = size_t(blockIdx.x) * NT + threadIdx.x;
  if (o >= total) return;
  const int iz = int(o % KZ), iy = int((o / KZ) % K), x = int(o / (size_t(KZ) * K));
  const cpx<R>* col = B + size_t(x) * N * KZ + iz;
  A[o] = full_sum_fwd<R>([&](int y) { return col[size_t(y) * KZ]; }, kint(iy, M), N, tw);
}

// x pass of the forward: C[ix][iy][kz] = sum_x A[x][iy][kz] e^{-i kx x}.
template <typename R>
__global__ __launch_bounds__(NT) void g_fwd_x(const cpx<R>* __restrict__ A, cpx<R>* __restrict__ C,
                                              const cpx<R>* __restrict__ tw, int N, int M) {
  const int K = 2 * M + 1, KZ = M + 1;
  const size_t total = size_t(K) * K * KZ, plane = size_t(K) * KZ;
  const size_t o = size_t(blockIdx.x) * NT + threadIdx.x;
  if (o >= total) return;
  const int ix = int(o / plane);
  const cpx<R>* col = A + o % plane;
  C[o] = full_sum_fwd<R>([&](int x) { return col[size_t(x) * plane]; }, kint(ix, M), N, tw);
}

// nl (+)= a * t over n3 points.
template <typename R>
__global__ __launch_bounds__(NT) void g_mul_acc(R* __restrict__ nl, const R* __restrict__ a,
                                                const R* __restrict__ t, size_t n3, int first) {
  const size_t i = size_t(blockIdx.x) * NT + threadIdx.x;
  if (i >= n3) return;
  nl[i] = (first ? R(0) : nl[i]) + a[i] * t[i];
}

// out = a x b (three components each, stride n3).
template <typename R>
__global__ __launch_bounds__(NT) void g_cross(const R* __restrict__ a, const R* __restrict__ b, R* __restrict__ out,
                                              size_t n3) {
  const size_t i = size_t(blockIdx.x) * NT + threadIdx.x;
  if (i >= n3) return;
  const R a0 = a[i], a1 = a[n3 + i], a2 = a[2 * n3 + i];
  const R b0 = b[i], b1 = b[n3 + i], b2 = b[2 * n3 + i];
  out[i] = a1 * b2 - a2 * b1;
  out[n3 + i] = a2 * b0 - a0 * b2;
  out[2 * n3 + i] = a0 * b1 - a1 * b0;
}

// Per-block max |P| over n3 points (grid-stride), one value per block.
template <typename R>
__global__ __launch_bounds__(NT) void g_maxabs(const R* __restrict__ P, size_t n3, R* __restrict__ bmax) {
  R m = R(0);
  for (size_t i = size_t(blockIdx.x) * NT + threadIdx.x; i < n3; i += size_t(gridDim.x) * NT) m = fmax(m, fabs(P[i]));
  __shared__ R red[NT];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = NT / 2; w > 0; w /= 2) {
    if (int(threadIdx.x) < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) bmax[blockIdx.x] = red[0];
}

// Leray projection P(v) = v - k (k . v) / k^2 (k = 0 kept as is).
template <typename R>
__device__ inline void leray(cpx<R>* v, R kx, R ky, R kz) {
  const R k2 = kx * kx + ky * ky + kz * kz;
  const R r = k2 == R(0) ? R(0) : R(1) / k2;
  const cpx<R> kd = {(kx * v[0].x + ky * v[1].x + kz * v[2].x) * r, (kx * v[0].y + ky * v[1].y + kz * v[2].y) * r};
  const R k[3] = {kx, ky, kz};
  for (int c = 0; c < 3; ++c) v[c] = {v[c].x - k[c] * kd.x, v[c].y - k[c] * kd.y};
}

// Right-hand side assembly in the compact layout.  Hydro: ru = -P(NL) -
// nu k^2 U.  MHD (LOR, CR non-null): ru += s P(LOR), rb = P(i k x CR) - eta k^2 B.
template <typename R>
__global__ __launch_bounds__(NT) void g_combine(const cpx<R>* __restrict__ NL, const cpx<R>* __restrict__ LOR,
                                                const cpx<R>* __restrict__ CR, const cpx<R>* __restrict__ U,
                                                const cpx<R>* __restrict__ Bm, cpx<R>* __restrict__ ru,
                                                cpx<R>* __restrict__ rb, R nu, R eta, R sign, R ks, int M) {
  const int K = 2 * M + 1, KZ = M + 1;
  const size_t cs = size_t(K) * K * KZ;
  const size_t pt = size_t(blockIdx.x) * NT + threadIdx.x;
  if (pt >= cs) return;
  const int iz = int(pt % KZ), iy = int((pt / KZ) % K), ix = int(pt / (size_t(KZ) * K));
  const R kx = ks * R(kint(ix, M)), ky = ks * R(kint(iy, M)), kz = ks * R(iz);
  const R k2 = kx * kx + ky * ky + kz * kz;
  cpx<R> a[3], l[3];
  for (int c = 0; c < 3; ++c) a[c] = NL[c * cs + pt];
  leray<R>(a, kx, ky, kz);
  if (LOR) {
    for (int c = 0; c < 3; ++c) l[c] = LOR[c * cs + pt];
    leray<R>(l, kx, ky, kz);
  }
  for (int c = 0; c < 3; ++c) {
    const cpx<R> u = U[c * cs + pt];
    cpx<R> r = {-a[c].x - nu * k2 * u.x, -a[c].y - nu * k2 * u.y};
    if (LOR) r = {r.x + sign * l[c].x, r.y + sign * l[c].y};
    ru[c * cs + pt] = r;
  }
  if (CR) {
    cpx<R> x[3], ind[3];
    for (int c = 0; c < 3; ++c) x[c] = CR[c * cs + pt];
    const R k[3] = {kx, ky, kz};
    for (int c = 0; c < 3; ++c) {
      const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      const cpx<R> t = {k[c1] * x[c2].x - k[c2] * x[c1].x, k[c1] * x[c2].y - k[c2] * x[c1].y};
      ind[c] = {-t.y, t.x};  // i (k x CR)_c
    }
    leray<R>(ind, kx, ky, kz);
    for (int c = 0; c < 3; ++c) {
      const cpx<R> b = Bm[c * cs + pt];
      rb[c * cs + pt] = {ind[c].x - eta * k2 * b.x, ind[c].y - eta * k2 * b.y};
    }
  }
}

// One RK4 stage update over n complex values, k = rhs of the stage input:
//   stage 0: u1 = x0 + dt/2 k,  acc = x0 + dt/6 k
//   stage 1: u1 = x0 + dt/2 k,  acc += dt/3 k
//   stage 2: u1 = x0 + dt k,    acc += dt/3 k
//   stage 3: x0 = acc + dt/6 k
template <typename R>
__global__ __launch_bounds__(NT) void g_rk4(cpx<R>* __restrict__ x0, cpx<R>* __restrict__ u1,
                                            cpx<R>* __restrict__ acc, const cpx<R>* __restrict__ k, size_t n,
                                            int stage, R dt) {
  const size_t i = size_t(blockIdx.x) * NT + threadIdx.x;
  if (i >= n) return;
  const cpx<R> kk = k[i];
  if (stage == 3) {
    const R h = dt / R(6);
    x0[i] = {acc[i].x + h * kk.x, acc[i].y + h * kk.y};
    return;
  }
  const cpx<R> x = x0[i];
  const R h = stage == 2 ? dt : dt / R(2);
  u1[i] = {x.x + h * kk.x, x.y + h * kk.y};
  const R w = stage == 0 ? dt / R(6) : dt / R(3);
  const cpx<R> base = stage == 0 ? x : acc[i];
  acc[i] = {base.x + w * kk.x, base.y + w * kk.y};
}

inline unsigned blocks(size_t n) { return unsigned((n + NT - 1) / NT); }

// Workspace of a generic plan (byte offsets in the tg_plan fields):
//   offA  x-pass buffer  N K KZ complex       offB  y-pass buffer N N KZ complex
//   offP  physical fields (7 hydro, 13 MHD) x N^3 real
//   offC  compact temporaries (3 hydro, 9 MHD) x K K KZ
//   offU0 stage right-hand side, offU1 stage input, offAcc accumulator (3 or 6 components each)
template <typename R>
struct Ws {
  cpx<R>* A;
  cpx<R>* B;
  R* P;
  cpx<R>* C;
  cpx<R>* k;
  cpx<R>* u1;
  cpx<R>* acc;
  double* part;
  R* maxu;
  R* maxb;
  double* out;
};

template <typename R>
Ws<R> carve_ws(const tg_plan* p, void* ws) {
  char* w = static_cast<char*>(ws);
  Ws<R> q;
  q.A = reinterpret_cast<cpx<R>*>(w + p->offA);
  q.B = reinterpret_cast<cpx<R>*>(w + p->offB);
  q.P = reinterpret_cast<R*>(w + p->offP);
  q.C = reinterpret_cast<cpx<R>*>(w + p->offC);
  q.k = reinterpret_cast<cpx<R>*>(w + p->offU0);
  q.u1 = reinterpret_cast<cpx<R>*>(w + p->offU1);
  q.acc = reinterpret_cast<cpx<R>*>(w + p->offAcc);
  q.part = reinterpret_cast<double*>(w + p->offPart);
  q.maxu = reinterpret_cast<R*>(w + p->offMax);
  q.maxb = q.maxu + p->nMax;
  q.out = reinterpret_cast<double*>(w + p->offOut);
  return q;
}

template <typename R>
struct GenOps {
  // Physical field P (N^3) of operand (op, c, d) of the compact spectrum S.
  static int inv1(const tg_plan* p, const cpx<R>* S, int op, int c, int d, R* P, const Ws<R>& w, hipStream_t s) {
    const int N = p->n, M = p->M, K = p->K, KZ = p->KZ;
    const cpx<R>* tw = static_cast<const cpx<R>*>(p->tw);
    const double n3 = double(N) * N * N;
    hipLaunchKernelGGL((g_inv_x<R>), dim3(blocks(size_t(N) * K * KZ)), dim3(NT), 0, s, S, op, c, d, w.A, tw, N, M,
                       R(p->ks));
    hipLaunchKernelGGL((g_inv_y<R>), dim3(blocks(size_t(N) * N * KZ)), dim3(NT), 0, s, w.A, w.B, tw, N, M);
    hipLaunchKernelGGL((g_inv_z<R>), dim3(blocks(size_t(N) * N * N)), dim3(NT), 0, s, w.B, P, tw, N, M,
                       R(1.0 / n3));
    TG_LAUNCHED();
    return TG_OK;
  }
  // Compact spectrum C of the physical field P.
  static int fwd1(const tg_plan* p, const R* P, cpx<R>* C, const Ws<R>& w, hipStream_t s) {
    const int N = p->n, M = p->M, K = p->K, KZ = p->KZ;
    const cpx<R>* tw = static_cast<const cpx<R>*>(p->tw);
    hipLaunchKernelGGL((g_fwd_z<R>), dim3(blocks(size_t(N) * N * KZ)), dim3(NT), 0, s, P, w.B, tw, N, M);
    hipLaunchKernelGGL((g_fwd_y<R>), dim3(blocks(size_t(N) * K * KZ)), dim3(NT), 0, s, w.B, w.A, tw, N, M);
    hipLaunchKernelGGL((g_fwd_x<R>), dim3(blocks(size_t(K) * K * KZ)), dim3(NT), 0, s, w.A, C, tw, N, M);
    TG_LAUNCHED();
    return TG_OK;
  }

  static int forward(const tg_plan* p, const R* phys, cpx<R>* spec, void* ws, hipStream_t s) {
    const Ws<R> w = carve_ws<R>(p, ws);
    const size_t n3 = size_t(p->n) * p->n * p->n;
    for (int c = 0; c < 3; ++c) TG_RET(fwd1(p, phys + c * n3, spec + c * p->compact, w, s));
    return TG_OK;
  }

  static int inverse(const tg_plan* p, const cpx<R>* spec, int op, R* phys, void* ws, hipStream_t s) {
    const Ws<R> w = carve_ws<R>(p, ws);
    const size_t n3 = size_t(p->n) * p->n * p->n;
    if (op == TG_OP_DIV) return inv1(p, spec, OP_DIV, 0, 0, phys, w, s);
    for (int c = 0; c < 3; ++c) {
      if (op == TG_OP_GRAD) TG_RET(inv1(p, spec, OP_DERIV, 0, c, phys + c * n3, w, s));
      else TG_RET(inv1(p, spec, op == TG_OP_CURL ? OP_CURL : OP_FIELD, c, 0, phys + c * n3, w, s));
    }
    return TG_OK;
  }

  // F[u . grad u] of the compact velocity u into the compact array NL;
  // Pu holds the physical u on return.  Physical scratch: Pu (3), T (1), W (3).
  static int advect(const tg_plan* p, const cpx<R>* u, R* Pu, R* T, R* W, cpx<R>* NL, const Ws<R>& w,
                    hipStream_t s) {
    const size_t n3 = size_t(p->n) * p->n * p->n;
    for (int c = 0; c < 3; ++c) TG_RET(inv1(p, u, OP_FIELD, c, 0, Pu + c * n3, w, s));
    for (int c = 0; c < 3; ++c) {
      for (int d = 0; d < 3; ++d) {
        TG_RET(inv1(p, u, OP_DERIV, c, d, T, w, s));
        hipLaunchKernelGGL((g_mul_acc<R>), dim3(blocks(n3)), dim3(NT), 0, s, W + c * n3, Pu + d * n3, T, n3,
                           int(d == 0));
      }
      TG_LAUNCHED();
    }
    for (int c = 0; c < 3; ++c) TG_RET(fwd1(p, W + c * n3, NL + c * p->compact, w, s));
    return TG_OK;
  }

  static int rhs_into(const tg_plan* p, const cpx<R>* u, const cpx<R>* b, double nu, double eta, double sign,
                      cpx<R>* ru, cpx<R>* rb, const Ws<R>& w, hipStream_t s) {
    const size_t n3 = size_t(p->n) * p->n * p->n, cs = p->compact;
    R* Pu = w.P;
    R* T = Pu + 3 * n3;
    R* W = T + n3;
    cpx<R>* NL = w.C;
    TG_RET(advect(p, u, Pu, T, W, NL, w, s));
    cpx<R>* LOR = nullptr;
    cpx<R>* CR = nullptr;
    if (b) {
      R* Pb = W + 3 * n3;
      R* Pj = Pb + 3 * n3;
      LOR = NL + 3 * cs;
      CR = LOR + 3 * cs;
      for (int c = 0; c < 3; ++c) {
        TG_RET(inv1(p, b, OP_FIELD, c, 0, Pb + c * n3, w, s));
        TG_RET(inv1(p, b, OP_CURL, c, 0, Pj + c * n3, w, s));
      }
      hipLaunchKernelGGL((g_cross<R>), dim3(blocks(n3)), dim3(NT), 0, s, Pj, Pb, W, n3);
      TG_LAUNCHED();
      for (int c = 0; c < 3; ++c) TG_RET(fwd1(p, W + c * n3, LOR + c * cs, w, s));
      hipLaunchKernelGGL((g_cross<R>), dim3(blocks(n3)), dim3(NT), 0, s, Pu, Pb, W, n3);
      TG_LAUNCHED();
      for (int c = 0; c < 3; ++c) TG_RET(fwd1(p, W + c * n3, CR + c * cs, w, s));
    }
    hipLaunchKernelGGL((g_combine<R>), dim3(blocks(cs)), dim3(NT), 0, s, NL, LOR, CR, u, b, ru, rb, R(nu), R(eta),
                       R(sign), R(p->ks), p->M);
    TG_LAUNCHED();
    return TG_OK;
  }

  static int rhs(const tg_plan* p, const cpx<R>* u, const cpx<R>* b, double nu, double eta, double sign,
                 cpx<R>* ru, cpx<R>* rb, void* ws, hipStream_t s) {
    return rhs_into(p, u, p->mhd ? b : nullptr, nu, eta, sign, ru, p->mhd ? rb : nullptr, carve_ws<R>(p, ws), s);
  }

  static int rk4(const tg_plan* p, cpx<R>* u, cpx<R>* b, int nsteps, double dt, double nu, double eta, double sign,
                 void* ws, hipStream_t s) {
    const Ws<R> w = carve_ws<R>(p, ws);
    const size_t cs = p->compact, nv = 3 * cs;
    const bool mhd = p->mhd != 0;
    cpx<R>* ku = w.k;
    cpx<R>* kb = w.k + nv;
    for (int step = 0; step < nsteps; ++step) {
      for (int st = 0; st < 4; ++st) {
        const cpx<R>* iu = st == 0 ? u : w.u1;
        const cpx<R>* ib = st == 0 ? b : w.u1 + nv;
        TG_RET(rhs_into(p, iu, mhd ? ib : nullptr, nu, eta, sign, ku, mhd ? kb : nullptr, w, s));
        hipLaunchKernelGGL((g_rk4<R>), dim3(blocks(nv)), dim3(NT), 0, s, u, w.u1, w.acc, ku, nv, st, R(dt));
        if (mhd)
          hipLaunchKernelGGL((g_rk4<R>), dim3(blocks(nv)), dim3(NT), 0, s, b, w.u1 + nv, w.acc + nv, kb, nv, st,
                             R(dt));
        TG_LAUNCHED();
      }
    }
    return TG_OK;
  }

  static int diagnostics(const tg_plan* p, const cpx<R>* u, const cpx<R>* b, int divmax, double* out, void* ws,
                         hipStream_t s) {
    const Ws<R> w = carve_ws<R>(p, ws);
    const size_t n3 = size_t(p->n) * p->n * p->n;
    hipLaunchKernelGGL((k_diag<R>), dim3(p->nDiagBlk), dim3(DIAG_NT), 0, s, u, b, w.part, p->ks, p->M);
    TG_LAUNCHED();
    if (divmax) {
      TG_RET(inv1(p, u, OP_DIV, 0, 0, w.P, w, s));
      hipLaunchKernelGGL((g_maxabs<R>), dim3(p->nMax), dim3(NT), 0, s, w.P, n3, w.maxu);
      if (b) {
        TG_RET(inv1(p, b, OP_DIV, 0, 0, w.P, w, s));
        hipLaunchKernelGGL((g_maxabs<R>), dim3(p->nMax), dim3(NT), 0, s, w.P, n3, w.maxb);
      }
      TG_LAUNCHED();
    }
    hipLaunchKernelGGL((k_diag_final<R>), dim3(1), dim3(64), 0, s, w.part, p->nDiagBlk, divmax ? w.maxu : nullptr,
                       (divmax && b) ? w.maxb : nullptr, divmax ? p->nMax : 0, w.out);
    TG_LAUNCHED();
    double host[NDIAG + 2];
    TG_HIP(hipMemcpyAsync(host, w.out, sizeof(host), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    const double n6 = double(n3) * double(n3);
    out[0] = 0.5 * host[0] / n6;
    out[1] = 0.5 * host[1] / n6;
    out[2] = host[2] / n6;
    out[3] = 0.5 * host[3] / n6;
    for (int q = 4; q < 8; ++q) out[q] = host[q] / n6;
    out[8] = host[NDIAG];
    out[9] = host[NDIAG + 1];
    out[10] = host[8];
    return TG_OK;
  }

  static int unsupported_vjp(const tg_plan* p, const cpx<R>*, const cpx<R>*, double, cpx<R>*, void*, hipStream_t) {
    return fail(TG_ERR_UNSUPPORTED, "adjoint: N=" + std::to_string(p->n) + " has no compiled plan");
  }
  static int unsupported_fwd(const tg_plan* p, const cpx<R>*, int, double, double, int, int, double*, void*, void*,
                             hipStream_t) {
    return fail(TG_ERR_UNSUPPORTED, "adjoint: N=" + std::to_string(p->n) + " has no compiled plan");
  }
  static int unsupported_bwd(const tg_plan* p, int, double, double, int, cpx<R>*, double*, void*, void*,
                             hipStream_t) {
    return fail(TG_ERR_UNSUPPORTED, "adjoint: N=" + std::to_string(p->n) + " has no compiled plan");
  }
  // The allocation-free diagnostics need the compiled passes (the C ABI
  // rejects generic plans before asking for the scratch size).
  static size_t no_scratch(const tg_plan*, int) { return 0; }
  static int unsupported_diag_dev(const tg_plan* p, const cpx<R>*, const cpx<R>*, int, double*, void*, hipStream_t) {
    return fail(TG_ERR_UNSUPPORTED, "diagnostics on device: N=" + std::to_string(p->n) + " has no compiled plan");
  }
};

template <typename R>
OpsTable<R> make_generic_table() {
  using O = GenOps<R>;
  OpsTable<R> t;
  t.n = 0;  // any N
  t.kzp = 0;
  t.nmax = 1024;
  t.forward = &O::forward;
  t.inverse = &O::inverse;
  t.rhs = &O::rhs;
  t.rk4 = &O::rk4;
  t.vjp = &O::unsupported_vjp;
  t.adj_forward = &O::unsupported_fwd;
  t.adj_backward = &O::unsupported_bwd;
  t.diagnostics = &O::diagnostics;
  t.diag_scratch_bytes = &O::no_scratch;
  t.diagnostics_dev = &O::unsupported_diag_dev;
  return t;
}

}  // namespace gen
}  // namespace tg

extern const tg::OpsTable<double> tg_ops_64_generic;
extern const tg::OpsTable<float> tg_ops_32_generic;
const tg::OpsTable<double> tg_ops_64_generic = tg::gen::make_generic_table<double>();
const tg::OpsTable<float> tg_ops_32_generic = tg::gen::make_generic_table<float>();
