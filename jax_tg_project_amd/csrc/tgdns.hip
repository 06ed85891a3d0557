// libtgdns.so: host side of the C ABI declared in include/tgdns.h.
// Plans own twiddle tables; callers own every other buffer.  The kernels and
// pass sequences of each (precision, grid size) live in their own translation
// unit (tg_inst.hip); this file dispatches to them through their tables.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "tg_common.hpp"
#include "tg_ops.hpp"
#include "tg_batch.hpp"

using namespace tg;

namespace tg {

namespace {
thread_local std::string g_err;
}

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace tg

// Tables exported by the tg_inst.hip objects, one per (precision, N).
#define TG_DECL_TABLES(n)                                                                  \
  extern const tg::OpsTable<double> tg_ops_64_##n;                                          \
  extern const tg::OpsTable<float> tg_ops_32_##n;                                           \
  extern const tg::EnsTable<double> tg_ens_64_##n;                                          \
  extern const tg::EnsTable<float> tg_ens_32_##n;
TG_GRID_SIZES(TG_DECL_TABLES)
#undef TG_DECL_TABLES
// Mixed-size ensemble launches (tg_ens_mixed.hip).
extern const tg::EnsMixTable<double> tg_ens_mix_64;
extern const tg::EnsMixTable<float> tg_ens_mix_32;
// Every other size: the direct-sum path of tg_generic.hip.
extern const tg::OpsTable<double> tg_ops_64_generic;
extern const tg::OpsTable<float> tg_ops_32_generic;

namespace {

#define TG_LIST_SIZE(n) n,
constexpr int kSizes[] = {TG_GRID_SIZES(TG_LIST_SIZE)};
#undef TG_LIST_SIZE
// Largest grid a plan accepts (the generic path indexes N^3 points in 64 bits
// and phases k n in 64 bits; the bound keeps a workspace request sane).
constexpr int kMaxGeneric = 2048;

template <typename R>
const OpsTable<R>* ops_table(int n);
template <>
const OpsTable<double>* ops_table<double>(int n) {
#define TG_CASE(s) \
  if (n == s) return &tg_ops_64_##s;
  TG_GRID_SIZES(TG_CASE)
#undef TG_CASE
  return nullptr;
}
template <>
const OpsTable<float>* ops_table<float>(int n) {
#define TG_CASE(s) \
  if (n == s) return &tg_ops_32_##s;
  TG_GRID_SIZES(TG_CASE)
#undef TG_CASE
  return nullptr;
}
template <typename R>
const EnsTable<R>* ens_table(int n);
template <>
const EnsTable<double>* ens_table<double>(int n) {
#define TG_CASE(s) \
  if (n == s) return tg_ens_64_##s.n ? &tg_ens_64_##s : nullptr;
  TG_GRID_SIZES(TG_CASE)
#undef TG_CASE
  return nullptr;
}
template <>
const EnsTable<float>* ens_table<float>(int n) {
#define TG_CASE(s) \
  if (n == s) return tg_ens_32_##s.n ? &tg_ens_32_##s : nullptr;
  TG_GRID_SIZES(TG_CASE)
#undef TG_CASE
  return nullptr;
}

template <typename R>
const EnsMixTable<R>* mix_table();
template <>
const EnsMixTable<double>* mix_table<double>() { return &tg_ens_mix_64; }
template <>
const EnsMixTable<float>* mix_table<float>() { return &tg_ens_mix_32; }

// Call c(table, (R*)nullptr) with the plan's entry-point table.
template <typename Call>
int with_table(const tg_plan* p, Call&& c) {
  if (p->prec == 64) return c(*static_cast<const OpsTable<double>*>(p->ops), (double*)nullptr);
  return c(*static_cast<const OpsTable<float>*>(p->ops), (float*)nullptr);
}

int check_plan(const tg_plan* p) {
  if (!p) return fail(TG_ERR_ARG, "plan is NULL");
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_version(void) { return 2; }

int tg_supported_sizes(int* out, int cap) {
  const int n = int(sizeof(kSizes) / sizeof(kSizes[0]));
  for (int i = 0; i < n && i < cap; ++i) out[i] = kSizes[i];
  return n;
}

int tg_plan_create(int n, int precision_bits, int mhd, double domain_length, tg_plan** out) {
  if (!out) return fail(TG_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (precision_bits != 64 && precision_bits != 32)
    return fail(TG_ERR_ARG, "precision_bits must be 64 or 32");
  if (n < 4 || n > kMaxGeneric)
    return fail(TG_ERR_UNSUPPORTED, "grid size N=" + std::to_string(n) + " outside [4, " +
                                        std::to_string(kMaxGeneric) + "]");
  const void* ops = precision_bits == 64 ? static_cast<const void*>(ops_table<double>(n))
                                         : static_cast<const void*>(ops_table<float>(n));
  const bool generic = ops == nullptr;
  if (generic)
    ops = precision_bits == 64 ? static_cast<const void*>(&tg_ops_64_generic)
                               : static_cast<const void*>(&tg_ops_32_generic);
  if (!(domain_length > 0.0)) return fail(TG_ERR_ARG, "domain_length must be positive");
  tg_plan* p = new tg_plan{};
  p->n = n;
  p->prec = precision_bits;
  p->mhd = mhd ? 1 : 0;
  p->ops = ops;
  p->generic = generic;
  p->M = n / 3;
  p->K = 2 * p->M + 1;
  p->KZ = p->M + 1;
  p->ks = 2.0 * M_PI / domain_length;
  if (domain_length == 2.0 * M_PI) p->ks = 1.0;
  p->L = domain_length;
  p->cb = precision_bits == 64 ? 16 : 8;
  p->compact = size_t(p->K) * p->K * p->KZ;
  const size_t N = n;
  p->nDiagBlk = 512;
  size_t o = 0;
  if (generic) {
    // x-pass and y-pass buffers of one field, the physical fields of the
    // RHS (hydro u, T, W; MHD also b, j), compact temporaries (F[u.grad u];
    // MHD also F[j x b], F[u x b]) and the three RK4 vectors (rhs, stage
    // input, accumulator) in the dense compact layout
    const size_t sA = N * p->K * p->KZ * p->cb, sB = N * N * p->KZ * p->cb;
    const size_t sP = (p->mhd ? 13 : 7) * N * N * N * (p->cb / 2);
    const size_t sC = (p->mhd ? 9 : 3) * p->compact * p->cb;
    const size_t sU = (p->mhd ? 6 : 3) * p->compact * p->cb;
    p->nMax = precision_bits == 64 ? tg_ops_64_generic.nmax : tg_ops_32_generic.nmax;
    p->offA = o; o += align256(sA);
    p->offB = o; o += align256(sB);
    p->offP = o; o += align256(sP);
    p->offC = o; o += align256(sC);
    p->offU0 = o; o += align256(sU);
    p->offU1 = o; o += align256(sU);
    p->offAcc = o; o += align256(sU);
  } else {
    const size_t fi = p->mhd ? 12 : 9;  // inverse fields: hydro RHS 6, adjoint 9, MHD 12
    const size_t fo = 6;                 // z-pass outputs: hydro 3, adjoint and MHD 6
    // A holds the inverse-x output (fi x N K KZ) and later the z output (fo x N N KZ)
    const size_t kzpad = precision_bits == 64 ? size_t(ops_table<double>(n)->kzp) : size_t(ops_table<float>(n)->kzp);
    const size_t sA = std::max(fi * N * p->K, fo * N * N) * kzpad * p->cb;
    const size_t sB = fi * N * N * kzpad * p->cb;
    // padded state vectors (row pitch kzpad) for the RK4 loop
    const size_t sU = (p->mhd ? 2 : 1) * 3 * size_t(p->K) * p->K * kzpad * p->cb;
    p->nMax = precision_bits == 64 ? ops_table<double>(n)->nmax : ops_table<float>(n)->nmax;
    p->offA = o; o += align256(sA);
    p->offB = o; o += align256(sB);
    p->offU0 = o; o += align256(sU);
    p->offU1 = o; o += align256(sU);
    p->offAcc = o; o += align256(sU);
  }
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
  info[TG_INFO_GENERIC] = p->generic ? 1 : 0;
  return TG_OK;
}

int tg_forward(tg_plan* p, const void* phys, void* spec, void* ws, void* stream) {
  TG_RET(check_plan(p));
  if (!phys || !spec || !ws) return fail(TG_ERR_ARG, "tg_forward: NULL buffer");
  return with_table(p, [&](const auto& T, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return T.forward(p, static_cast<const R*>(phys), static_cast<cpx<R>*>(spec), ws,
                                  static_cast<hipStream_t>(stream));
  });
}

int tg_inverse(tg_plan* p, const void* spec, int op, void* phys, void* ws, void* stream) {
  TG_RET(check_plan(p));
  if (!phys || !spec || !ws) return fail(TG_ERR_ARG, "tg_inverse: NULL buffer");
  if (op != TG_OP_FIELD && op != TG_OP_CURL && op != TG_OP_DIV && op != TG_OP_GRAD)
    return fail(TG_ERR_ARG, "tg_inverse: op must be TG_OP_FIELD, TG_OP_CURL, TG_OP_DIV or TG_OP_GRAD");
  return with_table(p, [&](const auto& T, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return T.inverse(p, static_cast<const cpx<R>*>(spec), op, static_cast<R*>(phys), ws,
                                  static_cast<hipStream_t>(stream));
  });
}

int tg_rhs(tg_plan* p, const void* u_hat, const void* b_hat, double nu, double eta, double lorentz_sign,
           void* ru_hat, void* rb_hat, void* ws, void* stream) {
  TG_RET(check_plan(p));
  if (!u_hat || !ru_hat || !ws) return fail(TG_ERR_ARG, "tg_rhs: NULL buffer");
  if (p->mhd && (!b_hat || !rb_hat)) return fail(TG_ERR_ARG, "tg_rhs: MHD plan needs b_hat and rb_hat");
  return with_table(p, [&](const auto& T, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return T.rhs(p, static_cast<const cpx<R>*>(u_hat), static_cast<const cpx<R>*>(b_hat), nu, eta,
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
  return with_table(p, [&](const auto& T, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return T.rk4(p, static_cast<cpx<R>*>(u_hat), static_cast<cpx<R>*>(b_hat), nsteps, dt, nu, eta,
                              lorentz_sign, ws, static_cast<hipStream_t>(stream));
  });
}

int tg_diagnostics(tg_plan* p, const void* u_hat, const void* b_hat, int with_divmax, double* out, void* ws,
                   void* stream) {
  TG_RET(check_plan(p));
  if (!u_hat || !out || !ws) return fail(TG_ERR_ARG, "tg_diagnostics: NULL buffer");
  return with_table(p, [&](const auto& T, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return T.diagnostics(p, static_cast<const cpx<R>*>(u_hat), static_cast<const cpx<R>*>(b_hat),
                                      with_divmax, out, ws, static_cast<hipStream_t>(stream));
  });
}

int tg_vjp(tg_plan* p, const void* x_hat, const void* a_hat, double nu, void* out_hat, void* ws, void* stream) {
  TG_RET(check_plan(p));
  if (p->mhd) return fail(TG_ERR_UNSUPPORTED, "tg_vjp: hydro plans only");
  if (!x_hat || !a_hat || !out_hat || !ws) return fail(TG_ERR_ARG, "tg_vjp: NULL buffer");
  return with_table(p, [&](const auto& T, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return T.vjp(p, static_cast<const cpx<R>*>(x_hat), static_cast<const cpx<R>*>(a_hat), nu,
                              static_cast<cpx<R>*>(out_hat), ws, static_cast<hipStream_t>(stream));
  });
}

int tg_adjoint_workspace(tg_plan* p, int nsteps, int seg, long long* bytes) {
  TG_RET(check_plan(p));
  if (nsteps < 1 || seg < 1 || !bytes) return fail(TG_ERR_ARG, "tg_adjoint_workspace: need nsteps, seg >= 1");
  if (p->generic) return fail(TG_ERR_UNSUPPORTED, "adjoint: N=" + std::to_string(p->n) + " has no compiled plan");
  const size_t kzpad = p->prec == 64 ? size_t(ops_table<double>(p->n)->kzp) : size_t(ops_table<float>(p->n)->kzp);
  const size_t vb = align256(3 * size_t(p->K) * p->K * kzpad * p->cb);  // line-padded vector
  const size_t nck = (nsteps + seg - 1) / seg;
  size_t b = (4 + nck + 4 * size_t(seg)) * vb + align256(size_t(p->K) * p->KZ * 8 * sizeof(double));
  if (nck > 1) b += (1 + 4 * size_t(seg)) * vb + align256(p->wsBytes);  // overlapped recompute
  const size_t sb = align256(6 * size_t(p->n) * p->n * kzpad * p->cb);
  if (adj_reg_z_rt(p->n))  // kept inverse-y output (u, w) per stage, per stage buffer
    b += (nck > 1 ? 2 : 1) * 4 * size_t(seg) * sb;
  for (int k = 2; k < kAdjKeptSegments; ++k)  // a stage buffer per further kept segment
    if (nck > size_t(k)) b += 4 * size_t(seg) * (vb + (adj_reg_z_rt(p->n) ? sb : 0));
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
  return with_table(p, [&](const auto& T, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return T.adj_forward(p, static_cast<const cpx<R>*>(x0_hat), nsteps, dt, nu, objective, seg, J,
                                      ws, adj_ws, static_cast<hipStream_t>(stream));
  });
}

int tg_adjoint_backward(tg_plan* p, int nsteps, double dt, double nu, int seg, void* grad_hat, double* dJdnu,
                        void* ws, void* adj_ws, void* stream) {
  TG_RET(check_plan(p));
  if (p->mhd) return fail(TG_ERR_UNSUPPORTED, "adjoint: hydro plans only");
  if (!grad_hat || !dJdnu || !ws || !adj_ws) return fail(TG_ERR_ARG, "tg_adjoint_backward: NULL buffer");
  if (nsteps < 1 || seg < 1) return fail(TG_ERR_ARG, "adjoint: need nsteps, seg >= 1");
  return with_table(p, [&](const auto& T, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    return T.adj_backward(p, nsteps, dt, nu, seg, static_cast<cpx<R>*>(grad_hat), dJdnu, ws, adj_ws,
                                       static_cast<hipStream_t>(stream));
  });
}

int tg_plan_variants(tg_plan* p, const int* set, int* get, int* count) {
  TG_RET(check_plan(p));
  const int nv = (!p->generic && p->n >= 128) ? NVAR : 1;
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

int tg_plan_recompute_variants(tg_plan* p, const int* set, int* get) {
  TG_RET(check_plan(p));
  const int nv = (!p->generic && p->n >= 128) ? NVAR : 1;
  if (set) {
    if (set[0] == -1) {
      for (int i = 0; i < TG_NPASS; ++i) p->rvar[i] = -1;
    } else {
      for (int i = 0; i < TG_NPASS; ++i)
        if (set[i] < 0 || set[i] >= nv) return fail(TG_ERR_ARG, "tg_plan_recompute_variants: variant out of range");
      for (int i = 0; i < TG_NPASS; ++i) p->rvar[i] = set[i];
    }
  }
  if (get)
    for (int i = 0; i < TG_NPASS; ++i) get[i] = p->rvar[0] >= 0 ? p->rvar[i] : p->var[i];
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
  std::vector<void*> dense, padded;  // per instance: caller's state, working copy
  std::vector<int> order;            // descriptor slot -> instance (grouped by size)
  std::vector<int> cls_n, cls_first, cls_count;  // launch classes (one grid size each) over descriptor slots
  std::vector<size_t> cls_prefix_off;            // offset of each class's EP_COUNT x (count+1) table
  // mixed group: the contiguous descriptor slots of the kMixSizes instances
  // (largest size first), with its own prefix tables; `mixed` selects one
  // launch per pass for the group instead of one per class
  int mix_first = 0, mix_count = 0;
  size_t mix_prefix_off = 0;
  bool mixed = false;
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

// The four stages of one RK4 step of size class c on stream s.
template <typename R>
int batch_class_step(tg_batch* b, int c, double dt, hipStream_t s) {
  const EnsTable<R>* t = ens_table<R>(b->cls_n[c]);
  if (!t) return fail(TG_ERR_UNSUPPORTED, "ensemble size class");
  const int first = b->cls_first[c], cnt = b->cls_count[c];
  const auto* inst = static_cast<const EnsInst<R>*>(b->d_inst) + first;
  const size_t po = b->cls_prefix_off[c];
  for (int st = 0; st < 4; ++st) TG_RET(t->stage(inst, b->d_prefix + po, b->prefix.data() + po, cnt, st, dt, s));
  return TG_OK;
}

// The four stages of one RK4 step of the mixed group on stream s.
template <typename R>
int batch_mix_step(tg_batch* b, double dt, hipStream_t s) {
  const auto* inst = static_cast<const EnsInst<R>*>(b->d_inst) + b->mix_first;
  const size_t po = b->mix_prefix_off;
  for (int st = 0; st < 4; ++st)
    TG_RET(mix_table<R>()->stage(inst, b->d_prefix + po, b->prefix.data() + po, b->mix_count, st, dt, s));
  return TG_OK;
}

// Launch units of one step: the mixed group (when selected) and every class
// outside it, or every class.  Unit -1 is the mixed group.
std::vector<int> batch_units(const tg_batch* b) {
  std::vector<int> u;
  const bool mix = b->mixed && b->mix_count > 0;
  if (mix) u.push_back(-1);
  for (int c = 0; c < int(b->cls_n.size()); ++c)
    if (!mix || !is_mix_size(b->cls_n[c])) u.push_back(c);
  return u;
}

template <typename R>
int batch_unit_step(tg_batch* b, int unit, double dt, hipStream_t s) {
  return unit < 0 ? batch_mix_step<R>(b, dt, s) : batch_class_step<R>(b, unit, dt, s);
}

// One RK4 step of every instance.  Launch units are independent: with
// `branches` (the graph capture) unit u >= 1 runs on its own stream forked
// from and joined back to s, so the captured graph holds one parallel
// branch per unit; without, the units run one after the other on s.
template <typename R>
int batch_launch_step(tg_batch* b, double dt, hipStream_t s, bool branches) {
  const std::vector<int> units = batch_units(b);
  const int nu = int(units.size());
  if (!branches || nu == 1) {
    for (int u : units) TG_RET(batch_unit_step<R>(b, u, dt, s));
    return TG_OK;
  }
  TG_HIP(hipEventRecord(b->fork[0], s));
  for (int k = 1; k < nu; ++k) {
    TG_HIP(hipStreamWaitEvent(b->branch[k - 1], b->fork[0], 0));
    TG_RET(batch_unit_step<R>(b, units[k], dt, b->branch[k - 1]));
    TG_HIP(hipEventRecord(b->fork[k], b->branch[k - 1]));
  }
  TG_RET(batch_unit_step<R>(b, units[0], dt, s));
  for (int k = 1; k < nu; ++k) TG_HIP(hipStreamWaitEvent(s, b->fork[k], 0));
  return TG_OK;
}

template <typename R>
int batch_bind(tg_batch* b, void* state, void* ws) {
  std::vector<EnsInst<R>> h(b->ninst);
  char* w = static_cast<char*>(ws);
  cpx<R>* st = static_cast<cpx<R>*>(state);
  b->dense.assign(b->ninst, nullptr);
  b->padded.assign(b->ninst, nullptr);
  for (int slot = 0; slot < b->ninst; ++slot) {
    const int i = b->order[slot];
    const int N = b->n[i];
    const EnsTable<R>* t = ens_table<R>(N);
    const size_t K = 2 * (N / 3) + 1;
    const size_t vec = 3 * K * size_t(t->pa) * sizeof(cpx<R>);
    const size_t sA = size_t(t->a_elems) * sizeof(cpx<R>);
    char* base = w + b->ws_off[i];
    EnsInst<R>& I = h[slot];
    I.n = N;
    I.nu = R(b->nu[i]);
    I.ks = R(b->ks);
    I.s0 = reinterpret_cast<cpx<R>*>(base);
    I.u1 = reinterpret_cast<cpx<R>*>(base + align256(vec));
    I.acc = reinterpret_cast<cpx<R>*>(base + 2 * align256(vec));
    I.A = reinterpret_cast<cpx<R>*>(base + 3 * align256(vec));
    I.B = reinterpret_cast<cpx<R>*>(base + 3 * align256(vec) + align256(sA));
    I.tw = static_cast<const cpx<R>*>(b->tw_by_size[ens_size_index(N)]);
    b->dense[i] = st + b->state_off[i];
    b->padded[i] = I.s0;
  }
  TG_HIP(hipMemcpy(b->d_inst, h.data(), h.size() * sizeof(EnsInst<R>), hipMemcpyHostToDevice));
  b->bound = true;
  return TG_OK;
}

// Copy every instance's caller-owned dense state into its plane-padded
// working copy (to_padded) or back.
int batch_copy_state(tg_batch* b, bool to_padded, hipStream_t s) {
  for (int i = 0; i < b->ninst; ++i) {
    const int N = b->n[i];
    const size_t K = 2 * (N / 3) + 1, KZ = N / 3 + 1;
    const size_t pa = b->prec == 64 ? size_t(ens_table<double>(N)->pa) : size_t(ens_table<float>(N)->pa);
    const size_t row = K * KZ * b->cb, prow = pa * b->cb;
    if (to_padded)
      TG_HIP(hipMemcpy2DAsync(b->padded[i], prow, b->dense[i], row, row, 3 * K, hipMemcpyDeviceToDevice, s));
    else
      TG_HIP(hipMemcpy2DAsync(b->dense[i], row, b->padded[i], prow, row, 3 * K, hipMemcpyDeviceToDevice, s));
  }
  return TG_OK;
}

}  // namespace

namespace {
// The steps of tg_batch_rk4_steps on the instances' working copies.
int batch_steps(tg_batch* b, int nsteps, double dt, int use_graph, hipStream_t s) {
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
    const size_t ncls = batch_units(b).size();
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
    b->state_off[i + 1] = b->state_off[i] + (long long)(3 * K * K * KZ);
    size_t pa, ae, be;
    if (precision_bits == 64) {
      const EnsTable<double>* t = ens_table<double>(n[i]);
      pa = t->pa; ae = t->a_elems; be = t->b_elems;
    } else {
      const EnsTable<float>* t = ens_table<float>(n[i]);
      pa = t->pa; ae = t->a_elems; be = t->b_elems;
    }
    const size_t vec = 3 * K * pa * b->cb;  // state, stage input, accumulator
    b->ws_off[i + 1] = b->ws_off[i] + 3 * align256(vec) + align256(ae * b->cb) + align256(be * b->cb);
  }

  // descriptor slots grouped by size, largest first (the mixable sizes are
  // then contiguous, and the largest blocks of a mixed launch go first)
  for (int k = kNumEnsSizes - 1; k >= 0; --k) {
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
  // A class holding more than 40% of the work (instances x N^3) is split in
  // two halves, each its own launch sequence (and graph branch): while one
  // half runs its instruction-bound z pass the other streams through a
  // memory-bound pass.
  {
    double total = 0.0;
    for (int i = 0; i < ninst; ++i) total += double(n[i]) * n[i] * n[i];
    for (size_t c = 0; c < b->cls_n.size(); ++c) {
      const double w = double(b->cls_count[c]) * b->cls_n[c] * b->cls_n[c] * b->cls_n[c];
      if (b->cls_count[c] > 1 && w > 0.4 * total) {
        const int h = b->cls_count[c] / 2;
        b->cls_n.insert(b->cls_n.begin() + c + 1, b->cls_n[c]);
        b->cls_first.insert(b->cls_first.begin() + c + 1, b->cls_first[c] + h);
        b->cls_count.insert(b->cls_count.begin() + c + 1, b->cls_count[c] - h);
        b->cls_count[c] = h;
        ++c;
      }
    }
  }
  for (size_t c = 0; c < b->cls_n.size(); ++c) {
    const int cnt = b->cls_count[c];
    b->cls_prefix_off.push_back(b->prefix.size());
    for (int pass = 0; pass < EP_COUNT; ++pass) {
      const int per = precision_bits == 64 ? ens_table<double>(b->cls_n[c])->blocks[pass]
                                           : ens_table<float>(b->cls_n[c])->blocks[pass];
      for (int j = 0; j <= cnt; ++j) b->prefix.push_back(j * per);
    }
  }
  // the mixed group's prefix tables: every kMixSizes instance, in slot order
  {
    int first = -1, cnt = 0;
    for (int slot = 0; slot < ninst; ++slot)
      if (is_mix_size(n[b->order[slot]])) {
        if (first < 0) first = slot;
        ++cnt;
      }
    if (cnt > 0) {
      b->mix_first = first;
      b->mix_count = cnt;
      b->mix_prefix_off = b->prefix.size();
      for (int pass = 0; pass < EP_COUNT; ++pass) {
        int acc = 0;
        b->prefix.push_back(0);
        for (int j = 0; j < cnt; ++j) {
          const int N = n[b->order[first + j]];
          acc += precision_bits == 64 ? ens_table<double>(N)->blocks[pass] : ens_table<float>(N)->blocks[pass];
          b->prefix.push_back(acc);
        }
      }
      b->mixed = true;
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

int tg_batch_set_launch(tg_batch* b, int mixed) {
  if (!b) return fail(TG_ERR_ARG, "batch is NULL");
  const bool m = mixed != 0;
  if (m != b->mixed && b->exec) {
    TG_HIP(hipGraphExecDestroy(b->exec));
    b->exec = nullptr;
  }
  b->mixed = m;
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
  if (nsteps == 0) return TG_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  TG_RET(batch_copy_state(b, true, s));
  TG_RET(batch_steps(b, nsteps, dt, use_graph, s));
  return batch_copy_state(b, false, s);
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free diagnostics
// tg_diagnostics_noalloc keeps every intermediate in this file-scope device
// array (one copy per GPU, held until the process ends).  Host code reaches
// it only through the per-device address from hipGetSymbolAddress.
namespace {

constexpr size_t kScratchBytes = size_t(24) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// Address of `scratch` on the current device (looked up once per device).
int scratch_ptr(void** out) {
  int dev = 0;
  TG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(TG_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    TG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(TG_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_diagnostics_noalloc(tg_plan* p, const void* u_hat, const void* b_hat, int with_divmax, double* out_dev,
                           void* stream) {
  TG_RET(check_plan(p));
  if (!u_hat || !out_dev) return fail(TG_ERR_ARG, "tg_diagnostics_noalloc: NULL buffer");
  if (p->generic)
    return fail(TG_ERR_UNSUPPORTED, "tg_diagnostics_noalloc: N=" + std::to_string(p->n) +
                                        " has no compiled plan; use tg_diagnostics");
  return with_table(p, [&](const auto& T, auto* rt) {
    using R = std::remove_pointer_t<decltype(rt)>;
    const size_t need = T.diag_scratch_bytes(p, with_divmax);
    if (need > kScratchBytes)
      return fail(TG_ERR_UNSUPPORTED, "tg_diagnostics_noalloc: N=" + std::to_string(p->n) + " needs " +
                                          std::to_string(need) + " bytes of scratch" +
                                          (with_divmax ? " with with_divmax" : "") + ", more than the " +
                                          std::to_string(kScratchBytes) + " it holds; use tg_diagnostics");
    void* sp = nullptr;
    TG_RET(scratch_ptr(&sp));
    return T.diagnostics_dev(p, static_cast<const cpx<R>*>(u_hat), static_cast<const cpx<R>*>(b_hat), with_divmax,
                             out_dev, sp, static_cast<hipStream_t>(stream));
  });
}

}  // extern "C"

// ---------------------------------------------------------------- checkpoints
namespace {

struct CkptHeader {
  char magic[8];
  int32_t version, n, prec, mhd;
  double L, time, dt, nu, eta, lorentz;
  int64_t step, compact;
  uint64_t checksum;
  char pad[TG_CKPT_HEADER_BYTES - 8 - 4 * 4 - 6 * 8 - 2 * 8 - 8];
};
static_assert(sizeof(CkptHeader) == TG_CKPT_HEADER_BYTES, "checkpoint header is 128 bytes");
constexpr char kCkptMagic[8] = {'T', 'G', 'D', 'N', 'S', 'C', 'K', '1'};

uint64_t fnv1a(const unsigned char* p, size_t n, uint64_t h = 1469598103934665603ull) {
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

int read_header(const char* path, CkptHeader* h, std::vector<unsigned char>* payload) {
  if (!path) return fail(TG_ERR_ARG, "checkpoint path is NULL");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(TG_ERR_ARG, std::string("cannot open checkpoint ") + path);
  const size_t got = fread(h, 1, sizeof(*h), f);
  if (got != sizeof(*h) || memcmp(h->magic, kCkptMagic, 8) != 0 || h->version != 1) {
    fclose(f);
    return fail(TG_ERR_ARG, std::string(path) + " is not a libtgdns checkpoint (bad header)");
  }
  const size_t cb = h->prec == 64 ? 16 : 8;
  const size_t bytes = size_t(h->compact) * 3 * cb * (h->mhd ? 2 : 1);
  std::vector<unsigned char> buf(bytes);
  const size_t pn = bytes ? fread(buf.data(), 1, bytes, f) : 0;
  const bool extra = fgetc(f) != EOF;
  fclose(f);
  if (pn != bytes || extra) return fail(TG_ERR_ARG, std::string(path) + ": payload size does not match the header");
  if (fnv1a(buf.data(), bytes) != h->checksum) return fail(TG_ERR_ARG, std::string(path) + ": checksum mismatch");
  if (payload) payload->swap(buf);
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_checkpoint_write(tg_plan* p, const char* path, const void* u_hat, const void* b_hat,
                        const tg_checkpoint_meta* meta, void* stream) {
  TG_RET(check_plan(p));
  if (!path || !u_hat || !meta) return fail(TG_ERR_ARG, "tg_checkpoint_write: NULL argument");
  if (p->mhd && !b_hat) return fail(TG_ERR_ARG, "tg_checkpoint_write: MHD plan needs b_hat");
  const size_t vec = 3 * p->compact * p->cb;
  std::vector<unsigned char> buf(vec * (p->mhd ? 2 : 1));
  hipStream_t s = static_cast<hipStream_t>(stream);
  TG_HIP(hipMemcpyAsync(buf.data(), u_hat, vec, hipMemcpyDeviceToHost, s));
  if (p->mhd) TG_HIP(hipMemcpyAsync(buf.data() + vec, b_hat, vec, hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  CkptHeader h{};
  memcpy(h.magic, kCkptMagic, 8);
  h.version = 1;
  h.n = p->n;
  h.prec = p->prec;
  h.mhd = p->mhd;
  h.L = p->L;
  h.time = meta->time;
  h.dt = meta->dt;
  h.nu = meta->nu;
  h.eta = meta->eta;
  h.lorentz = meta->lorentz_sign;
  h.step = meta->step;
  h.compact = (int64_t)p->compact;
  h.checksum = fnv1a(buf.data(), buf.size());
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return fail(TG_ERR_ARG, "cannot create " + tmp);
  const bool ok = fwrite(&h, 1, sizeof(h), f) == sizeof(h) && fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  const bool closed = fclose(f) == 0;
  if (!ok || !closed || rename(tmp.c_str(), path) != 0) {
    remove(tmp.c_str());
    return fail(TG_ERR_ARG, std::string("writing checkpoint ") + path + " failed");
  }
  return TG_OK;
}

int tg_checkpoint_info(const char* path, int* n, int* precision_bits, int* mhd, double* domain_length,
                       tg_checkpoint_meta* meta) {
  CkptHeader h;
  TG_RET(read_header(path, &h, nullptr));
  if (n) *n = h.n;
  if (precision_bits) *precision_bits = h.prec;
  if (mhd) *mhd = h.mhd;
  if (domain_length) *domain_length = h.L;
  if (meta) *meta = tg_checkpoint_meta{h.step, h.time, h.dt, h.nu, h.eta, h.lorentz};
  return TG_OK;
}

int tg_checkpoint_read(tg_plan* p, const char* path, void* u_hat, void* b_hat, tg_checkpoint_meta* meta,
                       void* stream) {
  TG_RET(check_plan(p));
  if (!u_hat) return fail(TG_ERR_ARG, "tg_checkpoint_read: NULL u_hat");
  if (p->mhd && !b_hat) return fail(TG_ERR_ARG, "tg_checkpoint_read: MHD plan needs b_hat");
  CkptHeader h;
  std::vector<unsigned char> buf;
  TG_RET(read_header(path, &h, &buf));
  const double L = p->L;
  if (h.n != p->n || h.prec != p->prec || h.mhd != p->mhd || std::fabs(h.L - L) > 1e-12 * L)
    return fail(TG_ERR_ARG, std::string(path) + ": checkpoint of N=" + std::to_string(h.n) + ", " +
                                std::to_string(h.prec) + "-bit, mhd=" + std::to_string(h.mhd) +
                                " does not match the plan");
  const size_t vec = 3 * p->compact * p->cb;
  hipStream_t s = static_cast<hipStream_t>(stream);
  TG_HIP(hipMemcpyAsync(u_hat, buf.data(), vec, hipMemcpyHostToDevice, s));
  if (p->mhd) TG_HIP(hipMemcpyAsync(b_hat, buf.data() + vec, vec, hipMemcpyHostToDevice, s));
  TG_HIP(hipStreamSynchronize(s));
  if (meta) *meta = tg_checkpoint_meta{h.step, h.time, h.dt, h.nu, h.eta, h.lorentz};
  return TG_OK;
}

int tg_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return TG_ERR_ARG;
  snprintf(buf, n, "%s", g_err.c_str());
  return TG_OK;
}

}  // extern "C"
