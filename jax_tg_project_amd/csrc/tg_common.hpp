// Host-side definitions shared by the translation units of libtgdns.so: the
// plan object, error reporting and the launch-checking macros.  The kernels
// of each (precision, grid size) are compiled in their own translation unit
// (tg_inst.hip) and reached through a table of entry points (tg_ops.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/tgdns.h"

// Grid sizes with compiled plans: X(n) for each.  The build compiles
// tg_inst.hip once per (precision, size) of this list.
#define TG_GRID_SIZES(X) X(4) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

namespace tg {

// Record `msg` as the calling thread's last error and return `code`.
int fail(int code, const std::string& msg);

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Checkpoint segments of the adjoint whose stage states the forward sweep
// keeps (the last ones); the reverse sweep recomputes the others.  Each kept
// segment costs 4 seg stage states (and their inverse-y outputs) of
// workspace; at 128^3 FP64 with seg 15 going from two to three to four cut
// the sqrt(n) adjoint from 2.97x to 2.92x to 2.87x the forward time.
constexpr int kAdjKeptSegments = 4;
static_assert(kAdjKeptSegments >= 2, "the reverse sweep alternates two stage buffers");
inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace tg

#define TG_HIP(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ::tg::fail(TG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
  } while (0)

#define TG_LAUNCHED()                                                                     \
  do {                                                                                    \
    hipError_t e_ = hipGetLastError();                                                    \
    if (e_ != hipSuccess)                                                                 \
      return ::tg::fail(TG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

#define TG_RET(...)          \
  do {                       \
    int r_ = (__VA_ARGS__);  \
    if (r_) return r_;       \
  } while (0)

struct tg_plan {
  int n, prec, mhd;
  int M, K, KZ;
  double ks;          // 2 pi / L (exactly 1 for L = 2 pi)
  double L;           // domain length
  void* tw;           // N complex twiddles exp(-2 pi i j / N) in plan precision
  size_t cb;          // bytes per complex element
  size_t compact;     // complex elements per compact component
  // workspace layout (byte offsets); offP, offC only for generic plans
  size_t offA, offB, offU0, offU1, offAcc, offPart, offMax, offOut, wsBytes;
  size_t offP = 0, offC = 0;
  bool generic = false;  // no compiled kernels for n: direct-sum transforms (tg_generic.hip)
  int nDiagBlk, nMax;
  int var[TG_NPASS] = {0, 0, 0, 0, 0};  // tile variant per pass
  // tile variants of the adjoint recompute that runs beside the reverse
  // sweep (rvar[0] < 0: the plan's own var)
  int rvar[TG_NPASS] = {-1, -1, -1, -1, -1};
  // optional per-pass timing: pairs of events around each pass launch
  bool timing = false;
  std::vector<hipEvent_t> evpool;
  std::vector<int> evpass;   // pass id of each recorded pair
  size_t evused = 0;         // pairs recorded since the last read
  // adjoint: private stream (default priority) on which the reverse sweep
  // recomputes the next segment while the current one is reversed
  hipStream_t aux = nullptr;
  // one event per use: adjev[0] orders the sweep's start; segment c's
  // recompute-done and reverse-done events are adjev[1 + 2c], adjev[2 + 2c]
  // (an event is recorded once per sweep, so every wait names the record it
  // was enqueued after, however far the host runs ahead of the GPU)
  std::vector<hipEvent_t> adjev;
  const void* ops = nullptr;  // OpsTable<R> of (prec, n)
};
