// taylor_green_hip: the native command-line Taylor-Green solver on the
// gfx950 library (libtgdns.so, C ABI of include/tgdns.h).
//
//   taylor_green_hip [N] [dt] [Re] [total_time] [save_time] [options]
//
// Same positional arguments, defaults, validation, step count
// (int)(total/dt), save cadence max(1, (int)(save/dt)) and save line as the
// reference's standalone solver (cuda/taylor_green_cuda.cu:117-125); nu =
// 2 pi / Re.  It differs where that solver is wrong or lossy: the Leray
// projection has the right sign (the pruned kernels project n - k (k.n)/k^2)
// and the saved enstrophy and energy are FP64 Parseval sums of the spectral
// state instead of a host finite difference.
//
// Options (after the positional arguments):
//   --precision 32|64   kernel precision (default 32, the reference's float)
//   --checkpoint PATH   write the state to PATH at every save (atomically)
//   --restart PATH      continue from a checkpoint written by a run with the
//                       same N, precision and dt
// Exit status: 0 done, 1 bad arguments, 2 library / HIP error, 3 the state
// became non-finite.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tgdns.h"

namespace {

constexpr float kPiF = 3.14159265358979323846f;

void die_lib(const char* what) {
  char msg[512];
  tg_last_error(msg, sizeof(msg));
  fprintf(stderr, "taylor_green_hip: %s: %s\n", what, msg);
  exit(2);
}

void die_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, "taylor_green_hip: %s: %s\n", what, hipGetErrorString(e));
    exit(2);
  }
}

int usage(const char* prog) {
  fprintf(stderr, "Usage: %s [N] [dt] [Re] [total_time] [save_time] [--precision 32|64] [--checkpoint PATH] "
                  "[--restart PATH]\n", prog);
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<const char*> pos;
  int prec = 32;
  std::string ckpt, restart;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--precision" && i + 1 < argc) {
      prec = atoi(argv[++i]);
    } else if (a == "--checkpoint" && i + 1 < argc) {
      ckpt = argv[++i];
    } else if (a == "--restart" && i + 1 < argc) {
      restart = argv[++i];
    } else if (a.size() > 1 && a[0] == '-' && a[1] == '-') {
      return usage(argv[0]);
    } else {
      pos.push_back(argv[i]);
    }
  }
  const int n = pos.size() > 0 ? atoi(pos[0]) : 128;
  const float dt = pos.size() > 1 ? float(atof(pos[1])) : .005f;
  const float re = pos.size() > 2 ? float(atof(pos[2])) : 1000.f;
  const float total = pos.size() > 3 ? float(atof(pos[3])) : 1.f;
  const float save = pos.size() > 4 ? float(atof(pos[4])) : .05f;
  if (n < 4 || dt <= 0 || re <= 0 || total <= 0 || save <= 0 || pos.size() > 5 || (prec != 32 && prec != 64))
    return usage(argv[0]);
  const double nu = prec == 32 ? double(2.0f * kPiF / re) : 2.0 * M_PI / double(re);
  const int steps = int(total / dt), every = std::max(1, int(save / dt));
  printf("HIP Taylor-Green: N=%d Re=%.1f steps=%d\n", n, re, steps);
  fflush(stdout);

  tg_plan* plan = nullptr;
  if (tg_plan_create(n, prec, 0, 2.0 * M_PI, &plan)) die_lib("plan");
  long long info[TG_INFO_COUNT];
  if (tg_plan_query(plan, info)) die_lib("plan query");
  const size_t rb = prec == 64 ? 8 : 4, cb = 2 * rb;
  const size_t n3 = size_t(n) * n * n;
  void *state = nullptr, *ws = nullptr;
  die_hip(hipMalloc(&state, 3 * size_t(info[TG_INFO_COMPACT_ELEMS]) * cb), "hipMalloc(state)");
  die_hip(hipMalloc(&ws, size_t(info[TG_INFO_WS_BYTES])), "hipMalloc(workspace)");

  int start = 0;
  if (!restart.empty()) {
    int cn = 0, cp = 0, cm = 0;
    double cl = 0.0;
    tg_checkpoint_meta meta{};
    if (tg_checkpoint_info(restart.c_str(), &cn, &cp, &cm, &cl, &meta)) die_lib("restart");
    if (cn != n || cp != prec || cm != 0 || meta.dt != double(dt)) {
      fprintf(stderr, "taylor_green_hip: %s holds N=%d, %d-bit, dt=%g; this run is N=%d, %d-bit, dt=%g\n",
              restart.c_str(), cn, cp, meta.dt, n, prec, double(dt));
      return 1;
    }
    if (tg_checkpoint_read(plan, restart.c_str(), state, nullptr, &meta, nullptr)) die_lib("restart");
    start = int(meta.step);
  } else {
    // u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 on x_i = 2 pi i / N
    std::vector<double> h(3 * n3, 0.0);
    for (int x = 0; x < n; ++x)
      for (int y = 0; y < n; ++y)
        for (int z = 0; z < n; ++z) {
          const double X = 2.0 * M_PI * x / n, Y = 2.0 * M_PI * y / n, Z = 2.0 * M_PI * z / n;
          const size_t q = (size_t(x) * n + y) * n + z;
          h[q] = std::sin(X) * std::cos(Y) * std::cos(Z);
          h[n3 + q] = -std::cos(X) * std::sin(Y) * std::cos(Z);
        }
    void* phys = nullptr;
    die_hip(hipMalloc(&phys, 3 * n3 * rb), "hipMalloc(initial field)");
    if (prec == 64) {
      die_hip(hipMemcpy(phys, h.data(), 3 * n3 * rb, hipMemcpyHostToDevice), "copy initial field");
    } else {
      std::vector<float> hf(h.begin(), h.end());
      die_hip(hipMemcpy(phys, hf.data(), 3 * n3 * rb, hipMemcpyHostToDevice), "copy initial field");
    }
    if (tg_forward(plan, phys, state, ws, nullptr)) die_lib("forward transform");
    die_hip(hipDeviceSynchronize(), "initial transform");
    die_hip(hipFree(phys), "hipFree");
  }

  int step = start;
  double d[TG_DIAG_COUNT];
  while (step < steps) {
    const int next = std::min(steps, (step / every + 1) * every);
    if (tg_rk4_steps(plan, state, nullptr, next - step, double(dt), nu, 0.0, 0.0, ws, nullptr)) die_lib("rk4");
    step = next;
    if (tg_diagnostics(plan, state, nullptr, 0, d, ws, nullptr)) die_lib("diagnostics");
    const float t = step * dt;
    printf("Save time step: %d, time: %.6f, enstrophy: %.8e, kinetic energy: %.8e\n", step, t, d[1], d[0]);
    fflush(stdout);
    if (d[10] > 0.0) {
      fprintf(stderr, "taylor_green_hip: %d non-finite values in the state at step %d: the run diverged\n",
              int(d[10]), step);
      return 3;
    }
    if (!ckpt.empty()) {
      const tg_checkpoint_meta meta{step, double(t), double(dt), nu, 0.0, 0.0};
      if (tg_checkpoint_write(plan, ckpt.c_str(), state, nullptr, &meta, nullptr)) die_lib("checkpoint");
    }
  }
  die_hip(hipDeviceSynchronize(), "final synchronisation");
  die_hip(hipFree(state), "hipFree");
  die_hip(hipFree(ws), "hipFree");
  tg_plan_destroy(plan);
  return 0;
}
