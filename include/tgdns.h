/* tgdns.h - C ABI of the MI355X pseudo-spectral Taylor-Green DNS library
 * (libtgdns.so).
 *
 * State layout ("compact spectrum"): for a grid of N^3 points, M = N/3,
 * K = 2M+1, KZ = M+1, each vector component is a C-order array
 * [K][K][KZ] of complex numbers (interleaved re,im; double for precision 64,
 * float for 32) holding exactly the Fourier modes the reference keeps after
 * its inclusive 2/3 cube mask (|k_i| <= N/3), kz >= 0 only (the field is
 * real, so kz < 0 follows by Hermitian symmetry).  Compact index i maps to
 * the signed wavenumber i (i <= M) or i-K (i > M), times 2*pi/L.  A vector
 * field is 3 such components back to back.  Transforms follow numpy/JAX
 * conventions: forward unnormalised, inverse divided by N^3.
 *
 * Physical fields are C-order [3][N][N][N] real arrays (component, x, y, z),
 * the reference's own layout.
 *
 * Conventions: every entry point returns 0 on success and a nonzero TG_ERR_*
 * code otherwise; tg_last_error() then returns the message (thread-local).
 * Every pointer argument except `out`/`info` host arrays is a DEVICE pointer
 * owned by the caller; the library owns only the plan (twiddle tables).
 * `stream` is a hipStream_t (NULL = default stream); work is enqueued on it
 * and the call returns without synchronising, except tg_diagnostics, which
 * copies its scalars to the host and synchronises the stream.
 * A plan is not thread-safe: use one plan per host thread / stream.
 * `ws` is a device workspace of at least tg_plan_query(...)[TG_INFO_WS_BYTES]
 * bytes; its contents need not persist between calls.
 */
#ifndef TGDNS_H
#define TGDNS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TG_OK 0
#define TG_ERR_ARG 1
#define TG_ERR_HIP 2
#define TG_ERR_UNSUPPORTED 3

#define TG_INFO_N 0
#define TG_INFO_K 1
#define TG_INFO_KZ 2
#define TG_INFO_COMPACT_ELEMS 3 /* complex elements per component: K*K*KZ */
#define TG_INFO_WS_BYTES 4
#define TG_INFO_PRECISION 5
#define TG_INFO_MHD 6
#define TG_INFO_GENERIC 7 /* 1: N has no compiled kernels; the plan runs the direct-sum path */
#define TG_INFO_COUNT 8

typedef struct tg_plan tg_plan;

/* Library ABI version (monotone integer).  2: tg_diagnostics writes
 * TG_DIAG_COUNT values, TG_OP_GRAD. */
int tg_version(void);

/* Grid sizes with compiled plans, written to out[0..cap); returns the count.
 * Plans of every other N in [4, 2048] run a direct-sum path (O(N^4) per
 * transform, the reference's advective nonlinear term; no adjoint, no
 * ensembles, no tg_diagnostics_noalloc). */
int tg_supported_sizes(int* out, int cap);

/* Create a plan for an N^3 grid of side `domain_length` (2*pi in every
 * reference configuration), 4 <= N <= 2048 (the reference accepts any N >= 4;
 * sizes outside tg_supported_sizes run the direct-sum path).
 * precision_bits: 64 or 32.  mhd: 0 = hydro (velocity only), 1 = MHD
 * (velocity + magnetic field).
 * Replaces the reference's per-call setup: make_wavenumbers /
 * make_dealias_mask (jax/taylor_green_jax.py:55-67) and
 * SpectralGrid.__init__ (cuda/taylor_green_cuda/grid.py:5-13). */
int tg_plan_create(int n, int precision_bits, int mhd, double domain_length, tg_plan** out);
int tg_plan_destroy(tg_plan* plan);

/* info[TG_INFO_COUNT] receives the TG_INFO_* values. */
int tg_plan_query(const tg_plan* plan, long long* info);

/* Physical [3][N][N][N] real -> compact spectrum (3 components), i.e.
 * mask * fftn(u) restricted to the kept modes.
 * Replaces fftn(u) * dealias_mask (jax/taylor_green_jax.py:160-161) and
 * SpectralGrid.to_spectral (cuda/taylor_green_cuda/grid.py:31-34). */
int tg_forward(tg_plan* plan, const void* phys, void* spec, void* ws, void* stream);

#define TG_OP_FIELD 0 /* the field itself          -> phys [3][N][N][N] */
#define TG_OP_CURL 1  /* i k x field (vorticity)   -> phys [3][N][N][N] */
#define TG_OP_DIV 2   /* i k . field (divergence)  -> phys [N][N][N]    */
#define TG_OP_GRAD 3  /* i k f of ONE scalar component (spec holds a single
                         [K][K][KZ] component) -> phys [3][N][N][N]: the
                         x, y and z derivatives                          */

/* Compact spectrum (3 components; 1 for TG_OP_GRAD) -> physical real
 * field(s): ifftn(...).real of the Hermitian full-cube spectrum of the
 * field, its curl, its divergence or the gradient of a scalar.
 * Replaces ifftn(...).real (jax/taylor_green_jax.py:185),
 * SpectralGrid.to_real (cuda/taylor_green_cuda/grid.py:36-37), vorticity
 * (jax/taylor_green_jax.py:99-109; MHD spectral_curl operators.py:24-25) and
 * spectral_divergence (jax/taylor_green_jax.py:90-96; MHD
 * divergence_from_hat operators.py:50-51) and spectral_gradient
 * (jax/taylor_green_jax.py:82-87). */
int tg_inverse(tg_plan* plan, const void* spec, int op, void* phys, void* ws, void* stream);

/* Right-hand side in the compact spectrum.
 * hydro: r_u = P(u x omega) - nu k^2 u   (== spectral_rhs,
 *        jax/taylor_green_jax.py:158-185, for band-limited u);
 * MHD:   r_u = P(u x omega + lorentz_sign * j x b) - nu k^2 u,
 *        r_b = P(i k x (u x b)) - eta k^2 b
 *        (== MhdSpectralOperator.rhs, cuda/taylor_green_mhd/operators.py:70-107,
 *        with lorentz_sign = -1; +1 is cuda/README_MHD.md:9).
 * b_hat / rb_hat are ignored (may be NULL) for hydro plans. */
int tg_rhs(tg_plan* plan, const void* u_hat, const void* b_hat, double nu, double eta,
           double lorentz_sign, void* ru_hat, void* rb_hat, void* ws, void* stream);

/* nsteps classic RK4 steps in place on the compact state.
 * Replaces advance_one_step (jax/taylor_green_jax.py:188-196),
 * MhdRK4Integrator.step (cuda/taylor_green_mhd/integrator.py:14-32) and the
 * MHD shim advance_one_step (cuda/taylor_green_mhd.py:49-59). */
int tg_rk4_steps(tg_plan* plan, void* u_hat, void* b_hat, int nsteps, double dt, double nu,
                 double eta, double lorentz_sign, void* ws, void* stream);

/* Diagnostics of the compact state, FP64 compensated sums (also for FP32
 * plans).  out[0..TG_DIAG_COUNT):
 *  0 E_k = 1/2 <|u|^2>                 (kinetic_energy, jax/...:199-201)
 *  1 Z   = 1/2 <|curl u|^2>            (enstrophy, jax/...:112-115)
 *  2 <|grad u|^2>  (x nu = viscous_dissipation, jax/...:118-128)
 *  3 E_m = 1/2 <|b|^2>                 (cuda/taylor_green_mhd/diagnostics.py:23)
 *  4 H_c = <u . b>                     (diagnostics.py:25-29)
 *  5 <|div u|^2>   6 <|div b|^2>   7 <|curl b|^2>
 *  8 max |div u| on the grid, 9 max |div b| (only if with_divmax, else 0)
 * 10 number of stored real values of u_hat (and b_hat) that are NaN or
 *    infinite: nonzero once a run has diverged (the reference propagates
 *    NaN silently; the run loops of the Python layer raise on it)
 * Synchronises `stream`. */
#define TG_DIAG_COUNT 11
int tg_diagnostics(tg_plan* plan, const void* u_hat, const void* b_hat, int with_divmax,
                   double* out, void* ws, void* stream);

/* tg_diagnostics for callers that cannot allocate (no counterpart in the
 * reference): code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Same sums and out[] layout as tg_diagnostics, but
 * no workspace argument: every intermediate lives in a 24 MiB device array
 * inside the library (one per GPU, kept until the process ends), and the
 * TG_DIAG_COUNT results are written to the DEVICE array out_dev on
 * `stream` without a host copy or synchronisation.  Nothing needs clearing
 * between calls; calls that overlap on one GPU (other streams, other host
 * threads) share that array and race, so serialise them.  with_divmax needs
 * two single-field intermediates: N <= 128 in FP64 and FP32; larger grids
 * fail with TG_ERR_UNSUPPORTED (without with_divmax any grid works). */
int tg_diagnostics_noalloc(tg_plan* plan, const void* u_hat, const void* b_hat, int with_divmax, double* out_dev,
                           void* stream);

/* Per-pass timing hook (no counterpart in the reference: measurement only).
 * While enabled, every pass launch of tg_rhs / tg_rk4_steps is bracketed by
 * two HIP events recorded on the launch stream; tg_timing_read waits for the
 * last one and returns, per pass id, the summed elapsed milliseconds and the
 * number of launches since the previous read (or enable). */
#define TG_PASS_INV_X 0 /* inverse x + derived fields (u, curl u [, b, curl b]) */
#define TG_PASS_INV_Y 1 /* inverse y                                            */
#define TG_PASS_Z 2     /* z: C2R, nonlinear product, R2C                       */
#define TG_PASS_FWD_Y 3 /* forward y                                            */
#define TG_PASS_FWD_X 4 /* forward x + projection + viscous + RK4 stage         */
#define TG_NPASS 5
int tg_timing_enable(tg_plan* plan, int on);

/* Tile variants (no counterpart in the reference: tuning).  Each pass of
 * tg_rhs / tg_rk4_steps has `count` compiled tile shapes (3 for N >= 128,
 * else 1).  If `set` is non-NULL, set[TG_NPASS] selects the variant per
 * pass (TG_PASS_* order); if `get` is non-NULL it receives the current
 * selection.  Results do not depend on the variant beyond round-off order of
 * the same butterflies; the autotuner (jax_tg_project_amd/autotune.py)
 * times them and stores the fastest per (device, N, precision, equations). */
int tg_plan_variants(tg_plan* plan, const int* set, int* get, int* count);

/* Tile variants of the adjoint recompute (no counterpart in the reference:
 * tuning).  The reverse sweep of tg_adjoint_backward recomputes each earlier
 * checkpoint segment on a second stream while it reverses the current one;
 * those RHS passes may use their own variants, picked for running beside the
 * reverse sweep's kernels rather than alone.  set[0] = -1 returns them to
 * the plan's own (tg_plan_variants); `get` receives the effective set.  The
 * recomputed stages then differ from a recompute with the plan's variants
 * by round-off only. */
int tg_plan_recompute_variants(tg_plan* plan, const int* set, int* get);
int tg_timing_read(tg_plan* plan, double* ms, long long* count);

/* ---- Batched ensemble (no counterpart in the reference, which advances one
 * instance; each instance is the advance_one_step of
 * jax/taylor_green_jax.py:188-196 on its own grid).  Instances may have
 * different N (16, 32, 48, 64, 96, 128) and viscosities; one launch per pass
 * and grid size advances all instances of that size.  Instance i's compact state (3 components, plan
 * layout for its N) lives at complex-element offset state_offsets[i] of one
 * caller-owned state buffer; the workspace is one buffer of ws_bytes. */
typedef struct tg_batch tg_batch;
int tg_batch_create(int ninst, const int* n, const double* nu, int precision_bits, double domain_length,
                    tg_batch** out);
int tg_batch_destroy(tg_batch* batch);
/* state_offsets[ninst + 1] (complex elements) and the workspace size. */
int tg_batch_query(const tg_batch* batch, long long* state_offsets, long long* ws_bytes);
/* Record the state and workspace device pointers used by later steps. */
int tg_batch_bind(tg_batch* batch, void* state, void* ws);
/* Launch layout of a step.  mixed != 0 (the default when the batch holds
 * any of N = 32, 48, 64, 96): one launch per pass covers every instance of
 * those sizes, whose batched bodies share block shapes; mixed == 0: one
 * launch per pass and size class.  Other sizes always launch per class. */
int tg_batch_set_launch(tg_batch* batch, int mixed);
/* nsteps classic RK4 steps of every instance with time step dt.  With
 * use_graph != 0 one step (4 stages x 5 batched launches per launch unit:
 * the mixed group, or a size class) is captured once into a HIP graph
 * (again when dt or the layout changes), with one parallel branch per unit,
 * and replayed nsteps times. */
int tg_batch_rk4_steps(tg_batch* batch, int nsteps, double dt, int use_graph, void* stream);

/* ---- Adjoint (no counterpart in the reference, whose only gradient would be
 * jax.grad through advance_one_step, jax/taylor_green_jax.py:188-196).
 * Cotangents are unnormalised spectra of physical gradients under the
 * physical L2 inner product <f, g> = sum over grid points of f . g.
 * tg_vjp: out = J_rhs(x)^T a = mask F(w x q) + i k x mask F(q x u) - nu k^2 a,
 *   q = F^-1(P a), u = F^-1 x, w = curl u (hydro plans). */
int tg_vjp(tg_plan* plan, const void* x_hat, const void* a_hat, double nu, void* out_hat, void* ws, void* stream);
/* Bytes of the adjoint workspace adj_ws that tg_adjoint_forward and
 * tg_adjoint_backward need for (nsteps, seg). */
int tg_adjoint_workspace(tg_plan* plan, int nsteps, int seg, long long* bytes);
/* Forward sweep: J = E(T) (objective 0) or Z(T) (1) after nsteps RK4 steps of
 * size dt from x0_hat with viscosity nu, keeping a checkpoint every `seg`
 * steps (seg = ceil(sqrt(nsteps)) is the sqrt-n schedule) in adj_ws, and
 * the stage states of the last four segments (so the reverse sweep never
 * recomputes them; seg = nsteps keeps every stage and recomputes nothing).
 * Synchronises `stream` (J is returned to the host). */
int tg_adjoint_forward(tg_plan* plan, const void* x0_hat, int nsteps, double dt, double nu, int objective, int seg,
                       double* J, void* ws, void* adj_ws, void* stream);
/* Reverse sweep after tg_adjoint_forward with the same nsteps/dt/nu/seg and
 * adj_ws: each earlier segment's steps are recomputed from its checkpoint
 * once, keeping their stage states, then reversed with four vector-Jacobian
 * products per step.  grad_hat receives dJ/d(stored entries of x0_hat) as
 * d/dRe + i d/dIm (the convention of torch's complex .grad); *dJdnu the
 * viscosity derivative.  Synchronises `stream`. */
int tg_adjoint_backward(tg_plan* plan, int nsteps, double dt, double nu, int seg, void* grad_hat, double* dJdnu,
                        void* ws, void* adj_ws, void* stream);

/* ---- State dump and restart (no counterpart in the reference, which keeps
 * no state between runs; its default run, jax/taylor_green_jax.py:327-333,
 * is 12.5M steps).  A checkpoint file holds a 128-byte little-endian header
 *   "TGDNSCK1" | int32 version=1, n, precision_bits, mhd |
 *   double domain_length, time, dt, nu, eta, lorentz_sign |
 *   int64 step, compact_elems (K*K*KZ per component) | uint64 FNV-1a-64 of the payload | zero pad
 * followed by the payload: u_hat (3 compact components, plan precision),
 * then b_hat for MHD plans.  tg_checkpoint_write copies the device state to
 * the host, writes "<path>.tmp" and renames it over <path>, so a crash never
 * leaves a torn file; it synchronises `stream`. */
#define TG_CKPT_HEADER_BYTES 128
typedef struct tg_checkpoint_meta {
  long long step;
  double time, dt, nu, eta, lorentz_sign;
} tg_checkpoint_meta;
int tg_checkpoint_write(tg_plan* plan, const char* path, const void* u_hat, const void* b_hat,
                        const tg_checkpoint_meta* meta, void* stream);
/* Header of a checkpoint file (host only, no GPU needed): grid size,
 * precision, equations, domain length and the metadata.  Fails with
 * TG_ERR_ARG on a missing, short, foreign or corrupt (checksum) file. */
int tg_checkpoint_info(const char* path, int* n, int* precision_bits, int* mhd, double* domain_length,
                       tg_checkpoint_meta* meta);
/* Load a checkpoint into u_hat (and b_hat) of a plan with the same n,
 * precision, equations and domain length; synchronises `stream`. */
int tg_checkpoint_read(tg_plan* plan, const char* path, void* u_hat, void* b_hat, tg_checkpoint_meta* meta,
                       void* stream);

/* Copy the calling thread's last error message into buf (NUL-terminated). */
int tg_last_error(char* buf, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* TGDNS_H */
