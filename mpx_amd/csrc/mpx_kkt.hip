// C ABI of the primal-dual iLQR KKT hot path for gfx950 (declared in
// include/mpx_kkt.h).  Host side: argument checks, workspace carving and the
// launch sequence K1 (linearise) -> K2/K3 (LQR) -> K4 (line search + update).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mpx_kkt.h"
#include "common.h"
#include "k_linearize.h"
#include "k_lqr.h"
#include "k_lqr_wave.h"
#include "k_inst.h"
#include "k_scan.h"
#include "k_linesearch.h"
#include "k_trials.h"
#include "k_trials_mma.h"
#include "k_lin_mma.h"
#include "k_fwd_tanh.h"
#include "k_costate.h"
#include "k_adjoint.h"
#include "k_misc.h"
#include "k_caller.h"
#include "k_srbd.h"

using namespace mpxk;

namespace {

thread_local std::string g_err;

mpxkkt_status fail(mpxkkt_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

size_t dtype_size(int dt) { return dt == MPXKKT_F64 ? 8 : 4; }

int pow2_threads(int want, int lo, int hi) {
  int t = lo;
  while (t < want && t < hi) t <<= 1;
  return t;
}

template <typename T>
View<T> view(const mpxkkt_view& v) {
  View<T> r;
  r.p = static_cast<const T*>(v.ptr);
  r.sb = v.batch_stride;
  r.st = v.stage_stride;
  return r;
}

template <typename T>
Prob<T> make_prob(const mpxkkt_problem* pr) {
  Prob<T> P;
  P.B = pr->batch; P.n = pr->n; P.m = pr->m; P.N = pr->horizon; P.p = pr->n_eq;
  P.box = pr->has_box ? 1 : 0; P.model = pr->model;
  P.h = (T)pr->h; P.kappa = (T)pr->kappa;
  P.A = view<T>(pr->A); P.Bm = view<T>(pr->B); P.W = view<T>(pr->W);
  P.Q = view<T>(pr->Q); P.R = view<T>(pr->R); P.QN = view<T>(pr->QN);
  P.xref = view<T>(pr->xref); P.uref = view<T>(pr->uref);
  P.D = view<T>(pr->D); P.E = view<T>(pr->E); P.e = view<T>(pr->e);
  P.ulo = view<T>(pr->ulo); P.uhi = view<T>(pr->uhi); P.x0 = view<T>(pr->x0);
  P.fbar = view<T>(pr->fbar); P.M = view<T>(pr->M); P.q = view<T>(pr->q); P.r = view<T>(pr->r);
  P.g = view<T>(pr->g);
  return P;
}

int effective_segments(const mpxkkt_problem* pr, const mpxkkt_params* prm) {
  // the caller's linearisation (cross term M) runs the sequential Riccati recursion
  if (pr->model == MPXKKT_MODEL_CALLER) return 1;
  int P = prm->segments < 1 ? 1 : prm->segments;
  const int maxP = (pr->horizon + 1) / 2;  // at least two elements per segment
  if (P > maxP) P = maxP < 1 ? 1 : maxP;
  return P;
}

// The register-resident wave LQR kernels cover FP32 on the matrix cores with
// n + 1 <= 48, m <= 16, n_eq <= 16 and one segment (flags bits 0 and 1 turn
// them off); they never form the closed-loop maps F_t = A_t - B_t K_t.
bool wave_path(const mpxkkt_problem* pr, int flags, int P) {
  return pr->dtype == MPXKKT_F32 && (flags & 3) == 0 && P <= 1 && pr->n + 1 <= 48 && pr->m <= 16 &&
         pr->n_eq <= 16 && pr->model != MPXKKT_MODEL_CALLER;
}

template <typename T>
size_t carve(const mpxkkt_problem* pr, int P, bool wave, Ws<T>* ws, char* base) {
  const size_t B = pr->batch, n = pr->n, m = pr->m, N = pr->horizon, N1 = N + 1;
  size_t off = 0;
  auto take = [&](size_t cnt) -> T* {
    off = (off + 255) & ~(size_t)255;
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += cnt * sizeof(T);
    return p;
  };
  const bool tanh_model = pr->model == MPXKKT_MODEL_TANH;
  ws->Alin = take(tanh_model ? B * N * n * n : 0);
  ws->Blin = take(tanh_model ? B * n * m : 0);
  ws->sdh = take(tanh_model ? B * N * n : 0);
  ws->qh = take(B * N1 * n);
  ws->rh = take(B * N * m);
  ws->rd = take(B * N * m);
  ws->cdef = take(B * N1 * n);
  ws->S = take(B * N1 * n * n);
  ws->s = take(B * N1 * n);
  ws->K = take(B * N * m * n);
  ws->k = take(B * N * m);
  ws->F = take(wave ? 0 : B * N * n * n);
  ws->f = take(wave ? 0 : B * N * n);
  ws->dx = take(B * N1 * n);
  ws->du = take(B * N * m);
  ws->dv = take(B * N1 * n);
  ws->mpart = take(B * N1 * NMP);
  ws->fpart = take(B * N1 * NFP);
  ws->agg = take(P > 1 ? B * P * (3 * n * n + 2 * n) : 0);
  ws->fmap = take(P > 1 ? B * P * (n * n + n) : 0);
  ws->sval = take(P > 1 ? B * P * (n * n + n) : 0);
  ws->xs = take(P > 1 ? B * P * n : 0);
  ws->kpart = take(B * N1);
  ws->lsp = take(B * NLS);
  ws->tpart = take(B * N1 * 2 * MAX_TRIALS);
  ws->hq = take(B * N * 2 * (size_t)pr->n_eq);
  ws->gpart = take(B * N1 * 2);
  ws->last = take(B);
  ws->run = reinterpret_cast<int*>(take(B));
  ws->nrun = reinterpret_cast<int*>(take(1));
  const bool caller = pr->model == MPXKKT_MODEL_CALLER;
  ws->trial_g = take(caller ? B * MAX_TRIALS * N1 : 0);
  ws->trial_f = take(caller ? B * MAX_TRIALS * N * n : 0);
  return off + 256;
}

size_t workspace_bytes(const mpxkkt_problem* pr, const mpxkkt_params* prm) {
  const int P = effective_segments(pr, prm);
  const bool wave = wave_path(pr, prm->flags, P);
  if (pr->dtype == MPXKKT_F64) {
    Ws<double> w;
    return carve<double>(pr, P, wave, &w, nullptr);
  }
  Ws<float> w;
  return carve<float>(pr, P, wave, &w, nullptr);
}

mpxkkt_status check_view(const mpxkkt_view& v, const char* name, bool needed) {
  if (needed && v.ptr == nullptr) return fail(MPXKKT_ERR_ARG, std::string("null view: ") + name);
  if (v.batch_stride < 0 || v.stage_stride < 0)
    return fail(MPXKKT_ERR_ARG, std::string("negative stride: ") + name);
  return MPXKKT_OK;
}

mpxkkt_status check_problem(const mpxkkt_problem* pr) {
  if (!pr) return fail(MPXKKT_ERR_ARG, "null problem");
  if (pr->dtype != MPXKKT_F32 && pr->dtype != MPXKKT_F64) return fail(MPXKKT_ERR_ARG, "bad dtype");
  if (pr->model != MPXKKT_MODEL_LTV && pr->model != MPXKKT_MODEL_TANH && pr->model != MPXKKT_MODEL_CALLER)
    return fail(MPXKKT_ERR_ARG, "bad model");
  if (pr->batch < 1 || pr->batch > 65535) return fail(MPXKKT_ERR_ARG, "batch must be in [1, 65535]");
  if (pr->n < 1 || pr->m < 1 || pr->horizon < 1 || pr->n_eq < 0)
    return fail(MPXKKT_ERR_ARG, "dimensions must be positive");
  if (pr->n > 512 || pr->m > 64 || pr->n_eq > 512 || pr->horizon > 100000)
    return fail(MPXKKT_ERR_UNSUPPORTED, "dimension too large (n <= 512, m <= 64)");
  mpxkkt_status s;
  const bool eq = pr->n_eq > 0, box = pr->has_box != 0;
  if ((s = check_view(pr->A, "A", true)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->B, "B", true)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->W, "W", pr->model == MPXKKT_MODEL_TANH)) != MPXKKT_OK) return s;
  const bool caller = pr->model == MPXKKT_MODEL_CALLER;
  if ((s = check_view(pr->Q, "Q", true)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->R, "R", true)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->QN, "QN", true)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->xref, "xref", !caller)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->uref, "uref", !caller)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->fbar, "fbar", caller)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->M, "M", false)) != MPXKKT_OK) return s;
  if (!caller && (pr->M.ptr || pr->fbar.ptr || pr->q.ptr || pr->r.ptr || pr->g.ptr))
    return fail(MPXKKT_ERR_ARG, "fbar, M, q, r, g are views of MPXKKT_MODEL_CALLER only");
  if ((s = check_view(pr->q, "q", caller)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->r, "r", caller)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->g, "g", caller)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->D, "D", eq)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->E, "E", eq)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->e, "e", eq)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->ulo, "ulo", box)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->uhi, "uhi", box)) != MPXKKT_OK) return s;
  if ((s = check_view(pr->x0, "x0", true)) != MPXKKT_OK) return s;
  return MPXKKT_OK;
}

mpxkkt_status check_iterate(const mpxkkt_problem* pr, const mpxkkt_iterate* it) {
  if (!it || !it->X || !it->U || !it->V) return fail(MPXKKT_ERR_ARG, "null iterate array");
  if (pr->n_eq > 0 && !it->Y_eq) return fail(MPXKKT_ERR_ARG, "Y_eq required when n_eq > 0");
  if (pr->has_box && !it->Y_box) return fail(MPXKKT_ERR_ARG, "Y_box required when has_box");
  return MPXKKT_OK;
}

mpxkkt_status check_params(const mpxkkt_params* prm) {
  if (prm->max_trials < 1 || prm->max_trials > MAX_TRIALS)
    return fail(MPXKKT_ERR_ARG, "max_trials must be in [1, 16]");
  if (!(prm->beta > 0.0 && prm->beta < 1.0) || !(prm->tau > 0.0 && prm->tau < 1.0) || !(prm->mu_b > 0.0) ||
      !(prm->rho_c >= 0.0) || !(prm->hess_floor >= 0.0 && prm->hess_floor < 1.0))
    return fail(MPXKKT_ERR_ARG, "line-search / barrier parameters out of range");
  return MPXKKT_OK;
}

// LDS bytes per block of each device, queried once per device (any thread).
constexpr int kMaxDevices = 64;
std::mutex g_lds_mu;
int g_max_lds[kMaxDevices] = {0};

mpxkkt_status lds_limit(int* out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(MPXKKT_ERR_DEVICE, "hipGetDevice failed");
  std::lock_guard<std::mutex> lk(g_lds_mu);
  if (g_max_lds[dev] <= 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      return fail(MPXKKT_ERR_DEVICE, "cannot query LDS size");
    g_max_lds[dev] = v;
  }
  *out = g_max_lds[dev];
  return MPXKKT_OK;
}

// Workspace of mpxkkt_sqp_iterate_noalloc: a static device array, one copy per
// device, held while the process runs.  On the host the array's name is only a
// stand-in for its device copy: host code uses the address hipGetSymbolAddress
// returns, taken once per device with that device current.
static __device__ __align__(256) unsigned char scratch[MPXKKT_NOALLOC_WORKSPACE_BYTES];
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {nullptr};

mpxkkt_status scratch_address(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(MPXKKT_ERR_DEVICE, "hipGetDevice failed");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(MPXKKT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return MPXKKT_OK;
}

template <typename K>
mpxkkt_status prep_lds(K kernel, size_t bytes) {
  int lim = 0;
  mpxkkt_status s = lds_limit(&lim);
  if (s != MPXKKT_OK) return s;
  if (bytes > (size_t)lim)
    return fail(MPXKKT_ERR_UNSUPPORTED, "shape needs " + std::to_string(bytes) +
                                            " bytes of LDS per block, device has " + std::to_string(lim));
  if (bytes > 65536) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
      return fail(MPXKKT_ERR_DEVICE, "hipFuncSetAttribute failed");
  }
  return MPXKKT_OK;
}

mpxkkt_status launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MPXKKT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return MPXKKT_OK;
}

template <typename T>
Prm<T> make_prm(const mpxkkt_params* prm, int P) {
  Prm<T> r;
  r.rho0 = (T)prm->rho0; r.rho_max = (T)prm->rho_max; r.rho_c = (T)prm->rho_c;
  r.mu_b = (T)prm->mu_b; r.tau = (T)prm->tau; r.beta = (T)prm->beta; r.c1 = (T)prm->c1;
  r.slack = (T)prm->armijo_slack;
  r.hfloor = (T)prm->hess_floor;
  r.max_trials = std::max(1, std::min(prm->max_trials, MAX_TRIALS));
  int g = prm->trial_group <= 0 ? r.max_trials : prm->trial_group;
  if (prm->limited_memory) g = 1;
  r.group = std::max(1, std::min(g, r.max_trials));
  r.P = P;
  r.flags = prm->flags;
  return r;
}

template <typename T>
Iter<T> make_iter(const mpxkkt_iterate* it) {
  Iter<T> r;
  r.X = static_cast<T*>(it->X); r.U = static_cast<T*>(it->U); r.V = static_cast<T*>(it->V);
  r.Yeq = static_cast<T*>(it->Y_eq); r.Ybox = static_cast<T*>(it->Y_box);
  return r;
}

template <typename T>
LqrIn<T> make_lqr_in(const mpxkkt_problem* pr, const Ws<T>& ws) {
  LqrIn<T> in;
  if (pr->model == MPXKKT_MODEL_TANH) {
    const long long n = pr->n, m = pr->m, N = pr->horizon;
    in.A.p = ws.Alin; in.A.sb = N * n * n; in.A.st = n * n;
    in.Bm.p = ws.Blin; in.Bm.sb = n * m; in.Bm.st = 0;
  } else {
    in.A = view<T>(pr->A);
    in.Bm = view<T>(pr->B);
  }
  return in;
}

// The instance-per-block kernels (k_inst.h) take n, m, n_eq <= 64 when their
// LDS plan fits the device.
template <typename T>
bool use_inst(const mpxkkt_problem* pr, bool with_v) {
  if (pr->n > 64 || pr->m > 64 || pr->n_eq > 64) return false;
  const InstLds L = inst_lds(pr->n, pr->m, pr->horizon, pr->model == MPXKKT_MODEL_TANH, pr->Q.stage_stride == 0,
                             pr->R.stage_stride == 0, with_v);
  int lim = 0;
  if (lds_limit(&lim) != MPXKKT_OK) return false;
  return (size_t)L.total * sizeof(T) <= (size_t)lim;
}

int k1_threads(int n) { return n <= 64 ? 64 : 128; }

template <typename T, int MMAX>
mpxkkt_status run_scan(hipStream_t st, const mpxkkt_problem* pr, int nt, const Prob<T>& P, const Prm<T>& R,
                       const Iter<T>& I, const Ws<T>& W, const LqrIn<T>& in) {
  const int n = pr->n, m = pr->m, p = pr->n_eq, B = pr->batch, Pn = R.P;
  mpxkkt_status s;
  const size_t ric = ric_smem_elems<T>(n, m, p, nt) * sizeof(T);
  const size_t up = std::max(up_smem_elems<T>(n, m, p, nt) * sizeof(T), ric);
  if ((s = prep_lds(k_scan_up<T, MMAX>, up)) != MPXKKT_OK) return s;
  hipLaunchKernelGGL((k_scan_up<T, MMAX>), dim3(Pn, B), dim3(nt), up, st, P, R, I, W, in);
  if ((s = launch_check("k_scan_up")) != MPXKKT_OK) return s;
  const size_t mid = mid_smem_elems<T>(n, nt) * sizeof(T) + 64;
  if ((s = prep_lds(k_scan_mid<T>, mid)) != MPXKKT_OK) return s;
  hipLaunchKernelGGL(k_scan_mid<T>, dim3(B), dim3(nt), mid, st, P, R, W);
  if ((s = launch_check("k_scan_mid")) != MPXKKT_OK) return s;
  if ((s = prep_lds(k_scan_down<T, MMAX>, ric)) != MPXKKT_OK) return s;
  hipLaunchKernelGGL((k_scan_down<T, MMAX>), dim3(Pn - 1, B), dim3(nt), ric, st, P, R, I, W, in);
  if ((s = launch_check("k_scan_down")) != MPXKKT_OK) return s;
  const size_t fu = fwd_smem_elems<T>(n, m, nt) * sizeof(T);
  if ((s = prep_lds(k_fwd_up<T>, fu)) != MPXKKT_OK) return s;
  hipLaunchKernelGGL(k_fwd_up<T>, dim3(Pn - 1, B), dim3(nt), fu, st, P, R, I, W, in);
  if ((s = launch_check("k_fwd_up")) != MPXKKT_OK) return s;
  hipLaunchKernelGGL(k_fwd_mid<T>, dim3(B), dim3(64), 2 * n * sizeof(T), st, P, R, W);
  if ((s = launch_check("k_fwd_mid")) != MPXKKT_OK) return s;
  const size_t fd = ((size_t)2 * n * n + 2 * n + 8) * sizeof(T);
  if ((s = prep_lds(k_fwd_down<T>, fd)) != MPXKKT_OK) return s;
  hipLaunchKernelGGL(k_fwd_down<T>, dim3(Pn - 1, B), dim3(nt), fd, st, P, R, I, W, in);
  return launch_check("k_fwd_down");
}

template <typename T>
bool use_wave_lqr(const mpxkkt_problem* pr, const Prm<T>& R);

// k_fwd_wave for a stage-invariant B (cb) or not, with 128-bit tile loads when
// n % 4 == 0 (wide)
template <int NT>
void fwd_wave(hipStream_t st, dim3 gb, bool cb, bool wide, const Prob<float>& P, const Ws<float>& W,
              const LqrIn<float>& in) {
  if (cb && wide) hipLaunchKernelGGL((k_fwd_wave<NT, true, true>), gb, dim3(64), 0, st, P, W, in);
  else if (cb) hipLaunchKernelGGL((k_fwd_wave<NT, true, false>), gb, dim3(64), 0, st, P, W, in);
  else if (wide) hipLaunchKernelGGL((k_fwd_wave<NT, false, true>), gb, dim3(64), 0, st, P, W, in);
  else hipLaunchKernelGGL((k_fwd_wave<NT, false, false>), gb, dim3(64), 0, st, P, W, in);
}

// Rows of a view of row length L start 16-byte aligned (null views pass).
bool rows16(const mpxkkt_view& v, long long L) {
  return v.ptr == nullptr || ((reinterpret_cast<uintptr_t>(v.ptr) & 15) == 0 && v.batch_stride % 4 == 0 &&
                              v.stage_stride % 4 == 0 && L % 4 == 0);
}

template <typename T>
bool trial_rows16(const mpxkkt_problem* pr) {
  if constexpr (!IsF32<T>::value) {
    return false;
  } else {
    const long long n = pr->n, m = pr->m;
    return rows16(pr->A, n) && rows16(pr->W, n) && rows16(pr->B, m) && rows16(pr->Q, n) && rows16(pr->R, m) &&
           rows16(pr->QN, n) && (pr->n_eq == 0 || (rows16(pr->D, n) && rows16(pr->E, m)));
  }
}

// The matrix-core forms of the linearisation and of the trials
// (k_lin_mma.h, k_trials_mma.h): FP32, the tanh model with stage-invariant A,
// W, B, Q, R, at most three state tiles, one control tile and one tile of
// equality rows, and flags bit 4 clear (bit 4 keeps the row-per-lane forms,
// for comparison).  Returns the number of state tiles, 0 when they do not apply.
template <typename T>
int mma_tiles(const mpxkkt_problem* pr, int flags) {
  if (!IsF32<T>::value || pr->model != MPXKKT_MODEL_TANH || (flags & 16) != 0) return 0;
  if (pr->A.stage_stride != 0 || pr->W.stage_stride != 0 || pr->B.stage_stride != 0 || pr->Q.stage_stride != 0 ||
      pr->R.stage_stride != 0)
    return 0;
  if (pr->n > 48 || pr->m > 16 || pr->n_eq > 16) return 0;
  return (pr->n + 15) / 16;
}

// 128-bit loads and stores in the matrix-core forms: every row of the
// problem's views they read 16-byte aligned (n, m multiples of 4)
template <typename T>
bool mma_rows16(const mpxkkt_problem* pr) {
  return trial_rows16<T>(pr) && rows16(pr->xref, pr->n) && rows16(pr->uref, pr->m) &&
         (!pr->has_box || (rows16(pr->ulo, pr->m) && rows16(pr->uhi, pr->m)));
}

inline bool al16(const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; }

// The tanh model on the wave path after k_linearize_mma (flags bit 3 clear:
// it ran and wrote sd_t): the LQR kernels form f_x = I + h (A + kappa
// diag(sd_t) W) where they use it and the linearisation does not write it.
template <typename T>
bool fx_in_lqr(const mpxkkt_problem* pr, const Prm<T>& R) {
  return use_wave_lqr<T>(pr, R) && (R.flags & 8) == 0 && mma_tiles<T>(pr, R.flags) > 0;
}

// costates, control steps and slope partials: on the wave path the forward
// pass of the LQR solve and the costates -- for the tanh model after
// k_linearize_mma k_fwd_tanh and the adjoint sweep k_adjoint_tanh, else
// k_fwd_wave and k_costate_rows (lam = S x + s from the stored S) --, else
// parallel over stages after the dx chain (k_fwd_stage)
template <typename T>
mpxkkt_status run_fwd_stage(hipStream_t st, const mpxkkt_problem* pr, const Prob<T>& P, const Prm<T>& R,
                            const Iter<T>& I, const Ws<T>& W, const LqrIn<T>& in) {
  if constexpr (IsF32<T>::value) {
    if (use_wave_lqr<T>(pr, R)) {
      const int NT = (pr->n + 1 + 15) / 16;
      const dim3 gb(pr->batch);
      const bool cb = in.Bm.st == 0;   // B~ held for the whole chain
      // the tanh model after k_linearize_mma (flags bit 3 clear: it ran and wrote
      // sd_t): the forward chain on the vector ALU without A_t
      const int nf = fx_in_lqr<T>(pr, R) ? mma_tiles<T>(pr, R.flags) : 0;
      if (nf) {
        const bool wide = mma_rows16<T>(pr) && al16(W.K) && al16(W.cdef) && al16(W.dx);
        auto go = [&](auto kern) { hipLaunchKernelGGL(kern, gb, dim3(64), 0, st, P, W); };
        if (wide) {
          if (nf == 1) go(k_fwd_tanh<1, true, 4>);
          else if (nf == 2) go(k_fwd_tanh<2, true, 4>);
          else go(k_fwd_tanh<3, true, 2>);
        } else {
          if (nf == 1) go(k_fwd_tanh<1, false, 4>);
          else if (nf == 2) go(k_fwd_tanh<2, false, 4>);
          else go(k_fwd_tanh<3, false, 2>);
        }
        mpxkkt_status s = launch_check("k_fwd_tanh");
        if (s != MPXKKT_OK) return s;
        // the costates by the adjoint recursion (k_lqr_wave stored no S_t)
        auto goa = [&](auto kern) { hipLaunchKernelGGL(kern, gb, dim3(64), 0, st, P, R, I, W); };
        if (wide) {
          if (nf == 1) goa(k_adjoint_tanh<1, true>);
          else if (nf == 2) goa(k_adjoint_tanh<2, true>);
          else goa(k_adjoint_tanh<3, true>);
        } else {
          if (nf == 1) goa(k_adjoint_tanh<1, false>);
          else if (nf == 2) goa(k_adjoint_tanh<2, false>);
          else goa(k_adjoint_tanh<3, false>);
        }
        return launch_check("k_adjoint_tanh");
      } else {
        if (NT == 1) fwd_wave<1>(st, gb, cb, (pr->n & 3) == 0, P, W, in);
        else if (NT == 2) fwd_wave<2>(st, gb, cb, (pr->n & 3) == 0, P, W, in);
        else fwd_wave<3>(st, gb, cb, (pr->n & 3) == 0, P, W, in);
        mpxkkt_status s = launch_check("k_fwd_wave");
        if (s != MPXKKT_OK) return s;
      }
      // the costates: the stored tiles of S_t without the homogeneous coordinate
      const int nc = (pr->n + 15) / 16;
      const bool cw = (pr->n & 3) == 0 && al16(W.S) && al16(W.dx) && al16(W.qh);
      constexpr int spw = kCostateStages;
      const dim3 gcs((pr->horizon + 1 + spw - 1) / spw, pr->batch);
      auto gocs = [&](auto kern) { hipLaunchKernelGGL(kern, gcs, dim3(64), 0, st, P, I, W); };
      if (cw) {
        if (nc == 1) gocs(k_costate_rows<1, true, spw>);
        else if (nc == 2) gocs(k_costate_rows<2, true, spw>);
        else gocs(k_costate_rows<3, true, spw>);
      } else {
        if (nc == 1) gocs(k_costate_rows<1, false, spw>);
        else if (nc == 2) gocs(k_costate_rows<2, false, spw>);
        else gocs(k_costate_rows<3, false, spw>);
      }
      return launch_check("k_costate_rows");
    }
  }
  const int nt = 64;
  const size_t lds = stage_smem_elems<T>(pr->n, nt) * sizeof(T);
  mpxkkt_status s = prep_lds(k_fwd_stage<T>, lds);
  if (s != MPXKKT_OK) return s;
  hipLaunchKernelGGL(k_fwd_stage<T>, dim3(pr->horizon + 1, pr->batch), dim3(nt), lds, st, P, I, W);
  return launch_check("k_fwd_stage");
}

// K1: with flags bit 3 parallel over stages with shared constants
// (k_linearize_stage: lower latency for small batches) when n, m, n_eq <= 64
// and its LDS plan fits; else the matrix-core form (k_linearize_mma, 16 stages
// per wave, and k_linearize_term) where mma_tiles allows it; else one block
// per instance (k_linearize_inst, higher throughput for large batches) under
// the limits of k_linearize_stage; the general block kernel otherwise.
template <typename T>
mpxkkt_status run_linearize(hipStream_t st, const mpxkkt_problem* pr, const Prob<T>& P, const Prm<T>& R,
                            const Iter<T>& I, const Ws<T>& W) {
  const bool caller = pr->model == MPXKKT_MODEL_CALLER;
  if (!caller && (R.flags & 8) != 0 && pr->n <= 64 && pr->m <= 64 && pr->n_eq <= 64) {
    const InstLds L = inst_lds(pr->n, pr->m, pr->horizon, pr->model == MPXKKT_MODEL_TANH, pr->Q.stage_stride == 0,
                               pr->R.stage_stride == 0, false);
    const size_t lds =
        ((size_t)((L.X + 3) & ~3) + (size_t)kLinStageWaves * lin_stage_wave_elems(pr->n, pr->m)) * sizeof(T);
    int lim = 0;
    if (lds_limit(&lim) == MPXKKT_OK && lds <= (size_t)lim) {
      mpxkkt_status s = prep_lds(k_linearize_stage<T>, lds);
      if (s != MPXKKT_OK) return s;
      const dim3 grid((pr->horizon + kLinStageWaves) / kLinStageWaves, pr->batch);
      hipLaunchKernelGGL(k_linearize_stage<T>, grid, dim3(64 * kLinStageWaves), lds, st, P, R, I, W);
      return launch_check("k_linearize_stage");
    }
  }
  if constexpr (IsF32<T>::value) {
    const int nt = caller ? 0 : mma_tiles<T>(pr, R.flags);
    if (nt) {
      // 128-bit loads and stores when the iterate and the outputs are 16-byte aligned too
      const bool wide = mma_rows16<T>(pr) && al16(I.X) && al16(I.U) && al16(I.V) &&
                        (!pr->has_box || al16(I.Ybox)) && al16(W.qh) && al16(W.cdef) && al16(W.rh) &&
                        al16(W.rd) && al16(W.Alin) && al16(W.sdh);
      const dim3 grid((pr->horizon + 15) / 16, pr->batch);
      // f_x only for the LQR kernels that read it
      const int wfx = fx_in_lqr<T>(pr, R) ? 0 : 1;
      auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(64), 0, st, P, R, I, W, wfx); };
      if (wide) {
        if (nt == 1) go(k_linearize_mma<1, true>);
        else if (nt == 2) go(k_linearize_mma<2, true>);
        else go(k_linearize_mma<3, true>);
      } else {
        if (nt == 1) go(k_linearize_mma<1, false>);
        else if (nt == 2) go(k_linearize_mma<2, false>);
        else go(k_linearize_mma<3, false>);
      }
      mpxkkt_status s = launch_check("k_linearize_mma");
      if (s != MPXKKT_OK) return s;
      auto got = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(pr->batch), dim3(64), 0, st, P, I, W); };
      if (wide) {
        if (nt == 1) got(k_linearize_term<1, true>);
        else if (nt == 2) got(k_linearize_term<2, true>);
        else got(k_linearize_term<3, true>);
      } else {
        if (nt == 1) got(k_linearize_term<1, false>);
        else if (nt == 2) got(k_linearize_term<2, false>);
        else got(k_linearize_term<3, false>);
      }
      return launch_check("k_linearize_term");
    }
  }
  if (!caller && use_inst<T>(pr, false)) {
    const InstLds L = inst_lds(pr->n, pr->m, pr->horizon, pr->model == MPXKKT_MODEL_TANH, pr->Q.stage_stride == 0,
                               pr->R.stage_stride == 0, false);
    const size_t lds = (size_t)L.total * sizeof(T);
    mpxkkt_status s = prep_lds(k_linearize_inst<T>, lds);
    if (s != MPXKKT_OK) return s;
    hipLaunchKernelGGL(k_linearize_inst<T>, dim3(pr->batch), dim3(64 * kInstWaves), lds, st, P, R, I, W);
    return launch_check("k_linearize_inst");
  }
  const int nt = k1_threads(pr->n);
  const size_t lds = ((size_t)2 * pr->n + pr->m + pr->n_eq + NMP * nt) * sizeof(T);
  mpxkkt_status s = prep_lds(k_linearize<T>, lds);
  if (s != MPXKKT_OK) return s;
  dim3 grid(pr->horizon + 1, pr->batch);
  hipLaunchKernelGGL(k_linearize<T>, grid, dim3(nt), lds, st, P, R, I, W);
  return launch_check("k_linearize");
}

template <typename T>
bool use_wave_lqr(const mpxkkt_problem* pr, const Prm<T>& R) {
  return IsF32<T>::value && wave_path(pr, R.flags, R.P);
}

template <typename T>
mpxkkt_status run_lqr_wave(hipStream_t st, const mpxkkt_problem* pr, const Prob<T>& P, const Prm<T>& R,
                           const Iter<T>& I, const Ws<T>& W, const LqrIn<T>& in) {
  if constexpr (IsF32<T>::value) {
    const int NT = (pr->n + 1 + 15) / 16;
    // k steps of the last state tile: the n + 1 - 16 (NT - 1) offsets it holds, by 4
    const int ksn = (pr->n + 1 - 16 * (NT - 1) + 3) / 4;
    const bool cq = P.Q.st == 0 && P.R.st == 0 && in.Bm.st == 0;   // Q, B, R held for the whole recursion
    const dim3 gb(pr->batch);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, gb, dim3(64), 0, st, P, R, I, W, in); };
    if (cq) {
      // one instantiation per step count: the steps are unrolled matrix instructions
      auto cases = [&](auto FA, auto AS) {
        constexpr bool fa = decltype(FA)::value, as = decltype(AS)::value;
        switch (NT * 4 + ksn) {
          case 5: go(k_lqr_wave<1, true, 1, fa, as>); break;
          case 6: go(k_lqr_wave<1, true, 2, fa, as>); break;
          case 7: go(k_lqr_wave<1, true, 3, fa, as>); break;
          case 8: go(k_lqr_wave<1, true, 4, fa, as>); break;
          case 9: go(k_lqr_wave<2, true, 1, fa, as>); break;
          case 10: go(k_lqr_wave<2, true, 2, fa, as>); break;
          case 11: go(k_lqr_wave<2, true, 3, fa, as>); break;
          case 12: go(k_lqr_wave<2, true, 4, fa, as>); break;
          case 13: go(k_lqr_wave<3, true, 1, fa, as>); break;
          case 14: go(k_lqr_wave<3, true, 2, fa, as>); break;
          case 15: go(k_lqr_wave<3, true, 3, fa, as>); break;
          default: go(k_lqr_wave<3, true, 4, fa, as>); break;
        }
      };
      // n % 4 == 0: the homogeneous offset n alone in the last k step of the
      // state (split off the contractions); the tanh model after
      // k_linearize_mma: A~ formed from A, W and sd_t
      const bool as = pr->n % 4 == 0 && !(R.flags & 16);
      const bool fa = fx_in_lqr<T>(pr, R);
      if (fa && as) cases(std::true_type{}, std::true_type{});
      else if (fa) cases(std::true_type{}, std::false_type{});
      else if (as) cases(std::false_type{}, std::true_type{});
      else cases(std::false_type{}, std::false_type{});
    } else {   // stage-varying data: the generic step count (the steps past n + 1 add zeros)
      if (NT == 1) go(k_lqr_wave<1, false, 4, false, false>);
      else if (NT == 2) go(k_lqr_wave<2, false, 4, false, false>);
      else go(k_lqr_wave<3, false, 4, false, false>);
    }
    return launch_check("k_lqr_wave");
  } else {
    return fail(MPXKKT_ERR_UNSUPPORTED, "wave LQR kernel is FP32 only");
  }
}

template <typename T, int MMAX>
mpxkkt_status run_lqr_m(hipStream_t st, const mpxkkt_problem* pr, int nt, const Prob<T>& P, const Prm<T>& R,
                        const Iter<T>& I, const Ws<T>& W, const LqrIn<T>& in) {
  mpxkkt_status s;
  if (R.P > 1) {
    if ((s = run_scan<T, MMAX>(st, pr, nt, P, R, I, W, in)) != MPXKKT_OK) return s;
  } else if (use_wave_lqr<T>(pr, R)) {
    return run_lqr_wave(st, pr, P, R, I, W, in);
  } else {
    const size_t lds = ric_smem_elems<T>(pr->n, pr->m, pr->n_eq, nt, pr->M.ptr != nullptr) * sizeof(T);
    if ((s = prep_lds(k_lqr_seq<T, MMAX>, lds)) != MPXKKT_OK) return s;
    hipLaunchKernelGGL((k_lqr_seq<T, MMAX>), dim3(pr->batch), dim3(nt), lds, st, P, R, I, W, in);
    if ((s = launch_check("k_lqr_seq")) != MPXKKT_OK) return s;
  }
  return MPXKKT_OK;
}

template <typename T>
mpxkkt_status run_lqr(hipStream_t st, const mpxkkt_problem* pr, const mpxkkt_params* prm, const Prob<T>& P,
                      const Prm<T>& R, const Iter<T>& I, const Ws<T>& W, const LqrIn<T>& in) {
  const int nt = prm->block_threads > 0 ? pow2_threads(prm->block_threads, 64, 512) : 256;
  if (pr->m <= 16) return run_lqr_m<T, 16>(st, pr, nt, P, R, I, W, in);
  if (pr->m <= 32) return run_lqr_m<T, 32>(st, pr, nt, P, R, I, W, in);
  return run_lqr_m<T, 64>(st, pr, nt, P, R, I, W, in);
}

// Launch shape of k_trials: with `sh` (tanh model, stage-invariant Q and R,
// the shared constants fit) 8 waves per block share the instance's matrices
// in LDS, else 4 waves each read their rows from memory.
struct TrialShape {
  bool sh, v4;
  int mma;   // > 0: k_trials_mma<mma> for the stages < N (its state tiles)
  bool mma_wide;   // the problem's rows of k_trials_mma 16-byte aligned (n, m multiples of 4)
  bool hq;         // the equality rows and their slopes precomputed (the tanh model on the wave path)
  int waves;
  size_t lds;
  dim3 grid;
};

template <typename T>
TrialShape trial_shape(const mpxkkt_problem* pr, int flags) {
  TrialShape ts;
  const int N1 = pr->horizon + 1;
  ts.v4 = trial_rows16<T>(pr);
  int lim = 0;
  const size_t shared = (size_t)trial_shared_elems(pr->n, pr->m) + 8 * (size_t)trial_wave_elems(pr->n, pr->m);
  ts.sh = pr->model == MPXKKT_MODEL_TANH && pr->Q.stage_stride == 0 && pr->R.stage_stride == 0 &&
          lds_limit(&lim) == MPXKKT_OK && shared * sizeof(T) <= (size_t)lim;
  ts.waves = ts.sh ? 8 : kTrialWaves;
  ts.mma = ts.sh ? mma_tiles<T>(pr, flags) : 0;
  ts.mma_wide = ts.mma && mma_rows16<T>(pr);
  ts.hq = false;
  ts.lds = (ts.sh ? shared : (size_t)kTrialWaves * trial_wave_elems(pr->n, pr->m)) * sizeof(T);
  ts.grid = dim3((N1 + ts.waves - 1) / ts.waves, pr->batch);
  return ts;
}

template <typename T, int RPL, bool V4, bool SH>
mpxkkt_status launch_trials_k(hipStream_t st, const TrialShape& ts, const Prob<T>& P, const Prm<T>& R,
                              const Iter<T>& I, const Ws<T>& W, int j0, int j1) {
  mpxkkt_status s;
  if constexpr (IsF32<T>::value) {
    if (ts.mma) {
      // stages 0 .. N - 1 on the matrix cores, the terminal stage by one wave of k_trials
      const int N = P.N;
      const dim3 grid((N + 63) / 64, P.B);
      // 128-bit loads when the iterate and the step are 16-byte aligned too
      const bool wide = ts.mma_wide && al16(I.X) && al16(I.U) && al16(I.V) && al16(W.dx) && al16(W.du) &&
                        al16(W.dv);
      auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, P, R, I, W, j0, j1); };
      auto forms = [&](auto HQ) {
        constexpr bool hq = decltype(HQ)::value;
        if (wide) {
          if (ts.mma == 1) go(k_trials_mma<1, true, hq>);
          else if (ts.mma == 2) go(k_trials_mma<2, true, hq>);
          else go(k_trials_mma<3, true, hq>);
        } else {
          if (ts.mma == 1) go(k_trials_mma<1, false, hq>);
          else if (ts.mma == 2) go(k_trials_mma<2, false, hq>);
          else go(k_trials_mma<3, false, hq>);
        }
      };
      // after k_linearize_mma and k_adjoint_tanh the equality rows are in hq
      if (ts.hq) forms(std::true_type{});
      else forms(std::false_type{});
      if ((s = launch_check("k_trials_mma")) != MPXKKT_OK) return s;
      const size_t lds = (size_t)trial_wave_elems(P.n, P.m) * sizeof(T);
      if ((s = prep_lds(k_trials<T, RPL, V4, false>, lds)) != MPXKKT_OK) return s;
      hipLaunchKernelGGL((k_trials<T, RPL, V4, false>), dim3(1, P.B), dim3(64), lds, st, P, R, I, W, j0, j1, N);
      return launch_check("k_trials");
    }
  }
  if ((s = prep_lds(k_trials<T, RPL, V4, SH>, ts.lds)) != MPXKKT_OK) return s;
  hipLaunchKernelGGL((k_trials<T, RPL, V4, SH>), ts.grid, dim3(64 * ts.waves), ts.lds, st, P, R, I, W, j0, j1, 0);
  return launch_check("k_trials");
}

template <typename T, int RPL>
mpxkkt_status launch_trials(hipStream_t st, const TrialShape& ts, const Prob<T>& P, const Prm<T>& R,
                            const Iter<T>& I, const Ws<T>& W, int j0, int j1) {
  if (ts.sh) return launch_trials_k<T, RPL, false, true>(st, ts, P, R, I, W, j0, j1);
  if constexpr (IsF32<T>::value) {
    if (ts.v4) return launch_trials_k<T, RPL, true, false>(st, ts, P, R, I, W, j0, j1);
  }
  return launch_trials_k<T, RPL, false, false>(st, ts, P, R, I, W, j0, j1);
}

template <typename T>
mpxkkt_status run_trials(hipStream_t st, const mpxkkt_problem* pr, const TrialShape& ts, const Prob<T>& P,
                         const Prm<T>& R, const Iter<T>& I, const Ws<T>& W, int j0, int j1) {
  const int rows = std::max(pr->n, pr->n_eq);
  if (rows <= 64) return launch_trials<T, 1>(st, ts, P, R, I, W, j0, j1);
  if (rows <= 128) return launch_trials<T, 2>(st, ts, P, R, I, W, j0, j1);
  if (rows <= 256) return launch_trials<T, 4>(st, ts, P, R, I, W, j0, j1);
  return launch_trials<T, 8>(st, ts, P, R, I, W, j0, j1);
}

// The acceptance: k_ls_decide (the step, once per instance), then k_ls_update
// (the iterate and multipliers, one thread per entry).
template <typename T>
mpxkkt_status ls_accept(hipStream_t st, const mpxkkt_problem* pr, const Prob<T>& P, const Prm<T>& R,
                        const Iter<T>& I, const Ws<T>& W, const mpxkkt_stats* so) {
  const size_t B = pr->batch, n = pr->n, m = pr->m, p = pr->n_eq, N = pr->horizon, N1 = N + 1;
  T* a = so ? static_cast<T*>(so->alpha) : nullptr;
  T* r = so ? static_cast<T*>(so->rho) : nullptr;
  T* f0 = so ? static_cast<T*>(so->merit0) : nullptr;
  T* df = so ? static_cast<T*>(so->dmerit) : nullptr;
  int* stt = so ? so->status : nullptr;
  int* tr = so ? so->trials : nullptr;
  hipLaunchKernelGGL(k_ls_decide<T>, dim3((unsigned)B), dim3(64), 0, st, R, W, (int)N1, a, r, f0, df, stt, tr);
  mpxkkt_status s = launch_check("k_ls_decide");
  if (s != MPXKKT_OK) return s;
  const int nbx = (int)((B * N1 * n + 255) / 256), nbu = (int)((B * N * m + 255) / 256);
  const int nbp = (int)((B * N * p + 255) / 256);
  hipLaunchKernelGGL(k_ls_update<T>, dim3(nbx + nbu + nbp), dim3(256), 0, st, P, R, I, W, nbx, nbu);
  return launch_check("k_ls_update");
}

// K4: k_ls_setup, k_trials over [0, group) and then [group, max_trials) for the
// instances that accepted none of the first group, k_ls_decide and k_ls_update.  flags bit 2
// selects the single-kernel line search of k_inst.h / k_linesearch.h instead.
template <typename T>
mpxkkt_status run_linesearch(hipStream_t st, const mpxkkt_problem* pr, const Prob<T>& P, const Prm<T>& R,
                             const Iter<T>& I, const Ws<T>& W, const LqrIn<T>& in, const mpxkkt_stats* so) {
  if ((R.flags & 4) == 0) {
    mpxkkt_status s;
    const int B = pr->batch, N1 = pr->horizon + 1;
    const size_t setup_lds = (size_t)6 * 256 * sizeof(T);
    hipLaunchKernelGGL(k_ls_setup<T>, dim3(B), dim3(256), setup_lds, st, P, R, I, W);
    if ((s = launch_check("k_ls_setup")) != MPXKKT_OK) return s;
    TrialShape ts = trial_shape<T>(pr, R.flags);
    ts.hq = ts.mma && fx_in_lqr<T>(pr, R);
    if ((s = run_trials<T>(st, pr, ts, P, R, I, W, 0, R.group)) != MPXKKT_OK) return s;
    if (R.group < R.max_trials && (s = run_trials<T>(st, pr, ts, P, R, I, W, R.group, R.max_trials)) != MPXKKT_OK)
      return s;
    return ls_accept<T>(st, pr, P, R, I, W, so);
  }
  if (use_inst<T>(pr, true)) {
    const InstLds L = inst_lds(pr->n, pr->m, pr->horizon, pr->model == MPXKKT_MODEL_TANH, pr->Q.stage_stride == 0,
                               pr->R.stage_stride == 0, true);
    const size_t lds = (size_t)L.total * sizeof(T);
    mpxkkt_status s = prep_lds(k_linesearch_inst<T>, lds);
    if (s != MPXKKT_OK) return s;
    T* a = so ? static_cast<T*>(so->alpha) : nullptr;
    T* r = so ? static_cast<T*>(so->rho) : nullptr;
    T* f0 = so ? static_cast<T*>(so->merit0) : nullptr;
    T* df = so ? static_cast<T*>(so->dmerit) : nullptr;
    int* stt = so ? so->status : nullptr;
    int* tr = so ? so->trials : nullptr;
    hipLaunchKernelGGL(k_linesearch_inst<T>, dim3(pr->batch), dim3(64 * kInstWaves), lds, st, P, R, I, W, a, r, f0,
                       df, stt, tr);
    return launch_check("k_linesearch_inst");
  }
  const int nt = 256;
  const size_t lds = (size_t)MAX_TRIALS * nt * sizeof(T);
  mpxkkt_status s = prep_lds(k_linesearch<T>, lds);
  if (s != MPXKKT_OK) return s;
  T* a = so ? static_cast<T*>(so->alpha) : nullptr;
  T* r = so ? static_cast<T*>(so->rho) : nullptr;
  T* f0 = so ? static_cast<T*>(so->merit0) : nullptr;
  T* df = so ? static_cast<T*>(so->dmerit) : nullptr;
  int* stt = so ? so->status : nullptr;
  int* tr = so ? so->trials : nullptr;
  hipLaunchKernelGGL(k_linesearch<T>, dim3(pr->batch), dim3(nt), lds, st, P, R, I, W, in, a, r, f0, df, stt, tr);
  return launch_check("k_linesearch");
}

// phase-timing record (host-side; events on the caller's stream), shared by
// the threads of the process under g_prof_mu
struct PhaseEvents {
  hipEvent_t ev[MPXKKT_PHASES + 1];
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<PhaseEvents> g_prof;

void prof_mark(hipStream_t st, const PhaseEvents* pe, int k) {
  if (pe) (void)hipEventRecord(pe->ev[k], st);
}

// gated: instances whose run flag (workspace) is 0 skip every kernel -- the
// multi-iteration solve's converged instances; plain iterations run them all.
template <typename T>
mpxkkt_status sqp_iterate_t(hipStream_t st, const mpxkkt_problem* pr, const mpxkkt_params* prm,
                            mpxkkt_iterate* it, mpxkkt_stats* so, void* wsp, bool profile_ok = true,
                            bool gated = false) {
  const int Pseg = effective_segments(pr, prm);
  Ws<T> W;
  carve<T>(pr, Pseg, wave_path(pr, prm->flags, Pseg), &W, static_cast<char*>(wsp));
  if (!gated) W.run = nullptr;
  const Prob<T> P = make_prob<T>(pr);
  const Prm<T> R = make_prm<T>(prm, Pseg);
  const Iter<T> I = make_iter<T>(it);
  const LqrIn<T> in = make_lqr_in<T>(pr, W);
  mpxkkt_status s;
  PhaseEvents ev_store;
  const PhaseEvents* pe = nullptr;
  bool prof = false;
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    prof = g_prof_on && profile_ok;
  }
  if (prof) {
    for (int k = 0; k <= MPXKKT_PHASES; ++k)
      if (hipEventCreate(&ev_store.ev[k]) != hipSuccess) return fail(MPXKKT_ERR_DEVICE, "hipEventCreate failed");
    pe = &ev_store;
  }
  prof_mark(st, pe, 0);
  if ((s = run_linearize<T>(st, pr, P, R, I, W)) != MPXKKT_OK) return s;
  prof_mark(st, pe, 1);
  if ((s = run_lqr<T>(st, pr, prm, P, R, I, W, in)) != MPXKKT_OK) return s;
  prof_mark(st, pe, 2);
  if ((s = run_fwd_stage<T>(st, pr, P, R, I, W, in)) != MPXKKT_OK) return s;
  prof_mark(st, pe, 3);
  s = run_linesearch<T>(st, pr, P, R, I, W, in, so);
  prof_mark(st, pe, 4);
  if (prof) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.push_back(ev_store);
  }
  return s;
}

template <typename T>
mpxkkt_status kkt_t(hipStream_t st, const mpxkkt_problem* pr, const mpxkkt_params* prm, const mpxkkt_iterate* it,
                    void* out, void* wsp) {
  Ws<T> W;
  const int Pseg = effective_segments(pr, prm);
  carve<T>(pr, Pseg, wave_path(pr, prm->flags, Pseg), &W, static_cast<char*>(wsp));
  const Prob<T> P = make_prob<T>(pr);
  const Prm<T> R = make_prm<T>(prm, 1);
  const Iter<T> I = make_iter<T>(it);
  const int nt = k1_threads(pr->n);
  const size_t lds = ((size_t)3 * pr->n + pr->m + nt) * sizeof(T);
  mpxkkt_status s = prep_lds(k_kkt_stage<T>, lds);
  if (s != MPXKKT_OK) return s;
  hipLaunchKernelGGL(k_kkt_stage<T>, dim3(pr->horizon + 1, pr->batch), dim3(nt), lds, st, P, R, I, W);
  if ((s = launch_check("k_kkt_stage")) != MPXKKT_OK) return s;
  hipLaunchKernelGGL(k_kkt_reduce<T>, dim3((pr->batch + 255) / 256), dim3(256), 0, st, pr->batch,
                     pr->horizon + 1, (const T*)W.kpart, static_cast<T*>(out));
  return launch_check("k_kkt_reduce");
}

// Caller-dynamics iteration, first half: K1-K3 and the line-search set-up on
// the caller's linearisation, then the step to the caller.
template <typename T>
mpxkkt_status direction_t(hipStream_t st, const mpxkkt_problem* pr, const mpxkkt_params* prm,
                          mpxkkt_iterate* it, mpxkkt_direction* dir, void* wsp) {
  Ws<T> W;
  carve<T>(pr, 1, false, &W, static_cast<char*>(wsp));
  W.run = nullptr;
  const Prob<T> P = make_prob<T>(pr);
  const Prm<T> R = make_prm<T>(prm, 1);
  const Iter<T> I = make_iter<T>(it);
  const LqrIn<T> in = make_lqr_in<T>(pr, W);
  mpxkkt_status s;
  if ((s = run_linearize<T>(st, pr, P, R, I, W)) != MPXKKT_OK) return s;
  if ((s = run_lqr<T>(st, pr, prm, P, R, I, W, in)) != MPXKKT_OK) return s;
  if ((s = run_fwd_stage<T>(st, pr, P, R, I, W, in)) != MPXKKT_OK) return s;
  hipLaunchKernelGGL(k_ls_setup<T>, dim3(pr->batch), dim3(256), (size_t)6 * 256 * sizeof(T), st, P, R, I, W);
  if ((s = launch_check("k_ls_setup")) != MPXKKT_OK) return s;
  if (!dir->dX) return MPXKKT_OK;
  hipLaunchKernelGGL(k_export_dir<T>, dim3(pr->batch), dim3(256), 0, st, P, W, static_cast<T*>(dir->dX),
                     static_cast<T*>(dir->dU), static_cast<T*>(dir->alpha_max));
  return launch_check("k_export_dir");
}

// Second half: the trials' merits from the caller's evaluations, Armijo
// choice, multiplier and iterate update.
template <typename T>
mpxkkt_status accept_t(hipStream_t st, const mpxkkt_problem* pr, const mpxkkt_params* prm, mpxkkt_iterate* it,
                       int L, const void* tg, const void* tf, mpxkkt_stats* so, void* wsp) {
  Ws<T> W;
  carve<T>(pr, 1, false, &W, static_cast<char*>(wsp));
  W.run = nullptr;
  const Prob<T> P = make_prob<T>(pr);
  Prm<T> R = make_prm<T>(prm, 1);
  R.max_trials = L;
  R.group = L;
  const Iter<T> I = make_iter<T>(it);
  const int B = pr->batch, N1 = pr->horizon + 1;
  const dim3 grid((N1 + kTrialWaves - 1) / kTrialWaves, B);
  hipLaunchKernelGGL(k_trials_caller<T>, grid, dim3(64 * kTrialWaves), 0, st, P, R, I, W, L,
                     static_cast<const T*>(tg), static_cast<const T*>(tf));
  mpxkkt_status s = launch_check("k_trials_caller");
  if (s != MPXKKT_OK) return s;
  return ls_accept<T>(st, pr, P, R, I, W, so);
}

template <typename T>
Srbd<T> make_srbd(const mpxkkt_srbd* m) {
  Srbd<T> S;
  S.mass = (T)m->mass;
  S.dt = (T)m->dt;
  for (int k = 0; k < 9; ++k) { S.I[k] = (T)m->inertia[k]; S.Iinv[k] = (T)m->inertia_inv[k]; }
  S.W = view<T>(m->W);
  S.ref = view<T>(m->reference);
  S.par = view<T>(m->parameter);
  return S;
}

// One SQP iteration of the SRBD quadruped entirely on the device: its
// linearisation into the problem's caller views, the direction, the trial
// evaluations and the acceptance.
template <typename T>
mpxkkt_status srbd_iterate_t(hipStream_t st, const mpxkkt_problem* pr, const mpxkkt_srbd* model,
                             const mpxkkt_params* prm, mpxkkt_iterate* it, mpxkkt_stats* so, void* wsp) {
  Ws<T> W;
  carve<T>(pr, 1, false, &W, static_cast<char*>(wsp));
  W.run = nullptr;
  const Prob<T> P = make_prob<T>(pr);
  const Srbd<T> S = make_srbd<T>(model);
  const Iter<T> I = make_iter<T>(it);
  const int B = pr->batch, N1 = pr->horizon + 1;
  const int L = std::max(1, std::min(prm->max_trials, MAX_TRIALS));
  hipLaunchKernelGGL(k_srbd_linearize<T>, dim3((B * N1 + 63) / 64), dim3(64), 0, st, P, S, I);
  mpxkkt_status s = launch_check("k_srbd_linearize");
  if (s != MPXKKT_OK) return s;
  mpxkkt_direction dir;
  dir.dX = nullptr;
  // the step stays in the workspace (direction_t exports only when asked)
  if ((s = direction_t<T>(st, pr, prm, it, &dir, wsp)) != MPXKKT_OK) return s;
  hipLaunchKernelGGL(k_srbd_trials<T>, dim3((B * L * N1 + 63) / 64), dim3(64), 0, st, P, S, I, W,
                     (T)prm->beta, L);
  if ((s = launch_check("k_srbd_trials")) != MPXKKT_OK) return s;
  return accept_t<T>(st, pr, prm, it, L, W.trial_g, W.trial_f, so, wsp);
}

// SQP iterations until runOffline's stop rule (mpx/utils/mpc_wrapper.py:425-467)
// holds for every instance: after each iteration the per-stage (g_t, |c|^2)
// partials and the per-instance stop test; converged instances skip the
// kernels of later iterations.
template <typename T>
mpxkkt_status solve_t(hipStream_t st, const mpxkkt_problem* pr, const mpxkkt_params* prm,
                      const mpxkkt_solve_opts* o, mpxkkt_iterate* it, mpxkkt_stats* so, int32_t* iters,
                      void* cost, void* cc, void* wsp, int32_t* launched) {
  Ws<T> W;
  const int Pseg = effective_segments(pr, prm);
  carve<T>(pr, Pseg, wave_path(pr, prm->flags, Pseg), &W, static_cast<char*>(wsp));
  const Prob<T> P = make_prob<T>(pr);
  const Iter<T> I = make_iter<T>(it);
  const int B = pr->batch, N1 = pr->horizon + 1;
  const dim3 gB((B + 255) / 256);
  hipLaunchKernelGGL(k_solve_init<T>, gB, dim3(256), 0, st, B, W, iters);
  mpxkkt_status s = launch_check("k_solve_init");
  if (s != MPXKKT_OK) return s;
  const int nt = 64;
  const int total = o->max_iter + 1;
  int done = 0;
  for (int i = 0; i < total; ++i) {
    if ((s = sqp_iterate_t<T>(st, pr, prm, it, so, wsp, true, true)) != MPXKKT_OK) return s;
    hipLaunchKernelGGL(k_eval_partials<T>, dim3(N1, B), dim3(nt), 2 * nt * sizeof(T), st, P, I, W);
    if ((s = launch_check("k_eval_partials")) != MPXKKT_OK) return s;
    hipLaunchKernelGGL(k_stop<T>, gB, dim3(256), 0, st, B, N1, W, o->max_iter, (T)o->dcost_tol, (T)o->cc_tol,
                       iters, static_cast<T*>(cost), static_cast<T*>(cc));
    if ((s = launch_check("k_stop")) != MPXKKT_OK) return s;
    done = i + 1;
    if (o->check_every > 0 && done % o->check_every == 0 && done < total) {
      hipLaunchKernelGGL(k_count_running<T>, dim3(1), dim3(256), 0, st, B, W);
      if ((s = launch_check("k_count_running")) != MPXKKT_OK) return s;
      int running = 0;
      if (hipMemcpyAsync(&running, W.nrun, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return fail(MPXKKT_ERR_DEVICE, "reading the running-instance count failed");
      if (running == 0) break;
    }
  }
  if (launched) *launched = done;
  return MPXKKT_OK;
}

}  // namespace

extern "C" {

mpxkkt_status mpxkkt_default_params(mpxkkt_params* out) {
  if (!out) return fail(MPXKKT_ERR_ARG, "null params");
  std::memset(out, 0, sizeof(*out));
  out->rho0 = 1.0;
  out->rho_max = 1e8;
  out->rho_c = 10.0;
  out->mu_b = 1e-2;
  out->tau = 0.995;
  out->beta = 0.5;
  out->c1 = 1e-4;
  out->armijo_slack = 1e-6;
  out->max_trials = 10;
  out->limited_memory = 0;
  out->segments = 1;
  out->block_threads = 0;
  out->trial_group = 0;
  out->hess_floor = 1e-8;
  return MPXKKT_OK;
}

mpxkkt_status mpxkkt_workspace_size(const mpxkkt_problem* prob, const mpxkkt_params* prm, size_t* bytes) {
  mpxkkt_status s = check_problem(prob);
  if (s != MPXKKT_OK) return s;
  if (!prm || !bytes) return fail(MPXKKT_ERR_ARG, "null params or output");
  *bytes = workspace_bytes(prob, prm);
  return MPXKKT_OK;
}

mpxkkt_status mpxkkt_sqp_iterate(void* stream, const mpxkkt_problem* prob, const mpxkkt_params* prm,
                                 mpxkkt_iterate* it, mpxkkt_stats* stats, void* workspace,
                                 size_t workspace_bytes_) {
  mpxkkt_status s = check_problem(prob);
  if (s != MPXKKT_OK) return s;
  if (!prm) return fail(MPXKKT_ERR_ARG, "null params");
  if (prob->model == MPXKKT_MODEL_CALLER)
    return fail(MPXKKT_ERR_ARG, "caller dynamics: use mpxkkt_sqp_direction and mpxkkt_sqp_accept");
  if ((s = check_iterate(prob, it)) != MPXKKT_OK) return s;
  if ((s = check_params(prm)) != MPXKKT_OK) return s;
  const size_t need = workspace_bytes(prob, prm);
  if (!workspace || workspace_bytes_ < need)
    return fail(MPXKKT_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (prob->dtype == MPXKKT_F64) return sqp_iterate_t<double>(st, prob, prm, it, stats, workspace);
  return sqp_iterate_t<float>(st, prob, prm, it, stats, workspace);
}

mpxkkt_status mpxkkt_sqp_iterate_noalloc(void* stream, const mpxkkt_problem* prob, const mpxkkt_params* prm,
                                         mpxkkt_iterate* it, mpxkkt_stats* stats) {
  mpxkkt_status s = check_problem(prob);
  if (s != MPXKKT_OK) return s;
  if (!prm) return fail(MPXKKT_ERR_ARG, "null params");
  if (prob->model == MPXKKT_MODEL_CALLER)
    return fail(MPXKKT_ERR_ARG, "caller dynamics: use mpxkkt_sqp_direction and mpxkkt_sqp_accept");
  if ((s = check_iterate(prob, it)) != MPXKKT_OK) return s;
  if ((s = check_params(prm)) != MPXKKT_OK) return s;
  const size_t need = workspace_bytes(prob, prm);
  if (need > (size_t)MPXKKT_NOALLOC_WORKSPACE_BYTES)
    return fail(MPXKKT_ERR_WORKSPACE, "this problem needs " + std::to_string(need) +
                                          " bytes of workspace; the static workspace of "
                                          "mpxkkt_sqp_iterate_noalloc holds " +
                                          std::to_string((size_t)MPXKKT_NOALLOC_WORKSPACE_BYTES) +
                                          " (use mpxkkt_sqp_iterate with a caller workspace)");
  void* ws = nullptr;
  if ((s = scratch_address(&ws)) != MPXKKT_OK) return s;
  hipStream_t st = static_cast<hipStream_t>(stream);
  // no phase events here: creating them is a runtime allocation
  if (prob->dtype == MPXKKT_F64) return sqp_iterate_t<double>(st, prob, prm, it, stats, ws, false);
  return sqp_iterate_t<float>(st, prob, prm, it, stats, ws, false);
}

mpxkkt_status mpxkkt_sqp_direction(void* stream, const mpxkkt_problem* prob, const mpxkkt_params* prm,
                                   mpxkkt_iterate* it, mpxkkt_direction* dir, void* workspace,
                                   size_t workspace_bytes_) {
  mpxkkt_status s = check_problem(prob);
  if (s != MPXKKT_OK) return s;
  if (prob->model != MPXKKT_MODEL_CALLER)
    return fail(MPXKKT_ERR_ARG, "mpxkkt_sqp_direction takes MPXKKT_MODEL_CALLER problems");
  if (!prm || !dir || !dir->dX || !dir->dU || !dir->alpha_max)
    return fail(MPXKKT_ERR_ARG, "null params or direction output");
  if ((s = check_iterate(prob, it)) != MPXKKT_OK) return s;
  if ((s = check_params(prm)) != MPXKKT_OK) return s;
  const size_t need = workspace_bytes(prob, prm);
  if (!workspace || workspace_bytes_ < need)
    return fail(MPXKKT_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (prob->dtype == MPXKKT_F64) return direction_t<double>(st, prob, prm, it, dir, workspace);
  return direction_t<float>(st, prob, prm, it, dir, workspace);
}

mpxkkt_status mpxkkt_sqp_accept(void* stream, const mpxkkt_problem* prob, const mpxkkt_params* prm,
                                mpxkkt_iterate* it, int32_t n_trials, const void* trial_g, const void* trial_f,
                                mpxkkt_stats* stats, void* workspace, size_t workspace_bytes_) {
  mpxkkt_status s = check_problem(prob);
  if (s != MPXKKT_OK) return s;
  if (prob->model != MPXKKT_MODEL_CALLER)
    return fail(MPXKKT_ERR_ARG, "mpxkkt_sqp_accept takes MPXKKT_MODEL_CALLER problems");
  if (!prm || !trial_g || !trial_f) return fail(MPXKKT_ERR_ARG, "null params or trial evaluations");
  if (n_trials < 1 || n_trials > MAX_TRIALS) return fail(MPXKKT_ERR_ARG, "n_trials must be in [1, 16]");
  if ((s = check_iterate(prob, it)) != MPXKKT_OK) return s;
  if ((s = check_params(prm)) != MPXKKT_OK) return s;
  const size_t need = workspace_bytes(prob, prm);
  if (!workspace || workspace_bytes_ < need)
    return fail(MPXKKT_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (prob->dtype == MPXKKT_F64)
    return accept_t<double>(st, prob, prm, it, n_trials, trial_g, trial_f, stats, workspace);
  return accept_t<float>(st, prob, prm, it, n_trials, trial_g, trial_f, stats, workspace);
}

mpxkkt_status mpxkkt_srbd_iterate(void* stream, const mpxkkt_problem* prob, const mpxkkt_srbd* model,
                                  const mpxkkt_params* prm, mpxkkt_iterate* it, mpxkkt_stats* stats,
                                  void* workspace, size_t workspace_bytes_) {
  mpxkkt_status s = check_problem(prob);
  if (s != MPXKKT_OK) return s;
  if (prob->model != MPXKKT_MODEL_CALLER || prob->n != 13 || prob->m != 12)
    return fail(MPXKKT_ERR_ARG, "the SRBD model needs a MPXKKT_MODEL_CALLER problem with n = 13, m = 12");
  if (!prm || !model) return fail(MPXKKT_ERR_ARG, "null params or model");
  if (!model->W.ptr || !model->reference.ptr || !model->parameter.ptr)
    return fail(MPXKKT_ERR_ARG, "SRBD model needs W, reference and parameter views");
  if (model->W.stage_stride != 0 || model->W.batch_stride < 0 || model->reference.batch_stride < 0 ||
      model->reference.stage_stride < 0 || model->parameter.batch_stride < 0 || model->parameter.stage_stride < 0)
    return fail(MPXKKT_ERR_ARG, "bad SRBD view strides (W is per instance: stage stride 0)");
  if (!(model->mass > 0.0) || !(model->dt > 0.0)) return fail(MPXKKT_ERR_ARG, "mass and dt must be positive");
  if ((s = check_iterate(prob, it)) != MPXKKT_OK) return s;
  if ((s = check_params(prm)) != MPXKKT_OK) return s;
  const size_t need = workspace_bytes(prob, prm);
  if (!workspace || workspace_bytes_ < need)
    return fail(MPXKKT_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (prob->dtype == MPXKKT_F64) return srbd_iterate_t<double>(st, prob, model, prm, it, stats, workspace);
  return srbd_iterate_t<float>(st, prob, model, prm, it, stats, workspace);
}

mpxkkt_status mpxkkt_srbd_reference(void* stream, int32_t dtype, int32_t batch, int32_t horizon,
                                    const mpxkkt_srbd_gait* gait, void* t_timer, const void* x, const void* foot,
                                    const void* input, const void* contact, void* liftoff, void* reference,
                                    void* parameter, void* foot_ref_dot) {
  if (dtype != MPXKKT_F32 && dtype != MPXKKT_F64) return fail(MPXKKT_ERR_ARG, "bad dtype");
  if (batch < 1 || batch > (1 << 24) || horizon < 1 || horizon > 100000)
    return fail(MPXKKT_ERR_ARG, "batch and horizon must be positive");
  if (!gait || !t_timer || !x || !foot || !input || !contact || !liftoff || !reference || !parameter)
    return fail(MPXKKT_ERR_ARG, "null gait or array");
  if (!(gait->dt > 0.0) || !(gait->mass > 0.0) || !(gait->duty_factor > 0.0 && gait->duty_factor < 1.0) ||
      !(gait->step_freq > 0.0) || gait->advance_dt < 0.0)
    return fail(MPXKKT_ERR_ARG, "dt, mass, step_freq must be positive and duty_factor in (0, 1)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  auto launch = [&](auto zero) -> mpxkkt_status {
    using T = decltype(zero);
    SrbdGait<T> G;
    G.dt = (T)gait->dt; G.mass = (T)gait->mass; G.duty = (T)gait->duty_factor;
    G.step_freq = (T)gait->step_freq; G.step_height = (T)gait->step_height;
    G.clearance = (T)gait->clearance_speed; G.advance_dt = (T)gait->advance_dt;
    for (int k = 0; k < 12; ++k) G.foot0[k] = (T)gait->foot0[k];
    G.N = horizon; G.B = batch;
    G.t_timer = static_cast<T*>(t_timer); G.liftoff = static_cast<T*>(liftoff);
    G.x = static_cast<const T*>(x); G.foot = static_cast<const T*>(foot);
    G.input = static_cast<const T*>(input); G.contact = static_cast<const T*>(contact);
    G.ref = static_cast<T*>(reference); G.par = static_cast<T*>(parameter);
    G.foot_dot = static_cast<T*>(foot_ref_dot);
    hipLaunchKernelGGL(k_srbd_reference<T>, dim3((batch + 63) / 64), dim3(64), 0, st, G);
    return launch_check("k_srbd_reference");
  };
  if (dtype == MPXKKT_F64) return launch(0.0);
  return launch(0.0f);
}

mpxkkt_status mpxkkt_default_solve_opts(mpxkkt_solve_opts* out) {
  if (!out) return fail(MPXKKT_ERR_ARG, "null options");
  out->max_iter = 100;
  out->check_every = 0;
  out->dcost_tol = 1e-3;
  out->cc_tol = 1e-5;
  return MPXKKT_OK;
}

mpxkkt_status mpxkkt_solve(void* stream, const mpxkkt_problem* prob, const mpxkkt_params* prm,
                           const mpxkkt_solve_opts* opts, mpxkkt_iterate* it, mpxkkt_stats* stats,
                           int32_t* iters, void* cost, void* cc, void* workspace, size_t workspace_bytes_,
                           int32_t* launched) {
  mpxkkt_status s = check_problem(prob);
  if (s != MPXKKT_OK) return s;
  if (!prm || !opts || !iters) return fail(MPXKKT_ERR_ARG, "null params, options or iteration counts");
  if (prob->model == MPXKKT_MODEL_CALLER)
    return fail(MPXKKT_ERR_ARG, "caller dynamics: the solve loop needs the device model families");
  if ((s = check_iterate(prob, it)) != MPXKKT_OK) return s;
  if (opts->max_iter < 0 || opts->max_iter > 100000 || opts->check_every < 0)
    return fail(MPXKKT_ERR_ARG, "max_iter must be in [0, 100000] and check_every >= 0");
  if ((s = check_params(prm)) != MPXKKT_OK) return s;
  const size_t need = workspace_bytes(prob, prm);
  if (!workspace || workspace_bytes_ < need)
    return fail(MPXKKT_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (prob->dtype == MPXKKT_F64) return solve_t<double>(st, prob, prm, opts, it, stats, iters, cost, cc, workspace, launched);
  return solve_t<float>(st, prob, prm, opts, it, stats, iters, cost, cc, workspace, launched);
}

mpxkkt_status mpxkkt_kkt_residual(void* stream, const mpxkkt_problem* prob, const mpxkkt_params* prm,
                                  const mpxkkt_iterate* it, void* out, void* workspace,
                                  size_t workspace_bytes_) {
  mpxkkt_status s = check_problem(prob);
  if (s != MPXKKT_OK) return s;
  if (!prm || !out) return fail(MPXKKT_ERR_ARG, "null params or output");
  if ((s = check_iterate(prob, it)) != MPXKKT_OK) return s;
  const size_t need = workspace_bytes(prob, prm);
  if (!workspace || workspace_bytes_ < need)
    return fail(MPXKKT_ERR_WORKSPACE, "workspace needs " + std::to_string(need) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (prob->dtype == MPXKKT_F64) return kkt_t<double>(st, prob, prm, it, out, workspace);
  return kkt_t<float>(st, prob, prm, it, out, workspace);
}

mpxkkt_status mpxkkt_model_eval(void* stream, const mpxkkt_problem* prob, const void* X, const void* U, void* g,
                                void* c) {
  mpxkkt_status s = check_problem(prob);
  if (s != MPXKKT_OK) return s;
  if (!X || !U || !g || !c) return fail(MPXKKT_ERR_ARG, "null array");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nt = k1_threads(prob->n);
  dim3 grid(prob->horizon + 1, prob->batch);
  if (prob->dtype == MPXKKT_F64) {
    hipLaunchKernelGGL(k_model_eval<double>, grid, dim3(nt), nt * sizeof(double), st, make_prob<double>(prob),
                       (const double*)X, (const double*)U, (double*)g, (double*)c);
  } else {
    hipLaunchKernelGGL(k_model_eval<float>, grid, dim3(nt), nt * sizeof(float), st, make_prob<float>(prob),
                       (const float*)X, (const float*)U, (float*)g, (float*)c);
  }
  return launch_check("k_model_eval");
}

mpxkkt_status mpxkkt_shift_warmstart(void* stream, const mpxkkt_problem* prob, int32_t shift, mpxkkt_iterate* it,
                                     const void* u_reset, int32_t* reset) {
  mpxkkt_status s = check_problem(prob);
  if (s != MPXKKT_OK) return s;
  if ((s = check_iterate(prob, it)) != MPXKKT_OK) return s;
  if (shift < 0 || shift > prob->horizon) return fail(MPXKKT_ERR_ARG, "shift must be in [0, horizon]");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (prob->dtype == MPXKKT_F64) {
    hipLaunchKernelGGL(k_shift<double>, dim3(prob->batch), dim3(256), 0, st, make_prob<double>(prob), shift,
                       make_iter<double>(it), (const double*)u_reset, reset);
  } else {
    hipLaunchKernelGGL(k_shift<float>, dim3(prob->batch), dim3(256), 0, st, make_prob<float>(prob), shift,
                       make_iter<float>(it), (const float*)u_reset, reset);
  }
  return launch_check("k_shift");
}

mpxkkt_status mpxkkt_profile_enable(int32_t on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = on != 0;
  return MPXKKT_OK;
}

mpxkkt_status mpxkkt_profile_collect(double ms_per_phase[MPXKKT_PHASES], int32_t* calls) {
  if (!ms_per_phase || !calls) return fail(MPXKKT_ERR_ARG, "null output");
  for (int k = 0; k < MPXKKT_PHASES; ++k) ms_per_phase[k] = 0.0;
  mpxkkt_status s = MPXKKT_OK;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (auto& e : g_prof) {
    if (hipEventSynchronize(e.ev[MPXKKT_PHASES]) != hipSuccess)
      s = fail(MPXKKT_ERR_DEVICE, "hipEventSynchronize failed");
    for (int k = 0; k < MPXKKT_PHASES; ++k) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e.ev[k], e.ev[k + 1]) == hipSuccess) ms_per_phase[k] += ms;
    }
    for (int k = 0; k <= MPXKKT_PHASES; ++k) (void)hipEventDestroy(e.ev[k]);
  }
  *calls = (int32_t)g_prof.size();
  g_prof.clear();
  return s;
}

const char* mpxkkt_last_error(void) { return g_err.c_str(); }

int32_t mpxkkt_version(void) { return 100; }

}  // extern "C"
