/*
 * FP64 (and, with -DPDILQR_F32, FP32) C restatement of one primal-dual SQP iteration (TEST INFRASTRUCTURE and
 * CPU baseline only; the product never links this).  Same algorithm and
 * decisions D1-D10 as oracle/pdilqr.py (linearise + augmented Lagrangian of the
 * equality rows + interior-point box, sequential Riccati LQR, merit line search
 * and multiplier updates), for the two built-in dynamics families, batched over
 * instances with OpenMP.  The reference solver this follows is the un-vendored
 * primal_dual_ilqr submodule behind optimizers.mpc (mpx/utils/mpc_wrapper.py:74).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* real: double, or float with -DPDILQR_F32 (the same algorithm in FP32
 * arithmetic: the error model of the device's FP32 path). */
#ifdef PDILQR_F32
typedef float real;
/* FP32 library functions and constants in the FP32 build */
#define tanh tanhf
#define log logf
#define sqrt sqrtf
#define fabs fabsf
#else
typedef double real;
#endif
#define RC(x) ((real)(x))

typedef struct {
  int n, m, N, p, model, box, q_tv, r_tv;
  real h, kappa;
  const real *A, *B, *W, *Q, *R, *QN, *xref, *uref, *D, *E, *e, *ulo, *uhi, *x0;
} inst_t;

enum { P_RHO0, P_RHOMAX, P_RHOC, P_MUB, P_TAU, P_BETA, P_C1, P_SLACK, P_TRIALS, P_HFLOOR };

/* C = A (r x k) * B (k x c) */
static void mm(const real* A, const real* B, real* C, int r, int k, int c) {
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < c; ++j) {
      real s = 0;
      for (int l = 0; l < k; ++l) s += A[i * k + l] * B[l * c + j];
      C[i * c + j] = s;
    }
}

/* C = A^T (A: k x r) * B (k x c) */
static void mtm(const real* A, const real* B, real* C, int k, int r, int c) {
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < c; ++j) {
      real s = 0;
      for (int l = 0; l < k; ++l) s += A[l * r + i] * B[l * c + j];
      C[i * c + j] = s;
    }
}

/* in-place Cholesky (lower) with pivots max(|pivot|, floor |H_jj|) (decision
 * D2 of oracle/pdilqr.py); returns 0 on success */
static int chol(real* H, int m, real floor) {
  for (int j = 0; j < m; ++j) {
    real d = H[j * m + j];
    const real fl = floor * fabs(H[j * m + j]);
    for (int k = 0; k < j; ++k) d -= H[j * m + k] * H[j * m + k];
    d = fabs(d);
    if (!(d >= fl) && !isnan(d)) d = fl;
    if (!(d > 0)) return -1;
    d = sqrt(d);
    H[j * m + j] = d;
    for (int i = j + 1; i < m; ++i) {
      real v = H[i * m + j];
      for (int k = 0; k < j; ++k) v -= H[i * m + k] * H[j * m + k];
      H[i * m + j] = v / d;
    }
  }
  return 0;
}

/* solve (L L^T) X = G for the ncol columns of G (m x ncol) in place */
static void chol_solve(const real* L, real* G, int m, int ncol) {
  for (int c = 0; c < ncol; ++c) {
    for (int i = 0; i < m; ++i) {
      real v = G[i * ncol + c];
      for (int k = 0; k < i; ++k) v -= L[i * m + k] * G[k * ncol + c];
      G[i * ncol + c] = v / L[i * m + i];
    }
    for (int i = m - 1; i >= 0; --i) {
      real v = G[i * ncol + c];
      for (int k = i + 1; k < m; ++k) v -= L[k * m + i] * G[k * ncol + c];
      G[i * ncol + c] = v / L[i * m + i];
    }
  }
}

static const real* Qt(const inst_t* I, int t) { return I->Q + (I->q_tv ? (size_t)t * I->n * I->n : 0); }
static const real* Rt(const inst_t* I, int t) { return I->R + (I->r_tv ? (size_t)t * I->m * I->m : 0); }

static void dyn(const inst_t* I, const real* x, const real* u, int t, real* f, real* fx, real* wk) {
  const int n = I->n, m = I->m;
  if (I->model == 0) {
    const real* A = I->A + (size_t)t * n * n;
    const real* B = I->B + (size_t)t * n * m;
    for (int i = 0; i < n; ++i) {
      real v = 0;
      for (int j = 0; j < n; ++j) v += A[i * n + j] * x[j];
      for (int j = 0; j < m; ++j) v += B[i * m + j] * u[j];
      f[i] = v;
    }
    if (fx) memcpy(fx, A, sizeof(real) * n * n);
    return;
  }
  for (int i = 0; i < n; ++i) {
    real ax = 0, wx = 0, bu = 0;
    for (int j = 0; j < n; ++j) { ax += I->A[i * n + j] * x[j]; wx += I->W[i * n + j] * x[j]; }
    for (int j = 0; j < m; ++j) bu += I->B[i * m + j] * u[j];
    const real th = tanh(wx);
    wk[i] = RC(1.0) - th * th;
    f[i] = x[i] + I->h * (ax + bu + I->kappa * th);
  }
  if (fx)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        fx[i * n + j] = (i == j ? RC(1.0) : RC(0.0)) + I->h * (I->A[i * n + j] + I->kappa * wk[i] * I->W[i * n + j]);
}

static void fu_of(const inst_t* I, int t, real* fu) {
  const int n = I->n, m = I->m;
  if (I->model == 0) memcpy(fu, I->B + (size_t)t * n * m, sizeof(real) * n * m);
  else for (int k = 0; k < n * m; ++k) fu[k] = I->h * I->B[k];
}

/* stage AL / barrier cost (t < N) or terminal cost (t == N) */
static real al_cost(const inst_t* I, const real* x, const real* u, int t, const real* yeq,
                      const real* prm) {
  const int n = I->n, m = I->m, p = I->p, N = I->N;
  const real* xr = I->xref + (size_t)t * n;
  real val = 0;
  const real* Q = t == N ? I->QN : Qt(I, t);
  for (int i = 0; i < n; ++i) {
    real qx = 0;
    for (int j = 0; j < n; ++j) qx += Q[i * n + j] * (x[j] - xr[j]);
    val += RC(0.5) * (x[i] - xr[i]) * qx;
  }
  if (t == N) return val;
  const real* R = Rt(I, t);
  const real* ur = I->uref + (size_t)t * m;
  for (int i = 0; i < m; ++i) {
    real ru = 0;
    for (int j = 0; j < m; ++j) ru += R[i * m + j] * (u[j] - ur[j]);
    val += RC(0.5) * (u[i] - ur[i]) * ru;
  }
  for (int r = 0; r < p; ++r) {
    const real* D = I->D + ((size_t)t * p + r) * n;
    const real* E = I->E + ((size_t)t * p + r) * m;
    real hv = I->e[(size_t)t * p + r];
    for (int j = 0; j < n; ++j) hv += D[j] * x[j];
    for (int j = 0; j < m; ++j) hv += E[j] * u[j];
    val += yeq[r] * hv + RC(0.5) * prm[P_RHOC] * hv * hv;
  }
  if (I->box)
    for (int i = 0; i < m; ++i) val -= prm[P_MUB] * (log(I->uhi[i] - u[i]) + log(u[i] - I->ulo[i]));
  return val;
}

static real merit(const inst_t* I, const real* X, const real* U, const real* V, const real* Yeq,
                    real rho, const real* prm, real* wk) {
  const int n = I->n, m = I->m, N = I->N, p = I->p;
  real val = 0, *f = wk, *tmp = wk + n;
  for (int t = 0; t <= N; ++t) val += al_cost(I, X + (size_t)t * n, t < N ? U + (size_t)t * m : NULL, t,
                                              t < N ? Yeq + (size_t)t * p : NULL, prm);
  for (int i = 0; i < n; ++i) {
    const real c = I->x0[i] - X[i];
    val += V[i] * c + RC(0.5) * rho * c * c;
  }
  for (int t = 0; t < N; ++t) {
    dyn(I, X + (size_t)t * n, U + (size_t)t * m, t, f, NULL, tmp);
    for (int i = 0; i < n; ++i) {
      const real c = f[i] - X[(size_t)(t + 1) * n + i];
      val += V[(size_t)(t + 1) * n + i] * c + RC(0.5) * rho * c * c;
    }
  }
  return val;
}

/* one SQP iteration of one instance, in place; returns the accepted step */
static real sqp_one(const inst_t* I, real* X, real* U, real* V, real* Yeq, real* Ybox,
                      const real* prm) {
  const int n = I->n, m = I->m, N = I->N, p = I->p, N1 = N + 1;
  const real rc = prm[P_RHOC], mub = prm[P_MUB], tau = prm[P_TAU];
  const size_t nn = (size_t)n * n, nm = (size_t)n * m;
  const size_t tsz = (size_t)(n + m + 1) * (n + m + 1);  /* any n x n, m x (n+1), m x m temporary */
  real* mem = (real*)calloc((size_t)N * (2 * nn + 3 * nm + (size_t)m * m) + (size_t)N1 * (nn + 4 * n) +
                                    (size_t)N * 7 * m + 5 * tsz + 8 * (size_t)(n + m + 4),
                                sizeof(real));
  real* q = mem;
  real* Af = q;  q += (size_t)N * nn;      /* A_t */
  real* Bf = q;  q += (size_t)N * nm;      /* B_t */
  real* Kf = q;  q += (size_t)N * nm;      /* K_t (m x n) */
  real* Mh = q;  q += (size_t)N * nm;      /* M^ (n x m) */
  real* Qh = q;  q += (size_t)N * nn;      /* Q^ */
  real* Rh = q;  q += (size_t)N * m * m;   /* R^ */
  real* S = q;   q += (size_t)N1 * nn;
  real* s = q;   q += (size_t)N1 * n;
  real* qg = q;  q += (size_t)N1 * n;      /* gradients (row N: terminal) */
  real* cd = q;  q += (size_t)N1 * n;      /* constraint rows */
  real* dx = q;  q += (size_t)N1 * n;
  real* rg = q;  q += (size_t)N * m;
  real* kf = q;  q += (size_t)N * m;
  real* du = q;  q += (size_t)N * m;
  real* dY = q;  q += (size_t)N * 2 * m;
  real* Sl = q;  q += (size_t)N * 2 * m;
  real* t1 = q;  q += tsz;
  real* t2 = q;  q += tsz;
  real* t3 = q;  q += tsz;
  real* H = q;   q += tsz;
  real* G = q;   q += tsz;
  real* w = q;   q += 8 * (size_t)(n + m + 4);
  /* box: interior projection and positive multipliers */
  if (I->box) {
    for (int t = 0; t < N; ++t)
      for (int i = 0; i < m; ++i) {
        real* u = U + (size_t)t * m + i;
        const real lo = I->ulo[i], hi = I->uhi[i], wd = RC(1e-3) * (hi - lo);
        if (*u >= hi) *u = hi - wd;
        if (*u <= lo) *u = lo + wd;
        Sl[(size_t)t * 2 * m + i] = hi - *u;
        Sl[(size_t)t * 2 * m + m + i] = *u - lo;
        for (int k = 0; k < 2; ++k) {
          real* nu = Ybox + (size_t)t * 2 * m + k * m + i;
          if (!(*nu > 0)) *nu = mub / Sl[(size_t)t * 2 * m + k * m + i];
        }
      }
  }
  /* linearise + augment */
  for (int i = 0; i < n; ++i) cd[i] = I->x0[i] - X[i];
  for (int t = 0; t < N; ++t) {
    const real *x = X + (size_t)t * n, *u = U + (size_t)t * m;
    dyn(I, x, u, t, w, Af + (size_t)t * nn, w + n);
    fu_of(I, t, Bf + (size_t)t * nm);
    for (int i = 0; i < n; ++i) cd[(size_t)(t + 1) * n + i] = w[i] - X[(size_t)(t + 1) * n + i];
    const real* Q = Qt(I, t);
    const real* R = Rt(I, t);
    memcpy(Qh + (size_t)t * nn, Q, sizeof(real) * nn);
    memcpy(Rh + (size_t)t * m * m, R, sizeof(real) * m * m);
    for (int i = 0; i < n; ++i) {
      real v = 0;
      for (int j = 0; j < n; ++j) v += Q[i * n + j] * (x[j] - I->xref[(size_t)t * n + j]);
      qg[(size_t)t * n + i] = v;
    }
    for (int i = 0; i < m; ++i) {
      real v = 0;
      for (int j = 0; j < m; ++j) v += R[i * m + j] * (u[j] - I->uref[(size_t)t * m + j]);
      rg[(size_t)t * m + i] = v;
    }
    for (int r = 0; r < p; ++r) {
      const real* D = I->D + ((size_t)t * p + r) * n;
      const real* E = I->E + ((size_t)t * p + r) * m;
      real hv = I->e[(size_t)t * p + r];
      for (int j = 0; j < n; ++j) hv += D[j] * x[j];
      for (int j = 0; j < m; ++j) hv += E[j] * u[j];
      const real wv = Yeq[(size_t)t * p + r] + rc * hv;
      for (int i = 0; i < n; ++i) {
        qg[(size_t)t * n + i] += D[i] * wv;
        for (int j = 0; j < n; ++j) Qh[(size_t)t * nn + i * n + j] += rc * D[i] * D[j];
        for (int j = 0; j < m; ++j) Mh[(size_t)t * nm + i * m + j] += rc * D[i] * E[j];
      }
      for (int i = 0; i < m; ++i) {
        rg[(size_t)t * m + i] += E[i] * wv;
        for (int j = 0; j < m; ++j) Rh[(size_t)t * m * m + i * m + j] += rc * E[i] * E[j];
      }
    }
    if (I->box)
      for (int i = 0; i < m; ++i) {
        const real shi = Sl[(size_t)t * 2 * m + i], slo = Sl[(size_t)t * 2 * m + m + i];
        const real nhi = Ybox[(size_t)t * 2 * m + i], nlo = Ybox[(size_t)t * 2 * m + m + i];
        rg[(size_t)t * m + i] += mub / shi - mub / slo;
        Rh[(size_t)t * m * m + i * m + i] += nhi / shi + nlo / slo;
      }
  }
  for (int i = 0; i < n; ++i) {
    real v = 0;
    for (int j = 0; j < n; ++j) v += I->QN[i * n + j] * (X[(size_t)N * n + j] - I->xref[(size_t)N * n + j]);
    qg[(size_t)N * n + i] = v;
  }
  /* backward Riccati */
  memcpy(S + (size_t)N * nn, I->QN, sizeof(real) * nn);
  memcpy(s + (size_t)N * n, qg + (size_t)N * n, sizeof(real) * n);
  for (int t = N - 1; t >= 0; --t) {
    const real *A = Af + (size_t)t * nn, *B = Bf + (size_t)t * nm, *Sn = S + (size_t)(t + 1) * nn;
    real* Sc = w;
    for (int i = 0; i < n; ++i) {
      real v = s[(size_t)(t + 1) * n + i];
      for (int j = 0; j < n; ++j) v += Sn[i * n + j] * cd[(size_t)(t + 1) * n + j];
      Sc[i] = v;
    }
    mm(Sn, A, t1, n, n, n);          /* S A */
    mm(Sn, B, t2, n, n, m);          /* S B */
    mtm(B, t2, H, n, m, m);          /* B'SB */
    for (int k = 0; k < m * m; ++k) H[k] += Rh[(size_t)t * m * m + k];
    mtm(B, t1, G, n, m, n);          /* B'SA (m x n) -> columns 0..n-1 of [G | g] below */
    real* Gg = t3;                 /* m x (n+1) */
    for (int i = 0; i < m; ++i) {
      for (int j = 0; j < n; ++j) Gg[i * (n + 1) + j] = G[i * n + j] + Mh[(size_t)t * nm + j * m + i];
      real v = rg[(size_t)t * m + i];
      for (int k = 0; k < n; ++k) v += B[k * m + i] * Sc[k];
      Gg[i * (n + 1) + n] = v;
    }
    if (chol(H, m, prm[P_HFLOOR])) { free(mem); return NAN; }
    real* Y = G;                   /* keep G' = [G | g] for S update */
    memcpy(Y, Gg, sizeof(real) * m * (n + 1));
    chol_solve(H, Gg, m, n + 1);     /* [K | k] */
    for (int i = 0; i < m; ++i) {
      for (int j = 0; j < n; ++j) Kf[(size_t)t * nm + i * n + j] = Gg[i * (n + 1) + j];
      kf[(size_t)t * m + i] = Gg[i * (n + 1) + n];
    }
    real* St = S + (size_t)t * nn;
    mtm(A, t1, St, n, n, n);         /* A'SA */
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        real v = Qh[(size_t)t * nn + i * n + j] + St[i * n + j];
        for (int r = 0; r < m; ++r) v -= Y[r * (n + 1) + i] * Gg[r * (n + 1) + j];
        St[i * n + j] = v;
      }
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        const real a = RC(0.5) * (St[i * n + j] + St[j * n + i]);
        St[i * n + j] = St[j * n + i] = a;
      }
    for (int i = 0; i < n; ++i) {
      real v = qg[(size_t)t * n + i];
      for (int k = 0; k < n; ++k) v += A[k * n + i] * Sc[k];
      for (int r = 0; r < m; ++r) v -= Y[r * (n + 1) + i] * Gg[r * (n + 1) + n];
      s[(size_t)t * n + i] = v;
    }
  }
  /* forward pass, costates, slope terms */
  real d_f = 0, d_c = 0, lc = 0, cc = 0;
  memcpy(dx, cd, sizeof(real) * n);
  real* dV = (real*)malloc(sizeof(real) * N1 * n);
  for (int t = 0; t <= N; ++t) {
    const real* St = S + (size_t)t * nn;
    for (int i = 0; i < n; ++i) {
      real lam = s[(size_t)t * n + i];
      for (int j = 0; j < n; ++j) lam += St[i * n + j] * dx[(size_t)t * n + j];
      const real v = V[(size_t)t * n + i], c = cd[(size_t)t * n + i];
      dV[(size_t)t * n + i] = lam - v;
      d_c += (lam - 2 * v) * c;
      lc += lam * c;
      cc += c * c;
      d_f += qg[(size_t)t * n + i] * dx[(size_t)t * n + i];
    }
    if (t == N) break;
    for (int i = 0; i < m; ++i) {
      real v = kf[(size_t)t * m + i];
      for (int j = 0; j < n; ++j) v += Kf[(size_t)t * nm + i * n + j] * dx[(size_t)t * n + j];
      du[(size_t)t * m + i] = -v;
      d_f -= rg[(size_t)t * m + i] * v;
    }
    const real *A = Af + (size_t)t * nn, *B = Bf + (size_t)t * nm;
    for (int i = 0; i < n; ++i) {
      real v = cd[(size_t)(t + 1) * n + i];
      for (int j = 0; j < n; ++j) v += A[i * n + j] * dx[(size_t)t * n + j];
      for (int j = 0; j < m; ++j) v += B[i * m + j] * du[(size_t)t * m + j];
      dx[(size_t)(t + 1) * n + i] = v;
    }
  }
  const real kap = lc - d_f;
  real rho = prm[P_RHO0];
  if (cc > 0) { const real r2 = 2 * (d_f + d_c + RC(0.5) * kap) / cc; if (r2 > rho) rho = r2; }
  if (rho > prm[P_RHOMAX]) rho = prm[P_RHOMAX];
  const real dphi = d_f + d_c - rho * cc;
  real amax = 1, adual = 1;
  if (I->box)
    for (int t = 0; t < N; ++t)
      for (int i = 0; i < m; ++i) {
        const real d = du[(size_t)t * m + i];
        const real shi = Sl[(size_t)t * 2 * m + i], slo = Sl[(size_t)t * 2 * m + m + i];
        const real nhi = Ybox[(size_t)t * 2 * m + i], nlo = Ybox[(size_t)t * 2 * m + m + i];
        if (d > 0 && tau * shi / d < amax) amax = tau * shi / d;
        if (d < 0 && tau * slo / (-d) < amax) amax = tau * slo / (-d);
        const real dhi = mub / shi - nhi + nhi * d / shi, dlo = mub / slo - nlo - nlo * d / slo;
        dY[(size_t)t * 2 * m + i] = dhi;
        dY[(size_t)t * 2 * m + m + i] = dlo;
        if (dhi < 0 && -tau * nhi / dhi < adual) adual = -tau * nhi / dhi;
        if (dlo < 0 && -tau * nlo / dlo < adual) adual = -tau * nlo / dlo;
      }
  real* wk = (real*)malloc(sizeof(real) * 4 * (n + m + 4));
  const real phi0 = merit(I, X, U, V, Yeq, rho, prm, wk);
  const real slack = prm[P_SLACK] * (1 + fabs(phi0));
  real* Xt = (real*)malloc(sizeof(real) * (N1 * n * 2 + N * m));
  real *Vt = Xt + N1 * n, *Ut = Vt + N1 * n;
  real alpha = -1, a = amax;
  const int trials = (int)prm[P_TRIALS];
  for (int j = 0; j < trials; ++j, a *= prm[P_BETA]) {
    for (int k = 0; k < N1 * n; ++k) { Xt[k] = X[k] + a * dx[k]; Vt[k] = V[k] + a * dV[k]; }
    for (int k = 0; k < N * m; ++k) Ut[k] = U[k] + a * du[k];
    if (merit(I, Xt, Ut, Vt, Yeq, rho, prm, wk) <= phi0 + prm[P_C1] * a * dphi + slack) { alpha = a; break; }
  }
  if (alpha < 0) { alpha = amax; for (int j = 0; j < trials - 1; ++j) alpha *= prm[P_BETA]; }
  for (int k = 0; k < N1 * n; ++k) { X[k] += alpha * dx[k]; V[k] += alpha * dV[k]; }
  for (int k = 0; k < N * m; ++k) U[k] += alpha * du[k];
  for (int t = 0; t < N; ++t)
    for (int r = 0; r < p; ++r) {
      const real* D = I->D + ((size_t)t * p + r) * n;
      const real* E = I->E + ((size_t)t * p + r) * m;
      real hv = I->e[(size_t)t * p + r];
      for (int j = 0; j < n; ++j) hv += D[j] * X[(size_t)t * n + j];
      for (int j = 0; j < m; ++j) hv += E[j] * U[(size_t)t * m + j];
      Yeq[(size_t)t * p + r] += rc * hv;
    }
  if (I->box)
    for (int k = 0; k < N * 2 * m; ++k) Ybox[k] += adual * dY[k];
  free(Xt);
  free(wk);
  free(dV);
  free(mem);
  return alpha;
}

/* iters SQP iterations for B instances (arrays instance-major, as the device
 * layout); returns 0.  Q/R are [B][Tq][n][n] with q_tv = (Tq == N). */
int pdilqr_solve_batch(int B, int n, int m, int N, int p, int model, int box, int q_tv, int r_tv, real h,
                       real kappa, const real* A, const real* Bm, const real* W, const real* Q,
                       const real* R, const real* QN, const real* xref, const real* uref,
                       const real* D, const real* E, const real* e, const real* ulo, const real* uhi,
                       const real* x0, real* X, real* U, real* V, real* Yeq, real* Ybox,
                       const real* prm, int iters, real* alpha_out, int nthreads) {
  const size_t nA = model == 0 ? (size_t)N * n * n : (size_t)n * n;
  const size_t nB = model == 0 ? (size_t)N * n * m : (size_t)n * m;
  const size_t nQ = (size_t)(q_tv ? N : 1) * n * n, nR = (size_t)(r_tv ? N : 1) * m * m;
  (void)nthreads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int b = 0; b < B; ++b) {
    inst_t I;
    I.n = n; I.m = m; I.N = N; I.p = p; I.model = model; I.box = box; I.q_tv = q_tv; I.r_tv = r_tv;
    I.h = h; I.kappa = kappa;
    I.A = A + b * nA; I.B = Bm + b * nB; I.W = W ? W + (size_t)b * n * n : NULL;
    I.Q = Q + b * nQ; I.R = R + b * nR; I.QN = QN + (size_t)b * n * n;
    I.xref = xref + (size_t)b * (N + 1) * n; I.uref = uref + (size_t)b * N * m;
    I.D = p ? D + (size_t)b * N * p * n : NULL; I.E = p ? E + (size_t)b * N * p * m : NULL;
    I.e = p ? e + (size_t)b * N * p : NULL;
    I.ulo = box ? ulo + (size_t)b * m : NULL; I.uhi = box ? uhi + (size_t)b * m : NULL;
    I.x0 = x0 + (size_t)b * n;
    for (int k = 0; k < iters; ++k) {
      const real a = sqp_one(&I, X + (size_t)b * (N + 1) * n, U + (size_t)b * N * m, V + (size_t)b * (N + 1) * n,
                               p ? Yeq + (size_t)b * N * p : NULL, box ? Ybox + (size_t)b * N * 2 * m : NULL, prm);
      if (alpha_out) alpha_out[(size_t)b * iters + k] = a;
    }
  }
  return 0;
}
