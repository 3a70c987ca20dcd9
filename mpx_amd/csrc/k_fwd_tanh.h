// Forward pass of the LQR solve for the tanh model on the vector ALU, one wave
// per instance:
//   du_t = -(K_t dx_t + k_t),
//   dx_{t+1} = A_t dx_t + B du_t + c_{t+1},   A_t = I + h (A + kappa diag(sd_t) W)
// with sd_t = sech^2(W x_t) from k_linearize_mma and A, W, B stage-invariant.
// A_t is never read: A_t dx = (I + h A) dx + sd_t (h kappa W dx), with I + h A
// and h B in registers and h kappa W in LDS for the whole chain, so a stage reads
// only K_t, k_t, c_{t+1} and sd_t (about 1.7 KiB at n = 36, m = 12 instead of
// 7 KiB with A_t).  Lane (g, c) holds rows 16 i + c, columns 16 k + 4 g + s of
// the matrices (A-operand tiles) and dx's entries 16 k + 4 g + s; a product is
// 4 NT partial sums per lane and a sum over the four lanes of a column
// (col_sum4); the result's rows are gathered back into the column layout by
// ds_bpermute.  (On the matrix cores only one of the 16 columns of each
// product would be dx's: k_fwd_wave.)  The stage data runs R - 1 stages ahead
// in a ring of registers.
#pragma once
#include "k_lin_mma.h"

namespace mpxk {

// One stage's data, lane (g, c): K_t[c][16 k + 4 g + s] in register s of K[k]
// (0 past row m - 1 and column n - 1), k_t[c] (0 past m - 1), c_{t+1}[16 i + c]
// and sd_t[16 i + c] (0 past n - 1).
template <int NT>
struct FwdT {
  f4 K[NT];
  float k;
  float cn[NT];
  float sd[NT];
};

template <int NT, bool WIDE>
__device__ __forceinline__ void fwdt_load(FwdT<NT>& d, const Ws<float>& ws, int b, int t, int N, int n, int m,
                                          int c, const int (&off)[NT][4]) {
  f4 Kt[1][NT];
  load_aop_raw<NT, 1, WIDE>(Kt, ws.K + ((size_t)b * N + t) * m * n, n, m, off);
#pragma unroll
  for (int k = 0; k < NT; ++k) d.K[k] = Kt[0][k];
  d.k = rload(make_rsrc(ws.k + ((size_t)b * N + t) * m, m), c * 4);
  const rsrc_t rc = make_rsrc(ws.cdef + ((size_t)b * (N + 1) + t + 1) * n, n);
  const rsrc_t rs = make_rsrc(ws.sdh + ((size_t)b * N + t) * n, n);
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    d.cn[i] = rload(rc, (16 * i + c) * 4);
    d.sd[i] = rload(rs, (16 * i + c) * 4);
  }
}

__device__ __forceinline__ float lane_read(float x, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src * 4, __float_as_int(x)));
}

// (h kappa W stays in LDS, each lane reading back only its own entries, so
// that the kernel fits 128 VGPRs: four waves per SIMD, the whole batch of
// 4096 instances resident at once on 1024 SIMDs)
template <int NT, bool WIDE, int R>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WIDE ? 4 : 3))) k_fwd_tanh(Prob<float> P, Ws<float> ws) {
  __shared__ f4 wl[NT * NT * 64];
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  const float hs = P.h, hk = P.h * P.kappa;
  int off[NT][4];
  aop_offsets<NT>(off, n, g, c);
  // I + h A, h kappa W (rows 16 i + c, columns 16 k + 4 g + s) and h B
  // (columns 4 g + s), raw for now
  f4 Ah[NT][NT], Wh[NT][NT], Bh[NT];
  load_aop_raw<NT, NT, WIDE>(Ah, P.A.at(b, 0), n, n, off);
  load_aop_raw<NT, NT, WIDE>(Wh, P.W.at(b, 0), n, n, off);
  {
    const rsrc_t rB = make_rsrc(P.Bm.at(b, 0), n * m);
#pragma unroll
    for (int i = 0; i < NT; ++i) Bh[i] = vload4<WIDE>(rB, (16 * i + c) * m + 4 * g, 4 * g, m);
  }
  // dx_0 = c_init, entries 16 k + 4 g + s in every lane of row g
  f4 xs[NT];
  {
    const rsrc_t r0 = make_rsrc(ws.cdef + (size_t)b * N1 * n, n);
#pragma unroll
    for (int k = 0; k < NT; ++k) xs[k] = vload4<WIDE>(r0, 16 * k + 4 * g, 16 * k + 4 * g, n);
  }
  // the ring: stages 0 .. R - 2 now, stage t + R - 1 before stage t's step
  FwdT<NT> ring[R];
  static_for<0, R - 1>([&](auto J) {
    constexpr int j = decltype(J)::value;
    fwdt_load<NT, WIDE>(ring[j], ws, b, j < N ? j : N - 1, N, n, m, c, off);
  });
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int k = 0; k < NT; ++k) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int row = 16 * i + c, col = 16 * k + 4 * g + s;
        Ah[i][k][s] = (row == col && row < n ? 1.f : 0.f) + hs * Ah[i][k][s];
        Wh[i][k][s] = hk * Wh[i][k][s];
      }
      wl[(i * NT + k) * 64 + lane] = Wh[i][k];
    }
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int s = 0; s < 4; ++s) Bh[i][s] = hs * Bh[i][s];

  const rsrc_t rdx = make_rsrc(ws.dx + (size_t)b * N1 * n, N1 * n);
  const rsrc_t rdu = make_rsrc(ws.du + (size_t)b * N * m, N * m);
#pragma unroll
  for (int k = 0; k < NT; ++k) vstore4<WIDE>(rdx, 16 * k + 4 * g, 16 * k + 4 * g, n, c == 0, xs[k]);

  // stage t (t >= N: a padding stage of the last ring pass, nothing stored)
  auto step = [&](const FwdT<NT>& d, int t) {
    // du = -(K dx + k) on lane column c (every row g)
    float pk = 0.f;
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) pk += d.K[k][s] * xs[k][s];
    const float du = -(col_sum4(pk) + d.k);
    // du[4 g + s] into lane (g, c): lane 20 g + s holds it
    f4 dug;
#pragma unroll
    for (int s = 0; s < 4; ++s) dug[s] = lane_read(du, 20 * g + s);
    // dx_{t+1} rows 16 i + c: (I + h A) dx + sd (h kappa W dx) + h B du + c
    float y[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      float pa = 0.f, pw = 0.f;
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        const f4 w = wl[(i * NT + k) * 64 + lane];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          pa += Ah[i][k][s] * xs[k][s];
          pw += w[s] * xs[k][s];
        }
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) pa += Bh[i][s] * dug[s];
      y[i] = col_sum4(pa) + d.sd[i] * col_sum4(pw) + d.cn[i];
    }
    const bool st = t < N;
    rstore(rdu, st && g == 0 && c < m ? (t * m + c) * 4 : kOOB, du);
    // entries 16 k + 4 g + s of dx_{t+1}: lane 20 g + s holds row 16 k + 4 g + s
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) xs[k][s] = lane_read(y[k], 20 * g + s);
#pragma unroll
    for (int k = 0; k < NT; ++k)
      vstore4<WIDE>(rdx, (t + 1) * n + 16 * k + 4 * g, 16 * k + 4 * g, n, st && c == 0, xs[k]);
  };

  for (int t0 = 0; t0 < N; t0 += R) {
    static_for<0, R>([&](auto J) {
      constexpr int j = decltype(J)::value;
      const int t = t0 + j;
      const int tl = t + R - 1 < N ? t + R - 1 : N - 1;
      fwdt_load<NT, WIDE>(ring[(j + R - 1) % R], ws, b, tl, N, n, m, c, off);
      __builtin_amdgcn_sched_barrier(0);
      step(ring[j], t);
    });
  }
}

}  // namespace mpxk
