// csem_mixed_gmres.h -- FP32 inner FGMRES(m) of the mixed-precision solve
// (gfx950).
//
// One restart cycle of FGMRES (solver.c:403-512) on R right-hand sides at
// once: right-preconditioned Arnoldi that keeps the preconditioned vectors
// z_j = M^-1 v_j, Givens rotations on the Hessenberg matrix, the residual
// estimate |g_{j+1}|, back-substitution and x = Z y.  The vectors are FP32;
// every inner product is summed in FP64 in a fixed order (per chunk of
// RED_CHUNK elements, per thread, then a wave tree and the four waves in
// order), and the Hessenberg matrix, the rotations and g live in FP64 on
// the device, so a whole cycle replays from one graph without the host.
//
// Orthogonalisation is classical Gram-Schmidt with one reorthogonalisation
// pass (CGS2) where the reference uses modified Gram-Schmidt: a step forms
// its j+1 projections in one pass over the basis instead of j+1 dependent
// dot/axpy pairs, and the second pass restores the orthogonality MGS
// gives.  Three passes per step:
//   dot:  h_i = v_i^H w                      (reads w and v_0..v_j)
//   upd1: w' = w - sum h_i v_i, c_i = v_i^H w', |w'|^2   (one pass)
//   upd2: v_{j+1} = (w' - sum c_i v_i) / h_{j+1,j}       (one pass)
// with h_{j+1,j}^2 = |w'|^2 - sum |c_i|^2 (w'' is w' less its projection
// on an orthonormal basis).  Each cycle starts from x = 0 on the scaled
// true residual that the FP64 refinement hands over, which is what the
// reference's restart does with its recomputed residual (solver.c:420-425).
#pragma once
#include "csem_mixed_kernels.h"

#define GM_MMAX 16  // longest restart cycle
// Per right-hand side FP64 state, GM_HS doubles: Hessenberg column j at
// GM_H + 2 (GM_MMAX + 1) j (complex rows 0..j+1), the rotations (cos real,
// sin complex), g, y, the second-pass projections c and the factor that
// normalises the vector being formed.
enum {
  GM_H = 0,
  GM_CS = GM_H + 2 * (GM_MMAX + 1) * GM_MMAX,
  GM_SN = GM_CS + GM_MMAX,
  GM_G = GM_SN + 2 * GM_MMAX,
  GM_Y = GM_G + 2 * (GM_MMAX + 1),
  GM_C = GM_Y + 2 * GM_MMAX,
  GM_INV = GM_C + 2 * (GM_MMAX + 1),
  GM_BETA = GM_INV + 1,
  GM_HS = ((GM_BETA + 8) / 8) * 8
};
// partial sums per chunk and right-hand side: a norm and GM_MMAX + 1
// complex projections at most
#define GM_PS (2 * GM_MMAX + 4)

struct MxGm {
  long n;
  int nch, m, R;
  c2 *V;         // (m + 1) x R x n: basis vector i of RHS r at (i R + r) n
  c2 *Z;         // m x R x n: z_j = M^-1 v_j
  const c2 *r;   // R x n: the cycle's right-hand side
  c2 *x;         // R x n: the correction Z y
  double *part;  // R x nch x GM_PS
  double *sc;    // R x MS_N: MS_DONE, MS_USED, MS_TOL, MS_IT, MS_MAXIT, MS_REL
  double *hs;    // R x GM_HS
};

__device__ __forceinline__ const c2 *gm_v(const MxGm &g, int i, int r) { return g.V + ((long)i * g.R + r) * g.n; }
__device__ __forceinline__ bool gm_done(const MxGm &g, int r) { return g.sc[r * MS_N + MS_DONE] != 0.0; }

// Fold the first nq of acc[0..NQ) over the block (blockDim == RED_T): a
// tree within each wave (lanes l += l + h, h = 32..1), then the four waves'
// results in order; thread q < nq writes out[q].
template <int NQ>
__device__ __forceinline__ void gm_fold(const double *acc, int nq, double *out)
{
  __shared__ double sh[RED_T / 64][NQ];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    if (q < nq) {
      double v = acc[q];
      for (int h = 32; h >= 1; h >>= 1) v = v + __shfl_down(v, h, 64);
      if (l == 0) sh[w][q] = v;
    }
  }
  __syncthreads();
  if (t < nq) out[t] = ((sh[0][t] + sh[1][t]) + sh[2][t]) + sh[3][t];
}

// |r|^2 partials of the cycle's right-hand side
__global__ __launch_bounds__(RED_T) void k_gm_norm(MxGm g)
{
  const int r = blockIdx.y;
  if (gm_done(g, r)) return;
  double acc[1] = {0};
  const c2 *v = g.r + (long)r * g.n;
  const long base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int m = 0; m < RED_E; m++) {
    const long e = base + (long)m * RED_T;
    if (e < g.n) { const c2 q = v[e]; acc[0] += (double)q.x * q.x + (double)q.y * q.y; }
  }
  gm_fold<1>(acc, 1, g.part + ((long)r * g.nch + blockIdx.x) * GM_PS);
}

// v_0 = r / beta
__global__ void k_gm_v0(MxGm g)
{
  const int r = blockIdx.y;
  if (gm_done(g, r)) return;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= g.n) return;
  const float s = (float)g.hs[r * GM_HS + GM_INV];
  const c2 v = g.r[(long)r * g.n + e];
  g.V[(long)r * g.n + e] = cmk(v.x * s, v.y * s);
}

// first pass: h_i = v_i^H w for i <= j (w = A z_j, held in v_{j+1})
template <int MB>
__global__ __launch_bounds__(RED_T) void k_gm_dot(MxGm g, int j)
{
  const int r = blockIdx.y;
  if (gm_done(g, r)) return;
  double acc[2 * MB];
#pragma unroll
  for (int q = 0; q < 2 * MB; q++) acc[q] = 0;
  const c2 *w = gm_v(g, j + 1, r);
  const long base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int m = 0; m < RED_E; m++) {
    const long e = base + (long)m * RED_T;
    if (e < g.n) {
      const c2 wv = w[e];
#pragma unroll
      for (int i = 0; i < MB; i++)
        if (i <= j) mx_acc(acc[2 * i], acc[2 * i + 1], wv, gm_v(g, i, r)[e]);
    }
  }
  gm_fold<2 * MB>(acc, 2 * (j + 1), g.part + ((long)r * g.nch + blockIdx.x) * GM_PS);
}

// w' = w - sum_i h_i v_i (stored in place of w), and the partials of
// |w'|^2 (slot 0) and of c_i = v_i^H w' (slots 1 + 2i, 2 + 2i)
template <int MB>
__global__ __launch_bounds__(RED_T) void k_gm_upd1(MxGm g, int j)
{
  const int r = blockIdx.y;
  if (gm_done(g, r)) return;
  const double *hs = g.hs + r * GM_HS + GM_H + 2 * (GM_MMAX + 1) * j;
  c2 nh[MB];
#pragma unroll
  for (int i = 0; i < MB; i++) nh[i] = i <= j ? cmk(-(float)hs[2 * i], -(float)hs[2 * i + 1]) : cmk(0.f, 0.f);
  double acc[2 * MB + 1];
#pragma unroll
  for (int q = 0; q < 2 * MB + 1; q++) acc[q] = 0;
  c2 *w = g.V + ((long)(j + 1) * g.R + r) * g.n;
  const long base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int m = 0; m < RED_E; m++) {
    const long e = base + (long)m * RED_T;
    if (e < g.n) {
      c2 vi[MB];
      c2 wv = w[e];
#pragma unroll
      for (int i = 0; i < MB; i++)
        if (i <= j) {
          vi[i] = gm_v(g, i, r)[e];
          wv = cfma(wv, nh[i], vi[i]);
        }
      w[e] = wv;
      acc[0] += (double)wv.x * wv.x + (double)wv.y * wv.y;
#pragma unroll
      for (int i = 0; i < MB; i++)
        if (i <= j) mx_acc(acc[1 + 2 * i], acc[2 + 2 * i], wv, vi[i]);
    }
  }
  gm_fold<2 * MB + 1>(acc, 2 * (j + 1) + 1, g.part + ((long)r * g.nch + blockIdx.x) * GM_PS);
}

// v_{j+1} = (w' - sum_i c_i v_i) / h_{j+1,j}
template <int MB>
__global__ void k_gm_upd2(MxGm g, int j)
{
  const int r = blockIdx.y;
  if (gm_done(g, r)) return;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= g.n) return;
  const double *hs = g.hs + r * GM_HS;
  const float s = (float)hs[GM_INV];
  c2 *w = g.V + ((long)(j + 1) * g.R + r) * g.n;
  c2 wv = w[e];
#pragma unroll
  for (int i = 0; i < MB; i++)
    if (i <= j) wv = cfma(wv, cmk(-(float)hs[GM_C + 2 * i], -(float)hs[GM_C + 2 * i + 1]), gm_v(g, i, r)[e]);
  w[e] = cmk(wv.x * s, wv.y * s);
}

// Scalar work of one right-hand side (one block each): sum the chunk
// partials in order, then
//   stage 0: beta = |r|, g = (beta, 0, ...), the first normalisation;
//   stage 1: the first-pass projections into Hessenberg column j;
//   stage 2: add the second-pass projections, h_{j+1,j}, the previous
//            rotations on column j, a new rotation that zeroes h_{j+1,j},
//            g, the residual estimate |g_{j+1}| / beta and the end of the
//            cycle (reached MS_TOL, m steps, the budget, or breakdown).
__global__ __launch_bounds__(RED_T) void k_gm_fin(MxGm g, int j, int stage)
{
  __shared__ double d[GM_PS];
  const int r = blockIdx.x;
  double *sc = g.sc + r * MS_N;
  if (sc[MS_DONE] != 0.0) return;
  const int nq = stage == 0 ? 1 : stage == 1 ? 2 * (j + 1) : 2 * (j + 1) + 1;
  double acc[GM_PS];
#pragma unroll
  for (int q = 0; q < GM_PS; q++) acc[q] = 0;
  for (int c = threadIdx.x; c < g.nch; c += blockDim.x) {
    const double *p = g.part + ((long)r * g.nch + c) * GM_PS;
#pragma unroll
    for (int q = 0; q < GM_PS; q++)
      if (q < nq) acc[q] += p[q];
  }
  gm_fold<GM_PS>(acc, nq, d);
  __syncthreads();
  if (threadIdx.x != 0) return;
  double *hs = g.hs + r * GM_HS;
  if (stage == 0) {
    const double beta = sqrt(d[0]);
    hs[GM_BETA] = beta;
    for (int i = 0; i < 2 * (GM_MMAX + 1); i++) hs[GM_G + i] = 0.0;
    hs[GM_G] = beta;
    sc[MS_IT] = 0.0;
    sc[MS_REL] = 1.0;
    if (beta > 0.0) hs[GM_INV] = 1.0 / beta;
    else { hs[GM_INV] = 0.0; sc[MS_DONE] = 1.0; }
    return;
  }
  double *h = hs + GM_H + 2 * (GM_MMAX + 1) * j;
  if (stage == 1) {
    for (int i = 0; i <= j; i++) { h[2 * i] = d[2 * i]; h[2 * i + 1] = d[2 * i + 1]; }
    return;
  }
  double s2 = 0.0;
  for (int i = 0; i <= j; i++) {
    const double cr = d[1 + 2 * i], ci = d[2 + 2 * i];
    h[2 * i] += cr; h[2 * i + 1] += ci;
    hs[GM_C + 2 * i] = cr; hs[GM_C + 2 * i + 1] = ci;
    s2 += cr * cr + ci * ci;
  }
  const double hn2 = d[0] - s2;
  const double hn = hn2 > 0.0 ? sqrt(hn2) : 0.0;
  hs[GM_INV] = hn > 0.0 ? 1.0 / hn : 0.0;
  // the previous rotations on column j: [a; b] <- [cs a + sn b; -conj(sn) a + cs b]
  for (int i = 0; i < j; i++) {
    const double cs = hs[GM_CS + i], sr = hs[GM_SN + 2 * i], si = hs[GM_SN + 2 * i + 1];
    const double ar = h[2 * i], ai = h[2 * i + 1], br = h[2 * i + 2], bi = h[2 * i + 3];
    h[2 * i] = cs * ar + (sr * br - si * bi);
    h[2 * i + 1] = cs * ai + (sr * bi + si * br);
    h[2 * i + 2] = -(sr * ar + si * ai) + cs * br;
    h[2 * i + 3] = -(sr * ai - si * ar) + cs * bi;
  }
  // a rotation that zeroes h_{j+1,j} = hn (real, >= 0) below a = h_jj
  const double ar = h[2 * j], ai = h[2 * j + 1];
  const double am = hypot(ar, ai);
  double cs, sr, si;
  if (am == 0.0) {
    cs = 0.0; sr = 1.0; si = 0.0;
    h[2 * j] = hn; h[2 * j + 1] = 0.0;
  } else {
    const double nu = hypot(am, hn);
    cs = am / nu;
    sr = ar / am * (hn / nu); si = ai / am * (hn / nu);
    h[2 * j] = ar / am * nu; h[2 * j + 1] = ai / am * nu;
  }
  hs[GM_CS + j] = cs; hs[GM_SN + 2 * j] = sr; hs[GM_SN + 2 * j + 1] = si;
  double *gv = hs + GM_G;
  const double gr = gv[2 * j], gi = gv[2 * j + 1];
  gv[2 * j + 2] = -(sr * gr + si * gi);  // -conj(sn) g_j
  gv[2 * j + 3] = -(sr * gi - si * gr);
  gv[2 * j] = cs * gr; gv[2 * j + 1] = cs * gi;
  const double rel = hypot(gv[2 * j + 2], gv[2 * j + 3]) / hs[GM_BETA];
  sc[MS_IT] += 1.0;
  sc[MS_REL] = rel;
  if (rel < sc[MS_TOL] || sc[MS_IT] >= g.m || sc[MS_IT] >= sc[MS_MAXIT] || !(hn > 0.0) || !(rel == rel))
    sc[MS_DONE] = 1.0;
}

// y = H^-1 g over the cycle's k = MS_IT steps (upper triangular after the
// rotations), one thread per right-hand side taking part
__global__ void k_gm_solve(MxGm g)
{
  const int r = blockIdx.x;
  if (threadIdx.x != 0 || g.sc[r * MS_N + MS_USED] == 0.0) return;
  double *hs = g.hs + r * GM_HS;
  const int k = (int)g.sc[r * MS_N + MS_IT];
  for (int i = k - 1; i >= 0; i--) {
    double sr = hs[GM_G + 2 * i], si = hs[GM_G + 2 * i + 1];
    for (int l = i + 1; l < k; l++) {
      const double *hl = hs + GM_H + 2 * (GM_MMAX + 1) * l;
      const double yr = hs[GM_Y + 2 * l], yi = hs[GM_Y + 2 * l + 1];
      sr -= hl[2 * i] * yr - hl[2 * i + 1] * yi;
      si -= hl[2 * i] * yi + hl[2 * i + 1] * yr;
    }
    const double *hi = hs + GM_H + 2 * (GM_MMAX + 1) * i;
    const double dr = hi[2 * i], di = hi[2 * i + 1];
    const double den = dr * dr + di * di;
    hs[GM_Y + 2 * i] = den > 0.0 ? (sr * dr + si * di) / den : 0.0;
    hs[GM_Y + 2 * i + 1] = den > 0.0 ? (si * dr - sr * di) / den : 0.0;
  }
}

// x = sum_{i < k} y_i z_i
__global__ void k_gm_xupd(MxGm g)
{
  const int r = blockIdx.y;
  if (g.sc[r * MS_N + MS_USED] == 0.0) return;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= g.n) return;
  const double *hs = g.hs + r * GM_HS;
  const int k = (int)g.sc[r * MS_N + MS_IT];
  c2 acc = cmk(0.f, 0.f);
  for (int i = 0; i < k; i++)
    acc = cfma(acc, cmk((float)hs[GM_Y + 2 * i], (float)hs[GM_Y + 2 * i + 1]), g.Z[((long)i * g.R + r) * g.n + e]);
  g.x[(long)r * g.n + e] = acc;
}
