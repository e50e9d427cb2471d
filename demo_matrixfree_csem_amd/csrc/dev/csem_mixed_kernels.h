// csem_mixed_kernels.h -- FP32 kernels of the batched mixed-precision
// solve (gfx950).
//
// R right-hand sides share one model; RHS r has its own frequency.  Every
// FP32 vector is R blocks of n complex floats (block r at r*n).  These
// kernels approximate the FP64 operator and preconditioner to FP32
// accuracy: the outer FP64 iterative refinement (csem_mixed.hip) computes
// the true residual with the FP64 kernels, so nothing here has to round
// like the reference.  Divisions become multiplications by reciprocals and
// expressions are reassociated freely.
#pragma once
#include "csem_common.h"
#include "csem_kernels.h"

#define MX_RMAX 4  // right-hand sides per batch

// Per right-hand side scalar slots of the inner solve (device FP64), see
// the Krylov section.
enum { MS_RHO = 0, MS_ALPHA = 2, MS_OMEGA = 4, MS_BETA = 6, MS_RS0 = 8, MS_REL = 9, MS_DONE = 10, MS_TOL = 11,
       MS_IT = 12, MS_MAXIT = 13, MS_USED = 14, MS_N = 16 };
// A right-hand side whose inner solve is done (converged, or not taking
// part) needs no operator or preconditioner output: sc == NULL (a single
// application) computes every one.
__device__ __forceinline__ bool mx_live(const double *sc, int r) { return !sc || sc[r * MS_N + MS_DONE] == 0.0; }

__device__ __forceinline__ c2 cfma(c2 acc, c2 a, c2 b)  // acc + a*b
{
  return cmk(__builtin_fmaf(a.x, b.x, __builtin_fmaf(-a.y, b.y, acc.x)),
             __builtin_fmaf(a.x, b.y, __builtin_fmaf(a.y, b.x, acc.y)));
}
__device__ __forceinline__ c2 crfma(c2 acc, float a, c2 b)  // acc + a*b, a real
{
  return cmk(__builtin_fmaf(a, b.x, acc.x), __builtin_fmaf(a, b.y, acc.y));
}

// Per right-hand side constants: iwmu = 1/(i w mu0), iomu = i w mu0.
struct MxFreq {
  c2 iwmu[MX_RMAX], iomu[MX_RMAX];
};

// ---------------------------------------------------------------- operator
// FP32 model arrays: spacings and reciprocals (z arrays on the full grid),
// conductivities on the full grid.
struct MxGeo {
  int n1, n2, n3, n3f, kofs, airbc, fvm;
  long n, lEx, lEy;
  const float *d1, *d2, *d3, *d1s, *d2s, *d3s;  // full-grid z
  const float *r1, *r2, *r3, *r1s, *r2s, *r3s;
  const float *s11, *s22, *s33;
};

// y_r = A_r x_r for every r: the curl-curl of fdfd.c:86-205 in FP32.  H' =
// curl E (without 1/(i w mu0)) is formed once per face; the per-frequency
// factor multiplies the outer curl.  Ghost planes (airwave boundary) hold
// H' as well.  The schedule is the FP64 kernel's: a block is NW waves over
// a tile of rows and columns (wave w row j = j0-1+w, lane l column
// i = i0-1+l; waves 0 and NW-1 and lanes 0 and 63 are halos) marching over
// KC planes, fully unrolled; each lane prefetches its own E of all R
// right-hand sides two planes ahead and sigma (shared by the R right-hand
// sides) one plane ahead; j neighbours come through LDS, i neighbours
// through whole-wave lane shifts; tiles are numbered columns, rows, planes
// and consecutive tiles run on one XCD.
template <int R, int NW, int KC>
__global__ __launch_bounds__(64 * NW) void k_mx_apply(MxGeo g, MxFreq fr, const c2 *x, c2 *y, const c2 *hxg,
                                                      const c2 *hyg, int nti, int ntj, int ntk, const double *sc)
{
  constexpr int OI = 62, OJ = NW - 2;
  __shared__ c2 sEx[NW][R][64], sEz[NW][R][64], sHx[NW][R][64], sHz[NW][R][64];
  __shared__ float sD[KC][4];  // 1/d3s, 1/d3, d3, d3s of the tile's planes
  const int ntiles = nti * ntj * ntk, per = (ntiles + 7) / 8;
  const int T = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (T >= ntiles) return;
  const int ti = T % nti, tj = (T / nti) % ntj, tk = T / (nti * ntj);
  const int l = threadIdx.x, w = threadIdx.y;
  const int i = ti * OI - 1 + l, j = tj * OJ - 1 + w;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3, ko = g.kofs;
  const int k0 = tk * KC, k1 = min(k0 + KC, n3 + 1);
  const bool live = i >= 0 && i <= n1 && j >= 0 && j <= n2;
  const bool out = live && l >= 1 && l <= OI && w >= 1 && w <= OJ;
  const bool hx_ok = live && j < n2, hy_ok = live && i < n1, hz_ok = hy_ok && j < n2;
  const bool iin = i >= 1 && i <= n1 - 1, jin = j >= 1 && j <= n2 - 1;
  const int ic = min(max(i, 0), n1), jc = min(max(j, 0), n2);
  const float r1s_i = g.r1s[ic], r1_i = g.r1[ic], r2s_j = g.r2s[jc], r2_j = g.r2[jc];
  // fvm cell volumes without their z factor (Ex, Ey, Ez)
  const float vx = g.d1s[ic] * g.d2[jc], vy = g.d1[ic] * g.d2s[jc], vz = g.d1[ic] * g.d2[jc];
  const long px = (long)n1 * (n2 + 1), py = (long)(n1 + 1) * n2, pz = (long)(n1 + 1) * (n2 + 1);
  const long ox = min(ic, n1 - 1) + (long)n1 * jc, oy = ic + (long)(n1 + 1) * min(jc, n2 - 1),
             oz = ic + (long)(n1 + 1) * jc;
  const long n = g.n;
  const c2 z0 = cmk(0, 0);
  // sigma on the operator grid: the full grid offset by kofs planes
  const float *s11 = g.s11 + ko * px, *s22 = g.s22 + ko * py, *s33 = g.s33 + ko * pz;
  bool rhs_on[R];  // right-hand sides still iterating
#pragma unroll
  for (int r = 0; r < R; r++) rhs_on[r] = R == 1 || mx_live(sc, r);
  struct EV { c2 x[R], y[R], z[R]; };
  struct SV { float x, y, z; };
  auto ldE = [&](int k) {
    const long kc = max(k, 0), kxy = min(kc, (long)n3), kz = min(kc, (long)n3 - 1);
    EV e;
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (!rhs_on[r]) { e.x[r] = e.y[r] = e.z[r] = z0; continue; }
      e.x[r] = x[r * n + kxy * px + ox];
      e.y[r] = x[r * n + g.lEx + kxy * py + oy];
      e.z[r] = x[r * n + g.lEx + g.lEy + kz * pz + oz];
    }
    return e;
  };
  auto ldS = [&](int k) {
    SV v;
    v.x = s11[min(k, n3) * px + ox];
    v.y = s22[min(k, n3) * py + oy];
    v.z = s33[min(k, n3 - 1) * pz + oz];
    return v;
  };
  const int wu = min(w + 1, NW - 1), wd = max(w - 1, 0);
  if (w == 0 && l < 4 * KC) {
    const int kq = min(k0 + l / 4, n3) + ko, c = l % 4;
    const float *src = c == 0 ? g.r3s : c == 1 ? g.r3 : c == 2 ? g.d3 : g.d3s;
    sD[l / 4][c] = src[kq];
  }
  EV e[3];
  SV sg[2];
  e[2] = ldE(k0 - 1);
  e[0] = ldE(k0);
  e[1] = ldE(k0 + 1);
  sg[0] = ldS(k0);
  c2 hxp[R], hyp[R];
#pragma unroll
  for (int r = 0; r < R; r++) { hxp[r] = z0; hyp[r] = z0; }
  if (k0 == 0) {
    if (g.airbc) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (hx_ok) hxp[r] = hxg[r * (long)(n1 + 1) * n2 + ic + (long)(n1 + 1) * jc];
        if (hy_ok) hyp[r] = hyg[r * (long)n1 * (n2 + 1) + ic + (long)n1 * jc];
      }
    }
  } else {
    const float r3s = g.r3s[k0 - 1 + ko];
#pragma unroll
    for (int r = 0; r < R; r++) sEz[w][r][l] = e[2].z[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
      const c2 ez_up = sEz[wu][r][l], ez = e[2].z[r], ez_r = cdpp<DPP_WAVE_SHL1>(ez);
      const c2 hx = cmk((ez_up.x - ez.x) * r2s_j - (e[0].y[r].x - e[2].y[r].x) * r3s,
                        (ez_up.y - ez.y) * r2s_j - (e[0].y[r].y - e[2].y[r].y) * r3s);
      const c2 hy = cmk((e[0].x[r].x - e[2].x[r].x) * r3s - (ez_r.x - ez.x) * r1s_i,
                        (e[0].x[r].y - e[2].x[r].y) * r3s - (ez_r.y - ez.y) * r1s_i);
      hxp[r] = hx_ok ? hx : z0;
      hyp[r] = hy_ok ? hy : z0;
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < R; r++) { sEx[w][r][l] = e[0].x[r]; sEz[w][r][l] = e[0].z[r]; }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < KC; kk++) {
    const int k = k0 + kk;
    const EV cur = e[kk % 3], nxt = e[(kk + 1) % 3];
    e[(kk + 2) % 3] = ldE(k + 2);
    sg[(kk + 1) % 2] = ldS(k + 1);
    const float r3s_k = sD[kk][0], r3_k = sD[kk][1];
    c2 hx[R], hy[R], hz[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const c2 ez_up = sEz[wu][r][l], ex_up = sEx[wu][r][l];
      const c2 ey_r = cdpp<DPP_WAVE_SHL1>(cur.y[r]), ez_r = cdpp<DPP_WAVE_SHL1>(cur.z[r]);
      const c2 ex = cur.x[r], ey = cur.y[r], ez = cur.z[r];
      const c2 hx0 = cmk((ez_up.x - ez.x) * r2s_j - (nxt.y[r].x - ey.x) * r3s_k,
                         (ez_up.y - ez.y) * r2s_j - (nxt.y[r].y - ey.y) * r3s_k);
      const c2 hy0 = cmk((nxt.x[r].x - ex.x) * r3s_k - (ez_r.x - ez.x) * r1s_i,
                         (nxt.x[r].y - ex.y) * r3s_k - (ez_r.y - ez.y) * r1s_i);
      const c2 hz0 = cmk((ey_r.x - ey.x) * r1s_i - (ex_up.x - ex.x) * r2s_j,
                         (ey_r.y - ey.y) * r1s_i - (ex_up.y - ex.y) * r2s_j);
      hx[r] = (hx_ok && k < n3) ? hx0 : z0;
      hy[r] = (hy_ok && k < n3) ? hy0 : z0;
      hz[r] = hz_ok ? hz0 : z0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
      sHx[w][r][l] = hx[r]; sHz[w][r][l] = hz[r];
      sEx[w][r][l] = nxt.x[r]; sEz[w][r][l] = nxt.z[r];
    }
    __syncthreads();
    c2 hz_dn[R], hx_dn[R], hy_l[R], hz_l[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      hz_dn[r] = sHz[wd][r][l]; hx_dn[r] = sHx[wd][r][l];
      hy_l[r] = cdpp<DPP_WAVE_SHR1>(hy[r]); hz_l[r] = cdpp<DPP_WAVE_SHR1>(hz[r]);
    }
    if (out && k < k1) {
      const SV sv = sg[kk % 2];
      const bool kin = (k >= 1 && k <= n3 - 1) || (g.airbc && k == 0);
      const float d3_k = sD[kk][2], d3s_k = sD[kk][3];
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (!rhs_on[r]) continue;
        const c2 iw = fr.iwmu[r];
        c2 *yr = y + r * n;
        if (i < n1) {  // bx
          c2 v = z0;
          if (jin && kin) {
            const c2 cu = cmk((hz[r].x - hz_dn[r].x) * r2_j - (hy[r].x - hyp[r].x) * r3_k,
                              (hz[r].y - hz_dn[r].y) * r2_j - (hy[r].y - hyp[r].y) * r3_k);
            v = csub(cmul(cu, iw), crmul(sv.x, cur.x[r]));
            if (g.fvm) v = cmul(v, crmul(vx * d3_k, fr.iomu[r]));
          }
          yr[k * px + ox] = v;
        }
        if (j < n2) {  // by
          c2 v = z0;
          if (iin && kin) {
            const c2 cu = cmk((hx[r].x - hxp[r].x) * r3_k - (hz[r].x - hz_l[r].x) * r1_i,
                              (hx[r].y - hxp[r].y) * r3_k - (hz[r].y - hz_l[r].y) * r1_i);
            v = csub(cmul(cu, iw), crmul(sv.y, cur.y[r]));
            if (g.fvm) v = cmul(v, crmul(vy * d3_k, fr.iomu[r]));
          }
          yr[g.lEx + k * py + oy] = v;
        }
        if (k < n3) {  // bz
          c2 v = z0;
          if (iin && jin) {
            const c2 cu = cmk((hy[r].x - hy_l[r].x) * r1_i - (hx[r].x - hx_dn[r].x) * r2_j,
                              (hy[r].y - hy_l[r].y) * r1_i - (hx[r].y - hx_dn[r].y) * r2_j);
            v = csub(cmul(cu, iw), crmul(sv.z, cur.z[r]));
            if (g.fvm) v = cmul(v, crmul(vz * d3s_k, fr.iomu[r]));
          }
          yr[g.lEx + g.lEy + k * pz + oz] = v;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) { hxp[r] = hx[r]; hyp[r] = hy[r]; }
  }
}

// H'z = curl E on the plane k = 0 for every r (the airwave input).
__global__ void k_mx_hz0(MxGeo g, const c2 *x, c2 *hz, const double *sc)
{
  const int n1 = g.n1, n2 = g.n2, r = blockIdx.y;
  if (!mx_live(sc, r)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n1 * n2) return;
  const int i = t % n1, j = t / n1;
  const c2 *ex = x + r * g.n, *ey = x + r * g.n + g.lEx;
  const long a = i + (long)(n1 + 1) * j, b = i + (long)n1 * j;
  const float r1s = g.r1s[i], r2s = g.r2s[j];
  hz[r * (long)n1 * n2 + t] = cmk((ey[a + 1].x - ey[a].x) * r1s - (ex[b + n1].x - ex[b].x) * r2s,
                                  (ey[a + 1].y - ey[a].y) * r1s - (ex[b + n1].y - ex[b].y) * r2s);
}

// ----------------------------------------------------------- airwave FFT
// The FP64 path's Stockham passes (radix 4, then radix 2), in FP32 on LDS.
// Stockham FFT passes over `cols` interleaved transforms of length N in
// LDS (element n of transform c at n*cols + c), twiddles tw[n] = W_N^n from
// an LDS table.  A radix-16 pass is two radix-4 stages in registers
// (16 = 4 x 4, the inner twiddles W_16^(q m1) constants), so N = 1024
// takes three passes (16, 16, 4) with one barrier each (radix 4: five).
__device__ __forceinline__ void mx_r4(c2 &u0, c2 &u1, c2 &u2, c2 &u3, int dir)
{
  const c2 a0 = cadd(u0, u2), a1 = csub(u0, u2), a2 = cadd(u1, u3), a3 = csub(u1, u3);
  const c2 bb = dir < 0 ? cmk(a3.y, -a3.x) : cmk(-a3.y, a3.x);  // -i a3 forward, +i a3 inverse
  u0 = cadd(a0, a2); u1 = cadd(a1, bb); u2 = csub(a0, a2); u3 = csub(a1, bb);
}
template <int RD>
__device__ __forceinline__ void mx_fft_pass(const c2 *a, c2 *b, int N, int Ns, int cols, const c2 *tw, int dir)
{
  const int t = threadIdx.x, nt = blockDim.x;
  const int nb = N / RD;
  for (int w = t; w < nb * cols; w += nt) {
    const int c = w % cols, j = w / cols;
    const int k = j % Ns;
    const int o = (j / Ns) * RD * Ns + k;
    const int ts = k * (N / (RD * Ns));  // twiddle stride of this butterfly
    c2 u[RD];
#pragma unroll
    for (int r = 0; r < RD; r++) {
      u[r] = a[(j + r * nb) * cols + c];
      if (r != 0 && k != 0) {
        c2 tv = tw[r * ts];
        if (dir > 0) tv.y = -tv.y;
        u[r] = cmul(u[r], tv);
      }
    }
    if (RD == 16) {
      // radix 4 over p (r = q + 4p), twiddle W_16^(q m1), radix 4 over q
#pragma unroll
      for (int q = 0; q < 4; q++) mx_r4(u[q], u[q + 4], u[q + 8], u[q + 12], dir);
      const float c1 = 0.923879532511286756f, s1 = 0.382683432365089772f, h = 0.707106781186547524f;
      const float sg = dir < 0 ? -1.f : 1.f;
      // u[q + 4 m1] *= W_16^(q m1) = (cos, sg sin)(2 pi q m1 / 16)
      const c2 W1 = cmk(c1, sg * s1), W2 = cmk(h, sg * h), W3 = cmk(s1, sg * c1), W4 = cmk(0.f, sg),
               W6 = cmk(-h, sg * h), W9 = cmk(-c1, -sg * s1);
      u[5] = cmul(u[5], W1); u[9] = cmul(u[9], W2); u[13] = cmul(u[13], W3);
      u[6] = cmul(u[6], W2); u[10] = cmul(u[10], W4); u[14] = cmul(u[14], W6);
      u[7] = cmul(u[7], W3); u[11] = cmul(u[11], W6); u[15] = cmul(u[15], W9);
#pragma unroll
      for (int m1 = 0; m1 < 4; m1++) mx_r4(u[4 * m1], u[4 * m1 + 1], u[4 * m1 + 2], u[4 * m1 + 3], dir);
      // output m = m1 + 4 m2 sits in u[4 m1 + m2]
#pragma unroll
      for (int m1 = 0; m1 < 4; m1++)
#pragma unroll
        for (int m2 = 0; m2 < 4; m2++) b[(o + (m1 + 4 * m2) * Ns) * cols + c] = u[4 * m1 + m2];
    } else if (RD == 4) {
      mx_r4(u[0], u[1], u[2], u[3], dir);
#pragma unroll
      for (int m = 0; m < 4; m++) b[(o + m * Ns) * cols + c] = u[m];
    } else {
      b[o * cols + c] = cadd(u[0], u[1]);
      b[(o + Ns) * cols + c] = csub(u[0], u[1]);
    }
  }
}
// MAXR 16: fewest passes (the column pass: one butterfly per thread per
// pass); MAXR 4: four times the butterflies per pass (a single row per
// block keeps all its threads busy)
template <int MAXR>
__device__ __forceinline__ c2 *mx_fft(c2 *a, c2 *b, int N, int cols, const c2 *tw, int dir)
{
  for (int Ns = 1; Ns < N;) {
    if (MAXR >= 16 && (N / Ns) % 16 == 0) { mx_fft_pass<16>(a, b, N, Ns, cols, tw, dir); Ns *= 16; }
    else if ((N / Ns) % 4 == 0) { mx_fft_pass<4>(a, b, N, Ns, cols, tw, dir); Ns *= 4; }
    else { mx_fft_pass<2>(a, b, N, Ns, cols, tw, dir); Ns *= 2; }
    __syncthreads();
    c2 *tmp = a; a = b; b = tmp;
  }
  return a;
}
// the twiddle table into LDS (callers sync before the first pass)
__device__ __forceinline__ void mx_tw_lds(c2 *dst, const c2 *tw, int N)
{
  for (int e = threadIdx.x; e < N; e += blockDim.x) dst[e] = tw[e];
}

struct MxAir {
  int N, n1, n2, cw;                 // cw: columns per block of the column pass
  const c2 *hz;                      // R x n1*n2
  c2 *buf, *bufx, *bufy;             // R x N*N
  const c2 *sH1, *sH2, *tw;
  const int *jx, *jy;
  const float *wx, *wy;
  const int *rows;
  int nrowx, nrowy;
  const int *gxj1, *gxj2, *gyj1, *gyj2;
  const float *gxw1, *gxw2, *gyw1, *gyw2;
  c2 *hxg, *hyg;                     // R x ghost plane
  const double *sc;                  // inner-solve scalars: right-hand sides done are skipped (NULL: none)
};

// Bilinear H'z -> uniform row, forward row transform (airbc_init_free.c:137-153).
__global__ void k_mx_air_rows(MxAir a)
{
  extern __shared__ c2 smc[];
  const int N = a.N, i2 = blockIdx.x, r = blockIdx.y, n1 = a.n1;
  if (!mx_live(a.sc, r)) return;
  c2 *A = smc, *B = smc + N, *T = smc + 2 * N;
  mx_tw_lds(T, a.tw, N);
  const c2 *h = a.hz + r * (long)n1 * a.n2;
  const int j2 = a.jy[i2];
  const float w2 = a.wy[i2];
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) {
    const int j1 = a.jx[i1];
    c2 v = cmk(0, 0);
    if (j1 >= 0 && j2 >= 0) {
      const float w1 = a.wx[i1];
      const long b0 = j1 + (long)n1 * j2;
      v = crmul((1.f - w1) * (1.f - w2), h[b0]);
      v = crfma(v, w1 * (1.f - w2), h[b0 + 1]);
      v = crfma(v, (1.f - w1) * w2, h[b0 + n1]);
      v = crfma(v, w1 * w2, h[b0 + n1 + 1]);
    }
    A[i1] = v;
  }
  __syncthreads();
  c2 *res = mx_fft<4>(A, B, N, 1, T, -1);
  c2 *out = a.buf + r * (long)N * N + (long)N * i2;
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) out[i1] = res[i1];
}

// Column pass: forward transform of cw columns, then for sH1 -> bufx and
// sH2 -> bufy the product and the inverse column transform.  The forward
// spectrum is kept in registers (16 values per thread), so LDS holds the
// two Stockham buffers and the twiddles: N*cw must equal 16*blockDim.
// (The two products in separate blocks, each repeating the forward
// transform, made one solve alone 2% faster but three concurrent solves
// 0.8% slower: the repeated transform costs throughput.)
__global__ __launch_bounds__(256) void k_mx_air_cols(MxAir a)
{
  extern __shared__ c2 smc[];
  // column groups: consecutive groups on one XCD, so the neighbouring
  // strips of a row's cache lines are read and written through its L2
  const int ng = a.N / a.cw, per = (ng + 7) / 8, g = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (g >= ng) return;
  const int N = a.N, cw = a.cw, c0 = g * cw, r = blockIdx.y;
  if (!mx_live(a.sc, r)) return;
  const long off = r * (long)N * N;
  c2 *A = smc, *B = smc + N * cw, *T = smc + 2 * N * cw;
  const int t = threadIdx.x, nt = blockDim.x;
  mx_tw_lds(T, a.tw, N);
  for (int e = t; e < N * cw; e += nt) A[e] = a.buf[off + c0 + e % cw + (long)N * (e / cw)];
  __syncthreads();
  c2 *res = mx_fft<16>(A, B, N, cw, T, -1);
  c2 S[16];
#pragma unroll
  for (int q = 0; q < 16; q++) S[q] = res[t + q * nt];
  __syncthreads();
  for (int m = 0; m < 2; m++) {
    const c2 *mul = m == 0 ? a.sH1 : a.sH2;
    c2 *dst = (m == 0 ? a.bufx : a.bufy) + off;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = t + q * nt;
      A[e] = cmul(S[q], mul[c0 + e % cw + (long)N * (e / cw)]);
    }
    __syncthreads();
    c2 *o = mx_fft<16>(A, B, N, cw, T, +1);
    for (int e = t; e < N * cw; e += nt) dst[c0 + e % cw + (long)N * (e / cw)] = o[e];
    __syncthreads();
  }
}

// Inverse transform of the rows the back-interpolation reads.
__global__ void k_mx_air_rows_inv(MxAir a)
{
  extern __shared__ c2 smc[];
  const int N = a.N, w = blockIdx.x, r = blockIdx.y;
  if (!mx_live(a.sc, r)) return;
  c2 *buf = ((w < a.nrowx) ? a.bufx : a.bufy) + r * (long)N * N;
  const long row = a.rows[w];
  c2 *A = smc, *B = smc + N, *T = smc + 2 * N;
  mx_tw_lds(T, a.tw, N);
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) A[i1] = buf[i1 + N * row];
  __syncthreads();
  c2 *res = mx_fft<4>(A, B, N, 1, T, +1);
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) buf[i1 + N * row] = res[i1];
}

// Normalise by N*N and interpolate back to the ghost planes
// (airbc_init_free.c:162-207).
__global__ void k_mx_air_ghost(MxAir a)
{
  const int N = a.N, n1 = a.n1, n2 = a.n2, r = blockIdx.y;
  if (!mx_live(a.sc, r)) return;
  const float inn = 1.f / ((float)N * (float)N);
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = (n1 + 1) * n2, ny = n1 * (n2 + 1);
  const long off = r * (long)N * N;
  if (t < nx) {
    const int i1 = t % (n1 + 1), i2 = t / (n1 + 1);
    c2 v = cmk(0, 0);
    if (i1 < n1) {
      const int j1 = a.gxj1[i1], j2 = a.gxj2[i2];
      const float w1 = a.gxw1[i1], w2 = a.gxw2[i2];
      const c2 *b = a.bufx + off + j1 + (long)N * j2;
      v = crmul((1.f - w1) * (1.f - w2), b[0]);
      v = crfma(v, w1 * (1.f - w2), b[1]);
      v = crfma(v, (1.f - w1) * w2, b[N]);
      v = crfma(v, w1 * w2, b[N + 1]);
      v = crmul(inn, v);
    }
    a.hxg[r * (long)nx + t] = v;
  } else if (t < nx + ny) {
    t -= nx;
    const int i1 = t % n1, i2 = t / n1;
    c2 v = cmk(0, 0);
    if (i2 < n2) {
      const int j1 = a.gyj1[i1], j2 = a.gyj2[i2];
      const float w1 = a.gyw1[i1], w2 = a.gyw2[i2];
      const c2 *b = a.bufy + off + j1 + (long)N * j2;
      v = crmul((1.f - w1) * (1.f - w2), b[0]);
      v = crfma(v, w1 * (1.f - w2), b[1]);
      v = crfma(v, (1.f - w1) * w2, b[N]);
      v = crfma(v, w1 * w2, b[N + 1]);
      v = crmul(inn, v);
    }
    a.hyg[r * (long)ny + t] = v;
  }
}

// ----------------------------------------------------------------- LU-SGS
// The SGS sweeps of lusgs_apply (fdfd.c:287-510) in FP32 for R right-hand
// sides, with the same slices, diagonal wavefront and skewed per-cell
// layout as the FP64 kernels.  Written as a linear combination, the cell
// update of fdfd.c:357-369 (Ey, Ez alike) is
//   v = h*(beta*b - S + cAp*E(A+) + cAm*E(A-) + cBp*E(B+) + cBm*E(B-)),
// with h = iwmu/aii, beta = i w mu0 (1/vol when fvm), S = (P1 - P2)/dA +
// (Q3 - Q4)/dB the other components' terms, and 1-D coefficients
// c*p = 1/(ds[u] d[u]), c*m = 1/(ds[u-1] d[u]) of the component's own
// axes A and B (aii = (cAp+cAm+cBp+cBm)*iwmu - sigma).  The pre-pass folds
// everything but the new neighbours into T = beta*b - S + own (forward: the
// + neighbours are 0; backward: the - neighbours hold forward values), so
// the wavefront computes v = h*(T + cf*prev + cs*nbv): cf the fast axis'
// coefficient, cs the slow axis', prev the row's previous cell, nbv the
// neighbour row.
//
// Per-cell records are stored per right-hand side (r-major) in two planes
// of complex floats: T, written by each pre-pass, and h = iwmu/aii, which
// depends only on the frequency and is written once for a slot's
// frequencies (k_mx_hrec); a sweep step loads both, each wave's step two
// contiguous runs.
// Each slice of component C is a block of A + B + 1 diagonals of AP rows
// (AP = A rounded up to whole waves): diagonal d = s + f (2..A+B) at
// position d - 1, row s at s - 1.  The first and last diagonals and every
// slot off the slice's cells hold zero records, so a sweep step reads its
// wave's records from one wave-uniform base plus a fixed lane offset, and a
// lane on a slot off the slice computes exactly 0.
struct MxPad {
  int AP[3];
  long SL[3];    // records per slice
  long base[3];  // component offsets
  long n;        // records per right-hand side
};

struct MxLu {
  MxGeo g;
  SkewGeo sk;    // A, B, ns per component
  MxPad pk;
  const c2 *b;   // R x n right-hand sides
  c2 *y;         // R x n results
  c2 *W;         // R x fE state on the full grid (Ex | Ey | Ez)
  long fE, fEx, fEy;
  c2 *T;         // T per record slot, R x pk.n (every pre-pass)
  c2 *h;         // h = iwmu/aii per record slot, R x pk.n (once per frequency)
  c2 *V;         // a sweep's values per record slot, R x pk.n
  const float *c1m, *c1p, *c2m, *c2p, *c3m, *c3p;  // own-axis coefficients (z on the full grid)
  MxFreq fr;
  const double *sc;  // inner-solve scalars: right-hand sides done are skipped (NULL: none)
};

// Tile (x, y, z) of a pre-pass over component C launched as a 1-D grid of
// 8 * ceil(tiles / 8) blocks: tiles are numbered x fastest (Ex: slices,
// Ey/Ez: columns), then y, then z, and consecutive tiles run on one XCD
// (block b on XCD b mod 8), so tiles that share halo rows and planes share
// that XCD's L2.  False past the last tile.  (The copy k_mx_unskew keeps a
// 3-D grid: its tiles that share cache lines are the anti-diagonal
// neighbours (y - 1, z + 1), a multiple of 8 blocks apart in that order.)
template <class TL, int C>
__device__ __forceinline__ bool mx_tile_xyz(const SkewGeo &sk, int &x, int &y, int &z)
{
  const int A = sk.A[C], B = sk.B[C], ns = sk.ns[C];
  const int gx = (C == 0) ? (ns + TL::TQ - 1) / TL::TQ : (B + TL::TF - 1) / TL::TF;
  const int gy = (C == 0) ? (B + TL::TF - 1) / TL::TF : (A + TL::TS - 1) / TL::TS;
  const int gz = (C == 0) ? (A + TL::TS - 1) / TL::TS : (ns + TL::TQ - 1) / TL::TQ;
  const int nt = gx * gy * gz, per = (nt + 7) / 8;
  const int T = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (T >= nt) return false;
  x = T % gx; y = (T / gx) % gy; z = T / (gx * gy);
  return true;
}

#ifndef PRE_TQ
#define PRE_TQ 2  // slices per Ey/Ez pre-pass tile (the other components' planes are shared)
#endif
// Pre-pass tiles: TQ slices x TS rows x TF columns of a component's slices,
// read in the natural order (x fastest, so the W and b reads are whole
// rows) and written in the skewed order through LDS, so each diagonal's
// run of the tile is one contiguous store.  Ex: slices run along x, so a
// tile spans many slices; backward (which reads the W of all three
// components) 32 slices x 8 rows x 4 columns, forward (b only) 16 x 8 x 8,
// whose diagonal runs are twice as long (measured: 13% faster forward at
// three right-hand sides, 2% slower backward).  Ey, Ez: a tile is
// 2 slices x 8 rows x 64 columns.
template <int Q, int S, int F> struct MxTile3 { static constexpr int TQ = Q, TS = S, TF = F; };
template <int C, bool FWD> struct MxPreTile : MxTile3<PRE_TQ, 16 / PRE_TQ, 64> {};
template <> struct MxPreTile<0, false> : MxTile3<32, 8, 4> {};
template <> struct MxPreTile<0, true> : MxTile3<16, 8, 8> {};
// The candidates the autotuner times per model and batch size (V = 0 is
// the tile above).  Every candidate holds 1024 cells (four per thread) and,
// for Ex, a multiple of 256 / (TQ TF) rows, so the LDS arrays and the
// thread mapping are the same; a cell's record does not depend on the tile.
#define MX_PRE_NV 3
template <int C, bool FWD, int V> struct MxPreTileV : MxPreTile<C, FWD> {};
template <bool FWD> struct MxPreTileV<0, FWD, 1> : MxTile3<FWD ? 32 : 16, 8, FWD ? 4 : 8> {};
template <bool FWD> struct MxPreTileV<0, FWD, 2> : MxTile3<8, 8, 16> {};
template <int C, bool FWD> struct MxPreTileV<C, FWD, 1> : MxTile3<1, 16, 64> {};
template <int C, bool FWD> struct MxPreTileV<C, FWD, 2> : MxTile3<4, 4, 64> {};
// The copy from V (k_mx_unskew) reads diagonal runs and writes rows: its Ex
// tile trades row width for run length the other way, 16 x 8 x 8, or
// 8 x 8 x 16 for one right-hand side (measured 10% faster there, slower at
// three).
template <int C, int R> struct MxCopyTile : MxTile3<1, 16, 64> {};
template <int R> struct MxCopyTile<0, R> : MxTile3<(R == 1 ? 8 : 16), 8, (R == 1 ? 16 : 8)> {};

// W value of component C at (k, j, i) for right-hand side r.
#define MW0(r, k, j, i) a.W[(r) * a.fE + id1(n1, n2, k, j, i)]
#define MW1(r, k, j, i) a.W[(r) * a.fE + a.fEx + id2(n1, n2, k, j, i)]
#define MW2(r, k, j, i) a.W[(r) * a.fE + a.fEx + a.fEy + id3(n1, n2, k, j, i)]

// Pre-pass of sweep (C, FWD) for every r: the record T of a tile of cells.
// Each cell goes to LDS at its position in the tile's skewed order, with
// its global record index, so the store phase writes the records in runs.
template <int C, bool FWD, int R, int V>
__global__ __launch_bounds__(256) void k_mx_pre(MxLu a)
{
  using TL = MxPreTileV<C, FWD, V>;
  static_assert(TL::TQ * TL::TS * TL::TF == 1024 && (C != 0 || 256 % (TL::TQ * TL::TF) == 0), "pre-pass tile");
  constexpr int TQ = TL::TQ, TS = TL::TS, TF = TL::TF, NC = TQ * TS * TF;
  __shared__ c2 tile[R][NC];
  __shared__ int pidx[NC];  // global record index of each slot, -1 off the slices
  const int n1 = a.g.n1, n2 = a.g.n2, n3 = a.g.n3f;
  const int A = a.sk.A[C], B = a.sk.B[C], ns = a.sk.ns[C];
  const int t = threadIdx.x;
  int q0, s0, f0, bx, by, bz;
  if (!mx_tile_xyz<TL, C>(a.sk, bx, by, bz)) return;
  if (C == 0) { q0 = bx * TQ; f0 = 1 + by * TF; s0 = 1 + bz * TS; }
  else { f0 = 1 + bx * TF; s0 = 1 + by * TS; q0 = bz * TQ; }
  const float *r1 = a.g.r1, *r2 = a.g.r2, *r3 = a.g.r3, *r1s = a.g.r1s, *r2s = a.g.r2s, *r3s = a.g.r3s;
  const c2 z0 = cmk(0, 0);
  // compute: lanes along x
#pragma unroll
  for (int it = 0; it < NC / 256; it++) {
    int lq, ls, lf;
    if (C == 0) { lq = t % TQ; lf = (t / TQ) % TF; ls = t / (TQ * TF) + it * (256 / (TQ * TF)); }
    else { const int u = t + it * 256; lf = u % TF; ls = (u / TF) % TS; lq = u / (TS * TF); }
    const int q = q0 + lq, s = s0 + ls, f = f0 + lf;
    // the cell's slot: its tile's skewed order (1-based local diagonal dl)
    const int dl = ls + lf + 2;
    const int cell = lq * TS * TF + sk_off(dl, TS, TF) + (ls + 1 - max(1, dl - TF));
    if (q >= ns || s > A || f > B) {
      pidx[cell] = -1;
      continue;
    }
    pidx[cell] = (int)(a.pk.base[C] + q * a.pk.SL[C] + (long)(s + f - 1) * a.pk.AP[C] + (s - 1));
    int k, j, i;
    lu_kji<C>(q, s, f, k, j, i);
    const int kop = k - a.g.kofs;
    const long xo = (C == 0) ? id1(n1, n2, kop, j, i) : (C == 1) ? a.g.lEx + id2(n1, n2, kop, j, i)
                                                                   : a.g.lEx + a.g.lEy + id3(n1, n2, kop, j, i);
    const float vol = a.g.d1s[i] * a.g.d2[j] * a.g.d3[k];
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (R > 1 && !mx_live(a.sc, r)) continue;
      c2 S = z0, own = z0;
      if (C == 0) {
        if (!FWD) {  // Ey, Ez backward
          const c2 P1 = csub(MW1(r, k, j, i + 1), MW1(r, k, j, i)), P2 = csub(MW1(r, k, j - 1, i + 1), MW1(r, k, j - 1, i));
          const c2 Q3 = csub(MW2(r, k, j, i + 1), MW2(r, k, j, i)), Q4 = csub(MW2(r, k - 1, j, i + 1), MW2(r, k - 1, j, i));
          S = crmul(r1s[i], cadd(crmul(r2[j], csub(P1, P2)), crmul(r3[k], csub(Q3, Q4))));
          own = crfma(crmul(a.c2m[j], MW0(r, k, j - 1, i)), a.c3m[k], MW0(r, k - 1, j, i));
        }
      } else if (C == 1) {
        const c2 Q3 = csub(MW0(r, k, j + 1, i), MW0(r, k, j, i)), Q4 = csub(MW0(r, k, j + 1, i - 1), MW0(r, k, j, i - 1));
        c2 PP = z0;
        if (!FWD) PP = csub(csub(MW2(r, k, j + 1, i), MW2(r, k, j, i)), csub(MW2(r, k - 1, j + 1, i), MW2(r, k - 1, j, i)));
        S = crmul(r2s[j], cadd(crmul(r3[k], PP), crmul(r1[i], csub(Q3, Q4))));
        if (!FWD) own = crfma(crmul(a.c3m[k], MW1(r, k - 1, j, i)), a.c1m[i], MW1(r, k, j, i - 1));
      } else {
        const c2 P1 = csub(MW0(r, k + 1, j, i), MW0(r, k, j, i)), P2 = csub(MW0(r, k + 1, j, i - 1), MW0(r, k, j, i - 1));
        const c2 Q3 = csub(MW1(r, k + 1, j, i), MW1(r, k, j, i)), Q4 = csub(MW1(r, k + 1, j - 1, i), MW1(r, k, j - 1, i));
        S = crmul(r3s[k], cadd(crmul(r1[i], csub(P1, P2)), crmul(r2[j], csub(Q3, Q4))));
        if (!FWD) own = crfma(crmul(a.c1m[i], MW2(r, k, j, i - 1)), a.c2m[j], MW2(r, k, j - 1, i));
      }
      c2 bv = z0;
      if (kop >= 0) bv = a.b[r * a.g.n + xo];
      const c2 bb = a.g.fvm ? crmul(1.f / vol, bv) : cmul(a.fr.iomu[r], bv);
      tile[r][cell] = cadd(csub(bb, S), own);
    }
  }
  __syncthreads();
  // store: the tile's records in the skewed order (per slice, by diagonal,
  // rows contiguous within a diagonal)
#pragma unroll
  for (int it = 0; it < NC / 256; it++) {
    const int e = t + it * 256;
    const int p = pidx[e];
    if (p < 0) continue;
#pragma unroll
    for (int r = 0; r < R; r++)
      if (R == 1 || mx_live(a.sc, r)) a.T[r * a.pk.n + p] = tile[r][e];
  }
}

// The per-frequency half of the records of component C for r < R: h =
// iwmu/aii of every swept cell, aii = c iwmu - sigma with c the sum of the
// cell's own-axis coefficients (the fvm factor scales aij and aii alike and
// cancels).  One thread per record slot of a slice (coalesced stores); the
// slots off the slice keep their zeros.
template <int C>
__global__ __launch_bounds__(256) void k_mx_hrec(MxLu a, int R)
{
  const int A = a.sk.A[C], B = a.sk.B[C], AP = a.pk.AP[C];
  const int q = blockIdx.y;
  const long u = (long)blockIdx.x * 256 + threadIdx.x;
  if (u >= a.pk.SL[C]) return;
  const int s = (int)(u % AP) + 1, f = (int)(u / AP) + 1 - s;  // slot (d - 1) AP + s - 1, d = s + f
  if (s > A || f < 1 || f > B) return;
  int k, j, i;
  lu_kji<C>(q, s, f, k, j, i);
  const int n1 = a.g.n1, n2 = a.g.n2;
  c2 cs;
  if (C == 0) cs = cmk(a.c2p[j] + a.c2m[j] + a.c3p[k] + a.c3m[k], a.g.s11[id1(n1, n2, k, j, i)]);
  else if (C == 1) cs = cmk(a.c3p[k] + a.c3m[k] + a.c1p[i] + a.c1m[i], a.g.s22[id2(n1, n2, k, j, i)]);
  else cs = cmk(a.c1p[i] + a.c1m[i] + a.c2p[j] + a.c2m[j], a.g.s33[id3(n1, n2, k, j, i)]);
  const long p = a.pk.base[C] + (long)q * a.pk.SL[C] + u;
  for (int r = 0; r < R; r++) {
    const c2 iw = a.fr.iwmu[r];
    const c2 aii = cmk(cs.x * iw.x - cs.y, cs.x * iw.y);
    const float den = 1.f / (aii.x * aii.x + aii.y * aii.y);
    a.h[r * a.pk.n + p] = cmk((iw.x * aii.x + iw.y * aii.y) * den, (iw.y * aii.x - iw.x * aii.y) * den);
  }
}

// The values a sweep (C, FWD) left in V (one per record slot, each wave's
// step a contiguous run) copied to W on the full grid and, backward, to y
// on the operator grid, for every r: a tile of MxCopyTile cells is read in
// the skewed order through LDS (each diagonal's run of the tile
// contiguous) and written in the natural order (x fastest).  The last
// backward sweep (Ex) writes y only.
template <int C, bool FWD, int R>
__global__ __launch_bounds__(256) void k_mx_unskew(MxLu a)
{
  using TL = MxCopyTile<C, R>;
  constexpr int TQ = TL::TQ, TS = TL::TS, TF = TL::TF, NC = TQ * TS * TF;
  __shared__ c2 tile[R][NC];
  __shared__ int pidx[NC];  // record slot of each LDS slot, -1 off the slices
  const int n1 = a.g.n1, n2 = a.g.n2;
  const int A = a.sk.A[C], B = a.sk.B[C], ns = a.sk.ns[C], AP = a.pk.AP[C];
  const int t = threadIdx.x;
  int q0, s0, f0;
  if (C == 0) { q0 = blockIdx.x * TQ; f0 = 1 + blockIdx.y * TF; s0 = 1 + blockIdx.z * TS; }
  else { f0 = 1 + blockIdx.x * TF; s0 = 1 + blockIdx.y * TS; q0 = blockIdx.z * TQ; }
  // this thread's cells in the natural order (lanes along x) and their LDS slots
  int cell[NC / 256];
#pragma unroll
  for (int it = 0; it < NC / 256; it++) {
    int lq, ls, lf;
    if (C == 0) { lq = t % TQ; lf = (t / TQ) % TF; ls = t / (TQ * TF) + it * (256 / (TQ * TF)); }
    else { lq = 0; lf = t % TF; ls = t / TF + it * (256 / TF); }
    const int q = q0 + lq, s = s0 + ls, f = f0 + lf;
    const int dl = ls + lf + 2;
    cell[it] = lq * TS * TF + sk_off(dl, TS, TF) + (ls + 1 - max(1, dl - TF));
    pidx[cell[it]] = (q >= ns || s > A || f > B) ? -1
                     : (int)(a.pk.base[C] + q * a.pk.SL[C] + (long)(s + f - 1) * AP + (s - 1));
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NC / 256; it++) {
    const int e = t + it * 256;
    const int p = pidx[e];
    if (p < 0) continue;
#pragma unroll
    for (int r = 0; r < R; r++)
      if (R == 1 || mx_live(a.sc, r)) tile[r][e] = a.V[r * a.pk.n + p];
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NC / 256; it++) {
    if (pidx[cell[it]] < 0) continue;
    int lq, ls, lf;
    if (C == 0) { lq = t % TQ; lf = (t / TQ) % TF; ls = t / (TQ * TF) + it * (256 / (TQ * TF)); }
    else { lq = 0; lf = t % TF; ls = t / TF + it * (256 / TF); }
    int k, j, i;
    lu_kji<C>(q0 + lq, s0 + ls, f0 + lf, k, j, i);
    const long wo = (C == 0) ? id1(n1, n2, k, j, i) : (C == 1) ? a.fEx + id2(n1, n2, k, j, i)
                                                               : a.fEx + a.fEy + id3(n1, n2, k, j, i);
    const int kop = k - a.g.kofs;
    const long xo = (C == 0) ? id1(n1, n2, kop, j, i) : (C == 1) ? a.g.lEx + id2(n1, n2, kop, j, i)
                                                                   : a.g.lEx + a.g.lEy + id3(n1, n2, kop, j, i);
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (R > 1 && !mx_live(a.sc, r)) continue;
      const c2 v = tile[r][cell[it]];
      if (FWD || C != 0) a.W[r * a.fE + wo] = v;
      if (!FWD && kop >= 0) a.y[r * a.g.n + xo] = v;
    }
  }
}

// One colour of a red-black ordered symmetric Gauss-Seidel sweep of
// component C for every r: the splitting and cell update of the LU-SGS
// above (same h, same other-component terms S for the direction, same
// in-plane coefficients), with the cells of each 2-D slice visited red
// ((s + f) even) before black in the forward sweep and black before red in
// the backward one instead of lexicographically.  A cell of one colour
// couples in-plane only to the other colour, so each colour is one fully
// parallel pass: forward red x = h (beta b - S) (its neighbours come later),
// forward black and both backward colours x = h (beta b - S + sum of the
// four in-plane neighbours' current values).  Cells outside the swept
// range keep W = 0; the last pass (backward red) writes the zero rows of
// the operator-grid output.
template <int C, bool FWD, int COLOR, int R>
__global__ __launch_bounds__(256) void k_mx_rb(MxLu a)
{
  const int n1 = a.g.n1, n2 = a.g.n2, n3 = a.g.n3f;
  const int nx = (C == 0) ? n1 : n1 + 1, ny = (C == 1) ? n2 : n2 + 1, nz = (C == 2) ? n3 : n3 + 1;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= nx || j >= ny || k >= nz) return;
  int s, f;
  bool sw;
  if (C == 0) { s = k; f = j; sw = k >= 1 && k <= n3 - 1 && j >= 1 && j <= n2 - 1; }
  else if (C == 1) { s = k; f = i; sw = k >= 1 && k <= n3 - 1 && i >= 1 && i <= n1 - 1; }
  else { s = j; f = i; sw = j >= 1 && j <= n2 - 1 && i >= 1 && i <= n1 - 1; }
  const int kop = k - a.g.kofs;
  const long xo = (C == 0) ? id1(n1, n2, kop, j, i) : (C == 1) ? a.g.lEx + id2(n1, n2, kop, j, i)
                                                                 : a.g.lEx + a.g.lEy + id3(n1, n2, kop, j, i);
  const c2 z0 = cmk(0, 0);
  if (!sw) {
    if (!FWD && COLOR == 0 && kop >= 0) {
#pragma unroll
      for (int r = 0; r < R; r++) a.y[r * a.g.n + xo] = z0;
    }
    return;
  }
  if (((s + f) & 1) != COLOR) return;
  const float *r1 = a.g.r1, *r2 = a.g.r2, *r3 = a.g.r3, *r1s = a.g.r1s, *r2s = a.g.r2s, *r3s = a.g.r3s;
  const float vol = a.g.d1s[i] * a.g.d2[j] * a.g.d3[k];
  float c, sig, cam, cap, cbm, cbp;  // in-plane coefficients: fast axis -, +; slow axis -, +
  long wc, fa, sa;                   // own W index, fast- and slow-axis strides
  if (C == 0) {
    cam = a.c2m[j]; cap = a.c2p[j]; cbm = a.c3m[k]; cbp = a.c3p[k];
    sig = a.g.s11[id1(n1, n2, k, j, i)];
    wc = id1(n1, n2, k, j, i); fa = n1; sa = (long)n1 * (n2 + 1);
  } else if (C == 1) {
    cam = a.c1m[i]; cap = a.c1p[i]; cbm = a.c3m[k]; cbp = a.c3p[k];
    sig = a.g.s22[id2(n1, n2, k, j, i)];
    wc = a.fEx + id2(n1, n2, k, j, i); fa = 1; sa = (long)(n1 + 1) * n2;
  } else {
    cam = a.c1m[i]; cap = a.c1p[i]; cbm = a.c2m[j]; cbp = a.c2p[j];
    sig = a.g.s33[id3(n1, n2, k, j, i)];
    wc = a.fEx + a.fEy + id3(n1, n2, k, j, i); fa = 1; sa = n1 + 1;
  }
  c = cam + cap + cbm + cbp;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const c2 *Wr = a.W + r * a.fE;
    c2 S = z0;
    if (C == 0) {
      if (!FWD) {  // Ey, Ez (backward values)
        const c2 P1 = csub(MW1(r, k, j, i + 1), MW1(r, k, j, i)), P2 = csub(MW1(r, k, j - 1, i + 1), MW1(r, k, j - 1, i));
        const c2 Q3 = csub(MW2(r, k, j, i + 1), MW2(r, k, j, i)), Q4 = csub(MW2(r, k - 1, j, i + 1), MW2(r, k - 1, j, i));
        S = crmul(r1s[i], cadd(crmul(r2[j], csub(P1, P2)), crmul(r3[k], csub(Q3, Q4))));
      }
    } else if (C == 1) {
      const c2 Q3 = csub(MW0(r, k, j + 1, i), MW0(r, k, j, i)), Q4 = csub(MW0(r, k, j + 1, i - 1), MW0(r, k, j, i - 1));
      c2 PP = z0;
      if (!FWD) PP = csub(csub(MW2(r, k, j + 1, i), MW2(r, k, j, i)), csub(MW2(r, k - 1, j + 1, i), MW2(r, k - 1, j, i)));
      S = crmul(r2s[j], cadd(crmul(r3[k], PP), crmul(r1[i], csub(Q3, Q4))));
    } else {
      const c2 P1 = csub(MW0(r, k + 1, j, i), MW0(r, k, j, i)), P2 = csub(MW0(r, k + 1, j, i - 1), MW0(r, k, j, i - 1));
      const c2 Q3 = csub(MW1(r, k + 1, j, i), MW1(r, k, j, i)), Q4 = csub(MW1(r, k + 1, j - 1, i), MW1(r, k, j - 1, i));
      S = crmul(r3s[k], cadd(crmul(r1[i], csub(P1, P2)), crmul(r2[j], csub(Q3, Q4))));
    }
    c2 bv = z0;
    if (kop >= 0) bv = a.b[r * a.g.n + xo];
    const c2 bb = a.g.fvm ? crmul(1.f / vol, bv) : cmul(a.fr.iomu[r], bv);
    c2 u = csub(bb, S);
    if (!(FWD && COLOR == 0)) {
      const c2 nf = crfma(crmul(cam, Wr[wc - fa]), cap, Wr[wc + fa]);
      const c2 ns = crfma(crmul(cbm, Wr[wc - sa]), cbp, Wr[wc + sa]);
      u = cadd(u, cadd(nf, ns));
    }
    // h = iwmu/aii, aii = c iwmu - sigma (the fvm factor cancels)
    const c2 iw = a.fr.iwmu[r];
    const c2 aii = cmk(c * iw.x - sig, c * iw.y);
    const float den = 1.f / (aii.x * aii.x + aii.y * aii.y);
    const c2 h = cmk((iw.x * aii.x + iw.y * aii.y) * den, (iw.y * aii.x - iw.x * aii.y) * den);
    const c2 v = cmul(h, u);
    a.W[r * a.fE + wc] = v;
    if (!FWD && kop >= 0) a.y[r * a.g.n + xo] = v;
  }
}
#undef MW0
#undef MW1
#undef MW2

typedef unsigned v2u __attribute__((ext_vector_type(2)));
#define MX_OOB ((int)0x80000000)  // a raw-buffer offset past every buffer here (all shorter than 1 GiB)
// raw buffer resource over bytes [p, p + len) (gfx9 format word)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mx_rsrc(const void *p, long len)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)len, 0x00020000);
}

// Zero the cells (u, v), u < nu, v < nv, of a plane at base + u su + v sv
// whose u lies outside [ulo, uhi] or v outside [1, nv - 2].
__device__ __forceinline__ void mx_zero_frame(c2 *base, int nu, int nv, long su, long sv, int ulo, int uhi, int t,
                                              int nt)
{
  const c2 z0 = cmk(0, 0);
  const int lo = max(ulo, 0), hi = min(uhi, nu - 1);
  for (int u = 0; u < nu; u++) {
    if (u >= lo && u <= hi) u = hi;  // whole rows outside [lo, hi] only
    else
      for (int v = t; v < nv; v += nt) base[u * su + v * sv] = z0;
  }
  for (int w = t; w < 2 * (hi - lo + 1); w += nt) base[(lo + (w >> 1)) * su + ((w & 1) ? nv - 1 : 0) * sv] = z0;
}

// Sweep (C, FWD) of one slice for one right-hand side: block (slice, r),
// one thread per row s = 1 + t.  Step tau of a wave handles diagonal
// d = s + f; within the wave, the neighbour row's new value from the
// previous diagonal comes from the neighbouring lane by a whole-wave lane
// shift, and the shift's edge lane keeps the value handed over by the
// neighbouring wave, so a diagonal costs no LDS round trip and no barrier.
// Consecutive waves run MXS_L diagonals apart (in the sweep's direction the
// leading wave is the one whose rows come first): the row at a wave's edge
// goes through a small LDS ring, written at the step it is computed and read
// MXS_PD steps before it is used, and a barrier every MXS_BP steps orders
// every such write before its read (MXS_L >= MXS_BP + MXS_PD - 1; a slot is
// rewritten only after MXS_RING steps).  All waves run the same number of
// steps; a lane on a slot off the slice (before its row starts, after it
// ends, rows past A, the lag at either end) reads a zero record (or none:
// a raw-buffer load past the slice's records reads zero) and so computes 0.
// Where the values go.  The W and y cells of a diagonal are a row apart
// each, so a wave's step stored in the grid's order touches a cache line
// per lane; and a wave's vector-memory counter retires loads and stores in
// issue order, so a wave waiting for its prefetched record also waits for
// its earlier stores.  Ey and Ez (rows along x, contiguous in the grid): the
// computing waves put each step's values into an LDS staging area (two
// halves of MXS_BP steps), and one more wave of the block stores each half
// to W and y while the others fill the other half, eight lanes per row, so
// a store instruction writes eight rows' runs of MXS_BP values; a store
// wave that issued as many instructions as a computing wave would hold
// every barrier, so its offsets are precomputed per lane and blended.  Ex (rows a
// plane apart, runs only across slices): the values go to V in the records'
// slot order, one contiguous run per wave and step, and k_mx_unskew copies
// them to W and y through tiles that span slices.  Every value a step needs comes from registers
// filled steps ahead: the records (MXS_P steps), the fast-axis coefficient
// of the lane's column (MXS_PC steps, from an LDS table padded with zeros
// so that no index is clamped) and the edge value (MXS_PD steps).  Raw-buffer
// accesses keep the memory-operation count exact, so the compiler waits for
// each record just in time; the loop runs MXS_U steps per iteration.
// records in flight (steps ahead); with more the compiler's count of the
// loop's outstanding memory operations turns into a full wait at the back edge
#ifndef MXS_P
#define MXS_P 12
#endif
#ifndef MXS_U
#define MXS_U 96   // steps per loop iteration (the loop leaves at the first barrier past the last step)
#endif
#ifndef MXS_MASK
#define MXS_MASK 1                       // lanes off the slice skip their loads and stores
#endif
#define MXS_BP 8                         // steps between barriers
#define MXS_PD 4                         // steps ahead an edge value is read
#define MXS_PC 4                         // steps ahead a coefficient is read
#define MXS_L (MXS_BP + MXS_PD - 1)      // steps between consecutive waves
#define MXS_RING 32                      // ring slots per wave edge
static_assert(MXS_RING > MXS_L - MXS_PD + MXS_BP, "edge ring slot rewritten before it is read");
static_assert(MXS_U % MXS_RING == 0 && MXS_U % MXS_BP == 0 && MXS_U % MXS_P == 0 && MXS_U % MXS_PD == 0 &&
                  MXS_U % MXS_PC == 0 && (MXS_U / MXS_BP) % 2 == 0,
              "sweep unroll (an even number of staging halves per loop iteration)");
// LDS of a sweep block of nw computing waves: per wave an edge ring and
// MXS_RING + 64 slots the other lanes' ring stores land in, for Ey and Ez
// the store wave's staging (two halves of nw x MXS_BP steps x MXS_SP), then
// the coefficient table
__host__ __device__ __forceinline__ int mxs_table_off(int AP, int nw) { return AP + (nw - 1) * MXS_L + MXS_U + MXS_PC + 4; }
__host__ __device__ __forceinline__ int mxs_table_len(int A, int B, int AP, int nw)
{
  return mxs_table_off(AP, nw) + A + B + (nw - 1) * MXS_L + MXS_U + MXS_PC + 4;
}
// store waves of a sweep block of nw computing waves (Ey, Ez): one (a
// second one, on the SIMD a block of at most two computing waves leaves
// free, measured no faster)
__host__ __device__ constexpr int mxs_store_waves(int) { return 1; }
#define MXS_SP 72  // staging slots per step: 64 rows, padded so that the store wave's reads (eight
                   // steps of eight rows) fall in at most two-way bank conflicts
__host__ __device__ __forceinline__ long mxs_lds_bytes(int A, int B, int AP, int c)
{
  const int nw = AP / 64;
  return (long)nw * (2 * MXS_RING + 64) * 8 + (c != 0 ? 2L * nw * MXS_BP * MXS_SP * 8 : 0) +
         (long)mxs_table_len(A, B, AP, nw) * 4;
}
// the computing waves of a block of AP rows, and for Ey, Ez the store waves
template <int C, bool FWD, int NW>
__global__ __launch_bounds__(64 * (NW + 1)) void k_mx_sgs(MxLu a)
{
  extern __shared__ c2 lds[];
  const int nslice = a.sk.ns[C];
  const int per = (nslice + 7) / 8;
  const int q = (blockIdx.x % 8) * per + blockIdx.x / 8;  // consecutive slices on one XCD
  const int r = blockIdx.y;
  if (q >= nslice || !mx_live(a.sc, r)) return;
  const int n1 = a.g.n1, n2 = a.g.n2;
  constexpr bool SW = (C != 0);  // a store wave writes W and y (Ex: V and k_mx_unskew)
  constexpr int NSW = SW ? mxs_store_waves(NW) : 0;
  constexpr int nt = 64 * (NW + NSW), AP = 64 * NW, lagmax = (NW - 1) * MXS_L;
  constexpr int P = MXS_P;
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int A = a.sk.A[C], B = a.sk.B[C];
  const int nst = A + B - 1;
  const int OFF = mxs_table_off(AP, NW), TL = mxs_table_len(A, B, AP, NW);
  const int s = 1 + t, sc = min(s, A);
  c2 *ring = lds;                                  // [NW][MXS_RING]: the edge row a wave receives
  c2 *junk = lds + NW * MXS_RING;                  // [NW][MXS_RING + 64]
  c2 *stg = junk + NW * (MXS_RING + 64);           // SW: [2][NW][MXS_BP][MXS_SP], the store wave's halves
  float *cft = (float *)(stg + (SW ? 2 * NW * MXS_BP * MXS_SP : 0));  // fast-axis coefficient of column f at f + OFF
  {
    const float *g0 = (C == 0) ? (FWD ? a.c2m : a.c2p) : (FWD ? a.c1m : a.c1p);
    for (int u = t; u < TL; u += nt) {
      const int f = u - OFF;
      cft[u] = (f >= 1 && f <= B) ? g0[f] : 0.f;
    }
    for (int u = t; u < NW * MXS_RING; u += nt) ring[u] = cmk(0, 0);
  }
  const float *gs = (C == 2) ? (FWD ? a.c2m : a.c2p) : (FWD ? a.c3m : a.c3p);
  const float cs = gs[sc];
  // this slice's records (T and h planes) as raw buffers (the hardware
  // checks the vector offset: a load past the end reads zero, a store past
  // the end is dropped)
  const long rb0 = r * a.pk.n + a.pk.base[C] + (long)q * a.pk.SL[C];
  const __amdgpu_buffer_rsrc_t Tb = mx_rsrc(a.T + rb0, a.pk.SL[C] * (long)sizeof(c2));
  const __amdgpu_buffer_rsrc_t Hb = mx_rsrc(a.h + rb0, a.pk.SL[C] * (long)sizeof(c2));
  // the values go to V in the records' slot order: a wave's step is one
  // contiguous run (a lane off the slice stores its exact 0 into a slot no
  // reader uses, a lagged step's run lies outside the slice and is dropped)
  const __amdgpu_buffer_rsrc_t Vb = mx_rsrc(a.V + r * a.pk.n + a.pk.base[C] + (long)q * a.pk.SL[C],
                                            a.pk.SL[C] * (long)sizeof(c2));
  // lag of this wave and the edges: forward, wave w receives row 64w from
  // wave w-1's lane 63; backward, row 64w+65 from wave w+1's lane 0
  const int lag = (FWD ? wv : NW - 1 - wv) * MXS_L;
  const bool send = FWD ? (lane == 63 && wv < NW - 1) : (lane == 0 && wv > 0);
  // every lane stores to the ring each step, the sender into the
  // neighbour's edge ring and the others into this wave's junk slots; the
  // ring read address is held per lane (a lane-varying zero keeps it in a
  // vector register), so each access is an address plus an immediate offset
  const int z = __builtin_amdgcn_mbcnt_lo(0u, 0u);
  c2 *ring_in = ring + wv * MXS_RING + z;
  c2 *ring_st = send ? ring + (FWD ? wv + 1 : wv - 1) * MXS_RING : junk + wv * (MXS_RING + 64) + lane;
  c2 *stg_w = stg + wv * MXS_BP * MXS_SP + lane;  // SW: this lane's staging slot of step 0 of half 0
  const int nex = ((nst + lagmax + MXS_BP - 1) / MXS_BP) * MXS_BP;  // steps run
  const int ntk = ((nex + MXS_U - 1) / MXS_U) * MXS_U;
  // backward: the operator-grid cells of this slice outside the sweep range
  // are zero rows (every thread of the block takes a share)
  auto zero_frame = [&]() {
    const int kl = 1 - a.g.kofs, kh = a.g.n3f - 1 - a.g.kofs;  // operator planes of the sweep range
    c2 *y0 = a.y + r * a.g.n;
    if (C == 0) mx_zero_frame(y0 + q, a.g.n3 + 1, n2 + 1, (long)n1 * (n2 + 1), n1, kl, kh, t, nt);
    else if (C == 1) mx_zero_frame(y0 + a.g.lEx + (long)(n1 + 1) * q, a.g.n3 + 1, n1 + 1, (long)(n1 + 1) * n2, 1, kl, kh, t, nt);
    else if (q - a.g.kofs >= 0)
      mx_zero_frame(y0 + a.g.lEx + a.g.lEy + (long)(n1 + 1) * (n2 + 1) * (q - a.g.kofs), n2 + 1, n1 + 1, n1 + 1, 1, 1,
                    n2 - 1, t, nt);
  };
  if (SW && wv >= NW) {
    // the store wave: after the barrier that ends half h (steps MXS_BP h ..
    // MXS_BP h + MXS_BP - 1), store that half's values to W and (backward)
    // y while the computing waves fill the other half.  Lane l takes step
    // l % 8 of row 8 k + l / 8 of a wave: eight lanes write one row's run
    // of consecutive columns (cell (q, s, f) is at qb + s SS + f).
    const long PL = (C == 1) ? (long)(n1 + 1) * n2 : (long)(n1 + 1) * (n2 + 1);
    const long SS = (C == 2) ? (long)(n1 + 1) : PL;
    const long qb = (C == 1) ? (long)(n1 + 1) * q : PL * q;
    const long wofs = (C == 1) ? a.fEx : a.fEx + a.fEy;
    const long xofs = (C == 1) ? a.g.lEx : a.g.lEx + a.g.lEy;
    const long wlen = (C == 1) ? a.fEy : a.fE - a.fEx - a.fEy;
    const long ylen = (C == 1) ? a.g.lEy : a.g.n - a.g.lEx - a.g.lEy;
    const __amdgpu_buffer_rsrc_t Wb = mx_rsrc(a.W + r * a.fE + wofs, wlen * (long)sizeof(c2));
    const __amdgpu_buffer_rsrc_t Yb = mx_rsrc(a.y + r * a.g.n + xofs, ylen * (long)sizeof(c2));
    const int u = lane % MXS_BP, rl = lane / MXS_BP;
    static_assert(MXS_BP == 8, "eight lanes per row, eight rows per store");
    // per lane and (w, k), for half 0: f - 1 of the lane's cell (past any
    // column on a row past A) and the cell's byte offsets in W and y (1 GiB,
    // past the buffer whatever half, where the row has no y plane); half h
    // moves f by MXS_BP, forward or backward.  This wave issues about as many
    // instructions per half as a computing wave, so a store costs a handful
    // of 32-bit operations.
    // Store wave sw0 takes the groups of eight rows g = NSW j + sw0 (wave
    // g / 8, rows 8 (g % 8) .. of it).
    constexpr int NG = NW * 8 / (NSW > 0 ? NSW : 1);
    const int sw0 = wv - NW;
    int fm1[NG], wo[NG], yo[NG];
#pragma unroll
    for (int j = 0; j < NG; j++) {
      const int g = NSW * j + sw0, w = g / 8, k = g % 8;
      const int lw = (FWD ? w : NW - 1 - w) * MXS_L;
      const int d = FWD ? 2 + u - lw : A + B - u + lw;
      const int sw = 1 + 64 * w + 8 * k + rl, scw = min(sw, A), f = d - sw;
      const long cell = qb + scw * SS + f;
      const int kopw = ((C == 2) ? q : scw) - a.g.kofs;
      fm1[j] = sw <= A ? f - 1 : (1 << 30);
      wo[j] = (int)cell * (int)sizeof(c2);
      yo[j] = (FWD || kopw < 0) ? (1 << 30) : (int)(cell - (long)a.g.kofs * PL) * (int)sizeof(c2);
    }
    constexpr int FST = FWD ? MXS_BP : -MXS_BP, OST = FST * (int)sizeof(c2);
    __syncthreads();  // the computing waves' first barrier
    const int nhalf = nex / MXS_BP;
    for (int h = 0; h < nhalf; h++) {
      __syncthreads();
      const c2 *src = stg + (h & 1) * NW * MXS_BP * MXS_SP + u * MXS_SP + rl;
#pragma unroll
      for (int j = 0; j < NG; j++) {
        const int g = NSW * j + sw0;
        // a lane off the slice stores past the buffers (a blend, not a
        // branch: every lane issues both stores)
        const int m = -(int)((unsigned)(fm1[j] + h * FST) < (unsigned)B);
        const v2u v = __builtin_bit_cast(v2u, src[(g / 8) * MXS_BP * MXS_SP + 8 * (g % 8)]);
        __builtin_amdgcn_raw_buffer_store_b64(v, Wb, ((wo[j] + h * OST) & m) | (MX_OOB & ~m), 0, 0);
        if (!FWD)
          __builtin_amdgcn_raw_buffer_store_b64(v, Yb, ((yo[j] + h * OST) & m) | (MX_OOB & ~m), 0, 0);
      }
    }
    if (!FWD) zero_frame();
    return;
  }
  // table index of this lane's column at step tau: ib + tau forward, ib - tau backward
  const int ib = (FWD ? 2 - lag : A + B + lag) - s + OFF;
  // byte offset of this lane's slot at step tau in the T, h and V planes:
  // rv0 + tau * RSTEP
  constexpr int RSTEP = (FWD ? 1 : -1) * AP * (int)sizeof(c2);
  const int rv0 = (s - 1) * (int)sizeof(c2) + (FWD ? 1 - lag : nst + lag) * AP * (int)sizeof(c2);
  auto rec_at = [&](int vofs) -> float4 {
    const c2 T = __builtin_bit_cast(c2, __builtin_amdgcn_raw_buffer_load_b64(Tb, vofs, 0, 0));
    const c2 h = __builtin_bit_cast(c2, __builtin_amdgcn_raw_buffer_load_b64(Hb, vofs, 0, 0));
    return make_float4(T.x, T.y, h.x, h.y);
  };
  // a lane whose cell at step tau is off the slice (f - 1 = d - s - 1 outside
  // [0, B) or s > A) neither loads nor stores: its offset is past the
  // buffers (the loads read 0, the store is dropped), the same offset for
  // the records and V
  const int fm0 = s <= A ? (FWD ? 1 - lag : A + B + lag - 1) - s : (1 << 30);  // f - 1 at step 0
  auto rofs = [&](int tau, int ofs) {
    return (MXS_MASK == 0 || (unsigned)(FWD ? fm0 + tau : fm0 - tau) < (unsigned)B) ? ofs : MX_OOB;
  };
  float4 rec[P];  // rec[tau % P] holds step tau
  int rof[P];     // and its record offset (MX_OOB when masked)
#pragma unroll
  for (int u = 0; u < P; u++) {
    rof[u] = rofs(u, rv0 + u * RSTEP);
    rec[u] = rec_at(rof[u]);
  }
  __syncthreads();
  float cfr[MXS_PC];  // coefficient of step tau in cfr[tau % MXS_PC]
  c2 edg[MXS_PD];     // edge value of step tau in edg[tau % MXS_PD]
#pragma unroll
  for (int u = 0; u < MXS_PC; u++) cfr[u] = cft[FWD ? ib + u : ib - u];
#pragma unroll
  for (int u = 0; u < MXS_PD; u++) edg[u] = ring_in[(u - MXS_L - 1) & (MXS_RING - 1)];
  c2 prev = cmk(0, 0);
  for (int tk0 = 0; tk0 < ntk; tk0 += MXS_U) {
    const float *ct = cft + (FWD ? ib + tk0 : ib - tk0 - MXS_U - MXS_PC);
    const int rv = rv0 + (tk0 + P) * RSTEP;  // record offset of step tk0 + P
#pragma unroll
    for (int uu = 0; uu < MXS_U; uu++) {
      const float4 rc = rec[uu % P];
      const float cfv = cfr[uu % MXS_PC];
      // the neighbour row's value on the previous diagonal
      const c2 nbv = FWD ? cdpp_keep<DPP_WAVE_SHR1>(edg[uu % MXS_PD], prev)
                         : cdpp_keep<DPP_WAVE_SHL1>(edg[uu % MXS_PD], prev);
      const float ux = __builtin_fmaf(cfv, prev.x, rc.x), uy = __builtin_fmaf(cfv, prev.y, rc.y);
      const float wx = __builtin_fmaf(cs, nbv.x, ux), wy = __builtin_fmaf(cs, nbv.y, uy);
      prev = cmk(rc.z * wx - rc.w * wy, rc.z * wy + rc.w * wx);
      if (SW)  // MXS_U / MXS_BP is even: the half of step tk0 + uu is that of uu
        stg_w[((uu / MXS_BP) & 1) * NW * MXS_BP * MXS_SP + (uu % MXS_BP) * MXS_SP] = prev;
      else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, prev), Vb, rof[uu % P], 0, 0);
      ring_st[uu & (MXS_RING - 1)] = prev;
      if (uu % MXS_BP == MXS_BP - 1) {
        __syncthreads();
        if (tk0 + uu + 1 >= nex) goto swept;  // the same step for every wave
      }
      edg[uu % MXS_PD] = ring_in[(uu + MXS_PD - MXS_L - 1) & (MXS_RING - 1)];
      cfr[uu % MXS_PC] = ct[FWD ? uu + MXS_PC : MXS_U - uu];
      rof[uu % P] = rofs(tk0 + uu + P, rv + uu * RSTEP);
      rec[uu % P] = rec_at(rof[uu % P]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
swept:
  if (!FWD) zero_frame();
}

// ---------------------------------------------------------------- Krylov
// Batched F-BiCGStab (solver.c:93-178) on FP32 vectors.  Dot products
// accumulate in FP64 in the deterministic chunk order of the FP64 path (per
// right-hand side: chunk, then thread, then the block fold); the scalars
// and the convergence state live in FP64 on the device, so an iteration
// can be replayed from a graph without the host.  Per-RHS scalar slots:
// MS_USED marks the right-hand sides that take part in the current inner
// solve (their correction is added to the FP64 solution).


struct MxKry {
  long n;
  int nch;                 // reduction chunks per vector
  c2 *r, *r0, *p, *pp, *v, *s, *ss, *t, *x;  // R x n each
  double *part;            // R x nch x 4 partial sums
  double *sc;              // R x MS_N scalars
};

__device__ __forceinline__ bool mx_active(const MxKry &k, int r) { return k.sc[r * MS_N + MS_DONE] == 0.0; }

// a*conj(b) accumulated in double
__device__ __forceinline__ void mx_acc(double &re, double &im, c2 a, c2 b)
{
  re += (double)a.x * b.x + (double)a.y * b.y;
  im += (double)a.y * b.x - (double)a.x * b.y;
}
// Fold 4 doubles per thread over the block (blockDim == RED_T) into
// out[0..3], thread 0 writing: each value by block_fold's halving tree
// (v[t] += v[t+128], v[t] += v[t+64], then v[t] += v[t+h] in wave 0 for
// h = 32..1), the four trees sharing their barriers.
__device__ __forceinline__ void mx_fold4(double *acc, double *out, double *)
{
  __shared__ double f4[4][RED_T];
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; q++) f4[q][t] = acc[q];
  __syncthreads();
  if (t < 128) {
#pragma unroll
    for (int q = 0; q < 4; q++) f4[q][t] = f4[q][t] + f4[q][t + 128];
  }
  __syncthreads();
  if (t < 64) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      double v = f4[q][t] + f4[q][t + 64];
      for (int h = 32; h >= 1; h >>= 1) v = v + __shfl_down(v, h, 64);
      if (t == 0) out[q] = v;
    }
  }
  __syncthreads();
}

// init: r0 = p = r (r holds the inner right-hand side), x = 0;
// partials of <r, r0> and <r, r>
__global__ __launch_bounds__(RED_T) void k_mx_k_init(MxKry k)
{
  __shared__ double sh[RED_T];
  const int r = blockIdx.y;
  if (!mx_active(k, r)) return;
  double acc[4] = {0, 0, 0, 0};
  const long o = r * k.n, base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int m = 0; m < RED_E; m++) {
    const long e = base + (long)m * RED_T;
    if (e < k.n) {
      const c2 rv = k.r[o + e];
      k.r0[o + e] = rv; k.p[o + e] = rv; k.x[o + e] = cmk(0, 0);
      mx_acc(acc[0], acc[1], rv, rv);
      mx_acc(acc[2], acc[3], rv, rv);
    }
  }
  mx_fold4(acc, k.part + ((long)r * k.nch + blockIdx.x) * 4, sh);
}
// partials of <v, r0> (slot 0,1) and nothing else
__global__ __launch_bounds__(RED_T) void k_mx_k_dot_vr0(MxKry k)
{
  __shared__ double sh[RED_T];
  const int r = blockIdx.y;
  if (!mx_active(k, r)) return;
  double acc[4] = {0, 0, 0, 0};
  const long o = r * k.n, base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int m = 0; m < RED_E; m++) {
    const long e = base + (long)m * RED_T;
    if (e < k.n) mx_acc(acc[0], acc[1], k.v[o + e], k.r0[o + e]);
  }
  mx_fold4(acc, k.part + ((long)r * k.nch + blockIdx.x) * 4, sh);
}
// s = r - alpha v
__global__ void k_mx_k_s(MxKry k)
{
  const int r = blockIdx.y;
  if (!mx_active(k, r)) return;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= k.n) return;
  const double *sc = k.sc + r * MS_N;
  const c2 al = cmk((float)sc[MS_ALPHA], (float)sc[MS_ALPHA + 1]);
  const long o = r * k.n + e;
  const c2 v = k.v[o], rv = k.r[o];
  k.s[o] = cmk(rv.x - (al.x * v.x - al.y * v.y), rv.y - (al.x * v.y + al.y * v.x));
}
// partials of <s, t> (0,1) and <t, t> (2,3)
__global__ __launch_bounds__(RED_T) void k_mx_k_dot_st(MxKry k)
{
  __shared__ double sh[RED_T];
  const int r = blockIdx.y;
  if (!mx_active(k, r)) return;
  double acc[4] = {0, 0, 0, 0};
  const long o = r * k.n, base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int m = 0; m < RED_E; m++) {
    const long e = base + (long)m * RED_T;
    if (e < k.n) {
      const c2 tv = k.t[o + e];
      mx_acc(acc[0], acc[1], k.s[o + e], tv);
      mx_acc(acc[2], acc[3], tv, tv);
    }
  }
  mx_fold4(acc, k.part + ((long)r * k.nch + blockIdx.x) * 4, sh);
}
// x += alpha pp + omega ss; r = s - omega t; partials of <r, r0>, <r, r>
__global__ __launch_bounds__(RED_T) void k_mx_k_xr(MxKry k)
{
  __shared__ double sh[RED_T];
  const int r = blockIdx.y;
  if (!mx_active(k, r)) return;
  const double *sc = k.sc + r * MS_N;
  const c2 al = cmk((float)sc[MS_ALPHA], (float)sc[MS_ALPHA + 1]);
  const c2 om = cmk((float)sc[MS_OMEGA], (float)sc[MS_OMEGA + 1]);
  double acc[4] = {0, 0, 0, 0};
  const long o = r * k.n, base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int m = 0; m < RED_E; m++) {
    const long e = base + (long)m * RED_T;
    if (e < k.n) {
      const long q = o + e;
      k.x[q] = cfma(cfma(k.x[q], al, k.pp[q]), om, k.ss[q]);
      const c2 t = k.t[q], sv = k.s[q];
      const c2 rv = cmk(sv.x - (om.x * t.x - om.y * t.y), sv.y - (om.x * t.y + om.y * t.x));
      k.r[q] = rv;
      mx_acc(acc[0], acc[1], rv, k.r0[q]);
      mx_acc(acc[2], acc[3], rv, rv);
    }
  }
  mx_fold4(acc, k.part + ((long)r * k.nch + blockIdx.x) * 4, sh);
}
// p = r + beta (p - omega v)
__global__ void k_mx_k_p(MxKry k)
{
  const int r = blockIdx.y;
  if (!mx_active(k, r)) return;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= k.n) return;
  const double *sc = k.sc + r * MS_N;
  const c2 be = cmk((float)sc[MS_BETA], (float)sc[MS_BETA + 1]);
  const c2 om = cmk((float)sc[MS_OMEGA], (float)sc[MS_OMEGA + 1]);
  const long q = r * k.n + e;
  const c2 v = k.v[q], pv = k.p[q];
  const c2 d = cmk(pv.x - (om.x * v.x - om.y * v.y), pv.y - (om.x * v.y + om.y * v.x));
  k.p[q] = cfma(k.r[q], be, d);
}

// Sum the chunk partials of right-hand side r (fixed order) into out[0..3].
__device__ __forceinline__ void mx_sum_parts(const MxKry &k, int r, double *out, double *sh)
{
  double acc[4] = {0, 0, 0, 0};
  for (int c = threadIdx.x; c < k.nch; c += blockDim.x) {
    const double *pp = k.part + ((long)r * k.nch + c) * 4;
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] += pp[q];
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const double v = block_fold(acc[q], sh);
    if (threadIdx.x == 0) out[q] = v;
  }
}
__device__ __forceinline__ void zdiv_d(double ar, double ai, double br, double bi, double *o)
{
  const double den = br * br + bi * bi;
  o[0] = (ar * br + ai * bi) / den;
  o[1] = (ai * br - ar * bi) / den;
}
// Scalar steps, one block of RED_T threads per right-hand side.
// stage 0: after init (rho, rs0); 1: alpha; 2: omega; 3: rho_new, beta,
// relres, convergence.
__global__ __launch_bounds__(RED_T) void k_mx_k_fin(MxKry k, int stage)
{
  __shared__ double sh[RED_T];
  __shared__ double d[4];
  const int r = blockIdx.x;
  double *sc = k.sc + r * MS_N;
  if (sc[MS_DONE] != 0.0) return;
  mx_sum_parts(k, r, d, sh);
  __syncthreads();
  if (threadIdx.x != 0) return;
  double o[2];
  if (stage == 0) {
    sc[MS_RHO] = d[0]; sc[MS_RHO + 1] = d[1]; sc[MS_RS0] = d[2]; sc[MS_REL] = 1.0; sc[MS_IT] = 0;
    if (!(d[2] > 0.0)) sc[MS_DONE] = 1.0;
  } else if (stage == 1) {  // alpha = rho / <v, r0>
    zdiv_d(sc[MS_RHO], sc[MS_RHO + 1], d[0], d[1], o);
    sc[MS_ALPHA] = o[0]; sc[MS_ALPHA + 1] = o[1];
  } else if (stage == 2) {  // omega = <s, t> / Re<t, t>
    sc[MS_OMEGA] = d[0] / d[2]; sc[MS_OMEGA + 1] = d[1] / d[2];
  } else {
    const double rel = sqrt(d[2] / sc[MS_RS0]);
    sc[MS_REL] = rel;
    sc[MS_IT] += 1.0;
    // beta = (rho_new / rho_old) * alpha / omega
    double q1[2], q2[2];
    zdiv_d(d[0], d[1], sc[MS_RHO], sc[MS_RHO + 1], q1);
    const double ar = q1[0] * sc[MS_ALPHA] - q1[1] * sc[MS_ALPHA + 1];
    const double ai = q1[0] * sc[MS_ALPHA + 1] + q1[1] * sc[MS_ALPHA];
    zdiv_d(ar, ai, sc[MS_OMEGA], sc[MS_OMEGA + 1], q2);
    sc[MS_BETA] = q2[0]; sc[MS_BETA + 1] = q2[1];
    sc[MS_RHO] = d[0]; sc[MS_RHO + 1] = d[1];
    if (rel < sc[MS_TOL] || sc[MS_IT] >= sc[MS_MAXIT] || !(rel == rel)) sc[MS_DONE] = 1.0;
  }
}
