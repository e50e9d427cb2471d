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
                                                      const c2 *hyg, int nti, int ntj, int ntk)
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
  struct EV { c2 x[R], y[R], z[R]; };
  struct SV { float x, y, z; };
  auto ldE = [&](int k) {
    const long kc = max(k, 0), kxy = min(kc, (long)n3), kz = min(kc, (long)n3 - 1);
    EV e;
#pragma unroll
    for (int r = 0; r < R; r++) {
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
__global__ void k_mx_hz0(MxGeo g, const c2 *x, c2 *hz)
{
  const int n1 = g.n1, n2 = g.n2, r = blockIdx.y;
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
template <int RD>
__device__ __forceinline__ void mx_fft_pass(const c2 *a, c2 *b, int N, int Ns, int cols, const c2 *tw, int dir)
{
  const int t = threadIdx.x, nt = blockDim.x;
  const int nb = N / RD;
  for (int w = t; w < nb * cols; w += nt) {
    const int c = w % cols, j = w / cols;
    const int k = j % Ns;
    const int o = (j / Ns) * RD * Ns + k;
    c2 u[RD];
#pragma unroll
    for (int r = 0; r < RD; r++) {
      u[r] = a[(j + r * nb) * cols + c];
      if (r != 0 && k != 0) {
        c2 tv = tw[(long)r * k * (N / (RD * Ns))];
        if (dir > 0) tv.y = -tv.y;
        u[r] = cmul(u[r], tv);
      }
    }
    if (RD == 4) {
      const c2 a0 = cadd(u[0], u[2]), a1 = csub(u[0], u[2]), a2 = cadd(u[1], u[RD - 1]), a3 = csub(u[1], u[RD - 1]);
      const c2 bb = dir < 0 ? cmk(a3.y, -a3.x) : cmk(-a3.y, a3.x);
      b[o * cols + c] = cadd(a0, a2);
      b[(o + Ns) * cols + c] = cadd(a1, bb);
      b[(o + 2 * Ns) * cols + c] = csub(a0, a2);
      b[(o + 3 * Ns) * cols + c] = csub(a1, bb);
    } else {
      b[o * cols + c] = cadd(u[0], u[1]);
      b[(o + Ns) * cols + c] = csub(u[0], u[1]);
    }
  }
}
__device__ __forceinline__ c2 *mx_fft(c2 *a, c2 *b, int N, int cols, const c2 *tw, int dir)
{
  for (int Ns = 1; Ns < N;) {
    if ((N / Ns) % 4 == 0) { mx_fft_pass<4>(a, b, N, Ns, cols, tw, dir); Ns *= 4; }
    else { mx_fft_pass<2>(a, b, N, Ns, cols, tw, dir); Ns *= 2; }
    __syncthreads();
    c2 *tmp = a; a = b; b = tmp;
  }
  return a;
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
};

// Bilinear H'z -> uniform row, forward row transform (airbc_init_free.c:137-153).
__global__ void k_mx_air_rows(MxAir a)
{
  extern __shared__ c2 smc[];
  const int N = a.N, i2 = blockIdx.x, r = blockIdx.y, n1 = a.n1;
  c2 *A = smc, *B = smc + N;
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
  c2 *res = mx_fft(A, B, N, 1, a.tw, -1);
  c2 *out = a.buf + r * (long)N * N + (long)N * i2;
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) out[i1] = res[i1];
}

// Column pass: forward transform of cw columns, then for sH1 -> bufx and
// sH2 -> bufy the product and the inverse column transform.  The forward
// spectrum is kept in registers (16 values per thread), so LDS holds only
// the two Stockham buffers: N*cw must equal 16*blockDim.
__global__ __launch_bounds__(256) void k_mx_air_cols(MxAir a)
{
  extern __shared__ c2 smc[];
  // column groups: consecutive groups on one XCD, so the neighbouring
  // strips of a row's cache lines are read and written through its L2
  const int ng = a.N / a.cw, per = (ng + 7) / 8, g = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (g >= ng) return;
  const int N = a.N, cw = a.cw, c0 = g * cw, r = blockIdx.y;
  const long off = r * (long)N * N;
  c2 *A = smc, *B = smc + N * cw;
  const int t = threadIdx.x, nt = blockDim.x;
  for (int e = t; e < N * cw; e += nt) A[e] = a.buf[off + c0 + e % cw + (long)N * (e / cw)];
  __syncthreads();
  c2 *res = mx_fft(A, B, N, cw, a.tw, -1);
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
    c2 *o = mx_fft(A, B, N, cw, a.tw, +1);
    for (int e = t; e < N * cw; e += nt) dst[c0 + e % cw + (long)N * (e / cw)] = o[e];
    __syncthreads();
  }
}

// Inverse transform of the rows the back-interpolation reads.
__global__ void k_mx_air_rows_inv(MxAir a)
{
  extern __shared__ c2 smc[];
  const int N = a.N, w = blockIdx.x, r = blockIdx.y;
  c2 *buf = ((w < a.nrowx) ? a.bufx : a.bufy) + r * (long)N * N;
  const long row = a.rows[w];
  c2 *A = smc, *B = smc + N;
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) A[i1] = buf[i1 + N * row];
  __syncthreads();
  c2 *res = mx_fft(A, B, N, 1, a.tw, +1);
  for (int i1 = threadIdx.x; i1 < N; i1 += blockDim.x) buf[i1 + N * row] = res[i1];
}

// Normalise by N*N and interpolate back to the ghost planes
// (airbc_init_free.c:162-207).
__global__ void k_mx_air_ghost(MxAir a)
{
  const int N = a.N, n1 = a.n1, n2 = a.n2, r = blockIdx.y;
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
// Per-cell records (T, h) are stored per right-hand side (r-major) in the
// skewed order, one float4 at r*nsk + base[C] + q*A*B + p: h (z, w) is
// refreshed when the frequencies change, T (x, y) by each pre-pass, and a
// sweep step loads both with one 16-byte load.

struct MxLu {
  MxGeo g;
  SkewGeo sk;
  const c2 *b;   // R x n right-hand sides
  c2 *y;         // R x n results
  c2 *W;         // R x fE state on the full grid (Ex | Ey | Ez)
  long fE, fEx, fEy;
  float4 *th;    // (T, h = iwmu/aii) per swept cell, skewed Ex | Ey | Ez, per r
  long nsk;      // swept cells of all components (records per r)
  c2 *trash;     // stores of lanes off the diagonal, 2 * SGS_MAXT * MX_RMAX
  const float *c1m, *c1p, *c2m, *c2p, *c3m, *c3p;  // own-axis coefficients (z on the full grid)
  MxFreq fr;
};

// h of every swept cell of component C for every r.
template <int C, int R>
__global__ __launch_bounds__(256) void k_mx_diag(MxLu a)
{
  const int A = a.sk.A[C], B = a.sk.B[C], q = blockIdx.y;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= A * B) return;
  int lo = 2, hi = A + B;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sk_off(mid, A, B) <= p) lo = mid; else hi = mid - 1;
  }
  const int s = max(1, lo - B) + (p - sk_off(lo, A, B)), f = lo - s;
  int k, j, i;
  lu_kji<C>(q, s, f, k, j, i);
  const int n1 = a.g.n1, n2 = a.g.n2;
  float c, sig;
  if (C == 0) { c = a.c2p[j] + a.c2m[j] + a.c3p[k] + a.c3m[k]; sig = a.g.s11[id1(n1, n2, k, j, i)]; }
  else if (C == 1) { c = a.c3p[k] + a.c3m[k] + a.c1p[i] + a.c1m[i]; sig = a.g.s22[id2(n1, n2, k, j, i)]; }
  else { c = a.c1p[i] + a.c1m[i] + a.c2p[j] + a.c2m[j]; sig = a.g.s33[id3(n1, n2, k, j, i)]; }
  const long o = a.sk.base[C] + (long)q * A * B + p;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const c2 iw = a.fr.iwmu[r];
    // h = iwmu/aii (the fvm factor scales aij and aii alike and cancels)
    const c2 aii = cmk(c * iw.x - sig, c * iw.y);
    const float den = 1.f / (aii.x * aii.x + aii.y * aii.y);
    ((c2 *)a.th)[2 * (r * a.nsk + o) + 1] =
        cmk((iw.x * aii.x + iw.y * aii.y) * den, (iw.y * aii.x - iw.x * aii.y) * den);
  }
}

// Pre-pass tiles: TQ slices x TS rows x TF columns of a component's slices,
// read in the natural order (x fastest, so the W and b reads are whole
// rows) and written in the skewed order through LDS, so each diagonal's
// run of the tile is one contiguous store.  Ex: slices run along x, so a
// tile spans 16 slices; Ey, Ez: a tile is 16 rows x 64 columns of one slice.
template <int C> struct MxPreTile { static constexpr int TQ = 1, TS = 16, TF = 64; };
template <> struct MxPreTile<0> { static constexpr int TQ = 16, TS = 8, TF = 8; };

// W value of component C at (k, j, i) for right-hand side r.
#define MW0(r, k, j, i) a.W[(r) * a.fE + id1(n1, n2, k, j, i)]
#define MW1(r, k, j, i) a.W[(r) * a.fE + a.fEx + id2(n1, n2, k, j, i)]
#define MW2(r, k, j, i) a.W[(r) * a.fE + a.fEx + a.fEy + id3(n1, n2, k, j, i)]

// Pre-pass of sweep (C, FWD) for every r: T of a tile of cells.  Each cell
// goes to LDS at its position in the tile's skewed order, with its global
// record index, so the store phase is a straight copy.
template <int C, bool FWD, int R>
__global__ __launch_bounds__(256) void k_mx_pre(MxLu a)
{
  using TL = MxPreTile<C>;
  constexpr int TQ = TL::TQ, TS = TL::TS, TF = TL::TF, NC = TQ * TS * TF;
  __shared__ c2 tile[R][NC];
  __shared__ int pidx[NC];  // global record index of each slot, -1 off the slices
  const int n1 = a.g.n1, n2 = a.g.n2, n3 = a.g.n3f;
  const int A = a.sk.A[C], B = a.sk.B[C], ns = a.sk.ns[C];
  const long AB = (long)A * B;
  const int t = threadIdx.x;
  int q0, s0, f0;
  if (C == 0) { q0 = blockIdx.x * TQ; f0 = 1 + blockIdx.y * TF; s0 = 1 + blockIdx.z * TS; }
  else { f0 = 1 + blockIdx.x * TF; s0 = 1 + blockIdx.y * TS; q0 = blockIdx.z; }
  const float *r1 = a.g.r1, *r2 = a.g.r2, *r3 = a.g.r3, *r1s = a.g.r1s, *r2s = a.g.r2s, *r3s = a.g.r3s;
  const c2 z0 = cmk(0, 0);
  // compute: lanes along x
#pragma unroll
  for (int it = 0; it < NC / 256; it++) {
    int lq, ls, lf;
    if (C == 0) { lq = t % TQ; lf = (t / TQ) % TF; ls = t / (TQ * TF) + it * (256 / (TQ * TF)); }
    else { lq = 0; lf = t % TF; ls = t / TF + it * (256 / TF); }
    const int q = q0 + lq, s = s0 + ls, f = f0 + lf;
    // the cell's slot: its tile's skewed order (1-based local diagonal dl)
    const int dl = ls + lf + 2;
    const int cell = lq * TS * TF + sk_off(dl, TS, TF) + (ls + 1 - max(1, dl - TF));
    if (q >= ns || s > A || f > B) {
      pidx[cell] = -1;
      continue;
    }
    {
      const int d = s + f;
      pidx[cell] = (int)(a.sk.base[C] + q * AB + sk_off(d, A, B) + (s - max(1, d - B)));
    }
    int k, j, i;
    lu_kji<C>(q, s, f, k, j, i);
    const int kop = k - a.g.kofs;
    const long xo = (C == 0) ? id1(n1, n2, kop, j, i) : (C == 1) ? a.g.lEx + id2(n1, n2, kop, j, i)
                                                                   : a.g.lEx + a.g.lEy + id3(n1, n2, kop, j, i);
    const float vol = a.g.d1s[i] * a.g.d2[j] * a.g.d3[k];
#pragma unroll
    for (int r = 0; r < R; r++) {
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
  // store: the tile's cells in the skewed order (per slice, by diagonal,
  // rows contiguous within a diagonal)
#pragma unroll
  for (int it = 0; it < NC / 256; it++) {
    const int e = t + it * 256;
    const int p = pidx[e];
    if (p < 0) continue;
#pragma unroll
    for (int r = 0; r < R; r++) ((c2 *)a.th)[2 * (r * a.nsk + p)] = tile[r][e];
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

// Sweep (C, FWD) of one slice for one right-hand side: block (slice, r),
// one thread per row s = 1 + t.  Within a wave the new value of the
// neighbour row from the previous diagonal comes from the neighbouring
// lane by a whole-wave lane shift, so a diagonal costs no LDS round trip
// and no barrier.  Consecutive waves run MXS_L diagonals apart (in the
// sweep's direction the leading wave is the one whose rows come first):
// the row at a wave's edge is passed through a small LDS ring, written one
// diagonal after it is computed and read MXS_L diagonals later, and a
// barrier every MXS_L steps orders every such write before its read (a
// ring slot is reused only after 2 MXS_L steps).  All waves run the same
// number of steps (the sweep's diagonals plus the lags); a row off its
// diagonal computes on a clamped cell and stores to a trash row.  The
// records of the diagonals ahead stream into a register ring MXS_P deep
// (one load and at most two stores per diagonal keep that within the 63
// memory operations a wave can have outstanding); the compiler waits for
// every outstanding load at the loop's back edge, so the loop runs MXS_U
// steps per iteration.
#ifndef MXS_P
#define MXS_P 16
#endif
#ifndef MXS_U
#define MXS_U 32
#endif
#define MXS_L 8      // steps between consecutive waves; a barrier every MXS_L steps
#define MXS_RING 32  // ring slots per wave edge (> 2 MXS_L, a power of two)
template <int C, bool FWD>
__global__ __launch_bounds__(SGS_MAXT) void k_mx_sgs(MxLu a)
{
  extern __shared__ c2 lds[];
  const int nslice = a.sk.ns[C];
  const int per = (nslice + 7) / 8;
  const int q = (blockIdx.x % 8) * per + blockIdx.x / 8;  // consecutive slices on one XCD
  const int r = blockIdx.y;
  if (q >= nslice) return;
  const int n1 = a.g.n1, n2 = a.g.n2;
  const int t = threadIdx.x, nt = blockDim.x, nw = nt / 64, wv = t / 64, lane = t % 64;
  const int A = a.sk.A[C], B = a.sk.B[C], AB = A * B;
  const int s = 1 + t, sc = min(s, A);
  c2 *ring = lds;  // [nw][MXS_RING]: the edge row a wave receives
  float *cf = (float *)(lds + nw * MXS_RING);
  int *dtab = (int *)(cf + (B + 2));
  {
    const float *g0 = (C == 0) ? (FWD ? a.c2m : a.c2p) : (FWD ? a.c1m : a.c1p);
    for (int u = t; u <= B + 1; u += nt) cf[u] = g0[min(u, B)];
    for (int d = t; d <= A + B; d += nt) dtab[d] = d >= 2 ? sk_off(d, A, B) - max(1, d - B) : 0;
    for (int u = t; u < nw * MXS_RING; u += nt) ring[u] = cmk(0, 0);
  }
  const float *gs = (C == 2) ? (FWD ? a.c2m : a.c2p) : (FWD ? a.c3m : a.c3p);
  const float cs = gs[sc];
  const float4 *Rc = a.th + r * a.nsk + a.sk.base[C] + (long)q * AB;
  const long PL = (C == 0) ? (long)n1 * (n2 + 1) : (C == 1) ? (long)(n1 + 1) * n2 : (long)(n1 + 1) * (n2 + 1);
  const long SS = (C == 2) ? (long)(n1 + 1) : PL;
  const int FS = (C == 0) ? n1 : 1;
  const long nb = ((C == 0) ? (long)q : (C == 1) ? (long)(n1 + 1) * q : PL * q) + s * SS;
  const long wofs = (C == 0) ? 0 : (C == 1) ? a.fEx : a.fEx + a.fEy;
  const long xofs = (C == 0) ? 0 : (C == 1) ? a.g.lEx : a.g.lEx + a.g.lEy;
  const int kop = ((C == 2) ? q : s) - a.g.kofs;
  const long yb = xofs + nb - (long)a.g.kofs * PL;
  c2 *Wr = a.W + r * a.fE + wofs + nb;
  c2 *yr = a.y + r * a.g.n + yb;
  c2 *tw = a.trash + r * SGS_MAXT + t, *ty = a.trash + (MX_RMAX + r) * SGS_MAXT + t;
  // lag of this wave and the edges: forward, wave w receives row 64w from
  // wave w-1's lane 63; backward, row 64w+65 from wave w+1's lane 0
  const int lag = (FWD ? wv : nw - 1 - wv) * MXS_L;
  const bool recv = FWD ? (lane == 0 && wv > 0) : (lane == 63 && wv < nw - 1);
  const bool send = FWD ? (lane == 63 && wv < nw - 1) : (lane == 0 && wv > 0);
  c2 *ring_in = ring + wv * MXS_RING, *ring_out = ring + (FWD ? wv + 1 : wv - 1) * MXS_RING;
  const int dlo = 2, dhi = A + B, nst = dhi - dlo + 1;
  __syncthreads();
  auto load = [&](float4 &c, int st) {
    const int d = min(max(FWD ? dlo + st : dhi - st, dlo), dhi);
    c = Rc[min(max(dtab[d] + sc, 0), AB - 1)];
  };
  float4 rec[MXS_P];  // rec[tau % MXS_P] holds step tau - lag
#pragma unroll
  for (int u = 0; u < MXS_P; u++) load(rec[u], u - lag);
  c2 prev = cmk(0, 0);
  c2 edge = ring_in[(-MXS_L - 1) & (MXS_RING - 1)];  // the edge value for step 0
  const int ntk = ((nst + (nw - 1) * MXS_L + MXS_U - 1) / MXS_U) * MXS_U;
  for (int tk0 = 0; tk0 < ntk; tk0 += MXS_U) {
#pragma unroll
    for (int uu = 0; uu < MXS_U; uu++) {
      const int tau = tk0 + uu, u = uu % MXS_P, st = tau - lag;
      const int d = FWD ? dlo + st : dhi - st;
      const int f = d - s;
      const float cfv = cf[min(max(f, 0), B + 1)];
      const bool ok = s <= A && f >= 1 && f <= B;
      const long o = (long)f * FS;
      // the neighbour row on the previous diagonal
      c2 nbv = FWD ? cdpp<DPP_WAVE_SHR1>(prev) : cdpp<DPP_WAVE_SHL1>(prev);
      nbv.x = recv ? edge.x : nbv.x;
      nbv.y = recv ? edge.y : nbv.y;
      edge = ring_in[(tau - MXS_L) & (MXS_RING - 1)];  // for the next step
      const c2 w = cadd(cmk(rec[u].x, rec[u].y), crfma(crmul(cfv, prev), cs, nbv));
      const c2 v = cmul(cmk(rec[u].z, rec[u].w), w);
      if (FWD || C != 0) *(ok ? Wr + o : tw) = v;
      if (!FWD) *(ok && kop >= 0 ? yr + o : ty) = v;
      prev.x = ok ? v.x : prev.x;
      prev.y = ok ? v.y : prev.y;
      if (send) ring_out[tau & (MXS_RING - 1)] = prev;
      load(rec[u], st + MXS_P);
      if (uu % MXS_L == MXS_L - 1) __syncthreads();
    }
  }
  if (!FWD) {  // operator-grid cells of this slice outside the sweep range are zero rows
    const int n3 = a.g.n3f;
    const c2 z0 = cmk(0, 0);
    c2 *y0 = a.y + r * a.g.n;
    if (C == 0) {
      const int nk = a.g.n3 + 1;
      for (int w = t; w < nk * (n2 + 1); w += nt) {
        const int kq = w / (n2 + 1), j = w % (n2 + 1), kf = kq + a.g.kofs;
        if (!(kf >= 1 && kf <= n3 - 1 && j >= 1 && j <= n2 - 1)) y0[id1(n1, n2, kq, j, q)] = z0;
      }
    } else if (C == 1) {
      const int nk = a.g.n3 + 1;
      for (int w = t; w < nk * (n1 + 1); w += nt) {
        const int kq = w / (n1 + 1), i = w % (n1 + 1), kf = kq + a.g.kofs;
        if (!(kf >= 1 && kf <= n3 - 1 && i >= 1 && i <= n1 - 1)) y0[a.g.lEx + id2(n1, n2, kq, q, i)] = z0;
      }
    } else if (q - a.g.kofs >= 0) {
      for (int w = t; w < (n2 + 1) * (n1 + 1); w += nt) {
        const int jj = w / (n1 + 1), i = w % (n1 + 1);
        if (!(jj >= 1 && jj <= n2 - 1 && i >= 1 && i <= n1 - 1))
          y0[a.g.lEx + a.g.lEy + id3(n1, n2, q - a.g.kofs, jj, i)] = z0;
      }
    }
  }
}

// ---------------------------------------------------------------- Krylov
// Batched F-BiCGStab (solver.c:93-178) on FP32 vectors.  Dot products
// accumulate in FP64 in the deterministic chunk order of the FP64 path (per
// right-hand side: chunk, then thread, then the block fold); the scalars
// and the convergence state live in FP64 on the device, so an iteration
// can be replayed from a graph without the host.  Per-RHS scalar slots:
// MS_USED marks the right-hand sides that take part in the current inner
// solve (their correction is added to the FP64 solution).
enum { MS_RHO = 0, MS_ALPHA = 2, MS_OMEGA = 4, MS_BETA = 6, MS_RS0 = 8, MS_REL = 9, MS_DONE = 10, MS_TOL = 11,
       MS_IT = 12, MS_MAXIT = 13, MS_USED = 14, MS_N = 16 };

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
// fold 4 doubles per thread over the block into part[r][chunk][0..3]
__device__ __forceinline__ void mx_fold4(double *acc, double *out, double *sh)
{
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const double v = block_fold(acc[q], sh);
    if (threadIdx.x == 0) out[q] = v;
  }
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
