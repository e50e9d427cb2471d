// csem_mixed.hip -- batched mixed-precision solve on one model (gfx950).
//
// R <= MX_RMAX right-hand sides (sources x frequencies) of one model are
// solved together.  The outer loop is FP64 iterative refinement: the true
// residual b - A x uses the FP64 operator (the drop-in path's kernels) and
// the solution accumulates in FP64; each correction comes from an FP32
// inner solve (LU-SGS preconditioned F-BiCGStab, solver.c:93-178) on all
// right-hand sides at once, with FP64 dot products.  One inner iteration is
// captured in a hipGraph and replayed; the host looks at the convergence
// flags only every few iterations.  The operator's tile shape is picked by
// timing the candidates on the model's own grid (the autotuner below).
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <vector>
// the FP64 kernels live in csem_dev.hip; this unit needs their types only
#define CSEM_TYPES_ONLY
#include "csem_internal.h"
#include "csem_mixed_kernels.h"

void kt_pause(int paused);

struct MixState {
  int R;       // right-hand sides the buffers hold
  long n, fE;
  // FP32 model
  float *d1, *d2, *d3, *d1s, *d2s, *d3s, *r1, *r2, *r3, *r1s, *r2s, *r3s;
  float *s11, *s22, *s33, *c1m, *c1p, *c2m, *c2p, *c3m, *c3p;
  c2 *sH1, *sH2, *tw;
  float *wx, *wy, *gxw1, *gxw2, *gyw1, *gyw2;
  int cw;
  c2 *buf, *bufx, *bufy, *hz0, *hxg, *hyg;
  // LU-SGS
  c2 *W, *T, *h, *trash;
  float hkey[MX_RMAX][2];
  int hR;
  // Krylov
  c2 *kv;
  double *part, *sc;
  int nch;
  // refinement
  z2 *res, *tmp64;
  double *norm;  // device scratch for FP64 dots
  float *scale;  // R inverse residual norms
  // autotuned operator tile: rows per block and planes per block
  int ap_tj[MX_RMAX + 1], ap_kc[MX_RMAX + 1];
  // captured inner iteration (a graph cannot be captured from the legacy
  // default stream: NULL callers run on this stream of default priority)
  hipStream_t own;
  hipGraphExec_t gexec;
  int g_R;
  hipStream_t g_stream;
  MxFreq g_fr;  // the frequencies baked into the captured launches
  std::vector<void *> allocs;
  long bytes;
};

template <class T>
static int mx_dalloc(MixState *x, T **p, long count)
{
  long bytes = count * (long)sizeof(T);
  if (bytes <= 0) bytes = 256;
  hipError_t e = hipMalloc((void **)p, bytes);
  if (e != hipSuccess) return csem_fail(-2, "hipMalloc(%ld): %s", bytes, hipGetErrorString(e));
  x->allocs.push_back((void *)*p);
  x->bytes += bytes;
  return 0;
}

void mixed_free(csem_model *m)
{
  MixState *x = m->mx;
  if (!x) return;
  if (x->gexec) (void)hipGraphExecDestroy(x->gexec);
  if (x->own) (void)hipStreamDestroy(x->own);
  for (void *p : x->allocs) (void)hipFree(p);
  delete x;
  m->mx = NULL;
}

// ------------------------------------------------------------ conversions
__global__ void k_mx_cvt(long n, const double *a, float *o, int recip)
{
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) o[e] = recip ? (float)(1.0 / a[e]) : (float)a[e];
}
__global__ void k_mx_cvtz(long n, const z2 *a, c2 *o)
{
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) o[e] = cmk((float)a[e].x, (float)a[e].y);
}
// own-axis LU-SGS coefficients: cm[u] = 1/(ds[u-1] d[u]), cp[u] = 1/(ds[u] d[u])
__global__ void k_mx_coef(int n, const double *ds, const double *d, float *cm, float *cp)
{
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u > n) return;
  cm[u] = u >= 1 ? (float)(1.0 / (ds[u - 1] * d[u])) : 0.f;
  cp[u] = (float)(1.0 / (ds[u] * d[u]));
}
// FP64 <-> FP32 between the refinement and the inner solve
__global__ void k_mx_to32(long n, const z2 *res, const float *scale, c2 *r32)
{
  const int r = blockIdx.y;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const z2 v = res[r * n + e];
  const float s = scale[r];
  r32[r * n + e] = cmk((float)v.x * s, (float)v.y * s);
}
__global__ void k_mx_acc64(long n, z2 *x, const c2 *xin, const double *sc, const float *scale)
{
  const int r = blockIdx.y;
  if (sc[r * MS_N + MS_USED] == 0.0) return;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double s = 1.0 / (double)scale[r];
  const c2 v = xin[r * n + e];
  z2 &o = x[r * n + e];
  o = zmk(o.x + s * (double)v.x, o.y + s * (double)v.y);
}

#define GE(n) (unsigned)(((n) + 255) / 256), 256, 0, s

static MxGeo mx_geo(const csem_model *m, const MixState *x)
{
  MxGeo g;
  g.n1 = m->n1; g.n2 = m->n2; g.n3 = m->n3; g.n3f = m->n3f; g.kofs = m->kofs; g.airbc = m->airbc; g.fvm = m->fvm;
  g.n = m->lE; g.lEx = m->lEx; g.lEy = m->lEy;
  g.d1 = x->d1; g.d2 = x->d2; g.d3 = x->d3; g.d1s = x->d1s; g.d2s = x->d2s; g.d3s = x->d3s;
  g.r1 = x->r1; g.r2 = x->r2; g.r3 = x->r3; g.r1s = x->r1s; g.r2s = x->r2s; g.r3s = x->r3s;
  g.s11 = x->s11; g.s22 = x->s22; g.s33 = x->s33;
  return g;
}

static int mx_state(csem_model *m, int R, MixState **out)
{
  if (R < 1 || R > MX_RMAX) return csem_fail(-3, "batch of %d right-hand sides outside [1, %d]", R, MX_RMAX);
  if (m->mx && m->mx->R >= R) { *out = m->mx; return 0; }
  mixed_free(m);
  MixState *x = new MixState();
  m->mx = x;
  x->R = R; x->n = m->lE; x->fE = m->fE;
  const int n1 = m->n1, n2 = m->n2, n3f = m->n3f;
  const hipStream_t s = 0;
  int rc;
#define MA(p, cnt) if ((rc = mx_dalloc(x, &x->p, cnt))) goto fail
#define CV(dst, src, cnt, rcp) MA(dst, cnt); k_mx_cvt<<<GE(cnt)>>>(cnt, m->src, x->dst, rcp)
  CV(d1, d1, n1 + 1, 0); CV(d2, d2, n2 + 1, 0); CV(d3, d3, n3f + 1, 0);
  CV(d1s, d1s, n1 + 1, 0); CV(d2s, d2s, n2 + 1, 0); CV(d3s, d3s, n3f + 1, 0);
  CV(r1, d1, n1 + 1, 1); CV(r2, d2, n2 + 1, 1); CV(r3, d3, n3f + 1, 1);
  CV(r1s, d1s, n1 + 1, 1); CV(r2s, d2s, n2 + 1, 1); CV(r3s, d3s, n3f + 1, 1);
  CV(s11, s11, m->fEx, 0); CV(s22, s22, m->fEy, 0); CV(s33, s33, m->fEz, 0);
  MA(c1m, n1 + 1); MA(c1p, n1 + 1); MA(c2m, n2 + 1); MA(c2p, n2 + 1); MA(c3m, n3f + 1); MA(c3p, n3f + 1);
  k_mx_coef<<<GE(n1 + 1)>>>(n1, m->d1s, m->d1, x->c1m, x->c1p);
  k_mx_coef<<<GE(n2 + 1)>>>(n2, m->d2s, m->d2, x->c2m, x->c2p);
  k_mx_coef<<<GE(n3f + 1)>>>(n3f, m->d3s, m->d3, x->c3m, x->c3p);
  if (m->airbc) {
    const int N = m->nfft;
    const long NN = (long)N * N;
    if (N < 64 || 4096 % N) { rc = csem_fail(-3, "mixed path needs nfft in [64, 4096] (got %d)", N); goto fail; }
    x->cw = 4096 / N;  // 16 values per thread of a 256-thread column block
    if (hipFuncSetAttribute((const void *)k_mx_air_cols, hipFuncAttributeMaxDynamicSharedMemorySize, 65536) != hipSuccess) {
      rc = csem_fail(-1, "LDS attribute");
      goto fail;
    }
    MA(sH1, NN); MA(sH2, NN); MA(tw, N);
    k_mx_cvtz<<<GE(NN)>>>(NN, m->sH1, x->sH1);
    k_mx_cvtz<<<GE(NN)>>>(NN, m->sH2, x->sH2);
    k_mx_cvtz<<<GE(N)>>>(N, m->tw, x->tw);
    CV(wx, wx, N, 0); CV(wy, wy, N, 0);
    CV(gxw1, gxw1, n1, 0); CV(gxw2, gxw2, n2, 0); CV(gyw1, gyw1, n1, 0); CV(gyw2, gyw2, n2, 0);
    MA(buf, NN * R); MA(bufx, NN * R); MA(bufy, NN * R);
    MA(hz0, (long)n1 * n2 * R); MA(hxg, (long)(n1 + 1) * n2 * R); MA(hyg, (long)n1 * (n2 + 1) * R);
  }
  {
    long mx = 0;
    for (int c = 0; c < 3; c++) mx = std::max(mx, (long)m->g.ns[c] * m->g.A[c] * m->g.B[c]);
    MA(W, m->fE * R); MA(T, mx * R); MA(h, m->nsk * R); MA(trash, 2L * SGS_MAXT * MX_RMAX);
    x->nch = (int)((m->lE + RED_CHUNK - 1) / RED_CHUNK);
    MA(kv, 9L * R * m->lE); MA(part, (long)R * x->nch * 4); MA(sc, (long)R * MS_N);
    MA(res, R * m->lE); MA(tmp64, m->lE); MA(norm, 4 * R); MA(scale, R);
  }
#undef CV
#undef MA
  if (hipMemsetAsync(x->W, 0, m->fE * R * sizeof(c2), s) != hipSuccess) { rc = csem_fail(-1, "memset"); goto fail; }
  if (hipMemsetAsync(x->sc, 0, (long)R * MS_N * sizeof(double), s) != hipSuccess) { rc = csem_fail(-1, "memset"); goto fail; }
  if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) {
    rc = csem_fail(-1, "mixed-precision setup failed");
    goto fail;
  }
  x->hR = 0;
  if (hipStreamCreateWithFlags(&x->own, hipStreamNonBlocking) != hipSuccess) {
    x->own = NULL;
    rc = csem_fail(-1, "stream");
    goto fail;
  }
  *out = x;
  return 0;
fail:
  mixed_free(m);
  return rc;
}

static MxFreq mx_freq(const csem_freq *f, int R)
{
  MxFreq fr;
  memset(&fr, 0, sizeof fr);
  for (int r = 0; r < R; r++) {
    fr.iwmu[r] = cmk((float)f[r].iwmu_re, (float)f[r].iwmu_im);
    fr.iomu[r] = cmk((float)f[r].iomu_re, (float)f[r].iomu_im);
  }
  return fr;
}

// ---------------------------------------------------------------- operator
static MxAir mx_air(const csem_model *m, const MixState *x)
{
  MxAir a;
  a.N = m->nfft; a.n1 = m->n1; a.n2 = m->n2; a.cw = x->cw;
  a.hz = x->hz0; a.buf = x->buf; a.bufx = x->bufx; a.bufy = x->bufy;
  a.sH1 = x->sH1; a.sH2 = x->sH2; a.tw = x->tw;
  a.jx = m->jx; a.jy = m->jy; a.wx = x->wx; a.wy = x->wy;
  a.rows = m->rows; a.nrowx = m->nrowx; a.nrowy = m->nrowy;
  a.gxj1 = m->gxj1; a.gxj2 = m->gxj2; a.gyj1 = m->gyj1; a.gyj2 = m->gyj2;
  a.gxw1 = x->gxw1; a.gxw2 = x->gxw2; a.gyw1 = x->gyw1; a.gyw2 = x->gyw2;
  a.hxg = x->hxg; a.hyg = x->hyg;
  return a;
}

template <int R, int TJ>
static void mx_apply_launch(const csem_model *m, const MixState *x, const MxFreq &fr, const c2 *in, c2 *out, int kc,
                            hipStream_t s)
{
  dim3 blk(64, TJ + 1), grd((m->n1 + 1 + 62) / 63, (m->n2 + 1 + TJ - 1) / TJ, (m->n3 + 1 + kc - 1) / kc);
  k_mx_apply<R, TJ><<<grd, blk, 0, s>>>(mx_geo(m, x), fr, in, out, x->hxg, x->hyg, kc);
}
// rows per block: 2, 4 or 8 (static LDS 3*2*R*(TJ+1)*64*8 B stays under
// 64 KB: TJ = 8 only for R <= 2)
template <int R>
static void mx_apply_tile(const csem_model *m, const MixState *x, const MxFreq &fr, const c2 *in, c2 *out, int tj,
                          int kc, hipStream_t s)
{
  if (tj == 2) mx_apply_launch<R, 2>(m, x, fr, in, out, kc, s);
  else if (tj == 4 || R > 2) mx_apply_launch<R, 4>(m, x, fr, in, out, kc, s);
  else mx_apply_launch<R, (R > 2 ? 4 : 8)>(m, x, fr, in, out, kc, s);
}
static void mx_apply_kernel(const csem_model *m, const MixState *x, int R, const MxFreq &fr, const c2 *in, c2 *out,
                            int tj, int kc, hipStream_t s)
{
  switch (R) {
    case 1: mx_apply_tile<1>(m, x, fr, in, out, tj, kc, s); break;
    case 2: mx_apply_tile<2>(m, x, fr, in, out, tj, kc, s); break;
    case 3: mx_apply_tile<3>(m, x, fr, in, out, tj, kc, s); break;
    default: mx_apply_tile<4>(m, x, fr, in, out, tj, kc, s); break;
  }
}

// Autotuner: time every (rows per block, planes per block) candidate of the
// FP32 operator on this model and batch size, keep the fastest.  Tile shape
// changes no arithmetic: every output is formed by the same expression.
static int mx_tune(csem_model *m, MixState *x, int R, const MxFreq &fr, const c2 *in, c2 *out, hipStream_t s)
{
  if (x->ap_tj[R]) return 0;
  const int TJS[2] = {R > 2 ? 2 : 4, R > 2 ? 4 : 8}, KCS[3] = {4, 8, 16};
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return csem_fail(-1, "event");
  float best = 1e30f;
  int bt = 4, bk = 8;
  kt_pause(1);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 3; b++) {
      mx_apply_kernel(m, x, R, fr, in, out, TJS[a], KCS[b], s);  // warm
      (void)hipEventRecord(e0, s);
      for (int q = 0; q < 3; q++) mx_apply_kernel(m, x, R, fr, in, out, TJS[a], KCS[b], s);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) { best = ms; bt = TJS[a]; bk = KCS[b]; }
    }
  kt_pause(0);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  x->ap_tj[R] = bt;
  x->ap_kc[R] = bk;
  HIP_OK(hipGetLastError());
  return 0;
}

// y = A x for the R right-hand sides, FP32, with the airwave boundary.
static int mx_apply(csem_model *m, MixState *x, int R, const MxFreq &fr, const c2 *in, c2 *out, hipStream_t s)
{
  if (m->airbc) {
    const int nh = m->n1 * m->n2, N = m->nfft;
    MxAir a = mx_air(m, x);
    KT(KT_M_HZ0, s, 24.0 * nh * R, (k_mx_hz0<<<dim3((nh + 255) / 256, R), 256, 0, s>>>(mx_geo(m, x), in, x->hz0)));
    KT(KT_M_AIR_ROWS, s, 8.0 * N * N * R,
       (k_mx_air_rows<<<dim3(N, R), 256, 2 * (size_t)N * sizeof(c2), s>>>(a)));
    KT(KT_M_AIR_COLS, s, 40.0 * N * N * R,
       (k_mx_air_cols<<<dim3(N / x->cw, R), 256, 2 * (size_t)N * x->cw * sizeof(c2), s>>>(a)));
    KT(KT_M_AIR_RINV, s, 16.0 * N * (m->nrowx + m->nrowy) * R,
       (k_mx_air_rows_inv<<<dim3(m->nrowx + m->nrowy, R), 256, 2 * (size_t)N * sizeof(c2), s>>>(a)));
    const int ng = (m->n1 + 1) * m->n2 + m->n1 * (m->n2 + 1);
    KT(KT_M_AIR_GHOST, s, 8.0 * ng * R, (k_mx_air_ghost<<<dim3((ng + 255) / 256, R), 256, 0, s>>>(a)));
  }
  // algorithmic bytes: read E and write y per right-hand side, sigma once
  KT(KT_M_APPLY, s, (16.0 * R + 4.0) * m->lE,
     mx_apply_kernel(m, x, R, fr, in, out, x->ap_tj[R] ? x->ap_tj[R] : 4, x->ap_kc[R] ? x->ap_kc[R] : 8, s));
  HIP_OK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------- LU-SGS
static MxLu mx_lu(const csem_model *m, MixState *x, const MxFreq &fr, const c2 *b, c2 *y)
{
  MxLu a;
  a.g = mx_geo(m, x); a.sk = m->g;
  a.b = b; a.y = y; a.W = x->W; a.fE = m->fE; a.fEx = m->fEx; a.fEy = m->fEy;
  a.T = x->T; a.h = x->h; a.trash = x->trash;
  a.c1m = x->c1m; a.c1p = x->c1p; a.c2m = x->c2m; a.c2p = x->c2p; a.c3m = x->c3m; a.c3p = x->c3p;
  a.fr = fr;
  return a;
}

template <int R>
static void mx_lusgs_R(const csem_model *m, MixState *x, const MxLu &a, bool diag, bool sweep, hipStream_t s)
{
  const SkewGeo &g = m->g;
  dim3 gp[3], gq[3];
  for (int c = 0; c < 3; c++) gp[c] = dim3((unsigned)(((long)g.A[c] * g.B[c] + 255) / 256), (unsigned)g.ns[c]);
  if (diag) {
    k_mx_diag<0, R><<<gp[0], 256, 0, s>>>(a);
    k_mx_diag<1, R><<<gp[1], 256, 0, s>>>(a);
    k_mx_diag<2, R><<<gp[2], 256, 0, s>>>(a);
  }
  if (!sweep) return;
  gq[0] = dim3((m->n1 + 63) / 64, (m->n2 + 1 + 3) / 4, m->n3f + 1);
  gq[1] = dim3((m->n1 + 1 + 63) / 64, (m->n2 + 3) / 4, m->n3f + 1);
  gq[2] = dim3((m->n1 + 1 + 63) / 64, (m->n2 + 1 + 3) / 4, m->n3f);
  const dim3 bq(64, 4);
  int tb[3];
  size_t lds[3];
  unsigned gs[3];
  double nc[3];
  for (int c = 0; c < 3; c++) {
    tb[c] = ((g.A[c] + 63) / 64) * 64;
    lds[c] = 2 * (size_t)(tb[c] + 2) * R * sizeof(c2) + (size_t)(g.B[c] + 2) * sizeof(float) +
             (size_t)(g.A[c] + g.B[c] + 1) * sizeof(int);
    gs[c] = 8 * ((g.ns[c] + 7) / 8);
    nc[c] = (double)g.ns[c] * g.A[c] * g.B[c];
  }
  // algorithmic bytes per swept cell and right-hand side: read b, write
  // the value, plus the other components read (as the FP64 accounting, in
  // FP32: forward Ex 16, Ey 24, Ez 32; backward 40)
  KT(KT_M_PRE, s, 0.0, (k_mx_pre<0, true, R><<<gq[0], bq, 0, s>>>(a)));
  KT(KT_M_SGS, s, 16.0 * R * nc[0], (k_mx_sgs<0, true, R><<<gs[0], tb[0], lds[0], s>>>(a)));
  KT(KT_M_PRE, s, 0.0, (k_mx_pre<1, true, R><<<gq[1], bq, 0, s>>>(a)));
  KT(KT_M_SGS, s, 24.0 * R * nc[1], (k_mx_sgs<1, true, R><<<gs[1], tb[1], lds[1], s>>>(a)));
  KT(KT_M_PRE, s, 0.0, (k_mx_pre<2, true, R><<<gq[2], bq, 0, s>>>(a)));
  KT(KT_M_SGS, s, 32.0 * R * nc[2], (k_mx_sgs<2, true, R><<<gs[2], tb[2], lds[2], s>>>(a)));
  KT(KT_M_PRE, s, 0.0, (k_mx_pre<2, false, R><<<gq[2], bq, 0, s>>>(a)));
  KT(KT_M_SGS, s, 40.0 * R * nc[2], (k_mx_sgs<2, false, R><<<gs[2], tb[2], lds[2], s>>>(a)));
  KT(KT_M_PRE, s, 0.0, (k_mx_pre<1, false, R><<<gq[1], bq, 0, s>>>(a)));
  KT(KT_M_SGS, s, 40.0 * R * nc[1], (k_mx_sgs<1, false, R><<<gs[1], tb[1], lds[1], s>>>(a)));
  KT(KT_M_PRE, s, 0.0, (k_mx_pre<0, false, R><<<gq[0], bq, 0, s>>>(a)));
  KT(KT_M_SGS, s, 40.0 * R * nc[0], (k_mx_sgs<0, false, R><<<gs[0], tb[0], lds[0], s>>>(a)));
}

// the diagonal factors h depend on the batch's frequencies: recomputed when
// they change (never inside a captured iteration)
static bool mx_diag_stale(MixState *x, int R, const MxFreq &fr)
{
  bool stale = x->hR != R;
  for (int r = 0; r < R && !stale; r++)
    stale = x->hkey[r][0] != fr.iwmu[r].x || x->hkey[r][1] != fr.iwmu[r].y;
  if (stale) {
    x->hR = R;
    for (int r = 0; r < R; r++) { x->hkey[r][0] = fr.iwmu[r].x; x->hkey[r][1] = fr.iwmu[r].y; }
  }
  return stale;
}

// Refresh the diagonal factors for this batch's frequencies (outside any
// captured iteration), then sweep.
static int mx_diag(csem_model *m, MixState *x, int R, const MxFreq &fr, hipStream_t s)
{
  if (!mx_diag_stale(x, R, fr)) return 0;
  const MxLu a = mx_lu(m, x, fr, NULL, NULL);
  switch (R) {
    case 1: mx_lusgs_R<1>(m, x, a, true, false, s); break;
    case 2: mx_lusgs_R<2>(m, x, a, true, false, s); break;
    case 3: mx_lusgs_R<3>(m, x, a, true, false, s); break;
    default: mx_lusgs_R<4>(m, x, a, true, false, s); break;
  }
  HIP_OK(hipGetLastError());
  return 0;
}

static int mx_lusgs(csem_model *m, MixState *x, int R, const MxFreq &fr, const c2 *b, c2 *y, hipStream_t s)
{
  const MxLu a = mx_lu(m, x, fr, b, y);
  switch (R) {
    case 1: mx_lusgs_R<1>(m, x, a, false, true, s); break;
    case 2: mx_lusgs_R<2>(m, x, a, false, true, s); break;
    case 3: mx_lusgs_R<3>(m, x, a, false, true, s); break;
    default: mx_lusgs_R<4>(m, x, a, false, true, s); break;
  }
  HIP_OK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- Krylov
static MxKry mx_kry(const csem_model *m, MixState *x)
{
  MxKry k;
  const long n = m->lE, R = x->R;
  k.n = n; k.nch = x->nch;
  c2 *v = x->kv;
  k.r = v; k.r0 = v + R * n; k.p = v + 2 * R * n; k.pp = v + 3 * R * n; k.v = v + 4 * R * n;
  k.s = v + 5 * R * n; k.ss = v + 6 * R * n; k.t = v + 7 * R * n; k.x = v + 8 * R * n;
  k.part = x->part; k.sc = x->sc;
  return k;
}

// one F-BiCGStab iteration for every active right-hand side (solver.c:139-177)
static int mx_iteration(csem_model *m, MixState *x, int R, const MxFreq &fr, hipStream_t s)
{
  const MxKry k = mx_kry(m, x);
  const dim3 gr(x->nch, R), ge((unsigned)((k.n + 255) / 256), R);
  int rc;
  if ((rc = mx_lusgs(m, x, R, fr, k.p, k.pp, s))) return rc;
  if ((rc = mx_apply(m, x, R, fr, k.pp, k.v, s))) return rc;
  KT(KT_M_DOT, s, 16.0 * R * k.n, (k_mx_k_dot_vr0<<<gr, RED_T, 0, s>>>(k)));
  KT(KT_M_DOT, s, 0.0, (k_mx_k_fin<<<R, RED_T, 0, s>>>(k, 1)));
  KT(KT_M_VEC, s, 24.0 * R * k.n, (k_mx_k_s<<<ge, 256, 0, s>>>(k)));
  if ((rc = mx_lusgs(m, x, R, fr, k.s, k.ss, s))) return rc;
  if ((rc = mx_apply(m, x, R, fr, k.ss, k.t, s))) return rc;
  KT(KT_M_DOT, s, 16.0 * R * k.n, (k_mx_k_dot_st<<<gr, RED_T, 0, s>>>(k)));
  KT(KT_M_DOT, s, 0.0, (k_mx_k_fin<<<R, RED_T, 0, s>>>(k, 2)));
  KT(KT_M_VEC, s, 64.0 * R * k.n, (k_mx_k_xr<<<gr, RED_T, 0, s>>>(k)));
  KT(KT_M_DOT, s, 0.0, (k_mx_k_fin<<<R, RED_T, 0, s>>>(k, 3)));
  KT(KT_M_VEC, s, 32.0 * R * k.n, (k_mx_k_p<<<ge, 256, 0, s>>>(k)));
  HIP_OK(hipGetLastError());
  return 0;
}

// Capture one iteration as a graph (kernel launches only: the frequencies
// and every scalar live in device memory or in the captured arguments).
static int mx_graph(csem_model *m, MixState *x, int R, const MxFreq &fr, hipStream_t s)
{
  if (x->gexec && x->g_R == R && x->g_stream == s && !memcmp(&x->g_fr, &fr, sizeof fr)) return 0;
  if (x->gexec) { (void)hipGraphExecDestroy(x->gexec); x->gexec = NULL; }
  hipGraph_t g;
  kt_pause(1);
  HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = mx_iteration(m, x, R, fr, s);
  hipError_t e = hipStreamEndCapture(s, &g);
  kt_pause(0);
  if (rc) return rc;
  if (e != hipSuccess) return csem_fail(-1, "graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&x->gexec, g, NULL, NULL, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { x->gexec = NULL; return csem_fail(-1, "graph instantiate: %s", hipGetErrorString(e)); }
  x->g_R = R;
  x->g_stream = s;
  x->g_fr = fr;
  return 0;
}

// ------------------------------------------------------------ refinement
extern "C" int csem_mixed_solve(csem_model *m, int R, const csem_freq *f, void *xv, const void *bv, const double *tol,
                                const csem_mixed_opts *opt, int *iters, double *relres, void *stream)
{
  MixState *x;
  int rc = mx_state(m, R, &x);
  if (rc) return rc;
  // NULL: run on the model's own stream, after the caller's queued work
  if (!stream) HIP_OK(hipDeviceSynchronize());
  hipStream_t s = stream ? (hipStream_t)stream : x->own;
  stream = (void *)s;
  const long n = m->lE;
  const MxFreq fr = mx_freq(f, R);
  const csem_mixed_opts dflt = {2000, 1e-2, 40, 8, 1};
  const csem_mixed_opts &o = opt ? *opt : dflt;
  z2 *X = (z2 *)xv;
  const z2 *B = (const z2 *)bv;
  const MxKry k = mx_kry(m, x);
  // scratch reuse: r32 is the Krylov r vector, the correction is k.x
  std::vector<double> bn(R), rn(R), hsc((size_t)R * MS_N);
  std::vector<int> used(R, 0);
  std::vector<float> scl(R);
  for (int r = 0; r < R; r++) {
    if ((rc = csem_dot(n, B + r * n, B + r * n, x->norm + 4 * r, stream))) return rc;
  }
  HIP_OK(hipMemsetAsync(X, 0, (size_t)R * n * sizeof(z2), s));
  HIP_OK(hipMemcpyAsync(x->res, B, (size_t)R * n * sizeof(z2), hipMemcpyDeviceToDevice, s));
  {
    std::vector<double> nb(4 * R);
    HIP_OK(hipMemcpyAsync(nb.data(), x->norm, 4 * R * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (int r = 0; r < R; r++) bn[r] = sqrt(nb[4 * r]);
  }
  if ((rc = mx_tune(m, x, R, fr, k.r, k.v, s))) return rc;  // operands: any R x n FP32 buffers
  if ((rc = mx_diag(m, x, R, fr, s))) return rc;
  for (int r = 0; r < R; r++) { rn[r] = bn[r]; relres[r] = 1.0; iters[r] = 0; }
  for (int outer = 0; outer <= o.max_outer; outer++) {
    // which right-hand sides still need a correction
    bool any = false;
    for (int r = 0; r < R; r++) {
      double *sc = &hsc[(size_t)r * MS_N];
      memset(sc, 0, MS_N * sizeof(double));
      const bool act = bn[r] > 0 && relres[r] >= tol[r] && iters[r] < o.max_inner && outer < o.max_outer;
      sc[MS_DONE] = act ? 0.0 : 1.0;
      sc[MS_USED] = act ? 1.0 : 0.0;
      // reduce this residual by what is still needed (with a margin), at most to inner_tol
      sc[MS_TOL] = std::max(o.inner_tol, 0.3 * tol[r] / std::max(relres[r], 1e-300));
      sc[MS_MAXIT] = o.max_inner - iters[r];
      scl[r] = act ? (float)(1.0 / rn[r]) : 0.f;
      any |= act;
    }
    if (!any) break;
    HIP_OK(hipMemcpyAsync(x->sc, hsc.data(), hsc.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(x->scale, scl.data(), R * sizeof(float), hipMemcpyHostToDevice, s));
    KT(KT_M_REFINE, s, 24.0 * R * n, (k_mx_to32<<<dim3((unsigned)((n + 255) / 256), R), 256, 0, s>>>(n, x->res, x->scale, k.r)));
    KT(KT_M_DOT, s, 32.0 * R * n, (k_mx_k_init<<<dim3(x->nch, R), RED_T, 0, s>>>(k)));
    KT(KT_M_DOT, s, 0.0, (k_mx_k_fin<<<R, RED_T, 0, s>>>(k, 0)));
    if (o.use_graph && (rc = mx_graph(m, x, R, fr, s))) return rc;
    // inner iterations, convergence checked every check_every
    for (;;) {
      for (int q = 0; q < std::max(1, o.check_every); q++) {
        if (o.use_graph) HIP_OK(hipGraphLaunch(x->gexec, s));
        else if ((rc = mx_iteration(m, x, R, fr, s))) return rc;
      }
      HIP_OK(hipMemcpyAsync(hsc.data(), x->sc, hsc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      bool all = true;
      for (int r = 0; r < R; r++) all &= hsc[(size_t)r * MS_N + MS_DONE] != 0.0;
      if (all) break;
    }
    for (int r = 0; r < R; r++)
      if (hsc[(size_t)r * MS_N + MS_USED] != 0.0) iters[r] += (int)hsc[(size_t)r * MS_N + MS_IT];
    // x += |res| * correction (FP64), then the true residual in FP64
    KT(KT_M_REFINE, s, 32.0 * R * n, (k_mx_acc64<<<dim3((unsigned)((n + 255) / 256), R), 256, 0, s>>>(n, X, k.x, x->sc, x->scale)));
    for (int r = 0; r < R; r++) {
      if (hsc[(size_t)r * MS_N + MS_USED] == 0.0) continue;
      if ((rc = csem_apply(m, &f[r], X + r * n, x->tmp64, stream))) return rc;
      if ((rc = csem_vec_sub(n, B + r * n, x->tmp64, x->res + r * n, stream))) return rc;
      if ((rc = csem_dot(n, x->res + r * n, x->res + r * n, x->norm + 4 * r, stream))) return rc;
    }
    std::vector<double> nb(4 * R);
    HIP_OK(hipMemcpyAsync(nb.data(), x->norm, 4 * R * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (int r = 0; r < R; r++) {
      if (hsc[(size_t)r * MS_N + MS_USED] == 0.0) continue;
      rn[r] = sqrt(nb[4 * r]);
      relres[r] = bn[r] > 0 ? rn[r] / bn[r] : 0.0;
    }
  }
  HIP_OK(hipGetLastError());
  int conv = 0;
  for (int r = 0; r < R; r++) conv += relres[r] < tol[r];
  return conv;
}

extern "C" int csem_apply32(csem_model *m, int R, const csem_freq *f, const void *xv, void *yv, void *stream)
{
  MixState *x;
  int rc = mx_state(m, R, &x);
  if (rc) return rc;
  return mx_apply(m, x, R, mx_freq(f, R), (const c2 *)xv, (c2 *)yv, (hipStream_t)stream);
}

extern "C" int csem_lusgs32(csem_model *m, int R, const csem_freq *f, const void *xv, void *yv, void *stream)
{
  MixState *x;
  int rc = mx_state(m, R, &x);
  if (rc) return rc;
  const MxFreq fr = mx_freq(f, R);
  if ((rc = mx_diag(m, x, R, fr, (hipStream_t)stream))) return rc;
  return mx_lusgs(m, x, R, fr, (const c2 *)xv, (c2 *)yv, (hipStream_t)stream);
}

extern "C" int csem_mixed_tuning(csem_model *m, int R, int *rows, int *planes)
{
  if (!m->mx || R < 1 || R > MX_RMAX) return -1;
  *rows = m->mx->ap_tj[R];
  *planes = m->mx->ap_kc[R];
  return 0;
}

extern "C" long csem_mixed_bytes(const csem_model *m) { return m->mx ? m->mx->bytes : 0; }
