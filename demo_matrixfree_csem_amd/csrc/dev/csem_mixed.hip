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
#include <mutex>
#include <vector>
// the FP64 kernels live in csem_dev.hip; this unit needs their types only
#define CSEM_TYPES_ONLY
#include "csem_internal.h"
#include "csem_mixed_kernels.h"
#include "csem_mixed_gmres.h"

void kt_pause(int paused);

// FP32 copy of the model, shared by every solve on it.
struct MixState {
  float *d1, *d2, *d3, *d1s, *d2s, *d3s, *r1, *r2, *r3, *r1s, *r2s, *r3s;
  float *s11, *s22, *s33, *c1m, *c1p, *c2m, *c2p, *c3m, *c3p;
  c2 *sH1, *sH2, *tw;
  float *wx, *wy, *gxw1, *gxw2, *gyw1, *gyw2;
  int cw;       // default columns per block of the FFT column pass
  int cw_max;   // the largest candidate (the kernel's LDS attribute is set for it)
  // autotuned per batch size: the operator tile (rows per block, planes per
  // block), the six LU-SGS pre-passes' tiles (candidate index, in launch
  // order: Ex, Ey, Ez forward, Ez, Ey, Ex backward) and the FFT column
  // pass's columns per block
  int ap_tj[MX_RMAX + 1], ap_kc[MX_RMAX + 1];
  int pre_v[MX_RMAX + 1][6];
  int air_cw[MX_RMAX + 1];
  std::mutex mu;                 // guards slots and tuning
  std::vector<struct MixSlot *> slots;
  std::vector<void *> allocs;
  long bytes;
};

// Buffers of one solve in flight: solves on several host threads (each on
// its own stream) use different slots of the same model concurrently.
struct MixSlot {
  int R;      // right-hand sides the buffers hold
  int busy;
  c2 *buf, *bufx, *bufy, *hz0, *hxg, *hyg;  // FP32 airwave scratch, R of each
  AirBufs a64;                              // FP64 airwave scratch (true residual)
  c2 *W, *V;                                // LU-SGS state; a sweep's values in record-slot order
  c2 *tt, *hh;                              // LU-SGS records per record slot and r: T, h = iwmu/aii
  c2 h_iw[MX_RMAX];                         // the iwmu each right-hand side's h was computed for
  char h_ok[MX_RMAX];                       // whether it was
  int pc;                                   // preconditioner of the solve in flight (csem_mixed_opts)
  c2 *kv;                                   // Krylov vectors
  c2 *xbest;                                // F-BiCGStab: the correction at the best recursive residual
  c2 *gmV, *gmZ;                            // FGMRES basis and preconditioned basis (on first use)
  double *gmpart, *gmhs;
  int gm_m;                                 // restart length the FGMRES buffers hold
  double *part, *sc;
  int nch;
  z2 *res, *tmp64;                          // refinement
  z2 *xprev;                                // the solutions before the last correction
  double *npart, *norm;                     // FP64 residual norms
  float *scale;
  // captured inner iteration (a graph cannot be captured from the legacy
  // default stream: NULL callers run on this stream of default priority)
  hipStream_t own;
  hipGraphExec_t gexec;
  int g_R;
  hipStream_t g_stream;
  MxFreq g_fr;  // the frequencies baked into the captured launches
  int g_pc;     // and the preconditioner
  int g_kind;   // 0: one F-BiCGStab iteration; m > 0: one FGMRES(m) cycle
  std::vector<void *> allocs;
  long bytes;
};

template <class S, class T>
static int mx_dalloc(S *x, T **p, long count)
{
  long bytes = count * (long)sizeof(T);
  if (bytes <= 0) bytes = 256;
  hipError_t e = dev_malloc((void **)p, bytes);
  if (e != hipSuccess) return csem_fail(-2, "hipMalloc(%ld): %s", bytes, hipGetErrorString(e));
  x->allocs.push_back((void *)*p);
  x->bytes += bytes;
  return 0;
}

static void slot_free(MixSlot *q)
{
  if (q->gexec) (void)hipGraphExecDestroy(q->gexec);
  if (q->own) (void)hipStreamDestroy(q->own);
  for (void *p : q->allocs) dev_free(p);
  delete q;
}

void mixed_free(csem_model *m)
{
  MixState *x = m->mx;
  if (!x) return;
  for (MixSlot *q : x->slots) slot_free(q);
  for (void *p : x->allocs) dev_free(p);
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

// Record slots of the FP32 LU-SGS (MxPad): per slice A + B + 1 diagonals of
// A rounded up to whole waves rows.
static MxPad mx_pad(const csem_model *m)
{
  MxPad p;
  long off = 0;
  for (int c = 0; c < 3; c++) {
    p.AP[c] = ((m->g.A[c] + 63) / 64) * 64;
    p.SL[c] = (long)(m->g.A[c] + m->g.B[c] + 1) * p.AP[c];
    p.base[c] = off;
    off += p.SL[c] * m->g.ns[c];
  }
  p.n = off;
  return p;
}

#define MA(o, p, cnt) if ((rc = mx_dalloc(o, &o->p, cnt))) goto fail

// Guards the creation of a model's mixed-precision state (first use of a
// model from several threads) and readers of m->mx.
static std::mutex g_mx_init_mu;

// The FP32 model, once per model (callers hold g_mx_init_mu).
static int mx_model(csem_model *m, MixState **out)
{
  if (m->mx) { *out = m->mx; return 0; }
  MixState *x = new MixState();
  for (int r = 0; r <= MX_RMAX; r++)
    for (int k = 0; k < 6; k++) x->pre_v[r][k] = -1;  // untuned: the default tiles
  const int n1 = m->n1, n2 = m->n2, n3f = m->n3f;
  const hipStream_t s = 0;
  int rc;
#define CV(dst, src, cnt, rcp) MA(x, dst, cnt); k_mx_cvt<<<GE(cnt)>>>(cnt, m->src, x->dst, rcp)
  CV(d1, d1, n1 + 1, 0); CV(d2, d2, n2 + 1, 0); CV(d3, d3, n3f + 1, 0);
  CV(d1s, d1s, n1 + 1, 0); CV(d2s, d2s, n2 + 1, 0); CV(d3s, d3s, n3f + 1, 0);
  CV(r1, d1, n1 + 1, 1); CV(r2, d2, n2 + 1, 1); CV(r3, d3, n3f + 1, 1);
  CV(r1s, d1s, n1 + 1, 1); CV(r2s, d2s, n2 + 1, 1); CV(r3s, d3s, n3f + 1, 1);
  CV(s11, s11, m->fEx, 0); CV(s22, s22, m->fEy, 0); CV(s33, s33, m->fEz, 0);
  MA(x, c1m, n1 + 1); MA(x, c1p, n1 + 1); MA(x, c2m, n2 + 1); MA(x, c2p, n2 + 1); MA(x, c3m, n3f + 1);
  MA(x, c3p, n3f + 1);
  k_mx_coef<<<GE(n1 + 1)>>>(n1, m->d1s, m->d1, x->c1m, x->c1p);
  k_mx_coef<<<GE(n2 + 1)>>>(n2, m->d2s, m->d2, x->c2m, x->c2p);
  k_mx_coef<<<GE(n3f + 1)>>>(n3f, m->d3s, m->d3, x->c3m, x->c3p);
  if (m->airbc) {
    const int N = m->nfft;
    const long NN = (long)N * N;
    if (N < 64 || 4096 % N) { rc = csem_fail(-3, "mixed path needs nfft in [64, 4096] (got %d)", N); goto fail; }
    x->cw = N >= 2048 ? 1 : 2048 / N;  // columns per column-pass block of N cw / 16 threads, 16 values each
    x->cw_max = 4096 / N;              // candidates: N cw from 1024 to 4096 (64 to 256 threads)
    const int lc = (int)((2 * (size_t)N * x->cw_max + N) * sizeof(c2)), lr = (int)(3 * (size_t)N * sizeof(c2));
    if (hipFuncSetAttribute((const void *)k_mx_air_cols, hipFuncAttributeMaxDynamicSharedMemorySize, lc) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_mx_air_rows, hipFuncAttributeMaxDynamicSharedMemorySize, lr) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_mx_air_rows_inv, hipFuncAttributeMaxDynamicSharedMemorySize, lr) != hipSuccess) {
      rc = csem_fail(-1, "LDS attribute");
      goto fail;
    }
    MA(x, sH1, NN); MA(x, sH2, NN); MA(x, tw, N);
    k_mx_cvtz<<<GE(NN)>>>(NN, m->sH1, x->sH1);
    k_mx_cvtz<<<GE(NN)>>>(NN, m->sH2, x->sH2);
    k_mx_cvtz<<<GE(N)>>>(N, m->tw, x->tw);
    CV(wx, wx, N, 0); CV(wy, wy, N, 0);
    CV(gxw1, gxw1, n1, 0); CV(gxw2, gxw2, n2, 0); CV(gyw1, gyw1, n1, 0); CV(gyw2, gyw2, n2, 0);
  }
#undef CV
  if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) {
    rc = csem_fail(-1, "mixed-precision setup failed");
    goto fail;
  }
  m->mx = x;
  *out = x;
  return 0;
fail:
  for (void *p : x->allocs) dev_free(p);
  delete x;
  return rc;
}

static int slot_new(csem_model *m, int R, MixSlot **out)
{
  {
    // the FP32 sweeps address W, y and a slice's records as raw buffers of
    // under 1 GiB each (32-bit offsets, and a masked lane's offset past them)
    const MxPad pk = mx_pad(m);
    long big = std::max(std::max(m->fEx, m->fEy), m->fE - m->fEx - m->fEy) * (long)sizeof(c2);
    big = std::max(big, std::max(std::max(m->lEx, m->lEy), m->lE - m->lEx - m->lEy) * (long)sizeof(c2));
    for (int c = 0; c < 3; c++) big = std::max(big, pk.SL[c] * (long)sizeof(float4));
    if (pk.n >= (1L << 31) || big >= (1L << 30) - (1L << 24))
      return csem_fail(-3, "model too large for the FP32 LU-SGS addressing (%ld records, %ld-byte component)", pk.n, big);
  }
  MixSlot *q = new MixSlot();
  q->R = R;
  const int n1 = m->n1, n2 = m->n2;
  int rc;
  if (m->airbc) {
    const long NN = (long)m->nfft * m->nfft;
    MA(q, buf, NN * R); MA(q, bufx, NN * R); MA(q, bufy, NN * R);
    MA(q, hz0, (long)n1 * n2 * R); MA(q, hxg, (long)(n1 + 1) * n2 * R); MA(q, hyg, (long)n1 * (n2 + 1) * R);
    MA(q, a64.buf, NN); MA(q, a64.bufx, NN); MA(q, a64.bufy, NN);
    MA(q, a64.hz0, (long)n1 * n2); MA(q, a64.hxg, (long)(n1 + 1) * n2); MA(q, a64.hyg, (long)n1 * (n2 + 1));
  }
  {
    MA(q, W, m->fE * R); MA(q, tt, mx_pad(m).n * R); MA(q, hh, mx_pad(m).n * R); MA(q, V, mx_pad(m).n * R);
    q->nch = (int)((m->lE + RED_CHUNK - 1) / RED_CHUNK);
    MA(q, kv, 9L * R * m->lE); MA(q, part, (long)R * q->nch * 4); MA(q, sc, (long)R * MS_N);
    MA(q, xbest, R * m->lE);
    MA(q, res, R * m->lE); MA(q, tmp64, m->lE); MA(q, npart, (long)R * q->nch); MA(q, norm, R);
    MA(q, xprev, R * m->lE);
    MA(q, scale, R);
  }
  // records of slots off the slices stay zero (the pre-passes write the cells' records)
  if (dev_memset(q->W, 0, m->fE * R * sizeof(c2)) != hipSuccess ||
      dev_memset(q->tt, 0, mx_pad(m).n * R * sizeof(c2)) != hipSuccess ||
      dev_memset(q->hh, 0, mx_pad(m).n * R * sizeof(c2)) != hipSuccess ||
      dev_memset(q->sc, 0, (long)R * MS_N * sizeof(double)) != hipSuccess) {
    rc = csem_fail(-1, "memset");
    goto fail;
  }
  if (hipStreamCreateWithFlags(&q->own, hipStreamNonBlocking) != hipSuccess) {
    q->own = NULL;
    rc = csem_fail(-1, "stream");
    goto fail;
  }
  *out = q;
  return 0;
fail:
  slot_free(q);
  return rc;
}
#undef MA

// Take a free slot holding at least R right-hand sides (a new one if none).
static int slot_acquire(csem_model *m, MixState *x, int R, MixSlot **out)
{
  if (R < 1 || R > MX_RMAX) return csem_fail(-3, "batch of %d right-hand sides outside [1, %d]", R, MX_RMAX);
  {
    std::lock_guard<std::mutex> lk(x->mu);
    for (MixSlot *q : x->slots)
      if (!q->busy && q->R >= R) { q->busy = 1; *out = q; return 0; }
  }
  MixSlot *q;
  int rc = slot_new(m, R, &q);
  if (rc) return rc;
  q->busy = 1;
  std::lock_guard<std::mutex> lk(x->mu);
  x->slots.push_back(q);
  *out = q;
  return 0;
}
static void slot_release(MixState *x, MixSlot *q)
{
  std::lock_guard<std::mutex> lk(x->mu);
  q->busy = 0;
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
static MxAir mx_air(const csem_model *m, const MixState *x, const MixSlot *q, int R)
{
  MxAir a;
  a.sc = nullptr;
  a.N = m->nfft; a.n1 = m->n1; a.n2 = m->n2; a.cw = x->air_cw[R] ? x->air_cw[R] : x->cw;
  a.hz = q->hz0; a.buf = q->buf; a.bufx = q->bufx; a.bufy = q->bufy;
  a.sH1 = x->sH1; a.sH2 = x->sH2; a.tw = x->tw;
  a.jx = m->jx; a.jy = m->jy; a.wx = x->wx; a.wy = x->wy;
  a.rows = m->rows; a.nrowx = m->nrowx; a.nrowy = m->nrowy;
  a.gxj1 = m->gxj1; a.gxj2 = m->gxj2; a.gyj1 = m->gyj1; a.gyj2 = m->gyj2;
  a.gxw1 = x->gxw1; a.gxw2 = x->gxw2; a.gyw1 = x->gyw1; a.gyw2 = x->gyw2;
  a.hxg = q->hxg; a.hyg = q->hyg;
  return a;
}

// Operator tiles: NW waves (NW - 2 output rows) x KC planes.  LDS holds
// 4 * NW * R * 64 complex floats: NW = 8 for R <= 3, at most 6 for R = 4.
template <int R, int NW, int KC>
static void mx_apply_launch(const csem_model *m, const MixState *x, const MixSlot *q, const MxFreq &fr, const c2 *in,
                            c2 *out, hipStream_t s, const double *sc, double bytes)
{
  const int nti = (m->n1 + 1 + 61) / 62, ntj = (m->n2 + 1 + NW - 3) / (NW - 2), ntk = (m->n3 + 1 + KC - 1) / KC;
  const int ntiles = nti * ntj * ntk;
  kt_launch(KT_M_APPLY, s, bytes, k_mx_apply<R, NW, KC>, 8 * ((ntiles + 7) / 8), dim3(64, NW), 0, mx_geo(m, x), fr,
            in, out, q->hxg, q->hyg, nti, ntj, ntk, sc);
}
template <int R>
static void mx_apply_tile(const csem_model *m, const MixState *x, const MixSlot *q, const MxFreq &fr, const c2 *in,
                          c2 *out, int rows, int kc, hipStream_t s, const double *sc, double bytes)
{
  constexpr int NB = R > 3 ? 6 : 8;  // the larger block
#define MX_AP(KC)                                                                         \
  do {                                                                                    \
    if (rows + 2 >= NB) mx_apply_launch<R, NB, KC>(m, x, q, fr, in, out, s, sc, bytes);     \
    else mx_apply_launch<R, NB - 2, KC>(m, x, q, fr, in, out, s, sc, bytes);                \
  } while (0)
  switch (kc) {
    case 10: MX_AP(10); break;
    case 12: MX_AP(12); break;
    case 16: MX_AP(16); break;
    default: MX_AP(8); break;
  }
#undef MX_AP
}
static void mx_apply_kernel(const csem_model *m, const MixState *x, const MixSlot *q, int R, const MxFreq &fr,
                            const c2 *in, c2 *out, int rows, int kc, hipStream_t s, const double *sc = nullptr,
                            double bytes = 0.0)
{
  switch (R) {
    case 1: mx_apply_tile<1>(m, x, q, fr, in, out, rows, kc, s, sc, bytes); break;
    case 2: mx_apply_tile<2>(m, x, q, fr, in, out, rows, kc, s, sc, bytes); break;
    case 3: mx_apply_tile<3>(m, x, q, fr, in, out, rows, kc, s, sc, bytes); break;
    default: mx_apply_tile<4>(m, x, q, fr, in, out, rows, kc, s, sc, bytes); break;
  }
}
// output rows per block the operator offers at batch size R
static void mx_tile_rows(int R, int rows[2]) { rows[0] = R > 3 ? 2 : 4; rows[1] = R > 3 ? 4 : 6; }

// Autotuner: time every (rows per block, planes per block) candidate of the
// FP32 operator on this model and batch size, keep the fastest.  Tile shape
// changes no arithmetic: every output is formed by the same expression.
static int mx_tune(csem_model *m, MixState *x, MixSlot *q, int R, const MxFreq &fr, const c2 *in, c2 *out,
                   hipStream_t s)
{
  std::lock_guard<std::mutex> lk(x->mu);
  if (x->ap_tj[R]) return 0;
  int TJS[2];
  mx_tile_rows(R, TJS);
  const int KCS[4] = {8, 10, 12, 16};
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return csem_fail(-1, "event");
  float best = 1e30f;
  int bt = TJS[1], bk = 8;
  kt_pause(1);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 4; b++) {
      mx_apply_kernel(m, x, q, R, fr, in, out, TJS[a], KCS[b], s);  // warm
      (void)hipEventRecord(e0, s);
      for (int w = 0; w < 3; w++) mx_apply_kernel(m, x, q, R, fr, in, out, TJS[a], KCS[b], s);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) { best = ms; bt = TJS[a]; bk = KCS[b]; }
    }
  kt_pause(0);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  x->ap_kc[R] = bk;
  x->ap_tj[R] = bt;
  HIP_OK(hipGetLastError());
  return 0;
}

// y = A x for the R right-hand sides, FP32, with the airwave boundary.
static int mx_apply(csem_model *m, MixState *x, MixSlot *q, int R, const MxFreq &fr, const c2 *in, c2 *out,
                    hipStream_t s, const double *sc = nullptr)
{
  if (m->airbc) {
    const int nh = m->n1 * m->n2, N = m->nfft;
    MxAir a = mx_air(m, x, q, R);
    a.sc = sc;
    kt_launch(KT_M_HZ0, s, 24.0 * nh * kt_rhs(R), k_mx_hz0, dim3((nh + 255) / 256, R), 256, 0, mx_geo(m, x), in,
              q->hz0, sc);
    kt_launch(KT_M_AIR_ROWS, s, 8.0 * N * N * kt_rhs(R), k_mx_air_rows, dim3(N, R), 256,
              3 * (size_t)N * sizeof(c2), a);
    kt_launch(KT_M_AIR_COLS, s, 40.0 * N * N * kt_rhs(R), k_mx_air_cols, dim3(8 * ((N / a.cw + 7) / 8), R),
              N * a.cw / 16, (2 * (size_t)N * a.cw + N) * sizeof(c2), a);
    kt_launch(KT_M_AIR_RINV, s, 16.0 * N * (m->nrowx + m->nrowy) * kt_rhs(R), k_mx_air_rows_inv,
              dim3(m->nrowx + m->nrowy, R), 256, 3 * (size_t)N * sizeof(c2), a);
    const int ng = (m->n1 + 1) * m->n2 + m->n1 * (m->n2 + 1);
    kt_launch(KT_M_AIR_GHOST, s, 40.0 * ng * kt_rhs(R), k_mx_air_ghost, dim3((ng + 255) / 256, R), 256, 0, a);
  }
  // algorithmic bytes: read E and write y per right-hand side, sigma once
  const int tj = x->ap_tj[R] ? x->ap_tj[R] : 6, kc = x->ap_kc[R] ? x->ap_kc[R] : 8;
  mx_apply_kernel(m, x, q, R, fr, in, out, tj, kc, s, sc, (16.0 * kt_rhs(R) + 4.0) * m->lE);
  HIP_OK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------- LU-SGS
static MxLu mx_lu(const csem_model *m, const MixState *x, MixSlot *q, const MxFreq &fr, const c2 *b, c2 *y)
{
  MxLu a;
  a.g = mx_geo(m, x); a.sk = m->g;
  a.b = b; a.y = y; a.W = q->W; a.fE = m->fE; a.fEx = m->fEx; a.fEy = m->fEy;
  a.T = q->tt; a.h = q->hh; a.V = q->V; a.pk = mx_pad(m);
  a.c1m = x->c1m; a.c1p = x->c1p; a.c2m = x->c2m; a.c2p = x->c2p; a.c3m = x->c3m; a.c3p = x->c3p;
  a.fr = fr;
  a.sc = nullptr;
  return a;
}

// blocks of a tile kernel over component c (mx_tile_xyz): Ex tiles span
// slices, Ey and Ez tiles TQ slices; a 1-D grid, a multiple of the 8 XCDs
template <class TL>
static dim3 mx_tile_grid(const SkewGeo &g, int c)
{
  long n;
  if (c == 0) n = (long)((g.ns[0] + TL::TQ - 1) / TL::TQ) * ((g.B[0] + TL::TF - 1) / TL::TF) * ((g.A[0] + TL::TS - 1) / TL::TS);
  else n = (long)((g.B[c] + TL::TF - 1) / TL::TF) * ((g.A[c] + TL::TS - 1) / TL::TS) * ((g.ns[c] + TL::TQ - 1) / TL::TQ);
  return dim3((unsigned)(8 * ((n + 7) / 8)));
}

// the same tiles as a 3-D grid (x, y, z as in mx_tile_xyz)
template <class TL>
static dim3 mx_tile_grid3(const SkewGeo &g, int c)
{
  if (c == 0)
    return dim3((unsigned)((g.ns[0] + TL::TQ - 1) / TL::TQ), (unsigned)((g.B[0] + TL::TF - 1) / TL::TF),
                (unsigned)((g.A[0] + TL::TS - 1) / TL::TS));
  return dim3((unsigned)((g.B[c] + TL::TF - 1) / TL::TF), (unsigned)((g.A[c] + TL::TS - 1) / TL::TS),
              (unsigned)((g.ns[c] + TL::TQ - 1) / TL::TQ));
}

// the sweep for slices of tb / 64 waves
template <int C, bool FWD>
static void mx_sgs_launch(dim3 g, int tb, size_t lds, hipStream_t s, const MxLu &a, double bytes)
{
  const int nt = tb + (C != 0 ? 64 * mxs_store_waves(tb / 64) : 0);  // Ey, Ez: and the store waves
  switch (tb / 64) {
    case 1: kt_launch(KT_M_SGS, s, bytes, k_mx_sgs<C, FWD, 1>, g, nt, lds, a); break;
    case 2: kt_launch(KT_M_SGS, s, bytes, k_mx_sgs<C, FWD, 2>, g, nt, lds, a); break;
    case 3: kt_launch(KT_M_SGS, s, bytes, k_mx_sgs<C, FWD, 3>, g, nt, lds, a); break;
    default: kt_launch(KT_M_SGS, s, bytes, k_mx_sgs<C, FWD, 4>, g, nt, lds, a); break;
  }
}

// one pre-pass launch with tile candidate v (MxPreTileV)
template <int C, bool FWD, int R>
static void mx_pre_launch(const SkewGeo &g, int v, hipStream_t s, const MxLu &a, double bytes)
{
  switch (v) {
    case 1: kt_launch(KT_M_PRE, s, bytes, k_mx_pre<C, FWD, R, 1>, mx_tile_grid<MxPreTileV<C, FWD, 1>>(g, C), 256, 0, a); break;
    case 2: kt_launch(KT_M_PRE, s, bytes, k_mx_pre<C, FWD, R, 2>, mx_tile_grid<MxPreTileV<C, FWD, 2>>(g, C), 256, 0, a); break;
    default: kt_launch(KT_M_PRE, s, bytes, k_mx_pre<C, FWD, R, 0>, mx_tile_grid<MxPreTileV<C, FWD, 0>>(g, C), 256, 0, a); break;
  }
}
// pre-pass number kp of an application (Ex, Ey, Ez forward, Ez, Ey, Ex
// backward) with tile candidate v
template <int R>
static void mx_pre_k(const SkewGeo &g, int kp, int v, hipStream_t s, const MxLu &a, double bytes)
{
  switch (kp) {
    case 0: mx_pre_launch<0, true, R>(g, v, s, a, bytes); break;
    case 1: mx_pre_launch<1, true, R>(g, v, s, a, bytes); break;
    case 2: mx_pre_launch<2, true, R>(g, v, s, a, bytes); break;
    case 3: mx_pre_launch<2, false, R>(g, v, s, a, bytes); break;
    case 4: mx_pre_launch<1, false, R>(g, v, s, a, bytes); break;
    default: mx_pre_launch<0, false, R>(g, v, s, a, bytes); break;
  }
}
static void mx_pre_any(const SkewGeo &g, int R, int kp, int v, hipStream_t s, const MxLu &a)
{
  switch (R) {
    case 1: mx_pre_k<1>(g, kp, v, s, a, 0.0); break;
    case 2: mx_pre_k<2>(g, kp, v, s, a, 0.0); break;
    case 3: mx_pre_k<3>(g, kp, v, s, a, 0.0); break;
    default: mx_pre_k<4>(g, kp, v, s, a, 0.0); break;
  }
}

template <int R>
static void mx_lusgs_R(const csem_model *m, const MxLu &a, const int *pv, hipStream_t s)
{
  const SkewGeo &g = m->g;
  const MxPad &pk = a.pk;
  // tile grids of the copies (MxCopyTile, 3-D)
  const dim3 gu[3] = {mx_tile_grid3<MxCopyTile<0, R>>(g, 0), mx_tile_grid3<MxCopyTile<1, R>>(g, 1),
                      mx_tile_grid3<MxCopyTile<2, R>>(g, 2)};
  int tb[3];
  size_t lds[3];
  dim3 gs[3];
  double nc[3];
  for (int c = 0; c < 3; c++) {
    tb[c] = pk.AP[c];
    lds[c] = (size_t)mxs_lds_bytes(g.A[c], g.B[c], pk.AP[c], c);
    gs[c] = dim3(8 * ((g.ns[c] + 7) / 8), R);
    nc[c] = (double)g.ns[c] * g.A[c] * g.B[c];
  }
  // bytes per swept cell and right-hand side each kernel must move: the
  // pre-pass reads b and the other components' W values it differences
  // and writes T; the sweep reads T and h and, for Ey and Ez, writes the
  // value to W and (backward) y; for Ex it writes the value to V, and the
  // copy reads V and writes W (forward) or y (backward)
  const double r = kt_rhs(R);
  mx_pre_launch<0, true, R>(g, pv[0], s, a, 16.0 * r * nc[0]);
  mx_sgs_launch<0, true>(gs[0], tb[0], lds[0], s, a, 24.0 * r * nc[0]);
  kt_launch(KT_M_UNSKEW, s, 16.0 * r * nc[0], k_mx_unskew<0, true, R>, gu[0], 256, 0, a);
  mx_pre_launch<1, true, R>(g, pv[1], s, a, 24.0 * r * nc[1]);
  mx_sgs_launch<1, true>(gs[1], tb[1], lds[1], s, a, 24.0 * r * nc[1]);
  mx_pre_launch<2, true, R>(g, pv[2], s, a, 32.0 * r * nc[2]);
  mx_sgs_launch<2, true>(gs[2], tb[2], lds[2], s, a, 24.0 * r * nc[2]);
  mx_pre_launch<2, false, R>(g, pv[3], s, a, 40.0 * r * nc[2]);
  mx_sgs_launch<2, false>(gs[2], tb[2], lds[2], s, a, 32.0 * r * nc[2]);
  mx_pre_launch<1, false, R>(g, pv[4], s, a, 40.0 * r * nc[1]);
  mx_sgs_launch<1, false>(gs[1], tb[1], lds[1], s, a, 32.0 * r * nc[1]);
  mx_pre_launch<0, false, R>(g, pv[5], s, a, 40.0 * r * nc[0]);
  mx_sgs_launch<0, false>(gs[0], tb[0], lds[0], s, a, 24.0 * r * nc[0]);
  kt_launch(KT_M_UNSKEW, s, 16.0 * r * nc[0], k_mx_unskew<0, false, R>, gu[0], 256, 0, a);
}

// The red-black ordered sweeps: forward red, black per component (Ex, Ey,
// Ez), backward black, red (Ez, Ey, Ex).  Bytes per swept cell and
// right-hand side each pass must move (half the cells per pass): b, the
// other components' W values, the own component's neighbours (except the
// forward red pass), the value written (and y backward).
template <int R>
static void mx_rbsgs_R(const csem_model *m, const MxLu &a, hipStream_t s)
{
  const SkewGeo &g = m->g;
  dim3 gq[3];
  gq[0] = dim3((m->n1 + 63) / 64, (m->n2 + 1 + 3) / 4, m->n3f + 1);
  gq[1] = dim3((m->n1 + 1 + 63) / 64, (m->n2 + 3) / 4, m->n3f + 1);
  gq[2] = dim3((m->n1 + 1 + 63) / 64, (m->n2 + 1 + 3) / 4, m->n3f);
  const dim3 bq(64, 4);
  double h[3];
  for (int c = 0; c < 3; c++) h[c] = 0.5 * g.ns[c] * g.A[c] * g.B[c] * kt_rhs(R);
  kt_launch(KT_M_SGS, s, 16.0 * h[0], k_mx_rb<0, true, 0, R>, gq[0], bq, 0, a);
  kt_launch(KT_M_SGS, s, 24.0 * h[0], k_mx_rb<0, true, 1, R>, gq[0], bq, 0, a);
  kt_launch(KT_M_SGS, s, 24.0 * h[1], k_mx_rb<1, true, 0, R>, gq[1], bq, 0, a);
  kt_launch(KT_M_SGS, s, 32.0 * h[1], k_mx_rb<1, true, 1, R>, gq[1], bq, 0, a);
  kt_launch(KT_M_SGS, s, 32.0 * h[2], k_mx_rb<2, true, 0, R>, gq[2], bq, 0, a);
  kt_launch(KT_M_SGS, s, 40.0 * h[2], k_mx_rb<2, true, 1, R>, gq[2], bq, 0, a);
  kt_launch(KT_M_SGS, s, 48.0 * h[2], k_mx_rb<2, false, 1, R>, gq[2], bq, 0, a);
  kt_launch(KT_M_SGS, s, 48.0 * h[2], k_mx_rb<2, false, 0, R>, gq[2], bq, 0, a);
  kt_launch(KT_M_SGS, s, 48.0 * h[1], k_mx_rb<1, false, 1, R>, gq[1], bq, 0, a);
  kt_launch(KT_M_SGS, s, 48.0 * h[1], k_mx_rb<1, false, 0, R>, gq[1], bq, 0, a);
  kt_launch(KT_M_SGS, s, 48.0 * h[0], k_mx_rb<0, false, 1, R>, gq[0], bq, 0, a);
  kt_launch(KT_M_SGS, s, 48.0 * h[0], k_mx_rb<0, false, 0, R>, gq[0], bq, 0, a);
}

static void mx_rbsgs_dispatch(const csem_model *m, int R, const MxLu &a, hipStream_t s)
{
  switch (R) {
    case 1: mx_rbsgs_R<1>(m, a, s); break;
    case 2: mx_rbsgs_R<2>(m, a, s); break;
    case 3: mx_rbsgs_R<3>(m, a, s); break;
    default: mx_rbsgs_R<4>(m, a, s); break;
  }
}

static void mx_lusgs_dispatch(const csem_model *m, int R, const MxLu &a, const int *pv, hipStream_t s)
{
  switch (R) {
    case 1: mx_lusgs_R<1>(m, a, pv, s); break;
    case 2: mx_lusgs_R<2>(m, a, pv, s); break;
    case 3: mx_lusgs_R<3>(m, a, pv, s); break;
    default: mx_lusgs_R<4>(m, a, pv, s); break;
  }
}

// The records' per-frequency half (h) of a slot for the frequencies fr,
// computed when a right-hand side's frequency differs from the one its h
// holds.  A solve calls this before it captures an iteration, so a
// captured graph never holds it.
static int mx_hrec(csem_model *m, MixState *x, MixSlot *q, int R, const MxFreq &fr, hipStream_t s)
{
  bool need = false;
  for (int r = 0; r < R; r++)
    need = need || !q->h_ok[r] || q->h_iw[r].x != fr.iwmu[r].x || q->h_iw[r].y != fr.iwmu[r].y;
  if (!need) return 0;
  const MxLu a = mx_lu(m, x, q, fr, nullptr, nullptr);
  const SkewGeo &g = m->g;
  k_mx_hrec<0><<<dim3((unsigned)((a.pk.SL[0] + 255) / 256), g.ns[0]), 256, 0, s>>>(a, R);
  k_mx_hrec<1><<<dim3((unsigned)((a.pk.SL[1] + 255) / 256), g.ns[1]), 256, 0, s>>>(a, R);
  k_mx_hrec<2><<<dim3((unsigned)((a.pk.SL[2] + 255) / 256), g.ns[2]), 256, 0, s>>>(a, R);
  HIP_OK(hipGetLastError());
  for (int r = 0; r < R; r++) { q->h_iw[r] = fr.iwmu[r]; q->h_ok[r] = 1; }
  return 0;
}

// sc: the inner solve's scalars (right-hand sides done are skipped), or NULL
static int mx_lusgs(csem_model *m, MixState *x, MixSlot *q, int R, const MxFreq &fr, const c2 *b, c2 *y,
                    hipStream_t s, const double *sc = nullptr)
{
  MxLu a = mx_lu(m, x, q, fr, b, y);
  a.sc = sc;
  int rc;
  if (q->pc != CSEM_PRECOND_RBSGS && (rc = mx_hrec(m, x, q, R, fr, s))) return rc;
  if (q->pc == CSEM_PRECOND_RBSGS) mx_rbsgs_dispatch(m, R, a, s);
  else mx_lusgs_dispatch(m, R, a, x->pre_v[R], s);
  HIP_OK(hipGetLastError());
  return 0;
}

// Autotuner, second part: each of the six LU-SGS pre-passes over its tile
// candidates (MxPreTileV) and the FFT column pass over its columns per
// block, timed on this model and batch size; the fastest are kept.  Neither
// choice changes a result: a pre-pass record and a column's transform are
// the same expressions whatever the tile.  The pre-passes only write the
// slot's records (which every application rewrites before a sweep reads
// them) and the column pass only the slot's transform scratch.
static int mx_tune_lu(csem_model *m, MixState *x, MixSlot *q, int R, const MxFreq &fr, const c2 *in, c2 *out,
                      hipStream_t s)
{
  std::lock_guard<std::mutex> lk(x->mu);
  if (x->pre_v[R][0] >= 0 && x->air_cw[R]) return 0;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return csem_fail(-1, "event");
  auto timed = [&](auto launch) {
    launch();  // warm
    (void)hipEventRecord(e0, s);
    for (int w = 0; w < 3; w++) launch();
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return ms;
  };
  kt_pause(1);
  const MxLu a = mx_lu(m, x, q, fr, in, out);
  for (int kp = 0; kp < 6; kp++) {
    float best = 1e30f;
    int bv = 0;
    for (int v = 0; v < MX_PRE_NV; v++) {
      const float ms = timed([&] { mx_pre_any(m->g, R, kp, v, s, a); });
      if (ms < best) { best = ms; bv = v; }
    }
    x->pre_v[R][kp] = bv;
  }
  int bc = x->cw;
  if (m->airbc) {
    const int N = m->nfft;
    MxAir ar = mx_air(m, x, q, R);
    float best = 1e30f;
    for (int c = std::max(1, 1024 / N); c <= x->cw_max; c *= 2) {
      ar.cw = c;
      const float ms = timed([&] {
        k_mx_air_cols<<<dim3(8 * ((N / c + 7) / 8), R), N * c / 16, (2 * (size_t)N * c + N) * sizeof(c2), s>>>(ar);
      });
      if (ms < best) { best = ms; bc = c; }
    }
  }
  x->air_cw[R] = bc;
  kt_pause(0);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIP_OK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- Krylov
static MxKry mx_kry(const csem_model *m, const MixSlot *q)
{
  MxKry k;
  const long n = m->lE, R = q->R;
  k.n = n; k.nch = q->nch;
  c2 *v = q->kv;
  k.r = v; k.r0 = v + R * n; k.p = v + 2 * R * n; k.pp = v + 3 * R * n; k.v = v + 4 * R * n;
  k.s = v + 5 * R * n; k.ss = v + 6 * R * n; k.t = v + 7 * R * n; k.x = v + 8 * R * n;
  k.part = q->part; k.sc = q->sc;
  return k;
}

// one F-BiCGStab iteration for every active right-hand side (solver.c:139-177)
static int mx_iteration(csem_model *m, MixState *x, MixSlot *q, int R, const MxFreq &fr, hipStream_t s)
{
  const MxKry k = mx_kry(m, q);
  const dim3 gr(q->nch, R), ge((unsigned)((k.n + 255) / 256), R);
  int rc;
  // in a batch, the operator and preconditioner skip the right-hand sides
  // whose inner solve is done (one alone has nothing to skip)
  const double *live = R > 1 ? k.sc : nullptr;
  if ((rc = mx_lusgs(m, x, q, R, fr, k.p, k.pp, s, live))) return rc;
  if ((rc = mx_apply(m, x, q, R, fr, k.pp, k.v, s, live))) return rc;
  const double r = kt_rhs(R);
  kt_launch(KT_M_DOT, s, 16.0 * r * k.n, k_mx_k_dot_vr0, gr, RED_T, 0, k);
  kt_launch(KT_M_DOT, s, 0.0, k_mx_k_fin, R, RED_T, 0, k, 1);
  kt_launch(KT_M_VEC, s, 24.0 * r * k.n, k_mx_k_s, ge, 256, 0, k);
  if ((rc = mx_lusgs(m, x, q, R, fr, k.s, k.ss, s, live))) return rc;
  if ((rc = mx_apply(m, x, q, R, fr, k.ss, k.t, s, live))) return rc;
  kt_launch(KT_M_DOT, s, 16.0 * r * k.n, k_mx_k_dot_st, gr, RED_T, 0, k);
  kt_launch(KT_M_DOT, s, 0.0, k_mx_k_fin, R, RED_T, 0, k, 2);
  kt_launch(KT_M_VEC, s, 64.0 * r * k.n, k_mx_k_xr, gr, RED_T, 0, k);
  kt_launch(KT_M_DOT, s, 0.0, k_mx_k_fin, R, RED_T, 0, k, 3);
  kt_launch(KT_M_VEC, s, 32.0 * r * k.n, k_mx_k_p, ge, 256, 0, k);
  HIP_OK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- FGMRES
static void slot_drop(MixSlot *q, void *p, long bytes)
{
  if (!p) return;
  auto it = std::find(q->allocs.begin(), q->allocs.end(), p);
  if (it != q->allocs.end()) q->allocs.erase(it);
  dev_free(p);
  q->bytes -= std::max(bytes, 256L);
}

// FGMRES(M) buffers of a slot: M + 1 basis vectors and M preconditioned
// ones per right-hand side, the chunk partials and the Hessenberg state
static int gm_alloc(const csem_model *m, MixSlot *q, int M)
{
  if (q->gm_m >= M) return 0;
  const long n = m->lE, R = q->R, mo = q->gm_m;
  if (mo) {
    slot_drop(q, q->gmV, (mo + 1) * R * n * (long)sizeof(c2));
    slot_drop(q, q->gmZ, mo * R * n * (long)sizeof(c2));
    slot_drop(q, q->gmpart, R * q->nch * GM_PS * (long)sizeof(double));
    slot_drop(q, q->gmhs, R * GM_HS * (long)sizeof(double));
  }
  q->gm_m = 0;
  q->gmV = q->gmZ = NULL;
  q->gmpart = q->gmhs = NULL;
  int rc;
  if ((rc = mx_dalloc(q, &q->gmV, (M + 1) * R * n)) || (rc = mx_dalloc(q, &q->gmZ, (long)M * R * n)) ||
      (rc = mx_dalloc(q, &q->gmpart, R * q->nch * GM_PS)) || (rc = mx_dalloc(q, &q->gmhs, R * GM_HS)))
    return rc;
  q->gm_m = M;
  return 0;
}

static MxGm mx_gmv(const csem_model *m, const MixSlot *q, int R, int M)
{
  const MxKry k = mx_kry(m, q);
  MxGm g;
  g.n = m->lE; g.nch = q->nch; g.m = M; g.R = R;
  g.V = q->gmV; g.Z = q->gmZ; g.r = k.r; g.x = k.x;
  g.part = q->gmpart; g.sc = q->sc; g.hs = q->gmhs;
  return g;
}

// the three orthogonalisation passes of step j and their scalar steps
// (bytes: the vectors each pass reads and writes)
template <int MB>
static void gm_orth(const MxGm &g, int R, int j, dim3 gr, dim3 ge, double rr, hipStream_t s)
{
  const double n = (double)g.n, nb = j + 2.0;
  kt_launch(KT_M_ORTH, s, 8.0 * nb * rr * n, k_gm_dot<MB>, gr, RED_T, 0, g, j);
  kt_launch(KT_M_ORTH, s, 0.0, k_gm_fin, R, RED_T, 0, g, j, 1);
  kt_launch(KT_M_ORTH, s, 8.0 * (nb + 1) * rr * n, k_gm_upd1<MB>, gr, RED_T, 0, g, j);
  kt_launch(KT_M_ORTH, s, 0.0, k_gm_fin, R, RED_T, 0, g, j, 2);
  kt_launch(KT_M_ORTH, s, 8.0 * (nb + 1) * rr * n, k_gm_upd2<MB>, ge, 256, 0, g, j);
}

// One FGMRES(M) cycle from x = 0 on the inner right-hand side k.r for
// every right-hand side taking part (solver.c:426-505 with CGS2); the
// correction goes to k.x and MS_IT holds the steps taken.
static int mx_gm_cycle(csem_model *m, MixState *x, MixSlot *q, int R, const MxFreq &fr, int M, hipStream_t s)
{
  const MxGm g = mx_gmv(m, q, R, M);
  const long n = m->lE;
  const dim3 gr(q->nch, R), ge((unsigned)((n + 255) / 256), R);
  const double *live = R > 1 ? q->sc : nullptr;
  const double rr = kt_rhs(R);
  int rc;
  kt_launch(KT_M_ORTH, s, 8.0 * rr * n, k_gm_norm, gr, RED_T, 0, g);
  kt_launch(KT_M_ORTH, s, 0.0, k_gm_fin, R, RED_T, 0, g, 0, 0);
  kt_launch(KT_M_VEC, s, 16.0 * rr * n, k_gm_v0, ge, 256, 0, g);
  for (int j = 0; j < M; j++) {
    c2 *vj = q->gmV + (long)j * R * n, *zj = q->gmZ + (long)j * R * n, *w = q->gmV + (long)(j + 1) * R * n;
    if ((rc = mx_lusgs(m, x, q, R, fr, vj, zj, s, live))) return rc;
    if ((rc = mx_apply(m, x, q, R, fr, zj, w, s, live))) return rc;
    if (j < 4) gm_orth<4>(g, R, j, gr, ge, rr, s);
    else if (j < 8) gm_orth<8>(g, R, j, gr, ge, rr, s);
    else gm_orth<16>(g, R, j, gr, ge, rr, s);
  }
  kt_launch(KT_M_ORTH, s, 0.0, k_gm_solve, R, 64, 0, g);
  kt_launch(KT_M_VEC, s, 8.0 * (M + 1) * rr * n, k_gm_xupd, ge, 256, 0, g);
  HIP_OK(hipGetLastError());
  return 0;
}

// Capture one F-BiCGStab iteration (kind 0) or one FGMRES(kind) cycle as a
// graph (kernel launches only: the frequencies and every scalar live in
// device memory or in the captured arguments).
static int mx_graph(csem_model *m, MixState *x, MixSlot *q, int R, const MxFreq &fr, hipStream_t s, int kind)
{
  if (q->gexec && q->g_R == R && q->g_stream == s && q->g_pc == q->pc && q->g_kind == kind &&
      !memcmp(&q->g_fr, &fr, sizeof fr))
    return 0;
  if (q->gexec) { (void)hipGraphExecDestroy(q->gexec); q->gexec = NULL; }
  hipGraph_t g;
  int rc;
  hipError_t e;
  {
    std::lock_guard<std::mutex> lk(g_devwide_mu);
    kt_pause(1);
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      kt_pause(0);
      return csem_fail(-1, "graph capture: %s", hipGetErrorString(e));
    }
    rc = kind ? mx_gm_cycle(m, x, q, R, fr, kind, s) : mx_iteration(m, x, q, R, fr, s);
    e = hipStreamEndCapture(s, &g);
    kt_pause(0);
  }
  if (rc) return rc;
  if (e != hipSuccess) return csem_fail(-1, "graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&q->gexec, g, NULL, NULL, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { q->gexec = NULL; return csem_fail(-1, "graph instantiate: %s", hipGetErrorString(e)); }
  q->g_R = R;
  q->g_stream = s;
  q->g_fr = fr;
  q->g_pc = q->pc;
  q->g_kind = kind;
  return 0;
}

// ------------------------------------------------------------ refinement
// res_r = b_r - w (w = A x_r in FP64) and the chunk partials of |res_r|^2
__global__ __launch_bounds__(RED_T) void k_mx_res64(long n, const z2 *b, const z2 *w, z2 *res, double *npart, int nch)
{
  __shared__ double sh[RED_T];
  double acc = 0;
  const long base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int m = 0; m < RED_E; m++) {
    const long e = base + (long)m * RED_T;
    if (e < n) {
      const z2 v = zsub(b[e], w[e]);
      res[e] = v;
      acc += v.x * v.x + v.y * v.y;
    }
  }
  const double v = block_fold(acc, sh);
  if (threadIdx.x == 0) npart[blockIdx.x] = v;
}
// |v_r|^2 for every r (chunk partials then one block per r)
__global__ __launch_bounds__(RED_T) void k_mx_norm_part(long n, const z2 *v, double *npart, int nch)
{
  __shared__ double sh[RED_T];
  const int r = blockIdx.y;
  double acc = 0;
  const long base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int m = 0; m < RED_E; m++) {
    const long e = base + (long)m * RED_T;
    if (e < n) { const z2 q = v[r * n + e]; acc += q.x * q.x + q.y * q.y; }
  }
  const double s = block_fold(acc, sh);
  if (threadIdx.x == 0) npart[(long)r * nch + blockIdx.x] = s;
}
__global__ __launch_bounds__(RED_T) void k_mx_norm_fin(const double *npart, int nch, double *out)
{
  __shared__ double sh[RED_T];
  const int r = blockIdx.x;
  double acc = 0;
  for (int c = threadIdx.x; c < nch; c += RED_T) acc += npart[(long)r * nch + c];
  const double s = block_fold(acc, sh);
  if (threadIdx.x == 0) out[r] = s;
}

#define MX_STALL_ITERS 100  // default of csem_mixed_opts.stall_iters

static int mixed_solve_slot(csem_model *m, MixState *x, MixSlot *q, int R, const csem_freq *f, z2 *X, const z2 *B,
                            const double *tol, const csem_mixed_opts &o, int *iters, double *relres, hipStream_t s)
{
  const long n = m->lE;
  const MxFreq fr = mx_freq(f, R);
  const MxKry k = mx_kry(m, q);
  const int nch = q->nch;
  const unsigned gn = (unsigned)((n + 255) / 256);
  std::vector<double> bn(R), rn(R), hsc((size_t)R * MS_N), nb(R);
  std::vector<char> stalled(R, 0);  // a correction that did not lower the true residual was rejected
  std::vector<char> cut(R, 0);      // this step's F-BiCGStab inner solve was cut for stagnation
  std::vector<char> kept(R, 0);     // and its best iterate is in xbest
  std::vector<float> scl(R);
  int rc;
  if (o.precond != CSEM_PRECOND_LUSGS && o.precond != CSEM_PRECOND_RBSGS)
    return csem_fail(-3, "unknown preconditioner %d", o.precond);
  if (o.krylov != CSEM_KRYLOV_FBICGSTAB && o.krylov != CSEM_KRYLOV_FGMRES)
    return csem_fail(-3, "unknown inner Krylov method %d", o.krylov);
  const int M = o.restart == 0 ? 10 : o.restart;
  if (M < 1 || M > GM_MMAX) return csem_fail(-3, "FGMRES restart %d outside [1, %d]", M, GM_MMAX);
  q->pc = o.precond;
  // |b_r|, x = 0, res = b
  k_mx_norm_part<<<dim3(nch, R), RED_T, 0, s>>>(n, B, q->npart, nch);
  k_mx_norm_fin<<<R, RED_T, 0, s>>>(q->npart, nch, q->norm);
  HIP_OK(hipMemsetAsync(X, 0, (size_t)R * n * sizeof(z2), s));
  HIP_OK(hipMemcpyAsync(q->res, B, (size_t)R * n * sizeof(z2), hipMemcpyDeviceToDevice, s));
  HIP_OK(hipMemcpyAsync(nb.data(), q->norm, R * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  for (int r = 0; r < R; r++) { bn[r] = sqrt(nb[r]); rn[r] = bn[r]; relres[r] = 1.0; iters[r] = 0; }
  // operands: any R x n FP32 buffers
  if ((rc = mx_tune(m, x, q, R, fr, k.r, k.v, s)) || (rc = mx_tune_lu(m, x, q, R, fr, k.r, k.v, s))) return rc;
  if (o.precond == CSEM_PRECOND_LUSGS && (rc = mx_hrec(m, x, q, R, fr, s))) return rc;
  // Phase 0 refines with the inner method asked for.  A right-hand side
  // whose F-BiCGStab inner solve was cut for stagnation and whose
  // correction the guard still rejected goes on in phase 1 with FGMRES(M)
  // cycles from its best solution: GMRES minimises the residual over its
  // Krylov space, so a cycle cannot diverge where BiCGStab in FP32 did.
  std::vector<char> on(R, 1), cutrej(R, 0);
  for (int phase = 0; phase < 2; phase++) {
    const bool gm = phase == 1 || o.krylov == CSEM_KRYLOV_FGMRES;
    if (phase == 1) {
      bool any1 = false;
      for (int r = 0; r < R; r++) {
        on[r] = cutrej[r] && iters[r] < o.max_inner && relres[r] >= tol[r];
        if (on[r]) stalled[r] = 0;
        any1 |= on[r] != 0;
      }
      if (!any1) break;
    }
    if (gm && (rc = gm_alloc(m, q, M))) return rc;
    // FGMRES: each refinement step is one restart cycle of M steps, bounded
    // by the inner budget alone
    const int max_outer = gm ? o.max_inner + 1 : o.max_outer;
    for (int outer = 0; outer <= max_outer; outer++) {
      // which right-hand sides still need a correction
      bool any = false;
      for (int r = 0; r < R; r++) {
        double *sc = &hsc[(size_t)r * MS_N];
        memset(sc, 0, MS_N * sizeof(double));
        const bool act = on[r] && bn[r] > 0 && relres[r] >= tol[r] && iters[r] < o.max_inner && outer < max_outer &&
                         !stalled[r];
        sc[MS_DONE] = act ? 0.0 : 1.0;
        sc[MS_USED] = act ? 1.0 : 0.0;
        // reduce this residual by what is still needed (with a margin), at most to inner_tol
        sc[MS_TOL] = std::max(o.inner_tol, 0.3 * tol[r] / std::max(relres[r], 1e-300));
        sc[MS_MAXIT] = o.max_inner - iters[r];
        scl[r] = act ? (float)(1.0 / rn[r]) : 0.f;
        any |= act;
      }
      if (!any) break;
      int nact = 0;
      for (int r = 0; r < R; r++) nact += hsc[(size_t)r * MS_N + MS_DONE] == 0.0;
      g_kt_live = nact;
      HIP_OK(hipMemcpyAsync(q->sc, hsc.data(), hsc.size() * sizeof(double), hipMemcpyHostToDevice, s));
      HIP_OK(hipMemcpyAsync(q->scale, scl.data(), R * sizeof(float), hipMemcpyHostToDevice, s));
      kt_launch(KT_M_REFINE, s, 24.0 * R * n, k_mx_to32, dim3(gn, R), 256, 0, n, q->res, q->scale, k.r);
      if (gm) {
        // one FGMRES(M) cycle on the scaled residual
        if (o.use_graph) {
          if ((rc = mx_graph(m, x, q, R, fr, s, M))) return rc;
          HIP_OK(hipGraphLaunch(q->gexec, s));
        } else if ((rc = mx_gm_cycle(m, x, q, R, fr, M, s))) {
          return rc;
        }
        HIP_OK(hipMemcpyAsync(hsc.data(), q->sc, hsc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
      } else {
        kt_launch(KT_M_DOT, s, 32.0 * kt_rhs(R) * n, k_mx_k_init, dim3(nch, R), RED_T, 0, k);
        kt_launch(KT_M_DOT, s, 0.0, k_mx_k_fin, R, RED_T, 0, k, 0);
        if (o.use_graph && (rc = mx_graph(m, x, q, R, fr, s, 0))) return rc;
        // inner iterations, convergence checked every check_every; an inner
        // solve whose recursive residual has not improved by 1% in
        // stall_iters iterations (FP32 stagnation on an ill-conditioned
        // system) is ended; its iterate at each 1% improvement is kept as
        // a second candidate correction
        std::vector<double> best(R, 1.0);
        std::vector<double> best_it(R, 0.0);
        std::fill(kept.begin(), kept.end(), 0);
        std::fill(cut.begin(), cut.end(), 0);
        const int stall = o.stall_iters == 0 ? MX_STALL_ITERS : o.stall_iters;
        for (;;) {
          for (int w = 0; w < std::max(1, o.check_every); w++) {
            if (o.use_graph) HIP_OK(hipGraphLaunch(q->gexec, s));
            else if ((rc = mx_iteration(m, x, q, R, fr, s))) return rc;
          }
          HIP_OK(hipMemcpyAsync(hsc.data(), q->sc, hsc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
          HIP_OK(hipStreamSynchronize(s));
          bool all = true;
          for (int r = 0; r < R; r++) {
            double *sc = &hsc[(size_t)r * MS_N];
            if (sc[MS_DONE] == 0.0) {
              if (sc[MS_REL] < 0.99 * best[r]) {
                best[r] = sc[MS_REL];
                best_it[r] = sc[MS_IT];
                if (stall > 0) {
                  HIP_OK(hipMemcpyAsync(q->xbest + r * n, k.x + r * n, n * sizeof(c2), hipMemcpyDeviceToDevice, s));
                  kept[r] = 1;
                }
              } else if (stall > 0 && sc[MS_IT] - best_it[r] >= stall) {
                const double one = 1.0;
                sc[MS_DONE] = 1.0;
                cut[r] = 1;
                HIP_OK(hipMemcpyAsync(q->sc + (size_t)r * MS_N + MS_DONE, &one, sizeof one, hipMemcpyHostToDevice, s));
                HIP_OK(hipStreamSynchronize(s));
              }
            }
            all &= sc[MS_DONE] != 0.0;
          }
          int nlive = 0;
          for (int r = 0; r < R; r++) nlive += hsc[(size_t)r * MS_N + MS_DONE] == 0.0;
          g_kt_live = nlive;
          if (all) break;
        }
      }
      for (int r = 0; r < R; r++)
        if (hsc[(size_t)r * MS_N + MS_USED] != 0.0) iters[r] += (int)hsc[(size_t)r * MS_N + MS_IT];
      // x += |res| * correction (FP64), then the true residual in FP64
      HIP_OK(hipMemcpyAsync(q->xprev, X, (size_t)R * n * sizeof(z2), hipMemcpyDeviceToDevice, s));
      kt_launch(KT_M_REFINE, s, 32.0 * R * n, k_mx_acc64, dim3(gn, R), 256, 0, n, X, k.x, q->sc, q->scale);
      for (int r = 0; r < R; r++) {
        if (hsc[(size_t)r * MS_N + MS_USED] == 0.0) continue;
        if ((rc = run_apply_bufs(m, &f[r], X + r * n, q->tmp64, s, q->a64))) return rc;
        kt_launch(KT_M_REFINE, s, 48.0 * n, k_mx_res64, nch, RED_T, 0, n, B + r * n, q->tmp64, q->res + r * n,
                  q->npart + (long)r * nch, nch);
      }
      k_mx_norm_fin<<<R, RED_T, 0, s>>>(q->npart, nch, q->norm);
      HIP_OK(hipMemcpyAsync(nb.data(), q->norm, R * sizeof(double), hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      for (int r = 0; r < R; r++) {
        if (hsc[(size_t)r * MS_N + MS_USED] == 0.0) continue;
        double rnew = sqrt(nb[r]);
        if (!(rnew < rn[r]) && !gm && cut[r] && kept[r]) {
          // the stagnating inner solve's last iterate did not help: try the
          // iterate of its best recursive residual
          HIP_OK(hipMemcpyAsync(X + r * n, q->xprev + r * n, n * sizeof(z2), hipMemcpyDeviceToDevice, s));
          k_mx_acc64<<<dim3(gn, 1), 256, 0, s>>>(n, X + r * n, q->xbest + r * n, q->sc + r * MS_N, q->scale + r);
          if ((rc = run_apply_bufs(m, &f[r], X + r * n, q->tmp64, s, q->a64))) return rc;
          k_mx_res64<<<nch, RED_T, 0, s>>>(n, B + r * n, q->tmp64, q->res + r * n, q->npart + (long)r * nch, nch);
          k_mx_norm_fin<<<1, RED_T, 0, s>>>(q->npart + (long)r * nch, nch, q->norm + r);
          HIP_OK(hipMemcpyAsync(&nb[r], q->norm + r, sizeof(double), hipMemcpyDeviceToHost, s));
          HIP_OK(hipStreamSynchronize(s));
          rnew = sqrt(nb[r]);
        }
        if (!(rnew < rn[r])) {
          // the FP32 inner solve did not help (stagnation or divergence on an
          // ill-conditioned system): keep the previous solution and residual
          // and stop refining this right-hand side (after a cut F-BiCGStab
          // solve, FGMRES takes it over in phase 1); it is reported as not
          // converged if nothing reaches tol
          stalled[r] = 1;
          cutrej[r] = !gm && cut[r];
          HIP_OK(hipMemcpyAsync(X + r * n, q->xprev + r * n, n * sizeof(z2), hipMemcpyDeviceToDevice, s));
          if ((rc = run_apply_bufs(m, &f[r], X + r * n, q->tmp64, s, q->a64))) return rc;
          k_mx_res64<<<nch, RED_T, 0, s>>>(n, B + r * n, q->tmp64, q->res + r * n, q->npart + (long)r * nch, nch);
          continue;
        }
        rn[r] = rnew;
        relres[r] = bn[r] > 0 ? rn[r] / bn[r] : 0.0;
      }
    }
    if (gm) break;
  }
  g_kt_live = 0;
  HIP_OK(hipGetLastError());
  int conv = 0;
  for (int r = 0; r < R; r++) conv += relres[r] < tol[r];
  return conv;
}

extern "C" int csem_mixed_solve(csem_model *m, int R, const csem_freq *f, void *xv, const void *bv, const double *tol,
                                const csem_mixed_opts *opt, int *iters, double *relres, void *stream)
{
  MixState *x;
  MixSlot *q;
  int rc;
  {
    std::lock_guard<std::mutex> lk(g_mx_init_mu);
    if ((rc = mx_model(m, &x))) return rc;
  }
  if ((rc = slot_acquire(m, x, R, &q))) return rc;
  // NULL: run on the slot's own stream, after the caller's work queued on
  // the null stream (not the whole device: other solves keep running)
  if (!stream) {
    hipError_t e = hipStreamSynchronize(0);
    if (e != hipSuccess) { slot_release(x, q); return csem_fail(-1, "sync: %s", hipGetErrorString(e)); }
  }
  const hipStream_t s = stream ? (hipStream_t)stream : q->own;
  const csem_mixed_opts dflt = {2000, 1e-2, 40, 8, 1, CSEM_PRECOND_LUSGS, 0, CSEM_KRYLOV_FBICGSTAB, 0};
  rc = mixed_solve_slot(m, x, q, R, f, (z2 *)xv, (const z2 *)bv, tol, opt ? *opt : dflt, iters, relres, s);
  slot_release(x, q);
  return rc;
}

// The FP32 operator / preconditioner alone (tests, custom drivers).
static int mx_single(csem_model *m, int R, const csem_freq *f, const void *xv, void *yv, void *stream, bool lu)
{
  MixState *x;
  MixSlot *q;
  int rc;
  {
    std::lock_guard<std::mutex> lk(g_mx_init_mu);
    if ((rc = mx_model(m, &x))) return rc;
  }
  if ((rc = slot_acquire(m, x, R, &q))) return rc;
  const hipStream_t s = (hipStream_t)stream;
  const MxFreq fr = mx_freq(f, R);
  q->pc = CSEM_PRECOND_LUSGS;
  if (lu) {
    rc = mx_lusgs(m, x, q, R, fr, (const c2 *)xv, (c2 *)yv, s);
  } else {
    rc = mx_apply(m, x, q, R, fr, (const c2 *)xv, (c2 *)yv, s);
  }
  if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = csem_fail(-1, "sync");
  slot_release(x, q);
  return rc;
}

extern "C" int csem_apply32(csem_model *m, int R, const csem_freq *f, const void *xv, void *yv, void *stream)
{
  return mx_single(m, R, f, xv, yv, stream, false);
}

extern "C" int csem_lusgs32(csem_model *m, int R, const csem_freq *f, const void *xv, void *yv, void *stream)
{
  return mx_single(m, R, f, xv, yv, stream, true);
}

extern "C" int csem_mixed_tuning(csem_model *m, int R, int *rows, int *planes)
{
  MixState *x;
  {
    std::lock_guard<std::mutex> lk(g_mx_init_mu);
    x = m->mx;
  }
  if (!x || R < 1 || R > MX_RMAX) return -1;
  std::lock_guard<std::mutex> lk(x->mu);
  if (!x->ap_tj[R]) return -1;
  *rows = x->ap_tj[R];
  *planes = x->ap_kc[R];
  return 0;
}

extern "C" int csem_mixed_set_tile(csem_model *m, int R, int rows, int planes)
{
  MixState *x;
  int rc;
  if (R < 1 || R > MX_RMAX) return csem_fail(-3, "batch of %d right-hand sides outside [1, %d]", R, MX_RMAX);
  int ok_rows[2];
  mx_tile_rows(R, ok_rows);
  if (!(rows == ok_rows[0] || rows == ok_rows[1]) || !(planes == 8 || planes == 10 || planes == 12 || planes == 16))
    return csem_fail(-3, "tile %d x %d not available for %d right-hand sides", rows, planes, R);
  {
    std::lock_guard<std::mutex> lk(g_mx_init_mu);
    if ((rc = mx_model(m, &x))) return rc;
  }
  std::lock_guard<std::mutex> lk(x->mu);
  x->ap_kc[R] = planes;
  x->ap_tj[R] = rows;
  return 0;
}

extern "C" int csem_mixed_tuning_lu(csem_model *m, int R, int *pre, int *cols)
{
  MixState *x;
  {
    std::lock_guard<std::mutex> lk(g_mx_init_mu);
    x = m->mx;
  }
  if (!x || R < 1 || R > MX_RMAX) return -1;
  std::lock_guard<std::mutex> lk(x->mu);
  if (x->pre_v[R][0] < 0 || !x->air_cw[R]) return -1;
  for (int k = 0; k < 6; k++) pre[k] = x->pre_v[R][k];
  *cols = x->air_cw[R];
  return 0;
}

extern "C" int csem_mixed_set_lu_tiles(csem_model *m, int R, const int *pre, int cols)
{
  MixState *x;
  int rc;
  if (R < 1 || R > MX_RMAX) return csem_fail(-3, "batch of %d right-hand sides outside [1, %d]", R, MX_RMAX);
  for (int k = 0; k < 6; k++)
    if (pre[k] < 0 || pre[k] >= MX_PRE_NV) return csem_fail(-3, "pre-pass tile candidate %d outside [0, %d)", pre[k], MX_PRE_NV);
  {
    std::lock_guard<std::mutex> lk(g_mx_init_mu);
    if ((rc = mx_model(m, &x))) return rc;
  }
  if (m->airbc) {
    const int N = m->nfft;
    if (cols < std::max(1, 1024 / N) || cols > x->cw_max || (cols & (cols - 1)))
      return csem_fail(-3, "FFT column pass: %d columns per block not available for a %d-point transform", cols, N);
  }
  std::lock_guard<std::mutex> lk(x->mu);
  for (int k = 0; k < 6; k++) x->pre_v[R][k] = pre[k];
  x->air_cw[R] = m->airbc ? cols : x->cw;
  return 0;
}

// device bytes slot_new allocates for R right-hand sides (admission control)
extern "C" long csem_mixed_slot_bytes(const csem_model *m, int R)
{
  const long n = m->lE, c2b = sizeof(c2), zb = sizeof(z2);
  const long nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  long b = (m->fE * R + 3L * mx_pad(m).n * R + 10L * R * n) * c2b + (long)R * nch * 4 * 8 +
           (long)R * MS_N * 8 + (2 * R * n + n) * zb + (long)R * nch * 8 + R * 8 + R * 4;
  if (m->airbc) {
    const long NN = (long)m->nfft * m->nfft, p1 = (long)m->n1 * m->n2, p2 = (long)(m->n1 + 1) * m->n2,
               p3 = (long)m->n1 * (m->n2 + 1);
    b += (3 * NN + p1 + p2 + p3) * R * c2b + (3 * NN + p1 + p2 + p3) * zb;
  }
  return b;
}

extern "C" int csem_mem_info(long *free_bytes, long *total_bytes)
{
  size_t f = 0, t = 0;
  HIP_OK(hipMemGetInfo(&f, &t));
  *free_bytes = (long)f;
  *total_bytes = (long)t;
  return 0;
}

extern "C" long csem_mixed_bytes(const csem_model *m)
{
  MixState *x;
  {
    std::lock_guard<std::mutex> lk(g_mx_init_mu);
    x = m->mx;
  }
  if (!x) return 0;
  std::lock_guard<std::mutex> lk(x->mu);  // slots may be added by a running solve
  long b = x->bytes;
  for (const MixSlot *q : x->slots) b += q->bytes;
  return b;
}
