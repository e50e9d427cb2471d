// csem_dev.hip -- device model, launchers and the device-resident FP64
// BiCGStab of the MI355X solve path (C-ABI in include/csem_dev.h).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include <algorithm>
#include <mutex>

#include "csem_internal.h"

std::mutex g_devwide_mu;

static char g_err[1024];

extern "C" int csem_fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char *csem_last_error(void) { return g_err; }

#define PI_REF 3.14159265358979323846264

// ------------------------------------------------------- kernel timers
// Optional HIP-event timing of every launch of the hot kernels, on the
// stream they run on, with the algorithmic bytes each launch moves.
static const char *KT_NAMES[KT_N] = {
  "k_apply", "k_hz0", "k_air_rows_fwd", "k_air_cols", "k_air_rows_inv", "k_air_ghost",
  "k_sgs_ex<fwd>", "k_sgs_ey<fwd>", "k_sgs_ez<fwd>", "k_sgs_ez<bwd>", "k_sgs_ey<bwd>", "k_sgs_ex<bwd>",
  "k_dot_partial", "k_dot_final", "k_bicg_s", "k_bicg_xr", "k_bicg_p", "k_lu_pre", "k_lu_smith",
  "k_mx_apply", "k_mx_hz0", "k_mx_air_rows", "k_mx_air_cols", "k_mx_air_rows_inv", "k_mx_air_ghost",
  "k_mx_lu_pre", "k_mx_sgs", "k_mx_krylov", "k_mx_dot", "k_mx_refine", "k_mx_unskew", "k_gm_orth"};

struct KtPend {
  int id;
  hipEvent_t a, b;
  double bytes;
};
static struct {
  int on;
  std::vector<hipEvent_t> pool;
  std::vector<KtPend> pend, loops;  // loops: bytes holds the launch count
  double ms[KT_N], bytes[KT_N], loop_ms[KT_N];
  long cnt[KT_N], loop_cnt[KT_N];
  int rep_n[KT_N], rep_left[KT_N];  // csem_kernel_repeat
} g_kt;

static hipEvent_t kt_event()
{
  hipEvent_t e;
  if (!g_kt.pool.empty()) {
    e = g_kt.pool.back();
    g_kt.pool.pop_back();
    return e;
  }
  if (hipEventCreate(&e) != hipSuccess) return NULL;
  return e;
}

static std::mutex g_kt_mu;  // solves on several host threads record concurrently
static thread_local int g_kt_paused;
thread_local int g_kt_live;
// timing is paused while this thread captures a stream into a graph
void kt_pause(int paused) { g_kt_paused = paused; }

bool kt_events(hipEvent_t *a, hipEvent_t *b)
{
  if (!g_kt.on || g_kt_paused) return false;
  std::lock_guard<std::mutex> lk(g_kt_mu);
  *a = kt_event();
  if (!*a) return false;
  *b = kt_event();
  if (!*b) {
    g_kt.pool.push_back(*a);
    return false;
  }
  return true;
}

void kt_record(int id, hipEvent_t a, hipEvent_t b, double bytes)
{
  std::lock_guard<std::mutex> lk(g_kt_mu);
  g_kt.pend.push_back({id, a, b, bytes});
}

// kernels whose output does not feed their own input: repeating a launch
// leaves every result unchanged (the inverse row transforms work in place
// and the Krylov kernels update their vectors, so they are not here)
static bool kt_repeatable(int id)
{
  switch (id) {
    case KT_APPLY: case KT_HZ0: case KT_AIR_ROWS_FWD: case KT_AIR_COLS: case KT_AIR_GHOST:
    case KT_LU_PRE: case KT_LU_SMITH: case KT_DOT:
    case KT_SGS_EX_F: case KT_SGS_EY_F: case KT_SGS_EZ_F: case KT_SGS_EZ_B: case KT_SGS_EY_B: case KT_SGS_EX_B:
    case KT_M_APPLY: case KT_M_HZ0: case KT_M_AIR_ROWS: case KT_M_AIR_COLS: case KT_M_AIR_GHOST:
    case KT_M_PRE: case KT_M_SGS: case KT_M_UNSKEW:
      return true;
    default:
      return false;
  }
}

int kt_repeat(int id, hipEvent_t *a, hipEvent_t *b)
{
  if (!g_kt.on || g_kt_paused) return 0;
  std::lock_guard<std::mutex> lk(g_kt_mu);
  if (g_kt.rep_left[id] <= 0 || g_kt.rep_n[id] <= 0) return 0;
  *a = kt_event();
  if (!*a) return 0;
  *b = kt_event();
  if (!*b) {
    g_kt.pool.push_back(*a);
    return 0;
  }
  g_kt.rep_left[id]--;
  return g_kt.rep_n[id];
}

void kt_record_loop(int id, hipEvent_t a, hipEvent_t b, int n)
{
  std::lock_guard<std::mutex> lk(g_kt_mu);
  g_kt.loops.push_back({id, a, b, (double)n});
}

static void kt_collect()
{
  std::lock_guard<std::mutex> lk(g_kt_mu);
  for (auto &p : g_kt.pend) {
    float ms = 0;
    (void)hipEventSynchronize(p.b);
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    g_kt.ms[p.id] += ms;
    g_kt.bytes[p.id] += p.bytes;
    g_kt.cnt[p.id]++;
    g_kt.pool.push_back(p.a);
    g_kt.pool.push_back(p.b);
  }
  g_kt.pend.clear();
  for (auto &p : g_kt.loops) {
    float ms = 0;
    (void)hipEventSynchronize(p.b);
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    g_kt.loop_ms[p.id] += ms;
    g_kt.loop_cnt[p.id] += (long)p.bytes;
    g_kt.pool.push_back(p.a);
    g_kt.pool.push_back(p.b);
  }
  g_kt.loops.clear();
}

extern "C" int csem_kernel_timing(int on)
{
  kt_collect();
  std::lock_guard<std::mutex> lk(g_kt_mu);
  g_kt.on = on;
  for (int i = 0; i < KT_N; i++) {
    g_kt.ms[i] = 0; g_kt.bytes[i] = 0; g_kt.cnt[i] = 0;
    g_kt.loop_ms[i] = 0; g_kt.loop_cnt[i] = 0;
    g_kt.rep_n[i] = g_kt.rep_left[i] = 0;
  }
  return 0;
}

extern "C" int csem_kernel_repeat(int id, int reps, int loops)
{
  if (id < 0 || id >= KT_N || !kt_repeatable(id) || reps < 0 || loops < 0)
    return csem_fail(-1, "csem_kernel_repeat: kernel %d cannot repeat (its output feeds its input) or bad count", id);
  std::lock_guard<std::mutex> lk(g_kt_mu);
  g_kt.rep_n[id] = reps;
  g_kt.rep_left[id] = loops;
  return 0;
}

extern "C" int csem_kernel_loop_stats(int id, double *total_ms, long *count)
{
  if (id < 0 || id >= KT_N) return -1;
  kt_collect();
  *total_ms = g_kt.loop_ms[id];
  *count = g_kt.loop_cnt[id];
  return 0;
}

extern "C" int csem_kernel_stats(int id, const char **name, double *total_ms, long *count, double *bytes)
{
  if (id < 0 || id >= KT_N) return -1;
  kt_collect();
  *name = KT_NAMES[id];
  *total_ms = g_kt.ms[id];
  *count = g_kt.cnt[id];
  *bytes = g_kt.bytes[id];
  return KT_N;
}

extern "C" int csem_freq_consts(float freq, csem_freq *out)
{
  // I*2.*PI*f*mu0 (main.c:95) has a zero real part and imaginary part
  // ((2*PI)*f)*mu0; 1./that goes through complex division (fdfd.c:100).
  double w = ((2. * PI_REF) * (double)freq) * (4. * PI_REF * 1e-7);
  z2 iomu = zmk(0.0, w);
  z2 iwmu = zdiv(zmk(1.0, 0.0), iomu);
  out->iomu_re = iomu.x; out->iomu_im = iomu.y;
  out->iwmu_re = iwmu.x; out->iwmu_im = iwmu.y;
  return 0;
}

static bool g_attr_done = false;
static int set_lds_attrs()
{
  if (g_attr_done) return 0;
  HIP_OK(hipFuncSetAttribute((const void *)k_air_cols<2>, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
  HIP_OK(hipFuncSetAttribute((const void *)k_fft_cols<2>, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
  HIP_OK(hipFuncSetAttribute((const void *)k_air_rows_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
  HIP_OK(hipFuncSetAttribute((const void *)k_air_rows_inv, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
  HIP_OK(hipFuncSetAttribute((const void *)k_fft_rows, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
  g_attr_done = true;
  return 0;
}

// Planes per operator tile.  The tiles of one launch run in rounds of
// (CUs x resident blocks per CU); a tile costs its planes plus about two
// planes of prologue (the plane below it and the LDS set-up), so pick the
// depth with the fewest planes-times-rounds.  The depth changes no value.
template <int KC>
static int apply_cost(const csem_model *m, int ncu, long *cost)
{
  int bpc = 0;
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_apply<KC>, 64 * AP_NW, 0));
  const long tiles = (long)((m->n1 + 1 + AP_OI - 1) / AP_OI) * ((m->n2 + 1 + AP_OJ - 1) / AP_OJ) *
                     ((m->n3 + 1 + KC - 1) / KC);
  const long slots = (long)ncu * (bpc > 0 ? bpc : 1);
  *cost = ((tiles + slots - 1) / slots) * (KC + 2);
  return 0;
}
static int pick_apply_planes(csem_model *m)
{
  int dev = 0, ncu = 0;
  HIP_OK(hipGetDevice(&dev));
  HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int kcs[5] = {6, 8, 10, 12, 16};
  long cost[5];
  int rc;
  if ((rc = apply_cost<6>(m, ncu, &cost[0])) || (rc = apply_cost<8>(m, ncu, &cost[1])) ||
      (rc = apply_cost<10>(m, ncu, &cost[2])) || (rc = apply_cost<12>(m, ncu, &cost[3])) ||
      (rc = apply_cost<16>(m, ncu, &cost[4])))
    return rc;
  int best = 1;
  for (int c = 0; c < 5; c++)
    if (cost[c] < cost[best]) best = c;
  m->kc = kcs[best];
  return 0;
}

static ApplyArgs apply_args(const csem_model *m, const csem_freq *f, const z2 *x, z2 *y);

// Every plane depth gives the same values (a tile computes each cell in the
// same order whatever its depth), so the depth is picked by speed: each
// candidate's k_apply is timed on this model (one warm and three timed
// launches on zeroed vectors of the model's length) once, at creation,
// before any other thread can reach the model.  The cost model's pick
// above stays when the two scratch vectors do not fit.  The launches go on
// the null stream under the device-wide lock, which graph captures on
// other threads hold too.
template <int KC>
static float time_apply_kc(csem_model *m, const z2 *x, z2 *y, hipEvent_t e0, hipEvent_t e1)
{
  const csem_freq f = {0.0, 1.0, 0.0, -1.0};
  m->kc = KC;
  const ApplyArgs a = apply_args(m, &f, x, y);
  const int ntiles = a.nti * a.ntj * a.ntk;
  const dim3 blk(64, AP_NW), grd(8 * ((ntiles + 7) / 8));
  k_apply<KC><<<grd, blk>>>(a);
  (void)hipEventRecord(e0, 0);
  for (int w = 0; w < 3; w++) k_apply<KC><<<grd, blk>>>(a);
  (void)hipEventRecord(e1, 0);
  float ms = 1e30f;
  if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) ms = 1e30f;
  return ms;
}
static int time_apply_planes(csem_model *m)
{
  const long bytes = m->lE * (long)sizeof(z2);
  z2 *x = nullptr, *y = nullptr;
  if (dev_malloc((void **)&x, bytes) != hipSuccess || dev_malloc((void **)&y, bytes) != hipSuccess) {
    (void)hipGetLastError();
    if (x) dev_free(x);
    return 0;
  }
  const int pick = m->kc;
  int kc = pick;
  hipError_t e = hipSuccess;
  {
    std::lock_guard<std::mutex> lk(g_devwide_mu);
    hipEvent_t e0, e1;
    e = hipMemset(x, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) {
      if ((e = hipEventCreate(&e1)) == hipSuccess) {
        const int kcs[5] = {6, 8, 10, 12, 16};
        const float t[5] = {time_apply_kc<6>(m, x, y, e0, e1), time_apply_kc<8>(m, x, y, e0, e1),
                            time_apply_kc<10>(m, x, y, e0, e1), time_apply_kc<12>(m, x, y, e0, e1),
                            time_apply_kc<16>(m, x, y, e0, e1)};
        float best = 1e30f;
        for (int c = 0; c < 5; c++)
          if (t[c] < best) { best = t[c]; kc = kcs[c]; }
        (void)hipEventDestroy(e1);
      }
      (void)hipEventDestroy(e0);
    }
    if (e == hipSuccess) e = hipGetLastError();
  }
  m->kc = e == hipSuccess ? kc : pick;
  dev_free(x);
  dev_free(y);
  if (e != hipSuccess) return csem_fail(-1, "timing the operator tile depths: %s", hipGetErrorString(e));
  return 0;
}

extern "C" void csem_model_destroy(csem_model *m)
{
  if (!m) return;
  mixed_free(m);
  for (void *p : m->allocs) dev_free(p);
  delete m;
}

// ------------------------------------------------- device-side model setup
#define CSEM_PI 3.14159265358979323846264  // the host's pi (cstd.h:35-37)
// extend_model.c:108-175: the extended cells, conductivity for 11 and 22
// and resistivity for 33, from the core cells (float, x fastest): nb
// lateral padding columns and rows repeat the edge cells; the nb planes on
// top are air; bottom padding plane n3-1-k copies row k of the last core
// plane (quirk Q2: extend_model.c:169-173 index [k] where [j] is meant),
// a padding row, so the last plane's first core row.
struct CellArgs {
  int nx, ny, nz, nb, n1, n2, n3;
  const float *rho[3];
  double rho_air;
  double *c[3];
};
__global__ void k_setup_cells(CellArgs a)
{
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)a.n1 * a.n2 * a.n3) return;
  const int i = (int)(e % a.n1), j = (int)((e / a.n1) % a.n2), k = (int)(e / ((long)a.n1 * a.n2));
  const int nb = a.nb;
  const bool air = nb > 0 && k < nb;
  int kc = k - nb, jc = j - nb;
  if (nb > 0 && k >= a.n3 - nb) {
    kc = a.nz - 1;
    jc = (a.n3 - 1 - k) - nb;
  }
  kc = min(max(kc, 0), a.nz - 1);
  jc = min(max(jc, 0), a.ny - 1);
  const int ic = min(max(i - nb, 0), a.nx - 1);
  const long src = ((long)kc * a.ny + jc) * a.nx + ic;
  for (int q = 0; q < 3; q++) {
    const double r = air ? a.rho_air : (double)a.rho[q][src];
    a.c[q][e] = (q == 2) ? r : 1. / r;
  }
}

// extend_model.c:180-252: the edge averages, each edge from the four cells
// around it weighted by their half-widths, then the boundary edges copied
// from the nearest interior ones (the copies compose to the interior value
// at the clamped indices).  Same expressions as the host, no contraction.
struct SigArgs {
  int n1, n2, n3;
  const double *d1, *d2, *d3, *d1s, *d2s, *d3s;
  const double *c11, *c22, *r33;
  double *s11, *s22, *s33;
};
__global__ void k_setup_sigma(SigArgs a)
{
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  const long l11 = (long)n1 * (n2 + 1) * (n3 + 1), l22 = (long)(n1 + 1) * n2 * (n3 + 1),
             l33 = (long)(n1 + 1) * (n2 + 1) * n3;
#define CC(c, k, j, i) (c)[((long)(k) * n2 + (j)) * n1 + (i)]
  if (e < l11) {
    const int i = (int)(e % n1), j = (int)((e / n1) % (n2 + 1)), k = (int)(e / ((long)n1 * (n2 + 1)));
    const int kk = min(max(k, 1), n3 - 1), jj = min(max(j, 1), n2 - 1);
    a.s11[e] = 0.25 *
               ((a.d2s[jj] * CC(a.c11, kk, jj, i) + a.d2s[jj - 1] * CC(a.c11, kk, jj - 1, i)) * a.d3s[kk] +
                (a.d2s[jj] * CC(a.c11, kk - 1, jj, i) + a.d2s[jj - 1] * CC(a.c11, kk - 1, jj - 1, i)) * a.d3s[kk - 1]) /
               (a.d2[jj] * a.d3[kk]);
  } else if (e < l11 + l22) {
    const long f = e - l11;
    const int i = (int)(f % (n1 + 1)), j = (int)((f / (n1 + 1)) % n2), k = (int)(f / ((long)(n1 + 1) * n2));
    const int kk = min(max(k, 1), n3 - 1), ii = min(max(i, 1), n1 - 1);
    a.s22[f] = 0.25 *
               ((a.d1s[ii] * CC(a.c22, kk, j, ii) + a.d1s[ii - 1] * CC(a.c22, kk, j, ii - 1)) * a.d3s[kk] +
                (a.d1s[ii] * CC(a.c22, kk - 1, j, ii) + a.d1s[ii - 1] * CC(a.c22, kk - 1, j, ii - 1)) * a.d3s[kk - 1]) /
               (a.d1[ii] * a.d3[kk]);
  } else if (e < l11 + l22 + l33) {
    const long f = e - l11 - l22;
    const int i = (int)(f % (n1 + 1)), j = (int)((f / (n1 + 1)) % (n2 + 1)), k = (int)(f / ((long)(n1 + 1) * (n2 + 1)));
    const int jj = min(max(j, 1), n2 - 1), ii = min(max(i, 1), n1 - 1);
    const double t =
        0.25 *
        ((a.d1s[ii] * CC(a.r33, k, jj, ii) + a.d1s[ii - 1] * CC(a.r33, k, jj, ii - 1)) * a.d2s[jj] +
         (a.d1s[ii] * CC(a.r33, k, jj - 1, ii) + a.d1s[ii - 1] * CC(a.r33, k, jj - 1, ii - 1)) * a.d2s[jj - 1]) /
        (a.d1[ii] * a.d2[jj]);
    a.s33[f] = 1. / t;
  }
#undef CC
}

// airbc_init_free.c:69-84: the upward-continuation and derivative
// multipliers, in the C99 complex arithmetic the host expression performs:
// exp(-kz dz/2) * I * kx / (kz + 1e-15) scales componentwise, then the
// product with cexp(-I kx d1s[0]/2) = (cos, sin) of ((-kx) 0.5) d1s[0].
// exp, sin and cos are the device's (within a few ulp of the host's).
__global__ void k_setup_mult(int N, double dkx, double dky, double d3fft, double h1, double h2, z2 *sH1, z2 *sH2)
{
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)N * N) return;
  const int i1 = (int)(e % N), i2 = (int)(e / N);
  const double ky = (i2 <= N / 2) ? i2 * dky : (i2 - N) * dky;
  const double kx = (i1 <= N / 2) ? i1 * dkx : (i1 - N) * dkx;
  const double kz = sqrt(kx * kx + ky * ky);
  const double ez = exp(-kz * 0.5 * d3fft), den = kz + 1.e-15;
  for (int c = 0; c < 2; c++) {
    const double k = c ? ky : kx, h = c ? h2 : h1;
    const double ar = ((ez * 0.0) * k) / den, ai = ((ez * 1.0) * k) / den;
    const double th = ((-1.0 * k) * 0.5) * h;
    const double cr = cos(th), ci = sin(th);
    (c ? sH2 : sH1)[e] = zmk(ar * cr - ai * ci, ar * ci + ai * cr);
  }
}

static int model_create(const csem_model_desc *d, const csem_model_raw *raw, csem_model **out);

extern "C" int csem_model_create(const csem_model_desc *d, csem_model **out) { return model_create(d, NULL, out); }

extern "C" int csem_model_create_raw(const csem_model_desc *d, const csem_model_raw *raw, csem_model **out)
{
  *out = NULL;
  if (!raw || raw->nx + 2 * d->nb != d->n1 || raw->ny + 2 * d->nb != d->n2 || raw->nz + 2 * d->nb != d->n3f)
    return csem_fail(-3, "raw model %dx%dx%d does not match the extended grid %dx%dx%d (nb %d)", raw ? raw->nx : 0,
                     raw ? raw->ny : 0, raw ? raw->nz : 0, d->n1, d->n2, d->n3f, d->nb);
  // positive finite resistivities give finite positive conductivities
  const float *rh[3] = {raw->rho11, raw->rho22, raw->rho33};
  const long nc = (long)raw->nx * raw->ny * raw->nz;
  for (int c = 0; c < 3; c++)
    for (long e = 0; e < nc; e++)
      if (!(rh[c][e] > 0.f) || !std::isfinite(rh[c][e]))
        return csem_fail(-3, "resistivity must be finite and positive (component %d, cell %ld: %g)", c, e,
                         (double)rh[c][e]);
  if (d->nb > 0 && (!(raw->rho_air > 0.0) || !std::isfinite(raw->rho_air)))
    return csem_fail(-3, "air resistivity must be finite and positive (%g)", raw->rho_air);
  return model_create(d, raw, out);
}

// σ of a component copied to a host array, or the device-built ones from
// the extended cells
extern "C" int csem_model_sigma(const csem_model *m, int which, double *host)
{
  if (which < 0 || which > 2) return csem_fail(-3, "sigma component %d", which);
  const double *s = which == 0 ? m->s11 : which == 1 ? m->s22 : m->s33;
  const long n = which == 0 ? m->fEx : which == 1 ? m->fEy : m->fEz;
  std::lock_guard<std::mutex> lk(g_devwide_mu);
  HIP_OK(hipMemcpy(host, s, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int csem_model_multiplier(const csem_model *m, int which, double *host)
{
  if (!m->airbc || which < 0 || which > 1) return csem_fail(-3, "no multiplier %d", which);
  std::lock_guard<std::mutex> lk(g_devwide_mu);
  HIP_OK(hipMemcpy(host, which ? m->sH2 : m->sH1, (long)m->nfft * m->nfft * sizeof(z2), hipMemcpyDeviceToHost));
  return 0;
}

static int model_create(const csem_model_desc *d, const csem_model_raw *raw, csem_model **out)
{
  *out = NULL;
  if (d->n1 < 2 || d->n2 < 2 || d->n3f < 2) return csem_fail(-3, "grid too small: %d %d %d", d->n1, d->n2, d->n3f);
  if (d->airbc && d->nb < 1) return csem_fail(-3, "airbc=1 needs nb>0");
  if (d->n3f - 1 > SGS_MAXT || d->n2 - 1 > SGS_MAXT)
    return csem_fail(-3, "LU-SGS slices hold at most %d rows (n3f-1=%d, n2-1=%d)", SGS_MAXT, d->n3f - 1, d->n2 - 1);
  if (!raw) {
    // LU-SGS drops the reference's sigma*0 term, the identity for finite
    // non-negative conductivity
    const double *sg[3] = {d->sigma11, d->sigma22, d->sigma33};
    const long ln[3] = {(long)d->n1 * (d->n2 + 1) * (d->n3f + 1), (long)(d->n1 + 1) * d->n2 * (d->n3f + 1),
                        (long)(d->n1 + 1) * (d->n2 + 1) * d->n3f};
    for (int c = 0; c < 3; c++)
      for (long e = 0; e < ln[c]; e++)
        if (!(sg[c][e] >= 0.0) || !std::isfinite(sg[c][e]))
          return csem_fail(-3, "conductivity must be finite and non-negative (component %d, cell %ld: %g)", c, e,
                           sg[c][e]);
  }
  csem_model *m = new csem_model();
  int rc;
  m->n1 = d->n1; m->n2 = d->n2; m->n3f = d->n3f; m->nb = d->nb; m->airbc = d->airbc; m->fvm = d->fvm;
  m->kofs = d->airbc ? d->nb : 0;
  m->n3 = d->n3f - m->kofs;
  m->kc = 0;  // planes per operator tile: chosen below
  const int n1 = m->n1, n2 = m->n2, n3 = m->n3, n3f = m->n3f;
  m->lEx = (long)n1 * (n2 + 1) * (n3 + 1);
  m->lEy = (long)(n1 + 1) * n2 * (n3 + 1);
  m->lEz = (long)(n1 + 1) * (n2 + 1) * n3;
  m->lE = m->lEx + m->lEy + m->lEz;
  m->fEx = (long)n1 * (n2 + 1) * (n3f + 1);
  m->fEy = (long)(n1 + 1) * n2 * (n3f + 1);
  m->fEz = (long)(n1 + 1) * (n2 + 1) * n3f;
  m->fE = m->fEx + m->fEy + m->fEz;
#define UP(dst, src, cnt) if ((rc = upload(m, &m->dst, src, cnt))) goto fail
  UP(d1, d->d1, n1 + 1); UP(d2, d->d2, n2 + 1); UP(d3, d->d3, n3f + 1);
  UP(d1s, d->d1s, n1 + 1); UP(d2s, d->d2s, n2 + 1); UP(d3s, d->d3s, n3f + 1);
  {
    // correctly rounded reciprocals of every spacing, for ddiv
    std::vector<double> t;
    auto recip = [&](const double *v, int n) { t.assign(v, v + n); for (double &e : t) e = 1.0 / e; return t.data(); };
    UP(r1, recip(d->d1, n1 + 1), n1 + 1); UP(r2, recip(d->d2, n2 + 1), n2 + 1); UP(r3, recip(d->d3, n3f + 1), n3f + 1);
    UP(r1s, recip(d->d1s, n1 + 1), n1 + 1); UP(r2s, recip(d->d2s, n2 + 1), n2 + 1);
    UP(r3s, recip(d->d3s, n3f + 1), n3f + 1);
  }
  if (!raw) {
    UP(s11, d->sigma11, m->fEx); UP(s22, d->sigma22, m->fEy); UP(s33, d->sigma33, m->fEz);
  } else {
    // the cells and the edge averages on the device
    if ((rc = dalloc(m, &m->s11, m->fEx)) || (rc = dalloc(m, &m->s22, m->fEy)) || (rc = dalloc(m, &m->s33, m->fEz)))
      goto fail;
    const long nc = (long)raw->nx * raw->ny * raw->nz, ne = (long)n1 * n2 * n3f;
    float *rho = NULL;
    double *cel = NULL;
    hipError_t e = dev_malloc((void **)&rho, 3 * nc * (long)sizeof(float));
    if (e == hipSuccess) e = dev_malloc((void **)&cel, 3 * ne * (long)sizeof(double));
    if (e == hipSuccess) e = dev_upload(rho, raw->rho11, nc * sizeof(float));
    if (e == hipSuccess) e = dev_upload(rho + nc, raw->rho22, nc * sizeof(float));
    if (e == hipSuccess) e = dev_upload(rho + 2 * nc, raw->rho33, nc * sizeof(float));
    if (e == hipSuccess) {
      CellArgs ca;
      ca.nx = raw->nx; ca.ny = raw->ny; ca.nz = raw->nz; ca.nb = d->nb; ca.n1 = n1; ca.n2 = n2; ca.n3 = n3f;
      ca.rho[0] = rho; ca.rho[1] = rho + nc; ca.rho[2] = rho + 2 * nc;
      ca.rho_air = raw->rho_air;
      ca.c[0] = cel; ca.c[1] = cel + ne; ca.c[2] = cel + 2 * ne;
      SigArgs sa;
      sa.n1 = n1; sa.n2 = n2; sa.n3 = n3f;
      sa.d1 = m->d1; sa.d2 = m->d2; sa.d3 = m->d3; sa.d1s = m->d1s; sa.d2s = m->d2s; sa.d3s = m->d3s;
      sa.c11 = cel; sa.c22 = cel + ne; sa.r33 = cel + 2 * ne;
      sa.s11 = m->s11; sa.s22 = m->s22; sa.s33 = m->s33;
      // null-stream launches and a device-wide wait: under the lock that
      // graph captures on other threads hold
      std::lock_guard<std::mutex> lk(g_devwide_mu);
      k_setup_cells<<<(unsigned)((ne + 255) / 256), 256>>>(ca);
      k_setup_sigma<<<(unsigned)((m->fE + 255) / 256), 256>>>(sa);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    if (rho) dev_free(rho);
    if (cel) dev_free(cel);
    if (e != hipSuccess) { rc = csem_fail(-1, "device setup of the conductivities: %s", hipGetErrorString(e)); goto fail; }
  }
  {
    SkewGeo &g = m->g;
    g.A[0] = n3f - 1; g.B[0] = n2 - 1; g.ns[0] = n1;
    g.A[1] = n3f - 1; g.B[1] = n1 - 1; g.ns[1] = n2;
    g.A[2] = n2 - 1;  g.B[2] = n1 - 1; g.ns[2] = n3f;
    long off = 0, mx = 0, ncell = 0;
    for (int c = 0; c < 3; c++) {
      g.AP[c] = 64 * ((g.A[c] + 63) / 64);
      g.SL[c] = (long)(g.A[c] + g.B[c] - 1) * g.AP[c];
      const long nc = (long)g.ns[c] * g.SL[c];
      g.base[c] = off;
      off += nc;
      mx = std::max(mx, nc);
      ncell += (long)g.ns[c] * g.A[c] * g.B[c];
    }
    m->nsk = ncell;
    if ((rc = dalloc(m, &m->smith, off))) goto fail;
    if ((rc = dalloc(m, &m->rec, mx))) goto fail;
  }
  // cells outside every sweep range stay zero for the pre-passes
  if ((rc = dalloc(m, &m->W, m->fE))) goto fail;
  if (dev_memset(m->W, 0, m->fE * sizeof(z2)) != hipSuccess) { rc = csem_fail(-1, "memset W"); goto fail; }
  m->smith_ok = 0;
  m->nch = (m->fE + RED_CHUNK - 1) / RED_CHUNK;
  if ((rc = dalloc(m, &m->partials, m->nch * 4))) goto fail;
  if ((rc = dalloc(m, &m->scal, 64))) goto fail;
  if (dev_memset(m->scal, 0, 64 * sizeof(double)) != hipSuccess) { rc = csem_fail(-1, "memset scal"); goto fail; }
  if (m->airbc) {
    const int N = d->nfft;
    if (N < 4 || (N & (N - 1)) || N > 2048) { rc = csem_fail(-3, "nfft=%d must be a power of two in [4, 2048]", N); goto fail; }
    m->nfft = N;
    long NN = (long)N * N;
    if (!raw) {
      UP(sH1, d->sH1, NN); UP(sH2, d->sH2, NN);
    } else {
      if ((rc = dalloc(m, &m->sH1, NN)) || (rc = dalloc(m, &m->sH2, NN))) goto fail;
      const double dkx = 2.0 * CSEM_PI / (raw->d1fft * N), dky = 2.0 * CSEM_PI / (raw->d2fft * N);
      hipError_t e;
      {
        std::lock_guard<std::mutex> lk(g_devwide_mu);
        k_setup_mult<<<(unsigned)((NN + 255) / 256), 256>>>(N, dkx, dky, raw->d3fft, d->d1s[0], d->d2s[0], m->sH1,
                                                            m->sH2);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
      }
      if (e != hipSuccess) { rc = csem_fail(-1, "device setup of the multipliers: %s", hipGetErrorString(e)); goto fail; }
    }
    UP(tw, d->twiddle, N);
    UP(jx, d->jx, N); UP(jy, d->jy, N); UP(wx, d->wx, N); UP(wy, d->wy, N);
    UP(gxj1, d->gxj1, n1); UP(gxj2, d->gxj2, n2); UP(gxw1, d->gxw1, n1); UP(gxw2, d->gxw2, n2);
    UP(gyj1, d->gyj1, n1); UP(gyj2, d->gyj2, n2); UP(gyw1, d->gyw1, n1); UP(gyw2, d->gyw2, n2);
    // rows of the inverse transforms that the back-interpolation reads
    std::vector<int> rx, ry;
    for (int i2 = 0; i2 < n2; i2++) { rx.push_back(d->gxj2[i2]); rx.push_back(d->gxj2[i2] + 1); }
    for (int i2 = 0; i2 < n2; i2++) { ry.push_back(d->gyj2[i2]); ry.push_back(d->gyj2[i2] + 1); }
    std::sort(rx.begin(), rx.end()); rx.erase(std::unique(rx.begin(), rx.end()), rx.end());
    std::sort(ry.begin(), ry.end()); ry.erase(std::unique(ry.begin(), ry.end()), ry.end());
    for (int r : rx) if (r < 0 || r >= N) { rc = csem_fail(-3, "ghost row %d outside transform", r); goto fail; }
    for (int r : ry) if (r < 0 || r >= N) { rc = csem_fail(-3, "ghost row %d outside transform", r); goto fail; }
    for (int i1 = 0; i1 < n1; i1++)
      if (d->gxj1[i1] < 0 || d->gxj1[i1] + 1 >= N || d->gyj1[i1] < 0 || d->gyj1[i1] + 1 >= N) {
        rc = csem_fail(-3, "ghost column outside transform"); goto fail;
      }
    for (int q = 0; q < N; q++)
      if ((d->jx[q] >= 0 && d->jx[q] + 1 >= n1) || (d->jy[q] >= 0 && d->jy[q] + 1 >= n2)) {
        rc = csem_fail(-3, "forward interpolation index outside Hz plane"); goto fail;
      }
    m->nrowx = (int)rx.size(); m->nrowy = (int)ry.size();
    std::vector<int> rows(rx);
    rows.insert(rows.end(), ry.begin(), ry.end());
    UP(rows, rows.data(), (long)rows.size());
    if ((rc = dalloc(m, &m->buf, NN))) goto fail;
    if ((rc = dalloc(m, &m->bufx, NN))) goto fail;
    if ((rc = dalloc(m, &m->bufy, NN))) goto fail;
    if ((rc = dalloc(m, &m->hz0, (long)n1 * n2))) goto fail;
    if ((rc = dalloc(m, &m->hxg, (long)(n1 + 1) * n2))) goto fail;
    if ((rc = dalloc(m, &m->hyg, (long)n1 * (n2 + 1)))) goto fail;
  }
#undef UP
  if ((rc = set_lds_attrs())) goto fail;
  if ((rc = pick_apply_planes(m))) goto fail;
  if ((rc = time_apply_planes(m))) goto fail;
  *out = m;
  return 0;
fail:
  csem_model_destroy(m);
  return rc;
}

extern "C" long csem_model_len(const csem_model *m) { return m->lE; }
extern "C" long csem_model_full_len(const csem_model *m) { return m->fE; }
extern "C" long csem_model_bytes(const csem_model *m) { return m->bytes; }
extern "C" int csem_model_apply_planes(const csem_model *m) { return m->kc; }
extern "C" int csem_model_set_apply_planes(csem_model *m, int planes)
{
  if (planes != 6 && planes != 8 && planes != 10 && planes != 12 && planes != 16)
    return csem_fail(-3, "operator planes per tile %d: 6, 8, 10, 12 or 16", planes);
  m->kc = planes;
  return 0;
}

static inline hipStream_t S(void *s) { return (hipStream_t)s; }

static ApplyArgs apply_args(const csem_model *m, const csem_freq *f, const z2 *x, z2 *y)
{
  ApplyArgs a;
  const int n1 = m->n1, n2 = m->n2, ks = m->kofs;
  a.x = x; a.y = y;
  a.s11 = m->s11 + (long)ks * n1 * (n2 + 1);
  a.s22 = m->s22 + (long)ks * (n1 + 1) * n2;
  a.s33 = m->s33 + (long)ks * (n1 + 1) * (n2 + 1);
  a.d1 = m->d1; a.d2 = m->d2; a.d3 = m->d3 + ks;
  a.d1s = m->d1s; a.d2s = m->d2s; a.d3s = m->d3s + ks;
  a.r1 = m->r1; a.r2 = m->r2; a.r3 = m->r3 + ks;
  a.r1s = m->r1s; a.r2s = m->r2s; a.r3s = m->r3s + ks;
  a.hxg = m->hxg; a.hyg = m->hyg;
  a.iwmu = zmk(f->iwmu_re, f->iwmu_im);
  a.iomu = zmk(f->iomu_re, f->iomu_im);
  a.n1 = n1; a.n2 = n2; a.n3 = m->n3; a.fvm = m->fvm; a.airbc = m->airbc;
  a.nti = (n1 + 1 + AP_OI - 1) / AP_OI; a.ntj = (n2 + 1 + AP_OJ - 1) / AP_OJ; a.ntk = (m->n3 + 1 + m->kc - 1) / m->kc;
  a.lEx = m->lEx; a.lEy = m->lEy;
  return a;
}

static AirArgs air_args(const csem_model *m, const AirBufs &b)
{
  AirArgs a;
  a.N = m->nfft; a.n1 = m->n1; a.n2 = m->n2;
  a.hz = b.hz0; a.buf = b.buf; a.bufx = b.bufx; a.bufy = b.bufy;
  a.sH1 = m->sH1; a.sH2 = m->sH2; a.tw = m->tw;
  a.jx = m->jx; a.jy = m->jy; a.wx = m->wx; a.wy = m->wy;
  a.rows = m->rows; a.nrowx = m->nrowx; a.nrowy = m->nrowy;
  a.gxj1 = m->gxj1; a.gxj2 = m->gxj2; a.gyj1 = m->gyj1; a.gyj2 = m->gyj2;
  a.gxw1 = m->gxw1; a.gxw2 = m->gxw2; a.gyw1 = m->gyw1; a.gyw2 = m->gyw2;
  a.hxg = b.hxg; a.hyg = b.hyg;
  return a;
}

// the model's own scratch (the public entry points use it)
static AirBufs model_bufs(const csem_model *m)
{
  return AirBufs{m->hz0, m->buf, m->bufx, m->bufy, m->hxg, m->hyg};
}

static int run_airbc(csem_model *m, hipStream_t s, const AirBufs &b)
{
  AirArgs a = air_args(m, b);
  const int N = m->nfft;
  const size_t row_lds = 2 * (size_t)N * sizeof(z2);
  const double NN = (double)N * N;
  kt_launch(KT_AIR_ROWS_FWD, s, 16 * NN, k_air_rows_fwd, N, 256, row_lds, a);
  kt_launch(KT_AIR_COLS, s, 80 * NN, k_air_cols<2>, 8 * ((N / 2 + 7) / 8), 256, 3 * 2 * (size_t)N * sizeof(z2), a);
  kt_launch(KT_AIR_ROWS_INV, s, 32.0 * N * (m->nrowx + m->nrowy), k_air_rows_inv, m->nrowx + m->nrowy, 256,
            row_lds, a);
  const int ng = (m->n1 + 1) * m->n2 + m->n1 * (m->n2 + 1);
  // bytes per ghost value: the four transform points it interpolates and the store
  kt_launch(KT_AIR_GHOST, s, 80.0 * ng, k_air_ghost, (ng + 255) / 256, 256, 0, a);
  HIP_OK(hipGetLastError());
  return 0;
}

// y = A x with the airwave scratch b: concurrent callers on one model
// pass scratch of their own.
int run_apply_bufs(csem_model *m, const csem_freq *f, const z2 *x, z2 *y, hipStream_t s, const AirBufs &b)
{
  ApplyArgs a = apply_args(m, f, x, y);
  a.hxg = b.hxg; a.hyg = b.hyg;
  if (m->airbc) {
    const int nh = m->n1 * m->n2;
    kt_launch(KT_HZ0, s, 48.0 * nh, k_hz0, (nh + 255) / 256, 256, 0, a, b.hz0);  // Ex, Ey planes, Hz
    int rc = run_airbc(m, s, b);
    if (rc) return rc;
  }
  const int ntiles = a.nti * a.ntj * a.ntk;
  const dim3 blk(64, AP_NW), grd(8 * ((ntiles + 7) / 8));
  switch (m->kc) {
    case 6: kt_launch(KT_APPLY, s, 40.0 * m->lE, k_apply<6>, grd, blk, 0, a); break;
    case 10: kt_launch(KT_APPLY, s, 40.0 * m->lE, k_apply<10>, grd, blk, 0, a); break;
    case 12: kt_launch(KT_APPLY, s, 40.0 * m->lE, k_apply<12>, grd, blk, 0, a); break;
    case 16: kt_launch(KT_APPLY, s, 40.0 * m->lE, k_apply<16>, grd, blk, 0, a); break;
    default: kt_launch(KT_APPLY, s, 40.0 * m->lE, k_apply<8>, grd, blk, 0, a); break;
  }
  HIP_OK(hipGetLastError());
  return 0;
}

static int run_apply(csem_model *m, const csem_freq *f, const z2 *x, z2 *y, hipStream_t s)
{
  return run_apply_bufs(m, f, x, y, s, model_bufs(m));
}

// the sweep instance for a slice of nw waves (1..4: at most SGS_MAXT rows)
template <int C, bool FWD>
static void (*sgs2_kernel(bool fvm, int nw))(LuArgs)
{
  switch (nw) {
    case 1: return fvm ? k_sgs2<C, FWD, true, 1> : k_sgs2<C, FWD, false, 1>;
    case 2: return fvm ? k_sgs2<C, FWD, true, 2> : k_sgs2<C, FWD, false, 2>;
    case 3: return fvm ? k_sgs2<C, FWD, true, 3> : k_sgs2<C, FWD, false, 3>;
    default: return fvm ? k_sgs2<C, FWD, true, 4> : k_sgs2<C, FWD, false, 4>;
  }
}

// Dynamic LDS above 64 KB needs the kernel's limit raised first: once per
// sweep instance, to the largest request seen.
static int sg2_lds_attr(const void *k, size_t bytes)
{
  static std::mutex mu;
  static std::vector<std::pair<const void *, size_t>> raised;
  if (bytes <= 64 * 1024) return 0;
  std::lock_guard<std::mutex> lk(mu);
  for (auto &p : raised)
    if (p.first == k && p.second >= bytes) return 0;
  if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
    return csem_fail(-1, "LU-SGS sweep: %zu bytes of LDS not available", bytes);
  for (auto &p : raised)
    if (p.first == k) { p.second = bytes; return 0; }
  raised.push_back({k, bytes});
  return 0;
}

static int run_lusgs(csem_model *m, const csem_freq *f, const z2 *x, z2 *y, hipStream_t s)
{
  LuArgs a;
  a.x = x; a.y = y; a.W = m->W; a.rec = m->rec; a.smith = m->smith;
  a.s11 = m->s11; a.s22 = m->s22; a.s33 = m->s33;
  a.d1 = m->d1; a.d2 = m->d2; a.d3 = m->d3; a.d1s = m->d1s; a.d2s = m->d2s; a.d3s = m->d3s;
  a.r1 = m->r1; a.r2 = m->r2; a.r3 = m->r3; a.r1s = m->r1s; a.r2s = m->r2s; a.r3s = m->r3s;
  a.iwmu = zmk(f->iwmu_re, f->iwmu_im);
  a.iomu = zmk(f->iomu_re, f->iomu_im);
  a.fvm = m->fvm; a.n1 = m->n1; a.n2 = m->n2; a.n3f = m->n3f; a.kofs = m->kofs; a.n3op = m->n3;
  a.fEx = m->fEx; a.fEy = m->fEy; a.lEx = m->lEx; a.lEy = m->lEy;
  a.g = m->g;
  const SkewGeo &g = m->g;
  dim3 gp[3], gq[3];
  for (int c = 0; c < 3; c++) gp[c] = dim3((unsigned)(((long)g.A[c] * g.B[c] + 255) / 256), (unsigned)g.ns[c]);
  // pre-pass tiles (LuPreTile): Ex 16 slices x 4 rows x 8 columns, Ey and
  // Ez 8 rows x 64 columns, 256 threads
  gq[0] = dim3((g.ns[0] + 15) / 16, (g.B[0] + 7) / 8, (g.A[0] + 3) / 4);
  gq[1] = dim3((g.B[1] + 63) / 64, (g.A[1] + 7) / 8, g.ns[1]);
  gq[2] = dim3((g.B[2] + 63) / 64, (g.A[2] + 7) / 8, g.ns[2]);
  const dim3 bq(256);
  if (!m->smith_ok || memcmp(&m->smith_f, f, sizeof *f)) {  // diagonal of this frequency
    kt_launch(KT_LU_SMITH, s, 48.0 * m->nsk, k_lu_smith<0>, gp[0], 256, 0, a);
    kt_launch(KT_LU_SMITH, s, 0.0, k_lu_smith<1>, gp[1], 256, 0, a);
    kt_launch(KT_LU_SMITH, s, 0.0, k_lu_smith<2>, gp[2], 256, 0, a);
    m->smith_f = *f;
    m->smith_ok = 1;
  }
  int tb[3];
  size_t lds[3];
  unsigned gs[3];
  for (int c = 0; c < 3; c++) {
    tb[c] = g.AP[c];  // one thread per row slot: the sweep's slots per diagonal
    lds[c] = (size_t)sg2_lds_bytes(g.A[c], g.B[c], tb[c] / 64);
    gs[c] = 8 * ((g.ns[c] + 7) / 8);
    if (lds[c] > 160 * 1024)
      return csem_fail(-3, "LU-SGS sweep of %d x %d cells needs %zu bytes of LDS", g.A[c], g.B[c], lds[c]);
  }
  double nc[3];
  for (int c = 0; c < 3; c++) nc[c] = (double)g.ns[c] * g.A[c] * g.B[c];
  // the fvm scaling is a template argument: a runtime branch in the sweep
  // loop would make the compiler drain the prefetched loads every step
#define SWEEP(C, F) sgs2_kernel<C, F>(m->fvm != 0, tb[C] / 64)
  {
    int rc = 0;
    if ((rc = sg2_lds_attr((const void *)SWEEP(0, true), lds[0])) || (rc = sg2_lds_attr((const void *)SWEEP(0, false), lds[0])) ||
        (rc = sg2_lds_attr((const void *)SWEEP(1, true), lds[1])) || (rc = sg2_lds_attr((const void *)SWEEP(1, false), lds[1])) ||
        (rc = sg2_lds_attr((const void *)SWEEP(2, true), lds[2])) || (rc = sg2_lds_attr((const void *)SWEEP(2, false), lds[2])))
      return rc;
  }
  // bytes per swept cell each kernel must move: the pre-pass reads b and
  // the other components' W values it differences and writes the record
  // (only b for Ex forward, not the two +0 units for Ey forward); the sweep
  // reads the record units it uses and the diagonal and writes W (not in
  // the last sweep) and y (backward sweeps).  The
  // records' round trip is the price of the parallel pre-pass: the whole
  // application needs 144 bytes per unknown (forward Ex 40, Ey 56, Ez 72,
  // backward 88 each: b, sigma, the values read, the value written).
  kt_launch(KT_LU_PRE, s, 32.0 * nc[0], k_lu_pre<0, true>, gq[0], bq, 0, a);
  kt_launch(KT_SGS_EX_F, s, 64.0 * nc[0], SWEEP(0, true), gs[0], tb[0] + 64, lds[0], a);
  kt_launch(KT_LU_PRE, s, 80.0 * nc[1], k_lu_pre<1, true>, gq[1], bq, 0, a);
  kt_launch(KT_SGS_EY_F, s, 96.0 * nc[1], SWEEP(1, true), gs[1], tb[1] + 64, lds[1], a);
  kt_launch(KT_LU_PRE, s, 128.0 * nc[2], k_lu_pre<2, true>, gq[2], bq, 0, a);
  kt_launch(KT_SGS_EZ_F, s, 128.0 * nc[2], SWEEP(2, true), gs[2], tb[2] + 64, lds[2], a);
  kt_launch(KT_LU_PRE, s, 144.0 * nc[2], k_lu_pre<2, false>, gq[2], bq, 0, a);
  kt_launch(KT_SGS_EZ_B, s, 144.0 * nc[2], SWEEP(2, false), gs[2], tb[2] + 64, lds[2], a);
  kt_launch(KT_LU_PRE, s, 144.0 * nc[1], k_lu_pre<1, false>, gq[1], bq, 0, a);
  kt_launch(KT_SGS_EY_B, s, 144.0 * nc[1], SWEEP(1, false), gs[1], tb[1] + 64, lds[1], a);
  kt_launch(KT_LU_PRE, s, 144.0 * nc[0], k_lu_pre<0, false>, gq[0], bq, 0, a);
  kt_launch(KT_SGS_EX_B, s, 128.0 * nc[0], SWEEP(0, false), gs[0], tb[0] + 64, lds[0], a);
#undef SWEEP
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int csem_apply(csem_model *m, const csem_freq *f, const void *x, void *y, void *stream)
{
  return run_apply(m, f, (const z2 *)x, (z2 *)y, S(stream));
}

extern "C" int csem_lusgs(csem_model *m, const csem_freq *f, const void *x, void *y, void *stream)
{
  return run_lusgs(m, f, (const z2 *)x, (z2 *)y, S(stream));
}

extern "C" int csem_curl_h(csem_model *m, const csem_freq *f, const void *x, void *h, void *stream)
{
  ApplyArgs a = apply_args(m, f, (const z2 *)x, NULL);
  long tot = (long)(m->n1 + 1) * (m->n2 + 1) * (m->n3 + 1);
  long lHx = (long)(m->n1 + 1) * m->n2 * m->n3, lHy = (long)m->n1 * (m->n2 + 1) * m->n3;
  k_curl_h<<<(unsigned)((tot + 255) / 256), 256, 0, S(stream)>>>(a, (z2 *)h, lHx, lHy);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int csem_airbc(csem_model *m, const void *hz0, void *hxg, void *hyg, void *stream)
{
  if (!m->airbc) return csem_fail(-3, "model has no airwave boundary");
  hipStream_t s = S(stream);
  HIP_OK(hipMemcpyAsync(m->hz0, hz0, (size_t)m->n1 * m->n2 * sizeof(z2), hipMemcpyDeviceToDevice, s));
  int rc = run_airbc(m, s, model_bufs(m));
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(hxg, m->hxg, (size_t)(m->n1 + 1) * m->n2 * sizeof(z2), hipMemcpyDeviceToDevice, s));
  HIP_OK(hipMemcpyAsync(hyg, m->hyg, (size_t)m->n1 * (m->n2 + 1) * sizeof(z2), hipMemcpyDeviceToDevice, s));
  return 0;
}

extern "C" int csem_fft2(csem_model *m, void *data, int dir, void *stream)
{
  if (!m->airbc) return csem_fail(-3, "model has no transform tables");
  const int N = m->nfft;
  hipStream_t s = S(stream);
  z2 *d = (z2 *)data;
  if (dir < 0) {
    k_fft_rows<<<N, 256, 2 * (size_t)N * sizeof(z2), s>>>(d, N, m->tw, -1);
    k_fft_cols<2><<<N / 2, 256, 2 * 2 * (size_t)N * sizeof(z2), s>>>(d, N, m->tw, -1);
  } else {
    k_fft_cols<2><<<N / 2, 256, 2 * 2 * (size_t)N * sizeof(z2), s>>>(d, N, m->tw, +1);
    k_fft_rows<<<N, 256, 2 * (size_t)N * sizeof(z2), s>>>(d, N, m->tw, +1);
  }
  HIP_OK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------- dots

static int dot_launch(long n, const z2 *a0, const z2 *b0, const z2 *a1, const z2 *b1, double *partials,
                      double *out, hipStream_t s)
{
  DotArgs g;
  g.n = n; g.a[0] = a0; g.b[0] = b0; g.a[1] = a1; g.b[1] = b1; g.partials = partials;
  long nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  // no chunk (n = 0): the fold alone writes the zero sum
  if (a1) {
    if (nch) k_dot_partial<2><<<(unsigned)nch, RED_T, 0, s>>>(g);
    k_dot_final<2><<<1, RED_T, 0, s>>>(partials, nch, out);
  } else {
    if (nch) k_dot_partial<1><<<(unsigned)nch, RED_T, 0, s>>>(g);
    k_dot_final<1><<<1, RED_T, 0, s>>>(partials, nch, out);
  }
  HIP_OK(hipGetLastError());
  return 0;
}

static int dot_partial_launch(long n, const z2 *a0, const z2 *b0, const z2 *a1, const z2 *b1, double *partials,
                              hipStream_t s)
{
  DotArgs g;
  g.n = n; g.a[0] = a0; g.b[0] = b0; g.a[1] = a1; g.b[1] = b1; g.partials = partials;
  long nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  // bytes: each distinct vector read once
  const void *v[4] = {a0, b0, a1, b1};
  int nv = 0;
  for (int i = 0; i < 4; i++) {
    bool seen = !v[i];
    for (int j = 0; j < i; j++) seen |= v[j] == v[i];
    nv += !seen;
  }
  if (a1) kt_launch(KT_DOT, s, 16.0 * nv * n, k_dot_partial<2>, (unsigned)nch, RED_T, 0, g);
  else kt_launch(KT_DOT, s, 16.0 * nv * n, k_dot_partial<1>, (unsigned)nch, RED_T, 0, g);
  HIP_OK(hipGetLastError());
  return 0;
}

// Reference summation order (cdotprod.c:18): one wavefront per dot product
// adds the terms strictly left to right; every lane runs the same dependent
// chain, fed by the 64 products of the current group through __shfl.
// Block b of the launch computes dot product b.
__global__ __launch_bounds__(64) void k_dot_serial(DotArgs g, double *out)
{
  const z2 *A = g.a[blockIdx.x], *B = g.b[blockIdx.x];
  const int l = threadIdx.x;
  const long n = g.n;
  double sr = 0.0, si = 0.0;
  for (long base = 0; base < n; base += 64) {
    const long e = base + l;
    double tr = 0, ti = 0;
    if (e < n) {
      z2 u = A[e], w = B[e];
      double c = w.x, d = -w.y;
      tr = u.x * c - u.y * d;
      ti = u.x * d + u.y * c;
    }
    if (n - base >= 64) {
#pragma unroll
      for (int q = 0; q < 64; q++) {
        sr = sr + __shfl(tr, q, 64);
        si = si + __shfl(ti, q, 64);
      }
    } else {
      for (int q = 0; q < (int)(n - base); q++) {
        sr = sr + __shfl(tr, q, 64);
        si = si + __shfl(ti, q, 64);
      }
    }
  }
  if (l == 0) {
    out[2 * blockIdx.x] = sr;
    out[2 * blockIdx.x + 1] = si;
  }
}

static int g_red_serial = 0;

extern "C" int csem_set_reduction(int serial)
{
  g_red_serial = serial ? 1 : 0;
  return 0;
}

static double *g_dot_partials = NULL;
static long g_dot_cap = 0;

extern "C" int csem_dot(long n, const void *a, const void *b, void *out, void *stream)
{
  if (g_red_serial) {
    DotArgs g;
    g.n = n; g.a[0] = (const z2 *)a; g.b[0] = (const z2 *)b; g.a[1] = NULL; g.b[1] = NULL; g.partials = NULL;
    k_dot_serial<<<1, 64, 0, S(stream)>>>(g, (double *)out);
    HIP_OK(hipGetLastError());
    return 0;
  }
  long nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  if (nch > g_dot_cap) {
    if (g_dot_partials) dev_free(g_dot_partials);
    g_dot_cap = 0;
    HIP_OK(dev_malloc((void **)&g_dot_partials, nch * 4 * sizeof(double)));
    g_dot_cap = nch;
  }
  return dot_launch(n, (const z2 *)a, (const z2 *)b, NULL, NULL, g_dot_partials, (double *)out, S(stream));
}

// csem_dot without allocating: the chunk partial sums go to a fixed device
// array instead of a buffer allocated on first use.  `scratch` names the
// array in device code only; host code uses the address hipGetSymbolAddress
// returns for the current device (each device holds its own copy).
#define CSEM_DOT_NOALLOC_CHUNKS 16384  // 16384 chunks of RED_CHUNK elements: n <= 2^26
static __device__ __align__(256) double scratch[CSEM_DOT_NOALLOC_CHUNKS * 4];
static double *g_scratch_dev[64];
static std::mutex g_scratch_mu;

static int scratch_ptr(double **out)
{
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return csem_fail(-3, "csem_dot_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = NULL;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return csem_fail(-1, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[dev] = (double *)p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

extern "C" int csem_dot_noalloc(long n, const void *a, const void *b, void *out, void *stream)
{
  if (n < 0 || n > (long)CSEM_DOT_NOALLOC_CHUNKS * RED_CHUNK)
    return csem_fail(-3, "csem_dot_noalloc: n=%ld outside [0, %ld]", n, (long)CSEM_DOT_NOALLOC_CHUNKS * RED_CHUNK);
  if (g_red_serial) {
    DotArgs g;
    g.n = n; g.a[0] = (const z2 *)a; g.b[0] = (const z2 *)b; g.a[1] = NULL; g.b[1] = NULL; g.partials = NULL;
    k_dot_serial<<<1, 64, 0, S(stream)>>>(g, (double *)out);
    HIP_OK(hipGetLastError());
    return 0;
  }
  double *part = NULL;
  int rc = scratch_ptr(&part);
  if (rc) return rc;
  return dot_launch(n, (const z2 *)a, (const z2 *)b, NULL, NULL, part, (double *)out, S(stream));
}

// ---------------------------------------------------------- BiCGStab

// Device scalar slots (doubles)
enum { SC_RHO = 0, SC_ALPHA = 2, SC_OMEGA = 4, SC_BETA = 6, SC_RS0 = 8, SC_RS = 9, SC_RELRES = 10,
       SC_CONV = 11, SC_TOL = 12, SC_DOT = 16 };

// r = b - v; p = r; r0 = r; partials of <r,r0>, <r,r> (solver.c:108-115)
__global__ __launch_bounds__(RED_T) void k_bicg_init(long n, const z2 *b, const z2 *v, z2 *r, z2 *p, z2 *r0,
                                                     double *partials)
{
  __shared__ double sh[RED_T];
  double acc[4] = {0, 0, 0, 0};
  const long base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int mm = 0; mm < RED_E; mm++) {
    long e = base + (long)mm * RED_T;
    if (e < n) {
      z2 rr = zsub(b[e], v[e]);
      r[e] = rr; p[e] = rr; r0[e] = rr;
      double c = rr.x, d = -rr.y;
      acc[0] = acc[0] + (rr.x * c - rr.y * d);
      acc[1] = acc[1] + (rr.x * d + rr.y * c);
      acc[2] = acc[2] + (rr.x * c - rr.y * d);
      acc[3] = acc[3] + (rr.x * d + rr.y * c);
    }
  }
  chunk_fold_store<2>(acc, partials, sh);
}

__global__ __launch_bounds__(RED_T) void k_bicg_fin_init(const double *partials, long nch, double *sc)
{
  __shared__ double sh[RED_T];
  if (nch >= 0) final_fold<2>(partials, nch, sc + SC_DOT, sh);  // nch < 0: dots already in sc
  if (threadIdx.x == 0) {
    sc[SC_RHO] = sc[SC_DOT];
    sc[SC_RHO + 1] = sc[SC_DOT + 1];
    double rs = sc[SC_DOT + 2];
    sc[SC_RS] = rs;
    sc[SC_RS0] = rs;
    sc[SC_RELRES] = sqrt(rs / rs);
    sc[SC_CONV] = 0;
  }
}

// alpha = rho_old / <v, r0> (solver.c:145-146)
__global__ __launch_bounds__(RED_T) void k_bicg_fin_alpha(const double *partials, long nch, double *sc)
{
  __shared__ double sh[RED_T];
  if (nch >= 0) final_fold<1>(partials, nch, sc + SC_DOT, sh);  // nch < 0: dots already in sc
  if (threadIdx.x == 0) {
    z2 al = zdiv(zmk(sc[SC_RHO], sc[SC_RHO + 1]), zmk(sc[SC_DOT], sc[SC_DOT + 1]));
    sc[SC_ALPHA] = al.x;
    sc[SC_ALPHA + 1] = al.y;
  }
}

// omega = <s,t> / Re<t,t> (solver.c:151)
__global__ __launch_bounds__(RED_T) void k_bicg_fin_omega(const double *partials, long nch, double *sc)
{
  __shared__ double sh[RED_T];
  if (nch >= 0) final_fold<2>(partials, nch, sc + SC_DOT, sh);  // nch < 0: dots already in sc
  if (threadIdx.x == 0) {
    double tt = sc[SC_DOT + 2];
    sc[SC_OMEGA] = sc[SC_DOT] / tt;
    sc[SC_OMEGA + 1] = sc[SC_DOT + 1] / tt;
  }
}

// relres = sqrt(Re<r,r>/rs0); beta = (rho_new/rho_old)*alpha/omega
// (solver.c:157-168)
__global__ __launch_bounds__(RED_T) void k_bicg_fin_conv(const double *partials, long nch, double *sc)
{
  __shared__ double sh[RED_T];
  if (nch >= 0) final_fold<2>(partials, nch, sc + SC_DOT, sh);  // nch < 0: dots already in sc
  if (threadIdx.x == 0) {
    double rs = sc[SC_DOT];
    double relres = sqrt(rs / sc[SC_RS0]);
    sc[SC_RS] = rs;
    sc[SC_RELRES] = relres;
    if (relres < sc[SC_TOL]) {
      sc[SC_CONV] = 1;
    } else {
      z2 rho_new = zmk(sc[SC_DOT + 2], sc[SC_DOT + 3]);
      z2 rho_old = zmk(sc[SC_RHO], sc[SC_RHO + 1]);
      z2 al = zmk(sc[SC_ALPHA], sc[SC_ALPHA + 1]), om = zmk(sc[SC_OMEGA], sc[SC_OMEGA + 1]);
      z2 be = zdiv(zmul(zdiv(rho_new, rho_old), al), om);
      sc[SC_BETA] = be.x;
      sc[SC_BETA + 1] = be.y;
      sc[SC_RHO] = rho_new.x;
      sc[SC_RHO + 1] = rho_new.y;
    }
  }
}

// s = r - alpha v (solver.c:147)
__global__ void k_bicg_s(long n, const z2 *r, const z2 *v, z2 *s, const double *sc)
{
  long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  z2 al = zmk(sc[SC_ALPHA], sc[SC_ALPHA + 1]);
  s[e] = zsub(r[e], zmul(al, v[e]));
}

// x += alpha xa + omega xb; r = s - omega t; partials of <r,r>, <r,r0>
// (solver.c:153-164)
__global__ __launch_bounds__(RED_T) void k_bicg_xr(long n, z2 *x, const z2 *xa, const z2 *xb, const z2 *s,
                                                   const z2 *t, z2 *r, const z2 *r0, const double *sc,
                                                   double *partials)
{
  __shared__ double sh[RED_T];
  double acc[4] = {0, 0, 0, 0};
  const z2 al = zmk(sc[SC_ALPHA], sc[SC_ALPHA + 1]), om = zmk(sc[SC_OMEGA], sc[SC_OMEGA + 1]);
  const long base = (long)blockIdx.x * RED_CHUNK + threadIdx.x;
  for (int mm = 0; mm < RED_E; mm++) {
    long e = base + (long)mm * RED_T;
    if (e < n) {
      x[e] = zadd(x[e], zadd(zmul(al, xa[e]), zmul(om, xb[e])));
      z2 rr = zsub(s[e], zmul(om, t[e]));
      r[e] = rr;
      double c = rr.x, d = -rr.y;
      acc[0] = acc[0] + (rr.x * c - rr.y * d);
      acc[1] = acc[1] + (rr.x * d + rr.y * c);
      z2 w = r0[e];
      c = w.x; d = -w.y;
      acc[2] = acc[2] + (rr.x * c - rr.y * d);
      acc[3] = acc[3] + (rr.x * d + rr.y * c);
    }
  }
  chunk_fold_store<2>(acc, partials, sh);
}

// p = r + beta (p - omega v) (solver.c:166)
__global__ void k_bicg_p(long n, z2 *p, const z2 *r, const z2 *v, const double *sc)
{
  long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  z2 be = zmk(sc[SC_BETA], sc[SC_BETA + 1]), om = zmk(sc[SC_OMEGA], sc[SC_OMEGA + 1]);
  p[e] = zadd(r[e], zmul(be, zsub(p[e], zmul(om, v[e]))));
}

// In the reference summation order the dot products go straight into the
// scalar block and the final kernels skip their fold (chunk count -1).
static long serial_dots(long n, const z2 *a0, const z2 *b0, const z2 *a1, const z2 *b1, double *sc, hipStream_t s,
                        long nch)
{
  if (!g_red_serial) return nch;
  DotArgs g;
  g.n = n; g.a[0] = a0; g.b[0] = b0; g.a[1] = a1; g.b[1] = b1; g.partials = NULL;
  k_dot_serial<<<a1 ? 2 : 1, 64, 0, s>>>(g, sc + SC_DOT);
  return -1;
}

static int ensure_kv(csem_model *m, int nv)
{
  if (m->nkv >= nv) return 0;
  if (m->kv) {
    dev_free(m->kv);
    m->allocs.erase(std::remove(m->allocs.begin(), m->allocs.end(), (void *)m->kv), m->allocs.end());
    m->bytes -= m->nkv * m->lE * (long)sizeof(z2);
    m->kv = NULL;
    m->nkv = 0;
  }
  int rc = dalloc(m, &m->kv, (long)nv * m->lE);
  if (rc) return rc;
  m->nkv = nv;
  return 0;
}

extern "C" int csem_op_apply(void *, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream)
{
  if (!m) return csem_fail(-4, "csem_op_apply: no model");
  return run_apply(m, f, (const z2 *)x, (z2 *)y, S(stream));
}

extern "C" int csem_op_lusgs(void *, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream)
{
  if (!m) return csem_fail(-4, "csem_op_lusgs: no model");
  return run_lusgs(m, f, (const z2 *)x, (z2 *)y, S(stream));
}

// ------------------------------------------------- A-phi preconditioner
//
// aphi_apply (fdfd.c:513-821): b_phi = -div x; one symmetric Gauss-Seidel
// pass from zero over the A-phi splitting, ax, ay, az, phi forward then
// phi, az, ay, ax backward; y = A + grad(phi).  Each row is the
// reference's expression tree (the arithmetic helpers of csem_common.h, IEEE
// division, Smith's complex division).  Each component's rows couple only
// within 2-D slices (ax: fixed i, rows (k, j); ay: fixed j, (k, i); az: fixed
// k, (j, i)) and phi's within the 3-D grid, so the lexicographic order is
// kept by sweeping diagonals (slow + fast = d) of all slices at once, and
// phi by planes k + j + i = d: every cell then reads its predecessors' new
// values and its successors' old ones, as in the serial loops.  A cell's
// own value, which the reference zeroes before forming its row, enters as
// zero.  One launch per diagonal.

struct AphiArgs {
  int n1, n2, n3, fvm;
  long lEx, lEy;
  const double *d1, *d2, *d3, *d1s, *d2s, *d3s;  // d3, d3s at the operator's first plane
  const double *s11, *s22, *s33;                 // conductivities at the operator's first plane
  const double *S11, *S22, *S33;                 // the same arrays from the full grid's first plane
  z2 iwmu, iomu;
  const z2 *bx, *by, *bz;
  z2 *ax, *ay, *az, *v, *bv;
};

__device__ __forceinline__ long a_idv(const AphiArgs &a, int k, int j, int i)
{
  return i + (long)(a.n1 + 1) * (j + (long)(a.n2 + 1) * k);
}
__device__ __forceinline__ z2 zneg(z2 x) { return zmk(-x.x, -x.y); }

// fdfd.c:554-566, plus zeroing A and phi (fdfd.c:570-573)
__global__ void k_aphi_init(AphiArgs a, long lE, long lv)
{
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int n1 = a.n1, n2 = a.n2;
  if (e < lE) a.ax[e] = zmk(0.0, 0.0);  // ax, ay, az are y
  if (e >= lv) return;
  a.v[e] = zmk(0.0, 0.0);
  const int i = (int)(e % (n1 + 1)), j = (int)((e / (n1 + 1)) % (n2 + 1)), k = (int)(e / ((long)(n1 + 1) * (n2 + 1)));
  z2 b = zmk(0.0, 0.0);
  if (k >= 1 && k < a.n3 && j >= 1 && j < n2 && i >= 1 && i < n1) {
    b = zadd(b, zdivr(zneg(zsub(a.bx[id1(n1, n2, k, j, i)], a.bx[id1(n1, n2, k, j, i - 1)])), a.d1[i]));
    b = zadd(b, zdivr(zneg(zsub(a.by[id2(n1, n2, k, j, i)], a.by[id2(n1, n2, k, j - 1, i)])), a.d2[j]));
    b = zadd(b, zdivr(zneg(zsub(a.bz[id3(n1, n2, k, j, i)], a.bz[id3(n1, n2, k - 1, j, i)])), a.d3[k]));
  }
  a.bv[e] = b;
}

__device__ __forceinline__ void aphi_fvm(const AphiArgs &a, int k, int j, int i, z2 &aij, z2 &aii)
{
  if (a.fvm) {
    const double vol = a.d1s[i] * a.d2[j] * a.d3[k];
    const z2 vq = zrmul(vol, a.iomu);
    aij = zmul(aij, vq);
    aii = zmul(aii, vq);
  }
}

// fdfd.c:585-604 / 768-787
__device__ void aphi_row_ax(const AphiArgs &a, int k, int j, int i)
{
  const int n1 = a.n1, n2 = a.n2, kp1 = k + 1, km1 = k - 1, jp1 = j + 1, jm1 = j - 1;
  const int ip1 = i + 1 < n1 ? i + 1 : n1;
  const z2 Z = zmk(0.0, 0.0);
  const z2 *ax = a.ax, *ay = a.ay, *az = a.az, *v = a.v;
  z2 t1 = zsub(zdivr(zsub(ay[id2(n1, n2, k, j, ip1)], ay[id2(n1, n2, k, j, i)]), a.d1s[i]),
               zdivr(zsub(ax[id1(n1, n2, k, jp1, i)], Z), a.d2s[j]));
  z2 t2 = zsub(zdivr(zsub(ay[id2(n1, n2, k, jm1, ip1)], ay[id2(n1, n2, k, jm1, i)]), a.d1s[i]),
               zdivr(zsub(Z, ax[id1(n1, n2, k, jm1, i)]), a.d2s[jm1]));
  z2 t3 = zsub(zdivr(zsub(ax[id1(n1, n2, kp1, j, i)], Z), a.d3s[k]),
               zdivr(zsub(az[id3(n1, n2, k, j, ip1)], az[id3(n1, n2, k, j, i)]), a.d1s[i]));
  z2 t4 = zsub(zdivr(zsub(Z, ax[id1(n1, n2, km1, j, i)]), a.d3s[km1]),
               zdivr(zsub(az[id3(n1, n2, km1, j, ip1)], az[id3(n1, n2, km1, j, i)]), a.d1s[i]));
  const double sg = a.s11[id1(n1, n2, k, j, i)];
  z2 aij = zsub(zmul(zsub(zdivr(zsub(t1, t2), a.d2[j]), zdivr(zsub(t3, t4), a.d3[k])), a.iwmu),
                zrmul(sg, zadd(Z, zdivr(zsub(v[a_idv(a, k, j, ip1)], v[a_idv(a, k, j, i)]), a.d1s[i]))));
  const double r = (1. / a.d2s[j] + 1. / a.d2s[jm1]) / a.d2[j] + (1. / a.d3s[k] + 1. / a.d3s[km1]) / a.d3[k];
  z2 aii = zsubr(zrmul(r, a.iwmu), sg);
  aphi_fvm(a, k, j, i, aij, aii);
  a.ax[id1(n1, n2, k, j, i)] = zdiv(zsub(a.bx[id1(n1, n2, k, j, i)], aij), aii);
}

// fdfd.c:606-627 / 746-766
__device__ void aphi_row_ay(const AphiArgs &a, int k, int j, int i)
{
  const int n1 = a.n1, n2 = a.n2, kp1 = k + 1, km1 = k - 1, jp1 = j + 1;
  const int ip1 = i + 1, im1 = i - 1;
  const z2 Z = zmk(0.0, 0.0);
  const z2 *ax = a.ax, *ay = a.ay, *az = a.az, *v = a.v;
  z2 t1 = zsub(zdivr(zsub(az[id3(n1, n2, k, jp1, i)], az[id3(n1, n2, k, j, i)]), a.d2s[j]),
               zdivr(zsub(ay[id2(n1, n2, kp1, j, i)], Z), a.d3s[k]));
  z2 t2 = zsub(zdivr(zsub(az[id3(n1, n2, km1, jp1, i)], az[id3(n1, n2, km1, j, i)]), a.d2s[j]),
               zdivr(zsub(Z, ay[id2(n1, n2, km1, j, i)]), a.d3s[km1]));
  z2 t3 = zsub(zdivr(zsub(ay[id2(n1, n2, k, j, ip1)], Z), a.d1s[i]),
               zdivr(zsub(ax[id1(n1, n2, k, jp1, i)], ax[id1(n1, n2, k, j, i)]), a.d2s[j]));
  z2 t4 = zsub(zdivr(zsub(Z, ay[id2(n1, n2, k, j, im1)]), a.d1s[im1]),
               zdivr(zsub(ax[id1(n1, n2, k, jp1, im1)], ax[id1(n1, n2, k, j, im1)]), a.d2s[j]));
  const double sg = a.s22[id2(n1, n2, k, j, i)];
  z2 aij = zsub(zmul(zsub(zdivr(zsub(t1, t2), a.d3[k]), zdivr(zsub(t3, t4), a.d1[i])), a.iwmu),
                zrmul(sg, zadd(Z, zdivr(zsub(v[a_idv(a, k, jp1, i)], v[a_idv(a, k, j, i)]), a.d2s[j]))));
  const double r = (1. / a.d3s[k] + 1. / a.d3s[km1]) / a.d3[k] + (1. / a.d1s[i] + 1. / a.d1s[im1]) / a.d1[i];
  z2 aii = zsubr(zrmul(r, a.iwmu), sg);
  aphi_fvm(a, k, j, i, aij, aii);
  a.ay[id2(n1, n2, k, j, i)] = zdiv(zsub(a.by[id2(n1, n2, k, j, i)], aij), aii);
}

// fdfd.c:628-649 / 724-744
__device__ void aphi_row_az(const AphiArgs &a, int k, int j, int i)
{
  const int n1 = a.n1, n2 = a.n2, kp1 = k + 1, jp1 = j + 1, jm1 = j - 1;
  const int ip1 = i + 1, im1 = i - 1;
  const z2 Z = zmk(0.0, 0.0);
  const z2 *ax = a.ax, *ay = a.ay, *az = a.az, *v = a.v;
  z2 t1 = zsub(zdivr(zsub(ax[id1(n1, n2, kp1, j, i)], ax[id1(n1, n2, k, j, i)]), a.d3s[k]),
               zdivr(zsub(az[id3(n1, n2, k, j, ip1)], Z), a.d1s[i]));
  z2 t2 = zsub(zdivr(zsub(ax[id1(n1, n2, kp1, j, im1)], ax[id1(n1, n2, k, j, im1)]), a.d3s[k]),
               zdivr(zsub(Z, az[id3(n1, n2, k, j, im1)]), a.d1s[im1]));
  z2 t3 = zsub(zdivr(zsub(az[id3(n1, n2, k, jp1, i)], Z), a.d2s[j]),
               zdivr(zsub(ay[id2(n1, n2, kp1, j, i)], ay[id2(n1, n2, k, j, i)]), a.d3s[k]));
  z2 t4 = zsub(zdivr(zsub(Z, az[id3(n1, n2, k, jm1, i)]), a.d2s[jm1]),
               zdivr(zsub(ay[id2(n1, n2, kp1, jm1, i)], ay[id2(n1, n2, k, jm1, i)]), a.d3s[k]));
  const double sg = a.s33[id3(n1, n2, k, j, i)];
  z2 aij = zsub(zmul(zsub(zdivr(zsub(t1, t2), a.d1[i]), zdivr(zsub(t3, t4), a.d2[j])), a.iwmu),
                zrmul(sg, zadd(Z, zdivr(zsub(v[a_idv(a, kp1, j, i)], v[a_idv(a, k, j, i)]), a.d3s[k]))));
  const double r = (1. / a.d1s[i] + 1. / a.d1s[im1]) / a.d1[i] + (1. / a.d2s[j] + 1. / a.d2s[jm1]) / a.d2[j];
  z2 aii = zsubr(zrmul(r, a.iwmu), sg);
  aphi_fvm(a, k, j, i, aij, aii);
  a.az[id3(n1, n2, k, j, i)] = zdiv(zsub(a.bz[id3(n1, n2, k, j, i)], aij), aii);
}

// fdfd.c:650-722; the conductivity behind the cell (t2, t4, t6) is read from
// the full grid's plane k as the reference does (emf->sigma*[k][j][i])
__device__ void aphi_row_v(const AphiArgs &a, int k, int j, int i)
{
  const int n1 = a.n1, n2 = a.n2, kp1 = k + 1, km1 = k - 1, jp1 = j + 1, jm1 = j - 1;
  const int ip1 = i + 1, im1 = i - 1;
  const z2 Z = zmk(0.0, 0.0);
  const z2 *ax = a.ax, *ay = a.ay, *az = a.az, *v = a.v;
  z2 t1 = zrmul(a.s11[id1(n1, n2, k, j, i)],
                zadd(ax[id1(n1, n2, k, j, i)], zdivr(zsub(v[a_idv(a, k, j, ip1)], Z), a.d1s[i])));
  z2 t2 = zrmul(a.S11[id1(n1, n2, k, j, im1)],
                zadd(ax[id1(n1, n2, k, j, im1)], zdivr(zsub(Z, v[a_idv(a, k, j, im1)]), a.d1s[im1])));
  z2 t3 = zrmul(a.s22[id2(n1, n2, k, j, i)],
                zadd(ay[id2(n1, n2, k, j, i)], zdivr(zsub(v[a_idv(a, k, jp1, i)], Z), a.d2s[j])));
  z2 t4 = zrmul(a.S22[id2(n1, n2, k, jm1, i)],
                zadd(ay[id2(n1, n2, k, jm1, i)], zdivr(zsub(Z, v[a_idv(a, k, jm1, i)]), a.d2s[jm1])));
  z2 t5 = zrmul(a.s33[id3(n1, n2, k, j, i)],
                zadd(az[id3(n1, n2, k, j, i)], zdivr(zsub(v[a_idv(a, kp1, j, i)], Z), a.d3s[k])));
  z2 t6 = zrmul(a.S33[id3(n1, n2, km1, j, i)],
                zadd(az[id3(n1, n2, km1, j, i)], zdivr(zsub(Z, v[a_idv(a, km1, j, i)]), a.d3s[km1])));
  z2 aij = zadd(zadd(zdivr(zsub(t1, t2), a.d1[i]), zdivr(zsub(t3, t4), a.d2[j])), zdivr(zsub(t5, t6), a.d3[k]));
  const double r = -((1. / a.d1s[i] + 1. / a.d1s[im1]) / a.d1[i] + (1. / a.d2s[j] + 1. / a.d2s[jm1]) / a.d2[j] +
                     (1. / a.d3s[k] + 1. / a.d3s[km1]) / a.d3[k]);
  z2 aii = zmk(r, 0.0);
  aphi_fvm(a, k, j, i, aij, aii);
  a.v[a_idv(a, k, j, i)] = zdiv(zsub(a.bv[a_idv(a, k, j, i)], aij), aii);
}

// One diagonal slow + fast = d of every 2-D slice of component C: thread
// (slice, slow) takes fast = d - slow when it is in range.  Slices: C=0 i,
// C=1 j, C=2 k; (slow, fast): (k, j), (k, i), (j, i).  lo*/hi* are the
// inclusive row ranges of fdfd.c's loops.
template <int C>
__global__ void k_aphi_diag(AphiArgs a, int d, int nslice, int lo_s, int hi_s, int lo_f, int hi_f)
{
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int slice = (int)(t % nslice), slow = lo_s + (int)(t / nslice), fast = d - slow;
  if (slow > hi_s || fast < lo_f || fast > hi_f) return;
  if (C == 0) aphi_row_ax(a, slow, fast, slice);
  else if (C == 1) aphi_row_ay(a, slow, slice, fast);
  else aphi_row_az(a, slice, slow, fast);
}

// One plane k + j + i = d of phi's rows (k, j, i in [1, n-1]).
__global__ void k_aphi_vplane(AphiArgs a, int d)
{
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int nj = a.n2 - 1;
  const int j = 1 + (int)(t % nj), k = 1 + (int)(t / nj), i = d - k - j;
  if (k > a.n3 - 1 || i < 1 || i > a.n1 - 1) return;
  aphi_row_v(a, k, j, i);
}

// y = A + grad(phi) in place (fdfd.c:793-820)
__global__ void k_aphi_grad(AphiArgs a, long lE)
{
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= lE) return;
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  int i, j, k;
  if (e < a.lEx) {
    i = (int)(e % n1); j = (int)((e / n1) % (n2 + 1)); k = (int)(e / ((long)n1 * (n2 + 1)));
    if (k >= 1 && k < n3 && j >= 1 && j < n2)
      a.ax[e] = zadd(a.ax[e], zdivr(zsub(a.v[a_idv(a, k, j, i + 1)], a.v[a_idv(a, k, j, i)]), a.d1s[i]));
  } else if (e < a.lEx + a.lEy) {
    const long q = e - a.lEx;
    i = (int)(q % (n1 + 1)); j = (int)((q / (n1 + 1)) % n2); k = (int)(q / ((long)(n1 + 1) * n2));
    if (k >= 1 && k < n3 && i >= 1 && i < n1)
      a.ax[e] = zadd(a.ax[e], zdivr(zsub(a.v[a_idv(a, k, j + 1, i)], a.v[a_idv(a, k, j, i)]), a.d2s[j]));
  } else {
    const long q = e - a.lEx - a.lEy;
    i = (int)(q % (n1 + 1)); j = (int)((q / (n1 + 1)) % (n2 + 1)); k = (int)(q / ((long)(n1 + 1) * (n2 + 1)));
    if (j >= 1 && j < n2 && i >= 1 && i < n1)
      a.ax[e] = zadd(a.ax[e], zdivr(zsub(a.v[a_idv(a, k + 1, j, i)], a.v[a_idv(a, k, j, i)]), a.d3s[k]));
  }
}

template <int C>
static int aphi_sweep(const AphiArgs &a, bool fwd, hipStream_t s)
{
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  int nslice, lo_s, hi_s, lo_f, hi_f;
  if (C == 0) { nslice = n1; lo_s = 1; hi_s = n3 - 1; lo_f = 1; hi_f = n2 - 1; }
  else if (C == 1) { nslice = n2; lo_s = 1; hi_s = n3 - 1; lo_f = 1; hi_f = n1 - 1; }
  else { nslice = n3; lo_s = 1; hi_s = n2 - 1; lo_f = 1; hi_f = n1 - 1; }
  if (hi_s < lo_s || hi_f < lo_f || nslice <= 0) return 0;
  const long nt = (long)nslice * (hi_s - lo_s + 1);
  const unsigned nb = (unsigned)((nt + 255) / 256);
  const int d0 = lo_s + lo_f, d1 = hi_s + hi_f;
  for (int q = 0; q <= d1 - d0; q++) k_aphi_diag<C><<<nb, 256, 0, s>>>(a, fwd ? d0 + q : d1 - q, nslice, lo_s, hi_s, lo_f, hi_f);
  HIP_OK(hipGetLastError());
  return 0;
}

static int aphi_vsweep(const AphiArgs &a, bool fwd, hipStream_t s)
{
  if (a.n1 < 2 || a.n2 < 2 || a.n3 < 2) return 0;
  const long nt = (long)(a.n3 - 1) * (a.n2 - 1);
  const unsigned nb = (unsigned)((nt + 255) / 256);
  const int d0 = 3, d1 = (a.n1 - 1) + (a.n2 - 1) + (a.n3 - 1);
  for (int q = 0; q <= d1 - d0; q++) k_aphi_vplane<<<nb, 256, 0, s>>>(a, fwd ? d0 + q : d1 - q);
  HIP_OK(hipGetLastError());
  return 0;
}

static int run_aphi(csem_model *m, const csem_freq *f, const z2 *x, z2 *y, hipStream_t s)
{
  const int n1 = m->n1, n2 = m->n2, n3 = m->n3, ks = m->kofs;
  const long lv = (long)(n1 + 1) * (n2 + 1) * (n3 + 1);
  int rc;
  if (!m->aphi_v && ((rc = dalloc(m, &m->aphi_v, lv)) || (rc = dalloc(m, &m->aphi_bv, lv)))) return rc;
  AphiArgs a;
  a.n1 = n1; a.n2 = n2; a.n3 = n3; a.fvm = m->fvm; a.lEx = m->lEx; a.lEy = m->lEy;
  a.d1 = m->d1; a.d2 = m->d2; a.d3 = m->d3 + ks; a.d1s = m->d1s; a.d2s = m->d2s; a.d3s = m->d3s + ks;
  a.s11 = m->s11 + (long)n1 * (n2 + 1) * ks;
  a.s22 = m->s22 + (long)(n1 + 1) * n2 * ks;
  a.s33 = m->s33 + (long)(n1 + 1) * (n2 + 1) * ks;
  a.S11 = m->s11; a.S22 = m->s22; a.S33 = m->s33;
  a.iwmu = zmk(f->iwmu_re, f->iwmu_im);
  a.iomu = zmk(f->iomu_re, f->iomu_im);
  a.bx = x; a.by = x + m->lEx; a.bz = x + m->lEx + m->lEy;
  a.ax = y; a.ay = y + m->lEx; a.az = y + m->lEx + m->lEy;
  a.v = m->aphi_v; a.bv = m->aphi_bv;
  const long ne = lv > m->lE ? lv : m->lE;
  k_aphi_init<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(a, m->lE, lv);
  HIP_OK(hipGetLastError());
  if ((rc = aphi_sweep<0>(a, true, s)) || (rc = aphi_sweep<1>(a, true, s)) || (rc = aphi_sweep<2>(a, true, s)) ||
      (rc = aphi_vsweep(a, true, s)) || (rc = aphi_vsweep(a, false, s)) || (rc = aphi_sweep<2>(a, false, s)) ||
      (rc = aphi_sweep<1>(a, false, s)) || (rc = aphi_sweep<0>(a, false, s)))
    return rc;
  k_aphi_grad<<<(unsigned)((m->lE + 255) / 256), 256, 0, s>>>(a, m->lE);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int csem_op_aphi(void *, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream)
{
  if (!m) return csem_fail(-4, "csem_op_aphi: no model");
  return run_aphi(m, f, (const z2 *)x, (z2 *)y, S(stream));
}

// Krylov work vectors, scalars and partial sums: the model's when the solve
// has one, otherwise the call's own (a solve over host operators only).
struct KrylovWork {
  z2 *kv;
  double *scal, *partials;
  void *own[3];
};

static int kwork_get(csem_model *m, long n, int nv, KrylovWork *w)
{
  memset(w, 0, sizeof *w);
  if (m) {
    int rc = ensure_kv(m, nv);
    if (rc) return rc;
    w->kv = m->kv; w->scal = m->scal; w->partials = m->partials;
    return 0;
  }
  const long nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  if (dev_malloc(&w->own[0], (long)nv * n * (long)sizeof(z2)) != hipSuccess ||
      dev_malloc(&w->own[1], 64 * sizeof(double)) != hipSuccess ||
      dev_malloc(&w->own[2], nch * 4 * (long)sizeof(double)) != hipSuccess)
    return csem_fail(-2, "Krylov work vectors (%ld x %d)", n, nv);
  HIP_OK(dev_memset(w->own[1], 0, 64 * sizeof(double)));
  w->kv = (z2 *)w->own[0]; w->scal = (double *)w->own[1]; w->partials = (double *)w->own[2];
  return 0;
}

static void kwork_release(KrylovWork *w)
{
  for (int q = 0; q < 3; q++)
    if (w->own[q]) dev_free(w->own[q]);
}

// Flexible right-preconditioned BiCGStab (solver.c:93-178) or, with M =
// NULL, BiCGStab (solver.c:16-87), all vectors and scalars resident.  The
// operators are called through their pointers; the library's own run as
// device kernels, a caller's host operator through host copies.
static int bicgstab_run(csem_model *m, const csem_freq *f, long n, z2 *x, const z2 *b, int niter, double tol,
                        csem_op A, void *actx, csem_op M, void *mctx, double *hist, hipStream_t s,
                        const KrylovWork &w)
{
  const long nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  int rc;
  z2 *r = w.kv, *r0 = r + n, *p = r0 + n, *pp = p + n, *v = pp + n, *sv = v + n, *t = sv + n;
  double *sc = w.scal, *pa = w.partials;
  const unsigned ge = (unsigned)((n + 255) / 256);
  double hs[16];
  HIP_OK(hipMemcpyAsync(sc + SC_TOL, &tol, sizeof(double), hipMemcpyHostToDevice, s));
  if ((rc = A(actx, m, f, x, v, s))) return rc;
  k_bicg_init<<<(unsigned)nch, RED_T, 0, s>>>(n, b, v, r, p, r0, pa);
  k_bicg_fin_init<<<1, RED_T, 0, s>>>(pa, serial_dots(n, r, r0, r, r, sc, s, nch), sc);
  HIP_OK(hipMemcpyAsync(hs, sc, 16 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (hs[SC_RS] == 0.) return 0;
  double relres = hs[SC_RELRES];
  if (hist) hist[0] = relres;
  int iter;
  for (iter = 0; iter < niter; ++iter) {
    if (hist) hist[iter] = relres;
    const z2 *pin = p;
    if (M) {
      if ((rc = M(mctx, m, f, p, pp, s))) return rc;
      pin = pp;
    }
    if ((rc = A(actx, m, f, pin, v, s))) return rc;
    if (!g_red_serial && (rc = dot_partial_launch(n, v, r0, NULL, NULL, pa, s))) return rc;
    k_bicg_fin_alpha<<<1, RED_T, 0, s>>>(pa, serial_dots(n, v, r0, NULL, NULL, sc, s, nch), sc);
    kt_launch(KT_BICG_S, s, 48.0 * n, k_bicg_s, ge, 256, 0, n, r, v, sv, sc);
    const z2 *xb;
    if (M) {
      if ((rc = M(mctx, m, f, sv, r, s))) return rc;
      if ((rc = A(actx, m, f, r, t, s))) return rc;
      if (!g_red_serial && (rc = dot_partial_launch(n, sv, t, t, t, pa, s))) return rc;
      xb = r;
    } else {
      if ((rc = A(actx, m, f, sv, t, s))) return rc;
      if (!g_red_serial && (rc = dot_partial_launch(n, t, sv, t, t, pa, s))) return rc;
      xb = sv;
    }
    const long nom = M ? serial_dots(n, sv, t, t, t, sc, s, nch) : serial_dots(n, t, sv, t, t, sc, s, nch);
    k_bicg_fin_omega<<<1, RED_T, 0, s>>>(pa, nom, sc);
    kt_launch(KT_BICG_XR, s, 128.0 * n, k_bicg_xr, (unsigned)nch, RED_T, 0, n, x, pin, xb, sv, t, r, r0, sc, pa);
    k_bicg_fin_conv<<<1, RED_T, 0, s>>>(pa, serial_dots(n, r, r, r, r0, sc, s, nch), sc);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(hs, sc, 16 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    kt_collect();
    relres = hs[SC_RELRES];
    if (hs[SC_CONV] != 0) {
      if (hist) hist[iter + 1] = relres;
      break;
    }
    kt_launch(KT_BICG_P, s, 64.0 * n, k_bicg_p, ge, 256, 0, n, p, r, v, sc);
  }
  HIP_OK(hipGetLastError());
  return iter;
}

extern "C" int csem_bicgstab_op(csem_model *m, const csem_freq *f, long n, void *xv, const void *bv, int niter,
                                double tol, csem_op A, void *actx, csem_op M, void *mctx, double *hist,
                                void *stream)
{
  if (!A) return csem_fail(-3, "csem_bicgstab_op: no operator");
  if (m && n != m->lE) return csem_fail(-3, "csem_bicgstab_op: n=%ld, the model's vectors hold %ld", n, m->lE);
  if (n <= 0) return csem_fail(-3, "csem_bicgstab_op: n=%ld", n);
  KrylovWork w;
  int rc = kwork_get(m, n, 7, &w);
  if (!rc) rc = bicgstab_run(m, f, n, (z2 *)xv, (const z2 *)bv, niter, tol, A, actx, M, mctx, hist, S(stream), w);
  if (!m) {
    (void)hipStreamSynchronize(S(stream));
    kwork_release(&w);
  }
  return rc;
}

extern "C" int csem_bicgstab(csem_model *m, const csem_freq *f, void *xv, const void *bv, int niter,
                             double tol, int precond, double *hist, void *stream)
{
  return csem_bicgstab_op(m, f, m->lE, xv, bv, niter, tol, csem_op_apply, NULL, precond ? csem_op_lusgs : NULL,
                          NULL, hist, stream);
}

// ------------------------------------------------- vector primitives

__global__ void k_vec_sub(long n, const z2 *a, const z2 *b, z2 *o)
{
  long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) o[e] = zsub(a[e], b[e]);
}
__global__ void k_vec_divr(long n, const z2 *a, double r, z2 *o)
{
  long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) o[e] = zdivr(a[e], r);
}
__global__ void k_vec_divz(long n, const z2 *a, z2 q, z2 *o)
{
  long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) o[e] = zdiv(a[e], q);
}
// y = y - h*x, h read from device memory (solver.c:447)
__global__ void k_vec_subscaled_dev(long n, z2 *y, const double *h, const z2 *x)
{
  long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) y[e] = zsub(y[e], zmul(zmk(h[0], h[1]), x[e]));
}
// y = y + c*x (solver.c:497)
__global__ void k_vec_addscaled(long n, z2 *y, z2 c, const z2 *x)
{
  long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) y[e] = zadd(y[e], zmul(c, x[e]));
}
__global__ void k_scatter_add(long nnz, const long *idx, const z2 *val, z2 *b)
{
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (long q = 0; q < nnz; q++) b[idx[q]] = zadd(b[idx[q]], val[q]);
}
// one thread per receiver: s = 0; s += ((sig*x)*wa)*wb)*wc for the 8
// corners in the order of inject_extract.c:151-158 / 199-206; then s*scale
__global__ void k_gather(int nrec, const long *idx, const double *wts, const double *sig, const double *scale,
                         const z2 *x, z2 *out)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  z2 s = zmk(0, 0);
  for (int q = 0; q < 8; q++) {
    long e = idx[8 * r + q];
    double sg = sig[8 * r + q];
    z2 v = zrmul(sg, x[e]);  // 1.0 for Ex/Ey, which multiply by nothing: exact
    v = zmulr(zmulr(zmulr(v, wts[24 * r + 3 * q]), wts[24 * r + 3 * q + 1]), wts[24 * r + 3 * q + 2]);
    s = zadd(s, v);
  }
  out[r] = zmulr(s, scale[r]);
}

#define GE(n) (unsigned)(((n) + 255) / 256), 256, 0, S(stream)

extern "C" int csem_vec_sub(long n, const void *a, const void *b, void *o, void *stream)
{
  k_vec_sub<<<GE(n)>>>(n, (const z2 *)a, (const z2 *)b, (z2 *)o);
  HIP_OK(hipGetLastError());
  return 0;
}
extern "C" int csem_vec_divr(long n, const void *a, double r, void *o, void *stream)
{
  k_vec_divr<<<GE(n)>>>(n, (const z2 *)a, r, (z2 *)o);
  HIP_OK(hipGetLastError());
  return 0;
}
extern "C" int csem_vec_divz(long n, const void *a, double re, double im, void *o, void *stream)
{
  k_vec_divz<<<GE(n)>>>(n, (const z2 *)a, zmk(re, im), (z2 *)o);
  HIP_OK(hipGetLastError());
  return 0;
}
extern "C" int csem_vec_subscaled_dev(long n, void *y, const void *h, const void *x, void *stream)
{
  k_vec_subscaled_dev<<<GE(n)>>>(n, (z2 *)y, (const double *)h, (const z2 *)x);
  HIP_OK(hipGetLastError());
  return 0;
}
extern "C" int csem_vec_addscaled(long n, void *y, double re, double im, const void *x, void *stream)
{
  k_vec_addscaled<<<GE(n)>>>(n, (z2 *)y, zmk(re, im), (const z2 *)x);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int csem_scatter_add(long nnz, const long *idx, const double *val, void *b, void *stream)
{
  if (nnz <= 0) return 0;
  long *di;
  z2 *dv;
  HIP_OK(dev_malloc((void **)&di, nnz * sizeof(long)));
  HIP_OK(dev_malloc((void **)&dv, nnz * sizeof(z2)));
  HIP_OK(hipMemcpyAsync(di, idx, nnz * sizeof(long), hipMemcpyHostToDevice, S(stream)));
  HIP_OK(hipMemcpyAsync(dv, val, nnz * sizeof(z2), hipMemcpyHostToDevice, S(stream)));
  k_scatter_add<<<1, 64, 0, S(stream)>>>(nnz, di, dv, (z2 *)b);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(S(stream)));
  dev_free(di);
  dev_free(dv);
  return 0;
}

extern "C" int csem_gather(int nrec, const long *idx, const double *wts, const double *sig, const double *scale,
                           const void *x, double *out, void *stream)
{
  if (nrec <= 0) return 0;
  char *blob;
  size_t bi = 8 * (size_t)nrec * sizeof(long), bw = 24 * (size_t)nrec * sizeof(double);
  size_t bs = 8 * (size_t)nrec * sizeof(double), bc = (size_t)nrec * sizeof(double), bo = (size_t)nrec * sizeof(z2);
  HIP_OK(dev_malloc((void **)&blob, bi + bw + bs + bc + bo));
  hipStream_t s = S(stream);
  HIP_OK(hipMemcpyAsync(blob, idx, bi, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(blob + bi, wts, bw, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(blob + bi + bw, sig, bs, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(blob + bi + bw + bs, scale, bc, hipMemcpyHostToDevice, s));
  k_gather<<<(nrec + 63) / 64, 64, 0, s>>>(nrec, (const long *)blob, (const double *)(blob + bi),
                                           (const double *)(blob + bi + bw), (const double *)(blob + bi + bw + bs),
                                           (const z2 *)x, (z2 *)(blob + bi + bw + bs + bc));
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out, blob + bi + bw + bs + bc, bo, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  dev_free(blob);
  return 0;
}

// ------------------------------------------------------ memory helpers

extern "C" int csem_malloc(void **p, long bytes)
{
  HIP_OK(dev_malloc(p, bytes > 0 ? bytes : 256));
  return 0;
}
extern "C" int csem_free(void *p)
{
  dev_free(p);
  return 0;
}
extern "C" int csem_memcpy_h2d(void *dst, const void *src, long bytes, void *stream)
{
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)));
  HIP_OK(hipStreamSynchronize(S(stream)));
  return 0;
}
extern "C" int csem_memcpy_d2h(void *dst, const void *src, long bytes, void *stream)
{
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)));
  HIP_OK(hipStreamSynchronize(S(stream)));
  return 0;
}
// A plain streaming copy, one 16-byte load and store per thread: the
// reference kernel against which hardware byte counters are scaled, and
// the bandwidth a streaming kernel reaches (measured: one element per
// thread over many blocks beats grid-stride loops and several elements
// per thread).
__global__ __launch_bounds__(256) void k_copy16(long n16, const float4 *src, float4 *dst)
{
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e < n16) dst[e] = src[e];
}
extern "C" int csem_copy(void *dst, const void *src, long bytes, void *stream)
{
  if (bytes <= 0) return 0;
  if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
    return csem_fail(-3, "csem_copy needs 16-byte aligned buffers and a multiple of 16 bytes");
  const long n16 = bytes / 16;
  if ((n16 + 255) / 256 > 0x7fffffffL) return csem_fail(-3, "csem_copy: %ld bytes is too large for one launch", bytes);
  k_copy16<<<(unsigned)((n16 + 255) / 256), 256, 0, S(stream)>>>(n16, (const float4 *)src, (float4 *)dst);
  HIP_OK(hipGetLastError());
  return 0;
}
extern "C" int csem_memset(void *p, int v, long bytes, void *stream)
{
  HIP_OK(hipMemsetAsync(p, v, bytes, S(stream)));
  return 0;
}
extern "C" int csem_sync(void *stream)
{
  HIP_OK(hipStreamSynchronize(S(stream)));
  return 0;
}
