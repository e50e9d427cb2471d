// csem_internal.h -- state shared by the FP64 and mixed-precision
// translation units of libcsem (not part of the public C ABI).
#pragma once
#include <cstdint>
#include <mutex>
#include <vector>
#include <hip/hip_ext.h>
#include "csem_kernels.h"
#include "../../../include/csem_dev.h"

// ------------------------------------------------ device-wide operations
// hipFree waits for the whole device, and synchronous copies and fills run
// on the null stream; issued while another thread captures a graph they
// can invalidate that capture.  Every such call in the library, every
// allocation, and every capture (mixed-precision solve slots) hold this
// lock, so solves on several threads can capture while others allocate,
// upload a model or free one.
extern std::mutex g_devwide_mu;
static inline void dev_free(void *p)
{
  std::lock_guard<std::mutex> lk(g_devwide_mu);
  (void)hipFree(p);
}
static inline hipError_t dev_malloc(void **p, long bytes)
{
  std::lock_guard<std::mutex> lk(g_devwide_mu);
  return hipMalloc(p, bytes);
}
static inline hipError_t dev_memset(void *p, int v, long bytes)
{
  std::lock_guard<std::mutex> lk(g_devwide_mu);
  return hipMemset(p, v, bytes);
}
static inline hipError_t dev_upload(void *dst, const void *src, long bytes)
{
  std::lock_guard<std::mutex> lk(g_devwide_mu);
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
}

// ------------------------------------------------------- kernel timers
// Optional HIP-event timing of every launch of the hot kernels, on the
// stream they run on, with the algorithmic bytes each launch moves.
enum {
  KT_APPLY = 0, KT_HZ0, KT_AIR_ROWS_FWD, KT_AIR_COLS, KT_AIR_ROWS_INV, KT_AIR_GHOST,
  KT_SGS_EX_F, KT_SGS_EY_F, KT_SGS_EZ_F, KT_SGS_EZ_B, KT_SGS_EY_B, KT_SGS_EX_B,
  KT_DOT, KT_DOT_FIN, KT_BICG_S, KT_BICG_XR, KT_BICG_P, KT_LU_PRE, KT_LU_SMITH,
  KT_M_APPLY, KT_M_HZ0, KT_M_AIR_ROWS, KT_M_AIR_COLS, KT_M_AIR_RINV, KT_M_AIR_GHOST,
  KT_M_PRE, KT_M_SGS, KT_M_VEC, KT_M_DOT, KT_M_REFINE, KT_M_UNSKEW, KT_M_ORTH, KT_N
};
// An event pair for the next launch when timing is on and not paused
// (graph capture pauses it), and the pair's record once launched.
bool kt_events(hipEvent_t *a, hipEvent_t *b);
void kt_record(int id, hipEvent_t a, hipEvent_t b, double bytes);
// Repeat loops (csem_kernel_repeat): the repeat count for this launch of
// kernel id (0: none) with the loop's event pair, and the loop's record.
int kt_repeat(int id, hipEvent_t *a, hipEvent_t *b);
void kt_record_loop(int id, hipEvent_t a, hipEvent_t b, int n);
// Every hot kernel is launched through here.  Timed launches go through
// hipExtLaunchKernelGGL with the launch's event pair; a launch that
// repeats is followed by its back-to-back copies between one more pair.
template <typename... P, typename... A>
static inline void kt_launch(int id, hipStream_t s, double bytes, void (*k)(P...), dim3 g, dim3 b, size_t shm,
                             A... args)
{
  hipEvent_t e0, e1;
  if (kt_events(&e0, &e1)) {
    hipExtLaunchKernelGGL(k, g, b, (std::uint32_t)shm, s, e0, e1, 0, static_cast<P>(args)...);
    kt_record(id, e0, e1, bytes);
    const int n = kt_repeat(id, &e0, &e1);
    if (n > 0) {
      (void)hipEventRecord(e0, s);
      for (int i = 0; i < n; i++) k<<<g, b, shm, s>>>(static_cast<P>(args)...);
      (void)hipEventRecord(e1, s);
      kt_record_loop(id, e0, e1, n);
    }
  } else {
    k<<<g, b, shm, s>>>(static_cast<P>(args)...);
  }
}
// Right-hand sides a batched launch works on, for the timers' byte counts:
// the batch kernels skip right-hand sides whose inner solve is done, and
// the host learns which are done at every convergence check, so it sets
// the count of the others here (0: all R).  A right-hand side that finishes
// between two checks is counted until the next one.
extern thread_local int g_kt_live;
static inline double kt_rhs(int R) { return g_kt_live > 0 && g_kt_live < R ? (double)g_kt_live : (double)R; }


struct MixState;
struct csem_model {
  int n1, n2, n3f, n3, nb, airbc, fvm, kofs, nfft, kc;
  long lEx, lEy, lEz, lE, fEx, fEy, fEz, fE;
  double *d1, *d2, *d3, *d1s, *d2s, *d3s;
  double *r1, *r2, *r3, *r1s, *r2s, *r3s;  // RN(1/d) of the spacings
  double *s11, *s22, *s33;
  SkewGeo g;       // LU-SGS sweep geometry
  long nsk;        // swept cells, all components
  z2 *W;           // LU-SGS state on the full grid
  LuRec *rec;      // LU-SGS pre-pass records of one component, skewed
  double4 *smith;  // per-frequency diagonal, skewed Ex | Ey | Ez
  csem_freq smith_f;
  int smith_ok;
  z2 *sH1, *sH2, *tw, *buf, *bufx, *bufy, *hz0, *hxg, *hyg;
  int *jx, *jy, *rows, nrowx, nrowy;
  double *wx, *wy;
  int *gxj1, *gxj2, *gyj1, *gyj2;
  double *gxw1, *gxw2, *gyw1, *gyw2;
  double *partials, *scal;
  long nch;
  z2 *kv;  // Krylov work vectors (lazily allocated)
  z2 *aphi_v, *aphi_bv;  // A-phi potential and its right-hand side (lazily allocated)
  int nkv;
  long bytes;
  std::vector<void *> allocs;
  MixState *mx;  // mixed-precision state, created on first use
};

template <class T>
static inline int dalloc(csem_model *m, T **p, long count)
{
  long bytes = count * (long)sizeof(T);
  if (bytes <= 0) bytes = 256;
  hipError_t e = dev_malloc((void **)p, bytes);
  if (e != hipSuccess) return csem_fail(-2, "hipMalloc(%ld): %s", bytes, hipGetErrorString(e));
  m->allocs.push_back((void *)*p);
  m->bytes += bytes;
  return 0;
}

template <class T>
static inline int upload(csem_model *m, T **p, const void *src, long count)
{
  int rc = dalloc(m, p, count);
  if (rc) return rc;
  HIP_OK(dev_upload(*p, src, count * sizeof(T)));
  return 0;
}

void mixed_free(csem_model *m);

// scratch of the airwave boundary of one FP64 operator application
struct AirBufs {
  z2 *hz0, *buf, *bufx, *bufy, *hxg, *hyg;
};
int run_apply_bufs(csem_model *m, const csem_freq *f, const z2 *x, z2 *y, hipStream_t s, const AirBufs &b);
