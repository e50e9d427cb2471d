// csem_acsr.hip -- the operator assembled as a CSR matrix on the device
// (the matrix-storage path of Acsr.c), as a brute-force cross-check of the
// matrix-free kernels.
//
// The matrix is iw mu0 times the curl-curl operator of fdfd.c on the full
// (explicit-air) grid, with truncated stencils on the boundary rows where
// fdfd_apply has zero rows: row r of component Ex / Ey / Ez holds its
// entries in the order and with the expressions of Acsr.c:22-693 (the same
// coefficient 1/(d d') per stencil group, the diagonal accumulated over the
// four own-axis groups in that order, minus i w mu0 sigma), then the
// diagonal.  Unlike Acsr.c, whose row pointer is set before each row's
// diagonal is stored (Acsr.c:425-428, 553-556, 680-683) so that every
// diagonal lands in the next row's range, each row's range here holds its
// own diagonal: the products are those of the assembled matrix.
//
// Build: a count kernel (one thread per row), an exclusive scan of the
// counts (hipCUB), a fill kernel.  Products: y = A x one thread per row,
// entries summed in row order; the symmetric Gauss-Seidel of
// Acsr.c:719-748 (y = 0, a forward sweep over the rows, then a backward
// one, each row y_r = (x_r - sum_{c != r} a_rc y_c) / a_rr in its entry
// order) runs as wavefronts: Ex rows couple to Ex rows only through
// (j +- 1, k +- 1), Ey through (k +- 1, i +- 1), Ez through (j +- 1,
// i +- 1), and every other coupling is to a whole component that the
// serial order finishes before (or starts after) this one, so the rows of
// one component with equal j + k (Ex), k + i (Ey) or j + i (Ez) are
// independent and see exactly the values the serial order gives them.
// One launch per such level.
#include <math.h>
#include <hipcub/hipcub.hpp>
#define CSEM_TYPES_ONLY
#include "csem_internal.h"

struct csem_csr {
  int n1, n2, n3;
  long lEx, lEy, lE;  // rows: Ex | Ey | Ez, each x fastest (id1/id2/id3)
  long nnz;
  double *d1, *d2, *d3, *d1s, *d2s, *d3s, *s11, *s22, *s33;
  int *row_ptr, *col;
  z2 *val;
  z2 iomu;
  std::vector<void *> allocs;
};

struct CsrArgs {
  int n1, n2, n3;
  long lEx, lEy, lE;
  const double *d1, *d2, *d3, *d1s, *d2s, *d3s, *s11, *s22, *s33;
  z2 iomu;
  const int *row_ptr;
  int *col;
  z2 *val;
};

static CsrArgs csr_args(const csem_csr *A)
{
  CsrArgs a;
  a.n1 = A->n1; a.n2 = A->n2; a.n3 = A->n3; a.lEx = A->lEx; a.lEy = A->lEy; a.lE = A->lE;
  a.d1 = A->d1; a.d2 = A->d2; a.d3 = A->d3; a.d1s = A->d1s; a.d2s = A->d2s; a.d3s = A->d3s;
  a.s11 = A->s11; a.s22 = A->s22; a.s33 = A->s33;
  a.iomu = A->iomu; a.row_ptr = A->row_ptr; a.col = A->col; a.val = A->val;
  return a;
}

// Enumerate row r's entries (column, value) in order; with col == NULL
// only count them.  Returns the count.
__device__ int csr_row(const CsrArgs &a, long r, int *col, z2 *val)
{
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  const long oy = a.lEx, oz = a.lEx + a.lEy;
  int e = 0;
  double aii = 0.0, t;
  auto put = [&](long c, double v) {
    if (col) { col[e] = (int)c; val[e] = zmk(v, 0.0); }
    e++;
  };
#define EX(k, j, i) ((long)(i) + (long)n1 * ((j) + (long)(n2 + 1) * (k)))
#define EY(k, j, i) (oy + (long)(i) + (long)(n1 + 1) * ((j) + (long)n2 * (k)))
#define EZ(k, j, i) (oz + (long)(i) + (long)(n1 + 1) * ((j) + (long)(n2 + 1) * (k)))
  int i, j, k;
  double sg;
  if (r < a.lEx) {
    i = (int)(r % n1); j = (int)((r / n1) % (n2 + 1)); k = (int)(r / ((long)n1 * (n2 + 1)));
    // Acsr.c:47-130 (counts), 305-428 (values)
    t = 1. / (a.d2[j] * a.d1s[i]);
    if (i + 1 <= n1 && j < n2) put(EY(k, j, i + 1), t);
    if (j < n2) put(EY(k, j, i), -t);
    t = 1. / (a.d2[j] * a.d2s[j]);
    if (j + 1 <= n2) put(EX(k, j + 1, i), -t);
    aii += t;
    t = 1. / (a.d2[j] * a.d1s[i]);
    if (i + 1 <= n1 && j >= 1) put(EY(k, j - 1, i + 1), -t);
    if (j >= 1) put(EY(k, j - 1, i), t);
    t = 1. / (a.d2[j] * a.d2s[j >= 1 ? j - 1 : 0]);
    if (j >= 1) put(EX(k, j - 1, i), -t);
    aii += t;
    t = 1. / (a.d3[k] * a.d3s[k]);
    if (k + 1 <= n3) put(EX(k + 1, j, i), -t);
    aii += t;
    t = 1. / (a.d3[k] * a.d1s[i]);
    if (i + 1 <= n1 && k < n3) put(EZ(k, j, i + 1), t);
    if (k < n3) put(EZ(k, j, i), -t);
    t = 1. / (a.d3[k] * a.d3s[k >= 1 ? k - 1 : 0]);
    aii += t;
    if (k >= 1) put(EX(k - 1, j, i), -t);
    t = 1. / (a.d3[k] * a.d1s[i]);
    if (i + 1 <= n1 && k >= 1) put(EZ(k - 1, j, i + 1), -t);
    if (k >= 1) put(EZ(k - 1, j, i), t);
    sg = a.s11[r];
  } else if (r < a.lEx + a.lEy) {
    const long q = r - a.lEx;
    i = (int)(q % (n1 + 1)); j = (int)((q / (n1 + 1)) % n2); k = (int)(q / ((long)(n1 + 1) * n2));
    // Acsr.c:436-556
    t = 1. / (a.d3[k] * a.d2s[j]);
    if (j + 1 <= n2 && k < n3) put(EZ(k, j + 1, i), t);
    if (k < n3) put(EZ(k, j, i), -t);
    t = 1. / (a.d3[k] * a.d3s[k]);
    if (k + 1 <= n3) put(EY(k + 1, j, i), -t);
    aii += t;
    t = 1. / (a.d3[k] * a.d2s[j]);
    if (j + 1 <= n2 && k >= 1) put(EZ(k - 1, j + 1, i), -t);
    if (k >= 1) put(EZ(k - 1, j, i), t);
    t = 1. / (a.d3[k] * a.d3s[k >= 1 ? k - 1 : 0]);
    aii += t;
    if (k >= 1) put(EY(k - 1, j, i), -t);
    t = 1. / (a.d1[i] * a.d1s[i]);
    if (i + 1 <= n1) put(EY(k, j, i + 1), -t);
    aii += t;
    t = 1. / (a.d1[i] * a.d2s[j]);
    if (i < n1 && j + 1 <= n2) put(EX(k, j + 1, i), t);
    if (i < n1) put(EX(k, j, i), -t);
    t = 1. / (a.d1[i] * a.d1s[i >= 1 ? i - 1 : 0]);
    aii += t;
    if (i >= 1) put(EY(k, j, i - 1), -t);
    t = 1. / (a.d1[i] * a.d2s[j]);
    if (i >= 1 && j + 1 <= n2) put(EX(k, j + 1, i - 1), -t);
    if (i >= 1) put(EX(k, j, i - 1), t);
    sg = a.s22[q];
  } else {
    const long q = r - a.lEx - a.lEy;
    i = (int)(q % (n1 + 1)); j = (int)((q / (n1 + 1)) % (n2 + 1)); k = (int)(q / ((long)(n1 + 1) * (n2 + 1)));
    // Acsr.c:562-683
    t = 1. / (a.d1[i] * a.d3s[k]);
    if (i < n1 && k + 1 <= n3) put(EX(k + 1, j, i), t);
    if (i < n1) put(EX(k, j, i), -t);
    t = 1. / (a.d1[i] * a.d1s[i]);
    if (i + 1 <= n1) put(EZ(k, j, i + 1), -t);
    aii += t;
    t = 1. / (a.d1[i] * a.d3s[k]);
    if (i >= 1 && k + 1 <= n3) put(EX(k + 1, j, i - 1), -t);
    if (i >= 1) put(EX(k, j, i - 1), t);
    t = 1. / (a.d1[i] * a.d1s[i >= 1 ? i - 1 : 0]);
    aii += t;
    if (i >= 1) put(EZ(k, j, i - 1), -t);
    t = 1. / (a.d2[j] * a.d2s[j]);
    if (j + 1 <= n2) put(EZ(k, j + 1, i), -t);
    aii += t;
    t = 1. / (a.d2[j] * a.d3s[k]);
    if (j < n2 && k + 1 <= n3) put(EY(k + 1, j, i), t);
    if (j < n2) put(EY(k, j, i), -t);
    t = 1. / (a.d2[j] * a.d2s[j >= 1 ? j - 1 : 0]);
    aii += t;
    if (j >= 1) put(EZ(k, j - 1, i), -t);
    t = 1. / (a.d2[j] * a.d3s[k]);
    if (j >= 1 && k + 1 <= n3) put(EY(k + 1, j - 1, i), -t);
    if (j >= 1) put(EY(k, j - 1, i), t);
    sg = a.s33[q];
  }
#undef EX
#undef EY
#undef EZ
  // the diagonal: aii - (i w mu0) sigma, C99 complex minus complex times real
  if (col) { col[e] = (int)r; val[e] = zmk(aii - a.iomu.x * sg, 0.0 - a.iomu.y * sg); }
  return e + 1;
}

__global__ void k_csr_count(CsrArgs a, int *cnt)
{
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r < a.lE) cnt[r] = csr_row(a, r, nullptr, nullptr);
  else if (r == a.lE) cnt[r] = 0;
}

__global__ void k_csr_fill(CsrArgs a)
{
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= a.lE) return;
  const int p = a.row_ptr[r];
  (void)csr_row(a, r, a.col + p, a.val + p);
}

// y = A x: the row's entries in order, y_r = 0 + a_r0 x_c0 + a_r1 x_c1 + ...
__global__ void k_csr_spmv(const int *row_ptr, const int *col, const z2 *val, long n, const z2 *x, z2 *y)
{
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= n) return;
  z2 s = zmk(0.0, 0.0);
  for (int e = row_ptr[r]; e < row_ptr[r + 1]; e++) s = zadd(s, zmul(val[e], x[col[e]]));
  y[r] = s;
}

// One wavefront level of a Gauss-Seidel sweep over component c: the rows
// (k, j, i) of c with u + v = L for its two coupled axes (Ex: u = j,
// v = k; Ey: u = i, v = k; Ez: u = i, v = j), each counted from the far
// end backward.  Thread t takes pair t / nw and w = t % nw along the free
// axis.
__global__ void k_csr_gs_level(CsrArgs a, const z2 *x, z2 *y, int c, int L, int bwd, int u0, int npair)
{
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  int nu, nv, nw;  // extents (largest index) of u, v and the count of the free axis
  if (c == 0) { nu = n2; nv = n3; nw = n1; }
  else if (c == 1) { nu = n1; nv = n3; nw = n2; }
  else { nu = n1; nv = n2; nw = n3; }
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)npair * nw) return;
  const int p = (int)(t / nw), w = (int)(t % nw);
  int u = u0 + p, v = L - u;
  if (bwd) { u = nu - u; v = nv - v; }
  long r;
  if (c == 0) r = (long)w + (long)n1 * (u + (long)(n2 + 1) * v);                       // Ex (k=v, j=u, i=w)
  else if (c == 1) r = a.lEx + u + (long)(n1 + 1) * (w + (long)n2 * v);              // Ey (k=v, j=w, i=u)
  else r = a.lEx + a.lEy + u + (long)(n1 + 1) * (v + (long)(n2 + 1) * w);           // Ez (k=w, j=v, i=u)
  z2 s = zmk(0.0, 0.0), aii = zmk(1.0, 0.0);
  for (int e = a.row_ptr[r]; e < a.row_ptr[r + 1]; e++) {
    const int cc = a.col[e];
    if (cc == r) aii = a.val[e];
    else s = zadd(s, zmul(a.val[e], y[cc]));
  }
  y[r] = zdiv(zsub(x[r], s), aii);
}

__global__ void k_csr_zero(long n, z2 *y)
{
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) y[e] = zmk(0.0, 0.0);
}

template <class T>
static int csr_alloc(csem_csr *A, T **p, long count)
{
  long bytes = count * (long)sizeof(T);
  if (bytes <= 0) bytes = 256;
  hipError_t e = dev_malloc((void **)p, bytes);
  if (e != hipSuccess) return csem_fail(-2, "hipMalloc(%ld): %s", bytes, hipGetErrorString(e));
  A->allocs.push_back((void *)*p);
  return 0;
}

extern "C" void csem_csr_destroy(csem_csr *A)
{
  if (!A) return;
  for (void *p : A->allocs) dev_free(p);
  delete A;
}

extern "C" int csem_csr_create(csem_csr **out, int n1, int n2, int n3, const double *d1, const double *d2,
                               const double *d3, const double *d1s, const double *d2s, const double *d3s,
                               const double *s11, const double *s22, const double *s33, double iomu_re,
                               double iomu_im)
{
  *out = NULL;
  if (n1 < 1 || n2 < 1 || n3 < 1) return csem_fail(-3, "grid %d x %d x %d", n1, n2, n3);
  csem_csr *A = new csem_csr();
  A->n1 = n1; A->n2 = n2; A->n3 = n3;
  A->lEx = (long)n1 * (n2 + 1) * (n3 + 1);
  A->lEy = (long)(n1 + 1) * n2 * (n3 + 1);
  A->lE = A->lEx + A->lEy + (long)(n1 + 1) * (n2 + 1) * n3;
  A->iomu = zmk(iomu_re, iomu_im);
  // at most 13 entries per row; the indices are 32-bit (as Acsr.c's)
  if (13 * A->lE >= (1L << 31)) { delete A; return csem_fail(-3, "grid too large for 32-bit CSR indices"); }
  int rc = 0;
  int *cnt = NULL;
  void *tmp = NULL;
  size_t tmp_bytes = 0;
  const long n = A->lE;
  // the build runs on a stream of its own and waits for that stream only:
  // a device-wide wait would invalidate another thread's graph capture
  hipStream_t bs = NULL;
  if (hipStreamCreateWithFlags(&bs, hipStreamNonBlocking) != hipSuccess) {
    delete A;
    return csem_fail(-1, "stream");
  }
  const unsigned nb = (unsigned)((n + 1 + 255) / 256);
  struct { double **dst; const double *src; long cnt; } up[] = {
      {&A->d1, d1, n1 + 1}, {&A->d2, d2, n2 + 1}, {&A->d3, d3, n3 + 1}, {&A->d1s, d1s, n1 + 1},
      {&A->d2s, d2s, n2 + 1}, {&A->d3s, d3s, n3 + 1}, {&A->s11, s11, A->lEx}, {&A->s22, s22, A->lEy},
      {&A->s33, s33, n - A->lEx - A->lEy}};
  for (auto &u : up) {
    if ((rc = csr_alloc(A, u.dst, u.cnt))) goto fail;
    if (dev_upload(*u.dst, u.src, u.cnt * sizeof(double)) != hipSuccess) { rc = csem_fail(-1, "upload"); goto fail; }
  }
  if ((rc = csr_alloc(A, &A->row_ptr, n + 1)) || (rc = csr_alloc(A, &cnt, n + 1))) goto fail;
  {
    CsrArgs a = csr_args(A);
    k_csr_count<<<nb, 256, 0, bs>>>(a, cnt);
    if (hipcub::DeviceScan::ExclusiveSum(NULL, tmp_bytes, cnt, A->row_ptr, (int)(n + 1), bs) != hipSuccess) {
      rc = csem_fail(-1, "scan size");
      goto fail;
    }
    if ((rc = csr_alloc(A, (char **)&tmp, (long)tmp_bytes))) goto fail;
    if (hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, A->row_ptr, (int)(n + 1), bs) != hipSuccess) {
      rc = csem_fail(-1, "scan");
      goto fail;
    }
    int total = 0;
    if (hipMemcpyAsync(&total, A->row_ptr + n, sizeof(int), hipMemcpyDeviceToHost, bs) != hipSuccess ||
        hipStreamSynchronize(bs) != hipSuccess) {
      rc = csem_fail(-1, "row count");
      goto fail;
    }
    A->nnz = total;
    if ((rc = csr_alloc(A, &A->col, A->nnz)) || (rc = csr_alloc(A, &A->val, A->nnz))) goto fail;
    a = csr_args(A);
    k_csr_fill<<<nb, 256, 0, bs>>>(a);
    if (hipStreamSynchronize(bs) != hipSuccess || hipGetLastError() != hipSuccess) {
      rc = csem_fail(-1, "CSR assembly failed");
      goto fail;
    }
  }
  (void)hipStreamDestroy(bs);
  *out = A;
  return 0;
fail:
  (void)hipStreamSynchronize(bs);
  (void)hipStreamDestroy(bs);
  csem_csr_destroy(A);
  return rc;
}

extern "C" long csem_csr_nnz(const csem_csr *A) { return A ? A->nnz : -1; }
extern "C" long csem_csr_rows(const csem_csr *A) { return A ? A->lE : -1; }

extern "C" int csem_csr_download(const csem_csr *A, int *row_ptr, int *col, void *val)
{
  if (!A) return csem_fail(-3, "no matrix");
  std::lock_guard<std::mutex> lk(g_devwide_mu);  // synchronous copies: not during another thread's capture
  if ((row_ptr && hipMemcpy(row_ptr, A->row_ptr, (A->lE + 1) * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) ||
      (col && hipMemcpy(col, A->col, A->nnz * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) ||
      (val && hipMemcpy(val, A->val, A->nnz * sizeof(z2), hipMemcpyDeviceToHost) != hipSuccess))
    return csem_fail(-1, "CSR download");
  return 0;
}

// y = A x on device vectors (ctx: the csem_csr; the model and frequency
// are not used: the matrix holds its own)
extern "C" int csem_op_csr(void *ctx, csem_model *, const csem_freq *, const void *x, void *y, void *stream)
{
  const csem_csr *A = (const csem_csr *)ctx;
  if (!A) return csem_fail(-3, "no matrix");
  k_csr_spmv<<<(unsigned)((A->lE + 255) / 256), 256, 0, (hipStream_t)stream>>>(A->row_ptr, A->col, A->val, A->lE,
                                                                               (const z2 *)x, (z2 *)y);
  HIP_OK(hipGetLastError());
  return 0;
}

// y = M^-1 x, the symmetric Gauss-Seidel of Acsr.c:719-748 on the matrix's rows
extern "C" int csem_op_csr_sgs(void *ctx, csem_model *, const csem_freq *, const void *x, void *y, void *stream)
{
  const csem_csr *A = (const csem_csr *)ctx;
  if (!A) return csem_fail(-3, "no matrix");
  const hipStream_t s = (hipStream_t)stream;
  const CsrArgs a = csr_args(A);
  // y = 0 by a fill kernel (a memset node is not safe in a caller's captured graph)
  k_csr_zero<<<(unsigned)((A->lE + 255) / 256), 256, 0, s>>>(A->lE, (z2 *)y);
  const int n1 = A->n1, n2 = A->n2, n3 = A->n3;
  // per component: the extents of its coupled axes (u, v) and free axis
  const int NU[3] = {n2, n1, n1}, NV[3] = {n3, n3, n2}, NW[3] = {n1, n2, n3};
  for (int pass = 0; pass < 2; pass++) {
    for (int q = 0; q < 3; q++) {
      const int c = pass == 0 ? q : 2 - q;
      const int nu = NU[c], nv = NV[c], nw = NW[c];
      for (int L = 0; L <= nu + nv; L++) {
        const int u0 = L > nv ? L - nv : 0, u1 = L < nu ? L : nu;
        const int npair = u1 - u0 + 1;
        const long nt = (long)npair * nw;
        k_csr_gs_level<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(a, (const z2 *)x, (z2 *)y, c, L, pass, u0, npair);
      }
    }
  }
  HIP_OK(hipGetLastError());
  return 0;
}
