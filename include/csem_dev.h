/* csem_dev.h -- device-level C-ABI of the MI355X solve path.
 *
 * Plain C types only: pointers, sizes, doubles.  Device buffers are
 * passed as `void *` device addresses; `stream` is a hipStream_t (NULL =
 * the null stream).  Complex data is interleaved (re, im) double pairs,
 * the layout of C99 `double _Complex` that the reference uses for its
 * vectors (emf.h:44-47), so the host entry points in csem.h hand their
 * buffers straight through.
 *
 * Vector layout (fdfd.c:13-21): E = [Ex | Ey | Ez], each x fastest,
 *   Ex(k,j,i) = i + n1*(j + (n2+1)*k)       k in [0,n3], j in [0,n2], i in [0,n1)
 *   Ey(k,j,i) = i + (n1+1)*(j + n2*k)       k in [0,n3], j in [0,n2), i in [0,n1]
 *   Ez(k,j,i) = i + (n1+1)*(j + (n2+1)*k)   k in [0,n3), j in [0,n2], i in [0,n1]
 * with n3 the operator depth (nz+nb under the airwave boundary, fdfd.c:49).
 *
 * All functions return 0 on success and a negative code on error; the
 * message is available from csem_last_error().
 */
#ifndef CSEM_DEV_H
#define CSEM_DEV_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct csem_model csem_model;

/* Everything the device needs to know about one extended model.  All
 * arrays are host arrays, copied at creation. */
typedef struct {
  int n1, n2, n3f;  /* extended grid: nx+2nb, ny+2nb, nz+2nb */
  int nb, airbc, fvm;
  const double *d1, *d2, *d3, *d1s, *d2s, *d3s; /* n+1 each, full grid (extend_model.c:8-21) */
  const double *sigma11, *sigma22, *sigma33;    /* edge conductivities, full grid (extend_model.c:180-252) */
  /* airwave boundary (airbc=1), airbc_init_free.c:24-109 */
  int nfft;                  /* transform size (square, power of two) */
  const double *sH1, *sH2;   /* nfft*nfft interleaved complex multipliers, [i2][i1] */
  const double *twiddle;     /* nfft interleaved complex, exp(-2 pi i t / nfft) */
  const int *jx, *jy;        /* forward interpolation cell per uniform index, -1 = outside */
  const double *wx, *wy;
  const int *gxj1, *gxj2;    /* ghost Hx: uniform cell per i1 in [0,n1), per i2 in [0,n2) */
  const double *gxw1, *gxw2;
  const int *gyj1, *gyj2;    /* ghost Hy: per i1 in [0,n1), per i2 in [0,n2) */
  const double *gyw1, *gyw2;
} csem_model_desc;

/* Per-frequency constants: iomu = I*2*pi*f*mu0 (main.c:95), iwmu = 1/iomu
 * (fdfd.c:100).  csem_freq_consts computes both exactly as the reference. */
typedef struct {
  double iomu_re, iomu_im, iwmu_re, iwmu_im;
} csem_freq;

const char *csem_last_error(void);
int csem_freq_consts(float freq, csem_freq *out);

int csem_model_create(const csem_model_desc *d, csem_model **out);
void csem_model_destroy(csem_model *m);

/* Device-side model setup (no counterpart in the reference; for servers
 * that build many models).  The raw model: core cell resistivities as
 * emf_init reads them (emf_init_free.c:120-150, float32, x fastest), the
 * air resistivity, and the airwave transform spacings airbc_init sets
 * (airbc_init_free.c:44-46).  csem_model_create_raw takes everything else
 * from d (sigma11..33 and sH1, sH2 are ignored) and builds on the device:
 * the extended cells (extend_model.c:108-175, quirk Q2 included) and the
 * edge-averaged conductivities (extend_model.c:180-252), bit for bit what
 * the host computes; and the airwave multipliers (airbc_init_free.c:69-84)
 * with the device's exp, sin and cos, within a few ulp of the host's.
 * csem_model_sigma / csem_model_multiplier copy a component (0, 1, 2 /
 * 0 = sH1, 1 = sH2) of any model to a host array of its full length. */
typedef struct {
  int nx, ny, nz;
  const float *rho11, *rho22, *rho33;
  double rho_air;
  double d1fft, d2fft, d3fft;
} csem_model_raw;
int csem_model_create_raw(const csem_model_desc *d, const csem_model_raw *raw, csem_model **out);
int csem_model_sigma(const csem_model *m, int which, double *host);
int csem_model_multiplier(const csem_model *m, int which, double *host);
/* operator vector length lE, LU-SGS (full grid) length, device bytes held */
long csem_model_len(const csem_model *m);
long csem_model_full_len(const csem_model *m);
long csem_model_bytes(const csem_model *m);

/* y = A x, fdfd.c:86-205 (fused: H is never stored). */
int csem_apply(csem_model *m, const csem_freq *f, const void *x, void *y, void *stream);
/* y = M^-1 x, LU-SGS, fdfd.c:287-510 */
int csem_lusgs(csem_model *m, const csem_freq *f, const void *x, void *y, void *stream);
/* H = curl E / (i w mu0) on faces, fdfd.c:125-149: Hx (n1+1)*n2*n3,
 * Hy n1*(n2+1)*n3, Hz n1*n2*(n3+1), stored contiguously Hx|Hy|Hz. */
int csem_curl_h(csem_model *m, const csem_freq *f, const void *x, void *h, void *stream);
/* ghost planes from Hz(z=0), airbc_init_free.c:130-209 */
int csem_airbc(csem_model *m, const void *hz0, void *hxg, void *hyg, void *stream);
/* 2-D transform of an nfft x nfft buffer in place, dir -1 forward (rows
 * then columns), +1 inverse unnormalised (columns then rows). */
int csem_fft2(csem_model *m, void *data, int dir, void *stream);
/* out[0..1] = sum a_i conj(b_i) in the fixed chunked order; out is a
 * device pointer to two doubles. */
int csem_dot(long n, const void *a, const void *b, void *out, void *stream);

/* csem_dot for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  No counterpart in the reference.
 * Same kernels and result as csem_dot, but the chunk partial sums live in
 * a fixed device array of 512 KiB per device (held until the process
 * ends) instead of a buffer allocated on first use; n above 2^26
 * (67,108,864) is rejected with an error.  Calls on one device share that
 * array: calls that overlap in time on one device (several streams or
 * host threads) race and must be ordered by the caller. */
int csem_dot_noalloc(long n, const void *a, const void *b, void *out, void *stream);

/* Summation order of every dot product (csem_dot and the solvers):
 * 0 (default) the fixed chunked parallel order; 1 the reference's serial
 * left-to-right order (cdotprod.c:18), one dependent add per element on a
 * single wavefront -- slow, for end-to-end bitwise comparison with the
 * reference. */
int csem_set_reduction(int serial);

/* Device-resident solves of A x = b.  x holds the initial guess on entry.
 * precond: 0 none, 1 LU-SGS.  Returns the iteration index at convergence
 * (as printed by "converged at iter=", solver.c:160) or niter, or < 0 on
 * error.  hist (host, may be NULL, niter+2 doubles) receives relres. */
int csem_bicgstab(csem_model *m, const csem_freq *f, void *x, const void *b, int niter,
                  double tol, int precond, double *hist, void *stream);
int csem_fgmres(csem_model *m, const csem_freq *f, void *x, const void *b, int nouter,
                double tol, int mrestart, int precond, double *hist, void *stream);
int csem_gmres(csem_model *m, const csem_freq *f, void *x, const void *b, int nouter,
               double tol, int mrestart, double *hist, void *stream);

/* ---- The operator as a CSR matrix (Acsr.c:22-748): a brute-force
 * cross-check of the matrix-free kernels.  Built on the device from host
 * grid arrays (the extended grid of extend_model_init: d1, d1s of n1 + 1
 * values, d2, d2s of n2 + 1, d3, d3s of n3 + 1, the edge conductivities in
 * the Ex / Ey / Ez layouts) at one frequency (i w mu0): iw mu0 times the
 * curl-curl operator, boundary rows with truncated stencils, each row's
 * entries in Acsr.c's order and its diagonal last.  Acsr.c sets each
 * row's pointer before storing the row's diagonal, which puts every
 * diagonal in the next row's range; here each row holds its own.
 * csem_op_csr (y = A x) and csem_op_csr_sgs (y = M^-1 x, the symmetric
 * Gauss-Seidel of Acsr.c:719-748 on the rows, in the serial order's
 * arithmetic) take the matrix as ctx. */
typedef struct csem_csr csem_csr;
int csem_csr_create(csem_csr **A, int n1, int n2, int n3, const double *d1, const double *d2, const double *d3,
                    const double *d1s, const double *d2s, const double *d3s, const double *s11,
                    const double *s22, const double *s33, double iomu_re, double iomu_im);
void csem_csr_destroy(csem_csr *A);
long csem_csr_nnz(const csem_csr *A);
long csem_csr_rows(const csem_csr *A);
/* host arrays of rows + 1, nnz and nnz complex doubles (any may be NULL) */
int csem_csr_download(const csem_csr *A, int *row_ptr, int *col, void *val);
int csem_op_csr(void *ctx, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream);
int csem_op_csr_sgs(void *ctx, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream);

/* ---- Operators as arguments.  The reference's Krylov drivers are generic
 * over op_t (solver.h:4): any y = Op x.  Here an operator acts on device
 * vectors (interleaved complex doubles) as op(ctx, model, freq, x, y,
 * stream) and returns 0 or a negative error code.  The library's own:
 *   csem_op_apply   y = A x            (fdfd_apply, fdfd.c:86-205)
 *   csem_op_lusgs   y = M^-1 x, LU-SGS (lusgs_apply, fdfd.c:287-510)
 *   csem_op_aphi    y = M^-1 x, A-phi  (aphi_apply, fdfd.c:513-821)
 *   csem_op_gmres   y = GMRES(m) x     (gmres_apply, gmres_init_free.c:58-137);
 *                   ctx is a csem_gmres_prec
 *   csem_op_host    y = fn(x) for a caller's host function fn of the
 *                   reference's op_t shape; ctx is a csem_host_op.  x is
 *                   copied to host memory, fn called, y copied back.
 * The device operators need the model; csem_op_host does not. */
typedef int (*csem_op)(void *ctx, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream);
int csem_op_apply(void *ctx, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream);
int csem_op_lusgs(void *ctx, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream);
int csem_op_aphi(void *ctx, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream);
int csem_op_gmres(void *ctx, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream);
int csem_op_host(void *ctx, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream);
typedef struct {
  void (*fn)(int n, void *x, void *y); /* host arrays of n interleaved complex doubles */
  int n;
  double *hx, *hy;                     /* host staging, 2n doubles each (allocated by the caller) */
} csem_host_op;
/* Inner GMRES(m) preconditioner on the operator (A, actx): a fixed m
 * Arnoldi steps from x = 0 (no convergence test), the least-squares
 * solution by Givens rotations and back substitution.  Device work
 * vectors are allocated on first use for vectors of length n and released
 * by csem_gmres_prec_release. */
typedef struct {
  int m;
  long n;
  csem_op A;
  void *actx;
  void *work; /* device, (m + 3) * n complex doubles + scalars; NULL until first use */
} csem_gmres_prec;
void csem_gmres_prec_release(csem_gmres_prec *p);
/* Krylov drivers over operators.  n is the vector length (csem_model_len
 * when m is given; m may be NULL when A and M are host operators).  M =
 * NULL: no preconditioner, i.e. bicgstab / gmres (solver.c:16-87,
 * 181-285); otherwise fbicgstab / fgmres (solver.c:93-178, 403-512).
 * Return values and hist as csem_bicgstab / csem_fgmres. */
int csem_bicgstab_op(csem_model *m, const csem_freq *f, long n, void *x, const void *b, int niter, double tol,
                     csem_op A, void *actx, csem_op M, void *mctx, double *hist, void *stream);
int csem_gmres_op(csem_model *m, const csem_freq *f, long n, void *x, const void *b, int nouter, double tol,
                  int mrestart, csem_op A, void *actx, csem_op M, void *mctx, double *hist, void *stream);

/* Sparse source term: b[idx[q]] += val[q] for q in order (val interleaved
 * complex, host arrays). */
int csem_scatter_add(long nnz, const long *idx, const double *val, void *b, void *stream);
/* Receiver interpolation: for each receiver r, out[r] = sum_{q<8}
 * (s[q] * x[idx[8r+q]]) * w1 * w2 * w3 in the order of
 * inject_extract.c:151-158, then times scale[r].  Host tables, device x,
 * host out (interleaved complex doubles). */
int csem_gather(int nrec, const long *idx, const double *wts, const double *sig,
                const double *scale, const void *x, double *out, void *stream);

/* Vector primitives the host-side GMRES uses (solver.c:423-499):
 * o = a - b; o = a / r (real r); o = a / (re + i im) (Smith); y -= h x
 * with h two doubles in device memory; y += (re + i im) x. */
int csem_vec_sub(long n, const void *a, const void *b, void *o, void *stream);
int csem_vec_divr(long n, const void *a, double r, void *o, void *stream);
int csem_vec_divz(long n, const void *a, double re, double im, void *o, void *stream);
int csem_vec_subscaled_dev(long n, void *y, const void *h, const void *x, void *stream);
int csem_vec_addscaled(long n, void *y, double re, double im, const void *x, void *stream);

/* ---- Batched mixed-precision solve (no counterpart in the reference: the
 * multi-right-hand-side, FP32-inner path for throughput callers).
 *
 * R (1..4) right-hand sides of one model, RHS r with frequency constants
 * f[r].  x and b are R device vectors of csem_model_len complex doubles,
 * back to back.  Solves A_r x_r = b_r from x = 0 to a true FP64 relative
 * residual |b - A x| / |b| < tol[r] by FP64 iterative refinement around an
 * FP32 LU-SGS + F-BiCGStab (or FGMRES) inner solve with FP64 dot products.  Returns
 * the number of right-hand sides that reached tol, or < 0 on error;
 * iters[r] receives the inner iterations spent, relres[r] the final true
 * relative residual (host arrays).  opt may be NULL for the defaults.
 * An F-BiCGStab inner solve whose recursive residual has made no new 1%
 * improvement in stall_iters iterations is ended early (FP32 stagnation);
 * its last iterate is the correction, or, if that does not lower the true
 * residual, the iterate of its best recursive residual, and the next
 * refinement step starts afresh from the new true residual; if neither
 * lowers it, that right-hand side goes on with FGMRES(restart) refinement
 * steps (below) from its best solution within the same budget.  With krylov = CSEM_KRYLOV_FGMRES each refinement
 * step is one FGMRES(restart) cycle from x = 0 on the scaled true
 * residual (classical Gram-Schmidt with one reorthogonalisation pass, the
 * Hessenberg matrix in FP64 on the device, one captured graph per cycle);
 * a cycle ends early once its residual estimate reaches the reduction the
 * step needs, and max_inner bounds the cycles' steps in total (max_outer
 * and check_every are not used).  Such a solve holds (2 restart + 1) R
 * more complex-float vectors than csem_mixed_slot_bytes states.  A refinement step that does not lower the true residual (the FP32 inner
 * solve stagnating or diverging on an ill-conditioned system) is undone
 * and refinement of that right-hand side stops: x keeps the best solution
 * found and relres[r] its residual, never above 1.  A caller that needs
 * tol then continues with the FP64 solver (csem_bicgstab) from that x,
 * asking it for tol[r] / relres[r]: its tolerance is relative to the
 * residual it starts from, as the reference's (solver.c:115, 158-159). */
typedef struct {
  int max_inner;     /* inner iteration budget per right-hand side (default 2000) */
  double inner_tol;  /* reduction asked of one inner solve (1e-2: FP32
                        BiCGStab stagnates when asked for much more) */
  int max_outer;     /* refinement steps (40) */
  int check_every;   /* inner iterations between convergence checks (8) */
  int use_graph;     /* replay a captured hipGraph of one inner iteration (1) */
  int precond;       /* FP32 preconditioner: CSEM_PRECOND_LUSGS, the reference's
                        lexicographic LU-SGS (fdfd.c:287-510), or
                        CSEM_PRECOND_RBSGS, the same symmetric Gauss-Seidel
                        splitting with the cells of each slice visited in
                        red-black order (fully parallel passes) */
  int stall_iters;   /* inner iterations without a new 1% residual
                        improvement that end an F-BiCGStab inner solve (0 =
                        the default, 100; < 0 = never) */
  int krylov;        /* inner method: CSEM_KRYLOV_FBICGSTAB (solver.c:93-178)
                        or CSEM_KRYLOV_FGMRES (solver.c:403-512) */
  int restart;       /* FGMRES restart length m (0 = 10, at most 16) */
} csem_mixed_opts;
#define CSEM_PRECOND_LUSGS 0
#define CSEM_PRECOND_RBSGS 1
#define CSEM_KRYLOV_FBICGSTAB 0
#define CSEM_KRYLOV_FGMRES 1
int csem_mixed_solve(csem_model *m, int R, const csem_freq *f, void *x, const void *b, const double *tol,
                     const csem_mixed_opts *opt, int *iters, double *relres, void *stream);
/* The FP32 operator and preconditioner on R complex-float vectors back to
 * back (length R * csem_model_len), for tests and custom drivers. */
int csem_apply32(csem_model *m, int R, const csem_freq *f, const void *x, void *y, void *stream);
int csem_lusgs32(csem_model *m, int R, const csem_freq *f, const void *x, void *y, void *stream);
/* Planes per tile of the FP64 operator (6, 8, 10, 12 or 16): the fastest on
 * the model, timed once when the model is created; every depth gives the
 * same values.  csem_model_set_apply_planes fixes another depth; it must
 * not run while another thread applies the model's operator. */
int csem_model_apply_planes(const csem_model *m);
int csem_model_set_apply_planes(csem_model *m, int planes);
/* Tile shape the autotuner picked for the FP32 operator at batch size R
 * (rows per block, planes per block); -1 before the first solve. */
int csem_mixed_tuning(csem_model *m, int R, int *rows, int *planes);
/* Fix the FP32 operator's tile at batch size R instead of timing the
 * candidates: output rows per block 4 or 6 (R <= 3), 2 or 4 (R = 4);
 * planes per block 8, 10, 12 or 16. */
int csem_mixed_set_tile(csem_model *m, int R, int rows, int planes);
/* The second part of the autotuner: the tile each of the six FP32 LU-SGS
 * pre-passes uses at batch size R (candidate 0, 1 or 2, in launch order Ex,
 * Ey, Ez forward, Ez, Ey, Ex backward) and the columns per block of the
 * FFT column pass; -1 before the first solve at that R.  A solve times the
 * candidates once per model and R; csem_mixed_set_lu_tiles fixes them
 * instead (cols a power of two with nfft * cols in [1024, 4096]).  No
 * choice changes a result. */
int csem_mixed_tuning_lu(csem_model *m, int R, int *pre, int *cols);
int csem_mixed_set_lu_tiles(csem_model *m, int R, const int *pre, int cols);
/* device bytes held by the mixed-precision state of a model (its FP32
 * copy and every solve slot), and the bytes one more slot for R
 * right-hand sides takes: what a scheduler admits solves against */
long csem_mixed_bytes(const csem_model *m);
long csem_mixed_slot_bytes(const csem_model *m, int R);
/* free and total device memory (hipMemGetInfo) */
int csem_mem_info(long *free_bytes, long *total_bytes);

/* Kernel timers: csem_kernel_timing(1) resets and starts HIP-event timing
 * of every hot-kernel launch (events recorded on the launch stream);
 * csem_kernel_stats(id, ...) returns the kernel's name, summed duration,
 * launch count and summed algorithmic bytes, and the number of kernel ids
 * (ids run 0 .. n-1), or -1 for an id out of range.  An event pair around
 * one launch also spans the dispatch and the event packets themselves
 * (several microseconds on top of the kernel).
 *
 * csem_kernel_repeat(id, reps, loops): while timing is on, each of the
 * next `loops` launches of kernel id is followed by `reps` more launches
 * of it with the same arguments, back to back between one event pair
 * (each id keeps its own setting until csem_kernel_timing resets them);
 * csem_kernel_loop_stats gives their summed time and launch count, so
 * time / count is the kernel's duration plus only the gap between two
 * dispatches.  Only kernels whose output does not feed their own input
 * may repeat (operators, airwave stages but the in-place inverse rows,
 * LU-SGS pre-passes, sweeps and copies, dot partials): a repeat leaves
 * every result unchanged.  Returns
 * 0, or -1 for any other id. */
int csem_kernel_timing(int on);
int csem_kernel_stats(int id, const char **name, double *total_ms, long *count, double *bytes);
int csem_kernel_repeat(int id, int reps, int loops);
int csem_kernel_loop_stats(int id, double *total_ms, long *count);

/* Device buffer helpers for hosts without their own allocator. */
int csem_malloc(void **p, long bytes);
int csem_free(void *p);
int csem_memcpy_h2d(void *dst, const void *src, long bytes, void *stream);
int csem_memcpy_d2h(void *dst, const void *src, long bytes, void *stream);
int csem_memset(void *p, int v, long bytes, void *stream);
/* dst = src on the device with a plain streaming kernel (16-byte aligned
 * buffers, bytes a multiple of 16): a copy of known size for scaling
 * hardware byte counters. */
int csem_copy(void *dst, const void *src, long bytes, void *stream);
int csem_sync(void *stream);

#ifdef __cplusplus
}
#endif
#endif
