/* csem_entry.c -- the reference's operator, preconditioner, airwave and
 * Krylov entry points, running on the MI355X device library.
 *
 * fdfd_init builds the device model from the extended model (one active
 * model per process, like the reference's static emf pointer, fdfd.c:23).
 * fdfd_apply / lusgs_apply / airbc_apply / cdotprod take host pointers, as
 * the reference's op_t calls do, and stage them through device buffers.
 * fbicgstab / bicgstab / fgmres / gmres recognise the device operators
 * (Aop == fdfd_apply, invMop == lusgs_apply) and then run the whole solve
 * device-resident: one upload of b and x, one download of x.
 * inject_source / extract_emf touch 8 grid values per source / receiver and
 * keep the reference's host-array interface.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/csem.h"

#define CSEM_PI 3.14159265358979323846264 /* cstd.h:35-37 */
#define CSEM_MU0 (4. * CSEM_PI * 1e-7)    /* cstd.h:38-40 */

void csem_set_status(int st);
/* records an error message for csem_last_error() and returns code (device library) */
int csem_fail(int code, const char *fmt, ...);

/* vector primitives exported by the device library */
int csem_vec_sub(long n, const void *a, const void *b, void *o, void *stream);
int csem_vec_divr(long n, const void *a, double r, void *o, void *stream);
int csem_vec_divz(long n, const void *a, double re, double im, void *o, void *stream);
int csem_vec_subscaled_dev(long n, void *y, const void *h, const void *x, void *stream);
int csem_vec_addscaled(long n, void *y, double re, double im, const void *x, void *stream);

static struct {
  emf_t *emf;
  csem_model *model;
  void *dx, *dy, *dh, *dg1, *dg2, *dred, *dhf;
  long nvec;
  void *px, *py; /* staging of gmres_apply / aphi_apply, apart from the operators' (a caller's
                    op_t inside the inner GMRES may call fdfd_apply, which stages through dx, dy) */
  long npvec;
  int last_iters;
  /* airwave host tables (airbc_init) */
  double *twiddle;
  int *jx, *jy, *gxj1, *gxj2, *gyj1, *gyj2;
  double *wx, *wy, *gxw1, *gxw2, *gyw1, *gyw2;
} G;

csem_model *csem_active_model(void) { return G.model; }
int csem_last_iterations(void) { return G.last_iters; }

static void die_dev(const char *what, int rc)
{
  csem_set_status(rc);
  err("%s failed (%d): %s", what, rc, csem_last_error());
}

#define DEV(call)                            \
  do {                                       \
    int rc_ = (call);                        \
    if (rc_) {                               \
      die_dev(#call, rc_);                   \
      return;                                \
    }                                        \
  } while (0)

static int is_pow2(int n) { return n >= 4 && (n & (n - 1)) == 0; }

/* airbc_init_free.c:24-109: transform size, uniform spacing, the
 * frequency-independent multipliers (when `multipliers`; the device setup
 * builds them itself) and the interpolation tables. */
static void airbc_setup(emf_t *emf, int multipliers)
{
  int N, i1, i2, j1, j2, n1 = emf->n1, n2 = emf->n2;
  double dkx, dky, kx, ky, kz, x1, x2, s1, s2;
  if (!getparint("n1fft", &emf->n1fft)) emf->n1fft = 1024;
  if (!getparint("n2fft", &emf->n2fft)) emf->n2fft = 1024;
  if (emf->n1fft != emf->n2fft || !is_pow2(emf->n1fft) || emf->n1fft > 1024)
    err("airwave transform must be square with a power-of-two size <= 1024 (n1fft=%d n2fft=%d)", emf->n1fft,
        emf->n2fft);
  N = emf->n1fft;
  emf->d1fft = (emf->x1[n1] - emf->x1[0]) / (N - 1);
  emf->d2fft = (emf->x2[n2] - emf->x2[0]) / (N - 1);
  emf->d3fft = emf->x3node[1] - emf->x3node[0];
  if (emf->verb) {
    printf("[n1fft, n2fft]=[%d, %d]\n", emf->n1fft, emf->n2fft);
    printf("[d1fft, d2fft, d3fft]=[%g, %g, %g]\n", emf->d1fft, emf->d2fft, emf->d3fft);
  }
  emf->sH1kxky = emf->sH2kxky = NULL;
  dkx = 2.0 * CSEM_PI / (emf->d1fft * N);
  dky = 2.0 * CSEM_PI / (emf->d2fft * N);
  if (multipliers) {
    emf->sH1kxky = alloc2complex(N, N);
    emf->sH2kxky = alloc2complex(N, N);
  }
  for (i2 = 0; multipliers && i2 < N; i2++) {
    ky = (i2 <= N / 2) ? i2 * dky : (i2 - N) * dky;
    for (i1 = 0; i1 < N; i1++) {
      kx = (i1 <= N / 2) ? i1 * dkx : (i1 - N) * dkx;
      kz = sqrt(kx * kx + ky * ky);
      /* upward continuation by half a cell and d/dx, d/dy of the
       * quasi-static air field, shifted half a cell for the staggering */
      emf->sH1kxky[i2][i1] = exp(-kz * 0.5 * emf->d3fft) * I * kx / (kz + 1.e-15);
      emf->sH2kxky[i2][i1] = exp(-kz * 0.5 * emf->d3fft) * I * ky / (kz + 1.e-15);
      emf->sH1kxky[i2][i1] *= cexp(-I * kx * 0.5 * emf->d1s[0]);
      emf->sH2kxky[i2][i1] *= cexp(-I * ky * 0.5 * emf->d2s[0]);
    }
  }
  G.twiddle = (double *)malloc(2 * N * sizeof(double));
  for (i1 = 0; i1 < N; i1++) {
    double a = -2.0 * CSEM_PI * (double)i1 / (double)N;
    G.twiddle[2 * i1] = cos(a);
    G.twiddle[2 * i1 + 1] = sin(a);
  }
  G.jx = alloc1int(N); G.jy = alloc1int(N); G.wx = alloc1double(N); G.wy = alloc1double(N);
  for (i1 = 0; i1 < N; i1++) {
    x1 = emf->x1s[0] + i1 * emf->d1fft;
    G.jx[i1] = -1;
    G.wx[i1] = 0;
    if (x1 < emf->x1s[n1 - 1]) {
      j1 = find_index(n1 + 1, emf->x1s, x1);
      G.jx[i1] = j1;
      G.wx[i1] = (x1 - emf->x1s[j1]) / (emf->x1s[j1 + 1] - emf->x1s[j1]);
    }
  }
  for (i2 = 0; i2 < N; i2++) {
    x2 = emf->x2s[0] + i2 * emf->d2fft;
    G.jy[i2] = -1;
    G.wy[i2] = 0;
    if (x2 < emf->x2s[n2 - 1]) {
      j2 = find_index(n2 + 1, emf->x2s, x2);
      G.jy[i2] = j2;
      G.wy[i2] = (x2 - emf->x2s[j2]) / (emf->x2s[j2 + 1] - emf->x2s[j2]);
    }
  }
  /* back-interpolation cells (airbc_init_free.c:168-175, 195-202) */
  G.gxj1 = alloc1int(n1); G.gxw1 = alloc1double(n1); G.gxj2 = alloc1int(n2); G.gxw2 = alloc1double(n2);
  G.gyj1 = alloc1int(n1); G.gyw1 = alloc1double(n1); G.gyj2 = alloc1int(n2); G.gyw2 = alloc1double(n2);
  for (i2 = 0; i2 < n2; i2++) {
    s2 = (emf->x2s[i2] - emf->x2s[0]);
    G.gxj2[i2] = s2 / emf->d2fft;
    G.gxw2[i2] = s2 / emf->d2fft - G.gxj2[i2];
    s2 = (emf->x2[i2] - emf->x2[0]);
    G.gyj2[i2] = s2 / emf->d2fft;
    G.gyw2[i2] = s2 / emf->d2fft - G.gyj2[i2];
  }
  for (i1 = 0; i1 < n1; i1++) {
    s1 = (emf->x1[i1] - emf->x1[0]);
    G.gxj1[i1] = s1 / emf->d1fft;
    G.gxw1[i1] = s1 / emf->d1fft - G.gxj1[i1];
    s1 = (emf->x1s[i1] - emf->x1s[0]);
    G.gyj1[i1] = s1 / emf->d1fft;
    G.gyw1[i1] = s1 / emf->d1fft - G.gyj1[i1];
  }
}

void airbc_init(emf_t *emf) { airbc_setup(emf, 1); }

void airbc_free(emf_t *emf)
{
  if (emf->sH1kxky) free2complex(emf->sH1kxky);
  if (emf->sH2kxky) free2complex(emf->sH2kxky);
  emf->sH1kxky = emf->sH2kxky = NULL;
  free(G.twiddle);
  free1int(G.jx); free1int(G.jy); free1double(G.wx); free1double(G.wy);
  free1int(G.gxj1); free1int(G.gxj2); free1int(G.gyj1); free1int(G.gyj2);
  free1double(G.gxw1); free1double(G.gxw2); free1double(G.gyw1); free1double(G.gyw2);
  G.twiddle = NULL;
}

static void ensure_staging(long n)
{
  if (G.nvec >= n) return;
  if (G.dx) csem_free(G.dx);
  if (G.dy) csem_free(G.dy);
  G.dx = G.dy = NULL;
  if (csem_malloc(&G.dx, n * 16) || csem_malloc(&G.dy, n * 16)) {
    die_dev("staging allocation", -2);
    return;
  }
  G.nvec = n;
}

/* Host half of fdfd_init (fdfd.c:30-63): operator truncation under the
 * airwave boundary, the airwave tables (with the multipliers when asked)
 * and the vector lengths. */
static void fdfd_prepare(emf_t *emf, int multipliers)
{
  G.emf = emf;
  if (emf->airbc) {
    emf->xx = emf->yy = NULL; /* LU-SGS scratch lives on the device */
    emf->n3 = emf->nz + emf->nb;
    airbc_setup(emf, multipliers);
  }
  emf->lEx = emf->n1 * (emf->n2 + 1) * (emf->n3 + 1);
  emf->lEy = (emf->n1 + 1) * emf->n2 * (emf->n3 + 1);
  emf->lEz = (emf->n1 + 1) * (emf->n2 + 1) * emf->n3;
  emf->lE = emf->lEx + emf->lEy + emf->lEz;
  emf->lHx = (emf->n1 + 1) * emf->n2 * emf->n3;
  emf->lHy = emf->n1 * (emf->n2 + 1) * emf->n2; /* fdfd.c:59 sizes the k extent with n2 */
  emf->lHz = emf->n1 * emf->n2 * (emf->n3 + 1);
  emf->lH = emf->lHx + emf->lHy + emf->lHz;
  emf->lv = (emf->n1 + 1) * (emf->n2 + 1) * (emf->n3 + 1);
}

void csem_fdfd_prepare(emf_t *emf) { fdfd_prepare(emf, 1); }

/* The device model's description from a prepared emf_t (without the
 * conductivities and multipliers when `arrays` is 0). */
static void model_desc(const emf_t *emf, csem_model_desc *dd, int arrays)
{
  csem_model_desc d;
  memset(&d, 0, sizeof d);
  d.n1 = emf->n1; d.n2 = emf->n2; d.n3f = emf->nz + 2 * emf->nb; d.nb = emf->nb;
  d.airbc = emf->airbc; d.fvm = emf->fvm;
  d.d1 = emf->d1; d.d2 = emf->d2; d.d3 = emf->d3; d.d1s = emf->d1s; d.d2s = emf->d2s; d.d3s = emf->d3s;
  if (arrays) {
    d.sigma11 = &emf->sigma11[0][0][0];
    d.sigma22 = &emf->sigma22[0][0][0];
    d.sigma33 = &emf->sigma33[0][0][0];
  }
  if (emf->airbc) {
    d.nfft = emf->n1fft;
    if (arrays) {
      d.sH1 = (const double *)&emf->sH1kxky[0][0];
      d.sH2 = (const double *)&emf->sH2kxky[0][0];
    }
    d.twiddle = G.twiddle;
    d.jx = G.jx; d.jy = G.jy; d.wx = G.wx; d.wy = G.wy;
    d.gxj1 = G.gxj1; d.gxj2 = G.gxj2; d.gxw1 = G.gxw1; d.gxw2 = G.gxw2;
    d.gyj1 = G.gyj1; d.gyj2 = G.gyj2; d.gyw1 = G.gyw1; d.gyw2 = G.gyw2;
  }
  *dd = d;
}

int csem_model_build_device(acq_t *acq, emf_t *emf, csem_model **out)
{
  csem_model_desc d;
  csem_model_raw r;
  int rc, c;
  *out = NULL;
  csem_extend_grid(acq, emf);
  const int n1 = emf->n1, n2 = emf->n2, n3f = emf->nz + 2 * emf->nb;
  fdfd_prepare(emf, 0);
  model_desc(emf, &d, 0);
  memset(&r, 0, sizeof r);
  r.nx = emf->nx; r.ny = emf->ny; r.nz = emf->nz;
  r.rho11 = &emf->rho11[0][0][0]; r.rho22 = &emf->rho22[0][0][0]; r.rho33 = &emf->rho33[0][0][0];
  r.rho_air = emf->rho_air;
  r.d1fft = emf->d1fft; r.d2fft = emf->d2fft; r.d3fft = emf->d3fft;
  if ((rc = csem_model_create_raw(&d, &r, out))) return rc;
  emf->sigma11 = alloc3double(n1, n2 + 1, n3f + 1);
  emf->sigma22 = alloc3double(n1 + 1, n2, n3f + 1);
  emf->sigma33 = alloc3double(n1 + 1, n2 + 1, n3f);
  double *dst[3] = {&emf->sigma11[0][0][0], &emf->sigma22[0][0][0], &emf->sigma33[0][0][0]};
  for (c = 0; c < 3; c++)
    if ((rc = csem_model_sigma(*out, c, dst[c]))) {
      csem_model_destroy(*out);
      *out = NULL;
      return rc;
    }
  return 0;
}

/* Device half: the model the device library solves on, built from the
 * extended model and the airwave tables csem_fdfd_prepare made. */
int csem_model_from_emf(const emf_t *emf, csem_model **out)
{
  csem_model_desc d;
  model_desc(emf, &d, 1);
  return csem_model_create(&d, out);
}

/* fdfd.c:30-63 */
void fdfd_init(emf_t *emf)
{
  csem_fdfd_prepare(emf);
  if (G.model) csem_model_destroy(G.model);
  G.model = NULL;
  DEV(csem_model_from_emf(emf, &G.model));
  csem_set_status(0);
}

void fdfd_free(emf_t *emf)
{
  if (emf->airbc) airbc_free(emf);
  if (G.model) csem_model_destroy(G.model);
  G.model = NULL;
  G.emf = NULL;
  if (G.dx) csem_free(G.dx);
  if (G.dy) csem_free(G.dy);
  if (G.dh) csem_free(G.dh);
  if (G.dg1) csem_free(G.dg1);
  if (G.dg2) csem_free(G.dg2);
  if (G.dred) csem_free(G.dred);
  if (G.dhf) csem_free(G.dhf);
  if (G.px) csem_free(G.px);
  if (G.py) csem_free(G.py);
  G.dx = G.dy = G.dh = G.dg1 = G.dg2 = G.dred = G.dhf = G.px = G.py = NULL;
  G.nvec = G.npvec = 0;
}

/* iomu = emf->I_omega_mu0 as the caller set it (main.c:95), iwmu = 1/iomu
 * by C complex division as fdfd.c:100 forms it. */
static csem_freq freq_of(const emf_t *emf)
{
  csem_freq f;
  csem_complex iwmu = 1. / emf->I_omega_mu0;
  f.iomu_re = creal(emf->I_omega_mu0);
  f.iomu_im = cimag(emf->I_omega_mu0);
  f.iwmu_re = creal(iwmu);
  f.iwmu_im = cimag(iwmu);
  return f;
}

static int need_model(const char *who)
{
  if (!G.model || !G.emf) {
    csem_set_status(-4);
    err("%s called before fdfd_init", who);
    return 0;
  }
  return 1;
}

/* fdfd.c:117-149: fdfd_apply leaves H = curl E / (i w mu0) of its input in
 * emf->xh, Hx at 0, Hy at lHx, Hz at lHx + lHy.  lHy is the reference's
 * size (fdfd.c:59, k extent n2): when n3 > n2 the reference's Hy runs into
 * Hz and past the array; here the Hy tail beyond lHy is dropped. */
static int store_h(const csem_freq *f)
{
  emf_t *e = G.emf;
  const long hx = e->lHx, hy = (long)e->n1 * (e->n2 + 1) * e->n3, hz = e->lHz;
  const long hyc = hy < e->lHy ? hy : e->lHy;
  int rc = G.dhf ? 0 : csem_malloc(&G.dhf, (hx + hy + hz) * 16);
  if (!rc) rc = csem_curl_h(G.model, f, G.dx, G.dhf, NULL);
  if (!rc) rc = csem_memcpy_d2h(e->xh, G.dhf, (hx + hyc) * 16, NULL);
  if (!rc) rc = csem_memcpy_d2h(e->xh + hx + e->lHy, (char *)G.dhf + (hx + hy) * 16, hz * 16, NULL);
  if (rc) die_dev("fdfd_apply: H to emf->xh", rc);
  return rc;
}

void fdfd_apply(int n, csem_complex *x, csem_complex *y)
{
  if (!need_model("fdfd_apply")) return;
  csem_freq f = freq_of(G.emf);
  ensure_staging(n);
  DEV(csem_memcpy_h2d(G.dx, x, (long)n * 16, NULL));
  DEV(csem_apply(G.model, &f, G.dx, G.dy, NULL));
  DEV(csem_memcpy_d2h(y, G.dy, (long)n * 16, NULL));
  if (G.emf->xh && store_h(&f)) return;
  csem_set_status(0);
}

void lusgs_apply(int n, csem_complex *x, csem_complex *y)
{
  if (!need_model("lusgs_apply")) return;
  csem_freq f = freq_of(G.emf);
  ensure_staging(n);
  DEV(csem_memcpy_h2d(G.dx, x, (long)n * 16, NULL));
  DEV(csem_lusgs(G.model, &f, G.dx, G.dy, NULL));
  DEV(csem_memcpy_d2h(y, G.dy, (long)n * 16, NULL));
  csem_set_status(0);
}

void airbc_apply(emf_t *emf, csem_complex *H3k, csem_complex *H1km1, csem_complex *H2km1)
{
  long nz = (long)emf->n1 * emf->n2, nx = (long)(emf->n1 + 1) * emf->n2, ny = (long)emf->n1 * (emf->n2 + 1);
  if (!need_model("airbc_apply")) return;
  if (!G.dh) DEV(csem_malloc(&G.dh, nz * 16));
  if (!G.dg1) DEV(csem_malloc(&G.dg1, nx * 16));
  if (!G.dg2) DEV(csem_malloc(&G.dg2, ny * 16));
  DEV(csem_memcpy_h2d(G.dh, H3k, nz * 16, NULL));
  DEV(csem_airbc(G.model, G.dh, G.dg1, G.dg2, NULL));
  DEV(csem_memcpy_d2h(H1km1, G.dg1, nx * 16, NULL));
  DEV(csem_memcpy_d2h(H2km1, G.dg2, ny * 16, NULL));
  csem_set_status(0);
}

/* sum a_i conj(b_i); the device sums in a fixed chunked order, so the
 * last bits can differ from a serial left-to-right sum */
csem_complex cdotprod(int n, csem_complex *a, csem_complex *b)
{
  double out[2] = {0, 0};
  ensure_staging(n);
  if (!G.dred && csem_malloc(&G.dred, 256)) {
    die_dev("cdotprod", -2);
    return 0;
  }
  if (csem_memcpy_h2d(G.dx, a, (long)n * 16, NULL) || csem_memcpy_h2d(G.dy, b, (long)n * 16, NULL) ||
      csem_dot(n, G.dx, G.dy, G.dred, NULL) || csem_memcpy_d2h(out, G.dred, 16, NULL)) {
    die_dev("cdotprod", -1);
    return 0;
  }
  csem_set_status(0);
  return out[0] + I * out[1];
}

/* iterate_*.txt and stdout lines of solver.c:40-54,127-141 */
static void log_history(const char *file, const char *tag, const double *hist, int nrow, int conv_iter)
{
  FILE *fp = fopen(file, "w");
  int it;
  if (fp) {
    fprintf(fp, "iteration \t relres\n");
    for (it = 0; it < nrow; it++) fprintf(fp, "%d \t %e\n", it, hist[it]);
    fclose(fp);
  }
  for (it = 0; it < nrow; it++) printf("%s iter=%d relres=%e\n", tag, it, hist[it]);
  if (conv_iter >= 0) printf("converged at iter=%d\n", conv_iter);
}

/* Restarted (flexible) GMRES with the vectors on the device and the small
 * Hessenberg / Givens recurrence on the host, in the arithmetic of
 * solver.c:403-512 (fgmres) and solver.c:181-285 (gmres).  A and M are
 * called through their pointers (M = NULL: Z = V). */
static int gmres_core(csem_model *model, const csem_freq *fq, long n, void *x, const void *b, int nouter,
                      double tol, int m, csem_op A, void *actx, csem_op M, void *mctx, int flexible, double *hist)
{
  const long nb = n * 16;
  const csem_freq f = *fq;
  void *w = NULL, *r = NULL, *V = NULL, *Z = NULL, *hd = NULL;
  csem_complex *h = (csem_complex *)calloc((size_t)(m + 1) * (m > 0 ? m : 1), sizeof(csem_complex));
  csem_complex *g = (csem_complex *)calloc(m + 2, sizeof(csem_complex));
  csem_complex *y = (csem_complex *)calloc(m + 2, sizeof(csem_complex));
  csem_complex *c = (csem_complex *)calloc(m + 2, sizeof(csem_complex));
  double *s = (double *)calloc(m + 2, sizeof(double));
  double dv[2 * 64 + 2];
  double beta, tmp, r0 = 1, relres;
  csem_complex ss;
  int i, j, iter, mm, conv = 0, steps = 0, rc = 0, mcur = m, early = 0;
  if (m > 64 || m < 1) {
    rc = -3;
    goto out;
  }
#define CK(call)       \
  if ((rc = (call))) { \
    goto out;          \
  }
  CK(csem_malloc(&w, nb)); CK(csem_malloc(&r, nb));
  CK(csem_malloc(&V, nb * (m + 1)));
  if (M) CK(csem_malloc(&Z, nb * m));
  CK(csem_malloc(&hd, (long)(2 * 64 + 2) * sizeof(double)));
#define VJ(k) ((char *)V + (long)(k) * nb)
#define ZJ(k) (M ? (char *)Z + (long)(k) * nb : VJ(k))
/* storage element (column j, row i) of the Hessenberg: fgmres keeps
 * h[j][i] (solver.c:412), gmres h[i][j] (solver.c:190) -- same numbers */
#define HH(row, col) h[(long)(col) * (m + 1) + (row)]
  for (iter = 0; iter < nouter; iter++) {
    CK(A(actx, model, &f, x, w, NULL));
    CK(csem_vec_sub(n, b, w, r, NULL));
    CK(csem_dot(n, r, r, hd, NULL));
    CK(csem_memcpy_d2h(dv, hd, 16, NULL));
    beta = sqrt(dv[0]);
    if (beta == 0.) break;
    CK(csem_vec_divr(n, r, beta, VJ(0), NULL));
    memset(g, 0, (m + 1) * sizeof(csem_complex));
    g[0] = beta;
    memset(h, 0, (size_t)(m + 1) * m * sizeof(csem_complex));
    if (iter == 0) {
      r0 = beta;
      if (hist) hist[0] = flexible ? 1.0 : beta / r0;
    }
    for (j = 0; j < mcur; j++) {
      if (M) CK(M(mctx, model, &f, VJ(j), ZJ(j), NULL));
      CK(A(actx, model, &f, ZJ(j), w, NULL));
      for (i = 0; i <= j; i++) { /* modified Gram-Schmidt, solver.c:445-448 */
        CK(csem_dot(n, w, VJ(i), (char *)hd + 16 * i, NULL));
        CK(csem_vec_subscaled_dev(n, w, (char *)hd + 16 * i, VJ(i), NULL));
      }
      CK(csem_dot(n, w, w, (char *)hd + 16 * (j + 1), NULL));
      CK(csem_memcpy_d2h(dv, hd, 16 * (j + 2), NULL));
      for (i = 0; i <= j; i++) HH(i, j) = dv[2 * i] + I * dv[2 * i + 1];
      HH(j + 1, j) = sqrt(dv[2 * (j + 1)]);
      if (cabs(HH(j + 1, j)) == 0.) {
        if (flexible) conv = 1;
        else mcur = j + 1;
        break;
      }
      CK(csem_vec_divz(n, w, creal(HH(j + 1, j)), cimag(HH(j + 1, j)), VJ(j + 1), NULL));
      for (i = 0; i < j; i++) {
        ss = conj(c[i]) * HH(i, j) + s[i] * HH(i + 1, j);
        HH(i + 1, j) = -s[i] * HH(i, j) + c[i] * HH(i + 1, j);
        HH(i, j) = ss;
      }
      tmp = sqrt(creal(HH(j, j) * conj(HH(j, j)) + HH(j + 1, j) * conj(HH(j + 1, j))));
      s[j] = creal(HH(j + 1, j)) / tmp;
      c[j] = HH(j, j) / tmp;
      if (flexible) {
        g[j + 1] = -s[j] * g[j];
        g[j] = conj(c[j]) * g[j];
        HH(j, j) = conj(c[j]) * HH(j, j) + s[j] * HH(j + 1, j);
      } else {
        HH(j, j) = conj(c[j]) * HH(j, j) + s[j] * HH(j + 1, j);
        HH(j + 1, j) = 0.;
        g[j + 1] = -s[j] * g[j];
        g[j] = conj(c[j]) * g[j];
      }
      relres = cabs(g[j + 1]) / r0;
      steps = iter * mcur + j + 1;
      if (hist) hist[steps] = relres;
      if (relres < tol) {
        conv = 1;
        if (!flexible) early = 1; /* solver.c:251 returns before updating x */
        break;
      }
    }
    if (early) break;
    mm = flexible ? (mcur < j + 1 ? mcur : j + 1) : mcur;
    y[mm - 1] = g[mm - 1] / HH(mm - 1, mm - 1);
    for (i = mm - 2; i >= 0; i--) {
      y[i] = g[i];
      for (j = i + 1; j < mm; j++) y[i] -= HH(i, j) * y[j];
      y[i] /= HH(i, i);
    }
    for (j = 0; j < mm; j++) CK(csem_vec_addscaled(n, x, creal(y[j]), cimag(y[j]), ZJ(j), NULL));
    if (conv && flexible) break;
  }
  rc = steps;
out:
#undef CK
#undef VJ
#undef ZJ
#undef HH
  if (w) csem_free(w);
  if (r) csem_free(r);
  if (V) csem_free(V);
  if (Z) csem_free(Z);
  if (hd) csem_free(hd);
  free(h); free(g); free(y); free(c); free(s);
  return rc;
}

int csem_fgmres(csem_model *m, const csem_freq *f, void *x, const void *b, int nouter, double tol, int mrestart,
                int precond, double *hist, void *stream)
{
  (void)stream;
  int rc = gmres_core(m, f, csem_model_len(m), x, b, nouter, tol, mrestart, csem_op_apply, NULL,
                      precond ? csem_op_lusgs : NULL, NULL, 1, hist);
  if (rc == -3) return csem_fail(-3, "restart length %d outside [1, 64]", mrestart);
  return rc;
}

int csem_gmres(csem_model *m, const csem_freq *f, void *x, const void *b, int nouter, double tol, int mrestart,
               double *hist, void *stream)
{
  (void)stream;
  int rc = gmres_core(m, f, csem_model_len(m), x, b, nouter, tol, mrestart, csem_op_apply, NULL, NULL, NULL, 0,
                      hist);
  if (rc == -3) return csem_fail(-3, "restart length %d outside [1, 64]", mrestart);
  return rc;
}

int csem_gmres_op(csem_model *m, const csem_freq *f, long n, void *x, const void *b, int nouter, double tol,
                  int mrestart, csem_op A, void *actx, csem_op M, void *mctx, double *hist, void *stream)
{
  (void)stream;
  if (!A) return csem_fail(-3, "csem_gmres_op: no operator");
  if (m && n != csem_model_len(m))
    return csem_fail(-3, "csem_gmres_op: n=%ld, the model's vectors hold %ld", n, csem_model_len(m));
  int rc = gmres_core(m, f, n, x, b, nouter, tol, mrestart, A, actx, M, mctx, M != NULL, hist);
  if (rc == -3) return csem_fail(-3, "restart length %d outside [1, 64]", mrestart);
  return rc;
}

/* ------------------------------------------------ operators as arguments */

/* A caller's host operator on device vectors: download x, call, upload y.
 * The host entry status is cleared first so that an error the operator
 * reports through csem_last_status() is seen here. */
int csem_op_host(void *ctx, csem_model *m, const csem_freq *f, const void *x, void *y, void *stream)
{
  csem_host_op *h = (csem_host_op *)ctx;
  const long bytes = (long)h->n * 16;
  int rc;
  (void)m; (void)f;
  if ((rc = csem_memcpy_d2h(h->hx, x, bytes, stream))) return rc;
  csem_set_status(0);
  h->fn(h->n, h->hx, h->hy);
  if ((rc = csem_last_status()) < 0) return rc;
  return csem_memcpy_h2d(y, h->hy, bytes, stream);
}

/* gmres_apply (gmres_init_free.c:58-137): y = GMRES(m) solution of A y = x
 * from y = 0 after exactly m Arnoldi steps, modified Gram-Schmidt, complex
 * Givens rotations (c complex, s real), back substitution, y = V_m z.  The
 * reference applies no breakdown test and no residual test here. */
int csem_op_gmres(void *ctx, csem_model *model, const csem_freq *f, const void *bv, void *xv, void *stream)
{
  csem_gmres_prec *P = (csem_gmres_prec *)ctx;
  const int m = P->m;
  const long n = P->n, nb = n * 16;
  csem_complex h[65 * 64], g[66], y[66], c[66], ss;
  double s[66], dv[2 * 64 + 2], beta, tmp;
  int i, j, rc;
  (void)stream;
  if (m < 1 || m > 64) return csem_fail(-3, "gmres_apply: inner dimension %d outside [1, 64]", m);
  if (!P->work && (rc = csem_malloc(&P->work, nb * (m + 2) + (long)(2 * 64 + 2) * sizeof(double)))) return rc;
  char *V = (char *)P->work, *w = V + nb * (m + 1), *hd = w + nb;
#define VJ(k) (V + (long)(k) * nb)
#define HH(row, col) h[(long)(col) * (m + 1) + (row)]
  if ((rc = csem_memset(xv, 0, nb, NULL))) return rc;
  if ((rc = csem_dot(n, bv, bv, hd, NULL)) || (rc = csem_memcpy_d2h(dv, hd, 16, NULL))) return rc;
  beta = sqrt(dv[0]); /* r = b (x = 0), gmres_init_free.c:80-82 */
  if (beta == 0.0) return 0;
  if ((rc = csem_vec_divr(n, bv, beta, VJ(0), NULL))) return rc;
  memset(g, 0, sizeof g);
  g[0] = beta;
  memset(h, 0, sizeof h);
  for (j = 0; j < m; j++) {
    if ((rc = P->A(P->actx, model, f, VJ(j), w, NULL))) return rc;
    for (i = 0; i <= j; i++) {
      if ((rc = csem_dot(n, w, VJ(i), hd + 16 * i, NULL))) return rc;
      if ((rc = csem_vec_subscaled_dev(n, w, hd + 16 * i, VJ(i), NULL))) return rc;
    }
    if ((rc = csem_dot(n, w, w, hd + 16 * (j + 1), NULL))) return rc;
    if ((rc = csem_memcpy_d2h(dv, hd, 16 * (j + 2), NULL))) return rc;
    for (i = 0; i <= j; i++) HH(i, j) = dv[2 * i] + I * dv[2 * i + 1];
    HH(j + 1, j) = sqrt(dv[2 * (j + 1)]);
    if ((rc = csem_vec_divz(n, w, creal(HH(j + 1, j)), cimag(HH(j + 1, j)), VJ(j + 1), NULL))) return rc;
    for (i = 0; i < j; i++) {
      ss = conj(c[i]) * HH(i, j) + s[i] * HH(i + 1, j);
      HH(i + 1, j) = -s[i] * HH(i, j) + c[i] * HH(i + 1, j);
      HH(i, j) = ss;
    }
    tmp = sqrt(creal(HH(j, j) * conj(HH(j, j)) + HH(j + 1, j) * conj(HH(j + 1, j))));
    s[j] = creal(HH(j + 1, j)) / tmp;
    c[j] = HH(j, j) / tmp;
    HH(j, j) = conj(c[j]) * HH(j, j) + s[j] * HH(j + 1, j);
    HH(j + 1, j) = 0.;
    g[j + 1] = -s[j] * g[j];
    g[j] = conj(c[j]) * g[j];
  }
  y[m - 1] = g[m - 1] / HH(m - 1, m - 1);
  for (i = m - 2; i >= 0; i--) {
    y[i] = g[i];
    for (j = i + 1; j < m; j++) y[i] -= HH(i, j) * y[j];
    y[i] /= HH(i, i);
  }
  for (j = 0; j < m; j++)
    if ((rc = csem_vec_addscaled(n, xv, creal(y[j]), cimag(y[j]), VJ(j), NULL))) return rc;
#undef VJ
#undef HH
  return 0;
}

void csem_gmres_prec_release(csem_gmres_prec *p)
{
  if (p && p->work) csem_free(p->work);
  if (p) p->work = NULL;
}

/* An op_t of the reference's shape, as an operator on device vectors: the
 * library's own entry points map to their device kernels, any other
 * function is called on host copies. */
typedef struct {
  csem_op op;
  void *ctx;
  csem_host_op host;
} opmap_t;

static struct {
  csem_gmres_prec prec;
  opmap_t inner;
  int on;
} GP; /* gmres_init / gmres_free state (gmres_init_free.c:27) */

static csem_csr *GA; /* Acsr_init / Acsr_free state (Acsr.c:19) */

static int map_op(op_t fn, int n, opmap_t *o)
{
  memset(o, 0, sizeof *o);
  if (!fn) return 0;
  if (fn == fdfd_apply) o->op = csem_op_apply;
  else if (fn == lusgs_apply) o->op = csem_op_lusgs;
  else if (fn == aphi_apply) o->op = csem_op_aphi;
  else if (fn == Acsr_apply || fn == Acsr_lusgs_apply) {
    if (!GA) return csem_fail(-4, "Acsr_apply used before Acsr_init");
    o->op = fn == Acsr_apply ? csem_op_csr : csem_op_csr_sgs;
    o->ctx = GA;
  }
  else if (fn == gmres_apply) {
    if (!GP.on) return csem_fail(-4, "gmres_apply used before gmres_init");
    o->op = csem_op_gmres;
    o->ctx = &GP.prec;
  } else {
    o->op = csem_op_host;
    o->host.fn = (void (*)(int, void *, void *))fn;
    o->host.n = n;
    o->host.hx = (double *)malloc((size_t)n * 16);
    o->host.hy = (double *)malloc((size_t)n * 16);
    if (!o->host.hx || !o->host.hy) return csem_fail(-2, "host staging of %d values", n);
    o->ctx = &o->host;
  }
  return 0;
}

static void unmap_op(opmap_t *o)
{
  free(o->host.hx);
  free(o->host.hy);
  o->host.hx = o->host.hy = NULL;
}

/* does the operator run on the active model's device kernels */
static int op_on_model(const opmap_t *o)
{
  if (o->op == csem_op_apply || o->op == csem_op_lusgs || o->op == csem_op_aphi) return 1;
  if (o->op == csem_op_gmres) return op_on_model(&GP.inner);
  return 0;
}

enum { K_BICGSTAB, K_FBICGSTAB, K_GMRES, K_FGMRES };

/* bicgstab / fbicgstab / gmres / fgmres (solver.c:16-512): x and b are
 * host arrays; the solve runs on the device with one upload of each and
 * one download of x.  Operators on the active model run as its kernels. */
static void krylov(const char *who, int kind, int n, csem_complex *xh, csem_complex *bh, op_t Aop, op_t invMop,
                   int niter, double tol, int m, int verb)
{
  opmap_t A, M;
  void *x = NULL, *b = NULL;
  const int gm = kind == K_GMRES || kind == K_FGMRES;
  const int nh = gm ? niter * m + 2 : niter + 2;
  double *hist = (double *)calloc(nh > 2 ? nh : 2, sizeof(double));
  csem_model *model = NULL;
  csem_freq f;
  int it, i, rc;
  memset(&f, 0, sizeof f);
  memset(&A, 0, sizeof A);
  memset(&M, 0, sizeof M);
  if (!Aop || ((kind == K_FBICGSTAB || kind == K_FGMRES) && !invMop)) {
    csem_set_status(-3);
    err("%s: no %s operator", who, Aop ? "preconditioner" : "system");
    goto out;
  }
  if ((rc = map_op(Aop, n, &A)) || (rc = map_op(kind == K_FBICGSTAB || kind == K_FGMRES ? invMop : NULL, n, &M))) {
    die_dev(who, rc);
    goto out;
  }
  if (op_on_model(&A) || (M.op && op_on_model(&M))) {
    if (!need_model(who)) goto out;
    if (n != csem_model_len(G.model)) {
      csem_set_status(-3);
      err("%s: n=%d but the operator's vectors hold %ld values", who, n, csem_model_len(G.model));
      goto out;
    }
    model = G.model;
  }
  if (G.emf) f = freq_of(G.emf);
  if ((rc = csem_malloc(&x, (long)n * 16)) || (rc = csem_malloc(&b, (long)n * 16)) ||
      (rc = csem_memcpy_h2d(x, xh, (long)n * 16, NULL)) || (rc = csem_memcpy_h2d(b, bh, (long)n * 16, NULL))) {
    die_dev("solver staging", rc);
    goto out;
  }
  if (gm)
    it = csem_gmres_op(model, &f, n, x, b, niter, tol, m, A.op, A.ctx, M.op, M.ctx, hist, NULL);
  else
    it = csem_bicgstab_op(model, &f, n, x, b, niter, tol, A.op, A.ctx, M.op, M.ctx, hist, NULL);
  if (it < 0) {
    die_dev(who, it);
    goto out;
  }
  if ((rc = csem_memcpy_d2h(xh, x, (long)n * 16, NULL))) {
    die_dev("solver download", rc);
    goto out;
  }
  G.last_iters = it;
  if (verb && !gm) {
    log_history(kind == K_FBICGSTAB ? "iterate_fbicgstab.txt" : "iterate_bicgstab.txt",
                kind == K_FBICGSTAB ? "f-bicgstab" : "bicgstab", hist, (it < niter ? it + 1 : niter),
                it < niter ? it : -1);
  } else if (verb) {
    const char *file = kind == K_FGMRES ? "iterate_fgmres.txt" : "iterate_gmres.txt";
    FILE *fp = fopen(file, "w");
    if (fp) {
      fprintf(fp, "iteration \t error\n");
      for (i = 0; i <= it; i++) fprintf(fp, "%d \t %e\n", i, hist[i]);
      fclose(fp);
    }
    for (i = 0; i <= it; i++) printf("%s iter=%d relres=%e\n", kind == K_FGMRES ? "f-gmres" : "gmres", i, hist[i]);
  }
  csem_set_status(0);
out:
  if (x) csem_free(x);
  if (b) csem_free(b);
  unmap_op(&A);
  unmap_op(&M);
  free(hist);
}

void fbicgstab(int n, csem_complex *x, csem_complex *b, op_t Aop, op_t invMop, int niter, double tol, int verb)
{
  krylov("fbicgstab", K_FBICGSTAB, n, x, b, Aop, invMop, niter, tol, 0, verb);
}

void bicgstab(int n, csem_complex *x, csem_complex *b, op_t Aop, int niter, double tol, int verb)
{
  krylov("bicgstab", K_BICGSTAB, n, x, b, Aop, NULL, niter, tol, 0, verb);
}

void fgmres(int n, csem_complex *x, csem_complex *b, op_t Aop, op_t invMop, int niter, double tol, int m, int verb)
{
  krylov("fgmres", K_FGMRES, n, x, b, Aop, invMop, niter, tol, m, verb);
}

void gmres(int n, csem_complex *x, csem_complex *b, op_t Aop, int niter, double tol, int m, int verb)
{
  krylov("gmres", K_GMRES, n, x, b, Aop, NULL, niter, tol, m, verb);
}

/* gmres_init_free.c:29-44: the inner-GMRES preconditioner (preco=2) on
 * Aop with m Arnoldi steps per application */
void gmres_init(int n, int m, op_t Aop)
{
  int rc;
  if (GP.on) gmres_free();
  memset(&GP, 0, sizeof GP);
  if ((rc = map_op(Aop, n, &GP.inner))) {
    die_dev("gmres_init", rc);
    return;
  }
  if (m < 1 || m > 64) {
    unmap_op(&GP.inner);
    csem_set_status(-3);
    err("gmres_init: inner dimension mp=%d outside [1, 64]", m);
    return;
  }
  GP.prec.m = m;
  GP.prec.n = n;
  GP.prec.A = GP.inner.op;
  GP.prec.actx = GP.inner.ctx;
  GP.on = 1;
  csem_set_status(0);
}

void gmres_free(void)
{
  if (!GP.on) return;
  csem_gmres_prec_release(&GP.prec);
  unmap_op(&GP.inner);
  memset(&GP, 0, sizeof GP);
}

/* op_t entry points that are preconditioners: host arrays in and out */
static void apply_staged(const char *who, csem_op op, void *ctx, int n, csem_complex *x, csem_complex *y,
                         int on_model)
{
  csem_freq f;
  int rc;
  memset(&f, 0, sizeof f);
  if (on_model) {
    if (!need_model(who)) return;
    if (n != csem_model_len(G.model)) {
      csem_set_status(-3);
      err("%s: n=%d but the operator's vectors hold %ld values", who, n, csem_model_len(G.model));
      return;
    }
  }
  if (G.emf) f = freq_of(G.emf);
  if (G.npvec < n) {
    if (G.px) csem_free(G.px);
    if (G.py) csem_free(G.py);
    G.px = G.py = NULL;
    G.npvec = 0;
    if ((rc = csem_malloc(&G.px, (long)n * 16)) || (rc = csem_malloc(&G.py, (long)n * 16))) {
      die_dev(who, rc);
      return;
    }
    G.npvec = n;
  }
  if ((rc = csem_memcpy_h2d(G.px, x, (long)n * 16, NULL)) ||
      (rc = op(ctx, on_model ? G.model : NULL, &f, G.px, G.py, NULL)) ||
      (rc = csem_memcpy_d2h(y, G.py, (long)n * 16, NULL))) {
    die_dev(who, rc);
    return;
  }
  csem_set_status(0);
}

/* gmres_init_free.c:58: gmres_apply(n, b, x) -- b in, x out */
void gmres_apply(int n, csem_complex *b, csem_complex *x)
{
  if (!GP.on) {
    csem_set_status(-4);
    err("gmres_apply called before gmres_init");
    return;
  }
  apply_staged("gmres_apply", csem_op_gmres, &GP.prec, n, b, x, op_on_model(&GP.inner));
}

/* fdfd.c:513-821 */
void aphi_apply(int n, csem_complex *x, csem_complex *y)
{
  apply_staged("aphi_apply", csem_op_aphi, NULL, n, x, y, 1);
}

/* The matrix-storage path of main.c's main2 (Acsr.c): the operator as a
 * CSR matrix on the device (csem_csr_create), a brute-force cross-check of
 * fdfd_apply / lusgs_apply.  Acsr_apply and Acsr_lusgs_apply are
 * csem_op_csr / csem_op_csr_sgs on host copies; passed to the Krylov
 * drivers they run on the device vectors.  The MUMPS direct solve of
 * main3 (Acoo.c) is not provided: each call reports an error. */
void Acsr_init(emf_t *emf, int ifreq)
{
  int rc;
  if (!emf || !emf->freqs || ifreq < 0 || ifreq >= emf->nfreq || !emf->d1 || !emf->d3s || !emf->sigma11 ||
      !emf->sigma22 || !emf->sigma33 || emf->n1 < 1 || emf->n2 < 1 || emf->n3 < 1) {
    csem_set_status(-3);
    err("Acsr_init: needs emf_init's frequencies and extend_model_init's grid and conductivities");
    return;
  }
  /* Acsr.c:34-39: the frequency's i w mu0 and the vector lengths */
  emf->I_omega_mu0 = I * 2. * CSEM_PI * emf->freqs[ifreq] * CSEM_MU0;
  emf->lEx = emf->n1 * (emf->n2 + 1) * (emf->n3 + 1);
  emf->lEy = (emf->n1 + 1) * emf->n2 * (emf->n3 + 1);
  emf->lEz = (emf->n1 + 1) * (emf->n2 + 1) * emf->n3;
  emf->lE = emf->lEx + emf->lEy + emf->lEz;
  if (GA) csem_csr_destroy(GA);
  GA = NULL;
  rc = csem_csr_create(&GA, emf->n1, emf->n2, emf->n3, emf->d1, emf->d2, emf->d3, emf->d1s, emf->d2s, emf->d3s,
                       &emf->sigma11[0][0][0], &emf->sigma22[0][0][0], &emf->sigma33[0][0][0],
                       creal(emf->I_omega_mu0), cimag(emf->I_omega_mu0));
  if (rc < 0) {
    die_dev("Acsr_init", rc);
    return;
  }
  if (emf->verb) printf("#nnz=%ld\n", csem_csr_nnz(GA));
  csem_set_status(0);
}

void Acsr_free(void)
{
  csem_csr_destroy(GA);
  GA = NULL;
}

static int acsr_ready(const char *who, int n)
{
  if (!GA) {
    csem_set_status(-4);
    err("%s called before Acsr_init", who);
    return 0;
  }
  if (n != csem_csr_rows(GA)) {
    csem_set_status(-3);
    err("%s: n=%d but the matrix has %ld rows", who, n, csem_csr_rows(GA));
    return 0;
  }
  return 1;
}

/* Acsr.c:704-716 */
void Acsr_apply(int n, csem_complex *x, csem_complex *y)
{
  if (acsr_ready("Acsr_apply", n)) apply_staged("Acsr_apply", csem_op_csr, GA, n, x, y, 0);
}

/* Acsr.c:719-748 */
void Acsr_lusgs_apply(int n, csem_complex *x, csem_complex *y)
{
  if (acsr_ready("Acsr_lusgs_apply", n)) apply_staged("Acsr_lusgs_apply", csem_op_csr_sgs, GA, n, x, y, 0);
}

static void not_provided(const char *who)
{
  csem_set_status(-5);
  err("%s: the MUMPS direct solve (Acoo.c) is not provided by this library", who);
}
void Acoo_init(emf_t *emf, int ifreq) { (void)emf; (void)ifreq; not_provided("Acoo_init"); }
void Acoo_free(void) { not_provided("Acoo_free"); }
void Acoo_solve(emf_t *emf) { (void)emf; not_provided("Acoo_solve"); }

/* ---------------------------------------------------------------- sources */

#define I1(k, j, i) ((i) + (long)emf->n1 * ((j) + (long)(emf->n2 + 1) * (k)))
#define I2(k, j, i) ((i) + (long)(emf->n1 + 1) * ((j) + (long)emf->n2 * (k)))
#define I3(k, j, i) ((i) + (long)(emf->n1 + 1) * ((j) + (long)(emf->n2 + 1) * (k)))

/* inject_extract.c:24-120.  With the airwave boundary the plane index is
 * shifted by nb before the conductivity lookup as well (the reference
 * reads sigma nb planes above the source cell, inject_extract.c:55-63). */
void inject_source(acq_t *acq, emf_t *emf, csem_complex *b)
{
  const int n1 = emf->nx + 2 * emf->nb, n2 = emf->ny + 2 * emf->nb, n3 = emf->nz + 2 * emf->nb;
  int ic, is, i, j, k, q;
  double w1, w2, w3, vol;
  csem_complex sv, *bx = b, *by = b + emf->lEx, *bz = b + emf->lEx + emf->lEy;
  for (ic = 0; ic < emf->nchsrc; ic++) {
    int ch = !strcmp(emf->chsrc[ic], "Ex") ? 0 : !strcmp(emf->chsrc[ic], "Ey") ? 1 : !strcmp(emf->chsrc[ic], "Ez") ? 2 : -1;
    if (ch < 0) continue;
    for (is = 0; is < acq->nsrc; is++) {
      const double *xa = ch == 0 ? emf->x1s : emf->x1, *ya = ch == 1 ? emf->x2s : emf->x2,
                   *za = ch == 2 ? emf->x3s : emf->x3;
      i = find_index(ch == 0 ? n1 : n1 + 1, (double *)xa, acq->src_x1[is]);
      j = find_index(ch == 1 ? n2 : n2 + 1, (double *)ya, acq->src_x2[is]);
      k = find_index(ch == 2 ? n3 : n3 + 1, (double *)za, acq->src_x3[is]);
      w1 = (acq->src_x1[is] - xa[i]) / (xa[i + 1] - xa[i]);
      w2 = (acq->src_x2[is] - ya[j]) / (ya[j + 1] - ya[j]);
      w3 = (acq->src_x3[is] - za[k]) / (za[k + 1] - za[k]);
      vol = (xa[i + 1] - xa[i]) * (ya[j + 1] - ya[j]) * (za[k + 1] - za[k]);
      sv = 1. / vol;
      if (emf->fvm) sv = emf->I_omega_mu0;
      if (ch != 2) sv /= emf->rho_water;
      if (emf->airbc) k -= emf->nb;
      for (q = 0; q < 8; q++) {
        int di = q & 1, dj = (q >> 1) & 1, dk = q >> 2;
        double f1 = di ? w1 : (1. - w1), f2 = dj ? w2 : (1. - w2), f3 = dk ? w3 : (1. - w3);
        if (ch == 0) bx[I1(k + dk, j + dj, i + di)] += sv / emf->sigma11[k + dk][j + dj][i + di] * f1 * f2 * f3;
        else if (ch == 1) by[I2(k + dk, j + dj, i + di)] += sv / emf->sigma22[k + dk][j + dj][i + di] * f1 * f2 * f3;
        else bz[I3(k + dk, j + dj, i + di)] += sv * f1 * f2 * f3;
      }
    }
  }
}

/* H = curl E / (i w mu0) on face (k, j, i) of the operator grid, component
 * c (0 Hx, 1 Hy, 2 Hz), with the expressions of fdfd.c:131-146; the
 * vertical spacings are those of operator plane k (fdfd.c:106-109). */
static csem_complex h_face(const emf_t *emf, const csem_complex *x, int c, int k, int j, int i, csem_complex iwmu)
{
  const csem_complex *Ex = x, *Ey = x + emf->lEx, *Ez = x + emf->lEx + emf->lEy;
  const double *d1s = emf->d1s, *d2s = emf->d2s, *d3s = emf->d3s + (emf->airbc ? emf->nb : 0);
  csem_complex t1, t2;
  if (c == 0) {
    t1 = (Ez[I3(k, j + 1, i)] - Ez[I3(k, j, i)]) / d2s[j];
    t2 = (Ey[I2(k + 1, j, i)] - Ey[I2(k, j, i)]) / d3s[k];
  } else if (c == 1) {
    t1 = (Ex[I1(k + 1, j, i)] - Ex[I1(k, j, i)]) / d3s[k];
    t2 = (Ez[I3(k, j, i + 1)] - Ez[I3(k, j, i)]) / d1s[i];
  } else {
    t1 = (Ey[I2(k, j, i + 1)] - Ey[I2(k, j, i)]) / d1s[i];
    t2 = (Ex[I1(k, j + 1, i)] - Ex[I1(k, j, i)]) / d2s[j];
  }
  return (t1 - t2) * iwmu;
}

/* Hx, Hy, Hz at a receiver (no counterpart in the reference, whose H
 * channels stay zero, inject_extract.c:123-213): H on the faces around the
 * receiver as fdfd_apply forms it (Hx at (x1, x2s, x3s), Hy at (x1s, x2,
 * x3s), Hz at (x1s, x2s, x3)), interpolated trilinearly in the order the E
 * channels use.  NaN when the eight faces are not all inside the operator
 * grid (a receiver within half a cell of its top or bottom). */
static csem_complex extract_h(const emf_t *emf, const csem_complex *x, int c, double rx, double ry, double rz)
{
  const int n1 = emf->nx + 2 * emf->nb, n2 = emf->ny + 2 * emf->nb, n3 = emf->nz + 2 * emf->nb;
  const double *xa = c == 0 ? emf->x1 : emf->x1s, *ya = c == 1 ? emf->x2 : emf->x2s, *za = c == 2 ? emf->x3 : emf->x3s;
  const int i = find_index(c == 0 ? n1 + 1 : n1, (double *)xa, rx);
  const int j = find_index(c == 1 ? n2 + 1 : n2, (double *)ya, ry);
  int k = find_index(c == 2 ? n3 + 1 : n3, (double *)za, rz);
  const double w1 = (rx - xa[i]) / (xa[i + 1] - xa[i]), w2 = (ry - ya[j]) / (ya[j + 1] - ya[j]),
               w3 = (rz - za[k]) / (za[k + 1] - za[k]);
  const csem_complex iwmu = 1. / emf->I_omega_mu0;
  csem_complex s = 0;
  int q;
  if (emf->airbc) k -= emf->nb;
  /* faces k and k+1: Hx, Hy exist on planes 0..n3-1, Hz on 0..n3 */
  if (k < 0 || k + 1 > (c == 2 ? emf->n3 : emf->n3 - 1)) return NAN;
  for (q = 0; q < 8; q++) {
    int di = q & 1, dj = (q >> 1) & 1, dk = q >> 2;
    double f1 = di ? w1 : (1. - w1), f2 = dj ? w2 : (1. - w2), f3 = dk ? w3 : (1. - w3);
    s += h_face(emf, x, c, k + dk, j + dj, i + di, iwmu) * f1 * f2 * f3;
  }
  return s;
}

/* inject_extract.c:123-213: trilinear Ex, Ey; for Ez the vertical current
 * sigma33*Ez is interpolated (at the unshifted plane index) and scaled by
 * rho_water.  Hx, Hy, Hz (beyond the reference, which leaves them zero)
 * come from extract_h. */
/* Receiver value of channel ch (0..5: Ex, Ey, Ez, Hx, Hy, Hz) at receiver
 * ir, in double precision (before the float32 store of
 * inject_extract.c:159,182,208). */
static csem_complex rec_value(const acq_t *acq, const emf_t *emf, const csem_complex *x, int ch, int ir)
{
  const int n1 = emf->nx + 2 * emf->nb, n2 = emf->ny + 2 * emf->nb, n3 = emf->nz + 2 * emf->nb;
  const csem_complex *Ex = x, *Ey = x + emf->lEx, *Ez = x + emf->lEx + emf->lEy;
  int i, j, k, kk, q;
  double w1, w2, w3;
  csem_complex s;
  if (ch >= 3) return extract_h(emf, x, ch - 3, acq->rec_x1[ir], acq->rec_x2[ir], acq->rec_x3[ir]);
  const double *xa = ch == 0 ? emf->x1s : emf->x1, *ya = ch == 1 ? emf->x2s : emf->x2,
               *za = ch == 2 ? emf->x3s : emf->x3;
  i = find_index(ch == 0 ? n1 : n1 + 1, (double *)xa, acq->rec_x1[ir]);
  j = find_index(ch == 1 ? n2 : n2 + 1, (double *)ya, acq->rec_x2[ir]);
  k = find_index(ch == 2 ? n3 : n3 + 1, (double *)za, acq->rec_x3[ir]);
  w1 = (acq->rec_x1[ir] - xa[i]) / (xa[i + 1] - xa[i]);
  w2 = (acq->rec_x2[ir] - ya[j]) / (ya[j + 1] - ya[j]);
  w3 = (acq->rec_x3[ir] - za[k]) / (za[k + 1] - za[k]);
  kk = k;
  if (emf->airbc) k -= emf->nb;
  s = 0;
  for (q = 0; q < 8; q++) {
    int di = q & 1, dj = (q >> 1) & 1, dk = q >> 2;
    double f1 = di ? w1 : (1. - w1), f2 = dj ? w2 : (1. - w2), f3 = dk ? w3 : (1. - w3);
    if (ch == 0) s += Ex[I1(k + dk, j + dj, i + di)] * f1 * f2 * f3;
    else if (ch == 1) s += Ey[I2(k + dk, j + dj, i + di)] * f1 * f2 * f3;
    else s += emf->sigma33[kk + dk][j + dj][i + di] * Ez[I3(k + dk, j + dj, i + di)] * f1 * f2 * f3;
  }
  return (ch == 2) ? s * emf->rho_water : s;
}

static int channel_of(const char *cn)
{
  return !strcmp(cn, "Ex") ? 0 : !strcmp(cn, "Ey") ? 1 : !strcmp(cn, "Ez") ? 2
       : !strcmp(cn, "Hx") ? 3 : !strcmp(cn, "Hy") ? 4 : !strcmp(cn, "Hz") ? 5 : -1;
}

void extract_emf_double(acq_t *acq, emf_t *emf, csem_complex *x, csem_complex *out)
{
  int ic, ir, outside;
  for (ic = 0; ic < emf->nchrec; ic++) {
    const int ch = channel_of(emf->chrec[ic]);
    outside = 0;
    for (ir = 0; ir < acq->nrec; ir++) {
      const csem_complex s = ch < 0 ? 0 : rec_value(acq, emf, x, ch, ir);
      outside += ch >= 3 && isnan(creal(s));
      out[(long)ic * acq->nrec + ir] = s;
    }
    if (outside)
      fprintf(stderr, "extract_emf: %d receiver(s) too close to the grid's top or bottom for %s (NaN)\n", outside,
              emf->chrec[ic]);
  }
}

void extract_emf(acq_t *acq, emf_t *emf, csem_complex *x, int ifreq)
{
  int ic, ir;
  csem_complex *v = (csem_complex *)malloc((size_t)emf->nchrec * acq->nrec * sizeof(csem_complex));
  if (!v) {
    csem_set_status(-2);
    err("extract_emf: out of host memory");
    return;
  }
  extract_emf_double(acq, emf, x, v);
  for (ic = 0; ic < emf->nchrec; ic++) {
    if (channel_of(emf->chrec[ic]) < 0) continue;
    for (ir = 0; ir < acq->nrec; ir++) emf->dcal_fd[ic][ifreq][ir] = v[(long)ic * acq->nrec + ir];
  }
  free(v);
}
