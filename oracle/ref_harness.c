/* ref_harness.c -- drives the reference's own routines, compiled from the
 * reference sources (see Makefile target `ref`), to produce golden data for
 * the oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * Usage: ref_harness <reference key=value flags> hmode=<mode> [hfreq=f]
 *        [hin=file] [hout=prefix]
 * Modes: grid, apply, lusgs, aphi, gmresprec, inject, extract, solve, acsr.  Binary outputs are
 * raw little-endian doubles / complex128 / complex64 named <prefix>_<what>.
 * airbc=1 is accepted for lusgs/aphi/inject/extract (the index shift only);
 * apply and solve need airbc=0 because airbc_apply lives in the file this
 * build cannot compile (no FFTW). */
#include "cstd.h"
#include "emf.h"
#include "acq.h"
#include "solver.h"

int iproc = 0, nproc = 1, ierr = 0; /* defined by main.c, which is not built */

void emf_init(emf_t *emf);
void acq_init(acq_t *acq, emf_t *emf);
void extend_model_init(acq_t *acq, emf_t *emf);
void fdfd_init(emf_t *emf);
void fdfd_apply(int n, complex *x, complex *y);
void lusgs_apply(int n, complex *x, complex *y);
void aphi_apply(int n, complex *x, complex *y);
void gmres_init(int n, int m, op_t Aop);
void gmres_free();
void gmres_apply(int n, complex *b, complex *x);
void inject_source(acq_t *acq, emf_t *emf, complex *b);
void extract_emf(acq_t *acq, emf_t *emf, complex *x, int ifreq);
void write_data(acq_t *acq, emf_t *emf, char *fname, float _Complex ***dcal_fd);
void Acsr_init(emf_t *emf, int ifreq);
void Acsr_free();
void Acsr_apply(int n, complex *x, complex *y);
void Acsr_lusgs_apply(int n, complex *x, complex *y);
/* Acsr.c:12-19: the matrix it assembles (a global of that file) */
typedef struct {
  int nrow, ncol, nnz;
  int *row_ptr, *col_ind;
  complex *val;
} ref_ccsr_t;
extern ref_ccsr_t *Acsr;
void bicgstab(int n, complex *x, complex *b, op_t Aop, int niter, double tol, int verb);
void gmres(int n, complex *x, complex *b, op_t Aop, int niter, double tol, int m, int verb);
void fbicgstab(int n, complex *x, complex *b, op_t Aop, op_t invMop, int niter, double tol, int verb);
void fgmres(int n, complex *x, complex *b, op_t Aop, op_t invMop, int niter, double tol, int m, int verb);

static char *prefix = "ref";

static void dump(const char *what, const void *p, size_t bytes)
{
  char fn[1024];
  FILE *f;
  snprintf(fn, sizeof fn, "%s_%s", prefix, what);
  f = fopen(fn, "wb");
  if (!f) err("cannot write %s", fn);
  fwrite(p, 1, bytes, f);
  fclose(f);
}

static void load(const char *fn, void *p, size_t bytes)
{
  FILE *f = fopen(fn, "rb");
  if (!f || fread(p, 1, bytes, f) != bytes) err("cannot read %s", fn);
  fclose(f);
}

int main(int argc, char **argv)
{
  emf_t *emf = calloc(1, sizeof(emf_t));
  acq_t *acq = calloc(1, sizeof(acq_t));
  char *mode = "grid", *fin = NULL;
  float f = 1.0f;
  int airbc, ifreq, ic, irec;
  complex *x, *y;

  initargs(argc, argv);
  getparstring("hmode", &mode);
  getparstring("hout", &prefix);
  getparstring("hin", &fin);
  getparfloat("hfreq", &f);
  /* main.c:62-69 */
  if (!getparint("mode", &emf->mode)) emf->mode = 0;
  if (!getparint("verb", &emf->verb)) emf->verb = 0;
  if (!getparint("niter", &emf->niter)) emf->niter = 1000;
  if (!getparint("preco", &emf->preco)) emf->preco = 1;
  if (!getparint("algopt", &emf->algopt)) emf->algopt = 1;
  if (!getparfloat("tol", &emf->tol)) emf->tol = 1e-8;
  if (!getparint("nrestart", &emf->nrestart)) emf->nrestart = 10;
  if (!getparint("mp", &emf->mp)) emf->mp = 30;

  emf_init(emf);
  acq_init(acq, emf);
  extend_model_init(acq, emf);
  airbc = emf->airbc;
  emf->airbc = 0;
  fdfd_init(emf);
  emf->airbc = airbc;
  if (airbc) { /* fdfd.c:41-49 and 54-62 without the airbc_init call */
    emf->lEx = emf->n1 * (emf->n2 + 1) * (emf->n3 + 1);
    emf->lEy = (emf->n1 + 1) * emf->n2 * (emf->n3 + 1);
    emf->lEz = (emf->n1 + 1) * (emf->n2 + 1) * emf->n3;
    emf->lE = emf->lEx + emf->lEy + emf->lEz;
    emf->xx = alloc1complex(emf->lE);
    emf->yy = alloc1complex(emf->lE);
    emf->n3 = emf->nz + emf->nb;
    emf->lEx = emf->n1 * (emf->n2 + 1) * (emf->n3 + 1);
    emf->lEy = (emf->n1 + 1) * emf->n2 * (emf->n3 + 1);
    emf->lEz = (emf->n1 + 1) * (emf->n2 + 1) * emf->n3;
    emf->lE = emf->lEx + emf->lEy + emf->lEz;
    emf->lHx = (emf->n1 + 1) * emf->n2 * emf->n3;
    emf->lHy = emf->n1 * (emf->n2 + 1) * emf->n2;
    emf->lHz = emf->n1 * emf->n2 * (emf->n3 + 1);
    emf->lH = emf->lHx + emf->lHy + emf->lHz;
  }
  emf->ne = emf->lE;
  emf->nh = emf->lH;
  emf->xe = alloc1complex(emf->ne);
  emf->xh = alloc1complex(emf->nh);
  emf->be = alloc1complex(emf->ne);
  memset(emf->xe, 0, emf->ne * sizeof(complex));
  memset(emf->xh, 0, emf->nh * sizeof(complex));
  memset(emf->be, 0, emf->ne * sizeof(complex));
  emf->I_omega_mu0 = I * 2. * PI * f * mu0; /* main.c:95 */
  x = alloc1complex(emf->ne);
  y = alloc1complex(emf->ne);
  if (fin) load(fin, x, emf->ne * sizeof(complex));

  if (!strcmp(mode, "grid")) {
    int n1 = emf->n1, n2 = emf->n2, n3 = emf->nz + 2 * emf->nb;
    double info[8];
    dump("x1", emf->x1, (n1 + 1) * sizeof(double));
    dump("x2", emf->x2, (n2 + 1) * sizeof(double));
    dump("x3", emf->x3, (n3 + 1) * sizeof(double));
    dump("x1s", emf->x1s, (n1 + 1) * sizeof(double));
    dump("x2s", emf->x2s, (n2 + 1) * sizeof(double));
    dump("x3s", emf->x3s, (n3 + 1) * sizeof(double));
    dump("d1", emf->d1, (n1 + 1) * sizeof(double));
    dump("d2", emf->d2, (n2 + 1) * sizeof(double));
    dump("d3", emf->d3, (n3 + 1) * sizeof(double));
    dump("d1s", emf->d1s, (n1 + 1) * sizeof(double));
    dump("d2s", emf->d2s, (n2 + 1) * sizeof(double));
    dump("d3s", emf->d3s, (n3 + 1) * sizeof(double));
    dump("sigma11", &emf->sigma11[0][0][0], (size_t)n1 * (n2 + 1) * (n3 + 1) * sizeof(double));
    dump("sigma22", &emf->sigma22[0][0][0], (size_t)(n1 + 1) * n2 * (n3 + 1) * sizeof(double));
    dump("sigma33", &emf->sigma33[0][0][0], (size_t)(n1 + 1) * (n2 + 1) * n3 * sizeof(double));
    info[0] = n1; info[1] = n2; info[2] = emf->n3; info[3] = emf->lextend;
    info[4] = emf->rho_water; info[5] = emf->rstretch; info[6] = emf->lE; info[7] = 0;
    dump("info", info, sizeof info);
  } else if (!strcmp(mode, "apply")) {
    if (airbc) err("apply needs airbc=0 in this build");
    fdfd_apply(emf->ne, x, y);
    dump("y", y, emf->ne * sizeof(complex));
    dump("xh", emf->xh, emf->nh * sizeof(complex));
  } else if (!strcmp(mode, "lusgs")) {
    lusgs_apply(emf->ne, x, y);
    dump("y", y, emf->ne * sizeof(complex));
  } else if (!strcmp(mode, "aphi")) {
    if (emf->preco != 3) err("aphi needs preco=3 (fdfd_init allocates its potential then)");
    aphi_apply(emf->ne, x, y);
    dump("y", y, emf->ne * sizeof(complex));
  } else if (!strcmp(mode, "gmresprec")) {
    if (airbc) err("gmresprec needs airbc=0 in this build");
    gmres_init(emf->ne, emf->mp, fdfd_apply);
    gmres_apply(emf->ne, x, y);
    gmres_free();
    dump("y", y, emf->ne * sizeof(complex));
  } else if (!strcmp(mode, "inject")) {
    inject_source(acq, emf, y);
    dump("b", y, emf->ne * sizeof(complex));
  } else if (!strcmp(mode, "extract")) {
    emf->dcal_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
    memset(&emf->dcal_fd[0][0][0], 0, acq->nrec * emf->nfreq * emf->nchrec * sizeof(float _Complex));
    extract_emf(acq, emf, x, 0);
    for (ic = 0; ic < emf->nchrec; ic++) {
      char what[32];
      snprintf(what, sizeof what, "rec%d", ic);
      dump(what, &emf->dcal_fd[ic][0][0], acq->nrec * sizeof(float _Complex));
    }
  } else if (!strcmp(mode, "solve")) { /* main.c:89-127 */
    if (airbc) err("solve needs airbc=0 in this build");
    emf->dcal_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
    memset(&emf->dcal_fd[0][0][0], 0, acq->nrec * emf->nfreq * emf->nchrec * sizeof(float _Complex));
    for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
      float tol = emf->tol / emf->freqs[ifreq];
      char what[32];
      emf->I_omega_mu0 = I * 2. * PI * emf->freqs[ifreq] * mu0;
      memset(emf->xe, 0, emf->ne * sizeof(complex));
      memset(emf->xh, 0, emf->nh * sizeof(complex));
      memset(emf->be, 0, emf->ne * sizeof(complex));
      inject_source(acq, emf, emf->be);
      printf("harness freq=%d\n", ifreq);
      if (emf->preco == 0) {
        if (emf->algopt == 1) bicgstab(emf->ne, emf->xe, emf->be, fdfd_apply, emf->niter, tol, emf->verb);
        if (emf->algopt == 2) gmres(emf->ne, emf->xe, emf->be, fdfd_apply, emf->niter / emf->nrestart, tol, emf->nrestart, emf->verb);
      } else if (emf->preco == 1) {
        if (emf->algopt == 1) fbicgstab(emf->ne, emf->xe, emf->be, fdfd_apply, lusgs_apply, emf->niter, tol, emf->verb);
        if (emf->algopt == 2) fgmres(emf->ne, emf->xe, emf->be, fdfd_apply, lusgs_apply, emf->niter / emf->nrestart, tol, emf->nrestart, emf->verb);
      } else if (emf->preco == 2) { /* main.c:110-115 */
        gmres_init(emf->ne, emf->mp, fdfd_apply);
        if (emf->algopt == 1) fbicgstab(emf->ne, emf->xe, emf->be, fdfd_apply, gmres_apply, emf->niter, tol, emf->verb);
        if (emf->algopt == 2) fgmres(emf->ne, emf->xe, emf->be, fdfd_apply, gmres_apply, emf->niter / emf->nrestart, tol, emf->nrestart, emf->verb);
        gmres_free();
      } else if (emf->preco == 3) {
        if (emf->algopt == 1) fbicgstab(emf->ne, emf->xe, emf->be, fdfd_apply, aphi_apply, emf->niter, tol, emf->verb);
        if (emf->algopt == 2) fgmres(emf->ne, emf->xe, emf->be, fdfd_apply, aphi_apply, emf->niter / emf->nrestart, tol, emf->nrestart, emf->verb);
      }
      snprintf(what, sizeof what, "x%d", ifreq);
      dump(what, emf->xe, emf->ne * sizeof(complex));
      extract_emf(acq, emf, emf->xe, ifreq);
    }
    write_data(acq, emf, "emf_0001.txt", emf->dcal_fd);
  } else if (!strcmp(mode, "acsr")) { /* main.c:183-196 at freqs[0] */
    if (airbc) err("acsr needs airbc=0");
    Acsr_init(emf, 0);
    dump("row_ptr", Acsr->row_ptr, (Acsr->nrow + 1) * sizeof(int));
    dump("col", Acsr->col_ind, Acsr->nnz * sizeof(int));
    dump("val", Acsr->val, Acsr->nnz * sizeof(complex));
    Acsr_apply(emf->ne, x, y);
    dump("y", y, emf->ne * sizeof(complex));
    Acsr_lusgs_apply(emf->ne, x, y);
    dump("ysgs", y, emf->ne * sizeof(complex));
    Acsr_free();
  } else {
    err("unknown hmode=%s", mode);
  }
  return 0;
}
