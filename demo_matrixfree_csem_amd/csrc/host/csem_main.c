/* csem_main.c -- the reference's driver (main.c:46-145) on the device
 * library: same key=value flags, same per-frequency loop (ascending
 * frequencies, tol/f in float, x0 = 0, inject -> solve -> extract), same
 * emf_0001.txt.  The solve time is measured with a double-precision clock
 * around the solver call only, as main.c:103-120 scopes it. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "../../../include/csem.h"

#define CSEM_PI 3.14159265358979323846264
#define CSEM_MU0 (4. * CSEM_PI * 1e-7)

static double now(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

int main(int argc, char **argv)
{
  emf_t *emf = (emf_t *)calloc(1, sizeof(emf_t));
  acq_t *acq = (acq_t *)calloc(1, sizeof(acq_t));
  int ifreq;
  float tol;
  double t0;
  initargs(argc, argv);
  if (!getparint("mode", &emf->mode)) emf->mode = 0;
  if (!getparint("verb", &emf->verb)) emf->verb = 1;
  if (!getparint("niter", &emf->niter)) emf->niter = 1000;
  if (!getparint("preco", &emf->preco)) emf->preco = 1;
  if (!getparint("algopt", &emf->algopt)) emf->algopt = 1;
  if (!getparfloat("tol", &emf->tol)) emf->tol = 1e-8;
  if (!getparint("nrestart", &emf->nrestart)) emf->nrestart = 10;
  if (!getparint("mp", &emf->mp)) emf->mp = 30;
  if (emf->preco < 0 || emf->preco > 3) err("preco=%d: 0 none, 1 LU-SGS, 2 inner GMRES, 3 A-phi", emf->preco);
  emf_init(emf);
  acq_init(acq, emf);
  extend_model_init(acq, emf);
  fdfd_init(emf);
  if (emf->verb) printf("dof=%d\n", emf->lE);
  emf->ne = emf->lE;
  emf->nh = emf->lH;
  emf->xe = alloc1complex(emf->ne);
  emf->xh = alloc1complex(emf->nh);
  emf->be = alloc1complex(emf->ne);
  emf->dcal_fd = alloc3complexf(acq->nrec, emf->nfreq, emf->nchrec);
  memset(&emf->dcal_fd[0][0][0], 0, (size_t)acq->nrec * emf->nfreq * emf->nchrec * sizeof(float _Complex));
  /* the receivers in double precision too (emf_0001_f64.txt): extract_emf
   * rounds them to float as the reference does */
  csem_complex *dval = (csem_complex *)calloc((size_t)emf->nchrec * emf->nfreq * acq->nrec, sizeof(csem_complex));
  csem_complex *done = (csem_complex *)calloc((size_t)emf->nchrec * acq->nrec, sizeof(csem_complex));
  for (ifreq = 0; ifreq < emf->nfreq; ifreq++) {
    tol = emf->tol / emf->freqs[ifreq];
    emf->I_omega_mu0 = I * 2. * CSEM_PI * emf->freqs[ifreq] * CSEM_MU0;
    printf("------freq=%g Hz, tol=%e--------\n", emf->freqs[ifreq], tol);
    memset(emf->xe, 0, emf->ne * sizeof(csem_complex));
    memset(emf->xh, 0, emf->nh * sizeof(csem_complex));
    memset(emf->be, 0, emf->ne * sizeof(csem_complex));
    inject_source(acq, emf, emf->be);
    t0 = now();
    if (emf->preco == 0) {
      if (emf->algopt == 1) bicgstab(emf->ne, emf->xe, emf->be, fdfd_apply, emf->niter, tol, emf->verb);
      if (emf->algopt == 2)
        gmres(emf->ne, emf->xe, emf->be, fdfd_apply, emf->niter / emf->nrestart, tol, emf->nrestart, emf->verb);
    } else {
      /* main.c:106-117: the preconditioner by preco */
      op_t M = emf->preco == 1 ? lusgs_apply : emf->preco == 2 ? gmres_apply : aphi_apply;
      if (emf->preco == 2) gmres_init(emf->ne, emf->mp, fdfd_apply);
      if (emf->algopt == 1) fbicgstab(emf->ne, emf->xe, emf->be, fdfd_apply, M, emf->niter, tol, emf->verb);
      if (emf->algopt == 2)
        fgmres(emf->ne, emf->xe, emf->be, fdfd_apply, M, emf->niter / emf->nrestart, tol, emf->nrestart,
               emf->verb);
      if (emf->preco == 2) gmres_free();
    }
    printf("elapsed time: %g sec\n", now() - t0);
    extract_emf(acq, emf, emf->xe, ifreq);
    extract_emf_double(acq, emf, emf->xe, done);
    for (int ic = 0; ic < emf->nchrec; ic++)
      memcpy(dval + ((long)ic * emf->nfreq + ifreq) * acq->nrec, done + (long)ic * acq->nrec,
             acq->nrec * sizeof(csem_complex));
  }
  write_data(acq, emf, "emf_0001.txt", emf->dcal_fd);
  write_data_double(acq, emf, "emf_0001_f64.txt", dval);
  free(dval);
  free(done);
  free1complex(emf->xe);
  free1complex(emf->xh);
  free1complex(emf->be);
  free3complexf(emf->dcal_fd);
  fdfd_free(emf);
  extend_model_free(emf);
  emf_free(emf);
  acq_free(acq);
  free(emf);
  free(acq);
  return 0;
}
