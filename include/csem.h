/* csem.h -- drop-in C entry points of demo_matrixfree_csem's single-rank
 * frequency-domain solve path, backed by the MI355X device library.
 *
 * A caller written against the reference (main.c:17-43 declares what it
 * uses) compiles against this header and links libcsem.so instead of the
 * reference objects.  Types keep the reference's field order and meaning so
 * that the same main() works unchanged; what runs underneath is the HIP
 * path in csem_dev.h.
 *
 * Replaces, entry point by entry point:
 *   emf_init / emf_free                 emf_init_free.c:14-184
 *   acq_init / acq_free                 acq_init_free.c:14-165
 *   extend_model_init / _free           extend_model.c:24-318
 *   fdfd_init / fdfd_free               fdfd.c:30-80
 *   fdfd_apply   (op_t)                 fdfd.c:86-205
 *   lusgs_apply  (op_t)                 fdfd.c:287-510
 *   airbc_init / airbc_free / _apply    airbc_init_free.c:24-209
 *   aphi_apply   (op_t)                 fdfd.c:513-821 (A-phi preconditioner, preco=3)
 *   gmres_init / gmres_free / _apply    gmres_init_free.c:29-137 (inner GMRES, preco=2)
 *   bicgstab / fbicgstab                solver.c:16-87, 93-178
 *   gmres / fgmres                      solver.c:181-285, 403-512
 *   Acsr_* / Acoo_*                     Acsr.c:22-748, Acoo.c (main.c:148-151, 221-223;
 *                                       not provided: they report an error)
 *   cdotprod                            cdotprod.c:13-20
 *   inject_source / extract_emf         inject_extract.c:24-213
 *   write_data                          write_data.c:15-39
 *   find_index                          find_index.c:9-30
 *   create_nugrid                       create_nugrid.c:16-40
 *   initargs / getpar* / countparval    cstd.c:936-1300 (SU getpar semantics)
 *   alloc1..3* / free1..3*              cstd.c:213-700 (contiguous arrays)
 *   err / warn                          cstd.c:1539-1568
 * and adds csem_last_status() (non-exiting error reporting).
 */
#ifndef CSEM_H
#define CSEM_H
#include <complex.h>
#include <stddef.h>
#include "csem_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex csem_complex;

/* op_t / rop_t, solver.h:4-5 */
typedef void (*op_t)(int, csem_complex *, csem_complex *);
typedef void (*rop_t)(int, double *, double *);

/* emf_t: same fields, order and types as emf.h:4-57 */
typedef struct {
  int mode, verb, reciprocity, fvm, airbc;
  int nfreq;
  float *freqs;
  float rho_air, lextend, rho_water, rhomin, rhomax;
  int nchsrc, nchrec;
  char **chsrc, **chrec;
  int istretch;
  float rstretch;
  int nx, ny, nz, nb;
  int n1, n2, n3;
  int lv, lE, lH, lEx, lEy, lEz, lHx, lHy, lHz;
  float *x1node, *x2node, *x3node;
  double *x1, *x2, *x3;
  double *x1s, *x2s, *x3s;
  double *d1, *d2, *d3;
  double *d1s, *d2s, *d3s;
  double d1min, d2min, d3min;
  double x1min, x2min, x3min;
  double x1max, x2max, x3max;
  int n1fft, n2fft;
  double d1fft, d2fft, d3fft;
  csem_complex **sH1kxky, **sH2kxky;
  float ***rho11, ***rho22, ***rho33;
  double ***sigma11, ***sigma22, ***sigma33;
  float _Complex ****E1, ****E2, ****E3;
  float _Complex ***dcal_fd;
  int ne, nh;
  csem_complex *xe, *xh, *be;
  csem_complex *v, *bv;
  csem_complex *xx, *yy;
  csem_complex I_omega_mu0;
  int preco, algopt, niter, nrestart, mp;
  float tol;
  int param_extended;
} emf_t;

/* acq_t: acq.h:3-18 */
typedef struct {
  int nsrc, nrec, nsubsrc, nsubrec;
  float lensrc, lenrec;
  float *src_x1, *src_x2, *src_x3, *src_azimuth, *src_dip;
  float *rec_x1, *rec_x2, *rec_x3, *rec_azimuth, *rec_dip;
  float x1min, x1max, x2min, x2max, x3min, x3max;
  int nsrc_total, nrec_total;
  int *shot_idx;
} acq_t;

extern int iproc, nproc, ierr; /* cstd.h:122; always 0, 1, 0 here */

void emf_init(emf_t *emf);
void emf_free(emf_t *emf);
void acq_init(acq_t *acq, emf_t *emf);
void acq_free(acq_t *acq);
void extend_model_init(acq_t *acq, emf_t *emf);
void extend_model_free(emf_t *emf);
void fdfd_init(emf_t *emf);
void fdfd_free(emf_t *emf);
void fdfd_apply(int n, csem_complex *x, csem_complex *y);
void lusgs_apply(int n, csem_complex *x, csem_complex *y);
void airbc_init(emf_t *emf);
void airbc_free(emf_t *emf);
void airbc_apply(emf_t *emf, csem_complex *H3k, csem_complex *H1km1, csem_complex *H2km1);
void aphi_apply(int n, csem_complex *x, csem_complex *y);
void gmres_init(int n, int m, op_t Aop);
void gmres_free(void);
void gmres_apply(int n, csem_complex *b, csem_complex *x);
/* The Krylov drivers accept any op_t.  fdfd_apply, lusgs_apply,
 * aphi_apply and gmres_apply (on fdfd_apply) run as device kernels and
 * the whole solve stays device-resident; any other function is called on
 * host copies of its input and output vectors. */
void bicgstab(int n, csem_complex *x, csem_complex *b, op_t Aop, int niter, double tol, int verb);
void gmres(int n, csem_complex *x, csem_complex *b, op_t Aop, int niter, double tol, int m, int verb);
void fbicgstab(int n, csem_complex *x, csem_complex *b, op_t Aop, op_t invMop, int niter, double tol, int verb);
void fgmres(int n, csem_complex *x, csem_complex *b, op_t Aop, op_t invMop, int niter, double tol, int m, int verb);
csem_complex cdotprod(int n, csem_complex *a, csem_complex *b);
void inject_source(acq_t *acq, emf_t *emf, csem_complex *b);
void extract_emf(acq_t *acq, emf_t *emf, csem_complex *x, int ifreq);
void write_data(acq_t *acq, emf_t *emf, char *fname, float _Complex ***dcal_fd);
/* No counterpart in the reference, which keeps receivers only as float32
 * (inject_extract.c:159,182,208): the double-precision values extract_emf
 * rounds, out[ic * nrec + ir] for every receiver channel ic, and a result
 * file in write_data's layout with them printed %.17e
 * (vals[(ic * nfreq + ifreq) * nrec + ir]). */
void extract_emf_double(acq_t *acq, emf_t *emf, csem_complex *x, csem_complex *out);
void write_data_double(acq_t *acq, emf_t *emf, char *fname, csem_complex *vals);
/* The reference's matrix-storage variants (main.c:148-151, 221-223: CSR
 * assembly with CSR LU-SGS, and a MUMPS direct solve, used only by its
 * main2 / main3).  Declared so that a reference main.c links; each call
 * reports an error through err() / csem_last_status(). */
void Acsr_init(emf_t *emf, int ifreq);
void Acsr_free(void);
void Acsr_apply(int n, csem_complex *x, csem_complex *y);
void Acsr_lusgs_apply(int n, csem_complex *x, csem_complex *y);
void Acoo_init(emf_t *emf, int ifreq);
void Acoo_free(void);
void Acoo_solve(emf_t *emf);
int find_index(int n, double *x, double val);
double create_nugrid(int n, double len, double dx, double *x, int istretch);

/* getpar (cstd.c): key=value argv parsing, last occurrence wins, comma
 * separated vectors, par=file parameter files */
void initargs(int argc, char **argv);
int getparint(char *name, int *p);
int getparfloat(char *name, float *p);
int getpardouble(char *name, double *p);
int getparstring(char *name, char **p);
int getparstringarray(char *name, char **p);
int getnparint(int n, char *name, int *p);
int getnparfloat(int n, char *name, float *p);
int getnpardouble(int n, char *name, double *p);
int getnparstring(int n, char *name, char **p);
int countparname(char *name);
int countparval(char *name);
int countnparval(int n, char *name);

void *alloc1(size_t n1, size_t size);
void **alloc2(size_t n1, size_t n2, size_t size);
void ***alloc3(size_t n1, size_t n2, size_t n3, size_t size);
void free1(void *p);
void free2(void **p);
void free3(void ***p);
int *alloc1int(size_t n1);
void free1int(int *p);
float *alloc1float(size_t n1);
float ***alloc3float(size_t n1, size_t n2, size_t n3);
void free1float(float *p);
void free3float(float ***p);
double *alloc1double(size_t n1);
double ***alloc3double(size_t n1, size_t n2, size_t n3);
void free1double(double *p);
void free3double(double ***p);
csem_complex *alloc1complex(size_t n1);
csem_complex **alloc2complex(size_t n1, size_t n2);
void free1complex(csem_complex *p);
void free2complex(csem_complex **p);
float _Complex ***alloc3complexf(size_t n1, size_t n2, size_t n3);
void free3complexf(float _Complex ***p);

void err(char *fmt, ...);
void warn(char *fmt, ...);

/* Status of the last host entry point (0 ok; < 0 error, message in
 * csem_last_error()).  err() still exits, as the reference does, unless
 * csem_set_exit_on_error(0) was called; then it records the error here. */
int csem_last_status(void);
void csem_set_exit_on_error(int on);

/* fdfd_init in two halves: the host half (lengths, operator truncation,
 * airwave tables) and the device model built from a prepared emf_t. */
void csem_fdfd_prepare(emf_t *emf);
int csem_model_from_emf(const emf_t *emf, csem_model **out);

/* extend_model_init in two halves: the 1-D half (extended and staggered
 * nodes and spacings, extend_model.c:24-106) alone. */
void csem_extend_grid(acq_t *acq, emf_t *emf);
/* Device-side setup in place of extend_model_init + fdfd_init (no
 * counterpart in the reference; for servers that build many models): after
 * emf_init and acq_init, the 1-D halves on the host, then the cells, the
 * edge-averaged conductivities and the airwave multipliers on the device
 * (csem_model_create_raw).  On return emf is as after extend_model_init and
 * csem_fdfd_prepare, with emf->sigma11..33 copied back from the device (for
 * inject_source and extract_emf) and no host multipliers; the model is the
 * caller's (csem_model_destroy), not the active one. */
int csem_model_build_device(acq_t *acq, emf_t *emf, csem_model **out);

/* The device model behind fdfd_init (NULL before it). */
csem_model *csem_active_model(void);
/* Iteration index reached by the last Krylov call ("converged at iter="). */
int csem_last_iterations(void);

#ifdef __cplusplus
}
#endif
#endif
