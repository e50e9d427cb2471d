/* csem_oracle.h -- CPU restatement of demo_matrixfree_csem's single-rank
 * frequency-domain solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against, and the CPU baseline that bench.py times.  It is
 * never linked into, loaded by or called from the product library.
 *
 * Every function restates a reference routine (file:line cited in
 * csem_oracle.c).  Complex values are C99 `double _Complex`, laid out as
 * (re, im) pairs, so numpy complex128 buffers can be passed directly.
 */
#ifndef CSEM_ORACLE_H
#define CSEM_ORACLE_H
#include <complex.h>

typedef double _Complex orc_c128;
typedef struct orc orc_t;

/* Reduction order used by the dot products.
 *   ORC_RED_SERIAL : left-to-right single accumulator (cdotprod.c:18)
 *   ORC_RED_CHUNKED: the device order -- chunks of 4096 elements, 256
 *                    per-thread sequential sums of 16 strided elements,
 *                    pairwise halving tree, then the same over partials. */
enum { ORC_RED_SERIAL = 0, ORC_RED_CHUNKED = 1 };

/* Build the extended model (extend_model.c:24-294) from the raw input model
 * (the float32 node/cell arrays emf_init reads, emf_init_free.c:77-166).
 * lextend < 0 selects the reference default (extend_model.c:43).
 * airbc=1 also performs fdfd_init's truncation and airbc_init
 * (fdfd.c:40-53, airbc_init_free.c:24-109) with an nfft x nfft transform. */
orc_t *orc_build(int nx, int ny, int nz,
                 const float *x1node, const float *x2node, const float *x3node,
                 const float *rho11, const float *rho22, const float *rho33,
                 int nb, float lextend, int istretch, float rho_air,
                 int airbc, int fvm, int nfft);
void orc_free(orc_t *m);

/* integer facts: 0 n1, 1 n2, 2 n3 (operator), 3 n3 full, 4 lEx, 5 lEy,
 * 6 lEz, 7 lE (operator), 8 full lEx, 9 full lEy, 10 full lEz, 11 full lE,
 * 12 nfft, 13 lHx, 14 lHy, 15 lHz */
long orc_info(const orc_t *m, int what);
/* float facts: 0 rho_water, 1 lextend, 2..7 rstretch (left,right,front,
 * rear,top,bottom), 8 d1fft, 9 d2fft, 10 d3fft */
double orc_finfo(const orc_t *m, int what);
/* named arrays: "x1","x2","x3","x1s","x2s","x3s","d1","d2","d3","d1s",
 * "d2s","d3s","sigma11","sigma22","sigma33","cell11","cell22","cell33"
 * (double; cells as extend_model.c:256-275 dumps them); "sH1","sH2","xh",
 * "Hxg","Hyg" (complex); "wx","wy" (double); "jx","jy" (int) */
void *orc_array(const orc_t *m, const char *name);

void orc_set_red(orc_t *m, int red);
/* I_omega_mu0 = I*2*PI*f*mu0 exactly as main.c:95 */
void orc_set_freq(orc_t *m, float freq);

void orc_fdfd_apply(orc_t *m, const orc_c128 *x, orc_c128 *y);
void orc_lusgs_apply(orc_t *m, const orc_c128 *x, orc_c128 *y);
/* aphi_apply (fdfd.c:513-821), the A-phi preconditioner of preco=3 */
void orc_aphi_apply(orc_t *m, const orc_c128 *x, orc_c128 *y);
/* gmres_apply (gmres_init_free.c:58-137), the inner-GMRES preconditioner of
 * preco=2 with orc_set_mp's dimension (default 30, main.c:69) */
void orc_gmres_prec(orc_t *m, const orc_c128 *b, orc_c128 *x);
void orc_set_mp(orc_t *m, int mp);
void orc_airbc_apply(orc_t *m, const orc_c128 *H3k, orc_c128 *H1km1, orc_c128 *H2km1);
/* One stage of orc_airbc_apply: 0 forward bilinear map Hz(n1*n2) ->
 * uniform N*N; 1 uniform N*N -> (Hx, Hy) on the uniform grid (DFT,
 * multipliers, inverse DFT / N^2; out and out2); 2 uniform -> ghost Hx
 * (n1+1)*n2; 3 uniform -> ghost Hy n1*(n2+1). */
void orc_airbc_stage(orc_t *m, int stage, const orc_c128 *in, orc_c128 *out, orc_c128 *out2);

/* 2-D transform used by the airwave boundary: data[i1 + n*i2].
 * dir=-1 forward (rows i1 then columns i2), dir=+1 inverse unnormalised
 * (columns then rows). */
void orc_fft2(orc_t *m, orc_c128 *data, int dir);

orc_c128 orc_dot(const orc_t *m, long n, const orc_c128 *a, const orc_c128 *b);

/* Krylov drivers (solver.c).  precond: 0 none, 1 LU-SGS.  Return the
 * iteration index at convergence (the number printed by "converged at
 * iter="), or niter when the loop ran out; hist (may be NULL, length
 * niter+1) receives relres per iteration. */
/* precond: 0 none, 1 LU-SGS, 2 inner GMRES, 3 A-phi (main.c:104-118) */
int orc_fbicgstab(orc_t *m, orc_c128 *x, const orc_c128 *b, int niter, double tol, int precond, double *hist);
int orc_fgmres(orc_t *m, orc_c128 *x, const orc_c128 *b, int nouter, double tol, int mrestart, int precond, double *hist);

int orc_gmres(orc_t *m, orc_c128 *x, const orc_c128 *b, int nouter, double tol, int mrestart, double *hist);

/* inject_source for one source and channel (0 Ex, 1 Ey, 2 Ez); adds into b */
void orc_inject(orc_t *m, int ch, float sx, float sy, float sz, orc_c128 *b);
/* extract_emf for one receiver, returning the double value before the
 * float32 store of inject_extract.c:159,182,208 */
orc_c128 orc_extract(orc_t *m, int ch, float rx, float ry, float rz, const orc_c128 *x);

/* Matrix storage (Acsr.c): the CSR matrix on the full grid at the current
 * frequency (returns nnz; fills row_ptr[rows + 1], col[nnz], val[nnz]
 * when non-NULL; each row's range holds its own diagonal, last), the
 * product and the symmetric Gauss-Seidel of Acsr.c:704-748. */
long orc_csr_build(const orc_t *m, int *row_ptr, int *col, orc_c128 *val);
void orc_csr_apply(long nrow, const int *row_ptr, const int *col, const orc_c128 *val, const orc_c128 *x,
                   orc_c128 *y);
void orc_csr_sgs(long nrow, const int *row_ptr, const int *col, const orc_c128 *val, const orc_c128 *x,
                 orc_c128 *y);

double orc_create_nugrid(int n, double len, double dx, double *x, int istretch);
int orc_find_index(int n, const double *x, double val);
#endif
