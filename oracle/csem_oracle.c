/* csem_oracle.c -- CPU restatement of demo_matrixfree_csem's single-rank
 * frequency-domain solve path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg are its only users.  The product library never links or
 * calls it.
 *
 * It follows the reference loop order and per-cell expression trees exactly,
 * written with C99 complex arithmetic compiled by gcc (no FMA contraction,
 * complex division through libgcc's Smith routine), so that it reproduces
 * the reference's floating-point results.  Two deliberate differences from
 * the reference, both documented where they occur:
 *   - the airwave 2-D DFT is a radix-4/radix-2 Stockham transform with a
 *     fixed pass order instead of FFTW (FFTW is not available here);
 *   - dot products can run in the device reduction order (ORC_RED_CHUNKED)
 *     as well as the reference's serial order (ORC_RED_SERIAL).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "csem_oracle.h"

#define ORC_PI 3.14159265358979323846264 /* cstd.h:35-37 */
#define ORC_MU0 (4. * ORC_PI * 1e-7)     /* cstd.h:38-40 */
#define MAXI(a, b) ((a) > (b) ? (a) : (b))
#define MINI(a, b) ((a) < (b) ? (a) : (b))

typedef double _Complex zc;

struct orc {
  int nx, ny, nz, nb, n1, n2, n3, n3f, airbc, fvm, istretch;
  float lextend, rho_air, rho_water, rstretch[6];
  double *x1, *x2, *x3, *x1s, *x2s, *x3s, *d1, *d2, *d3, *d1s, *d2s, *d3s;
  double *s11, *s22, *s33; /* edge conductivities on the full (air-extended) grid */
  double *c11, *c22, *c33; /* extended cell model as extend_model.c:256-275 dumps it */
  long lEx, lEy, lEz, lE;  /* operator vector (truncated grid when airbc) */
  long fEx, fEy, fEz, fE;  /* full grid, used by LU-SGS under airbc */
  long lHx, lHy, lHz, lH, hz_off;
  zc *xh, *xx, *yy, *Hxg, *Hyg;
  int nfft;
  double d1fft, d2fft, d3fft;
  zc *sH1, *sH2, *buf, *buf0, *buf2, *tw, *tmp;
  int *jx, *jy;
  double *wx, *wy;
  zc iomu, iwmu;
  int red;
  int mp;           /* inner GMRES dimension of the preco=2 preconditioner */
  zc *av, *abv;     /* A-phi potential and its right-hand side (preco=3) */
};

/* ------------------------------------------------------------------ grid */

/* find_index.c:9-30 */
int orc_find_index(int n, const double *x, double val)
{
  int low = 0, high = n - 1, i;
  while (low < high) {
    i = (low + high) / 2;
    if (x[i] <= val) low = i;
    if (x[i] > val) high = i;
    if (low == high || low == high - 1) break;
  }
  return low;
}

/* create_nugrid.c:16-40 */
double orc_create_nugrid(int n, double len, double dx, double *x, int istretch)
{
  int i, iter;
  double q, qq, eps = 1e-15;
  if (n * dx >= len || !istretch) {
    for (i = 0; i <= n; i++) x[i] = i * dx;
    return 1;
  }
  q = 1.1;
  for (iter = 0; iter < 100; ++iter) {
    qq = pow(len * (q - 1.) / dx + 1., 1. / n);
    if (fabs(qq - q) < eps) break;
    q = qq;
  }
  for (x[0] = 0, i = 1; i <= n; i++) x[i] = (pow(q, i) - 1.) * dx / (q - 1.);
  return q;
}

/* extend_model.c:8-21 */
static void stagger(int n, const double *x, double *xs, double *d, double *ds)
{
  int i;
  for (i = 0; i < n; i++) {
    ds[i] = x[i + 1] - x[i];
    xs[i] = 0.5 * (x[i] + x[i + 1]);
  }
  ds[n] = ds[n - 1];
  xs[n] = x[n] + 0.5 * ds[n - 1];
  for (i = 1; i <= n; i++) d[i] = xs[i] - xs[i - 1];
  d[0] = d[1];
}

static inline zc mk(double re, double im)
{
  union { zc z; double d[2]; } u;
  u.d[0] = re;
  u.d[1] = im;
  return u.z;
}

static double *dalloc(long n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }
static zc *zalloc(long n) { return (zc *)calloc((size_t)(n > 0 ? n : 1), sizeof(zc)); }

/* the stretched padding of one side, extend_model.c:76-106 */
static float pad_side(const orc_t *m, double dx, float *rs)
{
  double *tmp = dalloc(m->nb + 1);
  *rs = (float)orc_create_nugrid(m->nb, m->lextend, dx, tmp, m->istretch);
  free(tmp);
  return *rs;
}

static void fft_init(orc_t *m);

orc_t *orc_build(int nx, int ny, int nz,
                 const float *x1node, const float *x2node, const float *x3node,
                 const float *rho11, const float *rho22, const float *rho33,
                 int nb, float lextend, int istretch, float rho_air,
                 int airbc, int fvm, int nfft)
{
  orc_t *m = (orc_t *)calloc(1, sizeof(orc_t));
  m->mp = 30; /* main.c:69 default */
  int i, j, k, im1, jm1, km1;
  long c;
  double dx;
  float rs;
  m->nx = nx; m->ny = ny; m->nz = nz; m->nb = nb;
  m->airbc = airbc; m->fvm = fvm; m->istretch = istretch; m->rho_air = rho_air;
  m->nfft = nfft; m->red = ORC_RED_SERIAL;

  /* rho_water = min over rho11, rho22, rho33 (emf_init_free.c:152-166) */
  {
    float rmin = rho11[0];
    for (c = 0; c < (long)nx * ny * nz; c++) {
      rmin = MINI(rmin, rho11[c]);
      rmin = MINI(rmin, rho22[c]);
      rmin = MINI(rmin, rho33[c]);
    }
    m->rho_water = rmin;
  }

  /* extend_model.c:33-51 */
  m->n1 = nx + 2 * nb; m->n2 = ny + 2 * nb; m->n3f = nz + 2 * nb;
  m->lextend = (lextend < 0) ? ((nb > 0) ? 40e3 : 0) : lextend;
  if (nb > 0) {
    dx = x1node[nx] - x1node[0];
    m->lextend = MAXI(m->lextend, 2.2 * dx);
    dx = x2node[ny] - x2node[0];
    m->lextend = MAXI(m->lextend, 2.2 * dx);
    dx = x3node[nz] - x3node[0];
    m->lextend = MAXI(m->lextend, 2.2 * dx);
  }
  int n1 = m->n1, n2 = m->n2, n3 = m->n3f;
  m->x1 = dalloc(n1 + 1); m->x2 = dalloc(n2 + 1); m->x3 = dalloc(n3 + 1);
  m->x1s = dalloc(n1 + 1); m->x2s = dalloc(n2 + 1); m->x3s = dalloc(n3 + 1);
  m->d1 = dalloc(n1 + 1); m->d2 = dalloc(n2 + 1); m->d3 = dalloc(n3 + 1);
  m->d1s = dalloc(n1 + 1); m->d2s = dalloc(n2 + 1); m->d3s = dalloc(n3 + 1);
  for (i = 0; i <= nx; i++) m->x1[i + nb] = x1node[i];
  for (j = 0; j <= ny; j++) m->x2[j + nb] = x2node[j];
  for (k = 0; k <= nz; k++) m->x3[k + nb] = x3node[k];
  if (nb > 0) { /* extend_model.c:75-108 */
    dx = x1node[1] - x1node[0];
    rs = pad_side(m, dx, &m->rstretch[0]);
    for (i = 0; i < nb; i++) m->x1[nb - 1 - i] = m->x1[nb - i] - dx * pow(rs, i);
    dx = x1node[nx] - x1node[nx - 1];
    rs = pad_side(m, dx, &m->rstretch[1]);
    for (i = 0; i < nb; i++) m->x1[nb + nx + 1 + i] = m->x1[nb + nx + i] + dx * pow(rs, i);
    dx = x2node[1] - x2node[0];
    rs = pad_side(m, dx, &m->rstretch[2]);
    for (j = 0; j < nb; j++) m->x2[nb - 1 - j] = m->x2[nb - j] - dx * pow(rs, j);
    dx = x2node[ny] - x2node[ny - 1];
    rs = pad_side(m, dx, &m->rstretch[3]);
    for (j = 0; j < nb; j++) m->x2[nb + ny + 1 + j] = m->x2[nb + ny + j] + dx * pow(rs, j);
    dx = x3node[1] - x3node[0];
    rs = pad_side(m, dx, &m->rstretch[4]);
    for (k = 0; k < nb; k++) m->x3[nb - 1 - k] = m->x3[nb - k] - dx * pow(rs, k);
    dx = x3node[nz] - x3node[nz - 1];
    rs = pad_side(m, dx, &m->rstretch[5]);
    for (k = 0; k < nb; k++) m->x3[nb + nz + 1 + k] = m->x3[nb + nz + k] + dx * pow(rs, k);
  }
  stagger(n1, m->x1, m->x1s, m->d1, m->d1s);
  stagger(n2, m->x2, m->x2s, m->d2, m->d2s);
  stagger(n3, m->x3, m->x3s, m->d3, m->d3s);

  /* cell model on the extended grid, extend_model.c:126-177, [k][j][i] */
  long nc = (long)n1 * n2 * n3;
  double *sg11 = dalloc(nc), *sg22 = dalloc(nc), *rh33 = dalloc(nc);
#define C3(k, j, i) ((i) + (long)n1 * ((j) + (long)n2 * (k)))
#define R3(k, j, i) ((i) + (long)nx * ((j) + (long)ny * (k)))
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) {
        sg11[C3(k + nb, j + nb, i + nb)] = 1. / rho11[R3(k, j, i)];
        sg22[C3(k + nb, j + nb, i + nb)] = 1. / rho22[R3(k, j, i)];
        rh33[C3(k + nb, j + nb, i + nb)] = rho33[R3(k, j, i)];
      }
  if (nb > 0) {
    for (k = 0; k < n3; k++)
      for (j = 0; j < n2; j++)
        for (i = 0; i < nb; i++) {
          sg11[C3(k, j, i)] = sg11[C3(k, j, nb)];
          sg11[C3(k, j, n1 - 1 - i)] = sg11[C3(k, j, n1 - 1 - nb)];
          sg22[C3(k, j, i)] = sg22[C3(k, j, nb)];
          sg22[C3(k, j, n1 - 1 - i)] = sg22[C3(k, j, n1 - 1 - nb)];
          rh33[C3(k, j, i)] = rh33[C3(k, j, nb)];
          rh33[C3(k, j, n1 - 1 - i)] = rh33[C3(k, j, n1 - 1 - nb)];
        }
    for (k = 0; k < n3; k++)
      for (j = 0; j < nb; j++)
        for (i = 0; i < n1; i++) {
          sg11[C3(k, j, i)] = sg11[C3(k, nb, i)];
          sg11[C3(k, n2 - 1 - j, i)] = sg11[C3(k, n2 - 1 - nb, i)];
          sg22[C3(k, j, i)] = sg22[C3(k, nb, i)];
          sg22[C3(k, n2 - 1 - j, i)] = sg22[C3(k, n2 - 1 - nb, i)];
          rh33[C3(k, j, i)] = rh33[C3(k, nb, i)];
          rh33[C3(k, n2 - 1 - j, i)] = rh33[C3(k, n2 - 1 - nb, i)];
        }
    /* air on top; the bottom copies row [k] of the last core layer instead
     * of row [j] (extend_model.c:169,171,173) */
    for (k = 0; k < nb; k++)
      for (j = 0; j < n2; j++)
        for (i = 0; i < n1; i++) {
          sg11[C3(k, j, i)] = 1. / rho_air;
          sg11[C3(n3 - 1 - k, j, i)] = sg11[C3(n3 - 1 - nb, k, i)];
          sg22[C3(k, j, i)] = 1. / rho_air;
          sg22[C3(n3 - 1 - k, j, i)] = sg22[C3(n3 - 1 - nb, k, i)];
          rh33[C3(k, j, i)] = rho_air;
          rh33[C3(n3 - 1 - k, j, i)] = rh33[C3(n3 - 1 - nb, k, i)];
        }
  }
  /* edge averaging, extend_model.c:180-252 */
  m->s11 = dalloc((long)n1 * (n2 + 1) * (n3 + 1));
  m->s22 = dalloc((long)(n1 + 1) * n2 * (n3 + 1));
  m->s33 = dalloc((long)(n1 + 1) * (n2 + 1) * n3);
#define S1(k, j, i) m->s11[(i) + (long)n1 * ((j) + (long)(n2 + 1) * (k))]
#define S2(k, j, i) m->s22[(i) + (long)(n1 + 1) * ((j) + (long)n2 * (k))]
#define S3(k, j, i) m->s33[(i) + (long)(n1 + 1) * ((j) + (long)(n2 + 1) * (k))]
  double *d1 = m->d1, *d2 = m->d2, *d3 = m->d3, *d1s = m->d1s, *d2s = m->d2s, *d3s = m->d3s;
  for (k = 1; k < n3; k++) {
    km1 = k - 1;
    for (j = 1; j < n2; j++) {
      jm1 = j - 1;
      for (i = 0; i < n1; i++)
        S1(k, j, i) = 0.25 * ((d2s[j] * sg11[C3(k, j, i)] + d2s[jm1] * sg11[C3(k, jm1, i)]) * d3s[k] +
                              (d2s[j] * sg11[C3(km1, j, i)] + d2s[jm1] * sg11[C3(km1, jm1, i)]) * d3s[km1]) /
                      (d2[j] * d3[k]);
    }
  }
  for (j = 0; j <= n2; j++)
    for (i = 0; i < n1; i++) {
      S1(0, j, i) = S1(1, j, i);
      S1(n3, j, i) = S1(n3 - 1, j, i);
    }
  for (k = 0; k <= n3; k++)
    for (i = 0; i < n1; i++) {
      S1(k, 0, i) = S1(k, 1, i);
      S1(k, n2, i) = S1(k, n2 - 1, i);
    }
  for (k = 1; k < n3; k++) {
    km1 = k - 1;
    for (j = 0; j < n2; j++)
      for (i = 1; i < n1; i++) {
        im1 = i - 1;
        S2(k, j, i) = 0.25 * ((d1s[i] * sg22[C3(k, j, i)] + d1s[im1] * sg22[C3(k, j, im1)]) * d3s[k] +
                              (d1s[i] * sg22[C3(km1, j, i)] + d1s[im1] * sg22[C3(km1, j, im1)]) * d3s[km1]) /
                      (d1[i] * d3[k]);
      }
  }
  for (j = 0; j < n2; j++)
    for (i = 0; i <= n1; i++) {
      S2(0, j, i) = S2(1, j, i);
      S2(n3, j, i) = S2(n3 - 1, j, i);
    }
  for (k = 0; k <= n3; k++)
    for (j = 0; j < n2; j++) {
      S2(k, j, 0) = S2(k, j, 1);
      S2(k, j, n1) = S2(k, j, n1 - 1);
    }
  for (k = 0; k < n3; k++)
    for (j = 1; j < n2; j++) {
      jm1 = j - 1;
      for (i = 1; i < n1; i++) {
        im1 = i - 1;
        S3(k, j, i) = 0.25 * ((d1s[i] * rh33[C3(k, j, i)] + d1s[im1] * rh33[C3(k, j, im1)]) * d2s[j] +
                              (d1s[i] * rh33[C3(k, jm1, i)] + d1s[im1] * rh33[C3(k, jm1, im1)]) * d2s[jm1]) /
                      (d1[i] * d2[j]);
        S3(k, j, i) = 1. / S3(k, j, i);
      }
    }
  for (k = 0; k < n3; k++)
    for (j = 0; j <= n2; j++) {
      S3(k, j, 0) = S3(k, j, 1);
      S3(k, j, n1) = S3(k, j, n1 - 1);
    }
  for (k = 0; k < n3; k++)
    for (i = 0; i <= n1; i++) {
      S3(k, 0, i) = S3(k, 1, i);
      S3(k, n2, i) = S3(k, n2 - 1, i);
    }
#undef C3
#undef R3
  m->c11 = dalloc(nc); m->c22 = dalloc(nc); m->c33 = dalloc(nc);
  for (c = 0; c < nc; c++) {
    m->c11[c] = 1. / sg11[c];
    m->c22[c] = 1. / sg22[c];
    m->c33[c] = rh33[c];
  }
  free(sg11); free(sg22); free(rh33);

  /* fdfd_init (fdfd.c:40-62) */
  m->fEx = (long)n1 * (n2 + 1) * (n3 + 1);
  m->fEy = (long)(n1 + 1) * n2 * (n3 + 1);
  m->fEz = (long)(n1 + 1) * (n2 + 1) * n3;
  m->fE = m->fEx + m->fEy + m->fEz;
  m->n3 = airbc ? nz + nb : n3;
  {
    int o3 = m->n3;
    m->lEx = (long)n1 * (n2 + 1) * (o3 + 1);
    m->lEy = (long)(n1 + 1) * n2 * (o3 + 1);
    m->lEz = (long)(n1 + 1) * (n2 + 1) * o3;
    m->lE = m->lEx + m->lEy + m->lEz;
    m->lHx = (long)(n1 + 1) * n2 * o3;
    m->lHy = (long)n1 * (n2 + 1) * n2; /* fdfd.c:59 uses n2 for the k extent */
    m->lHz = (long)n1 * n2 * (o3 + 1);
    m->lH = m->lHx + m->lHy + m->lHz;
    /* The Hy block is indexed up to n1*(n2+1)*o3.  For o3 <= n2 (every
     * model the tests use) Hz sits where the reference puts it; for o3 > n2
     * the reference's Hy writes run into Hz and past the allocation, and
     * here Hz is moved past the real Hy extent instead. */
    long hy = MAXI(m->lHy, (long)n1 * (n2 + 1) * o3);
    m->hz_off = m->lHx + hy;
    m->xh = zalloc(m->lHx + hy + m->lHz);
  }
  if (airbc) {
    m->xx = zalloc(m->fE);
    m->yy = zalloc(m->fE);
    m->Hxg = zalloc((long)(n1 + 1) * n2);
    m->Hyg = zalloc((long)n1 * (n2 + 1));
    m->d3fft = x3node[1] - x3node[0];
    fft_init(m);
  }
  orc_set_freq(m, 1.0f);
  return m;
}

void orc_free(orc_t *m)
{
  if (!m) return;
  free(m->x1); free(m->x2); free(m->x3); free(m->x1s); free(m->x2s); free(m->x3s);
  free(m->d1); free(m->d2); free(m->d3); free(m->d1s); free(m->d2s); free(m->d3s);
  free(m->s11); free(m->s22); free(m->s33);
  free(m->c11); free(m->c22); free(m->c33);
  free(m->xh); free(m->xx); free(m->yy); free(m->Hxg); free(m->Hyg);
  free(m->av); free(m->abv);
  free(m->sH1); free(m->sH2); free(m->buf); free(m->buf0); free(m->buf2); free(m->tw); free(m->tmp);
  free(m->jx); free(m->jy); free(m->wx); free(m->wy);
  free(m);
}

long orc_info(const orc_t *m, int what)
{
  switch (what) {
  case 0: return m->n1;
  case 1: return m->n2;
  case 2: return m->n3;
  case 3: return m->n3f;
  case 4: return m->lEx;
  case 5: return m->lEy;
  case 6: return m->lEz;
  case 7: return m->lE;
  case 8: return m->fEx;
  case 9: return m->fEy;
  case 10: return m->fEz;
  case 11: return m->fE;
  case 12: return m->nfft;
  case 13: return m->lHx;
  case 14: return m->lHy;
  case 15: return m->lHz;
  case 16: return m->hz_off;
  }
  return -1;
}

double orc_finfo(const orc_t *m, int what)
{
  switch (what) {
  case 0: return m->rho_water;
  case 1: return m->lextend;
  case 2: case 3: case 4: case 5: case 6: case 7: return m->rstretch[what - 2];
  case 8: return m->d1fft;
  case 9: return m->d2fft;
  case 10: return m->d3fft;
  }
  return 0;
}

void *orc_array(const orc_t *m, const char *s)
{
  if (!strcmp(s, "x1")) return m->x1;
  if (!strcmp(s, "x2")) return m->x2;
  if (!strcmp(s, "x3")) return m->x3;
  if (!strcmp(s, "x1s")) return m->x1s;
  if (!strcmp(s, "x2s")) return m->x2s;
  if (!strcmp(s, "x3s")) return m->x3s;
  if (!strcmp(s, "d1")) return m->d1;
  if (!strcmp(s, "d2")) return m->d2;
  if (!strcmp(s, "d3")) return m->d3;
  if (!strcmp(s, "d1s")) return m->d1s;
  if (!strcmp(s, "d2s")) return m->d2s;
  if (!strcmp(s, "d3s")) return m->d3s;
  if (!strcmp(s, "sigma11")) return m->s11;
  if (!strcmp(s, "sigma22")) return m->s22;
  if (!strcmp(s, "sigma33")) return m->s33;
  if (!strcmp(s, "cell11")) return m->c11;
  if (!strcmp(s, "cell22")) return m->c22;
  if (!strcmp(s, "cell33")) return m->c33;
  if (!strcmp(s, "sH1")) return m->sH1;
  if (!strcmp(s, "sH2")) return m->sH2;
  if (!strcmp(s, "xh")) return m->xh;
  if (!strcmp(s, "Hxg")) return m->Hxg;
  if (!strcmp(s, "Hyg")) return m->Hyg;
  if (!strcmp(s, "wx")) return m->wx;
  if (!strcmp(s, "wy")) return m->wy;
  if (!strcmp(s, "jx")) return m->jx;
  if (!strcmp(s, "jy")) return m->jy;
  return NULL;
}

void orc_set_red(orc_t *m, int red) { m->red = red; }

void orc_set_freq(orc_t *m, float freq)
{
  m->iomu = I * 2. * ORC_PI * freq * ORC_MU0; /* main.c:95 */
  m->iwmu = 1. / m->iomu;                     /* fdfd.c:100 */
}

/* --------------------------------------------------------------- operator */

#define id1(k, j, i) ((i) + (long)n1 * ((j) + (long)(n2 + 1) * (k)))
#define id2(k, j, i) ((i) + (long)(n1 + 1) * ((j) + (long)n2 * (k)))
#define id3(k, j, i) ((i) + (long)(n1 + 1) * ((j) + (long)(n2 + 1) * (k)))
#define idx(k, j, i) ((i) + (long)(n1 + 1) * ((j) + (long)n2 * (k)))
#define idy(k, j, i) ((i) + (long)n1 * ((j) + (long)(n2 + 1) * (k)))
#define idz(k, j, i) ((i) + (long)n1 * ((j) + (long)n2 * (k)))
#define idHx(j, i) ((i) + (long)(n1 + 1) * (j))
#define idHy(j, i) ((i) + (long)n1 * (j))

/* fdfd_apply, fdfd.c:86-205 */
void orc_fdfd_apply(orc_t *m, const zc *x, zc *y)
{
  int n1 = m->n1, n2 = m->n2, n3 = m->n3, i, j, k, kstart;
  int ip1, jp1, kp1, im1, jm1, km1;
  zc t1, t2, iwmu = m->iwmu;
  kstart = m->airbc ? m->nb : 0;
  const double *d1 = m->d1, *d2 = m->d2, *d3 = &m->d3[kstart];
  const double *d1s = m->d1s, *d2s = m->d2s, *d3s = &m->d3s[kstart];
  const double *sigma11 = &m->s11[(long)kstart * n1 * (n2 + 1)];
  const double *sigma22 = &m->s22[(long)kstart * (n1 + 1) * n2];
  const double *sigma33 = &m->s33[(long)kstart * (n1 + 1) * (n2 + 1)];
  const zc *Ex = x, *Ey = x + m->lEx, *Ez = x + m->lEx + m->lEy;
  zc *Hx = m->xh, *Hy = m->xh + m->lHx, *Hz = m->xh + m->hz_off;
  zc *bx = y, *by = y + m->lEx, *bz = y + m->lEx + m->lEy;
  memset(y, 0, m->lE * sizeof(zc));

  for (k = 0; k <= n3; k++) {
    kp1 = k + 1;
    for (j = 0; j <= n2; j++) {
      jp1 = j + 1;
      for (i = 0; i <= n1; i++) {
        ip1 = i + 1;
        if (j < n2 && k < n3) {
          t1 = (Ez[id3(k, jp1, i)] - Ez[id3(k, j, i)]) / d2s[j];
          t2 = (Ey[id2(kp1, j, i)] - Ey[id2(k, j, i)]) / d3s[k];
          Hx[idx(k, j, i)] = (t1 - t2) * iwmu;
        }
        if (i < n1 && k < n3) {
          t1 = (Ex[id1(kp1, j, i)] - Ex[id1(k, j, i)]) / d3s[k];
          t2 = (Ez[id3(k, j, ip1)] - Ez[id3(k, j, i)]) / d1s[i];
          Hy[idy(k, j, i)] = (t1 - t2) * iwmu;
        }
        if (i < n1 && j < n2) {
          t1 = (Ey[id2(k, j, ip1)] - Ey[id2(k, j, i)]) / d1s[i];
          t2 = (Ex[id1(k, jp1, i)] - Ex[id1(k, j, i)]) / d2s[j];
          Hz[idz(k, j, i)] = (t1 - t2) * iwmu;
        }
      }
    }
  }

  if (m->airbc) {
    zc *Hxkm1 = m->Hxg, *Hykm1 = m->Hyg;
    orc_airbc_apply(m, Hz, Hxkm1, Hykm1);
    k = 0;
    for (j = 0; j < n2; j++) {
      jm1 = j - 1;
      for (i = 0; i < n1; i++) {
        im1 = i - 1;
        if (j > 0) {
          t1 = (Hz[idz(k, j, i)] - Hz[idz(k, jm1, i)]) / d2[j];
          t2 = (Hy[idy(k, j, i)] - Hykm1[idHy(j, i)]) / d3[k];
          bx[id1(k, j, i)] = t1 - t2 - sigma11[id1(k, j, i)] * Ex[id1(k, j, i)];
          if (m->fvm) bx[id1(k, j, i)] *= d1s[i] * d2[j] * d3[k] * m->iomu;
        }
        if (i > 0) {
          t1 = (Hx[idx(k, j, i)] - Hxkm1[idHx(j, i)]) / d3[k];
          t2 = (Hz[idz(k, j, i)] - Hz[idz(k, j, im1)]) / d1[i];
          by[id2(k, j, i)] = t1 - t2 - sigma22[id2(k, j, i)] * Ey[id2(k, j, i)];
          if (m->fvm) by[id2(k, j, i)] *= d1[i] * d2s[j] * d3[k] * m->iomu;
        }
      }
    }
  }

  for (k = 0; k < n3; k++) {
    km1 = k - 1;
    for (j = 0; j < n2; j++) {
      jm1 = j - 1;
      for (i = 0; i < n1; i++) {
        im1 = i - 1;
        if (j > 0 && k > 0) {
          t1 = (Hz[idz(k, j, i)] - Hz[idz(k, jm1, i)]) / d2[j];
          t2 = (Hy[idy(k, j, i)] - Hy[idy(km1, j, i)]) / d3[k];
          bx[id1(k, j, i)] = t1 - t2 - sigma11[id1(k, j, i)] * Ex[id1(k, j, i)];
          if (m->fvm) bx[id1(k, j, i)] *= d1s[i] * d2[j] * d3[k] * m->iomu;
        }
        if (i > 0 && k > 0) {
          t1 = (Hx[idx(k, j, i)] - Hx[idx(km1, j, i)]) / d3[k];
          t2 = (Hz[idz(k, j, i)] - Hz[idz(k, j, im1)]) / d1[i];
          by[id2(k, j, i)] = t1 - t2 - sigma22[id2(k, j, i)] * Ey[id2(k, j, i)];
          if (m->fvm) by[id2(k, j, i)] *= d1[i] * d2s[j] * d3[k] * m->iomu;
        }
        if (i > 0 && j > 0) {
          t1 = (Hy[idy(k, j, i)] - Hy[idy(k, j, im1)]) / d1[i];
          t2 = (Hx[idx(k, j, i)] - Hx[idx(k, jm1, i)]) / d2[j];
          bz[id3(k, j, i)] = t1 - t2 - sigma33[id3(k, j, i)] * Ez[id3(k, j, i)];
          if (m->fvm) bz[id3(k, j, i)] *= d1[i] * d2[j] * d3s[k] * m->iomu;
        }
      }
    }
  }
}

/* ------------------------------------------------------------- LU-SGS */

/* lusgs_apply, fdfd.c:287-510 */
void orc_lusgs_apply(orc_t *m, const zc *x, zc *y)
{
  int n1 = m->n1, n2 = m->n2, n3 = m->n3, i, j, k;
  int ip1, jp1, kp1, im1, jm1, km1;
  long lEx, lEy, lEz, lE;
  double vol;
  zc *Ex, *Ey, *Ez;
  const zc *bx, *by, *bz;
  zc t1, t2, t3, t4, aij_xj, aii, iwmu = m->iwmu;
  const double *d1 = m->d1, *d2 = m->d2, *d3 = m->d3, *d1s = m->d1s, *d2s = m->d2s, *d3s = m->d3s;
  const double *sigma11 = m->s11, *sigma22 = m->s22, *sigma33 = m->s33;
  if (m->airbc) {
    n3 += m->nb;
    lEx = (long)n1 * (n2 + 1) * (n3 + 1);
    lEy = (long)(n1 + 1) * n2 * (n3 + 1);
    lEz = (long)(n1 + 1) * (n2 + 1) * n3;
    lE = lEx + lEy + lEz;
    Ex = m->yy; Ey = m->yy + lEx; Ez = m->yy + lEx + lEy;
    memset(m->xx, 0, lE * sizeof(zc));
    memset(m->yy, 0, lE * sizeof(zc));
    memcpy(&m->xx[(long)n1 * (n2 + 1) * m->nb], x, m->lEx * sizeof(zc));
    memcpy(&m->xx[(long)(n1 + 1) * n2 * m->nb + lEx], &x[m->lEx], m->lEy * sizeof(zc));
    memcpy(&m->xx[(long)(n1 + 1) * (n2 + 1) * m->nb + lEx + lEy], &x[m->lEx + m->lEy], m->lEz * sizeof(zc));
    bx = m->xx; by = m->xx + lEx; bz = m->xx + lEx + lEy;
  } else {
    lEx = m->lEx; lEy = m->lEy;
    Ex = y; Ey = y + m->lEx; Ez = y + m->lEx + m->lEy;
    bx = x; by = x + m->lEx; bz = x + m->lEx + m->lEy;
    memset(y, 0, m->lE * sizeof(zc));
  }

#define FVM_SCALE()                                  \
  if (m->fvm) {                                      \
    vol = d1s[i] * d2[j] * d3[k];                    \
    aij_xj *= vol * m->iomu;                         \
    aii *= vol * m->iomu;                            \
  }
#define CELL_EX()                                                                                           \
  ip1 = MINI(i + 1, n1); im1 = MAXI(i - 1, 0);                                                              \
  Ex[id1(k, j, i)] = 0;                                                                                     \
  t1 = (Ey[id2(k, j, ip1)] - Ey[id2(k, j, i)]) / d1s[i] - (Ex[id1(k, jp1, i)] - Ex[id1(k, j, i)]) / d2s[j]; \
  t2 = (Ey[id2(k, jm1, ip1)] - Ey[id2(k, jm1, i)]) / d1s[i] -                                               \
       (Ex[id1(k, j, i)] - Ex[id1(k, jm1, i)]) / d2s[jm1];                                                  \
  t3 = (Ex[id1(kp1, j, i)] - Ex[id1(k, j, i)]) / d3s[k] - (Ez[id3(k, j, ip1)] - Ez[id3(k, j, i)]) / d1s[i]; \
  t4 = (Ex[id1(k, j, i)] - Ex[id1(km1, j, i)]) / d3s[km1] -                                                 \
       (Ez[id3(km1, j, ip1)] - Ez[id3(km1, j, i)]) / d1s[i];                                                \
  aij_xj = ((t1 - t2) / d2[j] - (t3 - t4) / d3[k]) * iwmu - sigma11[id1(k, j, i)] * Ex[id1(k, j, i)];       \
  aii = ((1. / d2s[j] + 1. / d2s[jm1]) / d2[j] + (1. / d3s[k] + 1. / d3s[km1]) / d3[k]) * iwmu -            \
        sigma11[id1(k, j, i)];                                                                              \
  FVM_SCALE();                                                                                              \
  Ex[id1(k, j, i)] = (bx[id1(k, j, i)] - aij_xj) / aii;
#define CELL_EY()                                                                                           \
  ip1 = MINI(i + 1, n1); im1 = MAXI(i - 1, 0);                                                              \
  Ey[id2(k, j, i)] = 0;                                                                                     \
  t1 = (Ez[id3(k, jp1, i)] - Ez[id3(k, j, i)]) / d2s[j] - (Ey[id2(kp1, j, i)] - Ey[id2(k, j, i)]) / d3s[k]; \
  t2 = (Ez[id3(km1, jp1, i)] - Ez[id3(km1, j, i)]) / d2s[j] -                                               \
       (Ey[id2(k, j, i)] - Ey[id2(km1, j, i)]) / d3s[km1];                                                  \
  t3 = (Ey[id2(k, j, ip1)] - Ey[id2(k, j, i)]) / d1s[i] - (Ex[id1(k, jp1, i)] - Ex[id1(k, j, i)]) / d2s[j]; \
  t4 = (Ey[id2(k, j, i)] - Ey[id2(k, j, im1)]) / d1s[im1] -                                                 \
       (Ex[id1(k, jp1, im1)] - Ex[id1(k, j, im1)]) / d2s[j];                                                \
  aij_xj = ((t1 - t2) / d3[k] - (t3 - t4) / d1[i]) * iwmu - sigma22[id2(k, j, i)] * Ey[id2(k, j, i)];       \
  aii = ((1. / d3s[k] + 1. / d3s[km1]) / d3[k] + (1. / d1s[i] + 1. / d1s[im1]) / d1[i]) * iwmu -            \
        sigma22[id2(k, j, i)];                                                                              \
  FVM_SCALE();                                                                                              \
  Ey[id2(k, j, i)] = (by[id2(k, j, i)] - aij_xj) / aii;
#define CELL_EZ()                                                                                           \
  ip1 = MINI(i + 1, n1); im1 = MAXI(i - 1, 0);                                                              \
  Ez[id3(k, j, i)] = 0;                                                                                     \
  t1 = (Ex[id1(kp1, j, i)] - Ex[id1(k, j, i)]) / d3s[k] - (Ez[id3(k, j, ip1)] - Ez[id3(k, j, i)]) / d1s[i]; \
  t2 = (Ex[id1(kp1, j, im1)] - Ex[id1(k, j, im1)]) / d3s[k] -                                               \
       (Ez[id3(k, j, i)] - Ez[id3(k, j, im1)]) / d1s[im1];                                                  \
  t3 = (Ez[id3(k, jp1, i)] - Ez[id3(k, j, i)]) / d2s[j] - (Ey[id2(kp1, j, i)] - Ey[id2(k, j, i)]) / d3s[k]; \
  t4 = (Ez[id3(k, j, i)] - Ez[id3(k, jm1, i)]) / d2s[jm1] -                                                 \
       (Ey[id2(kp1, jm1, i)] - Ey[id2(k, jm1, i)]) / d3s[k];                                                \
  aij_xj = ((t1 - t2) / d1[i] - (t3 - t4) / d2[j]) * iwmu - sigma33[id3(k, j, i)] * Ez[id3(k, j, i)];       \
  aii = ((1. / d1s[i] + 1. / d1s[im1]) / d1[i] + (1. / d2s[j] + 1. / d2s[jm1]) / d2[j]) * iwmu -            \
        sigma33[id3(k, j, i)];                                                                              \
  FVM_SCALE();                                                                                              \
  Ez[id3(k, j, i)] = (bz[id3(k, j, i)] - aij_xj) / aii;
#define KJ()                                 \
  kp1 = MINI(k + 1, n3); km1 = MAXI(k - 1, 0); \
  jp1 = MINI(j + 1, n2); jm1 = MAXI(j - 1, 0);

  /* forward: fdfd.c:347-424 */
  for (k = 1; k < n3; k++)
    for (j = 1; j < n2; j++)
      for (i = 0; i < n1; i++) { KJ(); CELL_EX(); }
  for (k = 1; k < n3; k++)
    for (j = 0; j < n2; j++)
      for (i = 1; i < n1; i++) { KJ(); CELL_EY(); }
  for (k = 0; k < n3; k++)
    for (j = 1; j < n2; j++)
      for (i = 1; i < n1; i++) { KJ(); CELL_EZ(); }
  /* backward: fdfd.c:426-503 */
  for (k = n3 - 1; k >= 0; k--)
    for (j = n2 - 1; j > 0; j--)
      for (i = n1 - 1; i > 0; i--) { KJ(); CELL_EZ(); }
  for (k = n3 - 1; k > 0; k--)
    for (j = n2 - 1; j >= 0; j--)
      for (i = n1 - 1; i > 0; i--) { KJ(); CELL_EY(); }
  for (k = n3 - 1; k > 0; k--)
    for (j = n2 - 1; j > 0; j--)
      for (i = n1 - 1; i >= 0; i--) { KJ(); CELL_EX(); }
#undef KJ
#undef CELL_EX
#undef CELL_EY
#undef CELL_EZ
#undef FVM_SCALE

  if (m->airbc) { /* fdfd.c:505-509 */
    memcpy(y, &m->yy[(long)n1 * (n2 + 1) * m->nb], m->lEx * sizeof(zc));
    memcpy(&y[m->lEx], &m->yy[(long)(n1 + 1) * n2 * m->nb + lEx], m->lEy * sizeof(zc));
    memcpy(&y[m->lEx + m->lEy], &m->yy[(long)(n1 + 1) * (n2 + 1) * m->nb + lEx + lEy], m->lEz * sizeof(zc));
  }
}

/* ------------------------------------------------------------ airwave FFT */

/* Twiddles w_t = exp(-2*pi*i*t/N), t = 0..N-1.  The inverse uses conj. */
static void fft_init(orc_t *m)
{
  int N = m->nfft, n1 = m->n1, n2 = m->n2, i1, i2, j1, j2;
  double dkx, dky, kx, ky, kz, x1, x2;
  long NN = (long)N * N;
  m->tw = zalloc(N);
  for (int t = 0; t < N; t++) {
    double a = -2.0 * ORC_PI * (double)t / (double)N;
    m->tw[t] = mk(cos(a), sin(a));
  }
  m->tmp = zalloc(2 * N);
  m->buf = zalloc(NN);
  m->buf0 = zalloc(NN);
  m->buf2 = zalloc(NN);
  m->sH1 = zalloc(NN);
  m->sH2 = zalloc(NN);
  m->jx = (int *)calloc(N, sizeof(int));
  m->jy = (int *)calloc(N, sizeof(int));
  m->wx = dalloc(N);
  m->wy = dalloc(N);

  /* airbc_init_free.c:32-34, 69-105 */
  m->d1fft = (m->x1[n1] - m->x1[0]) / (N - 1);
  m->d2fft = (m->x2[n2] - m->x2[0]) / (N - 1);
  dkx = 2.0 * ORC_PI / (m->d1fft * N);
  dky = 2.0 * ORC_PI / (m->d2fft * N);
  for (i2 = 0; i2 < N; i2++) {
    ky = (i2 <= N / 2) ? i2 * dky : (i2 - N) * dky;
    for (i1 = 0; i1 < N; i1++) {
      kx = (i1 <= N / 2) ? i1 * dkx : (i1 - N) * dkx;
      kz = sqrt(kx * kx + ky * ky);
      zc a = exp(-kz * 0.5 * m->d3fft) * I * kx / (kz + 1.e-15);
      zc b = exp(-kz * 0.5 * m->d3fft) * I * ky / (kz + 1.e-15);
      a *= cexp(-I * kx * 0.5 * m->d1s[0]);
      b *= cexp(-I * ky * 0.5 * m->d2s[0]);
      m->sH1[i1 + (long)N * i2] = a;
      m->sH2[i1 + (long)N * i2] = b;
    }
  }
  for (i1 = 0; i1 < N; i1++) {
    x1 = m->x1s[0] + i1 * m->d1fft;
    if (x1 < m->x1s[n1 - 1]) {
      j1 = orc_find_index(n1 + 1, m->x1s, x1);
      m->jx[i1] = j1;
      m->wx[i1] = (x1 - m->x1s[j1]) / (m->x1s[j1 + 1] - m->x1s[j1]);
    }
  }
  for (i2 = 0; i2 < N; i2++) {
    x2 = m->x2s[0] + i2 * m->d2fft;
    if (x2 < m->x2s[n2 - 1]) {
      j2 = orc_find_index(n2 + 1, m->x2s, x2);
      m->jy[i2] = j2;
      m->wy[i2] = (x2 - m->x2s[j2]) / (m->x2s[j2 + 1] - m->x2s[j2]);
    }
  }
}

static inline zc tw_mul(zc v, zc w)
{
  double a = creal(v), b = cimag(v), c = creal(w), d = cimag(w);
  double re = a * c - b * d, im = a * d + b * c;
  return mk(re, im);
}

/* One length-N transform of v (contiguous), Stockham autosort: radix-4
 * passes with Ns = 1, 4, 16, ... and a final radix-2 pass when log2(N) is
 * odd.  Pass with radix R and span Ns: for j in [0, N/R): k = j % Ns,
 * inputs v[j + r*N/R] times w^(r*k*N/(R*Ns)) (no multiply when r*k == 0),
 * R-point DFT, outputs to (j/Ns)*R*Ns + k + r*Ns. */
static void fft1(const orc_t *m, zc *v, int dir)
{
  int N = m->nfft, Ns, j, k, r;
  zc *a = v, *b = m->tmp, *t;
  for (Ns = 1; Ns < N;) {
    int R = (N / Ns) % 4 == 0 ? 4 : 2;
    int stride = N / R;
    for (j = 0; j < N / R; j++) {
      zc u[4];
      k = j % Ns;
      for (r = 0; r < R; r++) {
        u[r] = a[j + r * stride];
        if (r != 0 && k != 0) {
          zc w = m->tw[(long)r * k * (N / (R * Ns))];
          if (dir > 0) w = mk(creal(w), -cimag(w));
          u[r] = tw_mul(u[r], w);
        }
      }
      int o = (j / Ns) * R * Ns + k;
      if (R == 4) {
        double a0r = creal(u[0]) + creal(u[2]), a0i = cimag(u[0]) + cimag(u[2]);
        double a1r = creal(u[0]) - creal(u[2]), a1i = cimag(u[0]) - cimag(u[2]);
        double a2r = creal(u[1]) + creal(u[3]), a2i = cimag(u[1]) + cimag(u[3]);
        double a3r = creal(u[1]) - creal(u[3]), a3i = cimag(u[1]) - cimag(u[3]);
        /* forward multiplies a3 by -i, inverse by +i */
        double br_ = dir < 0 ? a3i : -a3i, bi_ = dir < 0 ? -a3r : a3r;
        b[o] = mk(a0r + a2r, a0i + a2i);
        b[o + Ns] = mk(a1r + br_, a1i + bi_);
        b[o + 2 * Ns] = mk(a0r - a2r, a0i - a2i);
        b[o + 3 * Ns] = mk(a1r - br_, a1i - bi_);
      } else {
        b[o] = mk(creal(u[0]) + creal(u[1]), cimag(u[0]) + cimag(u[1]));
        b[o + Ns] = mk(creal(u[0]) - creal(u[1]), cimag(u[0]) - cimag(u[1]));
      }
    }
    t = a; a = b; b = t;
    Ns *= R;
  }
  if (a != v) memcpy(v, a, N * sizeof(zc));
}

static void fft_rows(orc_t *m, zc *d, int dir)
{
  int N = m->nfft;
  for (int i2 = 0; i2 < N; i2++) fft1(m, d + (long)N * i2, dir);
}

static void fft_cols(orc_t *m, zc *d, int dir)
{
  int N = m->nfft;
  zc *col = (zc *)malloc(N * sizeof(zc));
  for (int i1 = 0; i1 < N; i1++) {
    for (int i2 = 0; i2 < N; i2++) col[i2] = d[i1 + (long)N * i2];
    fft1(m, col, dir);
    for (int i2 = 0; i2 < N; i2++) d[i1 + (long)N * i2] = col[i2];
  }
  free(col);
}

void orc_fft2(orc_t *m, zc *data, int dir)
{
  if (dir < 0) {
    fft_rows(m, data, -1);
    fft_cols(m, data, -1);
  } else {
    fft_cols(m, data, +1);
    fft_rows(m, data, +1);
  }
}

/* airbc_apply, airbc_init_free.c:130-209, in its four stages.  Each stage
 * is exported on its own (orc_airbc_stage) so that tests can check it
 * against what the mathematics fixes; orc_airbc_apply is their
 * composition, in the reference's order of operations. */

/* Stage 0 (airbc_init_free.c:137-152): Hz(z=0) at (x1s[j1], x2s[j2]),
 * j1 < n1, j2 < n2, bilinearly onto the uniform N x N grid
 * (x1s[0] + i1*d1fft, x2s[0] + i2*d2fft), zero outside the hull. */
static void air_forward_map(const orc_t *m, const zc *H3k, zc *U)
{
  int N = m->nfft, n1 = m->n1, n2 = m->n2, i1, i2, j1, j2;
  double x1, x2, w1, w2;
  for (i2 = 0; i2 < N; i2++) {
    x2 = m->x2s[0] + i2 * m->d2fft;
    for (i1 = 0; i1 < N; i1++) {
      x1 = m->x1s[0] + i1 * m->d1fft;
      if (x1 < m->x1s[n1 - 1] && x2 < m->x2s[n2 - 1]) {
        j1 = m->jx[i1]; j2 = m->jy[i2];
        w1 = m->wx[i1]; w2 = m->wy[i2];
        U[i1 + (long)N * i2] = (1. - w1) * (1. - w2) * H3k[j1 + (long)n1 * j2] +
                               w1 * (1. - w2) * H3k[j1 + 1 + (long)n1 * j2] +
                               (1. - w1) * w2 * H3k[j1 + (long)n1 * (j2 + 1)] +
                               w1 * w2 * H3k[j1 + 1 + (long)n1 * (j2 + 1)];
      } else
        U[i1 + (long)N * i2] = 0.;
    }
  }
}

/* Stage 1 (airbc_init_free.c:153-161, 182-188): forward 2-D DFT, times sH1
 * (resp. sH2), inverse DFT, divided by N*N.  U is overwritten by its
 * spectrum. */
static void air_spectral(orc_t *m, zc *U, zc *Hxu, zc *Hyu)
{
  int N = m->nfft;
  long NN = (long)N * N, q;
  orc_fft2(m, U, -1);
  for (q = 0; q < NN; q++) Hxu[q] = U[q] * m->sH1[q];
  orc_fft2(m, Hxu, +1);
  for (q = 0; q < NN; q++) Hxu[q] /= (N * N);
  for (q = 0; q < NN; q++) Hyu[q] = U[q] * m->sH2[q];
  orc_fft2(m, Hyu, +1);
  for (q = 0; q < NN; q++) Hyu[q] /= (N * N);
}

/* Stage 2 (airbc_init_free.c:162-180): ghost Hx at (x1[i1], x2s[i2]) from
 * the uniform grid, cell index (x - origin)/d truncated.  The column
 * i1 = n1 is never consumed by the operator (its k=0 Ey rows stop at
 * i1 = n1-1) and is set to zero: the reference reads it across the end of
 * a uniform row. */
static void air_back_hx(const orc_t *m, const zc *U, zc *H1km1)
{
  int N = m->nfft, n1 = m->n1, n2 = m->n2, i1, i2, j1, j2;
  double w1, w2, s1, s2;
  for (i2 = 0; i2 < n2; i2++) {
    s2 = (m->x2s[i2] - m->x2s[0]);
    j2 = s2 / m->d2fft;
    w2 = s2 / m->d2fft - j2;
    for (i1 = 0; i1 < n1; i1++) {
      s1 = (m->x1[i1] - m->x1[0]);
      j1 = s1 / m->d1fft;
      w1 = s1 / m->d1fft - j1;
      H1km1[i1 + (long)(n1 + 1) * i2] = (1. - w1) * (1. - w2) * U[j1 + (long)N * j2] +
                                        w1 * (1. - w2) * U[j1 + 1 + (long)N * j2] +
                                        (1. - w1) * w2 * U[j1 + (long)N * (j2 + 1)] +
                                        w1 * w2 * U[j1 + 1 + (long)N * (j2 + 1)];
    }
    H1km1[n1 + (long)(n1 + 1) * i2] = 0.;
  }
}

/* Stage 3 (airbc_init_free.c:189-207): ghost Hy at (x1s[i1], x2[i2]).  The
 * row i2 = n2 is never consumed (the k=0 Ex rows stop at i2 = n2-1) and
 * is set to zero: the reference reads it past the end of the buffer. */
static void air_back_hy(const orc_t *m, const zc *U, zc *H2km1)
{
  int N = m->nfft, n1 = m->n1, n2 = m->n2, i1, i2, j1, j2;
  double w1, w2, s1, s2;
  for (i2 = 0; i2 < n2; i2++) {
    s2 = (m->x2[i2] - m->x2[0]);
    j2 = s2 / m->d2fft;
    w2 = s2 / m->d2fft - j2;
    for (i1 = 0; i1 < n1; i1++) {
      s1 = (m->x1s[i1] - m->x1s[0]);
      j1 = s1 / m->d1fft;
      w1 = s1 / m->d1fft - j1;
      H2km1[i1 + (long)n1 * i2] = (1. - w1) * (1. - w2) * U[j1 + (long)N * j2] +
                                  w1 * (1. - w2) * U[j1 + 1 + (long)N * j2] +
                                  (1. - w1) * w2 * U[j1 + (long)N * (j2 + 1)] +
                                  w1 * w2 * U[j1 + 1 + (long)N * (j2 + 1)];
    }
  }
  for (i1 = 0; i1 < n1; i1++) H2km1[i1 + (long)n1 * n2] = 0.;
}

void orc_airbc_stage(orc_t *m, int stage, const zc *in, zc *out, zc *out2)
{
  long NN = (long)m->nfft * m->nfft;
  switch (stage) {
  case 0: air_forward_map(m, in, out); break;
  case 1:
    memcpy(m->buf0, in, NN * sizeof(zc));
    air_spectral(m, m->buf0, out, out2);
    break;
  case 2: air_back_hx(m, in, out); break;
  case 3: air_back_hy(m, in, out); break;
  }
}

void orc_airbc_apply(orc_t *m, const zc *H3k, zc *H1km1, zc *H2km1)
{
  air_forward_map(m, H3k, m->buf0);
  air_spectral(m, m->buf0, m->buf, m->buf2);
  air_back_hx(m, m->buf, H1km1);
  air_back_hy(m, m->buf2, H2km1);
}

/* ------------------------------------------------------------ reductions */

#define RED_T 256
#define RED_E 16

static inline void dot_term(zc a, zc b, double *re, double *im)
{
  /* a*conj(b) as gcc forms it: (ac - bd, ad + bc) with (c, d) = (b.re, -b.im) */
  double ar = creal(a), ai = cimag(a), c = creal(b), d = -cimag(b);
  *re = ar * c - ai * d;
  *im = ar * d + ai * c;
}

orc_c128 orc_dot(const orc_t *m, long n, const zc *a, const zc *b)
{
  long i;
  if (!m || m->red == ORC_RED_SERIAL) { /* cdotprod.c:13-20 */
    zc s = 0;
    for (i = 0; i < n; i++) s += a[i] * conj(b[i]);
    return s;
  }
  long chunk = (long)RED_T * RED_E, nch = (n + chunk - 1) / chunk;
  double *pr = dalloc(nch), *pi = dalloc(nch);
  double vr[RED_T], vi[RED_T];
  for (long c = 0; c < nch; c++) {
    for (int t = 0; t < RED_T; t++) {
      double sr = 0, si = 0, tr, ti;
      for (int q = 0; q < RED_E; q++) {
        i = c * chunk + (long)q * RED_T + t;
        if (i < n) {
          dot_term(a[i], b[i], &tr, &ti);
          sr += tr;
          si += ti;
        }
      }
      vr[t] = sr; vi[t] = si;
    }
    for (int h = RED_T / 2; h >= 1; h /= 2)
      for (int t = 0; t < h; t++) { vr[t] += vr[t + h]; vi[t] += vi[t + h]; }
    pr[c] = vr[0]; pi[c] = vi[0];
  }
  for (int t = 0; t < RED_T; t++) {
    double sr = 0, si = 0;
    for (long q = t; q < nch; q += RED_T) { sr += pr[q]; si += pi[q]; }
    vr[t] = sr; vi[t] = si;
  }
  for (int h = RED_T / 2; h >= 1; h /= 2)
    for (int t = 0; t < h; t++) { vr[t] += vr[t + h]; vi[t] += vi[t + h]; }
  free(pr); free(pi);
  return mk(vr[0], vi[0]);
}

/* ----------------------------------------------- other preconditioners */

#define idv(k, j, i) ((i) + (long)(n1 + 1) * ((j) + (long)(n2 + 1) * (k)))

/* aphi_apply, fdfd.c:513-821: the A-phi splitting E = A + grad(phi).
 * b_phi = -div(x); one forward and one backward Gauss-Seidel sweep on
 * (A, phi) from zero, in the order ax, ay, az, phi forward and phi, az,
 * ay, ax backward; then y = A + grad(phi).  A lives in y.  As written
 * there: every cell zeroes itself before its row is formed; the phi rows
 * read the conductivity behind the cell (t2, t4, t6) from the full grid
 * without the airwave boundary's plane offset (fdfd.c:656-661), the
 * others with it. */
void orc_aphi_apply(orc_t *m, const zc *x, zc *y)
{
  const int n1 = m->n1, n2 = m->n2, n3 = m->n3, kstart = m->airbc ? m->nb : 0;
  int i, j, k, ip1, jp1, kp1, im1, jm1, km1;
  zc t1, t2, t3, t4, t5, t6, aij_xj, aii, iwmu = m->iwmu;
  double vol;
  const double *d1 = m->d1, *d2 = m->d2, *d3 = &m->d3[kstart];
  const double *d1s = m->d1s, *d2s = m->d2s, *d3s = &m->d3s[kstart];
  const double *sigma11 = &m->s11[(long)n1 * (n2 + 1) * kstart];
  const double *sigma22 = &m->s22[(long)(n1 + 1) * n2 * kstart];
  const double *sigma33 = &m->s33[(long)(n1 + 1) * (n2 + 1) * kstart];
  const double *S11 = m->s11, *S22 = m->s22, *S33 = m->s33; /* emf->sigma*[k][j][i], full grid */
  const long lv = (long)(n1 + 1) * (n2 + 1) * (n3 + 1);
  zc *ax = y, *ay = y + m->lEx, *az = y + m->lEx + m->lEy;
  const zc *bx = x, *by = x + m->lEx, *bz = x + m->lEx + m->lEy;
  zc *Ex, *Ey, *Ez;
  if (!m->av) {
    m->av = zalloc(lv);
    m->abv = zalloc(lv);
  }
  zc *v = m->av, *bv = m->abv;
#define FVM_SCALE_A()                                \
  if (m->fvm) {                                      \
    vol = d1s[i] * d2[j] * d3[k];                    \
    aij_xj *= vol * m->iomu;                         \
    aii *= vol * m->iomu;                            \
  }
  /* 1. bv = -div(x) (fdfd.c:554-566) */
  memset(bv, 0, lv * sizeof(zc));
  for (k = 1; k < n3; k++) {
    km1 = k - 1;
    for (j = 1; j < n2; j++) {
      jm1 = j - 1;
      for (i = 1; i < n1; i++) {
        im1 = i - 1;
        bv[idv(k, j, i)] = 0;
        bv[idv(k, j, i)] += -(bx[id1(k, j, i)] - bx[id1(k, j, im1)]) / d1[i];
        bv[idv(k, j, i)] += -(by[id2(k, j, i)] - by[id2(k, jm1, i)]) / d2[j];
        bv[idv(k, j, i)] += -(bz[id3(k, j, i)] - bz[id3(km1, j, i)]) / d3[k];
      }
    }
  }
  memset(ax, 0, m->lEx * sizeof(zc));
  memset(ay, 0, m->lEy * sizeof(zc));
  memset(az, 0, m->lEz * sizeof(zc));
  memset(v, 0, lv * sizeof(zc));
#define ROW_AX()                                                                                                 \
  ip1 = MINI(i + 1, n1); im1 = MAXI(i - 1, 0);                                                                   \
  ax[id1(k, j, i)] = 0;                                                                                          \
  t1 = (ay[id2(k, j, ip1)] - ay[id2(k, j, i)]) / d1s[i] - (ax[id1(k, jp1, i)] - ax[id1(k, j, i)]) / d2s[j];      \
  t2 = (ay[id2(k, jm1, ip1)] - ay[id2(k, jm1, i)]) / d1s[i] - (ax[id1(k, j, i)] - ax[id1(k, jm1, i)]) / d2s[jm1]; \
  t3 = (ax[id1(kp1, j, i)] - ax[id1(k, j, i)]) / d3s[k] - (az[id3(k, j, ip1)] - az[id3(k, j, i)]) / d1s[i];      \
  t4 = (ax[id1(k, j, i)] - ax[id1(km1, j, i)]) / d3s[km1] - (az[id3(km1, j, ip1)] - az[id3(km1, j, i)]) / d1s[i]; \
  aij_xj = ((t1 - t2) / d2[j] - (t3 - t4) / d3[k]) * iwmu -                                                      \
           sigma11[id1(k, j, i)] * (ax[id1(k, j, i)] + (v[idv(k, j, ip1)] - v[idv(k, j, i)]) / d1s[i]);          \
  aii = ((1. / d2s[j] + 1. / d2s[jm1]) / d2[j] + (1. / d3s[k] + 1. / d3s[km1]) / d3[k]) * iwmu -                 \
        sigma11[id1(k, j, i)];                                                                                   \
  FVM_SCALE_A();                                                                                                 \
  ax[id1(k, j, i)] = (bx[id1(k, j, i)] - aij_xj) / aii;
#define ROW_AY()                                                                                                   \
  ip1 = MINI(i + 1, n1); im1 = MAXI(i - 1, 0);                                                                     \
  ay[id2(k, j, i)] = 0;                                                                                            \
  t1 = (az[id3(k, jp1, i)] - az[id3(k, j, i)]) / d2s[j] - (ay[id2(kp1, j, i)] - ay[id2(k, j, i)]) / d3s[k];        \
  t2 = (az[id3(km1, jp1, i)] - az[id3(km1, j, i)]) / d2s[j] - (ay[id2(k, j, i)] - ay[id2(km1, j, i)]) / d3s[km1];  \
  t3 = (ay[id2(k, j, ip1)] - ay[id2(k, j, i)]) / d1s[i] - (ax[id1(k, jp1, i)] - ax[id1(k, j, i)]) / d2s[j];        \
  t4 = (ay[id2(k, j, i)] - ay[id2(k, j, im1)]) / d1s[im1] - (ax[id1(k, jp1, im1)] - ax[id1(k, j, im1)]) / d2s[j];  \
  aij_xj = ((t1 - t2) / d3[k] - (t3 - t4) / d1[i]) * iwmu -                                                        \
           sigma22[id2(k, j, i)] * (ay[id2(k, j, i)] + (v[idv(k, jp1, i)] - v[idv(k, j, i)]) / d2s[j]);            \
  aii = ((1. / d3s[k] + 1. / d3s[km1]) / d3[k] + (1. / d1s[i] + 1. / d1s[im1]) / d1[i]) * iwmu -                   \
        sigma22[id2(k, j, i)];                                                                                     \
  FVM_SCALE_A();                                                                                                   \
  ay[id2(k, j, i)] = (by[id2(k, j, i)] - aij_xj) / aii;
#define ROW_AZ()                                                                                                   \
  ip1 = MINI(i + 1, n1); im1 = MAXI(i - 1, 0);                                                                     \
  az[id3(k, j, i)] = 0;                                                                                            \
  t1 = (ax[id1(kp1, j, i)] - ax[id1(k, j, i)]) / d3s[k] - (az[id3(k, j, ip1)] - az[id3(k, j, i)]) / d1s[i];        \
  t2 = (ax[id1(kp1, j, im1)] - ax[id1(k, j, im1)]) / d3s[k] - (az[id3(k, j, i)] - az[id3(k, j, im1)]) / d1s[im1];  \
  t3 = (az[id3(k, jp1, i)] - az[id3(k, j, i)]) / d2s[j] - (ay[id2(kp1, j, i)] - ay[id2(k, j, i)]) / d3s[k];        \
  t4 = (az[id3(k, j, i)] - az[id3(k, jm1, i)]) / d2s[jm1] - (ay[id2(kp1, jm1, i)] - ay[id2(k, jm1, i)]) / d3s[k];  \
  aij_xj = ((t1 - t2) / d1[i] - (t3 - t4) / d2[j]) * iwmu -                                                        \
           sigma33[id3(k, j, i)] * (az[id3(k, j, i)] + (v[idv(kp1, j, i)] - v[idv(k, j, i)]) / d3s[k]);            \
  aii = ((1. / d1s[i] + 1. / d1s[im1]) / d1[i] + (1. / d2s[j] + 1. / d2s[jm1]) / d2[j]) * iwmu -                   \
        sigma33[id3(k, j, i)];                                                                                     \
  FVM_SCALE_A();                                                                                                   \
  az[id3(k, j, i)] = (bz[id3(k, j, i)] - aij_xj) / aii;
#define ROW_V()                                                                                                    \
  ip1 = i + 1; im1 = i - 1;                                                                                        \
  v[idv(k, j, i)] = 0;                                                                                             \
  t1 = sigma11[id1(k, j, i)] * (ax[id1(k, j, i)] + (v[idv(k, j, ip1)] - v[idv(k, j, i)]) / d1s[i]);                 \
  t2 = S11[id1(k, j, im1)] * (ax[id1(k, j, im1)] + (v[idv(k, j, i)] - v[idv(k, j, im1)]) / d1s[im1]);               \
  t3 = sigma22[id2(k, j, i)] * (ay[id2(k, j, i)] + (v[idv(k, jp1, i)] - v[idv(k, j, i)]) / d2s[j]);                 \
  t4 = S22[id2(k, jm1, i)] * (ay[id2(k, jm1, i)] + (v[idv(k, j, i)] - v[idv(k, jm1, i)]) / d2s[jm1]);               \
  t5 = sigma33[id3(k, j, i)] * (az[id3(k, j, i)] + (v[idv(kp1, j, i)] - v[idv(k, j, i)]) / d3s[k]);                 \
  t6 = S33[id3(km1, j, i)] * (az[id3(km1, j, i)] + (v[idv(k, j, i)] - v[idv(km1, j, i)]) / d3s[km1]);               \
  aij_xj = ((t1 - t2) / d1[i] + (t3 - t4) / d2[j] + (t5 - t6) / d3[k]);                                            \
  aii = -((1. / d1s[i] + 1. / d1s[im1]) / d1[i] + (1. / d2s[j] + 1. / d2s[jm1]) / d2[j] +                          \
          (1. / d3s[k] + 1. / d3s[km1]) / d3[k]);                                                                  \
  FVM_SCALE_A();                                                                                                   \
  v[idv(k, j, i)] = (bv[idv(k, j, i)] - aij_xj) / aii;
  /* 2. forward Gauss-Seidel (fdfd.c:575-687) */
  for (k = 1; k < n3; k++) {
    kp1 = MINI(k + 1, n3); km1 = MAXI(k - 1, 0);
    for (j = 1; j < n2; j++) {
      jp1 = MINI(j + 1, n2); jm1 = MAXI(j - 1, 0);
      for (i = 0; i < n1; i++) { ROW_AX(); }
    }
  }
  for (k = 1; k < n3; k++) {
    kp1 = MINI(k + 1, n3); km1 = MAXI(k - 1, 0);
    for (j = 0; j < n2; j++) {
      jp1 = MINI(j + 1, n2); jm1 = MAXI(j - 1, 0);
      for (i = 1; i < n1; i++) { ROW_AY(); }
    }
  }
  for (k = 0; k < n3; k++) {
    kp1 = MINI(k + 1, n3); km1 = MAXI(k - 1, 0);
    for (j = 1; j < n2; j++) {
      jp1 = MINI(j + 1, n2); jm1 = MAXI(j - 1, 0);
      for (i = 1; i < n1; i++) { ROW_AZ(); }
    }
  }
  for (k = 1; k < n3; k++) {
    kp1 = k + 1; km1 = k - 1;
    for (j = 1; j < n2; j++) {
      jp1 = j + 1; jm1 = j - 1;
      for (i = 1; i < n1; i++) { ROW_V(); }
    }
  }
  /* backward Gauss-Seidel (fdfd.c:689-790) */
  for (k = n3 - 1; k >= 1; k--) {
    kp1 = k + 1; km1 = k - 1;
    for (j = n2 - 1; j >= 1; j--) {
      jp1 = j + 1; jm1 = j - 1;
      for (i = n1 - 1; i >= 1; i--) { ROW_V(); }
    }
  }
  for (k = n3 - 1; k >= 0; k--) {
    kp1 = MINI(k + 1, n3); km1 = MAXI(k - 1, 0);
    for (j = n2 - 1; j > 0; j--) {
      jp1 = MINI(j + 1, n2); jm1 = MAXI(j - 1, 0);
      for (i = n1 - 1; i > 0; i--) { ROW_AZ(); }
    }
  }
  for (k = n3 - 1; k > 0; k--) {
    kp1 = MINI(k + 1, n3); km1 = MAXI(k - 1, 0);
    for (j = n2 - 1; j >= 0; j--) {
      jp1 = MINI(j + 1, n2); jm1 = MAXI(j - 1, 0);
      for (i = n1 - 1; i > 0; i--) { ROW_AY(); }
    }
  }
  for (k = n3 - 1; k > 0; k--) {
    kp1 = MINI(k + 1, n3); km1 = MAXI(k - 1, 0);
    for (j = n2 - 1; j > 0; j--) {
      jp1 = MINI(j + 1, n2); jm1 = MAXI(j - 1, 0);
      for (i = n1 - 1; i >= 0; i--) { ROW_AX(); }
    }
  }
  /* 3. y = A + grad(phi) (fdfd.c:793-820), in place */
  Ex = y; Ey = y + m->lEx; Ez = y + m->lEx + m->lEy;
  for (k = 1; k < n3; k++)
    for (j = 1; j < n2; j++)
      for (i = 0; i < n1; i++) Ex[id1(k, j, i)] = ax[id1(k, j, i)] + (v[idv(k, j, i + 1)] - v[idv(k, j, i)]) / d1s[i];
  for (k = 1; k < n3; k++)
    for (j = 0; j < n2; j++)
      for (i = 1; i < n1; i++) Ey[id2(k, j, i)] = ay[id2(k, j, i)] + (v[idv(k, j + 1, i)] - v[idv(k, j, i)]) / d2s[j];
  for (k = 0; k < n3; k++)
    for (j = 1; j < n2; j++)
      for (i = 1; i < n1; i++) Ez[id3(k, j, i)] = az[id3(k, j, i)] + (v[idv(k + 1, j, i)] - v[idv(k, j, i)]) / d3s[k];
#undef ROW_AX
#undef ROW_AY
#undef ROW_AZ
#undef ROW_V
#undef FVM_SCALE_A
}
#undef idv

/* gmres_apply, gmres_init_free.c:58-137: y = m-step GMRES solution of
 * A y = x from y = 0 (no residual test, no breakdown test). */
void orc_gmres_prec(orc_t *m, const zc *b, zc *x)
{
  const long n = m->lE;
  const int mp = m->mp;
  long i, k;
  int j;
  double beta, tmp;
  zc ss;
  zc *w = zalloc(n), *r = zalloc(n), *v = zalloc(n * (mp + 1));
  zc *h = zalloc((long)(mp + 1) * mp), *g = zalloc(mp + 1), *y = zalloc(mp + 1), *c = zalloc(mp + 1);
  double *s = dalloc(mp + 1);
#define HG(a, b2) h[(long)(a) * mp + (b2)] /* h[a][b2], alloc2complex(m, m+1) */
  memset(x, 0, n * sizeof(zc));
  for (i = 0; i < n; i++) r[i] = b[i];
  beta = sqrt(creal(orc_dot(m, n, r, r)));
  if (beta == 0.0) goto out;
  for (i = 0; i < n; i++) v[i] = r[i] / beta;
  memset(g, 0, (mp + 1) * sizeof(zc));
  g[0] = beta;
  for (j = 0; j < mp; j++) {
    orc_fdfd_apply(m, &v[n * j], w);
    for (i = 0; i <= j; i++) {
      HG(i, j) = orc_dot(m, n, w, &v[n * i]);
      for (k = 0; k < n; k++) w[k] -= HG(i, j) * v[n * i + k];
    }
    HG(j + 1, j) = sqrt(creal(orc_dot(m, n, w, w)));
    for (i = 0; i < n; i++) v[n * (j + 1) + i] = w[i] / HG(j + 1, j);
    if (j > 0) {
      for (i = 0; i < j; i++) {
        ss = conj(c[i]) * HG(i, j) + s[i] * HG(i + 1, j);
        HG(i + 1, j) = -s[i] * HG(i, j) + c[i] * HG(i + 1, j);
        HG(i, j) = ss;
      }
    }
    tmp = sqrt(creal(HG(j, j) * conj(HG(j, j)) + HG(j + 1, j) * conj(HG(j + 1, j))));
    s[j] = creal(HG(j + 1, j)) / tmp;
    c[j] = HG(j, j) / tmp;
    HG(j, j) = conj(c[j]) * HG(j, j) + s[j] * HG(j + 1, j);
    HG(j + 1, j) = 0.;
    g[j + 1] = -s[j] * g[j];
    g[j] = conj(c[j]) * g[j];
  }
  y[mp - 1] = g[mp - 1] / HG(mp - 1, mp - 1);
  for (i = mp - 2; i >= 0; i--) {
    y[i] = g[i];
    for (j = i + 1; j < mp; j++) y[i] -= HG(i, j) * y[j];
    y[i] /= HG(i, i);
  }
  for (j = 0; j < mp; j++)
    for (i = 0; i < n; i++) x[i] += y[j] * v[n * j + i];
#undef HG
out:
  free(w); free(r); free(v); free(h); free(g); free(y); free(c); free(s);
}

void orc_set_mp(orc_t *m, int mp) { m->mp = mp; }

/* the preconditioner of preco = 1 (LU-SGS), 2 (inner GMRES), 3 (A-phi) */
static void orc_precond(orc_t *m, int precond, const zc *x, zc *y)
{
  if (precond == 2) orc_gmres_prec(m, x, y);
  else if (precond == 3) orc_aphi_apply(m, x, y);
  else orc_lusgs_apply(m, x, y);
}

/* --------------------------------------------------------------- Krylov */

/* bicgstab (solver.c:16-87) and fbicgstab (solver.c:93-178); the
 * unpreconditioned form is F-BiCGStab with M = I, whose vector updates
 * coincide with solver.c:56-76 once pp = p and the reused r equals s. */
int orc_fbicgstab(orc_t *m, zc *x, const zc *b, int niter, double tol, int precond, double *hist)
{
  long n = m->lE, i;
  int iter;
  double rs0, rs, relres;
  zc rho_old, rho_new, alpha, beta, omega;
  zc *r = zalloc(n), *r0 = zalloc(n), *p = zalloc(n), *pp = zalloc(n);
  zc *v = zalloc(n), *s = zalloc(n), *t = zalloc(n);
  orc_fdfd_apply(m, x, v);
  for (i = 0; i < n; i++) {
    r[i] = b[i] - v[i];
    p[i] = r[i];
    r0[i] = r[i];
  }
  rho_old = orc_dot(m, n, r, r0);
  rs = creal(orc_dot(m, n, r, r));
  rs0 = rs;
  if (rs == 0.) {
    iter = 0;
    goto done;
  }
  relres = sqrt(rs / rs0);
  if (hist) hist[0] = relres;
  for (iter = 0; iter < niter; ++iter) {
    if (hist) hist[iter] = relres;
    if (precond) {
      orc_precond(m, precond, p, pp);
      orc_fdfd_apply(m, pp, v);
    } else {
      orc_fdfd_apply(m, p, v);
    }
    alpha = orc_dot(m, n, v, r0);
    alpha = rho_old / alpha;
    for (i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
    if (precond) {
      orc_precond(m, precond, s, r);
      orc_fdfd_apply(m, r, t);
      omega = orc_dot(m, n, s, t) / creal(orc_dot(m, n, t, t));
      for (i = 0; i < n; i++) {
        x[i] += alpha * pp[i] + omega * r[i];
        r[i] = s[i] - omega * t[i];
      }
    } else {
      orc_fdfd_apply(m, s, t);
      /* solver.c:61 computes cdotprod(t, s) here, solver.c:151 (s, t) */
      omega = orc_dot(m, n, t, s) / creal(orc_dot(m, n, t, t));
      for (i = 0; i < n; i++) {
        x[i] += alpha * p[i] + omega * s[i];
        r[i] = s[i] - omega * t[i];
      }
    }
    rs = creal(orc_dot(m, n, r, r));
    relres = sqrt(rs / rs0);
    if (relres < tol) {
      if (hist) hist[iter + 1] = relres;
      break;
    }
    rho_new = orc_dot(m, n, r, r0);
    beta = (rho_new / rho_old) * alpha / omega;
    for (i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    rho_old = rho_new;
  }
done:
  free(r); free(r0); free(p); free(pp); free(v); free(s); free(t);
  return iter;
}

/* fgmres (solver.c:403-512); precond=0 runs the same recurrence with
 * z_j = v_j, i.e. right-preconditioned GMRES with M = I.  The reference's
 * unpreconditioned gmres (solver.c:181-285) orders its Hessenberg storage
 * differently but performs the same arithmetic. Returns the total count of
 * inner steps (the number printed as iter*m+j+1). */
int orc_fgmres(orc_t *m, zc *x, const zc *b, int nouter, double tol, int mr, int precond, double *hist)
{
  long n = m->lE, i;
  int j, k, mm, iter, conv = 0, steps = 0;
  double beta, tmp, r0 = 1, relres;
  zc ss;
  zc *w = zalloc(n), *r = zalloc(n);
  zc *v = zalloc(n * (mr + 1)), *z = zalloc(n * mr);
  zc *h = zalloc((long)(mr + 1) * mr), *g = zalloc(mr + 1), *y = zalloc(mr + 1), *c = zalloc(mr + 1);
  double *sn = dalloc(mr + 1);
#define H(a, b2) h[(long)(a) * (mr + 1) + (b2)]
  for (iter = 0; iter < nouter; iter++) {
    orc_fdfd_apply(m, x, w);
    for (i = 0; i < n; i++) r[i] = b[i] - w[i];
    beta = sqrt(creal(orc_dot(m, n, r, r)));
    if (beta == 0.) break;
    for (i = 0; i < n; i++) v[i] = r[i] / beta;
    memset(g, 0, (mr + 1) * sizeof(zc));
    g[0] = beta;
    memset(h, 0, (long)(mr + 1) * mr * sizeof(zc));
    if (iter == 0) {
      r0 = beta;
      if (hist) hist[0] = 1.0;
    }
    for (j = 0; j < mr; j++) {
      zc *zj = precond ? &z[n * j] : &v[n * j];
      if (precond) orc_precond(m, precond, &v[n * j], zj);
      orc_fdfd_apply(m, zj, w);
      for (i = 0; i <= j; i++) {
        H(j, i) = orc_dot(m, n, w, &v[n * i]);
        for (k = 0; k < n; k++) w[k] -= H(j, i) * v[n * i + k];
      }
      H(j, j + 1) = sqrt(creal(orc_dot(m, n, w, w)));
      if (cabs(H(j, j + 1)) == 0.) {
        conv = 1;
        break;
      }
      for (i = 0; i < n; i++) v[n * (j + 1) + i] = w[i] / H(j, j + 1);
      if (j > 0) {
        for (i = 0; i < j; i++) {
          ss = conj(c[i]) * H(j, i) + sn[i] * H(j, i + 1);
          H(j, i + 1) = -sn[i] * H(j, i) + c[i] * H(j, i + 1);
          H(j, i) = ss;
        }
      }
      tmp = sqrt(creal(H(j, j) * conj(H(j, j)) + H(j, j + 1) * conj(H(j, j + 1))));
      sn[j] = creal(H(j, j + 1)) / tmp;
      c[j] = H(j, j) / tmp;
      g[j + 1] = -sn[j] * g[j];
      g[j] = conj(c[j]) * g[j];
      H(j, j) = conj(c[j]) * H(j, j) + sn[j] * H(j, j + 1);
      relres = cabs(g[j + 1]) / r0;
      steps = iter * mr + j + 1;
      if (hist) hist[steps] = relres;
      if (relres < tol) {
        conv = 1;
        break;
      }
    }
    mm = MINI(mr, j + 1);
    y[mm - 1] = g[mm - 1] / H(mm - 1, mm - 1);
    for (i = mm - 2; i >= 0; i--) {
      y[i] = g[i];
      for (j = i + 1; j < mm; j++) y[i] -= H(j, i) * y[j];
      y[i] /= H(i, i);
    }
    for (j = 0; j < mm; j++) {
      const zc *zj = precond ? &z[n * j] : &v[n * j];
      for (i = 0; i < n; i++) x[i] += y[j] * zj[i];
    }
    if (conv) break;
  }
#undef H
  free(w); free(r); free(v); free(z); free(h); free(g); free(y); free(c); free(sn);
  return steps;
}

/* gmres (solver.c:181-285), unpreconditioned restarted GMRES(m).  Two
 * behaviours are kept as written there: a happy breakdown shrinks m for all
 * later restarts, and convergence returns before the solution update of the
 * current cycle (solver.c:251), so x keeps the previous cycle's value. */
int orc_gmres(orc_t *m, zc *x, const zc *b, int nouter, double tol, int mr, double *hist)
{
  long n = m->lE, i;
  int j, k, iter, steps = 0, mcap = mr;
  double beta, tmp, r0 = 1;
  zc ss;
  zc *w = zalloc(n), *r = zalloc(n), *v = zalloc(n * (mcap + 1));
  zc *h = zalloc((long)(mcap + 1) * mcap), *g = zalloc(mcap + 1), *y = zalloc(mcap + 1), *c = zalloc(mcap + 1);
  double *sn = dalloc(mcap + 1);
#define HH(a, b2) h[(long)(a) * mcap + (b2)]
  for (iter = 0; iter < nouter; iter++) {
    orc_fdfd_apply(m, x, w);
    for (i = 0; i < n; i++) r[i] = b[i] - w[i];
    beta = sqrt(creal(orc_dot(m, n, r, r)));
    if (beta == 0.0) goto out;
    for (i = 0; i < n; i++) v[i] = r[i] / beta;
    memset(g, 0, (mcap + 1) * sizeof(zc));
    g[0] = beta;
    memset(h, 0, (long)(mcap + 1) * mcap * sizeof(zc));
    if (iter == 0) {
      r0 = beta;
      if (hist) hist[0] = beta / r0;
    }
    for (j = 0; j < mr; j++) {
      orc_fdfd_apply(m, &v[n * j], w);
      for (i = 0; i <= j; i++) {
        HH(i, j) = orc_dot(m, n, w, &v[n * i]);
        for (k = 0; k < n; k++) w[k] -= HH(i, j) * v[n * i + k];
      }
      HH(j + 1, j) = sqrt(creal(orc_dot(m, n, w, w)));
      if (cabs(HH(j + 1, j)) == 0.0) {
        mr = j + 1;
        break;
      }
      for (i = 0; i < n; i++) v[n * (j + 1) + i] = w[i] / HH(j + 1, j);
      if (j > 0) {
        for (i = 0; i < j; i++) {
          ss = conj(c[i]) * HH(i, j) + sn[i] * HH(i + 1, j);
          HH(i + 1, j) = -sn[i] * HH(i, j) + c[i] * HH(i + 1, j);
          HH(i, j) = ss;
        }
      }
      tmp = sqrt(creal(HH(j, j) * conj(HH(j, j)) + HH(j + 1, j) * conj(HH(j + 1, j))));
      sn[j] = creal(HH(j + 1, j)) / tmp;
      c[j] = HH(j, j) / tmp;
      HH(j, j) = conj(c[j]) * HH(j, j) + sn[j] * HH(j + 1, j);
      HH(j + 1, j) = 0.;
      g[j + 1] = -sn[j] * g[j];
      g[j] = conj(c[j]) * g[j];
      tmp = cabs(g[j + 1]) / r0;
      steps = iter * mr + j + 1;
      if (hist) hist[steps] = tmp;
      if (tmp < tol) goto out;
    }
    y[mr - 1] = g[mr - 1] / HH(mr - 1, mr - 1);
    for (i = mr - 2; i >= 0; i--) {
      y[i] = g[i];
      for (j = i + 1; j < mr; j++) y[i] -= HH(i, j) * y[j];
      y[i] /= HH(i, i);
    }
    for (j = 0; j < mr; j++)
      for (i = 0; i < n; i++) x[i] += y[j] * v[n * j + i];
  }
out:
#undef HH
  free(w); free(r); free(v); free(h); free(g); free(y); free(c); free(sn);
  return steps;
}

/* ---------------------------------------------------- inject / extract */

/* inject_source, inject_extract.c:24-120 (one source, one channel) */
void orc_inject(orc_t *m, int ch, float sx, float sy, float sz, zc *b)
{
  int n1 = m->nx + 2 * m->nb, n2 = m->ny + 2 * m->nb, n3 = m->nz + 2 * m->nb, i, j, k;
  double w1, w2, w3, vol;
  zc src_vol;
  zc *bx = b, *by = b + m->lEx, *bz = b + m->lEx + m->lEy;
  const double *s11 = m->s11, *s22 = m->s22;
#define SG1(k, j, i) s11[id1(k, j, i)]
#define SG2(k, j, i) s22[id2(k, j, i)]
  if (ch == 0) {
    i = orc_find_index(n1, m->x1s, sx);
    j = orc_find_index(n2 + 1, m->x2, sy);
    k = orc_find_index(n3 + 1, m->x3, sz);
    w1 = (sx - m->x1s[i]) / (m->x1s[i + 1] - m->x1s[i]);
    w2 = (sy - m->x2[j]) / (m->x2[j + 1] - m->x2[j]);
    w3 = (sz - m->x3[k]) / (m->x3[k + 1] - m->x3[k]);
    vol = (m->x1s[i + 1] - m->x1s[i]) * (m->x2[j + 1] - m->x2[j]) * (m->x3[k + 1] - m->x3[k]);
    src_vol = 1. / vol;
    if (m->fvm) src_vol = m->iomu;
    src_vol /= m->rho_water;
    if (m->airbc) k -= m->nb; /* the conductivity below is then read nb planes up */
    bx[id1(k, j, i)] += src_vol / SG1(k, j, i) * (1. - w1) * (1. - w2) * (1. - w3);
    bx[id1(k, j, i + 1)] += src_vol / SG1(k, j, i + 1) * w1 * (1. - w2) * (1. - w3);
    bx[id1(k, j + 1, i)] += src_vol / SG1(k, j + 1, i) * (1. - w1) * w2 * (1. - w3);
    bx[id1(k, j + 1, i + 1)] += src_vol / SG1(k, j + 1, i + 1) * w1 * w2 * (1. - w3);
    bx[id1(k + 1, j, i)] += src_vol / SG1(k + 1, j, i) * (1. - w1) * (1. - w2) * w3;
    bx[id1(k + 1, j, i + 1)] += src_vol / SG1(k + 1, j, i + 1) * w1 * (1. - w2) * w3;
    bx[id1(k + 1, j + 1, i)] += src_vol / SG1(k + 1, j + 1, i) * (1. - w1) * w2 * w3;
    bx[id1(k + 1, j + 1, i + 1)] += src_vol / SG1(k + 1, j + 1, i + 1) * w1 * w2 * w3;
  } else if (ch == 1) {
    i = orc_find_index(n1 + 1, m->x1, sx);
    j = orc_find_index(n2, m->x2s, sy);
    k = orc_find_index(n3 + 1, m->x3, sz);
    w1 = (sx - m->x1[i]) / (m->x1[i + 1] - m->x1[i]);
    w2 = (sy - m->x2s[j]) / (m->x2s[j + 1] - m->x2s[j]);
    w3 = (sz - m->x3[k]) / (m->x3[k + 1] - m->x3[k]);
    vol = (m->x1[i + 1] - m->x1[i]) * (m->x2s[j + 1] - m->x2s[j]) * (m->x3[k + 1] - m->x3[k]);
    src_vol = 1. / vol;
    if (m->fvm) src_vol = m->iomu;
    src_vol /= m->rho_water;
    if (m->airbc) k -= m->nb;
    by[id2(k, j, i)] += src_vol / SG2(k, j, i) * (1. - w1) * (1. - w2) * (1. - w3);
    by[id2(k, j, i + 1)] += src_vol / SG2(k, j, i + 1) * w1 * (1. - w2) * (1. - w3);
    by[id2(k, j + 1, i)] += src_vol / SG2(k, j + 1, i) * (1. - w1) * w2 * (1. - w3);
    by[id2(k, j + 1, i + 1)] += src_vol / SG2(k, j + 1, i + 1) * w1 * w2 * (1. - w3);
    by[id2(k + 1, j, i)] += src_vol / SG2(k + 1, j, i) * (1. - w1) * (1. - w2) * w3;
    by[id2(k + 1, j, i + 1)] += src_vol / SG2(k + 1, j, i + 1) * w1 * (1. - w2) * w3;
    by[id2(k + 1, j + 1, i)] += src_vol / SG2(k + 1, j + 1, i) * (1. - w1) * w2 * w3;
    by[id2(k + 1, j + 1, i + 1)] += src_vol / SG2(k + 1, j + 1, i + 1) * w1 * w2 * w3;
  } else {
    i = orc_find_index(n1 + 1, m->x1, sx);
    j = orc_find_index(n2 + 1, m->x2, sy);
    k = orc_find_index(n3, m->x3s, sz);
    w1 = (sx - m->x1[i]) / (m->x1[i + 1] - m->x1[i]);
    w2 = (sy - m->x2[j]) / (m->x2[j + 1] - m->x2[j]);
    w3 = (sz - m->x3s[k]) / (m->x3s[k + 1] - m->x3s[k]);
    vol = (m->x1[i + 1] - m->x1[i]) * (m->x2[j + 1] - m->x2[j]) * (m->x3s[k + 1] - m->x3s[k]);
    src_vol = 1. / vol;
    if (m->fvm) src_vol = m->iomu;
    if (m->airbc) k -= m->nb;
    bz[id3(k, j, i)] += src_vol * (1. - w1) * (1. - w2) * (1. - w3);
    bz[id3(k, j, i + 1)] += src_vol * w1 * (1. - w2) * (1. - w3);
    bz[id3(k, j + 1, i)] += src_vol * (1. - w1) * w2 * (1. - w3);
    bz[id3(k, j + 1, i + 1)] += src_vol * w1 * w2 * (1. - w3);
    bz[id3(k + 1, j, i)] += src_vol * (1. - w1) * (1. - w2) * w3;
    bz[id3(k + 1, j, i + 1)] += src_vol * w1 * (1. - w2) * w3;
    bz[id3(k + 1, j + 1, i)] += src_vol * (1. - w1) * w2 * w3;
    bz[id3(k + 1, j + 1, i + 1)] += src_vol * w1 * w2 * w3;
  }
#undef SG1
#undef SG2
}

/* extract_emf, inject_extract.c:123-213 (one receiver, one channel) */
/* H at face (k, j, i) of the operator grid (c: 0 Hx, 1 Hy, 2 Hz) from E,
 * with the expressions of fdfd.c:131-146 and the vertical spacing of
 * operator plane k (fdfd.c:106-109). */
static zc orc_h_face(const orc_t *m, const zc *x, int c, int k, int j, int i)
{
  const int n1 = m->n1, n2 = m->n2;
  const zc *Ex = x, *Ey = x + m->lEx, *Ez = x + m->lEx + m->lEy;
  const double *d3s = &m->d3s[m->airbc ? m->nb : 0];
  zc t1, t2;
  if (c == 0) {
    t1 = (Ez[id3(k, j + 1, i)] - Ez[id3(k, j, i)]) / m->d2s[j];
    t2 = (Ey[id2(k + 1, j, i)] - Ey[id2(k, j, i)]) / d3s[k];
  } else if (c == 1) {
    t1 = (Ex[id1(k + 1, j, i)] - Ex[id1(k, j, i)]) / d3s[k];
    t2 = (Ez[id3(k, j, i + 1)] - Ez[id3(k, j, i)]) / m->d1s[i];
  } else {
    t1 = (Ey[id2(k, j, i + 1)] - Ey[id2(k, j, i)]) / m->d1s[i];
    t2 = (Ex[id1(k, j + 1, i)] - Ex[id1(k, j, i)]) / m->d2s[j];
  }
  return (t1 - t2) * m->iwmu;
}

/* extract_emf, inject_extract.c:123-213, for ch 0..2 (Ex, Ey, Ez).  ch 3..5
 * (Hx, Hy, Hz) have no counterpart in the reference (its H channels stay
 * zero): H is formed on the eight faces around the receiver as fdfd_apply
 * forms it (Hx at (x1, x2s, x3s), Hy at (x1s, x2, x3s), Hz at (x1s, x2s,
 * x3)) and interpolated trilinearly; NaN when a face lies outside the
 * operator grid. */
orc_c128 orc_extract(orc_t *m, int ch, float rx, float ry, float rz, const zc *x)
{
  if (ch >= 3) {
    const int c = ch - 3, N1 = m->nx + 2 * m->nb, N2 = m->ny + 2 * m->nb, N3 = m->nz + 2 * m->nb;
    const double *xa = c == 0 ? m->x1 : m->x1s, *ya = c == 1 ? m->x2 : m->x2s, *za = c == 2 ? m->x3 : m->x3s;
    int i = orc_find_index(c == 0 ? N1 + 1 : N1, xa, rx);
    int j = orc_find_index(c == 1 ? N2 + 1 : N2, ya, ry);
    int k = orc_find_index(c == 2 ? N3 + 1 : N3, za, rz);
    double w1 = (rx - xa[i]) / (xa[i + 1] - xa[i]);
    double w2 = (ry - ya[j]) / (ya[j + 1] - ya[j]);
    double w3 = (rz - za[k]) / (za[k + 1] - za[k]);
    zc s = 0;
    if (m->airbc) k -= m->nb;
    if (k < 0 || k + 1 > (c == 2 ? m->n3 : m->n3 - 1)) {
      return NAN;
    }
    s += orc_h_face(m, x, c, k, j, i) * (1. - w1) * (1. - w2) * (1. - w3);
    s += orc_h_face(m, x, c, k, j, i + 1) * w1 * (1. - w2) * (1. - w3);
    s += orc_h_face(m, x, c, k, j + 1, i) * (1. - w1) * w2 * (1. - w3);
    s += orc_h_face(m, x, c, k, j + 1, i + 1) * w1 * w2 * (1. - w3);
    s += orc_h_face(m, x, c, k + 1, j, i) * (1. - w1) * (1. - w2) * w3;
    s += orc_h_face(m, x, c, k + 1, j, i + 1) * w1 * (1. - w2) * w3;
    s += orc_h_face(m, x, c, k + 1, j + 1, i) * (1. - w1) * w2 * w3;
    s += orc_h_face(m, x, c, k + 1, j + 1, i + 1) * w1 * w2 * w3;
    return s;
  }
  int n1 = m->nx + 2 * m->nb, n2 = m->ny + 2 * m->nb, n3 = m->nz + 2 * m->nb, i, j, k, kk;
  double w1, w2, w3;
  const zc *Ex = x, *Ey = x + m->lEx, *Ez = x + m->lEx + m->lEy;
  zc s = 0;
  if (ch == 0) {
    i = orc_find_index(n1, m->x1s, rx);
    j = orc_find_index(n2 + 1, m->x2, ry);
    k = orc_find_index(n3 + 1, m->x3, rz);
    w1 = (rx - m->x1s[i]) / (m->x1s[i + 1] - m->x1s[i]);
    w2 = (ry - m->x2[j]) / (m->x2[j + 1] - m->x2[j]);
    w3 = (rz - m->x3[k]) / (m->x3[k + 1] - m->x3[k]);
    if (m->airbc) k -= m->nb;
    s += Ex[id1(k, j, i)] * (1. - w1) * (1. - w2) * (1. - w3);
    s += Ex[id1(k, j, i + 1)] * w1 * (1. - w2) * (1. - w3);
    s += Ex[id1(k, j + 1, i)] * (1. - w1) * w2 * (1. - w3);
    s += Ex[id1(k, j + 1, i + 1)] * w1 * w2 * (1. - w3);
    s += Ex[id1(k + 1, j, i)] * (1. - w1) * (1. - w2) * w3;
    s += Ex[id1(k + 1, j, i + 1)] * w1 * (1. - w2) * w3;
    s += Ex[id1(k + 1, j + 1, i)] * (1. - w1) * w2 * w3;
    s += Ex[id1(k + 1, j + 1, i + 1)] * w1 * w2 * w3;
  } else if (ch == 1) {
    i = orc_find_index(n1 + 1, m->x1, rx);
    j = orc_find_index(n2, m->x2s, ry);
    k = orc_find_index(n3 + 1, m->x3, rz);
    w1 = (rx - m->x1[i]) / (m->x1[i + 1] - m->x1[i]);
    w2 = (ry - m->x2s[j]) / (m->x2s[j + 1] - m->x2s[j]);
    w3 = (rz - m->x3[k]) / (m->x3[k + 1] - m->x3[k]);
    if (m->airbc) k -= m->nb;
    s += Ey[id2(k, j, i)] * (1. - w1) * (1. - w2) * (1. - w3);
    s += Ey[id2(k, j, i + 1)] * w1 * (1. - w2) * (1. - w3);
    s += Ey[id2(k, j + 1, i)] * (1. - w1) * w2 * (1. - w3);
    s += Ey[id2(k, j + 1, i + 1)] * w1 * w2 * (1. - w3);
    s += Ey[id2(k + 1, j, i)] * (1. - w1) * (1. - w2) * w3;
    s += Ey[id2(k + 1, j, i + 1)] * w1 * (1. - w2) * w3;
    s += Ey[id2(k + 1, j + 1, i)] * (1. - w1) * w2 * w3;
    s += Ey[id2(k + 1, j + 1, i + 1)] * w1 * w2 * w3;
  } else {
    const double *s33 = m->s33;
    i = orc_find_index(n1 + 1, m->x1, rx);
    j = orc_find_index(n2 + 1, m->x2, ry);
    k = orc_find_index(n3, m->x3s, rz);
    w1 = (rx - m->x1[i]) / (m->x1[i + 1] - m->x1[i]);
    w2 = (ry - m->x2[j]) / (m->x2[j + 1] - m->x2[j]);
    w3 = (rz - m->x3s[k]) / (m->x3s[k + 1] - m->x3s[k]);
    kk = k;
    if (m->airbc) k -= m->nb;
#define SG3(k, j, i) s33[id3(k, j, i)]
    s += SG3(kk, j, i) * Ez[id3(k, j, i)] * (1. - w1) * (1. - w2) * (1. - w3);
    s += SG3(kk, j, i + 1) * Ez[id3(k, j, i + 1)] * w1 * (1. - w2) * (1. - w3);
    s += SG3(kk, j + 1, i) * Ez[id3(k, j + 1, i)] * (1. - w1) * w2 * (1. - w3);
    s += SG3(kk, j + 1, i + 1) * Ez[id3(k, j + 1, i + 1)] * w1 * w2 * (1. - w3);
    s += SG3(kk + 1, j, i) * Ez[id3(k + 1, j, i)] * (1. - w1) * (1. - w2) * w3;
    s += SG3(kk + 1, j, i + 1) * Ez[id3(k + 1, j, i + 1)] * w1 * (1. - w2) * w3;
    s += SG3(kk + 1, j + 1, i) * Ez[id3(k + 1, j + 1, i)] * (1. - w1) * w2 * w3;
    s += SG3(kk + 1, j + 1, i + 1) * Ez[id3(k + 1, j + 1, i + 1)] * w1 * w2 * w3;
#undef SG3
    s = s * m->rho_water;
  }
  return s;
}

/* ------------------------------------------------------ matrix storage */

/* Acsr_init (Acsr.c:22-693) on the full grid at the current frequency:
 * iw mu0 times the curl-curl operator, every row (boundary rows with the
 * stencil truncated to the grid), its entries in Acsr.c's order and its
 * diagonal last.  Acsr.c writes row_ptr[r + 1] before the row's diagonal
 * (Acsr.c:425-428, 553-556, 680-683); here the range of row r includes
 * it, so row_ptr[r + 1] is one more than the reference's.  Returns nnz;
 * with non-NULL arrays (rows + 1, nnz, nnz) fills them. */
long orc_csr_build(const orc_t *m, int *row_ptr, int *col, orc_c128 *val)
{
  const int n1 = m->n1, n2 = m->n2, n3 = m->n3f;
  const long lEx = (long)n1 * (n2 + 1) * (n3 + 1), lEy = (long)(n1 + 1) * n2 * (n3 + 1);
  const double *d1 = m->d1, *d2 = m->d2, *d3 = m->d3, *d1s = m->d1s, *d2s = m->d2s, *d3s = m->d3s;
  long nnz = 0, row = 0;
  int i, j, k;
  double t;
  zc aii; /* complex, as Acsr.c's: its imaginary part is +0 before the diagonal's subtraction */
#define CX(k, j, i) ((long)(i) + (long)n1 * ((j) + (long)(n2 + 1) * (k)))
#define CY(k, j, i) (lEx + (long)(i) + (long)(n1 + 1) * ((j) + (long)n2 * (k)))
#define CZ(k, j, i) (lEx + lEy + (long)(i) + (long)(n1 + 1) * ((j) + (long)(n2 + 1) * (k)))
#define PUT(c, v) do { if (col) { col[nnz] = (int)(c); val[nnz] = (v); } nnz++; } while (0)
#define DIAG(s) do { PUT(row, aii - m->iomu * (s)); row++; if (row_ptr) row_ptr[row] = (int)nnz; } while (0)
  if (row_ptr) row_ptr[0] = 0;
  for (k = 0; k <= n3; k++)
    for (j = 0; j <= n2; j++)
      for (i = 0; i < n1; i++) { /* Ex rows, Acsr.c:305-428 */
        aii = 0;
        t = 1. / (d2[j] * d1s[i]);
        if (i + 1 <= n1 && j < n2) PUT(CY(k, j, i + 1), t);
        if (j < n2) PUT(CY(k, j, i), -t);
        t = 1. / (d2[j] * d2s[j]);
        if (j + 1 <= n2) PUT(CX(k, j + 1, i), -t);
        aii += t;
        t = 1. / (d2[j] * d1s[i]);
        if (i + 1 <= n1 && j - 1 >= 0) PUT(CY(k, j - 1, i + 1), -t);
        if (j - 1 >= 0) PUT(CY(k, j - 1, i), t);
        t = 1. / (d2[j] * d2s[MAXI(j - 1, 0)]);
        if (j - 1 >= 0) PUT(CX(k, j - 1, i), -t);
        aii += t;
        t = 1. / (d3[k] * d3s[k]);
        if (k + 1 <= n3) PUT(CX(k + 1, j, i), -t);
        aii += t;
        t = 1. / (d3[k] * d1s[i]);
        if (i + 1 <= n1 && k < n3) PUT(CZ(k, j, i + 1), t);
        if (k < n3) PUT(CZ(k, j, i), -t);
        t = 1. / (d3[k] * d3s[MAXI(k - 1, 0)]);
        aii += t;
        if (k - 1 >= 0) PUT(CX(k - 1, j, i), -t);
        t = 1. / (d3[k] * d1s[i]);
        if (i + 1 <= n1 && k - 1 >= 0) PUT(CZ(k - 1, j, i + 1), -t);
        if (k - 1 >= 0) PUT(CZ(k - 1, j, i), t);
        DIAG(m->s11[CX(k, j, i)]);
      }
  for (k = 0; k <= n3; k++)
    for (j = 0; j < n2; j++)
      for (i = 0; i <= n1; i++) { /* Ey rows, Acsr.c:436-556 */
        aii = 0;
        t = 1. / (d3[k] * d2s[j]);
        if (j + 1 <= n2 && k < n3) PUT(CZ(k, j + 1, i), t);
        if (k < n3) PUT(CZ(k, j, i), -t);
        t = 1. / (d3[k] * d3s[k]);
        if (k + 1 <= n3) PUT(CY(k + 1, j, i), -t);
        aii += t;
        t = 1. / (d3[k] * d2s[j]);
        if (j + 1 <= n2 && k - 1 >= 0) PUT(CZ(k - 1, j + 1, i), -t);
        if (k - 1 >= 0) PUT(CZ(k - 1, j, i), t);
        t = 1. / (d3[k] * d3s[MAXI(k - 1, 0)]);
        aii += t;
        if (k - 1 >= 0) PUT(CY(k - 1, j, i), -t);
        t = 1. / (d1[i] * d1s[i]);
        if (i + 1 <= n1) PUT(CY(k, j, i + 1), -t);
        aii += t;
        t = 1. / (d1[i] * d2s[j]);
        if (i < n1 && j + 1 <= n2) PUT(CX(k, j + 1, i), t);
        if (i < n1) PUT(CX(k, j, i), -t);
        t = 1. / (d1[i] * d1s[MAXI(i - 1, 0)]);
        aii += t;
        if (i - 1 >= 0) PUT(CY(k, j, i - 1), -t);
        t = 1. / (d1[i] * d2s[j]);
        if (i - 1 >= 0 && j + 1 <= n2) PUT(CX(k, j + 1, i - 1), -t);
        if (i - 1 >= 0) PUT(CX(k, j, i - 1), t);
        DIAG(m->s22[CY(k, j, i) - lEx]);
      }
  for (k = 0; k < n3; k++)
    for (j = 0; j <= n2; j++)
      for (i = 0; i <= n1; i++) { /* Ez rows, Acsr.c:562-683 */
        aii = 0;
        t = 1. / (d1[i] * d3s[k]);
        if (i < n1 && k + 1 <= n3) PUT(CX(k + 1, j, i), t);
        if (i < n1) PUT(CX(k, j, i), -t);
        t = 1. / (d1[i] * d1s[i]);
        if (i + 1 <= n1) PUT(CZ(k, j, i + 1), -t);
        aii += t;
        t = 1. / (d1[i] * d3s[k]);
        if (i - 1 >= 0 && k + 1 <= n3) PUT(CX(k + 1, j, i - 1), -t);
        if (i - 1 >= 0) PUT(CX(k, j, i - 1), t);
        t = 1. / (d1[i] * d1s[MAXI(i - 1, 0)]);
        aii += t;
        if (i - 1 >= 0) PUT(CZ(k, j, i - 1), -t);
        t = 1. / (d2[j] * d2s[j]);
        if (j + 1 <= n2) PUT(CZ(k, j + 1, i), -t);
        aii += t;
        t = 1. / (d2[j] * d3s[k]);
        if (j < n2 && k + 1 <= n3) PUT(CY(k + 1, j, i), t);
        if (j < n2) PUT(CY(k, j, i), -t);
        t = 1. / (d2[j] * d2s[MAXI(j - 1, 0)]);
        aii += t;
        if (j - 1 >= 0) PUT(CZ(k, j - 1, i), -t);
        t = 1. / (d2[j] * d3s[k]);
        if (j - 1 >= 0 && k + 1 <= n3) PUT(CY(k + 1, j - 1, i), -t);
        if (j - 1 >= 0) PUT(CY(k, j - 1, i), t);
        DIAG(m->s33[CZ(k, j, i) - lEx - lEy]);
      }
#undef CX
#undef CY
#undef CZ
#undef PUT
#undef DIAG
  return nnz;
}

/* Acsr_apply (Acsr.c:704-716), each row over its own range */
void orc_csr_apply(long nrow, const int *row_ptr, const int *col, const orc_c128 *val, const orc_c128 *x,
                   orc_c128 *y)
{
  long r;
  int e;
  for (r = 0; r < nrow; r++) {
    y[r] = 0;
    for (e = row_ptr[r]; e < row_ptr[r + 1]; e++) y[r] += val[e] * x[col[e]];
  }
}

/* Acsr_lusgs_apply (Acsr.c:719-748): y = 0, a forward Gauss-Seidel sweep
 * over the rows, then a backward one */
void orc_csr_sgs(long nrow, const int *row_ptr, const int *col, const orc_c128 *val, const orc_c128 *x,
                 orc_c128 *y)
{
  long r;
  int e, pass;
  memset(y, 0, nrow * sizeof(zc));
  for (pass = 0; pass < 2; pass++)
    for (r = pass ? nrow - 1 : 0; pass ? r >= 0 : r < nrow; r += pass ? -1 : 1) {
      zc s = 0, aii = 1;
      for (e = row_ptr[r]; e < row_ptr[r + 1]; e++) {
        if (col[e] == r) aii = val[e];
        else s += val[e] * y[col[e]];
      }
      y[r] = (x[r] - s) / aii;
    }
}
