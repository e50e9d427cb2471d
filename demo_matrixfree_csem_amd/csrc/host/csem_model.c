/* csem_model.c -- model and acquisition input, grid extension and edge
 * conductivities, and the receiver-data writer (host side of the boundary).
 *
 * The formulas are the reference's (the extended model must be the same
 * numbers for the solve to be the same): node coordinates and padding
 * (extend_model.c:33-124), cell fill including the bottom-padding row copy
 * of extend_model.c:169-173, edge averaging (extend_model.c:180-252); input
 * formats of emf_init_free.c:77-166 and acq_init_free.c:38-80; output format
 * of write_data.c:24-34.  Float/double conversions follow the reference's
 * field types (lextend, rstretch, rho_* are float in emf_t).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/csem.h"

void csem_set_status(int st);

#define CSEM_MAX(a, b) ((a) > (b) ? (a) : (b))
#define CSEM_MIN(a, b) ((a) < (b) ? (a) : (b))

/* find_index.c:9-30: k with x[k] <= val < x[k+1], by bisection */
int find_index(int n, double *x, double val)
{
  int lo = 0, hi = n - 1, mid;
  while (lo < hi) {
    mid = (lo + hi) / 2;
    if (x[mid] <= val) lo = mid;
    if (x[mid] > val) hi = mid;
    if (lo == hi || lo == hi - 1) break;
  }
  return lo;
}

/* create_nugrid.c:16-40: geometric padding of n cells spanning len from a
 * first cell dx; the ratio q solves q = (len (q-1)/dx + 1)^(1/n) by fixed
 * point iteration (start 1.1, |dq| < 1e-15, at most 100 steps). */
double create_nugrid(int n, double len, double dx, double *x, int istretch)
{
  double q = 1.1, qn;
  int it, i;
  if (n * dx >= len || !istretch) {
    for (i = 0; i <= n; i++) x[i] = i * dx;
    return 1;
  }
  for (it = 0; it < 100; ++it) {
    qn = pow(len * (q - 1.) / dx + 1., 1. / n);
    if (fabs(qn - q) < 1e-15) break;
    q = qn;
  }
  x[0] = 0;
  for (i = 1; i <= n; i++) x[i] = (pow(q, i) - 1.) * dx / (q - 1.);
  return q;
}

static int cmp_float_bits(const void *a, const void *b) { return *(const int *)a - *(const int *)b; }

static long file_floats(const char *fn)
{
  FILE *fp = fopen(fn, "rb");
  long sz;
  if (!fp) err("emf: unable to open %s", fn);
  if (!fp) return -1;
  fseek(fp, 0, SEEK_END);
  sz = ftell(fp);
  fclose(fp);
  return sz / (long)sizeof(float);
}

static void read_floats(const char *fn, float *dst, long n, const char *what)
{
  FILE *fp = fopen(fn, "rb");
  size_t got;
  if (!fp) err("emf: unable to open %s (%s=)", fn, what);
  if (!fp) return;
  got = fread(dst, sizeof(float), n, fp);
  fclose(fp);
  if ((long)got != n) err("emf: %s holds %d values, the grid needs %ld", what, (int)got, n);
}

/* emf_init_free.c:14-172 */
void emf_init(emf_t *emf)
{
  char *fx[3], *frho[3];
  int i;
  long n;
  csem_set_status(0);
  if (!(emf->nfreq = countparval("freqs"))) {
    err("emf: no frequencies given (freqs=f1,f2,...)");
    return;
  }
  emf->freqs = alloc1float(emf->nfreq);
  getparfloat("freqs", emf->freqs);
  /* ascending; the comparator compares the float bit patterns as int,
   * which orders positive frequencies numerically (emf_init_free.c:12,27) */
  qsort(emf->freqs, emf->nfreq, sizeof(float), cmp_float_bits);
  if ((emf->nchsrc = countparval("chsrc")) != 0) {
    emf->chsrc = (char **)alloc1(emf->nchsrc, sizeof(void *));
    getparstringarray("chsrc", emf->chsrc);
  } else {
    emf->nchsrc = 1;
    emf->chsrc = (char **)alloc1(1, sizeof(void *));
    emf->chsrc[0] = "Ex";
  }
  if ((emf->nchrec = countparval("chrec")) != 0) {
    emf->chrec = (char **)alloc1(emf->nchrec, sizeof(void *));
    getparstringarray("chrec", emf->chrec);
  } else {
    emf->nchrec = 1;
    emf->chrec = (char **)alloc1(1, sizeof(void *));
    emf->chrec[0] = "Ex";
  }
  if (!getparint("reciprocity", &emf->reciprocity)) emf->reciprocity = 0;
  if (!getparint("fvm", &emf->fvm)) emf->fvm = 0;
  if (!getparint("nb", &emf->nb)) emf->nb = 15;
  if (!getparfloat("rho_air", &emf->rho_air)) emf->rho_air = 1e6;
  if (!getparint("istretch", &emf->istretch)) emf->istretch = 1;
  if (!getparint("airbc", &emf->airbc)) emf->airbc = 0;
  if (emf->airbc && emf->nb == 0) {
    err("emf: the airwave boundary (airbc=1) needs padding cells below the model (nb >= 1)");
    return;
  }
  if (emf->verb) {
    printf("freqs=");
    for (i = 0; i < emf->nfreq; i++) printf("%g,", emf->freqs[i]);
    printf("\nfvm=%d airbc=%d istretch=%d rho_air=%g nb=%d\n", emf->fvm, emf->airbc, emf->istretch,
           emf->rho_air, emf->nb);
  }
  if (!getparstring("fx1", &fx[0])) {
    err("emf: missing node file fx1=");
    return;
  }
  if (!getparstring("fx2", &fx[1])) {
    err("emf: missing node file fx2=");
    return;
  }
  if (!getparstring("fx3", &fx[2])) {
    err("emf: missing node file fx3=");
    return;
  }
  emf->nx = (int)file_floats(fx[0]) - 1;
  emf->ny = (int)file_floats(fx[1]) - 1;
  emf->nz = (int)file_floats(fx[2]) - 1;
  if (csem_last_status() < 0) return;
  if (emf->nx < 1 || emf->ny < 1 || emf->nz < 1) {
    err("grid files must hold at least 2 nodes");
    return;
  }
  emf->x1node = alloc1float(emf->nx + 1);
  emf->x2node = alloc1float(emf->ny + 1);
  emf->x3node = alloc1float(emf->nz + 1);
  read_floats(fx[0], emf->x1node, emf->nx + 1, "fx1");
  read_floats(fx[1], emf->x2node, emf->ny + 1, "fx2");
  read_floats(fx[2], emf->x3node, emf->nz + 1, "fx3");
  if (csem_last_status() < 0) return;
  emf->rho11 = alloc3float(emf->nx, emf->ny, emf->nz);
  emf->rho22 = alloc3float(emf->nx, emf->ny, emf->nz);
  emf->rho33 = alloc3float(emf->nx, emf->ny, emf->nz);
  n = (long)emf->nx * emf->ny * emf->nz;
  if (!getparstring("frho11", &frho[0])) {
    err("emf: missing resistivity file frho11=");
    return;
  }
  if (!getparstring("frho22", &frho[1])) {
    err("emf: missing resistivity file frho22=");
    return;
  }
  if (!getparstring("frho33", &frho[2])) {
    err("emf: missing resistivity file frho33=");
    return;
  }
  read_floats(frho[0], &emf->rho11[0][0][0], n, "frho11");
  read_floats(frho[1], &emf->rho22[0][0][0], n, "frho22");
  read_floats(frho[2], &emf->rho33[0][0][0], n, "frho33");
  if (csem_last_status() < 0) return;
  {
    /* rho_water is the smallest resistivity of the input model */
    const float *a = &emf->rho11[0][0][0], *b = &emf->rho22[0][0][0], *c = &emf->rho33[0][0][0];
    long q;
    emf->rhomax = emf->rhomin = a[0];
    for (q = 0; q < n; q++) {
      emf->rhomax = CSEM_MAX(emf->rhomax, a[q]);
      emf->rhomin = CSEM_MIN(emf->rhomin, a[q]);
      emf->rhomax = CSEM_MAX(emf->rhomax, b[q]);
      emf->rhomin = CSEM_MIN(emf->rhomin, b[q]);
      emf->rhomax = CSEM_MAX(emf->rhomax, c[q]);
      emf->rhomin = CSEM_MIN(emf->rhomin, c[q]);
    }
    emf->rho_water = emf->rhomin;
  }
}

void emf_free(emf_t *emf)
{
  free1float(emf->freqs);
  free1float(emf->x1node);
  free1float(emf->x2node);
  free1float(emf->x3node);
  free3float(emf->rho11);
  free3float(emf->rho22);
  free3float(emf->rho33);
}

/* one whitespace table: a header line, then rows "x y z azimuth dip idx" */
static int read_table(const char *fn, float **cols, int cap)
{
  FILE *fp = fopen(fn, "r");
  float x, y, z, hd, pit;
  int idx, n = 0, c;
  if (!fp) err("acq: unable to open %s", fn);
  if (!fp) return 0;
  while ((c = fgetc(fp)) != EOF && c != '\n') {
  }
  while (fscanf(fp, "%f %f %f %f %f %d", &x, &y, &z, &hd, &pit, &idx) == 6) {
    if (n == cap) {
      cap *= 2;
      for (c = 0; c < 5; c++) cols[c] = (float *)realloc(cols[c], cap * sizeof(float));
    }
    cols[0][n] = x; cols[1][n] = y; cols[2][n] = z; cols[3][n] = hd; cols[4][n] = pit;
    n++;
  }
  fclose(fp);
  return n;
}

/* acq_init_free.c:14-148: the first source row only (nsrc = 1), all
 * receivers, bounds check against the input model. */
void acq_init(acq_t *acq, emf_t *emf)
{
  char *fsrc, *frec;
  float *sc[5], *rc[5];
  int i, ns, nr;
  if (!getparfloat("lensrc", &acq->lensrc)) acq->lensrc = 278;
  if (!getparfloat("lenrec", &acq->lenrec)) acq->lenrec = 8;
  if (!getparint("nsubsrc", &acq->nsubsrc)) acq->nsubsrc = 1;
  if (!getparint("nsubrec", &acq->nsubrec)) acq->nsubrec = 1;
  if (!getparstring("fsrc", &fsrc)) {
    err("acq: missing source table fsrc=");
    return;
  }
  if (!getparstring("frec", &frec)) {
    err("acq: missing receiver table frec=");
    return;
  }
  for (i = 0; i < 5; i++) {
    sc[i] = (float *)malloc(16 * sizeof(float));
    rc[i] = (float *)malloc(16 * sizeof(float));
  }
  ns = read_table(fsrc, sc, 16);
  nr = read_table(frec, rc, 16);
  if (csem_last_status() < 0) {
    for (i = 0; i < 5; i++) {
      free(sc[i]);
      free(rc[i]);
    }
    return;
  }
  if (ns < 1) {
    err("no source in %s", fsrc);
    return;
  }
  acq->nsrc = 1;
  acq->nrec = nr;
  acq->src_x1 = alloc1float(1); acq->src_x2 = alloc1float(1); acq->src_x3 = alloc1float(1);
  acq->src_azimuth = alloc1float(1); acq->src_dip = alloc1float(1);
  acq->src_x1[0] = sc[0][0]; acq->src_x2[0] = sc[1][0]; acq->src_x3[0] = sc[2][0];
  acq->src_azimuth[0] = sc[3][0]; acq->src_dip[0] = sc[4][0];
  acq->x1min = acq->x1max = acq->src_x1[0];
  acq->x2min = acq->x2max = acq->src_x2[0];
  acq->x3min = acq->x3max = acq->src_x3[0];
  acq->rec_x1 = alloc1float(nr); acq->rec_x2 = alloc1float(nr); acq->rec_x3 = alloc1float(nr);
  acq->rec_azimuth = alloc1float(nr); acq->rec_dip = alloc1float(nr);
  for (i = 0; i < nr; i++) {
    acq->rec_x1[i] = rc[0][i]; acq->rec_x2[i] = rc[1][i]; acq->rec_x3[i] = rc[2][i];
    acq->rec_azimuth[i] = rc[3][i]; acq->rec_dip[i] = rc[4][i];
    acq->x1min = CSEM_MIN(acq->x1min, acq->rec_x1[i]); acq->x1max = CSEM_MAX(acq->x1max, acq->rec_x1[i]);
    acq->x2min = CSEM_MIN(acq->x2min, acq->rec_x2[i]); acq->x2max = CSEM_MAX(acq->x2max, acq->rec_x2[i]);
    acq->x3min = CSEM_MIN(acq->x3min, acq->rec_x3[i]); acq->x3max = CSEM_MAX(acq->x3max, acq->rec_x3[i]);
  }
  for (i = 0; i < 5; i++) {
    free(sc[i]);
    free(rc[i]);
  }
  if (emf->x1node[0] > acq->x1min) {
    err("acq: a source or receiver lies left of the model in x (below x1node[0])");
    return;
  }
  if (emf->x1node[emf->nx] < acq->x1max) {
    err("acq: a source or receiver lies right of the model in x (beyond x1node[nx])");
    return;
  }
  if (emf->x2node[0] > acq->x2min) {
    err("acq: a source or receiver lies below the model range in y (below x2node[0])");
    return;
  }
  if (emf->x2node[emf->ny] < acq->x2max) {
    err("acq: a source or receiver lies beyond the model range in y (beyond x2node[ny])");
    return;
  }
  if (emf->x3node[0] > acq->x3min) {
    err("acq: a source or receiver lies above the model top (above x3node[0])");
    return;
  }
  if (emf->x3node[emf->nz] < acq->x3max) {
    err("acq: a source or receiver lies below the model bottom (beyond x3node[nz])");
    return;
  }
}

void acq_free(acq_t *acq)
{
  free(acq->src_x1); free(acq->src_x2); free(acq->src_x3); free(acq->src_azimuth); free(acq->src_dip);
  free(acq->rec_x1); free(acq->rec_x2); free(acq->rec_x3); free(acq->rec_azimuth); free(acq->rec_dip);
}

/* staggered coordinates of one axis (extend_model.c:8-21): xs at cell
 * centres, ds cell widths, d distances between centres */
static void stagger_axis(int n, const double *x, double *xs, double *d, double *ds)
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

/* nodes of one axis: input nodes in the middle, nb stretched cells on each
 * side (extend_model.c:72-106).  The stretching ratio is stored as float. */
static void extend_axis(emf_t *emf, const float *node, int nn, double *x)
{
  int nb = emf->nb, i;
  double dx;
  for (i = 0; i <= nn; i++) x[i + nb] = node[i];
  if (nb <= 0) return;
  double *scratch = alloc1double(nb + 1);
  dx = node[1] - node[0];
  emf->rstretch = create_nugrid(nb, emf->lextend, dx, scratch, emf->istretch);
  for (i = 0; i < nb; i++) x[nb - 1 - i] = x[nb - i] - dx * pow(emf->rstretch, i);
  dx = node[nn] - node[nn - 1];
  emf->rstretch = create_nugrid(nb, emf->lextend, dx, scratch, emf->istretch);
  for (i = 0; i < nb; i++) x[nb + nn + 1 + i] = x[nb + nn + i] + dx * pow(emf->rstretch, i);
  free1double(scratch);
}

/* The 1-D half of extend_model_init: extended and staggered nodes and
 * spacings (extend_model.c:24-106). */
void csem_extend_grid(acq_t *acq, emf_t *emf)
{
  const int nb = emf->nb;
  int i, j, k, n1, n2, n3;
  double dx;
  (void)acq;
  n1 = emf->n1 = emf->nx + 2 * nb;
  n2 = emf->n2 = emf->ny + 2 * nb;
  n3 = emf->n3 = emf->nz + 2 * nb;
  emf->d1min = emf->x1node[1] - emf->x1node[0];
  for (i = 0; i < emf->nx; i++) emf->d1min = CSEM_MIN(emf->d1min, emf->x1node[i + 1] - emf->x1node[i]);
  emf->d2min = emf->x2node[1] - emf->x2node[0];
  for (j = 0; j < emf->ny; j++) emf->d2min = CSEM_MIN(emf->d2min, emf->x2node[j + 1] - emf->x2node[j]);
  emf->d3min = emf->x3node[1] - emf->x3node[0];
  for (k = 0; k < emf->nz; k++) emf->d3min = CSEM_MIN(emf->d3min, emf->x3node[k + 1] - emf->x3node[k]);
  /* padding width: default 40 km, at least 2.2 times each model extent */
  if (!getparfloat("lextend", &emf->lextend)) emf->lextend = (nb > 0) ? 40e3 : 0;
  if (nb > 0) {
    dx = emf->x1node[emf->nx] - emf->x1node[0];
    emf->lextend = CSEM_MAX(emf->lextend, 2.2 * dx);
    dx = emf->x2node[emf->ny] - emf->x2node[0];
    emf->lextend = CSEM_MAX(emf->lextend, 2.2 * dx);
    dx = emf->x3node[emf->nz] - emf->x3node[0];
    emf->lextend = CSEM_MAX(emf->lextend, 2.2 * dx);
  }
  emf->x1 = alloc1double(n1 + 1); emf->x2 = alloc1double(n2 + 1); emf->x3 = alloc1double(n3 + 1);
  emf->x1s = alloc1double(n1 + 1); emf->x2s = alloc1double(n2 + 1); emf->x3s = alloc1double(n3 + 1);
  emf->d1 = alloc1double(n1 + 1); emf->d2 = alloc1double(n2 + 1); emf->d3 = alloc1double(n3 + 1);
  emf->d1s = alloc1double(n1 + 1); emf->d2s = alloc1double(n2 + 1); emf->d3s = alloc1double(n3 + 1);
  extend_axis(emf, emf->x1node, emf->nx, emf->x1);
  extend_axis(emf, emf->x2node, emf->ny, emf->x2);
  extend_axis(emf, emf->x3node, emf->nz, emf->x3);
  stagger_axis(n1, emf->x1, emf->x1s, emf->d1, emf->d1s);
  stagger_axis(n2, emf->x2, emf->x2s, emf->d2, emf->d2s);
  stagger_axis(n3, emf->x3, emf->x3s, emf->d3, emf->d3s);
}

void extend_model_init(acq_t *acq, emf_t *emf)
{
  const int nb = emf->nb;
  int i, j, k, n1, n2, n3;
  csem_extend_grid(acq, emf);
  n1 = emf->n1;
  n2 = emf->n2;
  n3 = emf->n3;

  /* cell conductivity (11, 22) and resistivity (33) on the extended grid */
  double ***c11 = alloc3double(n1, n2, n3), ***c22 = alloc3double(n1, n2, n3), ***r33 = alloc3double(n1, n2, n3);
  for (k = 0; k < emf->nz; k++)
    for (j = 0; j < emf->ny; j++)
      for (i = 0; i < emf->nx; i++) {
        c11[k + nb][j + nb][i + nb] = 1. / emf->rho11[k][j][i];
        c22[k + nb][j + nb][i + nb] = 1. / emf->rho22[k][j][i];
        r33[k + nb][j + nb][i + nb] = emf->rho33[k][j][i];
      }
  if (nb > 0) {
    double ***arr[3] = {c11, c22, r33};
    int a;
    for (a = 0; a < 3; a++) {
      double ***c = arr[a];
      for (k = 0; k < n3; k++)
        for (j = 0; j < n2; j++)
          for (i = 0; i < nb; i++) {
            c[k][j][i] = c[k][j][nb];
            c[k][j][n1 - 1 - i] = c[k][j][n1 - 1 - nb];
          }
      for (k = 0; k < n3; k++)
        for (j = 0; j < nb; j++)
          for (i = 0; i < n1; i++) {
            c[k][j][i] = c[k][nb][i];
            c[k][n2 - 1 - j][i] = c[k][n2 - 1 - nb][i];
          }
      /* air above; below, padding layer k copies row j=k of the last core
       * layer (extend_model.c:169,171,173 index [k] where [j] is meant) */
      for (k = 0; k < nb; k++)
        for (j = 0; j < n2; j++)
          for (i = 0; i < n1; i++) {
            c[k][j][i] = (a == 2) ? emf->rho_air : 1. / emf->rho_air;
            c[n3 - 1 - k][j][i] = c[n3 - 1 - nb][k][i];
          }
    }
  }
  /* edge averages (extend_model.c:180-252) */
  double ***s11 = emf->sigma11 = alloc3double(n1, n2 + 1, n3 + 1);
  double ***s22 = emf->sigma22 = alloc3double(n1 + 1, n2, n3 + 1);
  double ***s33 = emf->sigma33 = alloc3double(n1 + 1, n2 + 1, n3);
  const double *d1 = emf->d1, *d2 = emf->d2, *d3 = emf->d3, *d1s = emf->d1s, *d2s = emf->d2s, *d3s = emf->d3s;
  for (k = 1; k < n3; k++)
    for (j = 1; j < n2; j++)
      for (i = 0; i < n1; i++)
        s11[k][j][i] = 0.25 * ((d2s[j] * c11[k][j][i] + d2s[j - 1] * c11[k][j - 1][i]) * d3s[k] +
                               (d2s[j] * c11[k - 1][j][i] + d2s[j - 1] * c11[k - 1][j - 1][i]) * d3s[k - 1]) /
                       (d2[j] * d3[k]);
  for (j = 0; j <= n2; j++)
    for (i = 0; i < n1; i++) {
      s11[0][j][i] = s11[1][j][i];
      s11[n3][j][i] = s11[n3 - 1][j][i];
    }
  for (k = 0; k <= n3; k++)
    for (i = 0; i < n1; i++) {
      s11[k][0][i] = s11[k][1][i];
      s11[k][n2][i] = s11[k][n2 - 1][i];
    }
  for (k = 1; k < n3; k++)
    for (j = 0; j < n2; j++)
      for (i = 1; i < n1; i++)
        s22[k][j][i] = 0.25 * ((d1s[i] * c22[k][j][i] + d1s[i - 1] * c22[k][j][i - 1]) * d3s[k] +
                               (d1s[i] * c22[k - 1][j][i] + d1s[i - 1] * c22[k - 1][j][i - 1]) * d3s[k - 1]) /
                       (d1[i] * d3[k]);
  for (j = 0; j < n2; j++)
    for (i = 0; i <= n1; i++) {
      s22[0][j][i] = s22[1][j][i];
      s22[n3][j][i] = s22[n3 - 1][j][i];
    }
  for (k = 0; k <= n3; k++)
    for (j = 0; j < n2; j++) {
      s22[k][j][0] = s22[k][j][1];
      s22[k][j][n1] = s22[k][j][n1 - 1];
    }
  for (k = 0; k < n3; k++)
    for (j = 1; j < n2; j++)
      for (i = 1; i < n1; i++) {
        s33[k][j][i] = 0.25 * ((d1s[i] * r33[k][j][i] + d1s[i - 1] * r33[k][j][i - 1]) * d2s[j] +
                               (d1s[i] * r33[k][j - 1][i] + d1s[i - 1] * r33[k][j - 1][i - 1]) * d2s[j - 1]) /
                       (d1[i] * d2[j]);
        s33[k][j][i] = 1. / s33[k][j][i];
      }
  for (k = 0; k < n3; k++)
    for (j = 0; j <= n2; j++) {
      s33[k][j][0] = s33[k][j][1];
      s33[k][j][n1] = s33[k][j][n1 - 1];
    }
  for (k = 0; k < n3; k++)
    for (i = 0; i <= n1; i++) {
      s33[k][0][i] = s33[k][1][i];
      s33[k][n2][i] = s33[k][n2 - 1][i];
    }
  if (!getparint("param_extended", &emf->param_extended)) emf->param_extended = 0;
  if (emf->param_extended && emf->verb) { /* extend_model.c:254-289 */
    FILE *fp;
    long nc = (long)n1 * n2 * n3, q;
    double *a = &c11[0][0][0], *b = &c22[0][0][0];
    for (q = 0; q < nc; q++) {
      a[q] = 1. / a[q];
      b[q] = 1. / b[q];
    }
    const char *names[6] = {"frho11_extended", "frho22_extended", "frho33_extended",
                            "fx1_extended", "fx2_extended", "fx3_extended"};
    const void *data[6] = {a, b, &r33[0][0][0], emf->x1, emf->x2, emf->x3};
    const long cnt[6] = {nc, nc, nc, n1 + 1, n2 + 1, n3 + 1};
    for (q = 0; q < 6; q++) {
      fp = fopen(names[q], "wb");
      if (!fp) err("cannot write %s", names[q]);
      if (!fp) continue;
      fwrite(data[q], sizeof(double), cnt[q], fp);
      fclose(fp);
    }
  }
  free3double(c11);
  free3double(c22);
  free3double(r33);
}

void extend_model_free(emf_t *emf)
{
  free1double(emf->x1); free1double(emf->x2); free1double(emf->x3);
  free1double(emf->x1s); free1double(emf->x2s); free1double(emf->x3s);
  free1double(emf->d1); free1double(emf->d2); free1double(emf->d3);
  free1double(emf->d1s); free1double(emf->d2s); free1double(emf->d3s);
  free3double(emf->sigma11); free3double(emf->sigma22); free3double(emf->sigma33);
}

/* write_data.c:15-39: header, then rows ordered channel, frequency,
 * receiver, values as float printed with %e */
/* The same table with the receivers in double precision (%.17e, enough to
 * round-trip every double): vals[(ic * nfreq + ifr) * nrec + ir], as
 * extract_emf_double leaves them for each frequency. */
void write_data_double(acq_t *acq, emf_t *emf, char *fname, csem_complex *vals)
{
  FILE *fp = fopen(fname, "w");
  int ic, ifr, ir;
  if (!fp) err("write_data_double: unable to create %s", fname);
  if (!fp) return;
  fprintf(fp, "iTx \t iRx    ichrec  ifreq \t emf_real \t emf_imag\n");
  for (ic = 0; ic < emf->nchrec; ic++)
    for (ifr = 0; ifr < emf->nfreq; ifr++)
      for (ir = 0; ir < acq->nrec; ir++) {
        const csem_complex v = vals[((long)ic * emf->nfreq + ifr) * acq->nrec + ir];
        fprintf(fp, "%d \t %d \t %d \t %d \t %.17e \t %.17e\n", 1, ir + 1, ic + 1, ifr + 1, creal(v), cimag(v));
      }
  fclose(fp);
}

void write_data(acq_t *acq, emf_t *emf, char *fname, float _Complex ***dcal_fd)
{
  FILE *fp = fopen(fname, "w");
  int ic, ifr, ir;
  if (!fp) err("write_data: unable to create %s", fname);
  if (!fp) return;
  fprintf(fp, "iTx \t iRx    ichrec  ifreq \t emf_real \t emf_imag\n");
  for (ic = 0; ic < emf->nchrec; ic++)
    for (ifr = 0; ifr < emf->nfreq; ifr++)
      for (ir = 0; ir < acq->nrec; ir++) {
        float re = crealf(dcal_fd[ic][ifr][ir]), im = cimagf(dcal_fd[ic][ifr][ir]);
        fprintf(fp, "%d \t %d \t %d \t %d \t %e \t %e\n", 1, ir + 1, ic + 1, ifr + 1, re, im);
      }
  fclose(fp);
}
