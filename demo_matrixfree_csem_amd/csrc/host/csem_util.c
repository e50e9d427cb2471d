/* csem_util.c -- command-line parameters, contiguous arrays and error
 * reporting with the semantics the reference's callers rely on.
 *
 * Parameters (cstd.c:936-1300, SU getpar): argv holds "key=value" tokens;
 * tokens without '=' are ignored; the LAST occurrence of a key wins; a value
 * may be a comma-separated vector; "par=file" pulls in more key=value tokens
 * from a text file (whitespace separated, '#' starts a comment).
 * Arrays (cstd.c:213-700): allocN(n1, ..., nN) returns row pointers into
 * one contiguous block with n1 the fastest dimension.
 */
#include <ctype.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/csem.h"

int iproc = 0, nproc = 1, ierr = 0;

/* ------------------------------------------------------------ errors */

static int g_status = 0;
static int g_exit_on_error = 1;
int csem_fail(int code, const char *fmt, ...); /* device library: records the message */

int csem_last_status(void) { return g_status; }
void csem_set_exit_on_error(int on) { g_exit_on_error = on; }
void csem_set_status(int st) { g_status = st; }

void err(char *fmt, ...)
{
  va_list ap;
  char msg[1024];
  va_start(ap, fmt);
  vsnprintf(msg, sizeof msg, fmt, ap);
  va_end(ap);
  fprintf(stderr, "\n%s\n", msg);
  /* keep a specific code the caller set just before; the message goes
   * where csem_last_error() reads it */
  if (g_status >= 0) g_status = -100;
  csem_fail(g_status, "%s", msg);
  if (g_exit_on_error) exit(1);
}

void warn(char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "warning: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
}

/* ---------------------------------------------------------- getpar */

typedef struct {
  char *key, *val;
} kv_t;
static kv_t *g_kv = NULL;
static int g_nkv = 0, g_capkv = 0;

static void kv_push(const char *tok)
{
  const char *eq = strchr(tok, '=');
  if (!eq || eq == tok) return; /* tokens without '=' are ignored */
  if (g_nkv == g_capkv) {
    g_capkv = g_capkv ? 2 * g_capkv : 64;
    g_kv = (kv_t *)realloc(g_kv, g_capkv * sizeof(kv_t));
  }
  g_kv[g_nkv].key = strndup(tok, eq - tok);
  g_kv[g_nkv].val = strdup(eq + 1);
  g_nkv++;
}

static void kv_file(const char *fn)
{
  FILE *f = fopen(fn, "r");
  char tok[4096];
  int c, n = 0;
  if (!f) err("cannot open par file %s", fn);
  if (!f) return;
  for (;;) {
    c = fgetc(f);
    if (c == '#') {
      while (c != EOF && c != '\n') c = fgetc(f);
    }
    if (c == EOF || isspace(c)) {
      if (n) {
        tok[n] = 0;
        kv_push(tok);
        n = 0;
      }
      if (c == EOF) break;
      continue;
    }
    if (n < (int)sizeof tok - 1) tok[n++] = (char)c;
  }
  fclose(f);
}

void initargs(int argc, char **argv)
{
  int i;
  for (i = 0; i < g_nkv; i++) {
    free(g_kv[i].key);
    free(g_kv[i].val);
  }
  g_nkv = 0;
  for (i = 1; i < argc; i++) {
    if (!strncmp(argv[i], "par=", 4)) kv_file(argv[i] + 4);
    else kv_push(argv[i]);
  }
}

/* index of the n-th occurrence counted from the last (n=0: last) */
static int kv_find(int n, const char *name)
{
  int i, seen = 0;
  for (i = g_nkv - 1; i >= 0; i--)
    if (!strcmp(g_kv[i].key, name)) {
      if (seen == n) return i;
      seen++;
    }
  return -1;
}

int countparname(char *name)
{
  int i, c = 0;
  for (i = 0; i < g_nkv; i++)
    if (!strcmp(g_kv[i].key, name)) c++;
  return c;
}

int countnparval(int n, char *name)
{
  int i = kv_find(n, name), c = 1;
  const char *p;
  if (i < 0) return 0;
  for (p = g_kv[i].val; *p; p++)
    if (*p == ',') c++;
  return c;
}

int countparval(char *name) { return countnparval(0, name); }

/* parse the value list of one occurrence into typed storage */
static int getn(int n, const char *name, char type, void *ptr)
{
  int i = kv_find(n, name), c = 0;
  char *s, *tok, *save = NULL;
  if (i < 0) return 0;
  s = strdup(g_kv[i].val);
  for (tok = strtok_r(s, ",", &save); tok; tok = strtok_r(NULL, ",", &save), c++) {
    switch (type) {
    case 'i': ((int *)ptr)[c] = (int)strtol(tok, NULL, 10); break;
    case 'f': ((float *)ptr)[c] = (float)strtod(tok, NULL); break;
    case 'd': ((double *)ptr)[c] = strtod(tok, NULL); break;
    case 'a': ((char **)ptr)[c] = strdup(tok); break;
    }
  }
  free(s);
  return c;
}

int getnparint(int n, char *name, int *p) { return getn(n, name, 'i', p); }
int getnparfloat(int n, char *name, float *p) { return getn(n, name, 'f', p); }
int getnpardouble(int n, char *name, double *p) { return getn(n, name, 'd', p); }
int getparint(char *name, int *p) { return getn(0, name, 'i', p); }
int getparfloat(char *name, float *p) { return getn(0, name, 'f', p); }
int getpardouble(char *name, double *p) { return getn(0, name, 'd', p); }
int getparstringarray(char *name, char **p) { return getn(0, name, 'a', p); }

int getnparstring(int n, char *name, char **p)
{
  int i = kv_find(n, name);
  if (i < 0) return 0;
  *p = g_kv[i].val; /* whole value, commas included */
  return 1;
}
int getparstring(char *name, char **p) { return getnparstring(0, name, p); }

/* ---------------------------------------------------------- arrays */

void *alloc1(size_t n1, size_t size)
{
  void *p = malloc(n1 * size ? n1 * size : 1);
  if (!p) err("alloc1: out of memory");
  return p;
}
void free1(void *p) { free(p); }

void **alloc2(size_t n1, size_t n2, size_t size)
{
  size_t i2;
  void **p = (void **)malloc(n2 * sizeof(void *));
  if (!p) err("alloc2: out of memory");
  p[0] = alloc1(n1 * n2, size);
  for (i2 = 0; i2 < n2; i2++) p[i2] = (char *)p[0] + size * n1 * i2;
  return p;
}
void free2(void **p)
{
  free(p[0]);
  free(p);
}

void ***alloc3(size_t n1, size_t n2, size_t n3, size_t size)
{
  size_t i3, i2;
  void ***p = (void ***)malloc(n3 * sizeof(void **));
  if (!p) err("alloc3: out of memory");
  p[0] = (void **)malloc(n2 * n3 * sizeof(void *));
  p[0][0] = alloc1(n1 * n2 * n3, size);
  for (i3 = 0; i3 < n3; i3++) {
    p[i3] = p[0] + n2 * i3;
    for (i2 = 0; i2 < n2; i2++) p[i3][i2] = (char *)p[0][0] + size * n1 * (i2 + n2 * i3);
  }
  return p;
}
void free3(void ***p)
{
  free(p[0][0]);
  free(p[0]);
  free(p);
}

int *alloc1int(size_t n1) { return (int *)alloc1(n1, sizeof(int)); }
void free1int(int *p) { free1(p); }
float *alloc1float(size_t n1) { return (float *)alloc1(n1, sizeof(float)); }
void free1float(float *p) { free1(p); }
float ***alloc3float(size_t n1, size_t n2, size_t n3) { return (float ***)alloc3(n1, n2, n3, sizeof(float)); }
void free3float(float ***p) { free3((void ***)p); }
double *alloc1double(size_t n1) { return (double *)alloc1(n1, sizeof(double)); }
void free1double(double *p) { free1(p); }
double ***alloc3double(size_t n1, size_t n2, size_t n3) { return (double ***)alloc3(n1, n2, n3, sizeof(double)); }
void free3double(double ***p) { free3((void ***)p); }
csem_complex *alloc1complex(size_t n1) { return (csem_complex *)alloc1(n1, sizeof(csem_complex)); }
void free1complex(csem_complex *p) { free1(p); }
csem_complex **alloc2complex(size_t n1, size_t n2) { return (csem_complex **)alloc2(n1, n2, sizeof(csem_complex)); }
void free2complex(csem_complex **p) { free2((void **)p); }
float _Complex ***alloc3complexf(size_t n1, size_t n2, size_t n3)
{
  return (float _Complex ***)alloc3(n1, n2, n3, sizeof(float _Complex));
}
void free3complexf(float _Complex ***p) { free3((void ***)p); }
