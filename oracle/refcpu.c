/* ORACLE (test infrastructure only) -- see refcpu.h.
 *
 * Restatement of NuTRL's H*v and its setup.  Reference citations are given per
 * function; the structure follows the Fortran so that a reader can check the
 * two side by side, but the code is written fresh in C.
 */
#include "refcpu.h"

#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAX_CODE 72
static const char AN_CODE[] =
    "0123456789_$cdefghijklmnopqrstuvwxyz@ABCDEFGHIJKLMNMPQRSTUVWXYZ?a%^&=b-+#";

static __thread char g_err[512];
static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
const char *refcpu_last_error(void) { return g_err; }

/* ---------------------------------------------------------------- file helpers */
typedef struct {
  unsigned char *data;
  size_t size, pos;
} ffile_t;

static int ff_open(ffile_t *f, const char *path) {
  FILE *fp = fopen(path, "rb");
  if (!fp) { set_err("cannot open %s", path); return -1; }
  fseek(fp, 0, SEEK_END);
  long n = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  f->data = (unsigned char *)malloc(n > 0 ? n : 1);
  f->size = fread(f->data, 1, n, fp);
  f->pos = 0;
  fclose(fp);
  return 0;
}
static void ff_close(ffile_t *f) { free(f->data); f->data = NULL; }
/* next sequential-unformatted record: returns payload pointer and length */
static unsigned char *ff_rec(ffile_t *f, int *len) {
  if (f->pos + 4 > f->size) return NULL;
  int32_t n;
  memcpy(&n, f->data + f->pos, 4);
  if (n < 0 || f->pos + 8 + (size_t)n > f->size) return NULL;
  unsigned char *p = f->data + f->pos + 4;
  f->pos += 8 + (size_t)n;
  *len = n;
  return p;
}
static int32_t geti(const unsigned char *p, int k) { int32_t v; memcpy(&v, p + 4 * k, 4); return v; }

static void trim(char *s) {
  char *b = s;
  while (*b == ' ' || *b == '\t') b++;
  memmove(s, b, strlen(b) + 1);
  size_t n = strlen(s);
  while (n && (s[n - 1] == ' ' || s[n - 1] == '\n' || s[n - 1] == '\r' || s[n - 1] == '\t')) s[--n] = 0;
}

/* ---------------------------------------------------------------- angular momentum */
static double lfact(int n) { return lgamma((double)n + 1.0); }
static double binm(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  return floor(exp(lfact(n) - lfact(k) - lfact(n - k)) + 0.5);
}
static double angdelta(int a, int b, int c) {
  return exp(0.5 * (lfact((a + b - c) / 2) - lfact((a + b + c) / 2 + 1) + lfact((a - b + c) / 2) +
                    lfact((-a + b + c) / 2)));
}
/* src/NuCommon/NuCommon.f90:3268-3309 */
double refcpu_sixj(int a, int b, int c, int d, int e, int f) {
  if ((a + b) % 2 != c % 2 || (c + d) % 2 != e % 2 || (a + e) % 2 != f % 2 || (b + d) % 2 != f % 2)
    return 0.0;
  if (abs(a - b) > c || a + b < c || abs(c - d) > e || c + d < e || abs(a - e) > f || a + e < f ||
      abs(b - d) > f || b + d < f)
    return 0.0;
  double out = angdelta(a, e, f) / angdelta(a, b, c) * angdelta(b, d, f) * angdelta(c, d, e);
  int nlo = (a + b + c) / 2, t;
  t = (c + d + e) / 2; if (t > nlo) nlo = t;
  t = (b + d + f) / 2; if (t > nlo) nlo = t;
  t = (a + e + f) / 2; if (t > nlo) nlo = t;
  int nhi = (a + b + d + e) / 2;
  t = (b + c + e + f) / 2; if (t < nhi) nhi = t;
  t = (a + c + d + f) / 2; if (t < nhi) nhi = t;
  double s = 0.0;
  for (int n = nlo; n <= nhi; n++) {
    double term = (n % 2) ? -1.0 : 1.0;
    term *= binm(n + 1, n - (a + b + c) / 2);
    term *= binm((a + b - c) / 2, n - (c + d + e) / 2);
    term *= binm((a - b + c) / 2, n - (b + d + f) / 2);
    term *= binm((b - a + c) / 2, n - (a + e + f) / 2);
    s += term;
  }
  return s * out;
}
/* src/NuCommon/NuCommon.f90:3255-3265 */
static double racahW(int j1, int j2, int j3, int j4, int J, int K) {
  double ph = (((j1 + j2 + j3 + j4) / 2) & 1) ? -1.0 : 1.0;
  return refcpu_sixj(j1, j2, J, j4, j3, K) * ph;
}

/* ---------------------------------------------------------------- keys */
static int keycmp(const uint8_t *a, const uint8_t *b) { return memcmp(a, b, 8); }
/* exact binary search, 1-based result, 0 if absent */
static int kfind(const uint8_t *keys, int n, const uint8_t *k) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    int c = keycmp(keys + 8 * mid, k);
    if (c == 0) return mid + 1;
    if (c < 0) lo = mid + 1; else hi = mid - 1;
  }
  return 0;
}
/* src/NuCommonX/NuCommonX.f90:147-183 (float bisection, nint rounding) */
int refcpu_tfind(const uint8_t *keys, int n, const uint8_t *k) {
  if (n == 0) return 0;
  if (keycmp(keys + 8 * (n - 1), k) == 0) return n;
  float rn = (float)n / 2.0f, drn = rn;
  for (;;) {
    int in = (int)lroundf(rn);
    if (in < 1 || in > n) return 0; /* outside the array: the reference would read out of bounds */
    if (keycmp(keys + 8 * (in - 1), k) == 0) return in;
    drn = drn / 2.0f;
    if (fabsf(drn) < 0.5f) return 0;
    if (keycmp(k, keys + 8 * (in - 1)) > 0) rn = rn + drn; else rn = rn - drn;
  }
}

/* ---------------------------------------------------------------- data */
typedef struct {
  int nsp, minJ2, maxJ2, delJ2, words, proton, neutron, nJ;
  float xtag;
  int *nJT;      /* [nJ][nsp] : nJT[(J2/2)*nsp + p] */
  int *par;      /* parity per partition */
  int *nq;       /* oscillator quanta per partition: sum of cNN * occupation */
  int *d0;       /* d(J) per J2/2 */
  int *snJT0;    /* partition offsets per J: [(J2/2)*nsp + p] */
  int ndens, max_lm;
  uint8_t *keys; /* [ndens][8] */
  int64_t *pos;
  int *dim;
  uint32_t *mask; /* [ndens][nsp cols][words] */
  float *v32;
  double *v64;
  /* like-particle H */
  int *has_mat;  /* per J2/2 */
  int64_t *posm;
  uint32_t *mmask; /* [nJ][nsp][words] */
  float *m32;
  double *m64;
} side_t;

typedef struct {
  int na;
  int *a;       /* a-density index (0-based) */
  int *boff;    /* offsets into b arrays, na+1 */
} tab_t;

struct refcpu {
  int J2, basd, tparity, fp64, thr_on, nthr;
  int64_t N;
  int *bas_a, *bas_b;
  int64_t *posnv, *dimnv;
  uint8_t *lcoef;
  uint8_t *trunc;   /* N-hbar-omega window mask (NULL when not truncated) */
  uint8_t *bas_log; /* blocks with a partition pair in the window (NULL when not truncated) */
  side_t A, B;
  int azero, bzero, nslot, maxda, maxdb;
  tab_t *tab;      /* [basd][nslot], slot = ajr/2 */
  int *tb_b, *tb_h;
  double *tb_me64;
  float *tb_me32;
  int64_t ntb;
  int sh_a0, sh_a1, sh_b0, sh_b1; /* shell ranges (1-based, inclusive) */
  int *sj2;
};

static void side_free(side_t *s) {
  free(s->nJT); free(s->par); free(s->nq); free(s->d0); free(s->snJT0); free(s->keys); free(s->pos);
  free(s->dim); free(s->mask); free(s->v32); free(s->v64); free(s->has_mat); free(s->posm);
  free(s->mmask); free(s->m32); free(s->m64);
}

void refcpu_destroy(refcpu_t *h) {
  if (!h) return;
  side_free(&h->A);
  side_free(&h->B);
  if (h->tab) {
    for (int i = 0; i < h->basd * h->nslot; i++) { free(h->tab[i].a); free(h->tab[i].boff); }
    free(h->tab);
  }
  free(h->tb_b); free(h->tb_h); free(h->tb_me64); free(h->tb_me32);
  free(h->bas_a); free(h->bas_b); free(h->posnv); free(h->dimnv); free(h->lcoef); free(h->sj2);
  free(h->trunc); free(h->bas_log);
  free(h);
}

static inline int nJTs(const side_t *s, int J2, int p) { return s->nJT[(J2 / 2) * s->nsp + p]; }
static inline int bit_of(const uint32_t *col, int i) { return (col[i >> 5] >> (i & 31)) & 1u; }

/* .tri + .trd of one side: src/NuTRL/NuCoreX.f90:838-867,1145-1206 */
static int read_side(side_t *s, const char *opfile, const int *shell_par, const int *cNN, int nshells) {
  char path[1024];
  ffile_t f;
  int len;
  snprintf(path, sizeof path, "%s.tri", opfile);
  if (ff_open(&f, path)) return -1;
  unsigned char *r = ff_rec(&f, &len);
  if (!r || len < 36) { set_err("%s: bad header", path); ff_close(&f); return -2; }
  s->nsp = geti(r, 0); s->minJ2 = geti(r, 1); s->maxJ2 = geti(r, 2);
  s->proton = geti(r, 3) != 0; s->neutron = geti(r, 4) != 0;
  memcpy(&s->xtag, r + 20, 4);
  s->delJ2 = geti(r, 8);
  s->nJ = s->maxJ2 / 2 + 1;
  s->words = (s->nsp - 1) / 32 + 1;
  int nJfile = s->maxJ2 / 2 - s->minJ2 / 2 + 1;
  s->nJT = (int *)calloc((size_t)s->nJ * s->nsp, sizeof(int));
  r = ff_rec(&f, &len);
  if (!r || len < 4 * s->nsp * nJfile) { set_err("%s: bad nJT", path); ff_close(&f); return -2; }
  for (int j = 0; j < nJfile; j++)
    for (int p = 0; p < s->nsp; p++) s->nJT[(s->minJ2 / 2 + j) * s->nsp + p] = geti(r, j * s->nsp + p);
  r = ff_rec(&f, &len);
  if (!r) { set_err("%s: no partitions", path); ff_close(&f); return -2; }
  int ptdim = len / (2 * s->nsp);
  s->par = (int *)malloc(sizeof(int) * s->nsp);
  s->nq = (int *)calloc(s->nsp, sizeof(int));
  for (int p = 0; p < s->nsp; p++) {
    int par = 1;
    for (int i = 0; i < nshells && i < ptdim; i++) {
      int16_t occ;
      memcpy(&occ, r + 2 * (p * ptdim + i), 2);
      if (occ > 0)
        for (int k = 0; k < occ; k++) par *= shell_par[i];
      s->nq[p] += cNN[i] * occ; /* psum(cNN * spart), src/NuTRL/NuCoreX.f90:1010 */
    }
    s->par[p] = par;
  }
  ff_close(&f);
  s->d0 = (int *)calloc(s->nJ, sizeof(int));
  s->snJT0 = (int *)calloc((size_t)s->nJ * s->nsp, sizeof(int));
  for (int J2 = s->minJ2; J2 <= s->maxJ2; J2 += 2)
    for (int p = 0; p < s->nsp; p++) {
      if (p > 0) s->snJT0[(J2 / 2) * s->nsp + p] = s->snJT0[(J2 / 2) * s->nsp + p - 1] + nJTs(s, J2, p - 1);
      s->d0[J2 / 2] += nJTs(s, J2, p);
    }

  snprintf(path, sizeof path, "%s.trd", opfile);
  if (ff_open(&f, path)) return -1;
  r = ff_rec(&f, &len);
  if (!r || len < 12) { set_err("%s: bad header", path); ff_close(&f); return -2; }
  s->ndens = geti(r, 0);
  s->max_lm = geti(r, 1);
  int64_t total = geti(r, 2);
  s->keys = (uint8_t *)malloc(8 * (size_t)(s->ndens ? s->ndens : 1));
  s->pos = (int64_t *)calloc(s->ndens + 1, sizeof(int64_t));
  s->dim = (int *)calloc(s->ndens + 1, sizeof(int));
  s->mask = (uint32_t *)calloc((size_t)(s->ndens ? s->ndens : 1) * s->nsp * s->words, sizeof(uint32_t));
  s->v64 = (double *)calloc(total > 0 ? total : 1, sizeof(double));
  s->v32 = (float *)calloc(total > 0 ? total : 1, sizeof(float));
  if (s->ndens) {
    r = ff_rec(&f, &len);
    memcpy(s->keys, r, 8 * (size_t)s->ndens);
    r = ff_rec(&f, &len);
    for (int k = 0; k < s->ndens; k++) s->dim[k] = geti(r, k);
    int64_t off = 0;
    /* zero limits: 1e-11 for trda, 1e-10 for trdb (:1163, :1195); applied by caller via flag */
    for (int k = 0; k < s->ndens; k++) {
      s->pos[k] = off;
      if (s->dim[k] == 0) continue;
      r = ff_rec(&f, &len);
      if (!r) { set_err("%s: truncated", path); ff_close(&f); return -2; }
      int kind = len / s->dim[k];
      for (int i = 0; i < s->dim[k]; i++) {
        double v;
        if (kind == 8) memcpy(&v, r + 8 * i, 8);
        else { float x; memcpy(&x, r + 4 * i, 4); v = x; }
        s->v64[off + i] = v;
      }
      r = ff_rec(&f, &len);
      memcpy(s->mask + (size_t)k * s->nsp * s->words, r, 4 * (size_t)s->nsp * s->words);
      off += s->dim[k];
    }
  }
  ff_close(&f);
  return 0;
}

static void zero_small(side_t *s, double lim) {
  int64_t tot = s->ndens ? s->pos[s->ndens - 1] + s->dim[s->ndens - 1] : 0;
  for (int64_t i = 0; i < tot; i++) {
    if (fabs(s->v64[i]) < lim) s->v64[i] = 0.0;
    s->v32[i] = (float)s->v64[i];
  }
}

/* .mat/.sgt of one side: src/NuTRL/NuCoreX.f90:1216-1319, names NuModules.F90:553-597 */
static int fdiv(int a, int b) { return a / b; } /* C and Fortran both truncate toward zero */
static int read_mats(side_t *s, const char *side_nuc, int minT2, int maxT2, int nmp, char xtra) {
  s->has_mat = (int *)calloc(s->nJ, sizeof(int));
  s->posm = (int64_t *)calloc(s->nJ, sizeof(int64_t));
  s->mmask = (uint32_t *)calloc((size_t)s->nJ * s->nsp * s->words, sizeof(uint32_t));
  size_t cap = 1024, used = 0;
  s->m64 = (double *)malloc(cap * sizeof(double));
  for (int J2 = s->minJ2; J2 <= s->maxJ2; J2 += 2) {
    for (int T2 = minT2; T2 <= maxT2; T2 += 2) {
      char name[1024], path[1100];
      int k3 = fdiv(nmp, 2) + MAX_CODE / 2;
      int c1 = J2 / 2, c2 = T2 / 2;
      if (c1 < 0) c1 += MAX_CODE;
      if (c2 < 0) c2 += MAX_CODE;
      if (k3 < 0) k3 += MAX_CODE;
      snprintf(name, sizeof name, "%s%c%c%c%c", side_nuc, AN_CODE[c1], AN_CODE[c2], AN_CODE[k3], xtra);
      ffile_t fs, fm;
      snprintf(path, sizeof path, "%s.sgt", name);
      if (ff_open(&fs, path)) continue;
      int len;
      unsigned char *r = ff_rec(&fs, &len);
      r = r ? ff_rec(&fs, &len) : NULL;
      if (!r) { ff_close(&fs); continue; }
      int nseg = geti(r, 0);
      float xx;
      memcpy(&xx, r + 8, 4);
      ff_close(&fs);
      if (xx != s->xtag) { set_err("%s: mat file and proj file not consistent", path); return -3; }
      snprintf(path, sizeof path, "%s.mat", name);
      if (ff_open(&fm, path)) return -1;
      ff_rec(&fm, &len);
      s->has_mat[J2 / 2] = 1;
      s->posm[J2 / 2] = (int64_t)used;
      for (int g = 0; g < nseg; g++) {
        r = ff_rec(&fm, &len);
        if (!r) { set_err("%s: truncated", path); ff_close(&fm); return -2; }
        int dimpt = geti(r, 0), dimptr = geti(r, 1), npc = geti(r, 2), npr = geti(r, 3);
        int n = dimpt * dimptr;
        int kind = (len - 16 - 8 - 64) / (n ? n : 1);
        if (kind != 8) kind = 4;
        uint32_t *col = s->mmask + ((size_t)(J2 / 2) * s->nsp + (npc - 1)) * s->words;
        col[(npr - 1) >> 5] |= 1u << ((npr - 1) & 31);
        while (used + n > cap) { cap *= 2; s->m64 = (double *)realloc(s->m64, cap * sizeof(double)); }
        for (int i = 0; i < n; i++) {
          if (kind == 8) memcpy(&s->m64[used + i], r + 16 + 8 * i, 8);
          else { float x; memcpy(&x, r + 16 + 4 * i, 4); s->m64[used + i] = x; }
        }
        used += n;
      }
      ff_close(&fm);
      break; /* one T per J */
    }
  }
  s->m32 = (float *)malloc((used ? used : 1) * sizeof(float));
  for (size_t i = 0; i < used; i++) s->m32[i] = (float)s->m64[i];
  return 0;
}

/* ---------------------------------------------------------------- .nus */
typedef struct {
  int no_cI, no_cN, tparity, no_shells, sj2[64], sp[64], sn[64], sl[64], np_shells;
  int min_cJ2, max_cJ2, min_cT2, max_cT2, nmp, min_MSon, max_MSon, del_cJ2, del_cT2;
} nus_t;

static int read_nus(nus_t *n, const char *prefix) {
  char path[1024], line[4096];
  snprintf(path, sizeof path, "%s.nus", prefix);
  FILE *fp = fopen(path, "r");
  if (!fp) { set_err("cannot open %s", path); return -1; }
  char lines[64][512];
  int nl = 0;
  while (nl < 64 && fgets(line, sizeof line, fp)) {
    if (line[0] == '!') continue;
    strncpy(lines[nl], line, 511);
    lines[nl][511] = 0;
    nl++;
  }
  fclose(fp);
  int i = 4;
  if (nl < 10) { set_err("%s: too short", path); return -2; }
  sscanf(lines[i++], "%d %d", &n->no_cI, &n->no_cN);
  sscanf(lines[i++], "%d", &n->tparity);
  sscanf(lines[i++], "%d", &n->no_shells);
  char nuc[512];
  strcpy(nuc, lines[i++]);
  trim(nuc);
  n->np_shells = 0;
  for (int k = 0; k < n->no_shells && nuc[k] == 'p'; k++) n->np_shells++;
  for (int k = 0; k < n->no_shells; k++) {
    int a, b, c, d, e, f;
    if (sscanf(lines[i++], "%d %d %d %d %d %d", &a, &b, &c, &d, &e, &f) != 6) { set_err("%s: shells", path); return -2; }
    n->sj2[k] = d;
    n->sp[k] = f;
    n->sn[k] = b;
    n->sl[k] = c;
  }
  sscanf(lines[i++], "%d %d", &n->min_cJ2, &n->max_cJ2);
  sscanf(lines[i++], "%d %d %d", &n->min_cT2, &n->max_cT2, &n->nmp);
  sscanf(lines[i++], "%d %d", &n->min_MSon, &n->max_MSon);
  i += 4; /* shell limits */
  i += 1; /* output_control */
  n->del_cJ2 = 2; n->del_cT2 = 2;
  if (i < nl) {
    int a, b;
    if (sscanf(lines[i], "%d %d", &a, &b) == 2) { n->del_cJ2 = a; n->del_cT2 = b; }
  }
  if (n->del_cJ2 < 2 || (n->del_cJ2 & 1)) n->del_cJ2 = 2;
  if (n->del_cT2 < 2 || (n->del_cT2 & 1)) n->del_cT2 = 2;
  return 0;
}


/* ---------------------------------------------------------------- separable operator (NuTRLZ) */
/* .dop reader, list-directed: each READ starts on a new line and skips the rest
   of its last line (read by prd(), src/NuTRLZ/NuCoreXZ.f90:1637-1660). */
typedef struct { char **tok; int *ntok; int nl, ln; } ldin_t;

static int ld_open(ldin_t *r, const char *path) {
  FILE *fp = fopen(path, "r");
  if (!fp) { set_err("cannot open %s", path); return -1; }
  int cap = 256;
  r->tok = (char **)malloc(sizeof(char *) * cap);
  r->ntok = (int *)malloc(sizeof(int) * cap);
  r->nl = r->ln = 0;
  char line[4096];
  while (fgets(line, sizeof line, fp)) {
    if (r->nl == cap) {
      cap *= 2;
      r->tok = (char **)realloc(r->tok, sizeof(char *) * cap);
      r->ntok = (int *)realloc(r->ntok, sizeof(int) * cap);
    }
    for (char *c = line; *c; c++) if (*c == ',' || *c == '\t' || *c == '\r' || *c == '\n') *c = ' ';
    r->tok[r->nl] = strdup(line);
    r->ntok[r->nl] = 0;
    r->nl++;
  }
  fclose(fp);
  return 0;
}
static void ld_close(ldin_t *r) {
  for (int i = 0; i < r->nl; i++) free(r->tok[i]);
  free(r->tok); free(r->ntok);
}
static int ld_read(ldin_t *r, int n, double *v) {
  int got = 0;
  while (got < n) {
    if (r->ln >= r->nl) return -1;
    char *p = r->tok[r->ln++];
    while (got < n) {
      while (*p == ' ') p++;
      if (!*p) break;
      char buf[128];
      int k = 0;
      while (*p && *p != ' ' && k < 127) { buf[k++] = (*p == 'd' || *p == 'D') ? 'e' : *p; p++; }
      buf[k] = 0;
      char *end;
      v[got++] = strtod(buf, &end);
      if (end == buf || *end) return -1;
    }
  }
  return 0;
}

typedef struct { int s[4]; double me; } dline_t;
typedef struct { double e; int nl; dline_t *l; } dterm_t;
typedef struct { int lm, par, nt; dterm_t *t; } dsec_t;

static int read_dop(const char *prefix, int *max_lmo, dsec_t **secs_out, int *nsec_out) {
  char path[1100];
  snprintf(path, sizeof path, "%s.dop", prefix);
  ldin_t r;
  if (ld_open(&r, path)) return -1;
  double v[5];
  int rc = -2;
  dsec_t *secs = NULL;
  int nsec = 0;
  if (ld_read(&r, 1, v)) goto bad;
  *max_lmo = (int)v[0];
  nsec = 2 * (*max_lmo / 2 + 1);
  secs = (dsec_t *)calloc(nsec, sizeof(dsec_t));
  for (int s = 0; s < nsec; s++) {
    if (ld_read(&r, 2, v)) goto bad;
    secs[s].lm = (int)v[0];
    secs[s].par = (int)v[1];
    if (ld_read(&r, 1, v)) goto bad;
    secs[s].nt = (int)v[0];
    if (secs[s].nt < 0 || secs[s].nt > 255) goto bad;
    secs[s].t = (dterm_t *)calloc(secs[s].nt + 1, sizeof(dterm_t));
    for (int n = 0; n < secs[s].nt; n++) {
      dterm_t *t = &secs[s].t[n];
      if (ld_read(&r, 1, v)) goto bad;
      t->e = (double)(float)v[0]; /* real(kind=4):: e (:1617) */
      if (ld_read(&r, 1, v)) goto bad;
      t->nl = (int)v[0];
      if (t->nl < 0) goto bad;
      t->l = (dline_t *)calloc(t->nl + 1, sizeof(dline_t));
      for (int c = 0; c < t->nl; c++) {
        if (ld_read(&r, 5, v)) goto bad;
        for (int q = 0; q < 4; q++) t->l[c].s[q] = (int)v[q];
        t->l[c].me = v[4];
      }
    }
  }
  ld_close(&r);
  *secs_out = secs;
  *nsec_out = nsec;
  return 0;
bad:
  set_err("%s: malformed near line %d", path, r.ln);
  ld_close(&r);
  if (secs) {
    for (int s = 0; s < nsec; s++) {
      if (secs[s].t) for (int n = 0; n < secs[s].nt; n++) free(secs[s].t[n].l);
      free(secs[s].t);
    }
    free(secs);
  }
  return rc;
}
static void free_dop(dsec_t *secs, int nsec) {
  for (int s = 0; s < nsec; s++) {
    for (int n = 0; n < secs[s].nt; n++) free(secs[s].t[n].l);
    free(secs[s].t);
  }
  free(secs);
}

/* one merged density in the union-mask layout */
typedef struct {
  uint8_t key[8];
  uint32_t *mask;
  int dim;
  double *v64;
  float *v32;
} mdens_t;

/* coefor (src/NuXpct/NuCoreXE.f90:70-138; NuTRLZ has the same routine): walk the
   union of the old and new masks in next_ij order (column partition outer, row
   partition inner) and add, copy new, or copy old, block by block. */
static void coefor(const side_t *s, int J2l, int J2r, mdens_t *g, const uint32_t *nmask, const double *n64,
                   const float *n32) {
  const size_t mw = (size_t)s->nsp * s->words;
  uint32_t *um = (uint32_t *)calloc(mw, sizeof(uint32_t));
  for (size_t w = 0; w < mw; w++) um[w] = g->mask[w] | nmask[w];
  int dim = 0;
  for (int j = 0; j < s->nsp; j++)
    for (int i = 0; i < s->nsp; i++)
      if (bit_of(um + (size_t)j * s->words, i)) dim += nJTs(s, J2r, j) * nJTs(s, J2l, i);
  double *t64 = (double *)calloc(dim ? dim : 1, sizeof(double));
  float *t32 = (float *)calloc(dim ? dim : 1, sizeof(float));
  int po = 0, pn = 0, pon = 0;
  for (int j = 0; j < s->nsp; j++)
    for (int i = 0; i < s->nsp; i++) {
      if (!bit_of(um + (size_t)j * s->words, i)) continue;
      int d = nJTs(s, J2r, j) * nJTs(s, J2l, i);
      int in_o = bit_of(g->mask + (size_t)j * s->words, i), in_n = bit_of(nmask + (size_t)j * s->words, i);
      for (int q = 0; q < d; q++) {
        if (in_o && in_n) { t64[pon + q] = g->v64[po + q] + n64[pn + q]; t32[pon + q] = g->v32[po + q] + n32[pn + q]; }
        else if (in_n) { t64[pon + q] = n64[pn + q]; t32[pon + q] = n32[pn + q]; }
        else { t64[pon + q] = g->v64[po + q]; t32[pon + q] = g->v32[po + q]; }
      }
      if (in_o) po += d;
      if (in_n) pn += d;
      pon += d;
    }
  free(g->mask); free(g->v64); free(g->v32);
  g->mask = um; g->v64 = t64; g->v32 = t32; g->dim = dim;
}

static int mkey_cmp(const void *x, const void *y) {
  return memcmp(((const mdens_t *)x)->key, ((const mdens_t *)y)->key, 8);
}

/* prd() for one side (a: :1637-1695, b: :1759-1810).  Keys of the merged
   densities: (2Jl, 2Jr, n, 200+parity slot, 2lam); a side scaled by
   e*renorm*me, b side by me.  The side's density arrays are replaced. */
static int prd_side(side_t *s, const dsec_t *secs, int nsec, int first_pair, int scale_e, double renorm) {
  int cap = 64, nm = 0;
  mdens_t *md = (mdens_t *)calloc(cap, sizeof(mdens_t));
  const size_t mw = (size_t)s->nsp * s->words;
  const float renorm32 = (float)renorm;
  for (int jl = s->minJ2; jl <= s->maxJ2; jl += s->delJ2)
    for (int jr = s->minJ2; jr <= s->maxJ2; jr += s->delJ2)
      for (int sc = 0; sc < nsec; sc++) {
        const int lm = secs[sc].lm;
        if (!(jr >= abs(jl - lm) && jr <= jl + lm)) continue;
        for (int n = 0; n < secs[sc].nt; n++) {
          const dterm_t *t = &secs[sc].t[n];
          for (int c = 0; c < t->nl; c++) {
            const dline_t *l = &t->l[c];
            int s1 = first_pair ? l->s[0] : l->s[2], s2 = first_pair ? l->s[1] : l->s[3];
            uint8_t k[8] = {(uint8_t)jl, (uint8_t)jr, (uint8_t)s1, (uint8_t)s2, (uint8_t)lm, 0, 0, 0};
            int ds = kfind(s->keys, s->ndens, k);
            if (!ds) continue;
            ds--;
            const int dim = s->dim[ds];
            double *n64 = (double *)malloc(sizeof(double) * (dim ? dim : 1));
            float *n32 = (float *)malloc(sizeof(float) * (dim ? dim : 1));
            const double f64 = scale_e ? (t->e * renorm) * l->me : l->me;
            const float f32 = scale_e ? ((float)t->e * renorm32) * (float)l->me : (float)l->me;
            for (int q = 0; q < dim; q++) {
              n64[q] = f64 * s->v64[s->pos[ds] + q];
              n32[q] = f32 * s->v32[s->pos[ds] + q];
            }
            uint8_t mk[8] = {(uint8_t)jl, (uint8_t)jr, (uint8_t)(n + 1), (uint8_t)(200 + (secs[sc].par + 3) / 2),
                             (uint8_t)lm, 0, 0, 0};
            int g = -1;
            for (int q = nm - 1; q >= 0; q--)
              if (!memcmp(md[q].key, mk, 8)) { g = q; break; }
            if (g < 0) {
              if (nm == cap) { cap *= 2; md = (mdens_t *)realloc(md, sizeof(mdens_t) * cap); }
              g = nm++;
              memcpy(md[g].key, mk, 8);
              md[g].mask = (uint32_t *)calloc(mw, sizeof(uint32_t));
              memcpy(md[g].mask, s->mask + (size_t)ds * mw, sizeof(uint32_t) * mw);
              md[g].dim = dim;
              md[g].v64 = n64;
              md[g].v32 = n32;
            } else {
              coefor(s, jl, jr, &md[g], s->mask + (size_t)ds * mw, n64, n32);
              free(n64);
              free(n32);
            }
          }
        }
      }
  qsort(md, nm, sizeof(mdens_t), mkey_cmp);
  int64_t total = 0;
  for (int q = 0; q < nm; q++) total += md[q].dim;
  free(s->keys); free(s->pos); free(s->dim); free(s->mask); free(s->v64); free(s->v32);
  s->ndens = nm;
  s->keys = (uint8_t *)malloc(8 * (size_t)(nm ? nm : 1));
  s->pos = (int64_t *)calloc(nm + 1, sizeof(int64_t));
  s->dim = (int *)calloc(nm + 1, sizeof(int));
  s->mask = (uint32_t *)calloc((size_t)(nm ? nm : 1) * mw, sizeof(uint32_t));
  s->v64 = (double *)calloc(total ? total : 1, sizeof(double));
  s->v32 = (float *)calloc(total ? total : 1, sizeof(float));
  int64_t off = 0;
  for (int q = 0; q < nm; q++) {
    memcpy(s->keys + 8 * q, md[q].key, 8);
    s->pos[q] = off;
    s->dim[q] = md[q].dim;
    memcpy(s->mask + (size_t)q * mw, md[q].mask, sizeof(uint32_t) * mw);
    memcpy(s->v64 + off, md[q].v64, sizeof(double) * md[q].dim);
    memcpy(s->v32 + off, md[q].v32, sizeof(float) * md[q].dim);
    off += md[q].dim;
    free(md[q].mask); free(md[q].v64); free(md[q].v32);
  }
  free(md);
  return 0;
}

/* drdbh() of NuTRLZ (src/NuTRLZ/NuCoreXZ.f90:1536-1595): per (L, Ja') the merged
   a-densities over (2lam index, parity, n), each with (merged b-density, R, racah). */
static int drdbh_sep(refcpu_t *h, const dsec_t *secs, int nsec, int max_lmo, int max_sj2u, int limits) {
  const int totlms = max_lmo / 2 + 1;
  int max_elm = 0;
  for (int s = 0; s < nsec; s++) if (secs[s].nt > max_elm) max_elm = secs[s].nt;
  int *lmel = (int *)malloc(sizeof(int) * (size_t)totlms * (max_elm + 1) * 2);
  for (size_t q = 0; q < (size_t)totlms * (max_elm + 1) * 2; q++) lmel[q] = -1;
#define LMEL(t, n, p) lmel[((size_t)(t) * (max_elm + 1) + (n)) * 2 + (p)]
  for (int s = 0; s < nsec; s++)
    for (int n = 1; n <= secs[s].nt; n++) LMEL(secs[s].lm / 2, n, (secs[s].par + 3) / 2 - 1) = secs[s].lm;
  h->nslot = h->A.maxJ2 / 2 + 1;
  h->tab = (tab_t *)calloc((size_t)h->basd * h->nslot, sizeof(tab_t));
  size_t bcap = 1024;
  h->tb_b = (int *)malloc(sizeof(int) * bcap);
  h->tb_h = (int *)malloc(sizeof(int) * bcap);
  h->tb_me64 = (double *)malloc(sizeof(double) * bcap);
  int64_t nt = 0;
  for (int bsc = 0; bsc < h->basd; bsc++) {
    int ajl = h->bas_a[bsc], bjl = h->bas_b[bsc];
    for (int ajr = h->A.minJ2; ajr <= h->A.maxJ2; ajr += h->A.delJ2) {
      tab_t *t = &h->tab[bsc * h->nslot + ajr / 2];
      int acap = 16;
      t->a = (int *)malloc(sizeof(int) * acap);
      t->boff = (int *)malloc(sizeof(int) * (acap + 1));
      t->na = 0;
      for (int tl = 0; tl < totlms; tl++)
        for (int pp = 0; pp < 2; pp++)
          for (int n = 1; n <= max_elm; n++) {
            int lm = LMEL(tl, n, pp);
            if (lm == -1) continue;
            if (!(ajr >= abs(ajl - lm) && ajr <= ajl + lm)) continue;
            uint8_t ta[8] = {(uint8_t)ajl, (uint8_t)ajr, (uint8_t)n, (uint8_t)(201 + pp), (uint8_t)lm, 0, 0, 0};
            int atsc = kfind(h->A.keys, h->A.ndens, ta);
            if (!atsc) continue;
            if (t->na + 1 >= acap) {
              acap *= 2;
              t->a = (int *)realloc(t->a, sizeof(int) * acap);
              t->boff = (int *)realloc(t->boff, sizeof(int) * (acap + 1));
            }
            t->a[t->na] = atsc - 1;
            t->boff[t->na] = (int)nt;
            int bjrn = abs(bjl - lm) > abs(ajr - h->J2) ? abs(bjl - lm) : abs(ajr - h->J2);
            int bjrx = bjl + lm;
            if (ajr + h->J2 < bjrx) bjrx = ajr + h->J2;
            if (h->B.maxJ2 < bjrx) bjrx = h->B.maxJ2;
            bjrn = bjrn + (bjrn - h->B.minJ2) % h->B.delJ2;
            for (int bjr = bjrn; bjr <= bjrx; bjr += h->B.delJ2) {
              /* racah table entries: Jb' on the b grid, 2lam <= 2 max(2j), |W| above the limit (:1115-1133) */
              if (bjr < h->B.minJ2 || lm > 2 * max_sj2u) continue;
              double rcw = racahW(ajl, bjl, ajr, bjr, h->J2, lm);
              if (limits ? !(fabs(rcw) > 1e-6) : fabs(rcw) == 0.0) continue;
              double ph = ((((ajl + bjr - h->J2 - lm) / 2)) & 1) ? -1.0 : 1.0;
              double rcv = ph * rcw * sqrt((double)(lm + 1));
              if (!h->fp64) rcv = (double)(float)rcv;
              int R = -1;
              for (int q = 0; q < h->basd; q++)
                if (h->bas_a[q] == ajr && h->bas_b[q] == bjr) { R = q; break; }
              if (R < 0) { set_err("find error bh tables"); free(lmel); return -5; }
              if (!(bjr >= abs(bjl - lm) && bjr <= bjl + lm)) continue;
              uint8_t tb[8] = {(uint8_t)bjl, (uint8_t)bjr, (uint8_t)n, (uint8_t)(201 + pp), (uint8_t)lm, 0, 0, 0};
              int btsc = kfind(h->B.keys, h->B.ndens, tb);
              if (!btsc) continue;
              if ((size_t)nt + 1 >= bcap) {
                bcap *= 2;
                h->tb_b = (int *)realloc(h->tb_b, sizeof(int) * bcap);
                h->tb_h = (int *)realloc(h->tb_h, sizeof(int) * bcap);
                h->tb_me64 = (double *)realloc(h->tb_me64, sizeof(double) * bcap);
              }
              h->tb_b[nt] = btsc - 1;
              h->tb_h[nt] = R;
              h->tb_me64[nt] = rcv;
              nt++;
            }
            t->na++;
            t->boff[t->na] = (int)nt;
          }
      if (t->na == 0) t->boff[0] = (int)nt;
    }
  }
#undef LMEL
  free(lmel);
  h->ntb = nt;
  h->tb_me32 = (float *)malloc(sizeof(float) * (nt ? nt : 1));
  for (int64_t i = 0; i < nt; i++) h->tb_me32[i] = (float)h->tb_me64[i];
  return 0;
}

/* ---------------------------------------------------------------- create */
static int cmp2(const void *x, const void *y) {
  const int *a = (const int *)x, *b = (const int *)y;
  if (a[0] != b[0]) return a[0] - b[0];
  return a[1] - b[1];
}

static int create_impl(const char *nus_prefix, const char *opfile_a, const char *opfile_b,
                       const char *intfile, const char *dopfile, int fp64, int thresholds, int limits,
                       refcpu_t **out) {
  refcpu_t *h = (refcpu_t *)calloc(1, sizeof(refcpu_t));
  nus_t nus;
  int rc;
  *out = NULL;
  h->fp64 = fp64;
  h->thr_on = thresholds;
  h->nthr = omp_get_max_threads();
  if ((rc = read_nus(&nus, nus_prefix))) goto fail;
  h->J2 = nus.min_cJ2;
  h->tparity = nus.tparity;
  h->sj2 = (int *)malloc(sizeof(int) * 64);
  memcpy(h->sj2, nus.sj2, sizeof(int) * 64);
  int cNN[64] = {0}; /* 2n + l per shell (:814-817) */
  for (int k = 0; k < nus.no_shells; k++) cNN[k] = 2 * nus.sn[k] + nus.sl[k];
  if ((rc = read_side(&h->A, opfile_a, nus.sp, cNN, nus.no_shells))) goto fail;
  if ((rc = read_side(&h->B, opfile_b, nus.sp, cNN, nus.no_shells))) goto fail;
  /* zero limits: NuTRL 1e-11 / 1e-10 (:1163, :1195); NuTRLZ 1e-6 on both sides
     (src/NuTRLZ/NuCoreXZ.f90:1205,1237), applied when `limits` is set */
  zero_small(&h->A, dopfile && limits ? 1e-6 : 1e-11);
  zero_small(&h->B, dopfile && limits ? 1e-6 : 1e-10);
  {
    char typea = h->A.proton ? 'p' : 'n', typeb = h->B.proton ? 'p' : 'n';
    if (typea == typeb) { typea = 'n'; typeb = 'p'; }
    int nps = nus.np_shells;
    if (typea == 'p') { h->sh_a0 = 1; h->sh_a1 = nps; h->sh_b0 = nps + 1; h->sh_b1 = nus.no_shells; }
    else { h->sh_a0 = nps + 1; h->sh_a1 = nus.no_shells; h->sh_b0 = 1; h->sh_b1 = nps; }
    int no_cA, no_cB;
    if (typea == 'p') { no_cB = (nus.no_cN + nus.nmp) / 2; no_cA = (nus.no_cN - nus.nmp) / 2; }
    else { no_cB = (nus.no_cN - nus.nmp) / 2; no_cA = (nus.no_cN + nus.nmp) / 2; }
    h->azero = no_cA == 0;
    h->bzero = no_cB == 0;
    /* blocks: src/NuTRL/NuCoreX.f90:932-988 */
    int cap = 0;
    for (int ajl = h->A.minJ2; ajl <= h->A.maxJ2; ajl += h->A.delJ2)
      for (int bjl = h->B.minJ2; bjl <= h->B.maxJ2; bjl += 1) cap++;
    int *pairs = (int *)malloc(sizeof(int) * 2 * (cap + 1));
    int nb = 0;
    for (int ajl = h->A.minJ2; ajl <= h->A.maxJ2; ajl += h->A.delJ2) {
      int bjln = abs(ajl - h->J2) > h->B.minJ2 ? abs(ajl - h->J2) : h->B.minJ2;
      int bjlx = ajl + h->J2 < h->B.maxJ2 ? ajl + h->J2 : h->B.maxJ2;
      bjln = bjln + (bjln - h->B.minJ2) % h->B.delJ2;
      for (int bjl = bjln; bjl <= bjlx; bjl += h->B.delJ2) { pairs[2 * nb] = ajl; pairs[2 * nb + 1] = bjl; nb++; }
    }
    qsort(pairs, nb, 2 * sizeof(int), cmp2); /* ascending (char(2Ja)//char(2Jb)) strings */
    h->basd = nb;
    h->bas_a = (int *)malloc(sizeof(int) * (nb + 1));
    h->bas_b = (int *)malloc(sizeof(int) * (nb + 1));
    h->posnv = (int64_t *)malloc(sizeof(int64_t) * (nb + 1));
    h->dimnv = (int64_t *)malloc(sizeof(int64_t) * (nb + 1));
    int64_t N = 0;
    h->maxda = h->maxdb = 0;
    for (int b = 0; b < nb; b++) {
      h->bas_a[b] = pairs[2 * b];
      h->bas_b[b] = pairs[2 * b + 1];
      int da = h->A.d0[h->bas_a[b] / 2], db = h->B.d0[h->bas_b[b] / 2];
      h->posnv[b] = N;
      h->dimnv[b] = (int64_t)da * db;
      N += h->dimnv[b];
    }
    for (int J = h->A.minJ2; J <= h->A.maxJ2; J += 2) if (h->A.d0[J / 2] > h->maxda) h->maxda = h->A.d0[J / 2];
    for (int J = h->B.minJ2; J <= h->B.maxJ2; J += 2) if (h->B.d0[J / 2] > h->maxdb) h->maxdb = h->B.d0[J / 2];
    free(pairs);
    h->N = N;
    /* lcoef: state order b-part, b-state, a-part, a-state (:998-1030); with a
       window also trunc and bas_log (:990-1037) */
    int truncate = nus.min_MSon != 0 || nus.max_MSon != 999; /* :813 */
    h->lcoef = (uint8_t *)malloc(N ? N : 1);
    if (truncate) {
      h->trunc = (uint8_t *)calloc(N ? N : 1, 1);
      h->bas_log = (uint8_t *)calloc(nb ? nb : 1, 1);
    }
    int64_t nsc = 0;
    for (int b = 0; b < nb; b++)
      for (int bp = 0; bp < h->B.nsp; bp++)
        for (int bs = 0; bs < nJTs(&h->B, h->bas_b[b], bp); bs++)
          for (int ap = 0; ap < h->A.nsp; ap++) {
            int ntot = h->A.nq[ap] + h->B.nq[bp];
            int in = ntot >= nus.min_MSon && ntot <= nus.max_MSon;
            if (truncate && in) h->bas_log[b] = 1;
            for (int as = 0; as < nJTs(&h->A, h->bas_a[b], ap); as++) {
              int ok = h->tparity == 0 || h->A.par[ap] * h->B.par[bp] == h->tparity;
              if (truncate) h->trunc[nsc] = (uint8_t)(in && ok);
              h->lcoef[nsc++] = (uint8_t)ok;
            }
          }
    /* like-particle matrices */
    char xtra = intfile[strlen(intfile) - 1];
    int nmp_a = typea == 'n' ? no_cA : -no_cA;
    int nmp_b = typeb == 'n' ? no_cB : -no_cB;
    if ((rc = read_mats(&h->B, opfile_b, nus.min_cT2, nus.max_cT2, nmp_b, xtra))) goto fail;
    if ((rc = read_mats(&h->A, opfile_a, nus.min_cT2, nus.max_cT2, nmp_a, xtra))) goto fail;
    if (h->azero || h->bzero) { *out = h; return 0; }

    /* racahs (:1071-1127): sorted keys (ajl,bjl,ajr,bjr,lm) with |W|>0 */
    int max_sj2u = 0;
    for (int k = 0; k < nus.no_shells; k++) if (nus.sj2[k] > max_sj2u) max_sj2u = nus.sj2[k];
    /* spe: renorm (line 2), levels (line 3) */
    char path[1024], line[4096];
    snprintf(path, sizeof path, "%s.spe", intfile);
    FILE *fp = fopen(path, "r");
    if (!fp) { set_err("cannot open %s", path); rc = -1; goto fail; }
    double renorm = 1.0;
    if (!fgets(line, sizeof line, fp) || !fgets(line, sizeof line, fp)) { fclose(fp); set_err("%s: short", path); rc = -2; goto fail; }
    renorm = atof(line);
    fclose(fp);
    if (!fp64) renorm = (double)(float)renorm;
    /* .oph (:1133-1142) */
    ffile_t fo;
    snprintf(path, sizeof path, "%s.oph", intfile);
    if ((rc = ff_open(&fo, path))) goto fail;
    int len;
    unsigned char *r = ff_rec(&fo, &len);
    int opd = geti(r, 0), max_lmo = geti(r, 1);
    uint8_t *opn = (uint8_t *)malloc(8 * (size_t)(opd ? opd : 1));
    double *opmeV = (double *)malloc(sizeof(double) * (opd ? opd : 1));
    for (int k = 0; k < opd; k++) {
      r = ff_rec(&fo, &len);
      memcpy(opn + 8 * k, r, 8);
      double v;
      memcpy(&v, r + 8, 8);
      if (fabs(v) < 1e-10) v = 0.0;
      opmeV[k] = renorm * v;
    }
    ff_close(&fo);
    if (dopfile) {
      /* NuTRLZ: the .oph is read but the pn part comes from the .dop (prd, drdbh) */
      free(opn);
      free(opmeV);
      int dmax = 0, nsec = 0;
      dsec_t *secs = NULL;
      if ((rc = read_dop(dopfile, &dmax, &secs, &nsec))) goto fail;
      int a_first = h->sh_a0 == 1; /* typea == 'p': the proton pair is first on a line */
      prd_side(&h->A, secs, nsec, a_first, 1, renorm);
      prd_side(&h->B, secs, nsec, !a_first, 0, renorm);
      rc = drdbh_sep(h, secs, nsec, dmax, max_sj2u, limits);
      free_dop(secs, nsec);
      if (rc) goto fail;
      *out = h;
      return 0;
    }
    /* max_lm (:1320-1326) */
    int max_lambda = nus.del_cJ2;
    if (max_lambda <= 2 || (max_lambda & 1)) max_lambda = 2 * max_sj2u;
    int del_lm = nus.del_cT2;
    if (del_lm < 2 || (del_lm & 1)) del_lm = 2;
    int max_lm = h->A.max_lm;
    if (max_lmo < max_lm) max_lm = max_lmo;
    if (2 * max_sj2u < max_lm) max_lm = 2 * max_sj2u;
    if (max_lambda < max_lm) max_lm = max_lambda;

    /* drdbh (:1428-1577): per (L, ajr) the a-densities, each with (b-density, R, opmeV*racah) */
    h->nslot = h->A.maxJ2 / 2 + 1;
    h->tab = (tab_t *)calloc((size_t)h->basd * h->nslot, sizeof(tab_t));
    size_t bcap = 1024;
    h->tb_b = (int *)malloc(sizeof(int) * bcap);
    h->tb_h = (int *)malloc(sizeof(int) * bcap);
    h->tb_me64 = (double *)malloc(sizeof(double) * bcap);
    int64_t nt = 0;
    uint8_t key[8];
    for (int bsc = 0; bsc < h->basd; bsc++) {
      int ajl = h->bas_a[bsc], bjl = h->bas_b[bsc];
      for (int ajr = h->A.minJ2; ajr <= h->A.maxJ2; ajr += h->A.delJ2) {
        tab_t *t = &h->tab[bsc * h->nslot + ajr / 2];
        int acap = 16;
        t->a = (int *)malloc(sizeof(int) * acap);
        t->boff = (int *)malloc(sizeof(int) * (acap + 1));
        t->na = 0;
        for (int ashl = h->sh_a0; ashl <= h->sh_a1; ashl++)
          for (int ashr = h->sh_a0; ashr <= h->sh_a1; ashr++) {
            int lmn = abs(nus.sj2[ashl - 1] - nus.sj2[ashr - 1]);
            int lmx = nus.sj2[ashl - 1] + nus.sj2[ashr - 1];
            if (max_lm < lmx) lmx = max_lm;
            lmn = lmn + lmn % del_lm;
            for (int lm = lmn; lm <= lmx; lm += del_lm) {
              if (!(ajr >= abs(ajl - lm) && ajr <= ajl + lm)) continue;
              uint8_t ta[8] = {(uint8_t)ajl, (uint8_t)ajr, (uint8_t)ashl, (uint8_t)ashr, (uint8_t)lm, 0, 0, 0};
              int atsc = kfind(h->A.keys, h->A.ndens, ta);
              if (!atsc) continue;
              if (t->na + 1 >= acap) {
                acap *= 2;
                t->a = (int *)realloc(t->a, sizeof(int) * acap);
                t->boff = (int *)realloc(t->boff, sizeof(int) * (acap + 1));
              }
              t->a[t->na] = atsc - 1;
              t->boff[t->na] = (int)nt;
              int bjrn = abs(bjl - lm) > abs(ajr - h->J2) ? abs(bjl - lm) : abs(ajr - h->J2);
              int bjrx = bjl + lm;
              if (ajr + h->J2 < bjrx) bjrx = ajr + h->J2;
              if (h->B.maxJ2 < bjrx) bjrx = h->B.maxJ2;
              bjrn = bjrn + (bjrn - h->B.minJ2) % h->B.delJ2;
              for (int bjr = bjrn; bjr <= bjrx; bjr += h->B.delJ2) {
                double rcw = racahW(ajl, bjl, ajr, bjr, h->J2, lm);
                if (fabs(rcw) == 0.0) continue;
                double ph = ((((ajl + bjr - h->J2 - lm) / 2)) & 1) ? -1.0 : 1.0;
                double rcv = ph * rcw * sqrt((double)(lm + 1));
                if (!fp64) rcv = (double)(float)rcv;
                int R = -1;
                for (int q = 0; q < h->basd; q++)
                  if (h->bas_a[q] == ajr && h->bas_b[q] == bjr) { R = q; break; }
                if (R < 0) { set_err("find error bh tables"); rc = -5; free(opn); free(opmeV); goto fail; }
                for (int bshl = h->sh_b0; bshl <= h->sh_b1; bshl++)
                  for (int bshr = h->sh_b0; bshr <= h->sh_b1; bshr++) {
                    int k1 = nus.sj2[bshl - 1], k2 = nus.sj2[bshr - 1];
                    int blmn = abs(k1 - k2), blmx = k1 + k2;
                    if (max_lm < blmx) blmx = max_lm;
                    if (lm < blmn || lm > blmx) continue;
                    if (!(bjr >= abs(bjl - lm) && bjr <= bjl + lm)) continue;
                    uint8_t tb[8] = {(uint8_t)bjl, (uint8_t)bjr, (uint8_t)bshl, (uint8_t)bshr, (uint8_t)lm, 0, 0, 0};
                    int btsc = kfind(h->B.keys, h->B.ndens, tb);
                    if (!btsc) continue;
                    uint8_t op[8] = {(uint8_t)ashl, (uint8_t)ashr, (uint8_t)bshl, (uint8_t)bshr, (uint8_t)lm, 0, 0, 0};
                    int psc = kfind(opn, opd, op);
                    if (!psc) continue;
                    double me = opmeV[psc - 1];
                    if (me == 0.0) continue;
                    if ((size_t)nt + 1 >= bcap) {
                      bcap *= 2;
                      h->tb_b = (int *)realloc(h->tb_b, sizeof(int) * bcap);
                      h->tb_h = (int *)realloc(h->tb_h, sizeof(int) * bcap);
                      h->tb_me64 = (double *)realloc(h->tb_me64, sizeof(double) * bcap);
                    }
                    h->tb_b[nt] = btsc - 1;
                    h->tb_h[nt] = R;
                    h->tb_me64[nt] = fp64 ? me * rcv : (double)(float)(me * rcv);
                    nt++;
                  }
              }
              t->na++;
              t->boff[t->na] = (int)nt;
            }
          }
        if (t->na == 0) t->boff[0] = (int)nt;
      }
    }
    (void)key;
    h->ntb = nt;
    h->tb_me32 = (float *)malloc(sizeof(float) * (nt ? nt : 1));
    for (int64_t i = 0; i < nt; i++) h->tb_me32[i] = (float)h->tb_me64[i];
    free(opn);
    free(opmeV);
  }
  *out = h;
  return 0;
fail:
  refcpu_destroy(h);
  return rc ? rc : -1;
}

int refcpu_create(const char *nus_prefix, const char *opfile_a, const char *opfile_b,
                  const char *intfile, int fp64, int thresholds, refcpu_t **out) {
  return create_impl(nus_prefix, opfile_a, opfile_b, intfile, NULL, fp64, thresholds, 1, out);
}

int refcpu_create_sep(const char *nus_prefix, const char *opfile_a, const char *opfile_b,
                      const char *intfile, const char *dopfile, int fp64, int thresholds, int limits,
                      refcpu_t **out) {
  return create_impl(nus_prefix, opfile_a, opfile_b, intfile, dopfile, fp64, thresholds, limits, out);
}

int64_t refcpu_dim(const refcpu_t *h) { return h->N; }
int refcpu_basd(const refcpu_t *h) { return h->basd; }
int refcpu_block(const refcpu_t *h, int b, int *ja2, int *jb2, int64_t *pos, int64_t *dim) {
  if (b < 0 || b >= h->basd) return -1;
  *ja2 = h->bas_a[b]; *jb2 = h->bas_b[b]; *pos = h->posnv[b]; *dim = h->dimnv[b];
  return 0;
}
int refcpu_lcoef(const refcpu_t *h, uint8_t *mask) { memcpy(mask, h->lcoef, h->N); return 0; }
int refcpu_trunc(const refcpu_t *h, uint8_t *mask) {
  if (!h->trunc) { memcpy(mask, h->lcoef, h->N); return 0; }
  memcpy(mask, h->trunc, h->N);
  return 1;
}
int refcpu_set_threads(refcpu_t *h, int n) { h->nthr = n > 0 ? n : 1; return 0; }

/* ---------------------------------------------------------------- mask helpers */
/* crproj (src/NuCommonX/NuCommonX.f90:639-671): row and column projections */
static void crproj(const uint32_t *la, int nsr, int nsc, int words, uint32_t *lac, uint32_t *lar) {
  memset(lac, 0, 4 * (size_t)words);
  memset(lar, 0, 4 * (size_t)words);
  for (int c = 0; c < nsc; c++) {
    const uint32_t *col = la + (size_t)c * words;
    int nz = 0;
    for (int w = 0; w < words; w++) { lar[w] |= col[w]; if (col[w]) nz = 1; }
    if (nz) lac[c >> 5] |= 1u << (c & 31);
  }
  (void)nsr;
}
/* laor (:693-707) */
static void laor(uint32_t *dst, const uint32_t *src, int nsc, int words) {
  for (size_t i = 0; i < (size_t)nsc * words; i++) dst[i] |= src[i];
}

/* kernels, one instance per precision */
#define R float
#define SFX(x) x##_f32
#define VALS(s) ((s)->v32)
#define MATS(s) ((s)->m32)
#define TBME(h) ((h)->tb_me32)
#include "refcpu_kernels.h"
#undef R
#undef SFX
#undef VALS
#undef MATS
#undef TBME
#define R double
#define SFX(x) x##_f64
#define VALS(s) ((s)->v64)
#define MATS(s) ((s)->m64)
#define TBME(h) ((h)->tb_me64)
#include "refcpu_kernels.h"

int refcpu_apply(refcpu_t *h, const double *v_in, double *v_out) {
  if (h->fp64) return hv_f64(h, v_in, v_out);
  return hv_f32(h, v_in, v_out);
}

/* algorithmic work: the reference loop nest's sparse flops
   F_x = sum_(L,Ja',k) [ sum_(R,k') 2*dar*nnz(k') + 2*dbl*nnz(k) ],
   F_a = sum_L 2*db*(2*nnz_off(H_a) + nnz_diag(H_a)), F_b likewise. */
int refcpu_work(const refcpu_t *h, double *flops, int64_t *counts) {
  double f = 0.0;
  int64_t nnzA = 0, nnzB = 0, nnzHa = 0, nnzHb = 0;
  for (int k = 0; k < h->A.ndens; k++) nnzA += h->A.dim[k];
  for (int k = 0; k < h->B.ndens; k++) nnzB += h->B.dim[k];
  uint32_t *lac = (uint32_t *)calloc(h->A.words + h->B.words + 8, 4);
  uint32_t *lar = (uint32_t *)calloc(h->A.words + h->B.words + 8, 4);
  uint32_t *tmp = (uint32_t *)calloc((size_t)h->B.nsp * h->B.words + 8, 4);
  uint32_t *lbc = (uint32_t *)calloc(h->B.words + 8, 4), *lbr = (uint32_t *)calloc(h->B.words + 8, 4);
  if (!(h->azero || h->bzero)) {
    for (int bsc = 0; bsc < h->basd; bsc++)
      for (int s = 0; s < h->nslot; s++) {
        tab_t *t = &h->tab[bsc * h->nslot + s];
        int ajr = 2 * s;
        for (int n = 0; n < t->na; n++) {
          int atsc = t->a[n];
          if (t->boff[n + 1] == t->boff[n]) continue;
          crproj(h->A.mask + (size_t)atsc * h->A.nsp * h->A.words, h->A.nsp, h->A.nsp, h->A.words, lac, lar);
          memset(tmp, 0, 4 * (size_t)h->B.nsp * h->B.words);
          for (int q = t->boff[n]; q < t->boff[n + 1]; q++)
            laor(tmp, h->B.mask + (size_t)h->tb_b[q] * h->B.nsp * h->B.words, h->B.nsp, h->B.words);
          crproj(tmp, h->B.nsp, h->B.nsp, h->B.words, lbc, lbr);
          int dar = 0, dbl = 0;
          for (int i = 0; i < h->A.nsp; i++) if (bit_of(lac, i)) dar += nJTs(&h->A, ajr, i);
          for (int i = 0; i < h->B.nsp; i++) if (bit_of(lbr, i)) dbl += nJTs(&h->B, h->bas_b[bsc], i);
          for (int q = t->boff[n]; q < t->boff[n + 1]; q++) f += 2.0 * dar * h->B.dim[h->tb_b[q]];
          f += 2.0 * dbl * h->A.dim[atsc];
        }
      }
  }
  /* like-particle parts: for every block, every stored segment element touches a column */
  for (int pass = 0; pass < 2; pass++) {
    const side_t *s = pass ? &h->B : &h->A;
    for (int bsc = 0; bsc < h->basd; bsc++) {
      int J = pass ? h->bas_b[bsc] : h->bas_a[bsc];
      int other = pass ? h->A.d0[h->bas_a[bsc] / 2] : h->B.d0[h->bas_b[bsc] / 2];
      if (!s->has_mat[J / 2]) continue;
      const uint32_t *m = s->mmask + (size_t)(J / 2) * s->nsp * s->words;
      for (int j = 0; j < s->nsp; j++)
        for (int i = 0; i < s->nsp; i++)
          if (bit_of(m + (size_t)j * s->words, i) && s->par[i] == s->par[j]) {
            double n = (double)nJTs(s, J, i) * nJTs(s, J, j);
            f += 2.0 * other * n * (i == j ? 1.0 : 2.0);
          }
    }
    for (int J = s->minJ2; J <= s->maxJ2; J += 2) {
      if (!s->has_mat[J / 2]) continue;
      const uint32_t *m = s->mmask + (size_t)(J / 2) * s->nsp * s->words;
      int64_t cnt = 0;
      for (int j = 0; j < s->nsp; j++)
        for (int i = 0; i < s->nsp; i++)
          if (bit_of(m + (size_t)j * s->words, i) && s->par[i] == s->par[j]) cnt += (int64_t)nJTs(s, J, i) * nJTs(s, J, j);
      if (pass) nnzHb += cnt; else nnzHa += cnt;
    }
  }
  free(lac); free(lar); free(tmp); free(lbc); free(lbr);
  *flops = f;
  if (counts) { counts[0] = nnzA; counts[1] = nnzB; counts[2] = nnzHa; counts[3] = nnzHb; counts[4] = h->ntb; }
  return 0;
}
