/* ORACLE kernels, included twice by refcpu.c with R = float / double.
 * Restates src/NuCommonX/NuCommonX.f90 single_productm (:745-827),
 * single_product1 (:830-919), mulmtx (:1009-1104) and the drivers
 * multiplyx/multiplya/multiplyb of src/NuTRL/NuCoreX.f90:76-226. */

/* single_productm: gather V_R (masked a rows laco, masked b columns lbco) into
   B1 (dar x cols, col-major), then for each b-density of the group and each set
   bit (i=row b-partition of Jb, j=column b-partition of Jb'):
   B3[:, (i,l)] += (trdb*meV) * B1[:, (j,k)].
   On a truncated basis (:760-767, :807-812) the gather keeps the lcoef states
   only, F1 flags the gathered states inside the window (lbuffer1), an element
   adds only where its source is flagged, and F2 (lbuffer2) flags the elements
   of B3 that received a flagged source. */
static void SFX(single_productm)(const refcpu_t *h, const R *bveci, int64_t posi, const R *meV,
                                 int ajr, int bjl, int bjr, const uint32_t *laco,
                                 const uint32_t *lbco, const int *blocb, int dar, int dbl,
                                 int zero, R *B1, R *B3, int ntrb, const int *locb, uint8_t *F1,
                                 uint8_t *F2) {
  const side_t *A = &h->A, *B = &h->B;
  int nsa = A->nsp, nsb = B->nsp;
  const int trunc = h->trunc != NULL;
  int64_t p = posi;
  int64_t pp = 0;
  for (int j = 0; j < nsb; j++)
    for (int k = 0; k < nJTs(B, bjr, j); k++)
      for (int i = 0; i < nsa; i++) {
        int nr = nJTs(A, ajr, i);
        if (bit_of(lbco, j) && bit_of(laco, i) && nr != 0) {
          if (trunc) {
            for (int q = 0; q < nr; q++) {
              B1[pp + q] = h->lcoef[p + q] ? bveci[p + q] : (R)0;
              F1[pp + q] = (uint8_t)(h->lcoef[p + q] && h->trunc[p + q]);
            }
          } else {
            for (int q = 0; q < nr; q++) B1[pp + q] = bveci[p + q];
          }
          pp += nr;
        }
        p += nr;
      }
  int64_t dp = (int64_t)dbl * dar;
  if (zero)
    for (int64_t q = 0; q < dp; q++) B3[q] = (R)0;
  if (zero && trunc) memset(F2, 0, (size_t)dp);
  int *snJTb = (int *)alloca(sizeof(int) * (nsb + 1));
  snJTb[0] = 0;
  for (int j = 1; j < nsb; j++) snJTb[j] = snJTb[j - 1] + (bit_of(lbco, j - 1) ? dar * nJTs(B, bjr, j - 1) : 0);
  for (int ntb = 0; ntb < ntrb; ntb++) {
    int btsc = locb[ntb];
    int64_t pt = B->pos[btsc];
    const uint32_t *m = B->mask + (size_t)btsc * nsb * B->words;
    const R *trdb = VALS(B);
    for (int j = 0; j < nsb; j++)
      for (int i = 0; i < nsb; i++) {
        if (!bit_of(m + (size_t)j * B->words, i)) continue;
        int pu = snJTb[j];
        for (int k = 0; k < nJTs(B, bjr, j); k++) {
          int pm = blocb[i];
          for (int l = 0; l < nJTs(B, bjl, i); l++) {
            R me = trdb[pt++] * meV[ntb];
            if (!h->thr_on || fabs((double)me) >= 0.0000001) {
              if (trunc) {
                for (int q = 0; q < dar; q++)
                  if (F1[pu + q]) {
                    B3[pm + q] += me * B1[pu + q];
                    F2[pm + q] = 1;
                  }
              } else {
                for (int q = 0; q < dar; q++) B3[pm + q] += me * B1[pu + q];
              }
            }
            pm += dar;
          }
          pu += dar;
        }
      }
  }
  (void)dbl;
}

/* single_product1: transpose B3 (dar x dbl) to dbl-fastest, apply the a-density
   (rows i = a-partitions of Ja, columns j = a-partitions of Ja'), transpose back
   and scatter-add into the output block (masked a rows laro, b columns lbro).
   On a truncated basis the flags of B3 (F2) travel with it (F1) and an element
   adds only where flagged (:846, :885-889). */
static void SFX(single_product1)(const refcpu_t *h, R *bveco, int64_t poso, int ajl, int ajr,
                                 int bjl, int atsc, int dar, int dal, int dbl,
                                 const uint32_t *laco, const uint32_t *laro, const uint32_t *lbro,
                                 R *B2, const R *B3, R *B4, uint8_t *F1, const uint8_t *F2) {
  const side_t *A = &h->A, *B = &h->B;
  int nsa = A->nsp, nsb = B->nsp;
  const int trunc = h->trunc != NULL;
  for (int64_t q = 0; q < (int64_t)dal * dbl; q++) B4[q] = (R)0;
  int64_t p = 0;
  for (int j = 0; j < dbl; j++) {
    for (int r = 0; r < dar; r++) B2[j + (int64_t)dbl * r] = B3[p + r];
    if (trunc)
      for (int r = 0; r < dar; r++) F1[j + (int64_t)dbl * r] = F2[p + r];
    p += dar;
  }
  int *snJTal = (int *)alloca(sizeof(int) * (nsa + 1));
  int *snJTar = (int *)alloca(sizeof(int) * (nsa + 1));
  snJTal[0] = 0;
  snJTar[0] = 0;
  for (int j = 1; j < nsa; j++) {
    snJTal[j] = snJTal[j - 1] + (bit_of(laro, j - 1) ? dbl * nJTs(A, ajl, j - 1) : 0);
    snJTar[j] = snJTar[j - 1] + (bit_of(laco, j - 1) ? dbl * nJTs(A, ajr, j - 1) : 0);
  }
  int64_t pt = A->pos[atsc];
  const uint32_t *m = A->mask + (size_t)atsc * nsa * A->words;
  const R *trda = VALS(A);
  for (int j = 0; j < nsa; j++)
    for (int i = 0; i < nsa; i++) {
      if (!bit_of(m + (size_t)j * A->words, i)) continue;
      int pu = snJTar[j];
      for (int k = 0; k < nJTs(A, ajr, j); k++) {
        int pm = snJTal[i];
        for (int l = 0; l < nJTs(A, ajl, i); l++) {
          R me = trda[pt++];
          if (!h->thr_on || fabs((double)me) >= 0.0000001) {
            if (trunc) {
              for (int q = 0; q < dbl; q++)
                if (F1[pu + q]) B4[pm + q] += me * B2[pu + q];
            } else {
              for (int q = 0; q < dbl; q++) B4[pm + q] += me * B2[pu + q];
            }
          }
          pm += dbl;
        }
        pu += dbl;
      }
    }
  p = 0;
  for (int j = 0; j < dal; j++) {
    for (int c = 0; c < dbl; c++) B2[j + (int64_t)dal * c] = B4[p + c];
    p += dbl;
  }
  p = poso;
  int64_t pp = 0;
  for (int j = 0; j < nsb; j++)
    for (int k = 0; k < nJTs(B, bjl, j); k++)
      for (int i = 0; i < nsa; i++) {
        int nl = nJTs(A, ajl, i);
        if (bit_of(lbro, j) && bit_of(laro, i) && nl != 0) {
          for (int q = 0; q < nl; q++) bveco[p + q] += B2[pp + q];
          pp += nl;
        }
        p += nl;
      }
}

/* mulmtx: W_L += H_a V_L (trans) or W_L += V_L H_b^T (not trans) */
static void SFX(mulmtx)(const refcpu_t *h, R *bveco, const R *bveci, int64_t posb, int trans,
                        int ja, int jb, R *B1, R *B3) {
  const side_t *A = &h->A, *B = &h->B;
  int da = A->d0[ja / 2], db = B->d0[jb / 2];
  int64_t dp = (int64_t)da * db;
  if (trans) {
    const side_t *S = A;
    const R *mtx = MATS(S);
    int64_t p = posb;
    for (int j = 0; j < db; j++) {
      for (int r = 0; r < da; r++) B1[j + (int64_t)db * r] = bveci[p + r];
      p += da;
    }
    for (int64_t q = 0; q < dp; q++) B3[q] = (R)0;
    int64_t pt = S->posm[ja / 2];
    const uint32_t *m = S->mmask + (size_t)(ja / 2) * S->nsp * S->words;
    const int *sn = S->snJT0 + (ja / 2) * S->nsp;
    for (int j = 0; j < S->nsp; j++)
      for (int i = 0; i < S->nsp; i++) {
        if (!bit_of(m + (size_t)j * S->words, i)) continue;
        int64_t pu = (int64_t)db * sn[j];
        if (S->par[i] == S->par[j]) {
          for (int k = 0; k < nJTs(S, ja, j); k++) {
            int64_t pm = (int64_t)db * sn[i];
            for (int l = 0; l < nJTs(S, ja, i); l++) {
              R me = mtx[pt++];
              if (me != (R)0)
                for (int q = 0; q < db; q++) B3[pu + q] += me * B1[pm + q];
              if (i != j && me != (R)0)
                for (int q = 0; q < db; q++) B3[pm + q] += me * B1[pu + q];
              pm += db;
            }
            pu += db;
          }
        }
      }
    p = 0;
    for (int j = 0; j < da; j++) {
      for (int c = 0; c < db; c++) B1[j + (int64_t)da * c] = B3[p + c];
      p += db;
    }
    for (int64_t q = 0; q < dp; q++) bveco[posb + q] += B1[q];
  } else {
    const side_t *S = B;
    const R *mtx = MATS(S);
    int64_t pt = S->posm[jb / 2];
    const uint32_t *m = S->mmask + (size_t)(jb / 2) * S->nsp * S->words;
    const int *sn = S->snJT0 + (jb / 2) * S->nsp;
    for (int j = 0; j < S->nsp; j++)
      for (int i = 0; i < S->nsp; i++) {
        if (!bit_of(m + (size_t)j * S->words, i)) continue;
        int64_t pu = (int64_t)da * sn[j];
        if (S->par[j] == S->par[i]) {
          for (int k = 0; k < nJTs(S, jb, j); k++) {
            int64_t pm = (int64_t)da * sn[i];
            for (int l = 0; l < nJTs(S, jb, i); l++) {
              R me = mtx[pt++];
              if (me != (R)0)
                for (int q = 0; q < da; q++) bveco[posb + pu + q] += me * bveci[posb + pm + q];
              if (i != j && me != (R)0)
                for (int q = 0; q < da; q++) bveco[posb + pm + q] += me * bveci[posb + pu + q];
              pm += da;
            }
            pu += da;
          }
        }
      }
  }
}

static int SFX(hv)(refcpu_t *h, const double *v_in, double *v_out) {
  int64_t N = h->N;
  int nthr = h->nthr;
  R *coefR = (R *)malloc(sizeof(R) * (N ? N : 1));
  R *coefL = (R *)malloc(sizeof(R) * (N ? N : 1));
  R *coefX = (R *)calloc((size_t)nthr * (N ? N : 1), sizeof(R));
  for (int64_t i = 0; i < N; i++) coefR[i] = (R)v_in[i];
  size_t maxdimv = (size_t)(h->maxda > 0 ? h->maxda : 1) * (h->maxdb > 0 ? h->maxdb : 1);
  const side_t *A = &h->A, *B = &h->B;

  /* ---- multiplyx (src/NuTRL/NuCoreX.f90:128-226) */
  if (!(h->azero || h->bzero)) {
    int delj2 = (A->maxJ2 - A->minJ2) / A->delJ2 + 1;
    int loop = h->basd * delj2;
#pragma omp parallel num_threads(nthr)
    {
      int thrn = omp_get_thread_num();
      R *B1 = (R *)malloc(sizeof(R) * maxdimv);
      R *B2 = (R *)malloc(sizeof(R) * maxdimv);
      R *B3 = (R *)calloc(maxdimv, sizeof(R));
      uint8_t *F1 = (uint8_t *)calloc(maxdimv, 1), *F2 = (uint8_t *)calloc(maxdimv, 1);
      int wa = A->words, wb = B->words;
      uint32_t *lvac = (uint32_t *)calloc(wa, 4), *lvar = (uint32_t *)calloc(wa, 4);
      uint32_t *lvbc = (uint32_t *)calloc(wb, 4), *lvbr = (uint32_t *)calloc(wb, 4);
      uint32_t *nplrbo = (uint32_t *)calloc((size_t)B->nsp * wb, 4);
      uint32_t *nplrbop = (uint32_t *)calloc((size_t)B->nsp * wb, 4);
      int *blocb = (int *)calloc(B->nsp + 1, sizeof(int));
      int lcap = 256;
      int *locb = (int *)malloc(sizeof(int) * lcap);
      R *mev = (R *)malloc(sizeof(R) * lcap);
      R *X = coefX + (size_t)thrn * N;
#pragma omp for schedule(dynamic, 1)
      for (int lp = 0; lp < loop; lp++) {
        int bsc = lp / delj2;
        int dj2 = lp - bsc * delj2;
        if (h->bas_log && !h->bas_log[bsc]) continue; /* truncated basis (:155-157) */
        int ajl = h->bas_a[bsc], bjl = h->bas_b[bsc];
        int ajr = A->minJ2 + A->delJ2 * dj2;
        const tab_t *t = &h->tab[bsc * h->nslot + ajr / 2];
        for (int nsc = 0; nsc < t->na; nsc++) {
          int b0 = t->boff[nsc], b1 = t->boff[nsc + 1];
          if (b1 == b0) continue;
          if (b1 - b0 > lcap) {
            lcap = b1 - b0;
            locb = (int *)realloc(locb, sizeof(int) * lcap);
            mev = (R *)realloc(mev, sizeof(R) * lcap);
          }
          int atsc = t->a[nsc];
          crproj(A->mask + (size_t)atsc * A->nsp * wa, A->nsp, A->nsp, wa, lvac, lvar);
          memset(nplrbo, 0, 4 * (size_t)B->nsp * wb);
          memset(nplrbop, 0, 4 * (size_t)B->nsp * wb);
          for (int q = b0; q < b1; q++) laor(nplrbo, B->mask + (size_t)h->tb_b[q] * B->nsp * wb, B->nsp, wb);
          crproj(nplrbo, B->nsp, B->nsp, wb, lvbc, lvbr);
          /* condim (:721-743) */
          int dar = 0, dal = 0, dbl = 0;
          for (int i = 0; i < A->nsp; i++) if (bit_of(lvac, i)) dar += nJTs(A, ajr, i);
          for (int i = 0; i < A->nsp; i++) if (bit_of(lvar, i)) dal += nJTs(A, ajl, i);
          for (int i = 0; i < B->nsp; i++) {
            blocb[i] = dbl * dar;
            if (bit_of(lvbr, i)) dbl += nJTs(B, bjl, i);
          }
          int lscs = h->tb_h[b0];
          int bjrs = h->bas_b[lscs];
          int zero = 1, ntr = 0;
          for (int q = b0; q < b1; q++) {
            int btsc = h->tb_b[q], lsc = h->tb_h[q];
            int bjr = h->bas_b[lsc];
            R me = TBME(h)[q];
            if (lscs != lsc) {
              crproj(nplrbop, B->nsp, B->nsp, wb, lvbc, lvbr);
              SFX(single_productm)(h, coefR, h->posnv[lscs], mev, ajr, bjl, bjrs, lvac, lvbc, blocb,
                                   dar, dbl, zero, B1, B2, ntr, locb, F1, F2);
              ntr = 0;
              memset(nplrbop, 0, 4 * (size_t)B->nsp * wb);
              bjrs = bjr;
              lscs = lsc;
              zero = 0;
            }
            if (me != (R)0) {
              locb[ntr] = btsc;
              mev[ntr] = me;
              ntr++;
              laor(nplrbop, B->mask + (size_t)btsc * B->nsp * wb, B->nsp, wb);
            }
          }
          if (ntr != 0) {
            crproj(nplrbop, B->nsp, B->nsp, wb, lvbc, lvbr);
            SFX(single_productm)(h, coefR, h->posnv[lscs], mev, ajr, bjl, bjrs, lvac, lvbc, blocb,
                                 dar, dbl, zero, B1, B2, ntr, locb, F1, F2);
          }
          crproj(nplrbo, B->nsp, B->nsp, wb, lvbc, lvbr);
          SFX(single_product1)(h, X, h->posnv[bsc], ajl, ajr, bjl, atsc, dar, dal, dbl, lvac, lvar,
                               lvbr, B1, B2, B3, F1, F2);
        }
      }
      free(B1); free(B2); free(B3); free(F1); free(F2); free(lvac); free(lvar); free(lvbc); free(lvbr);
      free(nplrbo); free(nplrbop); free(blocb); free(locb); free(mev);
    }
  }
  for (int64_t i = 0; i < N; i++) coefL[i] = coefX[i];
  for (int t = 1; t < nthr; t++)
    for (int64_t i = 0; i < N; i++) coefL[i] += coefX[(size_t)t * N + i];

  /* ---- multiplya (:76-99) */
  if (!h->azero) {
#pragma omp parallel num_threads(nthr)
    {
      R *B1 = (R *)malloc(sizeof(R) * maxdimv);
      R *B3 = (R *)malloc(sizeof(R) * maxdimv);
#pragma omp for schedule(dynamic, 1)
      for (int bsc = 0; bsc < h->basd; bsc++) {
        int ja = h->bas_a[bsc], jb = h->bas_b[bsc];
        if (h->bas_log && !h->bas_log[bsc]) continue; /* :86-88 */
        if (!A->has_mat[ja / 2]) continue;
        SFX(mulmtx)(h, coefL, coefR, h->posnv[bsc], 1, ja, jb, B1, B3);
      }
      free(B1); free(B3);
    }
  }
  /* ---- multiplyb (:101-125) */
  if (!h->bzero) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthr)
    for (int bsc = 0; bsc < h->basd; bsc++) {
      int ja = h->bas_a[bsc], jb = h->bas_b[bsc];
      if (h->bas_log && !h->bas_log[bsc]) continue; /* :112-114 */
      if (!B->has_mat[jb / 2]) continue;
      SFX(mulmtx)(h, coefL, coefR, h->posnv[bsc], 0, ja, jb, NULL, NULL);
    }
  }
  for (int64_t i = 0; i < N; i++) v_out[i] = (double)coefL[i];
  free(coefR); free(coefL); free(coefX);
  return 0;
}
