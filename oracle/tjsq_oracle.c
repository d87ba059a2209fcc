/*
 * tjsq_oracle.c -- CPU restatement of the tjsq H_eff apply (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker the tests, smoke() and the
 * bench's cpu_baseline leg compare the HIP path against.  Nothing in the
 * product links or calls it.
 *
 * Each function follows one routine of DongNingSheng/Grand_SQL_tJ
 * (code/tjsq/<name>.f90, cited file:line) in loop order, search order ("first
 * match" of the reference's linear searches) and coefficient formula, so that
 * results agree with the reference to rounding of the GEMM summation order.
 * GEMMs are written out as plain loops (column-major, BLAS argument meaning).
 * Dead code of the reference (branches after unconditional `go to`, the
 * add_op = 0 branch) is not restated.
 */
#include "tjsq_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SU 2      /* spin step (common.f90:11, su = 2) */
#define NEIBT 12  /* neighbour slots per site (common.f90:34) */

/* ------------------------------------------------------------------ */
/* Factorials and the 6j symbol                                        */
/* ------------------------------------------------------------------ */

static double g_fact[302];
static int g_fact_ready = 0;

/* FACTRL, block.f90:1508-1516: fact(i) = fact(i-1) * float(i), i <= 142. */
static void factrl(void) {
  if (g_fact_ready) return;
  g_fact[0] = 1.0;
  for (int i = 1; i <= 142; ++i) g_fact[i] = g_fact[i - 1] * (double)(float)i;
  g_fact_ready = 1;
}

/* W6JS, block.f90:1564-1615: Racah single sum on doubled arguments,
 * evaluated in the reference's operation order. */
double oracle_w6js(int j1, int j2, int j3, int l1, int l2, int l3) {
  factrl();
  const double* F = g_fact;
  int ia = j1 + j2; if (ia < j3) return 0.0;
  int ib = j1 - j2; if (abs(ib) > j3) return 0.0;
  int ic = j1 + l2; if (ic < l3) return 0.0;
  int id = j1 - l2; if (abs(id) > l3) return 0.0;
  int ie = l1 + j2; if (ie < l3) return 0.0;
  int if_ = l1 - j2; if (abs(if_) > l3) return 0.0;
  int ig = l1 + l2; if (ig < j3) return 0.0;
  int ih = l1 - l2; if (abs(ih) > j3) return 0.0;
  int sum1 = ia + j3, sum2 = ic + l3, sum3 = ie + l3, sum4 = ig + j3;
  if (sum1 % 2 || sum2 % 2 || sum3 % 2 || sum4 % 2) return 0.0;
  int wmin = sum1; if (sum2 > wmin) wmin = sum2; if (sum3 > wmin) wmin = sum3; if (sum4 > wmin) wmin = sum4;
  int ii = ia + ig, ij = j2 + j3 + l2 + l3, ik = j3 + j1 + l3 + l1;
  int wmax = ii; if (ij < wmax) wmax = ij; if (ik < wmax) wmax = ik;
  double omega = 0.0;
  for (int w = wmin; w <= wmax; w += 2) {
    double denom = F[(w - sum1) / 2] * F[(w - sum2) / 2] * F[(w - sum3) / 2] * F[(w - sum4) / 2] *
                   F[(ii - w) / 2] * F[(ij - w) / 2] * F[(ik - w) / 2];
    if (w % 4 != 0) denom = -denom;
    omega = omega + F[w / 2 + 1] / denom;
  }
  double th1 = F[(ia - j3) / 2] * F[(j3 + ib) / 2] * F[(j3 - ib) / 2] / F[sum1 / 2 + 1];
  double th2 = F[(ic - l3) / 2] * F[(l3 + id) / 2] * F[(l3 - id) / 2] / F[sum2 / 2 + 1];
  double th3 = F[(ie - l3) / 2] * F[(l3 + if_) / 2] * F[(l3 - if_) / 2] / F[sum3 / 2 + 1];
  double th4 = F[(ig - j3) / 2] * F[(j3 + ih) / 2] * F[(j3 - ih) / 2] / F[sum4 / 2 + 1];
  double theta = th1 * th2 * th3 * th4;
  return omega * sqrt(theta);
}

/* (-1)**n for the integer exponents the kernels use. */
static double neg1pow(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }

/* ------------------------------------------------------------------ */
/* Block-sparse containers (views of the flat arrays)                  */
/* ------------------------------------------------------------------ */

typedef struct { int num_up, num_down, down_dif, row_dim, sdim; double* mat; } OSub;
typedef struct { int len, up_dif, down_dif, num; OSub* sub; int owns; } OBlock;
typedef struct { int bl_up, bl_down, st_up, st_down, spos, sdim; } OMapSub;
typedef struct { int new_up, new_down, dim, num; const OMapSub* sub; } OMap;
typedef struct { int num; OMap* sec; } OBasis;
typedef struct { int sys_up, sys_down, env_up, env_down, sys_dim, env_dim; double* vec; } OWSub;
typedef struct { int num; OWSub* sub; } OWave;

static void* xcalloc(size_t n, size_t s) {
  void* p = calloc(n ? n : 1, s ? s : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

static OBlock view_op(const int* hdr, const int* sub, const int64_t* off, const double* data) {
  OBlock b;
  b.len = hdr[0]; b.up_dif = hdr[1]; b.down_dif = hdr[2]; b.num = hdr[3]; b.owns = 0;
  b.sub = (OSub*)xcalloc((size_t)b.num, sizeof(OSub));
  for (int i = 0; i < b.num; ++i) {
    const int* s = sub + 5 * i;
    b.sub[i].num_up = s[0]; b.sub[i].num_down = s[1]; b.sub[i].down_dif = s[2];
    b.sub[i].row_dim = s[3]; b.sub[i].sdim = s[4];
    b.sub[i].mat = (double*)(data + off[i]);
  }
  return b;
}

static void free_block(OBlock* b) {
  if (!b->sub) return;
  if (b->owns)
    for (int i = 0; i < b->num; ++i) free(b->sub[i].mat);
  free(b->sub);
  b->sub = NULL; b->num = 0;
}

static OBasis view_basis(int nsec, const int* bsec, const int* bsub) {
  OBasis B; B.num = nsec; B.sec = (OMap*)xcalloc((size_t)nsec, sizeof(OMap));
  const OMapSub* s = (const OMapSub*)bsub;
  for (int i = 0; i < nsec; ++i) {
    B.sec[i].new_up = bsec[4 * i]; B.sec[i].new_down = bsec[4 * i + 1];
    B.sec[i].dim = bsec[4 * i + 2]; B.sec[i].num = bsec[4 * i + 3];
    B.sec[i].sub = s; s += B.sec[i].num;
  }
  return B;
}

static OWave view_wave(int nw, const int* wsec, double* data) {
  OWave W; W.num = nw; W.sub = (OWSub*)xcalloc((size_t)nw, sizeof(OWSub));
  size_t off = 0;
  for (int i = 0; i < nw; ++i) {
    const int* s = wsec + 6 * i;
    W.sub[i].sys_up = s[0]; W.sub[i].sys_down = s[1]; W.sub[i].env_up = s[2];
    W.sub[i].env_down = s[3]; W.sub[i].sys_dim = s[4]; W.sub[i].env_dim = s[5];
    W.sub[i].vec = data + off; off += (size_t)s[4] * (size_t)s[5];
  }
  return W;
}

/* Column-major GEMM with BLAS argument meaning: C = alpha*op(A)*op(B) + beta*C. */
static void gemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
                 const double* B, int ldb, double beta, double* C, int ldc) {
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < m; ++i) {
      double acc = 0.0;
      for (int l = 0; l < k; ++l) {
        double a = (ta == 'N') ? A[i + (size_t)l * lda] : A[l + (size_t)i * lda];
        double b = (tb == 'N') ? B[l + (size_t)j * ldb] : B[j + (size_t)l * ldb];
        acc += a * b;
      }
      double c = (beta == 0.0) ? 0.0 : beta * C[i + (size_t)j * ldc];
      C[i + (size_t)j * ldc] = c + alpha * acc;
    }
  }
}

/* first sub-block with ket (up, down) [and down_dif if dd_check] */
static int find_op(const OBlock* b, int up, int down, int dd_check, int dd) {
  for (int k = 0; k < b->num; ++k)
    if (b->sub[k].num_up == up && b->sub[k].num_down == down && (!dd_check || b->sub[k].down_dif == dd))
      return k;
  return -1;
}

static int find_sec(const OBasis* bs, int up, int down) {
  for (int x = 0; x < bs->num; ++x)
    if (bs->sec[x].new_up == up && bs->sec[x].new_down == down) return x;
  return -1;
}

static int find_wave(const OWave* w, int su, int sd, int eu, int ed) {
  for (int k = 0; k < w->num; ++k)
    if (w->sub[k].sys_up == su && w->sub[k].sys_down == sd && w->sub[k].env_up == eu && w->sub[k].env_down == ed)
      return k;
  return -1;
}

/* pointer to element (r, c) of a wave sector (0-based) */
#define WV(w, i, r, c) ((w)->sub[i].vec + (size_t)(c) * (w)->sub[i].sys_dim + (r))

/* ------------------------------------------------------------------ */
/* Site operators (Get_site_operator_su2_U0, dmrg_sub.f90:370-473)      */
/* ------------------------------------------------------------------ */

typedef struct { OBlock elec_up, elec_down, num_elec, st_sd, st_si; double vals[8]; OSub subs[8]; } SiteOps;

static void make_site_ops(SiteOps* s, int ub) {
  memset(s, 0, sizeof(*s));
  s->vals[0] = -sqrt(2.0);                   /* <1||c+||0> (dmrg_sub.f90:391) */
  s->vals[1] = -sqrt(2.0);                   /* st_elec_down (block_transfer copy) */
  s->vals[2] = 0.0; s->vals[3] = sqrt(2.0);  /* num_elec (:415,:423) */
  s->vals[4] = sqrt(1.5);                    /* st_sd (:442) */
  s->vals[5] = 1.0; s->vals[6] = sqrt(2.0);  /* st_si (:470-472) */
  OSub* q = s->subs;
  q[0] = (OSub){0, 0, 1, 1, 1, &s->vals[0]};
  s->elec_up = (OBlock){1, ub, 1, 1, &q[0], 0};
  q[1] = (OSub){ub, 1, -1, 1, 1, &s->vals[1]};
  s->elec_down = (OBlock){1, -ub, 1, 1, &q[1], 0};
  q[2] = (OSub){0, 0, 0, 1, 1, &s->vals[2]};
  q[3] = (OSub){ub, 1, 0, 1, 1, &s->vals[3]};
  s->num_elec = (OBlock){1, 0, 0, 2, &q[2], 0};
  q[4] = (OSub){ub, 1, 0, 1, 1, &s->vals[4]};
  s->st_sd = (OBlock){1, 0, SU, 1, &q[4], 0};
  q[5] = (OSub){0, 0, 0, 1, 1, &s->vals[5]};
  q[6] = (OSub){ub, 1, 0, 1, 1, &s->vals[6]};
  s->st_si = (OBlock){1, 0, 0, 2, &q[5], 0};
}

/* ------------------------------------------------------------------ */
/* The 14 apply kernels (get_sys.f90)                                  */
/* ------------------------------------------------------------------ */

/* block_vec_dia, get_sys.f90:1242-1289 */
static void k_block_vec_dia(const OBlock* bl, const OBasis* bs, const OWave* in, OWave* out, double coef) {
  for (int i = 0; i < in->num; ++i) {
    int env_dim = in->sub[i].env_dim;
    for (int x = 0; x < bs->num; ++x) {
      if (bs->sec[x].new_up != in->sub[i].sys_up || bs->sec[x].new_down != in->sub[i].sys_down) continue;
      for (int y = 0; y < bs->sec[x].num; ++y) {
        const OMapSub* m = &bs->sec[x].sub[y];
        double c1 = coef / sqrt(1.0 + m->bl_down);
        int k = find_op(bl, m->bl_up, m->bl_down, 0, 0);
        if (k < 0) continue;
        int ld = in->sub[i].sys_dim;
        /* Y(spos+1:spos+sdim, :) += c1 * A * X(spos+1:spos+sdim, :) */
        gemm('N', 'N', m->sdim, env_dim, m->sdim, c1, bl->sub[k].mat, m->sdim, WV(in, i, m->spos, 0), ld, 1.0,
             WV(out, i, m->spos, 0), ld);
      }
    }
  }
}

/* vec_block_dia, get_sys.f90:1293-1340 */
static void k_vec_block_dia(const OBlock* bl, const OBasis* bs, const OWave* in, OWave* out, double coef) {
  for (int i = 0; i < in->num; ++i) {
    int sys_dim = in->sub[i].sys_dim;
    for (int x = 0; x < bs->num; ++x) {
      if (bs->sec[x].new_up != in->sub[i].env_up || bs->sec[x].new_down != in->sub[i].env_down) continue;
      for (int y = 0; y < bs->sec[x].num; ++y) {
        const OMapSub* m = &bs->sec[x].sub[y];
        double c1 = coef / sqrt(1.0 + m->bl_down);
        int k = find_op(bl, m->bl_up, m->bl_down, 0, 0);
        if (k < 0) continue;
        gemm('N', 'T', sys_dim, m->sdim, m->sdim, c1, WV(in, i, 0, m->spos), sys_dim, bl->sub[k].mat, m->sdim, 1.0,
             WV(out, i, 0, m->spos), sys_dim);
      }
    }
  }
}

/* block_site_vec_dia, get_sys.f90:1460-1530 */
static void k_block_site_vec_dia(const OBlock* bl, const OBlock* st, const OBasis* bs, const OWave* in, OWave* out,
                                 double coef) {
  for (int i = 0; i < in->num; ++i) {
    int env_dim = in->sub[i].env_dim, ld = in->sub[i].sys_dim;
    for (int x = 0; x < bs->num; ++x) {
      if (bs->sec[x].new_up != in->sub[i].sys_up || bs->sec[x].new_down != in->sub[i].sys_down) continue;
      for (int y = 0; y < bs->sec[x].num; ++y) {
        const OMapSub* m = &bs->sec[x].sub[y];
        int b = find_op(bl, m->bl_up, m->bl_down, 0, 0);
        int s = find_op(st, m->st_up, m->st_down, 0, 0);
        if (b < 0 || s < 0) continue;
        double c = coef * st->sub[s].mat[0];
        c = c / sqrt((1.0 + m->bl_down) * (1.0 + m->st_down));
        gemm('N', 'N', m->sdim, env_dim, m->sdim, c, bl->sub[b].mat, m->sdim, WV(in, i, m->spos, 0), ld, 1.0,
             WV(out, i, m->spos, 0), ld);
      }
    }
  }
}

/* vec_block_site_dia, get_sys.f90:1534-1605 */
static void k_vec_block_site_dia(const OBlock* bl, const OBlock* st, const OBasis* bs, const OWave* in, OWave* out,
                                 double coef) {
  for (int i = 0; i < in->num; ++i) {
    int sys_dim = in->sub[i].sys_dim;
    for (int x = 0; x < bs->num; ++x) {
      if (bs->sec[x].new_up != in->sub[i].env_up || bs->sec[x].new_down != in->sub[i].env_down) continue;
      for (int y = 0; y < bs->sec[x].num; ++y) {
        const OMapSub* m = &bs->sec[x].sub[y];
        int b = find_op(bl, m->bl_up, m->bl_down, 0, 0);
        int s = find_op(st, m->st_up, m->st_down, 0, 0);
        if (b < 0 || s < 0) continue;
        double c = coef * st->sub[s].mat[0];
        c = c / sqrt((1.0 + m->bl_down) * (1.0 + m->st_down));
        gemm('N', 'T', sys_dim, m->sdim, m->sdim, c, WV(in, i, 0, m->spos), sys_dim, bl->sub[b].mat, m->sdim, 1.0,
             WV(out, i, 0, m->spos), sys_dim);
      }
    }
  }
}

/* Shared prologue of the "dia" two-side kernels: sector ids of wave sector i. */
static int two_side_secs(const OBasis* sbs, const OBasis* ebs, const OWSub* w, int* xs, int* xe) {
  *xs = find_sec(sbs, w->sys_up, w->sys_down);
  *xe = find_sec(ebs, w->env_up, w->env_down);
  return *xs >= 0 && *xe >= 0;
}

/* site_vec_site_dia, get_sys.f90:1609-1705 */
static void k_site_vec_site_dia(const OBlock* sst, const OBlock* est, const OBasis* sbs, const OBasis* ebs,
                                const OWave* in, OWave* out, double coef) {
  for (int i = 0; i < in->num; ++i) {
    int xs, xe;
    if (!two_side_secs(sbs, ebs, &in->sub[i], &xs, &xe)) continue;
    int ld = in->sub[i].sys_dim;
    for (int x = 0; x < sbs->sec[xs].num; ++x) {
      const OMapSub* ms = &sbs->sec[xs].sub[x];
      int s = find_op(sst, ms->st_up, ms->st_down, 0, 0);
      for (int y = 0; y < ebs->sec[xe].num; ++y) {
        const OMapSub* me = &ebs->sec[xe].sub[y];
        int e = find_op(est, me->st_up, me->st_down, 0, 0);
        if (s < 0 || e < 0) continue;
        double c = coef * sst->sub[s].mat[0] * est->sub[e].mat[0];
        c = c / sqrt((1.0 + ms->st_down) * (1.0 + me->st_down));
        for (int cc = 0; cc < me->sdim; ++cc)
          for (int r = 0; r < ms->sdim; ++r)
            *WV(out, i, ms->spos + r, me->spos + cc) += c * *WV(in, i, ms->spos + r, me->spos + cc);
        (void)ld;
      }
    }
  }
}

/* block_vec_site_dia, get_sys.f90:1709-1805 */
static void k_block_vec_site_dia(const OBlock* sbl, const OBlock* est, const OBasis* sbs, const OBasis* ebs,
                                 const OWave* in, OWave* out, double coef) {
  for (int i = 0; i < in->num; ++i) {
    int xs, xe;
    if (!two_side_secs(sbs, ebs, &in->sub[i], &xs, &xe)) continue;
    int ld = in->sub[i].sys_dim;
    for (int x = 0; x < sbs->sec[xs].num; ++x) {
      const OMapSub* ms = &sbs->sec[xs].sub[x];
      int b = find_op(sbl, ms->bl_up, ms->bl_down, 0, 0);
      for (int y = 0; y < ebs->sec[xe].num; ++y) {
        const OMapSub* me = &ebs->sec[xe].sub[y];
        int e = find_op(est, me->st_up, me->st_down, 0, 0);
        if (b < 0 || e < 0) continue;
        double c = coef * est->sub[e].mat[0];
        c = c / sqrt((1.0 + me->st_down) * (1.0 + ms->bl_down));
        gemm('N', 'N', ms->sdim, me->sdim, ms->sdim, c, sbl->sub[b].mat, ms->sdim, WV(in, i, ms->spos, me->spos), ld,
             1.0, WV(out, i, ms->spos, me->spos), ld);
      }
    }
  }
}

/* site_vec_block_dia, get_sys.f90:1809-1906 */
static void k_site_vec_block_dia(const OBlock* sst, const OBlock* ebl, const OBasis* sbs, const OBasis* ebs,
                                 const OWave* in, OWave* out, double coef) {
  for (int i = 0; i < in->num; ++i) {
    int xs, xe;
    if (!two_side_secs(sbs, ebs, &in->sub[i], &xs, &xe)) continue;
    int ld = in->sub[i].sys_dim;
    for (int x = 0; x < sbs->sec[xs].num; ++x) {
      const OMapSub* ms = &sbs->sec[xs].sub[x];
      int s = find_op(sst, ms->st_up, ms->st_down, 0, 0);
      for (int y = 0; y < ebs->sec[xe].num; ++y) {
        const OMapSub* me = &ebs->sec[xe].sub[y];
        int e = find_op(ebl, me->bl_up, me->bl_down, 0, 0);
        if (s < 0 || e < 0) continue;
        double c = coef * sst->sub[s].mat[0];
        c = c / sqrt((1.0 + me->bl_down) * (1.0 + ms->st_down));
        gemm('N', 'T', ms->sdim, me->sdim, me->sdim, c, WV(in, i, ms->spos, me->spos), ld, ebl->sub[e].mat, me->sdim,
             1.0, WV(out, i, ms->spos, me->spos), ld);
      }
    }
  }
}

/* block_vec_block_dia, get_sys.f90:1910-2012 */
static void k_block_vec_block_dia(const OBlock* sbl, const OBlock* ebl, const OBasis* sbs, const OBasis* ebs,
                                  const OWave* in, OWave* out, double coef) {
  for (int i = 0; i < in->num; ++i) {
    int xs, xe;
    if (!two_side_secs(sbs, ebs, &in->sub[i], &xs, &xe)) continue;
    int ld = in->sub[i].sys_dim;
    for (int x = 0; x < sbs->sec[xs].num; ++x) {
      const OMapSub* ms = &sbs->sec[xs].sub[x];
      int b = find_op(sbl, ms->bl_up, ms->bl_down, 0, 0);
      for (int y = 0; y < ebs->sec[xe].num; ++y) {
        const OMapSub* me = &ebs->sec[xe].sub[y];
        int e = find_op(ebl, me->bl_up, me->bl_down, 0, 0);
        if (b < 0 || e < 0) continue;
        double* mat = (double*)xcalloc((size_t)ms->sdim * me->sdim, sizeof(double));
        double c1 = coef / sqrt((1.0 + me->bl_down) * (1.0 + ms->bl_down));
        gemm('N', 'N', ms->sdim, me->sdim, ms->sdim, 1.0, sbl->sub[b].mat, ms->sdim, WV(in, i, ms->spos, me->spos), ld,
             0.0, mat, ms->sdim);
        gemm('N', 'T', ms->sdim, me->sdim, me->sdim, c1, mat, ms->sdim, ebl->sub[e].mat, me->sdim, 1.0,
             WV(out, i, ms->spos, me->spos), ld);
        free(mat);
      }
    }
  }
}

static double fermion_sign(char flag, int n) {
  if (flag == 'B') return 1.0;
  return (n % 2 == 0) ? 1.0 : -1.0;
}

/* block_site_vec_ndia, get_sys.f90:2016-2265 (live part :2049-2191) */
static void k_block_site_vec_ndia(const OBlock* bl, char fbl, const OBlock* st, char fst, const OBasis* bs,
                                  const OWave* in, OWave* out, double coef) {
  (void)fbl;
  int up_dif = bl->up_dif, d1 = bl->down_dif;
  if (up_dif != st->up_dif || d1 != st->down_dif) { fprintf(stderr, "oracle: QN error in block_site_vec\n"); abort(); }
  for (int i = 0; i < in->num; ++i) {
    int env_dim = in->sub[i].env_dim, ld = in->sub[i].sys_dim, sys_down = in->sub[i].sys_down;
    for (int k = 0; k < bs->num; ++k) {
      if (bs->sec[k].new_up != in->sub[i].sys_up || bs->sec[k].new_down != sys_down) continue;
      for (int x = 0; x < bs->sec[k].num; ++x) {
        const OMapSub* mo = &bs->sec[k].sub[x];
        for (int dd = -d1; dd <= d1; dd += SU) {
          for (int sdd = -d1; sdd <= d1; sdd += SU) {
            int b = find_op(bl, mo->bl_up, mo->bl_down, 1, dd);
            int nst_up = mo->st_up - up_dif, nst_down = mo->st_down - sdd;
            int s = find_op(st, nst_up, nst_down, 1, sdd);
            int nbl_up = mo->bl_up + up_dif, nbl_down = mo->bl_down + dd;
            if (b < 0 || s < 0) continue;
            int y = -1;
            for (int q = 0; q < bs->sec[k].num; ++q) {
              const OMapSub* mq = &bs->sec[k].sub[q];
              if (mq->bl_up == nbl_up && mq->bl_down == nbl_down && mq->st_up == nst_up && mq->st_down == nst_down) { y = q; break; }
            }
            if (y < 0) continue;
            const OMapSub* mn = &bs->sec[k].sub[y];
            double sign_st = fermion_sign(fst, mo->bl_down);
            int j1 = nbl_down, j2 = mo->bl_down, j3 = d1, j4 = mo->st_down, j5 = nst_down, j6 = sys_down;
            double c1 = oracle_w6js(j1, j2, j3, j4, j5, j6);
            if (c1 == 0.0) continue;
            c1 = c1 * neg1pow((j2 + j5 + j3 + j6) / 2) / sqrt(1.0 + d1);
            double ct = (1.0 * sign_st * c1) * coef * st->sub[s].mat[0];
            gemm('N', 'N', mn->sdim, env_dim, mo->sdim, ct, bl->sub[b].mat, mn->sdim, WV(in, i, mo->spos, 0), ld, 1.0,
                 WV(out, i, mn->spos, 0), ld);
            if (d1 == 1)
              gemm('T', 'N', mo->sdim, env_dim, mn->sdim, ct, bl->sub[b].mat, mn->sdim, WV(in, i, mn->spos, 0), ld, 1.0,
                   WV(out, i, mo->spos, 0), ld);
          }
        }
      }
    }
  }
}

/* vec_block_site_ndia, get_sys.f90:2269-2552 (live part :2303-2440) */
static void k_vec_block_site_ndia(const OBlock* bl, char fbl, const OBlock* st, char fst, const OBasis* bs,
                                  const OWave* in, OWave* out, double coef) {
  int up_dif = bl->up_dif, d1 = bl->down_dif;
  if (up_dif != st->up_dif || d1 != st->down_dif) { fprintf(stderr, "oracle: QN error in vec_block_site\n"); abort(); }
  for (int i = 0; i < in->num; ++i) {
    int sys_dim = in->sub[i].sys_dim, sys_down = in->sub[i].sys_down, env_down = in->sub[i].env_down;
    for (int k = 0; k < bs->num; ++k) {
      if (bs->sec[k].new_up != in->sub[i].env_up || bs->sec[k].new_down != env_down) continue;
      for (int x = 0; x < bs->sec[k].num; ++x) {
        const OMapSub* mo = &bs->sec[k].sub[x];
        for (int bdd = -d1; bdd <= d1; bdd += SU) {
          int b = find_op(bl, mo->bl_up, mo->bl_down, 1, bdd);
          for (int sdd = -d1; sdd <= d1; sdd += SU) {
            int nst_up = mo->st_up - up_dif, nst_down = mo->st_down - sdd;
            int s = find_op(st, nst_up, nst_down, 1, sdd);
            int nbl_up = mo->bl_up + up_dif, nbl_down = mo->bl_down + bdd;
            if (b < 0 || s < 0) continue;
            int y = -1;
            for (int q = 0; q < bs->sec[k].num; ++q) {
              const OMapSub* mq = &bs->sec[k].sub[q];
              if (mq->bl_up == nbl_up && mq->bl_down == nbl_down && mq->st_up == nst_up && mq->st_down == nst_down) { y = q; break; }
            }
            if (y < 0) continue;
            const OMapSub* mn = &bs->sec[k].sub[y];
            double sign_bl = fermion_sign(fbl, sys_down + nst_down);
            double sign_st = fermion_sign(fst, sys_down);
            int j1 = nbl_down, j2 = mo->bl_down, j3 = d1, j4 = mo->st_down, j5 = nst_down, j6 = env_down;
            double c1 = oracle_w6js(j1, j2, j3, j4, j5, j6);
            if (c1 == 0.0) continue;
            c1 = c1 * neg1pow((j2 + j5 + j3 + j6) / 2) / sqrt(1.0 + d1);
            double ct = (sign_bl * sign_st * c1) * coef * st->sub[s].mat[0];
            gemm('N', 'T', sys_dim, mn->sdim, mo->sdim, ct, WV(in, i, 0, mo->spos), sys_dim, bl->sub[b].mat, mn->sdim,
                 1.0, WV(out, i, 0, mn->spos), sys_dim);
            if (d1 == 1)
              gemm('N', 'N', sys_dim, mo->sdim, mn->sdim, ct, WV(in, i, 0, mn->spos), sys_dim, bl->sub[b].mat, mn->sdim,
                   1.0, WV(out, i, 0, mo->spos), sys_dim);
          }
        }
      }
    }
  }
}

/* search of an enlarged-basis sub-block: sector (up, down), then sub with the
 * given bl/st quantum numbers; any argument < -999 is not compared. */
#define ANYQ (-100000)
static int find_bs_sub(const OBasis* bs, int up, int down, int bl_up, int bl_down, int st_up, int st_down,
                       const OMapSub** out) {
  for (int k = 0; k < bs->num; ++k) {
    if (bs->sec[k].new_up != up || bs->sec[k].new_down != down) continue;
    for (int l = 0; l < bs->sec[k].num; ++l) {
      const OMapSub* m = &bs->sec[k].sub[l];
      if (bl_up != ANYQ && m->bl_up != bl_up) continue;
      if (bl_down != ANYQ && m->bl_down != bl_down) continue;
      if (st_up != ANYQ && m->st_up != st_up) continue;
      if (st_down != ANYQ && m->st_down != st_down) continue;
      *out = m;
      return k;
    }
  }
  return -1;
}

/* block_vec_block_ndia, get_sys.f90:2555-2981 (live part :2590-2834) */
static void k_block_vec_block_ndia(const OBlock* sbl, char fsl, const OBlock* ebl, char fel, const OBasis* sbs,
                                   const OBasis* ebs, const OWave* in, OWave* out, double coef) {
  (void)fsl;
  int up_dif = sbl->up_dif, d1 = sbl->down_dif;
  if (up_dif != ebl->up_dif || d1 != ebl->down_dif) { fprintf(stderr, "oracle: QN error in block_vec_block\n"); abort(); }
  for (int i = 0; i < in->num; ++i) {
    const OWSub* w = &in->sub[i];
    for (int x = 0; x < sbs->num; ++x) {
      if (sbs->sec[x].new_up != w->sys_up || sbs->sec[x].new_down != w->sys_down) continue;
      for (int yy = 0; yy < ebs->num; ++yy) {
        if (ebs->sec[yy].new_up != w->env_up || ebs->sec[yy].new_down != w->env_down) continue;
        for (int m = 0; m < sbs->sec[x].num; ++m) {
          for (int n = 0; n < ebs->sec[yy].num; ++n) {
            const OMapSub* ms = &sbs->sec[x].sub[m];
            const OMapSub* me = &ebs->sec[yy].sub[n];
            for (int dd = -d1; dd <= d1; dd += SU) {
              int b = find_op(sbl, ms->bl_up, ms->bl_down, 1, dd);
              int nsb_up = ms->bl_up + up_dif, nsb_down = ms->bl_down + dd;
              for (int edd = -d1; edd <= d1; edd += SU) {
                int neb_up = me->bl_up - up_dif, neb_down = me->bl_down - edd;
                int e = find_op(ebl, neb_up, neb_down, 1, edd);
                if (b < 0 || e < 0) continue;
                for (int nsd = abs(w->sys_down - d1); nsd <= w->sys_down + d1; nsd += SU) {
                  int ns_up = w->sys_up + up_dif, ne_up = w->env_up - up_dif, ned = nsd;
                  const OMapSub *mns = NULL, *mne = NULL;
                  if (find_bs_sub(sbs, ns_up, nsd, nsb_up, nsb_down, ms->st_up, ms->st_down, &mns) < 0) continue;
                  if (find_bs_sub(ebs, ne_up, ned, neb_up, neb_down, me->st_up, me->st_down, &mne) < 0) continue;
                  int k = find_wave(out, ns_up, nsd, ne_up, ned);
                  if (k < 0) continue;
                  double sign_el = fermion_sign(fel, w->sys_down + me->st_down);
                  int j1 = nsd, j2 = d1, j3 = w->sys_down, j4 = ms->bl_down, j5 = ms->st_down, j6 = nsb_down;
                  int j11 = ned, j22 = d1, j33 = w->env_down, j44 = me->bl_down, j55 = me->st_down, j66 = neb_down;
                  double c1 = oracle_w6js(j1, j2, j3, j4, j5, j6);
                  if (c1 == 0.0) continue;
                  double c11 = oracle_w6js(j11, j22, j33, j44, j55, j66);
                  if (c11 == 0.0) continue;
                  c1 = c1 * c11;
                  if (c1 == 0.0) continue;
                  c1 = c1 * sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + d1));
                  if (((j6 + j5 + j3 + j2) / 2 + (j66 + j55 + j33 + j22) / 2) % 2 == 1) c1 = -c1;
                  double ct = (1.0 * sign_el) * coef * c1;
                  int lds = in->sub[i].sys_dim, ldk = out->sub[k].sys_dim;
                  double* mat = (double*)xcalloc((size_t)mns->sdim * me->sdim, sizeof(double));
                  gemm('N', 'N', mns->sdim, me->sdim, ms->sdim, 1.0, sbl->sub[b].mat, mns->sdim,
                       WV(in, i, ms->spos, me->spos), lds, 0.0, mat, mns->sdim);
                  gemm('N', 'N', mns->sdim, mne->sdim, me->sdim, ct, mat, mns->sdim, ebl->sub[e].mat, me->sdim, 1.0,
                       WV(out, k, mns->spos, mne->spos), ldk);
                  free(mat);
                  if (d1 == 1) {
                    double* mt = (double*)xcalloc((size_t)ms->sdim * mne->sdim, sizeof(double));
                    gemm('T', 'N', ms->sdim, mne->sdim, mns->sdim, 1.0, sbl->sub[b].mat, mns->sdim,
                         WV(in, k, mns->spos, mne->spos), ldk, 0.0, mt, ms->sdim);
                    gemm('N', 'T', ms->sdim, me->sdim, mne->sdim, ct, mt, ms->sdim, ebl->sub[e].mat, me->sdim, 1.0,
                         WV(out, i, ms->spos, me->spos), lds);
                    free(mt);
                  }
                }
              }
            }
          }
        }
      }
    }
  }
}

/* block_vec_site_ndia, get_sys.f90:2984-3351 (live part :3020-3212) */
static void k_block_vec_site_ndia(const OBlock* sbl, char fsl, const OBlock* est, char fet, const OBasis* sbs,
                                  const OBasis* ebs, const OWave* in, OWave* out, double coef) {
  (void)fsl;
  int up_dif = sbl->up_dif, d1 = sbl->down_dif;
  if (up_dif != est->up_dif || d1 != est->down_dif) { fprintf(stderr, "oracle: QN error in block_vec_site\n"); abort(); }
  for (int i = 0; i < in->num; ++i) {
    const OWSub* w = &in->sub[i];
    for (int x = 0; x < sbs->num; ++x) {
      if (sbs->sec[x].new_up != w->sys_up || sbs->sec[x].new_down != w->sys_down) continue;
      for (int yy = 0; yy < ebs->num; ++yy) {
        if (ebs->sec[yy].new_up != w->env_up || ebs->sec[yy].new_down != w->env_down) continue;
        for (int m = 0; m < sbs->sec[x].num; ++m) {
          for (int n = 0; n < ebs->sec[yy].num; ++n) {
            const OMapSub* ms = &sbs->sec[x].sub[m];
            const OMapSub* me = &ebs->sec[yy].sub[n];
            for (int bdd = -sbl->down_dif; bdd <= sbl->down_dif; bdd += SU) {
              int nsb_up = ms->bl_up + up_dif, nsb_down = ms->bl_down + bdd;
              int b = find_op(sbl, ms->bl_up, ms->bl_down, 1, bdd);
              for (int sdd = -est->down_dif; sdd <= est->down_dif; sdd += SU) {
                int nes_up = me->st_up - up_dif, nes_down = me->st_down - sdd;
                int s = find_op(est, nes_up, nes_down, 1, sdd);
                int ns_up = w->sys_up + up_dif, ne_up = w->env_up - up_dif;
                if (b < 0 || s < 0) continue;
                for (int nsd = abs(w->sys_down - sbl->down_dif); nsd <= w->sys_down + sbl->down_dif; nsd += SU) {
                  int ned = nsd;
                  const OMapSub *mns = NULL, *mne = NULL;
                  if (find_bs_sub(sbs, ns_up, nsd, nsb_up, nsb_down, ANYQ, ms->st_down, &mns) < 0) continue;
                  if (find_bs_sub(ebs, ne_up, ned, ANYQ, me->bl_down, nes_up, nes_down, &mne) < 0) continue;
                  int k = find_wave(out, ns_up, nsd, ne_up, ned);
                  if (k < 0) continue;
                  double sign_et = fermion_sign(fet, w->sys_down);
                  int j1 = nsd, j2 = d1, j3 = w->sys_down, j4 = ms->bl_down, j5 = ms->st_down, j6 = nsb_down;
                  int j11 = ned, j22 = d1, j33 = w->env_down, j44 = me->st_down, j55 = me->bl_down, j66 = nes_down;
                  double c1 = oracle_w6js(j1, j2, j3, j4, j5, j6);
                  if (c1 == 0.0) continue;
                  double c11 = oracle_w6js(j11, j22, j33, j44, j55, j66);
                  if (c11 == 0.0) continue;
                  c1 = c1 * c11;
                  if (c1 == 0.0) continue;
                  c1 = c1 * sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + d1));
                  if (((j6 + j5 + j3 + j2) / 2 + (j55 + j44 + j11 + j22) / 2) % 2 != 0) c1 = -c1;
                  double ct = (1.0 * sign_et * c1) * coef * est->sub[s].mat[0];
                  int lds = in->sub[i].sys_dim, ldk = out->sub[k].sys_dim;
                  gemm('N', 'N', mns->sdim, me->sdim, ms->sdim, ct, sbl->sub[b].mat, mns->sdim,
                       WV(in, i, ms->spos, me->spos), lds, 1.0, WV(out, k, mns->spos, mne->spos), ldk);
                  if (d1 == 1)
                    gemm('T', 'N', ms->sdim, me->sdim, mns->sdim, ct, sbl->sub[b].mat, mns->sdim,
                         WV(in, k, mns->spos, mne->spos), ldk, 1.0, WV(out, i, ms->spos, me->spos), lds);
                }
              }
            }
          }
        }
      }
    }
  }
}

/* site_vec_block_ndia, get_sys.f90:3355-3757 (live part :3390-3605) */
static void k_site_vec_block_ndia(const OBlock* sst, char fst, const OBlock* ebl, char fel, const OBasis* sbs,
                                  const OBasis* ebs, const OWave* in, OWave* out, double coef) {
  int up_dif = sst->up_dif, d1 = sst->down_dif;
  if (up_dif != ebl->up_dif || d1 != ebl->down_dif) { fprintf(stderr, "oracle: QN error in site_vec_block\n"); abort(); }
  for (int i = 0; i < in->num; ++i) {
    const OWSub* w = &in->sub[i];
    for (int x = 0; x < sbs->num; ++x) {
      if (sbs->sec[x].new_up != w->sys_up || sbs->sec[x].new_down != w->sys_down) continue;
      for (int yy = 0; yy < ebs->num; ++yy) {
        if (ebs->sec[yy].new_up != w->env_up || ebs->sec[yy].new_down != w->env_down) continue;
        for (int m = 0; m < sbs->sec[x].num; ++m) {
          for (int n = 0; n < ebs->sec[yy].num; ++n) {
            const OMapSub* ms = &sbs->sec[x].sub[m];
            const OMapSub* me = &ebs->sec[yy].sub[n];
            for (int sdd = -sst->down_dif; sdd <= sst->down_dif; sdd += SU) {
              int nss_up = ms->st_up + up_dif, nss_down = ms->st_down + sdd;
              int s = find_op(sst, ms->st_up, ms->st_down, 1, sdd);
              for (int bdd = -ebl->down_dif; bdd <= ebl->down_dif; bdd += SU) {
                int neb_up = me->bl_up - up_dif, neb_down = me->bl_down - bdd;
                int e = find_op(ebl, neb_up, neb_down, 1, bdd);
                int ns_up = w->sys_up + up_dif, ne_up = w->env_up - up_dif;
                if (s < 0 || e < 0) continue;
                for (int nsd = abs(w->sys_down - sst->down_dif); nsd <= w->sys_down + sst->down_dif; nsd += SU) {
                  int ned = nsd;
                  const OMapSub *mns = NULL, *mne = NULL;
                  if (find_bs_sub(sbs, ns_up, nsd, ms->bl_up, ms->bl_down, nss_up, nss_down, &mns) < 0) continue;
                  if (find_bs_sub(ebs, ne_up, ned, neb_up, neb_down, me->st_up, me->st_down, &mne) < 0) continue;
                  int k = find_wave(out, ns_up, nsd, ne_up, ned);
                  if (k < 0) continue;
                  int j1 = nsd, j2 = d1, j3 = w->sys_down, j4 = ms->st_down, j5 = ms->bl_down, j6 = nss_down;
                  int j11 = ned, j22 = d1, j33 = w->env_down, j44 = me->bl_down, j55 = me->st_down, j66 = neb_down;
                  double c1 = oracle_w6js(j1, j2, j3, j4, j5, j6);
                  if (c1 == 0.0) continue;
                  double c11 = oracle_w6js(j11, j22, j33, j44, j55, j66);
                  if (c11 == 0.0) continue;
                  c1 = c1 * c11;
                  if (c1 == 0.0) continue;
                  c1 = c1 * sqrt((1.0 + j11) * (j33 + 1.0) / (1.0 + d1));
                  if (((j66 + j55 + j33 + j22) / 2 + (j5 + j4 + j1 + j2) / 2) % 2 != 0) c1 = -c1;
                  double sign_st = fermion_sign(fst, ms->bl_down);
                  double sign_el = fermion_sign(fel, w->sys_down + me->st_down);
                  double ct = (sign_st * sign_el * c1) * coef * sst->sub[s].mat[0];
                  int lds = in->sub[i].sys_dim, ldk = out->sub[k].sys_dim;
                  gemm('N', 'N', ms->sdim, mne->sdim, me->sdim, ct, WV(in, i, ms->spos, me->spos), lds,
                       ebl->sub[e].mat, me->sdim, 1.0, WV(out, k, mns->spos, mne->spos), ldk);
                  if (d1 == 1)
                    gemm('N', 'T', ms->sdim, me->sdim, mne->sdim, ct, WV(in, k, mns->spos, mne->spos), ldk,
                         ebl->sub[e].mat, me->sdim, 1.0, WV(out, i, ms->spos, me->spos), lds);
                }
              }
            }
          }
        }
      }
    }
  }
}

/* site_vec_site_ndia, get_sys.f90:3761-4124 (live part :3794-3994) */
static void k_site_vec_site_ndia(const OBlock* sst, char fst, const OBlock* est, char fet, const OBasis* sbs,
                                 const OBasis* ebs, const OWave* in, OWave* out, double coef) {
  int up_dif = sst->up_dif, d1 = sst->down_dif;
  if (up_dif != est->up_dif || d1 != est->down_dif) { fprintf(stderr, "oracle: QN error in site_vec_site\n"); abort(); }
  for (int i = 0; i < in->num; ++i) {
    const OWSub* w = &in->sub[i];
    for (int x = 0; x < sbs->num; ++x) {
      if (sbs->sec[x].new_up != w->sys_up || sbs->sec[x].new_down != w->sys_down) continue;
      for (int yy = 0; yy < ebs->num; ++yy) {
        if (ebs->sec[yy].new_up != w->env_up || ebs->sec[yy].new_down != w->env_down) continue;
        for (int m = 0; m < sbs->sec[x].num; ++m) {
          for (int n = 0; n < ebs->sec[yy].num; ++n) {
            const OMapSub* ms = &sbs->sec[x].sub[m];
            const OMapSub* me = &ebs->sec[yy].sub[n];
            for (int sdd1 = -sst->down_dif; sdd1 <= sst->down_dif; sdd1 += SU) {
              int s = find_op(sst, ms->st_up, ms->st_down, 1, sdd1);
              int nss_up = ms->st_up + up_dif, nss_down = ms->st_down + sdd1;
              for (int sdd2 = -est->down_dif; sdd2 <= est->down_dif; sdd2 += SU) {
                int nes_up = me->st_up - up_dif, nes_down = me->st_down - sdd2;
                int e = find_op(est, nes_up, nes_down, 1, sdd2);
                int ns_up = w->sys_up + up_dif, ne_up = w->env_up - up_dif;
                if (s < 0 || e < 0) continue;
                for (int nsd = abs(w->sys_down - sst->down_dif); nsd <= w->sys_down + sst->down_dif; nsd += SU) {
                  int ned = nsd;
                  const OMapSub *mns = NULL, *mne = NULL;
                  if (find_bs_sub(sbs, ns_up, nsd, ANYQ, ms->bl_down, nss_up, nss_down, &mns) < 0) continue;
                  if (find_bs_sub(ebs, ne_up, ned, ANYQ, me->bl_down, nes_up, nes_down, &mne) < 0) continue;
                  int k = find_wave(out, ns_up, nsd, ne_up, ned);
                  if (k < 0) continue;
                  double sign_st = fermion_sign(fst, ms->bl_down);
                  double sign_et = fermion_sign(fet, w->sys_down);
                  double ct = (sign_st * sign_et) * coef;
                  int j1 = nsd, j2 = d1, j3 = w->sys_down, j4 = ms->st_down, j5 = ms->bl_down, j6 = nss_down;
                  int j11 = ned, j22 = d1, j33 = w->env_down, j44 = me->st_down, j55 = me->bl_down, j66 = nes_down;
                  double c1 = oracle_w6js(j1, j2, j3, j4, j5, j6);
                  if (c1 == 0.0) continue;
                  double c11 = oracle_w6js(j11, j22, j33, j44, j55, j66);
                  if (c11 == 0.0) continue;
                  c1 = c1 * c11;
                  if (c1 == 0.0) continue;
                  c1 = c1 * sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + d1));
                  if (((j55 + j44 + j11 + j22) / 2 + (j5 + j4 + j1 + j2) / 2) % 2 != 0) c1 = -c1;
                  c1 = ct * c1 * sst->sub[s].mat[0] * est->sub[e].mat[0];
                  for (int cc = 0; cc < me->sdim; ++cc)
                    for (int r = 0; r < ms->sdim; ++r)
                      *WV(out, k, mns->spos + r, mne->spos + cc) += c1 * *WV(in, i, ms->spos + r, me->spos + cc);
                  if (d1 == 1)
                    for (int cc = 0; cc < me->sdim; ++cc)
                      for (int r = 0; r < ms->sdim; ++r)
                        *WV(out, i, ms->spos + r, me->spos + cc) += c1 * *WV(in, k, mns->spos + r, mne->spos + cc);
                }
              }
            }
          }
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* Lattice tables (dmrg_setup0 / dmrg_setup, dmrg_sub.f90:1-281)       */
/* ------------------------------------------------------------------ */

typedef struct {
  int nsite;         /* sites of the full lattice (arrays are sized nsite+1) */
  int* neib;         /* [(nsite+1)*NEIBT], neib[i*NEIBT + (k-1)] */
  int* neiba;        /* [(nsite+1)^2] */
  int* nposi; int* nposi_lat; int* nposi_dm;
  int* inb1;         /* [(nsite+1)^2], inb1[i*(nsite+1) + ns] */
  int* nleg11;       /* [nsite+1] */
  double jt[NEIBT + 1], jd[NEIBT + 1], jn[NEIBT + 1], esite;
} Lat;

/* Square-lattice neighbour table over nxc columns (dmrg_sub.f90:31-49, 87-143
 * with open = 1 and nly = 1). */
static void lat_square_neib(Lat* L, int nleg, int nxc) {
  int N = L->nsite;
  memset(L->neib, 0, sizeof(int) * (size_t)(N + 1) * NEIBT);
  memset(L->neiba, 0, sizeof(int) * (size_t)(N + 1) * (N + 1));
#define NB(i, k) L->neib[(size_t)(i) * NEIBT + (k) - 1]
#define NBA(i, j) L->neiba[(size_t)(i) * (N + 1) + (j)]
  for (int ix = 1; ix <= nxc; ++ix)
    for (int iy = 1; iy <= nleg; ++iy) {
      int ii = iy + (ix - 1) * nleg;
      int jy = ii - iy + (iy % nleg) + 1;
      if (iy != nleg || nleg != 2) { NB(ii, 1) = jy; NB(jy, 2) = ii; NBA(ii, jy) = 1; NBA(jy, ii) = 2; }
      int jx = ii + nleg;
      if (jx > nleg * nxc) jx = 0;  /* open = 1 */
      if (jx != 0) {
        NB(ii, 3) = jx; NB(jx, 4) = ii; NBA(ii, jx) = 3; NBA(jx, ii) = 4;
        jy = jx - 1; if (iy == 1) jy += nleg;
        if (jy != 0) { NB(ii, 5) = jy; NB(jy, 6) = ii; NBA(ii, jy) = 5; NBA(jy, ii) = 6; }
        jy = jx + 1; if (iy == nleg) jy -= nleg;
        if (jy != 0) { NB(ii, 7) = jy; NB(jy, 8) = ii; NBA(ii, jy) = 7; NBA(jy, ii) = 8; }
      }
    }
  /* dmrg_sub.f90:270-279: directions neibt1+1..neibt with zero couplings are
   * removed; for the square lattice they are never set. */
#undef NB
#undef NBA
}

static void lat_alloc(Lat* L, int nsite) {
  L->nsite = nsite;
  L->neib = (int*)xcalloc((size_t)(nsite + 1) * NEIBT, sizeof(int));
  L->neiba = (int*)xcalloc((size_t)(nsite + 1) * (nsite + 1), sizeof(int));
  L->nposi = (int*)xcalloc((size_t)nsite + 2, sizeof(int));
  L->nposi_lat = (int*)xcalloc((size_t)nsite + 2, sizeof(int));
  L->nposi_dm = (int*)xcalloc((size_t)nsite + 2, sizeof(int));
  L->inb1 = (int*)xcalloc((size_t)(nsite + 1) * (nsite + 1), sizeof(int));
  L->nleg11 = (int*)xcalloc((size_t)nsite + 1, sizeof(int));
}

static void lat_free(Lat* L) {
  free(L->neib); free(L->neiba); free(L->nposi); free(L->nposi_lat); free(L->nposi_dm);
  free(L->inb1); free(L->nleg11);
}

/* dmrg_setup0, dmrg_sub.f90:1-29 */
static void lat_setup0(Lat* L) {
  int N = L->nsite;
  for (int ns = 1; ns <= N - 2; ++ns) {
    int k0 = 0;
    for (int i1 = 1; i1 <= ns; ++i1) {
      L->inb1[(size_t)i1 * (N + 1) + ns] = 10000;
      int k00 = 0;
      for (int j = 1; j <= NEIBT; ++j)
        if (L->neib[(size_t)i1 * NEIBT + j - 1] >= ns) ++k00;
      if (k00 != 0) { ++k0; L->inb1[(size_t)i1 * (N + 1) + ns] = k0; }
    }
    L->nleg11[ns] = k0;
  }
}

/* dmrg_setup(ns, ns21), dmrg_sub.f90:31-281 (square, open = 1) */
static void lat_setup(Lat* L, int nleg, int ns, int ns21) {
  int nn = ns + ns21 + 2;
  int nxc = (nn - 1) / nleg + 1;
  int totsite = nleg * nxc;
  int N = L->nsite;
  memset(L->nposi, 0, sizeof(int) * (size_t)(N + 2));
  memset(L->nposi_lat, 0, sizeof(int) * (size_t)(N + 2));
  memset(L->nposi_dm, 0, sizeof(int) * (size_t)(N + 2));
  for (int i = 1; i <= ns + 1; ++i) { L->nposi[i] = 1; L->nposi_lat[i] = i; L->nposi_dm[i] = i; }
  L->nposi[ns + 1] = 2;
  for (int i = 1; i <= ns21 + 1; ++i) {
    L->nposi[ns + 1 + i] = 3;
    L->nposi_dm[totsite + 1 - i] = i + (ns + 1);
    L->nposi_lat[i + ns + 1] = totsite + 1 - i;
  }
  L->nposi[ns + 1 + ns21 + 1] = 4;
  lat_square_neib(L, nleg, nxc);
}

/* main.f90:35-47 couplings (uniform per direction) and the set-up sequence
 * dmrg_setup0; dmrg_setup(N/2-1, N/2-1); dmrg_setup0 (main.f90:50-52). */
static void lat_init(Lat* L, const oracle_lattice* lat) {
  int nsite = lat->ny * lat->nx_full;
  lat_alloc(L, nsite);
  for (int k = 0; k <= NEIBT; ++k) { L->jt[k] = 0; L->jd[k] = 0; L->jn[k] = 0; }
  for (int k = 1; k <= 4; ++k) { L->jt[k] = lat->jt[0]; L->jd[k] = lat->jd[0]; L->jn[k] = lat->jn[0]; }
  for (int k = 5; k <= 8; ++k) { L->jt[k] = lat->jt[1]; L->jd[k] = lat->jd[1]; L->jn[k] = lat->jn[1]; }
  L->esite = lat->esite;
  lat_setup0(L);
  lat_setup(L, lat->ny, nsite / 2 - 1, nsite / 2 - 1);
  lat_setup0(L);
}

#define NPOSI(L, j) ((L)->nposi[(L)->nposi_dm[(j)]])
#define NEIB(L, i, k) ((L)->neib[(size_t)(i) * NEIBT + (k) - 1])
#define INB1(L, i, ns) ((L)->inb1[(size_t)(i) * ((L)->nsite + 1) + (ns)])

/* ------------------------------------------------------------------ */
/* hmerge (get_sys.f90:4383-4593)                                      */
/* ------------------------------------------------------------------ */

typedef struct { OBlock elec_up, elec_down, spin_sd, num_sn; } OSet;

/* block_pass_info, block.f90:1741-1766: same structure, zero data */
static OBlock pass_info(const OBlock* old) {
  OBlock b; b.len = old->len; b.up_dif = old->up_dif; b.down_dif = old->down_dif; b.num = old->num; b.owns = 1;
  b.sub = (OSub*)xcalloc((size_t)b.num, sizeof(OSub));
  for (int i = 0; i < b.num; ++i) {
    b.sub[i] = old->sub[i];
    b.sub[i].mat = (double*)xcalloc((size_t)old->sub[i].row_dim * old->sub[i].sdim, sizeof(double));
  }
  return b;
}

/* block_add_block_two, block.f90:1782-1814: new += coef1 * block1, matching
 * sub-blocks by (num_up, num_down, down_dif) with a forward-only search. */
static void add_block_two(const OBlock* b1, double coef1, OBlock* nb) {
  int x1 = 0;
  for (int i = 0; i < nb->num; ++i) {
    int idx = -1;
    for (int x = x1; x < b1->num; ++x)
      if (b1->sub[x].num_up == nb->sub[i].num_up && b1->sub[x].num_down == nb->sub[i].num_down &&
          b1->sub[x].down_dif == nb->sub[i].down_dif) { idx = x; x1 = x + 1; break; }
    if (idx < 0) continue;
    size_t n = (size_t)nb->sub[i].row_dim * nb->sub[i].sdim;
    for (size_t q = 0; q < n; ++q) nb->sub[i].mat[q] = nb->sub[i].mat[q] + coef1 * b1->sub[idx].mat[q];
  }
}

/* block_transfer_trans, block.f90:1676-1706 */
static OBlock transfer_trans(const OBlock* in) {
  OBlock b; b.len = in->len; b.up_dif = -in->up_dif; b.down_dif = in->down_dif; b.num = in->num; b.owns = 1;
  b.sub = (OSub*)xcalloc((size_t)b.num, sizeof(OSub));
  for (int i = 0; i < b.num; ++i) {
    const OSub* s = &in->sub[i];
    b.sub[i].num_up = s->num_up + in->up_dif;
    b.sub[i].num_down = s->num_down + s->down_dif;
    b.sub[i].down_dif = -s->down_dif;
    b.sub[i].row_dim = s->sdim;
    b.sub[i].sdim = s->row_dim;
    b.sub[i].mat = (double*)xcalloc((size_t)s->row_dim * s->sdim, sizeof(double));
    for (int c = 0; c < s->sdim; ++c)
      for (int r = 0; r < s->row_dim; ++r) b.sub[i].mat[c + (size_t)r * s->sdim] = s->mat[r + (size_t)c * s->row_dim];
  }
  return b;
}

typedef struct {
  int coup_bs[3][3];  /* 1-based (a, b) */
  int coup_bb[3];
  int* inb;           /* inb[0..nsite+1] */
} Couplings;

/* integer part of hmerge: coup_bs, coup_bb, inb (get_sys.f90:4393-4405, 4477-4511) */
static void hmerge_tables(const Lat* L, int sys_len, int env_len, Couplings* C) {
  memset(C->coup_bs, 0, sizeof(C->coup_bs));
  int dmsite = sys_len + env_len + 2, syssite = sys_len + 1;
  for (int k = 1; k <= NEIBT; ++k) {
    int j = NEIB(L, syssite, k);
    if (NPOSI(L, j) == 3) C->coup_bs[2][1] = 1;
    if (NPOSI(L, j) == 1) C->coup_bs[1][1] = 1;
    j = NEIB(L, L->nposi_lat[dmsite], k);
    if (NPOSI(L, j) == 3) C->coup_bs[2][2] = 1;
    if (NPOSI(L, j) == 1) C->coup_bs[1][2] = 1;
  }
  int k11 = 0;
  for (int I = 1; I <= sys_len; ++I) {
    int k1 = 0;
    C->inb[I] = 0;
    for (int k = 1; k <= NEIBT; ++k)
      if (NPOSI(L, NEIB(L, I, k)) == 3) ++k1;
    C->inb[I] = k1;
    if (L->nleg11[env_len] < L->nleg11[sys_len]) C->inb[I] = 0;
    if (C->inb[I] != 0) ++k11;
  }
  C->coup_bb[1] = k11;
  C->inb[sys_len + 1] = 0;
  C->inb[sys_len + 1 + env_len + 1] = 0;
  for (int I = 1; I <= env_len; ++I) {
    int k1 = 0;
    C->inb[I + sys_len + 1] = 0;
    for (int k = 1; k <= NEIBT; ++k) {
      int J = NEIB(L, L->nposi_lat[I + sys_len + 1], k);
      if (NPOSI(L, J) == 1 && C->inb[J] == 0) ++k1;
    }
    C->inb[I + sys_len + 1] = k1;
    if (C->inb[I + sys_len + 1] != 0) ++k11;
  }
  C->coup_bb[2] = k11 - C->coup_bb[1];
}

typedef struct {
  OBlock ham;
  int nact;
  OSet* sub;     /* sub[1..nact] */
  OSet snc[3];   /* snc[1], snc[2] */
  int nsubs;
  OSet* sub_s;   /* sub_s[1..nsubs] */
} OModel;

static OModel view_model(int nops, const int* hdr, const int* sub, const int64_t* off, const double* data) {
  OModel M; memset(&M, 0, sizeof(M));
  int first = 0;
  const int* h = hdr;
  M.ham = view_op(h, sub + 5 * first, off + first, data);
  first += h[3]; h += 4;
  M.nact = (nops - 1) / 4;
  M.sub = (OSet*)xcalloc((size_t)M.nact + 1, sizeof(OSet));
  for (int i = 1; i <= M.nact; ++i) {
    OBlock* dst[4] = {&M.sub[i].elec_up, &M.sub[i].elec_down, &M.sub[i].spin_sd, &M.sub[i].num_sn};
    for (int q = 0; q < 4; ++q) { *dst[q] = view_op(h, sub + 5 * first, off + first, data); first += h[3]; h += 4; }
  }
  return M;
}

static void free_set(OSet* s) { free_block(&s->elec_up); free_block(&s->elec_down); free_block(&s->spin_sd); free_block(&s->num_sn); }

static void free_model(OModel* M) {
  free_block(&M->ham);
  for (int i = 1; i <= M->nact; ++i) free_set(&M->sub[i]);
  free(M->sub);
  free_set(&M->snc[1]); free_set(&M->snc[2]);
  for (int i = 1; i <= M->nsubs; ++i) free_set(&M->sub_s[i]);
  free(M->sub_s);
}


/* operator part of hmerge: snc and sub_s (get_sys.f90:4407-4475, 4513-4591) */
static void hmerge_ops(const Lat* L, OModel* sys, OModel* env, int sys_len, int env_len, const Couplings* C) {
  for (int ia = 1; ia <= 2; ++ia) {
    if (C->coup_bs[1][ia] == 1) {
      sys->snc[ia].elec_up = pass_info(&sys->sub[1].elec_up);
      sys->snc[ia].spin_sd = pass_info(&sys->sub[1].spin_sd);
      sys->snc[ia].num_sn = pass_info(&sys->sub[1].num_sn);
    }
    if (C->coup_bs[2][ia] == 1) {
      env->snc[ia].elec_up = pass_info(&env->sub[1].elec_up);
      env->snc[ia].elec_down = pass_info(&env->sub[1].elec_down);
      env->snc[ia].spin_sd = pass_info(&env->sub[1].spin_sd);
      env->snc[ia].num_sn = pass_info(&env->sub[1].num_sn);
    }
  }
  for (int ia = 1; ia <= 2; ++ia) {
    int j = sys_len + 1 + (env_len + 1) * (ia - 1);
    int sj = L->nposi_lat[j];
    for (int k1 = 1; k1 <= NEIBT; ++k1) {
      int Ib = NEIB(L, sj, k1);
      if (NPOSI(L, Ib) == 3) {
        int ib1 = INB1(L, L->nposi_dm[Ib] - sys_len - 1, env_len);
        if (C->coup_bs[2][ia] != 0) {
          double jt1 = L->jt[k1], jd1 = L->jd[k1], jn1 = L->jn[k1];
          if (jt1 != 0.0) {
            add_block_two(&env->sub[ib1].elec_up, jt1, &env->snc[ia].elec_up);
            add_block_two(&env->sub[ib1].elec_down, jt1, &env->snc[ia].elec_down);
          }
          if (jd1 != 0.0) add_block_two(&env->sub[ib1].spin_sd, jd1, &env->snc[ia].spin_sd);
          if (jn1 != 0.0) add_block_two(&env->sub[ib1].num_sn, jn1, &env->snc[ia].num_sn);
        }
      }
      if (NPOSI(L, Ib) == 1) {
        int ib1 = INB1(L, L->nposi_dm[Ib], sys_len);
        if (C->coup_bs[1][ia] == 1) {
          double jt1 = L->jt[k1], jd1 = L->jd[k1], jn1 = L->jn[k1];
          if (jt1 != 0.0) add_block_two(&sys->sub[ib1].elec_up, jt1, &sys->snc[ia].elec_up);
          if (jd1 != 0.0) add_block_two(&sys->sub[ib1].spin_sd, jd1, &sys->snc[ia].spin_sd);
          if (jn1 != 0.0) add_block_two(&sys->sub[ib1].num_sn, jn1, &sys->snc[ia].num_sn);
        }
      }
    }
  }
  env->nsubs = C->coup_bb[1];
  env->sub_s = (OSet*)xcalloc((size_t)env->nsubs + 1, sizeof(OSet));
  for (int i = 1; i <= C->coup_bb[1]; ++i) {
    env->sub_s[i].elec_up = pass_info(&env->sub[1].elec_up);
    env->sub_s[i].elec_down = pass_info(&env->sub[1].elec_down);
    env->sub_s[i].spin_sd = pass_info(&env->sub[1].spin_sd);
    env->sub_s[i].num_sn = pass_info(&env->sub[1].num_sn);
  }
  sys->nsubs = C->coup_bb[2];
  sys->sub_s = (OSet*)xcalloc((size_t)sys->nsubs + 1, sizeof(OSet));
  for (int i = 1; i <= C->coup_bb[2]; ++i) {
    sys->sub_s[i].elec_up = pass_info(&sys->sub[1].elec_up);
    sys->sub_s[i].spin_sd = pass_info(&sys->sub[1].spin_sd);
    sys->sub_s[i].num_sn = pass_info(&sys->sub[1].num_sn);
  }
  int jc = 0;
  for (int I = 1; I <= sys_len + env_len + 1; ++I) {
    if (C->inb[I] == 0) continue;
    int Ia = L->nposi_lat[I];
    ++jc;
    for (int k1 = 1; k1 <= NEIBT; ++k1) {
      int Ib = NEIB(L, Ia, k1);
      double jt1 = L->jt[k1], jd1 = L->jd[k1], jn1 = L->jn[k1];
      if (NPOSI(L, Ib) == 3 && I <= sys_len) {
        int ib1 = INB1(L, L->nposi_dm[Ib] - sys_len - 1, env_len);
        if (jt1 != 0.0) {
          add_block_two(&env->sub[ib1].elec_up, jt1, &env->sub_s[jc].elec_up);
          add_block_two(&env->sub[ib1].elec_down, jt1, &env->sub_s[jc].elec_down);
        }
        if (jd1 != 0.0) add_block_two(&env->sub[ib1].spin_sd, jd1, &env->sub_s[jc].spin_sd);
        if (jn1 != 0.0) add_block_two(&env->sub[ib1].num_sn, jn1, &env->sub_s[jc].num_sn);
      }
      if (NPOSI(L, Ib) == 1 && I > sys_len + 1) {
        if (C->inb[L->nposi_dm[Ib]] == 0) {
          int ib1 = INB1(L, Ib, sys_len);
          int jc1 = jc - C->coup_bb[1];
          if (jt1 != 0.0) add_block_two(&sys->sub[ib1].elec_up, jt1, &sys->sub_s[jc1].elec_up);
          if (jd1 != 0.0) add_block_two(&sys->sub[ib1].spin_sd, jd1, &sys->sub_s[jc1].spin_sd);
          if (jn1 != 0.0) add_block_two(&sys->sub[ib1].num_sn, jn1, &sys->sub_s[jc1].num_sn);
        }
      }
    }
  }
}

/* model_wave, get_sys.f90:4127-4240 (add_op = 1) */
static void model_wave(const Lat* L, const SiteOps* so, OModel* sys, OModel* env, int sys_len, int env_len,
                       const Couplings* C, const OBasis* sbs, const OBasis* ebs, const OWave* in, OWave* out) {
  for (int i = 0; i < out->num; ++i)
    memset(out->sub[i].vec, 0, sizeof(double) * (size_t)out->sub[i].sys_dim * out->sub[i].env_dim);
  const double cone = 1.0;
  k_block_vec_dia(&sys->ham, sbs, in, out, cone);
  k_vec_block_dia(&env->ham, ebs, in, out, cone);
  int is = L->nposi_lat[sys_len + 1], js = L->nposi_lat[sys_len + env_len + 2];
  if (L->esite != 0.0) k_site_vec_site_dia(&so->num_elec, &so->st_si, sbs, ebs, in, out, L->esite);
  if (L->esite != 0.0) k_site_vec_site_dia(&so->st_si, &so->num_elec, sbs, ebs, in, out, L->esite);
  int ii = L->neiba[(size_t)is * (L->nsite + 1) + js];
  if (ii != 0) {
    double jd1 = L->jd[ii], coefx = L->jt[ii], jn1 = L->jn[ii];
    if (coefx != 0.0) k_site_vec_site_ndia(&so->elec_up, 'F', &so->elec_up, 'F', sbs, ebs, in, out, coefx);
    if (jd1 != 0.0) k_site_vec_site_ndia(&so->st_sd, 'B', &so->st_sd, 'B', sbs, ebs, in, out, jd1);
    k_site_vec_site_dia(&so->num_elec, &so->num_elec, sbs, ebs, in, out, jn1);
  }
  if (C->coup_bs[1][1] == 1) {
    k_block_site_vec_ndia(&sys->snc[1].elec_up, 'F', &so->elec_up, 'F', sbs, in, out, cone);
    k_block_site_vec_ndia(&sys->snc[1].spin_sd, 'B', &so->st_sd, 'B', sbs, in, out, cone);
    k_block_site_vec_dia(&sys->snc[1].num_sn, &so->num_elec, sbs, in, out, cone);
  }
  if (C->coup_bs[1][2] == 1) {
    k_block_vec_site_ndia(&sys->snc[2].elec_up, 'F', &so->elec_up, 'F', sbs, ebs, in, out, cone);
    k_block_vec_site_ndia(&sys->snc[2].spin_sd, 'B', &so->st_sd, 'B', sbs, ebs, in, out, cone);
    k_block_vec_site_dia(&sys->snc[2].num_sn, &so->num_elec, sbs, ebs, in, out, cone);
  }
  if (C->coup_bs[2][1] == 1) {
    OBlock t1 = transfer_trans(&env->snc[1].elec_down);
    k_site_vec_block_ndia(&so->elec_up, 'F', &t1, 'F', sbs, ebs, in, out, cone);
    OBlock t = transfer_trans(&env->snc[1].spin_sd);
    k_site_vec_block_ndia(&so->st_sd, 'B', &t, 'B', sbs, ebs, in, out, cone);
    k_site_vec_block_dia(&so->num_elec, &env->snc[1].num_sn, sbs, ebs, in, out, cone);
    free_block(&t); free_block(&t1);
  }
  if (C->coup_bs[2][2] == 1) {
    k_vec_block_site_ndia(&env->snc[2].elec_up, 'F', &so->elec_up, 'F', ebs, in, out, cone);
    k_vec_block_site_ndia(&env->snc[2].spin_sd, 'B', &so->st_sd, 'B', ebs, in, out, cone);
    k_vec_block_site_dia(&env->snc[2].num_sn, &so->num_elec, ebs, in, out, cone);
  }
  int jc = 0;
  for (int i = 1; i <= sys_len + 1 + env_len; ++i) {
    if (C->inb[i] == 0) continue;
    ++jc;
    if (i <= sys_len) {
      int i1 = INB1(L, i, sys_len);
      OBlock t1 = transfer_trans(&env->sub_s[jc].elec_down);
      k_block_vec_block_ndia(&sys->sub[i1].elec_up, 'F', &t1, 'F', sbs, ebs, in, out, cone);
      OBlock t = transfer_trans(&env->sub_s[jc].spin_sd);
      k_block_vec_block_ndia(&sys->sub[i1].spin_sd, 'B', &t, 'B', sbs, ebs, in, out, cone);
      k_block_vec_block_dia(&sys->sub[i1].num_sn, &env->sub_s[jc].num_sn, sbs, ebs, in, out, cone);
      free_block(&t); free_block(&t1);
    } else {
      int i1 = INB1(L, i - sys_len - 1, env_len);
      int jc1 = jc - C->coup_bb[1];
      OBlock t1 = transfer_trans(&env->sub[i1].elec_down);
      k_block_vec_block_ndia(&sys->sub_s[jc1].elec_up, 'F', &t1, 'F', sbs, ebs, in, out, cone);
      OBlock t = transfer_trans(&env->sub[i1].spin_sd);
      k_block_vec_block_ndia(&sys->sub_s[jc1].spin_sd, 'B', &t, 'B', sbs, ebs, in, out, cone);
      k_block_vec_block_dia(&sys->sub_s[jc1].num_sn, &env->sub[i1].num_sn, sbs, ebs, in, out, cone);
      free_block(&t); free_block(&t1);
    }
  }
}

/* ------------------------------------------------------------------ */
/* Exported entry points                                               */
/* ------------------------------------------------------------------ */

int oracle_model_wave(const oracle_lattice* lat, int sys_len, int env_len, int sys_nops, const int* sys_hdr,
                      const int* sys_sub, const int64_t* sys_off, const double* sys_data, int env_nops,
                      const int* env_hdr, const int* env_sub, const int64_t* env_off, const double* env_data,
                      int sys_nsec, const int* sys_bsec, const int* sys_bsub, int env_nsec, const int* env_bsec,
                      const int* env_bsub, int nw, const int* wsec, const double* x, double* y) {
  Lat L; lat_init(&L, lat);
  if (sys_len + env_len + 2 > L.nsite) { lat_free(&L); return -1; }
  lat_setup(&L, lat->ny, sys_len, env_len);
  SiteOps so; make_site_ops(&so, lat->up_bias);
  OModel sys = view_model(sys_nops, sys_hdr, sys_sub, sys_off, sys_data);
  OModel env = view_model(env_nops, env_hdr, env_sub, env_off, env_data);
  OBasis sbs = view_basis(sys_nsec, sys_bsec, sys_bsub);
  OBasis ebs = view_basis(env_nsec, env_bsec, env_bsub);
  OWave in = view_wave(nw, wsec, (double*)x);
  OWave out = view_wave(nw, wsec, y);
  Couplings C; C.inb = (int*)xcalloc((size_t)L.nsite + 6, sizeof(int));
  hmerge_tables(&L, sys_len, env_len, &C);
  hmerge_ops(&L, &sys, &env, sys_len, env_len, &C);
  model_wave(&L, &so, &sys, &env, sys_len, env_len, &C, &sbs, &ebs, &in, &out);
  free(C.inb);
  free_model(&sys); free_model(&env);
  free(sbs.sec); free(ebs.sec); free(in.sub); free(out.sub);
  lat_free(&L);
  return 0;
}

int oracle_lattice_tables(const oracle_lattice* lat, int ns, int ns21, int* nposi, int* nposi_lat, int* nposi_dm,
                          int* neib, int* inb1_col_sys, int* inb1_col_env, int* nleg11_sys, int* nleg11_env) {
  Lat L; lat_init(&L, lat);
  if (ns + ns21 + 2 > L.nsite) { lat_free(&L); return -1; }
  lat_setup(&L, lat->ny, ns, ns21);
  int N = L.nsite;
  memcpy(nposi, L.nposi, sizeof(int) * (size_t)(N + 1));
  memcpy(nposi_lat, L.nposi_lat, sizeof(int) * (size_t)(N + 1));
  memcpy(nposi_dm, L.nposi_dm, sizeof(int) * (size_t)(N + 1));
  memcpy(neib, L.neib, sizeof(int) * (size_t)(N + 1) * NEIBT);
  for (int i = 0; i <= N; ++i) {
    inb1_col_sys[i] = INB1(&L, i, ns);
    inb1_col_env[i] = INB1(&L, i, ns21);
  }
  *nleg11_sys = L.nleg11[ns];
  *nleg11_env = L.nleg11[ns21];
  lat_free(&L);
  return 0;
}

int oracle_hmerge_tables(const oracle_lattice* lat, int sys_len, int env_len, int* coup_bs, int* coup_bb, int* inb) {
  Lat L; lat_init(&L, lat);
  if (sys_len + env_len + 2 > L.nsite) { lat_free(&L); return -1; }
  lat_setup(&L, lat->ny, sys_len, env_len);
  Couplings C; C.inb = (int*)xcalloc((size_t)L.nsite + 6, sizeof(int));
  hmerge_tables(&L, sys_len, env_len, &C);
  coup_bs[0] = C.coup_bs[1][1]; coup_bs[1] = C.coup_bs[1][2];
  coup_bs[2] = C.coup_bs[2][1]; coup_bs[3] = C.coup_bs[2][2];
  coup_bb[0] = C.coup_bb[1]; coup_bb[1] = C.coup_bb[2];
  for (int i = 0; i <= sys_len + env_len + 2; ++i) inb[i] = C.inb[i];
  free(C.inb);
  lat_free(&L);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Fusion bases                                                        */
/* ------------------------------------------------------------------ */

/* QuickSort_Integer (block.f90:212-239): exchange sort, 2S descending, then N
 * descending among equal 2S. */
static void sort_qn(int* a, int num) {
  for (int i = 0; i < num; ++i)
    for (int x = i + 1; x < num; ++x)
      if (a[2 * i + 1] < a[2 * x + 1]) {
        int t0 = a[2 * i], t1 = a[2 * i + 1];
        a[2 * i] = a[2 * x]; a[2 * i + 1] = a[2 * x + 1]; a[2 * x] = t0; a[2 * x + 1] = t1;
      }
  for (int i = 0; i < num; ++i)
    for (int x = i + 1; x < num; ++x)
      if (a[2 * i + 1] == a[2 * x + 1] && a[2 * i] < a[2 * x]) {
        int t0 = a[2 * i], t1 = a[2 * i + 1];
        a[2 * i] = a[2 * x]; a[2 * i + 1] = a[2 * x + 1]; a[2 * x] = t0; a[2 * x + 1] = t1;
      }
}

/* Get_Unique_QN (block.f90:165-198) */
static int unique_qn(const int* in, int* out, int num) {
  int cnt = 0, start = 0, sublen = 0, flag = 0;
  for (int i = 0; i < num; ++i) {
    for (int j = start; j < num; ++j)
      if (in[2 * j] == in[2 * i] && in[2 * j + 1] == in[2 * i + 1]) { flag = 1; ++sublen; }
    if (flag) {
      start += sublen;
      out[2 * cnt] = in[2 * i]; out[2 * cnt + 1] = in[2 * i + 1];
      ++cnt; sublen = 0; flag = 0;
    }
  }
  return cnt;
}

/* Get_basis, block.f90:9-104 */
int oracle_get_basis(int blen, int nb, const int* blk, int nsite, const int* site, int* sec, int* sub, int* nsub_total) {
  (void)blen;
  int cap = nb * nsite * 3 + 1;
  int* pre = (int*)xcalloc((size_t)cap * 2, sizeof(int));
  int* res = (int*)xcalloc((size_t)cap * 2, sizeof(int));
  int ij = 0;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nsite; ++j) {
      int s1 = abs(blk[3 * i + 1] - site[3 * j + 1]), s2 = blk[3 * i + 1] + site[3 * j + 1];
      for (int ss = s1; ss <= s2; ss += SU) { pre[2 * ij] = blk[3 * i] + site[3 * j]; pre[2 * ij + 1] = ss; ++ij; }
    }
  sort_qn(pre, ij);
  int nsec = unique_qn(pre, res, ij);
  int ns = 0;
  for (int i = 0; i < nsec; ++i) {
    int up = res[2 * i], down = res[2 * i + 1], dim = 0, cnt = 0;
    for (int x = 0; x < nb; ++x)
      for (int y = 0; y < nsite; ++y) {
        if (blk[3 * x] + site[3 * y] != up) continue;
        int s1 = abs(blk[3 * x + 1] - site[3 * y + 1]), s2 = blk[3 * x + 1] + site[3 * y + 1];
        if (down >= s1 && down <= s2) {
          int* q = sub + 6 * ns;
          q[0] = blk[3 * x]; q[1] = blk[3 * x + 1]; q[2] = site[3 * y]; q[3] = site[3 * y + 1];
          q[4] = dim; q[5] = blk[3 * x + 2] * site[3 * y + 2];
          dim += q[5]; ++ns; ++cnt;
        }
      }
    sec[4 * i] = up; sec[4 * i + 1] = down; sec[4 * i + 2] = dim; sec[4 * i + 3] = cnt;
  }
  *nsub_total = ns;
  free(pre); free(res);
  return nsec;
}

/* Get_super_basis, block.f90:110-161 */
int oracle_get_super_basis(int nsys, const int* sys_sec, int nenv, const int* env_sec, int total_up, int total_down,
                           int* wsec) {
  int n = 0;
  for (int x = 0; x < nsys; ++x)
    for (int y = 0; y < nenv; ++y) {
      if (env_sec[4 * y] != total_up - sys_sec[4 * x]) continue;
      int s1 = abs(env_sec[4 * y + 1] - sys_sec[4 * x + 1]), s2 = env_sec[4 * y + 1] + sys_sec[4 * x + 1];
      if (total_down >= s1 && total_down <= s2) {
        int* w = wsec + 6 * n;
        w[0] = sys_sec[4 * x]; w[1] = sys_sec[4 * x + 1]; w[2] = env_sec[4 * y]; w[3] = env_sec[4 * y + 1];
        w[4] = sys_sec[4 * x + 2]; w[5] = env_sec[4 * y + 2];
        ++n;
      }
    }
  return n;
}

/* ================================================================== */
/* Block update: density matrix, truncation, enlargement, projection   */
/* (den_mat.f90, get_sys.f90:4-762, block.f90:382-1503).  Dense        */
/* restatement: enlarged operators are formed explicitly, as in the    */
/* reference, then projected.                                          */
/* ================================================================== */

struct oracle_opset { int nops; OBlock* op; };

static OBlock ob_alloc(int len, int up_dif, int down_dif, int num) {
  OBlock b; b.len = len; b.up_dif = up_dif; b.down_dif = down_dif; b.num = num; b.owns = 1;
  b.sub = (OSub*)xcalloc((size_t)(num > 0 ? num : 1), sizeof(OSub));
  return b;
}
static void ob_set(OBlock* b, int i, int up, int down, int dd, int row, int col) {
  b->sub[i].num_up = up; b->sub[i].num_down = down; b->sub[i].down_dif = dd;
  b->sub[i].row_dim = row; b->sub[i].sdim = col;
  b->sub[i].mat = (double*)xcalloc((size_t)row * col + 1, sizeof(double));
}
/* block_transfer, block.f90:1710-1739 */
static OBlock ob_copy(const OBlock* in) {
  OBlock b = ob_alloc(in->len, in->up_dif, in->down_dif, in->num);
  for (int i = 0; i < in->num; ++i) {
    const OSub* s = &in->sub[i];
    ob_set(&b, i, s->num_up, s->num_down, s->down_dif, s->row_dim, s->sdim);
    memcpy(b.sub[i].mat, s->mat, sizeof(double) * (size_t)s->row_dim * s->sdim);
  }
  return b;
}
/* mat(r0:r0+h, c0:c0+w) += coef * src(h x w, leading dim lds), or src^T */
static void add_rect(OSub* d, int r0, int c0, int h, int w, double coef, const double* src, int lds, int trans) {
  for (int c = 0; c < w; ++c)
    for (int r = 0; r < h; ++r) {
      double v = trans ? src[c + (size_t)r * lds] : src[r + (size_t)c * lds];
      double* t = &d->mat[(size_t)(r0 + r) + (size_t)(c0 + c) * d->row_dim];
      *t = *t + coef * v;
    }
}

/* update_block_dia, block.f90:382-432 */
static OBlock o_update_block_dia(const OBlock* block, const OBasis* bs, int blen) {
  OBlock nb = ob_alloc(blen, 0, 0, bs->num);
  for (int i = 0; i < bs->num; ++i) ob_set(&nb, i, bs->sec[i].new_up, bs->sec[i].new_down, 0, bs->sec[i].dim, bs->sec[i].dim);
  for (int k = 0; k < nb.num; ++k)
    for (int x = 0; x < bs->sec[k].num; ++x) {
      const OMapSub* m = &bs->sec[k].sub[x];
      for (int i = 0; i < block->num; ++i)
        if (block->sub[i].num_up == m->bl_up && block->sub[i].num_down == m->bl_down) {
          double coef1 = sqrt((1.0 + nb.sub[k].num_down) / (1.0 + block->sub[i].num_down));
          add_rect(&nb.sub[k], m->spos, m->spos, m->sdim, m->sdim, coef1, block->sub[i].mat, block->sub[i].row_dim, 0);
          break;
        }
    }
  return nb;
}

/* update_site_dia, block.f90:437-498 */
static OBlock o_update_site_dia(const OBlock* site, const OBasis* bs, int blen) {
  OBlock nb = ob_alloc(blen, 0, 0, bs->num);
  for (int i = 0; i < bs->num; ++i) ob_set(&nb, i, bs->sec[i].new_up, bs->sec[i].new_down, 0, bs->sec[i].dim, bs->sec[i].dim);
  for (int k = 0; k < nb.num; ++k)
    for (int x = 0; x < bs->sec[k].num; ++x) {
      const OMapSub* m = &bs->sec[k].sub[x];
      for (int i = 0; i < site->num; ++i)
        if (site->sub[i].num_up == m->st_up && site->sub[i].num_down == m->st_down) {
          double coef1 = sqrt((1.0 + nb.sub[k].num_down) / (1.0 + site->sub[i].num_down));
          for (int y = 0; y < m->sdim; ++y) {
            double* t = &nb.sub[k].mat[(size_t)(m->spos + y) * (1 + (size_t)nb.sub[k].row_dim)];
            *t = *t + 1.0 * site->sub[i].mat[0] * coef1;
          }
          break;
        }
    }
  return nb;
}

/* steps <1>-<2> shared by the non-diagonal updates */
static OBlock ndia_structure(const OBasis* bs, int blen, int up_dif, int dd1) {
  int num = 0;
  for (int x = 0; x < bs->num; ++x)
    for (int dd = -dd1; dd <= dd1; dd += SU)
      if (find_sec(bs, bs->sec[x].new_up + up_dif, bs->sec[x].new_down + dd) >= 0) ++num;
  OBlock nb = ob_alloc(blen, up_dif, dd1, num);
  int id = 0;
  for (int x = 0; x < bs->num; ++x)
    for (int dd = -dd1; dd <= dd1; dd += SU) {
      int y = find_sec(bs, bs->sec[x].new_up + up_dif, bs->sec[x].new_down + dd);
      if (y >= 0) ob_set(&nb, id++, bs->sec[x].new_up, bs->sec[x].new_down, dd, bs->sec[y].dim, bs->sec[x].dim);
    }
  return nb;
}

/* update_block_ndia_sys (side 0, block.f90:503-715) / _env (side 1, :720-920) */
static OBlock o_update_block_ndia(const OBlock* block, const OBasis* bs, int blen, char flag, int side) {
  int up_dif = block->up_dif, dd1 = block->down_dif;
  OBlock nb = ndia_structure(bs, blen, up_dif, dd1);
  for (int i = 0; i < nb.num; ++i) {
    int rhs_up = nb.sub[i].num_up, rhs_down = nb.sub[i].num_down;
    int rid = find_sec(bs, rhs_up, rhs_down);
    int lhs_up = rhs_up + up_dif, lhs_down = rhs_down + nb.sub[i].down_dif;
    int lid = find_sec(bs, lhs_up, lhs_down);
    if (rid < 0 || lid < 0) continue;
    for (int x = 0; x < bs->sec[rid].num; ++x) {
      const OMapSub* r = &bs->sec[rid].sub[x];
      for (int sdd = -dd1; sdd <= dd1; sdd += SU) {
        int bid = -1, lsu = 0, lsd = 0;
        for (int y = 0; y < block->num; ++y)
          if (block->sub[y].num_up == r->bl_up && block->sub[y].num_down == r->bl_down && block->sub[y].down_dif == sdd) {
            bid = y; lsu = r->bl_up + up_dif; lsd = r->bl_down + sdd; break;
          }
        if (bid < 0) continue;
        const OMapSub* l = NULL;
        for (int y = 0; y < bs->sec[lid].num; ++y) {
          const OMapSub* q = &bs->sec[lid].sub[y];
          if (q->bl_up == lsu && q->bl_down == lsd && q->st_down == r->st_down) { l = q; break; }
        }
        if (!l) continue;
        int j1 = lhs_down, j2 = dd1, j3 = rhs_down, j4 = r->bl_down, j5 = r->st_down, j6 = lsd;
        double coef = oracle_w6js(j1, j2, j3, j4, j5, j6);
        if (side == 0) {
          coef = coef * sqrt((1.0 + j1) * (1.0 + j3));
          coef = coef * neg1pow((j6 + j5 + j3 + j2) / 2);
        } else {
          double signs = 1.0;
          if (flag == 'F') signs = (j5 % 2 == 0) ? 1.0 : -1.0;
          coef = coef * neg1pow((j6 + j5 + j3 + j2) / 2) * sqrt((1.0 + j1) * (1.0 + j3));
          if (coef != 0.0) coef = coef * signs;
        }
        if (coef != 0.0)
          add_rect(&nb.sub[i], l->spos, r->spos, l->sdim, r->sdim, coef, block->sub[bid].mat, block->sub[bid].row_dim, 0);
      }
    }
  }
  return nb;
}

/* update_site_ndia_sys (side 0, block.f90:925-1112) / _env (side 1, :1113-1305) */
static OBlock o_update_site_ndia(const OBlock* site, const OBasis* bs, int blen, char flag, int side) {
  int up_dif = site->up_dif, dd1 = site->down_dif;
  OBlock nb = ndia_structure(bs, blen, up_dif, dd1);
  for (int i = 0; i < nb.num; ++i) {
    int rhs_up = nb.sub[i].num_up, rhs_down = nb.sub[i].num_down;
    int rid = find_sec(bs, rhs_up, rhs_down);
    int lhs_up = rhs_up + up_dif, lhs_down = rhs_down + nb.sub[i].down_dif;
    int lid = find_sec(bs, lhs_up, lhs_down);
    if (rid < 0 || lid < 0) continue;
    for (int x = 0; x < bs->sec[rid].num; ++x) {
      const OMapSub* r = &bs->sec[rid].sub[x];
      for (int sdd = -dd1; sdd <= dd1; sdd += SU) {
        int sid = -1, lsu = 0, lsd = 0;
        for (int y = 0; y < site->num; ++y)
          if (site->sub[y].num_up == r->st_up && site->sub[y].num_down == r->st_down && site->sub[y].down_dif == sdd) {
            sid = y; lsu = r->st_up + up_dif; lsd = r->st_down + sdd; break;
          }
        if (sid < 0) continue;
        const OMapSub* l = NULL;
        for (int y = 0; y < bs->sec[lid].num; ++y) {
          const OMapSub* q = &bs->sec[lid].sub[y];
          if (q->st_up == lsu && q->st_down == lsd && q->bl_down == r->bl_down) { l = q; break; }
        }
        if (!l) continue;
        int j1 = lhs_down, j2 = dd1, j3 = rhs_down, j4 = r->st_down, j5 = r->bl_down, j6 = lsd;
        double coef = oracle_w6js(j1, j2, j3, j4, j5, j6);
        if (side == 0) {
          double signs = 1.0;
          if (flag == 'F') signs = (j5 % 2 == 0) ? 1.0 : -1.0;
          coef = coef * neg1pow((j5 + j4 + j1 + j2) / 2) * sqrt((1.0 + j1) * (1.0 + j3));
          if (coef != 0.0) coef = coef * signs;
        } else {
          coef = coef * sqrt((1.0 + j1) * (1.0 + j3)) * neg1pow((j5 + j4 + j1 + j2) / 2);
        }
        if (coef != 0.0)
          for (int y = 0; y < r->sdim; ++y) {
            double* t = &nb.sub[i].mat[(size_t)(l->spos + y) + (size_t)(r->spos + y) * nb.sub[i].row_dim];
            *t = *t + coef * 1.0 * site->sub[sid].mat[0];
          }
      }
    }
  }
  return nb;
}

/* Get_hamiltonian_ndia, get_sys.f90:583-760 */
static void o_ham_ndia(const OBlock* block, const OBlock* site, const OBasis* bs, double coef, OBlock* ham,
                       char fsign, char flags) {
  int up_dif = block->up_dif, dd1 = block->down_dif;
  for (int i = 0; i < ham->num; ++i) {
    int rhs_up = ham->sub[i].num_up, rhs_down = ham->sub[i].num_down;
    int rid = find_sec(bs, rhs_up, rhs_down);
    int lid = find_sec(bs, rhs_up, rhs_down + ham->sub[i].down_dif);
    if (rid < 0 || lid < 0) continue;
    for (int x = 0; x < bs->sec[rid].num; ++x) {
      const OMapSub* r = &bs->sec[rid].sub[x];
      for (int sdd = -dd1; sdd <= dd1; sdd += SU) {
        int bid = -1, lsu = 0, lsd = 0;
        for (int y = 0; y < block->num; ++y)
          if (block->sub[y].num_up == r->bl_up && block->sub[y].num_down == r->bl_down && block->sub[y].down_dif == sdd) {
            bid = y; lsu = r->bl_up + up_dif; lsd = r->bl_down + sdd; break;
          }
        if (bid < 0) continue;
        int lo = abs(r->st_down - site->down_dif), hi = r->st_down + site->down_dif;
        for (int nsd = lo; nsd <= hi; nsd += SU) {
          int st_dd = r->st_down - nsd, nsu = r->st_up - site->up_dif, sid = -1;
          for (int y = 0; y < site->num; ++y)
            if (site->sub[y].num_up == nsu && site->sub[y].num_down == nsd && site->sub[y].down_dif == st_dd) { sid = y; break; }
          const OMapSub* l = NULL;
          for (int y = 0; y < bs->sec[lid].num; ++y) {
            const OMapSub* q = &bs->sec[lid].sub[y];
            if (q->bl_up == lsu && q->bl_down == lsd && q->st_down == nsd) { l = q; break; }
          }
          if (!l || sid < 0) continue;
          int j1 = lsd, j2 = r->bl_down, j3 = dd1, j4 = r->st_down, j5 = nsd, j6 = rhs_down;
          double coef1 = oracle_w6js(j1, j2, j3, j4, j5, j6);
          coef1 = coef1 * sqrt((1.0 + j6) / (1.0 + j3)) * neg1pow((j2 + j5 + j3 + j6) / 2) * site->sub[sid].mat[0];
          double sign = 1.0;
          if (flags == 'S' && fsign == 'F') {
            if (r->bl_up % 2 == 1) sign = -1.0;
            if (r->bl_down % 2 == 1) sign = -1.0;
          }
          if (coef1 != 0.0) {
            coef1 = coef1 * coef * sign;
            add_rect(&ham->sub[i], l->spos, r->spos, l->sdim, r->sdim, coef1, block->sub[bid].mat, block->sub[bid].row_dim, 0);
            if (site->down_dif == 1)
              add_rect(&ham->sub[i], r->spos, l->spos, r->sdim, l->sdim, coef1, block->sub[bid].mat, block->sub[bid].row_dim, 1);
          }
        }
      }
    }
  }
}

/* Get_hamiltonian_dia, get_sys.f90:506-577 ('N','N', FlagConjg false) */
static void o_ham_dia(const OBlock* o1, const OBlock* o2, double coef, OBlock* ham) {
  for (int i = 0; i < ham->num; ++i) {
    int up = ham->sub[i].num_up, down = ham->sub[i].num_down, sdim = ham->sub[i].sdim;
    int a = find_op(o1, up, down, 0, 0), b = find_op(o2, up, down, 0, 0);
    if (a < 0 || b < 0) continue;
    double* m = (double*)xcalloc((size_t)sdim * sdim + 1, sizeof(double));
    double coef1 = coef / sqrt(1.0 + down);
    gemm('N', 'N', sdim, sdim, sdim, coef1, o1->sub[a].mat, sdim, o2->sub[b].mat, sdim, 0.0, m, sdim);
    for (size_t q = 0; q < (size_t)sdim * sdim; ++q) ham->sub[i].mat[q] = ham->sub[i].mat[q] + m[q];
    free(m);
  }
}

/* block_add_block_two, block.f90:1782-1814 */
static void o_add_two(const OBlock* b1, double coef1, OBlock* nb) { add_block_two(b1, coef1, nb); }

/* update_trun_dia, block.f90:1311-1358 */
static OBlock o_update_trun_dia(const OBlock* ori, const OBlock* trun) {
  OBlock eff = ob_alloc(ori->len, ori->up_dif, ori->down_dif, trun->num);
  for (int i = 0; i < trun->num; ++i)
    ob_set(&eff, i, trun->sub[i].num_up, trun->sub[i].num_down, 0, trun->sub[i].sdim, trun->sub[i].sdim);
  for (int i = 0; i < eff.num; ++i) {
    int ed = eff.sub[i].sdim;
    for (int x = 0; x < ori->num; ++x)
      if (ori->sub[x].num_up == eff.sub[i].num_up && ori->sub[x].num_down == eff.sub[i].num_down) {
        int od = ori->sub[x].row_dim;
        double* mid = (double*)xcalloc((size_t)od * ed + 1, sizeof(double));
        gemm('N', 'N', od, ed, od, 1.0, ori->sub[x].mat, od, trun->sub[i].mat, od, 0.0, mid, od);
        gemm('T', 'N', ed, ed, od, 1.0, trun->sub[i].mat, od, mid, od, 0.0, eff.sub[i].mat, ed);
        free(mid);
        break;
      }
  }
  return eff;
}

static int find_trun(const OBlock* t, int up, int down) {
  for (int y = 0; y < t->num; ++y)
    if (t->sub[y].num_up == up && t->sub[y].num_down == down) return y;
  return -1;
}

/* update_trun_ndia, block.f90:1360-1503 */
static OBlock o_update_trun_ndia(const OBlock* ori, const OBlock* trun) {
  int up_dif = ori->up_dif, dd1 = ori->down_dif, num = 0;
  for (int x = 0; x < trun->num; ++x)
    for (int dd = -dd1; dd <= dd1; dd += SU)
      if (find_trun(trun, trun->sub[x].num_up + up_dif, trun->sub[x].num_down + dd) >= 0) ++num;
  OBlock eff = ob_alloc(ori->len, up_dif, dd1, num);
  int id = 0;
  for (int x = 0; x < trun->num; ++x)
    for (int dd = -dd1; dd <= dd1; dd += SU) {
      int y = find_trun(trun, trun->sub[x].num_up + up_dif, trun->sub[x].num_down + dd);
      if (y >= 0) ob_set(&eff, id++, trun->sub[x].num_up, trun->sub[x].num_down, dd, trun->sub[y].sdim, trun->sub[x].sdim);
    }
  for (int i = 0; i < eff.num; ++i) {
    int ru = eff.sub[i].num_up, rd = eff.sub[i].num_down, dd = eff.sub[i].down_dif;
    int rid = find_trun(trun, ru, rd), lid = find_trun(trun, ru + up_dif, rd + dd), oid = -1;
    for (int x = 0; x < ori->num; ++x)
      if (ori->sub[x].num_up == ru && ori->sub[x].num_down == rd && ori->sub[x].down_dif == dd) { oid = x; break; }
    if (rid < 0 || lid < 0 || oid < 0) continue;
    int orow = ori->sub[oid].row_dim, ocol = ori->sub[oid].sdim;
    int rdim = trun->sub[rid].sdim, ldim = trun->sub[lid].sdim;
    double* mid = (double*)xcalloc((size_t)orow * rdim + 1, sizeof(double));
    gemm('N', 'N', orow, rdim, ocol, 1.0, ori->sub[oid].mat, orow, trun->sub[rid].mat, ocol, 0.0, mid, orow);
    gemm('T', 'N', ldim, rdim, orow, 1.0, trun->sub[lid].mat, orow, mid, orow, 0.0, eff.sub[i].mat, ldim);
    free(mid);
  }
  return eff;
}

static oracle_opset* opset_new(int nops) {
  oracle_opset* r = (oracle_opset*)xcalloc(1, sizeof(oracle_opset));
  r->nops = nops;
  r->op = (OBlock*)xcalloc((size_t)nops, sizeof(OBlock));
  return r;
}

int oracle_opset_sizes(const oracle_opset* r, int* nops, int* nsub, int64_t* ndata) {
  int ns = 0; int64_t nd = 0;
  for (int o = 0; o < r->nops; ++o)
    for (int i = 0; i < r->op[o].num; ++i) { ++ns; nd += (int64_t)r->op[o].sub[i].row_dim * r->op[o].sub[i].sdim; }
  *nops = r->nops; *nsub = ns; *ndata = nd;
  return 0;
}

void oracle_opset_copy(const oracle_opset* r, int* hdr, int* sub, int64_t* off, double* data) {
  int s = 0; int64_t nd = 0;
  for (int o = 0; o < r->nops; ++o) {
    const OBlock* b = &r->op[o];
    hdr[4 * o] = b->len; hdr[4 * o + 1] = b->up_dif; hdr[4 * o + 2] = b->down_dif; hdr[4 * o + 3] = b->num;
    for (int i = 0; i < b->num; ++i, ++s) {
      const OSub* q = &b->sub[i];
      sub[5 * s] = q->num_up; sub[5 * s + 1] = q->num_down; sub[5 * s + 2] = q->down_dif;
      sub[5 * s + 3] = q->row_dim; sub[5 * s + 4] = q->sdim;
      off[s] = nd;
      size_t n = (size_t)q->row_dim * q->sdim;
      memcpy(data + nd, q->mat, sizeof(double) * n);
      nd += (int64_t)n;
    }
  }
}

void oracle_opset_free(oracle_opset* r) {
  if (!r) return;
  for (int o = 0; o < r->nops; ++o) free_block(&r->op[o]);
  free(r->op);
  free(r);
}

/* Get_Hamiltonian_Sys (side 0, get_sys.f90:4-246) / Get_Hamiltonian_Env
 * (side 1, :247-478).  mshift = sys_len + 1 places the environment block's
 * sites on the chain (dmrg_sub.f90:692, 989). */
oracle_opset* oracle_get_hamiltonian(const oracle_lattice* lat, int side, int sys_len, int env_len, int pre_nops,
                                     const int* pre_hdr, const int* pre_sub, const int64_t* pre_off,
                                     const double* pre_data, int blen, int nsec, const int* bsec, const int* bsub,
                                     const int* trun_hdr, const int* trun_sub, const int64_t* trun_off,
                                     const double* trun_data) {
  Lat L; lat_init(&L, lat); lat_setup(&L, lat->ny, sys_len, env_len);
  SiteOps so; make_site_ops(&so, lat->up_bias);
  OModel pre = view_model(pre_nops, pre_hdr, pre_sub, pre_off, pre_data);
  OBasis bs = view_basis(nsec, bsec, bsub);
  OBlock trun = view_op(trun_hdr, trun_sub, trun_off, trun_data);
  int pre_len = side == 0 ? sys_len : env_len, mshift = sys_len + 1;
  int nleg_pre = L.nleg11[pre_len], nleg1 = L.nleg11[pre_len + 1];
  int nslot = 64, jnew = INB1(&L, pre_len + 1, pre_len + 1);
  if (pre.nact < nleg_pre || nleg_pre >= nslot || nleg1 >= nslot || jnew >= nslot) {
    free_model(&pre); free(bs.sec); free_block(&trun); lat_free(&L);
    return NULL;
  }
  OSet* nxt = (OSet*)xcalloc((size_t)nslot, sizeof(OSet));
  OBlock H = o_update_block_dia(&pre.ham, &bs, blen);
  for (int ij = 1; ij <= 5; ij += 2) {
    for (int k = 1; k <= NEIBT; ++k) {
      int II;
      if (side == 0) II = NEIB(&L, pre_len + 1, k);
      else II = L.nposi_dm[NEIB(&L, L.nposi_lat[pre_len + 1 + mshift], k)] - mshift;
      if (!(II <= pre_len && II > 0)) continue;
      int i1 = INB1(&L, II, pre_len);
      if (!(i1 > 0 && i1 <= nleg_pre)) continue;
      double coefx = L.jt[k], coefsd = L.jd[k];
      char fl = side == 0 ? 'S' : 'E';
      if (ij == 1 && coefx != 0.0) o_ham_ndia(&pre.sub[i1].elec_up, &so.elec_up, &bs, coefx, &H, 'F', fl);
      if (ij == 3 && coefsd != 0.0) o_ham_ndia(&pre.sub[i1].spin_sd, &so.st_sd, &bs, coefsd, &H, 'B', fl);
    }
    for (int y = 1; y <= pre_len; ++y) {
      int I1 = INB1(&L, y, pre_len);
      if (I1 > nleg_pre) continue;
      if (ij == 1) {
        nxt[I1].elec_up = o_update_block_ndia(&pre.sub[I1].elec_up, &bs, blen, 'F', side);
        nxt[I1].elec_down = o_update_block_ndia(&pre.sub[I1].elec_down, &bs, blen, 'F', side);
      } else if (ij == 3) {
        nxt[I1].spin_sd = o_update_block_ndia(&pre.sub[I1].spin_sd, &bs, blen, 'B', side);
      } else {
        nxt[I1].num_sn = o_update_block_dia(&pre.sub[I1].num_sn, &bs, blen);
      }
    }
    OBlock tmp, tmp1; memset(&tmp1, 0, sizeof(tmp1));
    if (ij == 1) {
      tmp = o_update_site_ndia(&so.elec_up, &bs, blen, 'F', side);
      tmp1 = o_update_site_ndia(&so.elec_down, &bs, blen, 'F', side);
    } else if (ij == 3) {
      tmp = o_update_site_ndia(&so.st_sd, &bs, blen, 'B', side);
    } else {
      tmp = o_update_site_dia(&so.num_elec, &bs, blen);
      for (int k = 1; k <= NEIBT; ++k) {
        int II;
        if (side == 0) II = NEIB(&L, pre_len + 1, k);
        else II = L.nposi_dm[NEIB(&L, L.nposi_lat[pre_len + 1 + mshift], k)] - mshift;
        if (!(II <= pre_len && II > 0)) continue;
        int i1 = INB1(&L, II, pre_len);
        if (!(i1 > 0 && i1 <= nleg_pre)) continue;
        o_ham_dia(&nxt[i1].num_sn, &tmp, L.jn[k], &H);
      }
      if (L.esite != 0.0) o_add_two(&tmp, L.esite, &H);
    }
    /* reorder (block_transfer copies) */
    for (int I = 1; I <= pre_len; ++I) {
      int I1 = INB1(&L, I, pre_len), J = INB1(&L, I, pre_len + 1);
      if (J <= nleg1 && J != I1) {
        if (ij == 1) {
          free_block(&nxt[J].elec_up); nxt[J].elec_up = ob_copy(&nxt[I1].elec_up);
          free_block(&nxt[J].elec_down); nxt[J].elec_down = ob_copy(&nxt[I1].elec_down);
        } else if (ij == 3) {
          free_block(&nxt[J].spin_sd); nxt[J].spin_sd = ob_copy(&nxt[I1].spin_sd);
        } else {
          free_block(&nxt[J].num_sn); nxt[J].num_sn = ob_copy(&nxt[I1].num_sn);
        }
      }
    }
    int j = INB1(&L, pre_len + 1, pre_len + 1);
    if (ij == 1) {
      free_block(&nxt[j].elec_up); nxt[j].elec_up = ob_copy(&tmp);
      free_block(&nxt[j].elec_down); nxt[j].elec_down = ob_copy(&tmp1);
    } else if (ij == 3) {
      free_block(&nxt[j].spin_sd); nxt[j].spin_sd = ob_copy(&tmp);
    } else {
      free_block(&nxt[j].num_sn); nxt[j].num_sn = ob_copy(&tmp);
    }
    for (int jj = nleg1 + 1; jj <= nleg_pre; ++jj) {
      if (ij == 1) { free_block(&nxt[jj].elec_up); free_block(&nxt[jj].elec_down); }
      else if (ij == 3) free_block(&nxt[jj].spin_sd);
      else free_block(&nxt[jj].num_sn);
    }
    free_block(&tmp); free_block(&tmp1);
  }
  oracle_opset* out = opset_new(1 + 4 * nleg1);
  out->op[0] = o_update_trun_dia(&H, &trun);
  for (int ii = 1; ii <= nleg1; ++ii) {
    out->op[1 + 4 * (ii - 1)] = o_update_trun_ndia(&nxt[ii].elec_up, &trun);
    out->op[2 + 4 * (ii - 1)] = o_update_trun_ndia(&nxt[ii].elec_down, &trun);
    out->op[3 + 4 * (ii - 1)] = o_update_trun_ndia(&nxt[ii].spin_sd, &trun);
    out->op[4 + 4 * (ii - 1)] = o_update_trun_dia(&nxt[ii].num_sn, &trun);
  }
  free_block(&H);
  for (int i = 0; i < nslot; ++i) free_set(&nxt[i]);
  free(nxt);
  free_model(&pre);
  free(bs.sec);
  free_block(&trun);
  lat_free(&L);
  return out;
}

/* update_trun_dia / update_trun_ndia of every operator (ndia[i] selects) */
oracle_opset* oracle_update_trun(int nops, const int* hdr, const int* sub, const int64_t* off, const double* data,
                                 const int* ndia, const int* trun_hdr, const int* trun_sub, const int64_t* trun_off,
                                 const double* trun_data) {
  OBlock trun = view_op(trun_hdr, trun_sub, trun_off, trun_data);
  oracle_opset* out = opset_new(nops);
  int first = 0;
  for (int o = 0; o < nops; ++o) {
    OBlock ori = view_op(hdr + 4 * o, sub + 5 * first, off + first, data);
    out->op[o] = (ndia && ndia[o]) ? o_update_trun_ndia(&ori, &trun) : o_update_trun_dia(&ori, &trun);
    first += hdr[4 * o + 3];
    free_block(&ori);
  }
  free_block(&trun);
  return out;
}

/* Get_density_matrix_sys/env_new (den_mat.f90:84-134 with :4-78): states
 * summed (wave_transfer + add), normalised (wave_normalize, block.f90:2418),
 * then rho = psi psi^T (side 0) or psi^T psi (side 1) per sector. */
oracle_opset* oracle_density_matrix(int side, int len, int nw, const int* wsec, int levels, const double* psi) {
  size_t D = 0;
  for (int i = 0; i < nw; ++i) D += (size_t)wsec[6 * i + 4] * wsec[6 * i + 5];
  double* v = (double*)xcalloc(D + 1, sizeof(double));
  memcpy(v, psi, sizeof(double) * D);
  for (int l = 1; l < levels; ++l)
    for (size_t q = 0; q < D; ++q) v[q] = v[q] + psi[(size_t)l * D + q];
  double nrm = 0.0;
  for (size_t q = 0; q < D; ++q) nrm += v[q] * v[q];
  nrm = sqrt(nrm);
  for (size_t q = 0; q < D; ++q) v[q] = v[q] / nrm;
  oracle_opset* out = opset_new(1);
  OBlock* b = &out->op[0];
  *b = ob_alloc(len, 0, 0, nw);
  size_t o = 0;
  for (int i = 0; i < nw; ++i) {
    const int* w = wsec + 6 * i;
    int ds = w[4], de = w[5];
    if (side == 0) {
      ob_set(b, i, w[0], w[1], 0, ds, ds);
      gemm('N', 'T', ds, ds, de, 1.0, v + o, ds, v + o, ds, 0.0, b->sub[i].mat, ds);
    } else {
      ob_set(b, i, w[2], w[3], 0, de, de);
      gemm('T', 'N', de, de, ds, 1.0, v + o, ds, v + o, ds, 0.0, b->sub[i].mat, de);
    }
    o += (size_t)ds * de;
  }
  free(v);
  return out;
}

/* QuickSort_Double, block.f90:288-332 (arr/idx 1-based inclusive ranges) */
static void qsort_double(double* arr, int* idx, int lower, int upper) {
  int i = lower, j = upper;
  int index = (int)((lower + upper) / 2.0);
  double mn = arr[index];
  while (i <= j) {
    while (arr[i] > mn) i = i + 1;
    while (arr[j] < mn) j = j - 1;
    if (i <= j) {
      double t = arr[i]; arr[i] = arr[j]; arr[j] = t;
      int ti = idx[i]; idx[i] = idx[j]; idx[j] = ti;
      i = i + 1; j = j - 1;
    }
  }
  if (lower < j) qsort_double(arr, idx, lower, j);
  if (i < upper) qsort_double(arr, idx, i, upper);
}

void oracle_quicksort_double(double* arr, int* idx, int n) {
  if (n > 0) qsort_double(arr - 1, idx - 1, 1, n);
}

/* Get_truncation_operator's selection (den_mat.f90:196-328) on given
 * ascending per-sector eigenvalues.  info = {num_state, trunc_err, Entropys,
 * entropy1, entropy2}; counts[k] = kept states of sector k; sorted/order
 * (may be NULL) receive the normalised descending spectrum and its
 * 1-based state indices. */
int oracle_truncation_select(int nsec, const int* den_sub, const double* eigval, int kept_min, int kept_max,
                             double trun_err, int* counts, double* info, double* sorted, int* order) {
  int dim = 0;
  for (int k = 0; k < nsec; ++k) dim += den_sub[5 * k + 4];
  double* ev = (double*)xcalloc((size_t)dim + 1, sizeof(double));
  int* idx = (int*)xcalloc((size_t)dim + 1, sizeof(int));
  int* eigs = (int*)xcalloc((size_t)dim + 1, sizeof(int));
  int* sec = (int*)xcalloc((size_t)dim + 1, sizeof(int));
  int p = 0;
  for (int k = 0; k < nsec; ++k)
    for (int i = 0; i < den_sub[5 * k + 4]; ++i, ++p) {
      ev[p] = eigval[p]; idx[p] = p + 1; eigs[p] = den_sub[5 * k + 1] + 1; sec[p] = k;
    }
  double tot_pm = 0.0, e1 = 0.0, e2 = 0.0;
  for (int i = 0; i < dim; ++i) tot_pm = tot_pm + ev[i] * eigs[i];
  for (int i = 0; i < dim; ++i)
    if (ev[i] >= 1.0E-15) { double v = ev[i] / eigs[i]; e1 = e1 - v * log(v) * eigs[i]; }
  for (int i = 0; i < dim; ++i) ev[i] = ev[i] / tot_pm;
  for (int i = 0; i < dim; ++i)
    if (ev[i] >= 1.0E-15) e2 = e2 - ev[i] * log(ev[i]) * eigs[i];
  for (int i = 0; i < dim; ++i) ev[i] = ev[i] * tot_pm;
  oracle_quicksort_double(ev, idx, dim);
  tot_pm = 0.0;
  for (int i = 0; i < dim; ++i) tot_pm = tot_pm + ev[i];
  for (int i = 0; i < dim; ++i) ev[i] = ev[i] / tot_pm;
  int num; double err = 0.0;
  if (dim <= kept_min) {
    num = dim;
  } else {
    for (int i = kept_min + 1; i <= dim; ++i) err = err + ev[i - 1];
    num = kept_min;
    while (err > trun_err && num < dim) {
      num = num + 1;
      err = err - ev[num - 1];
      if (err <= trun_err || num >= kept_max) break;
    }
  }
  double ent = 0.0;
  for (int i = 0; i < dim; ++i)
    if (ev[i] >= 1.0E-15) ent = ent - ev[i] * log(ev[i]);
  for (int k = 0; k < nsec; ++k) counts[k] = 0;
  for (int k = 0; k < num; ++k) counts[sec[idx[k] - 1]]++;
  info[0] = num; info[1] = err; info[2] = ent; info[3] = e1; info[4] = e2;
  if (sorted) memcpy(sorted, ev, sizeof(double) * (size_t)dim);
  if (order) memcpy(order, idx, sizeof(int) * (size_t)dim);
  free(ev); free(idx); free(eigs); free(sec);
  return num;
}

/* ------------------------------------------------------------------------
 * wave_trans_LR / wave_trans_RL (dmrg_sub.f90:1093-1238, 1241-1384) and
 * wave_normalize (block.f90:2397-2432)
 * ------------------------------------------------------------------------ */
static double wt_negpow(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }

/* first sub-block of operator (hdr, sub) with ket (up, down); -1 if none */
static int wt_first_sub(const int* hdr, const int* sub, int up, int down) {
  for (int k = 0; k < hdr[3]; ++k)
    if (sub[5 * k] == up && sub[5 * k + 1] == down) return k;
  return -1;
}

/* first sector of a basis with (up, down) holding sub-block (bl, st); the
 * sub-block's row in bsub is returned through *q */
static int wt_find_fused(int nsec, const int* bsec, const int* bsub, int up, int down, int blu, int bld, int stu,
                         int std_, int* q) {
  int first = 0;
  for (int y = 0; y < nsec; ++y) {
    const int* s = bsec + 4 * y;
    if (s[0] == up && s[1] == down)
      for (int z = 0; z < s[3]; ++z) {
        const int* m = bsub + 6 * (first + z);
        if (m[0] == blu && m[1] == bld && m[2] == stu && m[3] == std_) {
          *q = first + z;
          return y;
        }
      }
    first += s[3];
  }
  return -1;
}

int oracle_wave_trans(int dir, int sys_nsec, const int* sys_bsec, const int* sys_bsub, int env_nsec,
                      const int* env_bsec, const int* env_bsub, const int* st_hdr, const int* st_sub,
                      const int64_t* st_off, const double* st_data, const int* et_hdr, const int* et_sub,
                      const int64_t* et_off, const double* et_data, int nw_old, const int* wsec_old,
                      const double* x, int nw_new, const int* wsec_new, double* y) {
  int64_t* ooff = (int64_t*)malloc(sizeof(int64_t) * (nw_old > 0 ? nw_old : 1));
  int64_t* noff = (int64_t*)malloc(sizeof(int64_t) * (nw_new > 0 ? nw_new : 1));
  int64_t t = 0;
  for (int i = 0; i < nw_old; ++i) { ooff[i] = t; t += (int64_t)wsec_old[6 * i + 4] * wsec_old[6 * i + 5]; }
  int64_t dn = 0;
  for (int i = 0; i < nw_new; ++i) { noff[i] = dn; dn += (int64_t)wsec_new[6 * i + 4] * wsec_new[6 * i + 5]; }
  for (int64_t e = 0; e < dn; ++e) y[e] = 0.0;
  int rc = 0;
  for (int i = 0; i < nw_old && rc == 0; ++i) {
    const int* o = wsec_old + 6 * i;
    const int su = o[0], sd = o[1], eu = o[2], ed = o[3], osd = o[4], oed = o[5];
    /* the side whose truncation operator is keyed by the old sector, and the
       fused basis walked for the moving site */
    const int tid = dir == 0 ? wt_first_sub(st_hdr, st_sub, su, sd) : wt_first_sub(et_hdr, et_sub, eu, ed);
    if (tid < 0) continue;
    const int wnsec = dir == 0 ? env_nsec : sys_nsec;
    const int* wbsec = dir == 0 ? env_bsec : sys_bsec;
    const int* wbsub = dir == 0 ? env_bsub : sys_bsub;
    int first = 0;
    for (int k = 0; k < wnsec && rc == 0; first += wbsec[4 * k + 3], ++k) {
      const int* sec = wbsec + 4 * k;
      if (dir == 0 ? (sec[0] != eu || sec[1] != ed) : (sec[0] != su || sec[1] != sd)) continue;
      for (int xx = 0; xx < sec[3] && rc == 0; ++xx) {
        const int* m = wbsub + 6 * (first + xx);
        const int blu = m[0], bld = m[1], stu = m[2], std_ = m[3], opos = m[4], osub = m[5];
        const int other = dir == 0 ? ed : sd;   /* 2S that couples with the site on the receiving side */
        for (int nd = abs(dir == 0 ? sd - std_ : ed - std_); nd <= (dir == 0 ? sd : ed) + std_; nd += 2) {
          (void)other;
          int q = -1, w = -1, id2, y2;
          if (dir == 0) {
            const int nsu = su + stu;
            id2 = wt_first_sub(et_hdr, et_sub, blu, bld);
            y2 = wt_find_fused(sys_nsec, sys_bsec, sys_bsub, nsu, nd, su, sd, stu, std_, &q);
            for (int v = 0; v < nw_new; ++v) {
              const int* n = wsec_new + 6 * v;
              if (n[0] == nsu && n[1] == nd && n[2] == blu && n[3] == bld) { w = v; break; }
            }
          } else {
            const int neu = eu + stu;
            id2 = wt_first_sub(st_hdr, st_sub, blu, bld);
            y2 = wt_find_fused(env_nsec, env_bsec, env_bsub, neu, nd, eu, ed, stu, std_, &q);
            for (int v = 0; v < nw_new; ++v) {
              const int* n = wsec_new + 6 * v;
              if (n[0] == blu && n[1] == bld && n[2] == neu && n[3] == nd) { w = v; break; }
            }
          }
          if (id2 < 0 || y2 < 0 || w < 0) continue;
          const int* n = wsec_new + 6 * w;
          const int nsd_ = n[4], ned_ = n[5];
          const int npos = (dir == 0 ? sys_bsub : env_bsub)[6 * q + 4];
          const int nsub = (dir == 0 ? sys_bsub : env_bsub)[6 * q + 5];
          double c;
          if (dir == 0) {   /* j12 = sd, j3 = std_, j123 = nd, j4 = bld, j34 = ed (:1201-1214) */
            const int j12 = sd, j3 = std_, j123 = nd, j4 = bld, j34 = ed;
            c = wt_negpow((j12 + j3 + j123) / 2) / sqrt((1.0 + j12) * (1.0 + j123));
            if (c == 0.0) continue;
            c = c * sqrt((1.0 + j123) * (j34 + 1.0));
            if (((j12 + j3 + j4) / 2) % 2 != 0) c = -c;
            if (((j3 + j4 - j34) / 2) % 2 != 0) c = -c;
          } else {          /* j12 = ed, j3 = std_, j4 = bld, j123 = nd, j34 = sd (:1348-1361) */
            const int j12 = ed, j3 = std_, j4 = bld, j123 = nd, j34 = sd;
            c = wt_negpow((j34 + j4 + j3) / 2) / sqrt((1.0 + j12) * (1.0 + j4));
            if (c == 0.0) continue;
            c = c * sqrt((1.0 + j123) * (j34 + 1.0));
            if (((j4 + j3 + j12) / 2) % 2 != 0) c = -c;
            if (((j12 + j3 - j123) / 2) % 2 != 0) c = -c;
          }
          const int* A = dir == 0 ? st_sub + 5 * tid : st_sub + 5 * id2;   /* sys_trun sub-block */
          const int* B = dir == 0 ? et_sub + 5 * id2 : et_sub + 5 * tid;   /* env_trun sub-block */
          const double* Ad = st_data + (dir == 0 ? st_off[tid] : st_off[id2]);
          const double* Bd = et_data + (dir == 0 ? et_off[id2] : et_off[tid]);
          if (dir == 0) {
            /* sub = c A^T X(:, opos:opos+osub), A: osd x nsub; new(npos.., :) += sub B^T, B: ned_ x osub */
            if (A[3] != osd || A[4] != nsub || B[3] != ned_ || B[4] != osub) { rc = -1; break; }
            double* sub = (double*)calloc((size_t)nsub * osub + 1, sizeof(double));
            const double* X = x + ooff[i] + (int64_t)opos * osd;
            for (int cc = 0; cc < osub; ++cc)
              for (int r = 0; r < nsub; ++r) {
                double s = 0.0;
                for (int k2 = 0; k2 < osd; ++k2) s += Ad[k2 + (int64_t)r * osd] * X[k2 + (int64_t)cc * osd];
                sub[r + (int64_t)cc * nsub] = c * s;
              }
            double* Y = y + noff[w];
            for (int cc = 0; cc < ned_; ++cc)
              for (int r = 0; r < nsub; ++r) {
                double s = 0.0;
                for (int k2 = 0; k2 < osub; ++k2) s += sub[r + (int64_t)k2 * nsub] * Bd[cc + (int64_t)k2 * ned_];
                Y[npos + r + (int64_t)cc * nsd_] += s;
              }
            free(sub);
          } else {
            /* sub = c A X(opos:opos+osub, :), A: nsd_ x osub; new(:, npos..) += sub B, B: oed x nsub */
            if (A[3] != nsd_ || A[4] != osub || B[3] != oed || B[4] != nsub) { rc = -1; break; }
            double* sub = (double*)calloc((size_t)nsd_ * oed + 1, sizeof(double));
            const double* X = x + ooff[i] + opos;
            for (int cc = 0; cc < oed; ++cc)
              for (int r = 0; r < nsd_; ++r) {
                double s = 0.0;
                for (int k2 = 0; k2 < osub; ++k2) s += Ad[r + (int64_t)k2 * nsd_] * X[k2 + (int64_t)cc * osd];
                sub[r + (int64_t)cc * nsd_] = c * s;
              }
            double* Y = y + noff[w];
            for (int cc = 0; cc < nsub; ++cc)
              for (int r = 0; r < nsd_; ++r) {
                double s = 0.0;
                for (int k2 = 0; k2 < oed; ++k2) s += sub[r + (int64_t)k2 * nsd_] * Bd[k2 + (int64_t)cc * oed];
                Y[r + (int64_t)(npos + cc) * nsd_] += s;
              }
            free(sub);
          }
        }
      }
    }
  }
  if (rc == 0) {   /* wave_normalize: norm summed sector by sector, column by column */
    double nrm = 0.0;
    for (int v = 0; v < nw_new; ++v) {
      const int rows = wsec_new[6 * v + 4], cols = wsec_new[6 * v + 5];
      for (int cc = 0; cc < cols; ++cc) {
        double dcol = 0.0;
        for (int r = 0; r < rows; ++r) {
          const double a = y[noff[v] + r + (int64_t)cc * rows];
          dcol += a * a;
        }
        nrm += dcol;
      }
    }
    nrm = sqrt(nrm);
    for (int64_t e = 0; e < dn; ++e) y[e] = y[e] / nrm;
  }
  free(ooff);
  free(noff);
  return rc;
}
