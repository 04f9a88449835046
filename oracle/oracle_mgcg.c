/*
 * oracle_mgcg.c -- restatement of the multigrid-preconditioned conjugate
 * gradient solve of the pressure-correction system (an extension: the
 * reference solves that system with the 4-subdomain Gauss-Seidel solver of
 * src/modules/mod_solver.f90:124-189 and has no Krylov or multigrid solver;
 * its agglomeration levels, src/setup/mod_agglomeration.f90:331-561 and
 * src/setup/mod_mg_lvl_uns.f90:490-1001, are built but never used by a
 * solver).
 *
 * TEST INFRASTRUCTURE ONLY (the checker, not the product).
 *
 * The GPU implementation follows the same operation order, so the two are
 * bitwise equal; the order of every sum is stated below.
 *
 * Hierarchy (built once per mesh, from the cell centres and the CSR):
 *   - agglomeration by recursive longest-axis midpoint bisection of the cell
 *     centres (the splitting rule of the reference's oct-tree seeds,
 *     mod_agglomeration.f90:380-410): a node with more than `cap` cells is
 *     split on the axis of greatest extent (first axis on ties) at
 *     (min + max) / 2, points with coordinate < midpoint going left, both
 *     sides keeping increasing cell order; an empty side splits the node by
 *     halves.  A level-l aggregate is a maximal tree node with at most cap^l
 *     cells.  Aggregates are numbered in tree pre-order (left first), so each
 *     level's aggregates are unions of whole aggregates of the level below.
 *     Levels are added while the aggregate count exceeds `ncoarse`; a level
 *     of one aggregate or without coarsening is not added.
 *   - each level's graph: level 0 is the mesh CSR (interior neighbours);
 *     level l+1 has an edge I-J when an edge of level l joins them; rows
 *     list neighbours in increasing order.
 *   - colours: two, by the parity of the breadth-first distance from the
 *     least node of each connected part (neighbours in CSR order), when that
 *     is a proper colouring (the graphs of hex boxes and of their box-shaped
 *     aggregates are bipartite); otherwise greedy in index order (least
 *     colour none of the earlier neighbours has); the visiting order is by
 *     colour, then index.
 *   - Galerkin operator A_{l+1} = P^T A_l P for the piecewise-constant P:
 *     ap_I = (sum over members i of ap_i, in member order) minus (anb of the
 *     level-l entries inside I, in (member, CSR) order); anb_IJ = sum of the
 *     anb entries from I's members to J's members, in (member, CSR) order,
 *     starting from 0.
 *
 * V-cycle z = M r on level l (x_l starts at zero): one forward colour
 * Gauss-Seidel sweep x_i = (b_i + sum_k anb_k x_nb) / ap_i (sum in CSR
 * order, level 0 over every CSR entry with zero halo values), the residual
 * restricted as b_{l+1,I} = sum over members i of
 * ((b_i + sum_k anb_k x_nb) - ap_i x_i) from 0, the V-cycle on level l+1,
 * x_i += x_{l+1,agg(i)}, one backward colour sweep.  The coarsest level runs
 * MG_COARSE_SWEEPS forward + backward sweep pairs from zero.  Forward and
 * backward sweeps are adjoint, so M is symmetric.
 *
 * Conjugate gradients (FP64): r = b + sum anb x - ap x; res_i = rms(r);
 * stop when rms(r) <= res_i / 10 or after nit iterations (the reference's
 * stop rule, mod_solver.f90:271-322); each iteration z = M r, rho = r.z,
 * p = z (+ rho/rho_old p), q = A p with q_i = ap_i p_i - sum_k anb_k p_nb,
 * sigma = p.q, alpha = rho/sigma, x += alpha p, r -= alpha q.  A breakdown
 * (sigma or rho not positive) ends the solve before the update.  In the
 * mixed variant M runs in FP32 on FP32 copies of the coefficients and of r
 * (Galerkin sums in FP32), everything else in FP64.
 *
 * Reductions (dot products and norms) follow the GPU's blocked order: nblk
 * = min(1024, max(1, ceil(n / 4096))) blocks of chunk = ceil(n / nblk)
 * elements; lane t (of 256) of block x sums elements x*chunk + t + 256j in
 * increasing j from 0; the 256 lanes are combined by the pairwise tree
 * s[t] += s[t + h] for h = 128, 64, ..., 1; the block sums are combined the
 * same way (lane t summing blocks t, t + 256, ...).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PK_E(p) ((int)((unsigned)(p) >> 5))

/* ------------------------------------------------------------------------ */
/* blocked tree reductions                                                   */
/* ------------------------------------------------------------------------ */
void or_tree_blocks(int n, int *nblk, int *chunk) {
  int nb = (n + 4095) / 4096;
  if (nb < 1) nb = 1;
  if (nb > 1024) nb = 1024;
  *nblk = nb;
  *chunk = (n + nb - 1) / nb;
}

static double tree256(double *s) {
  for (int h = 128; h >= 1; h >>= 1)
    for (int t = 0; t < h; t++) s[t] = s[t] + s[t + h];
  return s[0];
}
static double tree256_max(double *s) {
  for (int h = 128; h >= 1; h >>= 1)
    for (int t = 0; t < h; t++) s[t] = fmax(s[t], s[t + h]);
  return s[0];
}

/* sum over i < n of a[i]*b[i] (b NULL: a[i]*a[i]) in the blocked order */
double or_tree_dot(const double *a, const double *b, int n) {
  int nblk, chunk;
  or_tree_blocks(n, &nblk, &chunk);
  double *part = malloc(sizeof(double) * (size_t)nblk);
  double s[256];
  for (int x = 0; x < nblk; x++) {
    const int lo = x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    for (int t = 0; t < 256; t++) {
      double acc = 0.0;
      for (int i = lo + t; i < hi; i += 256) acc = acc + a[i] * (b ? b[i] : a[i]);
      s[t] = acc;
    }
    part[x] = tree256(s);
  }
  for (int t = 0; t < 256; t++) {
    double acc = 0.0;
    for (int i = t; i < nblk; i += 256) acc = acc + part[i];
    s[t] = acc;
  }
  free(part);
  return tree256(s);
}

/* sum over i < n of a[i] in the blocked order */
double or_tree_sum(const double *a, int n) {
  int nblk, chunk;
  or_tree_blocks(n, &nblk, &chunk);
  double *part = malloc(sizeof(double) * (size_t)nblk);
  double s[256];
  for (int x = 0; x < nblk; x++) {
    const int lo = x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    for (int t = 0; t < 256; t++) {
      double acc = 0.0;
      for (int i = lo + t; i < hi; i += 256) acc = acc + a[i];
      s[t] = acc;
    }
    part[x] = tree256(s);
  }
  for (int t = 0; t < 256; t++) {
    double acc = 0.0;
    for (int i = t; i < nblk; i += 256) acc = acc + part[i];
    s[t] = acc;
  }
  free(part);
  return tree256(s);
}

/* max over i < n of |a[i]| in the same blocked order */
static double tree_absmax(const double *a, int n) {
  int nblk, chunk;
  or_tree_blocks(n, &nblk, &chunk);
  double *part = malloc(sizeof(double) * (size_t)nblk);
  double s[256];
  for (int x = 0; x < nblk; x++) {
    const int lo = x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    for (int t = 0; t < 256; t++) {
      double acc = 0.0;
      for (int i = lo + t; i < hi; i += 256) acc = fmax(acc, fabs(a[i]));
      s[t] = acc;
    }
    part[x] = tree256_max(s);
  }
  for (int t = 0; t < 256; t++) {
    double acc = 0.0;
    for (int i = t; i < nblk; i += 256) acc = fmax(acc, part[i]);
    s[t] = acc;
  }
  free(part);
  return tree256_max(s);
}

/* ------------------------------------------------------------------------ */
/* hierarchy                                                                 */
/* ------------------------------------------------------------------------ */
#define MG_MAX_LEVELS 16
#define MG_COARSE_SWEEPS 10

typedef struct {
  int n;          /* nodes */
  int *row, *col; /* CSR (level 0: the mesh CSR, 0-based slots; col -1 = halo) */
  int nnz;
  int ncol, *cptr, *cperm; /* colours: cperm[cptr[k] .. cptr[k+1]) */
  int *agg;                /* node -> aggregate on the next level (n) */
  /* the next level's Galerkin lists, indexed by next-level row / slot */
  int *mptr, *mem;   /* members of aggregate I (this level's nodes) */
  int *iptr, *ilist; /* slots of this level inside aggregate I */
  int *gptr, *glist; /* per next-level slot: this level's slots, in order */
} or_mglevel;

typedef struct or_mg {
  int nlev;
  or_mglevel lv[MG_MAX_LEVELS];
} or_mg;

typedef struct {
  int *ids;
  int n;
} idlist;

/* pre-order traversal state */
typedef struct {
  const double *xyz[3];
  int cap[MG_MAX_LEVELS]; /* cap[l] = maxagg^l, l >= 1 */
  int nl;                 /* levels considered */
  int *count;             /* aggregates numbered so far per level */
  int *aggid;             /* per level l (1..nl): fine cell -> level-l aggregate, nl*ne */
  int ne;
  int *tmp;
} bisect_t;

static void bisect_node(bisect_t *B, int *ids, int n, long parent_count) {
  /* aggregates this node forms */
  for (int l = 1; l <= B->nl; l++) {
    if (n <= B->cap[l] && parent_count > B->cap[l]) {
      const int a = B->count[l]++;
      for (int i = 0; i < n; i++) B->aggid[(size_t)(l - 1) * B->ne + ids[i]] = a;
    }
  }
  if (n <= B->cap[1]) return;
  double mn[3], mx[3];
  for (int a = 0; a < 3; a++) {
    mn[a] = B->xyz[a][ids[0]];
    mx[a] = mn[a];
  }
  for (int i = 1; i < n; i++)
    for (int a = 0; a < 3; a++) {
      const double v = B->xyz[a][ids[i]];
      if (v < mn[a]) mn[a] = v;
      if (v > mx[a]) mx[a] = v;
    }
  double d = 0.0;
  int ax = -1;
  for (int a = 0; a < 3; a++)
    if (mx[a] - mn[a] > d) {
      d = mx[a] - mn[a];
      ax = a;
    }
  int nl = 0;
  if (ax >= 0) {
    const double m = (mn[ax] + mx[ax]) / 2.0;
    int nr = 0;
    for (int i = 0; i < n; i++) {
      if (B->xyz[ax][ids[i]] < m)
        ids[nl++] = ids[i];
      else
        B->tmp[nr++] = ids[i];
    }
    memcpy(ids + nl, B->tmp, sizeof(int) * (size_t)nr);
  }
  if (nl == 0 || nl == n) nl = n / 2; /* degenerate: by halves */
  bisect_node(B, ids, nl, n);
  bisect_node(B, ids + nl, n - nl, n);
}

static void colour_level(or_mglevel *L, int interior_limit) {
  const int n = L->n;
  int *col = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *queue = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int ncol = 0, two = 1;
  for (int e = 0; e < n; e++) col[e] = -1;
  /* two colours by breadth-first parity from the least node of each part */
  for (int s = 0; s < n && two; s++) {
    if (col[s] >= 0) continue;
    col[s] = 0;
    int qh = 0, qt = 0;
    queue[qt++] = s;
    while (qh < qt && two) {
      const int e = queue[qh++];
      for (int k = L->row[e]; k < L->row[e + 1]; k++) {
        const int j = L->col[k];
        if (j < 0 || j >= interior_limit || j == e) continue;
        if (col[j] < 0) {
          col[j] = 1 - col[e];
          queue[qt++] = j;
        } else if (col[j] == col[e]) {
          two = 0;
          break;
        }
      }
    }
  }
  free(queue);
  if (two) {
    for (int e = 0; e < n; e++)
      if (col[e] + 1 > ncol) ncol = col[e] + 1;
  } else {
    for (int e = 0; e < n; e++) col[e] = 0;
  }
  /* otherwise greedy in index order */
  for (int e = 0; e < n && !two; e++) {
    unsigned long long used = 0;
    for (int k = L->row[e]; k < L->row[e + 1]; k++) {
      const int j = L->col[k];
      if (j >= 0 && j < interior_limit && j < e && col[j] < 64) used |= 1ull << col[j];
    }
    int c = 0;
    while (c < 64 && (used >> c & 1ull)) c++;
    col[e] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  L->ncol = ncol;
  L->cptr = calloc((size_t)ncol + 1, sizeof(int));
  L->cperm = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int e = 0; e < n; e++) L->cptr[col[e] + 1]++;
  for (int c = 0; c < ncol; c++) L->cptr[c + 1] += L->cptr[c];
  int *pos = malloc(sizeof(int) * (size_t)(ncol + 1));
  memcpy(pos, L->cptr, sizeof(int) * (size_t)(ncol + 1));
  for (int e = 0; e < n; e++) L->cperm[pos[col[e]]++] = e;
  free(pos);
  free(col);
}

static int cmp_int(const void *a, const void *b) {
  const int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* the next level from this level's aggregate map */
static void coarsen(or_mglevel *F, or_mglevel *C, int nagg) {
  const int n = F->n;
  C->n = nagg;
  F->mptr = calloc((size_t)nagg + 1, sizeof(int));
  F->mem = malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) F->mptr[F->agg[i] + 1]++;
  for (int a = 0; a < nagg; a++) F->mptr[a + 1] += F->mptr[a];
  {
    int *pos = malloc(sizeof(int) * (size_t)(nagg + 1));
    memcpy(pos, F->mptr, sizeof(int) * (size_t)(nagg + 1));
    for (int i = 0; i < n; i++) F->mem[pos[F->agg[i]]++] = i;
    free(pos);
  }
  /* coarse rows: sorted distinct neighbour aggregates */
  C->row = calloc((size_t)nagg + 1, sizeof(int));
  int cap = 64, *buf = malloc(sizeof(int) * cap);
  int nnz = 0, ccap = 1024;
  C->col = malloc(sizeof(int) * ccap);
  for (int I = 0; I < nagg; I++) {
    int nb = 0;
    for (int q = F->mptr[I]; q < F->mptr[I + 1]; q++) {
      const int i = F->mem[q];
      for (int k = F->row[i]; k < F->row[i + 1]; k++) {
        const int j = F->col[k];
        if (j < 0) continue;
        const int J = F->agg[j];
        if (J == I) continue;
        if (nb == cap) {
          cap *= 2;
          buf = realloc(buf, sizeof(int) * cap);
        }
        buf[nb++] = J;
      }
    }
    qsort(buf, (size_t)nb, sizeof(int), cmp_int);
    int u = 0;
    for (int t = 0; t < nb; t++)
      if (u == 0 || buf[t] != buf[u - 1]) buf[u++] = buf[t];
    if (nnz + u > ccap) {
      while (nnz + u > ccap) ccap *= 2;
      C->col = realloc(C->col, sizeof(int) * ccap);
    }
    memcpy(C->col + nnz, buf, sizeof(int) * (size_t)u);
    nnz += u;
    C->row[I + 1] = nnz;
  }
  C->nnz = nnz;
  free(buf);
  /* Galerkin lists: internal slots per aggregate, fine slots per coarse slot */
  F->iptr = calloc((size_t)nagg + 1, sizeof(int));
  F->gptr = calloc((size_t)nnz + 1, sizeof(int));
  for (int I = 0; I < nagg; I++)
    for (int q = F->mptr[I]; q < F->mptr[I + 1]; q++) {
      const int i = F->mem[q];
      for (int k = F->row[i]; k < F->row[i + 1]; k++) {
        const int j = F->col[k];
        if (j < 0) continue;
        const int J = F->agg[j];
        if (J == I) {
          F->iptr[I + 1]++;
        } else {
          int s = C->row[I];
          while (C->col[s] != J) s++;
          F->gptr[s + 1]++;
        }
      }
    }
  for (int I = 0; I < nagg; I++) F->iptr[I + 1] += F->iptr[I];
  for (int s = 0; s < nnz; s++) F->gptr[s + 1] += F->gptr[s];
  F->ilist = malloc(sizeof(int) * (size_t)(F->iptr[nagg] > 0 ? F->iptr[nagg] : 1));
  F->glist = malloc(sizeof(int) * (size_t)(F->gptr[nnz] > 0 ? F->gptr[nnz] : 1));
  int *ip = malloc(sizeof(int) * (size_t)(nagg + 1)), *gp = malloc(sizeof(int) * (size_t)(nnz + 1));
  memcpy(ip, F->iptr, sizeof(int) * (size_t)(nagg + 1));
  memcpy(gp, F->gptr, sizeof(int) * (size_t)(nnz + 1));
  for (int I = 0; I < nagg; I++)
    for (int q = F->mptr[I]; q < F->mptr[I + 1]; q++) {
      const int i = F->mem[q];
      for (int k = F->row[i]; k < F->row[i + 1]; k++) {
        const int j = F->col[k];
        if (j < 0) continue;
        const int J = F->agg[j];
        if (J == I) {
          F->ilist[ip[I]++] = k;
        } else {
          int s = C->row[I];
          while (C->col[s] != J) s++;
          F->glist[gp[s]++] = k;
        }
      }
    }
  free(ip);
  free(gp);
}

or_mg *or_mg_build(int ne, const int *ef2nb_idx, const int *ef2nb, const double *xc,
                   const double *yc, const double *zc, int maxagg, int ncoarse) {
  or_mg *M = calloc(1, sizeof(or_mg));
  /* level 0: the mesh CSR, 0-based slots, neighbours 0-based, halos -1 */
  or_mglevel *L0 = &M->lv[0];
  L0->n = ne;
  L0->row = malloc(sizeof(int) * (size_t)(ne + 1));
  for (int e = 0; e <= ne; e++) L0->row[e] = ef2nb_idx[e] - 1;
  L0->nnz = L0->row[ne];
  L0->col = malloc(sizeof(int) * (size_t)(L0->nnz > 0 ? L0->nnz : 1));
  for (int k = 0; k < L0->nnz; k++) {
    const int d = PK_E(ef2nb[k]);
    L0->col[k] = d <= ne ? d - 1 : -1;
  }
  colour_level(L0, ne);
  /* the bisection tree gives every level's aggregates at once */
  bisect_t B;
  memset(&B, 0, sizeof(B));
  B.xyz[0] = xc;
  B.xyz[1] = yc;
  B.xyz[2] = zc;
  B.ne = ne;
  B.nl = 0;
  long capl = 1;
  for (int l = 1; l < MG_MAX_LEVELS; l++) {
    capl *= maxagg;
    if (capl > (long)ne * maxagg || capl > (1L << 30)) break;
    B.cap[l] = (int)capl;
    B.nl = l;
  }
  B.count = calloc((size_t)B.nl + 1, sizeof(int));
  B.aggid = malloc(sizeof(int) * (size_t)B.nl * (size_t)ne);
  B.tmp = malloc(sizeof(int) * (size_t)ne);
  int *ids = malloc(sizeof(int) * (size_t)ne);
  for (int e = 0; e < ne; e++) ids[e] = e;
  bisect_node(&B, ids, ne, 2L * (1L << 30));
  M->nlev = 1;
  for (int l = 1; l <= B.nl; l++) {
    or_mglevel *F = &M->lv[l - 1];
    const int nagg = B.count[l];
    if (F->n <= ncoarse || nagg <= 1 || nagg >= F->n) break;
    /* node -> aggregate: via any member cell of the node */
    F->agg = malloc(sizeof(int) * (size_t)F->n);
    if (l == 1) {
      for (int e = 0; e < ne; e++) F->agg[e] = B.aggid[e];
    } else {
      const int *prev = B.aggid + (size_t)(l - 2) * ne, *cur = B.aggid + (size_t)(l - 1) * ne;
      for (int e = 0; e < ne; e++) F->agg[prev[e]] = cur[e];
    }
    coarsen(F, &M->lv[l], nagg);
    colour_level(&M->lv[l], nagg);
    M->nlev = l + 1;
  }
  free(ids);
  free(B.tmp);
  free(B.aggid);
  free(B.count);
  return M;
}

void or_mg_free(or_mg *M) {
  if (!M) return;
  for (int l = 0; l < MG_MAX_LEVELS; l++) {
    or_mglevel *L = &M->lv[l];
    void *p[] = {L->row, L->col, L->cptr, L->cperm, L->agg, L->mptr, L->mem,
                 L->iptr, L->ilist, L->gptr, L->glist};
    for (size_t k = 0; k < sizeof(p) / sizeof(p[0]); k++) free(p[k]);
  }
  free(M);
}

int or_mg_nlev(const or_mg *M) { return M ? M->nlev : 0; }

/* copy out one hierarchy array of a level ("n", "row", "col", "cptr",
   "cperm", "agg", "mptr", "mem", "iptr", "ilist", "gptr", "glist");
   returns the element count, or -1 */
int or_mg_get(const or_mg *M, int l, const char *name, int *dst, int n) {
  if (!M || l < 0 || l >= M->nlev) return -1;
  const or_mglevel *L = &M->lv[l];
  const int *src = NULL;
  int cnt = -1;
  const int last = l + 1 < M->nlev;
  const int nn = last ? M->lv[l + 1].n : 0, nnz = last ? M->lv[l + 1].nnz : 0;
  if (!strcmp(name, "n")) {
    if (dst && n >= 1) dst[0] = L->n;
    return 1;
  } else if (!strcmp(name, "row")) {
    src = L->row; cnt = L->n + 1;
  } else if (!strcmp(name, "col")) {
    src = L->col; cnt = L->nnz;
  } else if (!strcmp(name, "cptr")) {
    src = L->cptr; cnt = L->ncol + 1;
  } else if (!strcmp(name, "cperm")) {
    src = L->cperm; cnt = L->n;
  } else if (last && !strcmp(name, "agg")) {
    src = L->agg; cnt = L->n;
  } else if (last && !strcmp(name, "mptr")) {
    src = L->mptr; cnt = nn + 1;
  } else if (last && !strcmp(name, "mem")) {
    src = L->mem; cnt = L->n;
  } else if (last && !strcmp(name, "iptr")) {
    src = L->iptr; cnt = nn + 1;
  } else if (last && !strcmp(name, "ilist")) {
    src = L->ilist; cnt = L->iptr[nn];
  } else if (last && !strcmp(name, "gptr")) {
    src = L->gptr; cnt = nnz + 1;
  } else if (last && !strcmp(name, "glist")) {
    src = L->glist; cnt = L->gptr[nnz];
  }
  if (!src) return -1;
  if (dst) memcpy(dst, src, sizeof(int) * (size_t)(cnt < n ? cnt : n));
  return cnt;
}

/* ------------------------------------------------------------------------ */
/* numerics: Galerkin values, V-cycle (FP64 and FP32), conjugate gradients   */
/* ------------------------------------------------------------------------ */
/* One V-cycle variant per real type.  T is the type; the level-0 system is
   ap0/anb0 (CSR of the mesh), the coarse ones are built by Galerkin sums in
   T.  Vectors b, x per level; level 0 x has halo slots (col -1) read as 0. */
#define MG_DEFINE(T, SUF)                                                                     \
  typedef struct {                                                                            \
    T *ap[MG_MAX_LEVELS], *anb[MG_MAX_LEVELS], *b[MG_MAX_LEVELS], *x[MG_MAX_LEVELS];          \
  } mgval_##SUF;                                                                              \
  static void mg_galerkin_##SUF(const or_mg *M, mgval_##SUF *V) {                             \
    for (int l = 0; l + 1 < M->nlev; l++) {                                                   \
      const or_mglevel *F = &M->lv[l], *C = &M->lv[l + 1];                                    \
      for (int I = 0; I < C->n; I++) {                                                        \
        T acc = (T)0;                                                                         \
        for (int q = F->mptr[I]; q < F->mptr[I + 1]; q++) acc = acc + V->ap[l][F->mem[q]];    \
        for (int q = F->iptr[I]; q < F->iptr[I + 1]; q++) acc = acc - V->anb[l][F->ilist[q]]; \
        V->ap[l + 1][I] = acc;                                                                \
      }                                                                                       \
      for (int s = 0; s < C->nnz; s++) {                                                      \
        T acc = (T)0;                                                                         \
        for (int q = F->gptr[s]; q < F->gptr[s + 1]; q++) acc = acc + V->anb[l][F->glist[q]]; \
        V->anb[l + 1][s] = acc;                                                               \
      }                                                                                       \
    }                                                                                         \
  }                                                                                           \
  static T mg_rowsum_##SUF(const or_mglevel *L, const T *anb, const T *x, T b, int i) {       \
    T s = b;                                                                                  \
    for (int k = L->row[i]; k < L->row[i + 1]; k++) {                                         \
      const int j = L->col[k];                                                                \
      s = s + anb[k] * (j >= 0 ? x[j] : (T)0);                                                \
    }                                                                                         \
    return s;                                                                                 \
  }                                                                                           \
  static void mg_sweep_##SUF(const or_mglevel *L, const T *ap, const T *anb, const T *b,      \
                             T *x, int backward) {                                            \
    for (int cc = 0; cc < L->ncol; cc++) {                                                    \
      const int c = backward ? L->ncol - 1 - cc : cc;                                         \
      for (int q = L->cptr[c]; q < L->cptr[c + 1]; q++) {                                     \
        const int i = L->cperm[q];                                                            \
        x[i] = mg_rowsum_##SUF(L, anb, x, b[i], i) / ap[i];                                   \
      }                                                                                       \
    }                                                                                         \
  }                                                                                           \
  static void mg_vcycle_##SUF(const or_mg *M, mgval_##SUF *V, int l) {                        \
    const or_mglevel *L = &M->lv[l];                                                          \
    T *x = V->x[l];                                                                           \
    for (int i = 0; i < L->n; i++) x[i] = (T)0;                                               \
    if (l + 1 == M->nlev) {                                                                   \
      for (int k = 0; k < MG_COARSE_SWEEPS; k++) {                                            \
        mg_sweep_##SUF(L, V->ap[l], V->anb[l], V->b[l], x, 0);                                \
        mg_sweep_##SUF(L, V->ap[l], V->anb[l], V->b[l], x, 1);                                \
      }                                                                                       \
      return;                                                                                 \
    }                                                                                         \
    mg_sweep_##SUF(L, V->ap[l], V->anb[l], V->b[l], x, 0);                                    \
    const or_mglevel *C = &M->lv[l + 1];                                                      \
    for (int I = 0; I < C->n; I++) {                                                          \
      T acc = (T)0;                                                                           \
      for (int q = L->mptr[I]; q < L->mptr[I + 1]; q++) {                                     \
        const int i = L->mem[q];                                                              \
        const T t = mg_rowsum_##SUF(L, V->anb[l], x, V->b[l][i], i) - V->ap[l][i] * x[i];     \
        acc = acc + t;                                                                        \
      }                                                                                       \
      V->b[l + 1][I] = acc;                                                                   \
    }                                                                                         \
    mg_vcycle_##SUF(M, V, l + 1);                                                             \
    for (int i = 0; i < L->n; i++) x[i] = x[i] + V->x[l + 1][L->agg[i]];                      \
    mg_sweep_##SUF(L, V->ap[l], V->anb[l], V->b[l], x, 1);                                    \
  }                                                                                           \
  static void mgval_alloc_##SUF(const or_mg *M, mgval_##SUF *V) {                             \
    for (int l = 0; l < M->nlev; l++) {                                                       \
      const int n = M->lv[l].n, nnz = M->lv[l].nnz;                                           \
      V->ap[l] = malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));                                 \
      V->anb[l] = malloc(sizeof(T) * (size_t)(nnz > 0 ? nnz : 1));                            \
      V->b[l] = malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));                                  \
      V->x[l] = malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));                                  \
    }                                                                                         \
  }                                                                                           \
  static void mgval_free_##SUF(const or_mg *M, mgval_##SUF *V) {                              \
    for (int l = 0; l < M->nlev; l++) {                                                       \
      free(V->ap[l]);                                                                         \
      free(V->anb[l]);                                                                        \
      free(V->b[l]);                                                                          \
      free(V->x[l]);                                                                          \
    }                                                                                         \
  }

MG_DEFINE(double, d)
MG_DEFINE(float, f)

/* the hierarchy of a case, built on first use */
static or_mg *case_mg(or_case *c) {
  if (!c->mg)
    c->mg = or_mg_build(c->ne, c->ef2nb_idx, c->ef2nb, c->xc, c->yc, c->zc, 8, 64);
  return c->mg;
}

/* z = M r for the system ap / anb of the case (FP64 or FP32), exported for
   the tests (symmetry of M, Galerkin operators) */
void or_mg_apply(or_case *c, const double *ap, const double *anb, const double *r, double *z,
                 int fp32) {
  const or_mg *M = case_mg(c);
  const int ne = c->ne, L = M->lv[0].nnz;
  if (fp32) {
    mgval_f V;
    mgval_alloc_f(M, &V);
    for (int i = 0; i < ne; i++) V.ap[0][i] = (float)ap[i];
    for (int k = 0; k < L; k++) V.anb[0][k] = (float)anb[k];
    mg_galerkin_f(M, &V);
    for (int i = 0; i < ne; i++) V.b[0][i] = (float)r[i];
    mg_vcycle_f(M, &V, 0);
    for (int i = 0; i < ne; i++) z[i] = (double)V.x[0][i];
    mgval_free_f(M, &V);
  } else {
    mgval_d V;
    mgval_alloc_d(M, &V);
    memcpy(V.ap[0], ap, sizeof(double) * (size_t)ne);
    memcpy(V.anb[0], anb, sizeof(double) * (size_t)L);
    mg_galerkin_d(M, &V);
    memcpy(V.b[0], r, sizeof(double) * (size_t)ne);
    mg_vcycle_d(M, &V, 0);
    memcpy(z, V.x[0], sizeof(double) * (size_t)ne);
    mgval_free_d(M, &V);
  }
}

/* the coarse operator of level l (ap then anb), for the tests */
int or_mg_operator(or_case *c, const double *ap, const double *anb, int l, double *apl,
                   double *anbl) {
  const or_mg *M = case_mg(c);
  if (l < 0 || l >= M->nlev) return -1;
  mgval_d V;
  mgval_alloc_d(M, &V);
  memcpy(V.ap[0], ap, sizeof(double) * (size_t)c->ne);
  memcpy(V.anb[0], anb, sizeof(double) * (size_t)M->lv[0].nnz);
  mg_galerkin_d(M, &V);
  if (apl) memcpy(apl, V.ap[l], sizeof(double) * (size_t)M->lv[l].n);
  if (anbl) memcpy(anbl, V.anb[l], sizeof(double) * (size_t)M->lv[l].nnz);
  mgval_free_d(M, &V);
  return 0;
}

or_mg *or_case_mg(or_case *c) { return case_mg(c); }

/* multigrid-preconditioned conjugate gradients on the global pressure-
   correction system ap*phi - sum(anb*phi_nb) = b (cells 1..ne, the mesh
   CSR; halo values of phi stay as they are and enter with their anb, which
   calc_coef_p leaves at zero), the stop rule of solve_gs */
void or_mgcg_solve(or_case *c, const char *name, double *phi, int fp32) {
  const or_mg *M = case_mg(c);
  const or_mglevel *L0 = &M->lv[0];
  const int ne = c->ne, nnz = L0->nnz, nit = c->nit;
  const double *ap = c->ap, *anb = c->anb, *b = c->b;
  double *r = malloc(sizeof(double) * (size_t)ne), *z = malloc(sizeof(double) * (size_t)ne);
  double *p = malloc(sizeof(double) * (size_t)ne), *q = malloc(sizeof(double) * (size_t)ne);
  mgval_d Vd;
  mgval_f Vf;
  if (fp32) {
    mgval_alloc_f(M, &Vf);
    for (int i = 0; i < ne; i++) Vf.ap[0][i] = (float)ap[i];
    for (int k = 0; k < nnz; k++) Vf.anb[0][k] = (float)anb[k];
    mg_galerkin_f(M, &Vf);
  } else {
    mgval_alloc_d(M, &Vd);
    memcpy(Vd.ap[0], ap, sizeof(double) * (size_t)ne);
    memcpy(Vd.anb[0], anb, sizeof(double) * (size_t)nnz);
    mg_galerkin_d(M, &Vd);
  }
  /* r = b + sum anb phi_nb - ap phi (halos of phi included) */
  for (int i = 0; i < ne; i++) {
    double s = b[i];
    for (int k = L0->row[i]; k < L0->row[i + 1]; k++)
      s = s + anb[k] * phi[PK_E(c->ef2nb[k]) - 1];
    r[i] = s - ap[i] * phi[i];
  }
  const double res_i = sqrt(or_tree_dot(r, NULL, ne) / (double)ne);
  const double target = res_i / 10.0;
  double res_f = res_i, res_max = 0.0, rho_old = 0.0;
  int it = 0;
  int active = (0 < nit && res_f > target);
  while (active) {
    if (fp32) {
      for (int i = 0; i < ne; i++) Vf.b[0][i] = (float)r[i];
      mg_vcycle_f(M, &Vf, 0);
      for (int i = 0; i < ne; i++) z[i] = (double)Vf.x[0][i];
    } else {
      memcpy(Vd.b[0], r, sizeof(double) * (size_t)ne);
      mg_vcycle_d(M, &Vd, 0);
      memcpy(z, Vd.x[0], sizeof(double) * (size_t)ne);
    }
    const double rho = or_tree_dot(r, z, ne);
    if (it == 0) {
      for (int i = 0; i < ne; i++) p[i] = z[i];
    } else {
      const double beta = rho / rho_old;
      for (int i = 0; i < ne; i++) p[i] = z[i] + beta * p[i];
    }
    for (int i = 0; i < ne; i++) {
      double s = 0.0;
      for (int k = L0->row[i]; k < L0->row[i + 1]; k++) {
        const int j = L0->col[k];
        s = s + anb[k] * (j >= 0 ? p[j] : 0.0);
      }
      q[i] = ap[i] * p[i] - s;
    }
    const double sigma = or_tree_dot(p, q, ne);
    if (!(sigma > 0.0) || !(rho > 0.0)) break;
    const double alpha = rho / sigma;
    for (int i = 0; i < ne; i++) {
      phi[i] = phi[i] + alpha * p[i];
      r[i] = r[i] - alpha * q[i];
    }
    it = it + 1;
    res_f = sqrt(or_tree_dot(r, NULL, ne) / (double)ne);
    res_max = tree_absmax(r, ne);
    rho_old = rho;
    active = (it < nit && res_f > target);
  }
  or_log_line(c, name, it, res_i, res_f, res_max);
  if (fp32)
    mgval_free_f(M, &Vf);
  else
    mgval_free_d(M, &Vd);
  free(r);
  free(z);
  free(p);
  free(q);
}
