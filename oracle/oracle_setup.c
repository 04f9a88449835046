/*
 * oracle_setup.c -- restatement of the reference's mesh setup: face
 * connectivity, face/cell geometry, the oct-tree bisection partition, the
 * subdomain-local order and the subdomain systems.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, not the product).
 *
 * Parity: pinned against the reference compiled from its own sources
 * (oracle/_ref, driver in oracle/refdrv) on uniform and jittered boxes, and
 * against closed forms (closed cells, unit volume, nf = (6ne+nbf)/2).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PK(e, f) (((e) << 5) | (f))
#define PK_E(p) ((int)((unsigned)(p) >> 5))
#define PK_F(p) ((p) & 31)

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* CGNS hexahedron faces, src/modules/mod_util.f90:78-85 */
static const int HEX_FACES[6][4] = {{1, 4, 3, 2}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                    {3, 4, 8, 7}, {1, 5, 8, 4}, {5, 6, 7, 8}};

static int elem_nface(int t) { return t == OR_HEXA_8 ? 6 : (t == OR_QUAD_4 ? 1 : -1); }
static int elem_nvx(int t) { return t == OR_HEXA_8 ? 8 : (t == OR_QUAD_4 ? 4 : -1); }

/* face_vxlist, src/modules/mod_util.f90:1362-1426 (hex and quad only) */
static int face_vxlist(const int *ev, int t, int f, int *lst) {
  if (t == OR_HEXA_8) {
    for (int l = 0; l < 4; l++) lst[l] = ev[HEX_FACES[f - 1][l] - 1];
    return 4;
  }
  for (int l = 0; l < 4; l++) lst[l] = ev[l];
  return 4;
}

/* face_compare, src/modules/mod_util.f90:1450-1481 */
static int face_compare(int e1, int f1, int t1, const int *ev1, int e2, int f2, int t2,
                        const int *ev2) {
  int l1[4], l2[4];
  if (e1 == e2) return 0;
  int n1 = face_vxlist(ev1, t1, f1, l1), n2 = face_vxlist(ev2, t2, f2, l2);
  if (n1 != n2) return 0;
  int cnt = 0;
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++)
      if (l1[i] == l2[j]) { cnt++; break; }
  return cnt == n1;
}

or_case *or_case_new(int nvx, const double *x, const double *y, const double *z,
                     int nsec, const int *etype, const int *esec, const int *e2vx,
                     const int *bckind) {
  or_case *c = xcalloc(1, sizeof(or_case));
  c->nvx = nvx;
  c->nsec = nsec;
  c->x = xcalloc(nvx, 8); memcpy(c->x, x, 8 * (size_t)nvx);
  c->y = xcalloc(nvx, 8); memcpy(c->y, y, 8 * (size_t)nvx);
  c->z = xcalloc(nvx, 8); memcpy(c->z, z, 8 * (size_t)nvx);
  c->etype = xcalloc(nsec, 4); memcpy(c->etype, etype, 4 * (size_t)nsec);
  c->esec = xcalloc(2 * nsec, 4); memcpy(c->esec, esec, 8 * (size_t)nsec);
  c->bckind = xcalloc(nsec, 4); memcpy(c->bckind, bckind, 4 * (size_t)nsec);
  int ne = 0, nbf = 0, nfc = 0;
  /* cell_input.f90:55-72 */
  for (int s = 0; s < nsec; s++) {
    int cnt = esec[2 * s + 1] - esec[2 * s] + 1;
    if (etype[s] >= 10 && etype[s] <= 20) { ne += cnt; nfc += elem_nface(etype[s]) * cnt; }
    else nbf += cnt;
  }
  c->ne = ne; c->nbf = nbf; c->nelem = ne + nbf; c->nf = (nfc + nbf) / 2;
  c->L = 2 * c->nf - nbf;
  c->e2vx = xcalloc((size_t)8 * c->nelem, 4);
  memcpy(c->e2vx, e2vx, (size_t)32 * c->nelem);
  c->nlog = 0; c->caplog = 0; c->log = NULL;
  return c;
}

static int section_of(const or_case *c, int e) {
  for (int s = 0; s < c->nsec; s++)
    if (e < c->esec[2 * s + 1] + 1) return s;
  return c->nsec - 1;
}

/* find_element_nb, src/setup/mod_mg_lvl_uns.f90:283-488 */
int or_connectivity(or_case *c) {
  int ne = c->ne, nvx = c->nvx, nelem = c->nelem;
  c->ef2nb_idx = xcalloc(ne + 1, 4);
  c->ef2nb = xcalloc(2 * (size_t)c->L, 4);
  c->s2g = xcalloc(c->nf + 1, 4);
  c->bs = xcalloc(c->nbf, 4);
  int *g1 = c->ef2nb, *g2 = c->ef2nb + c->L;
  int *vx2e_idx = xcalloc(nvx + 2, 4);
  int fmax = 0;
  c->ef2nb_idx[0] = 1;
  for (int s = 0; s < c->nsec; s++)
    for (int e = c->esec[2 * s]; e <= c->esec[2 * s + 1]; e++) {
      int t = c->etype[s];
      if (e < ne + 1) c->ef2nb_idx[e] = c->ef2nb_idx[e - 1] + elem_nface(t);
      if (elem_nface(t) > fmax) fmax = elem_nface(t);
      for (int j = 0; j < elem_nvx(t); j++) vx2e_idx[c->e2vx[8 * (e - 1) + j] - 1]++;
    }
  /* exclusive prefix sum (1-based starts), mod_mg_lvl_uns.f90:316-324 */
  int tmp = vx2e_idx[0], emax = tmp;
  vx2e_idx[0] = 1;
  for (int v = 2; v <= nvx + 1; v++) {
    int tmp2 = vx2e_idx[v - 1];
    if (tmp2 > emax) emax = tmp2;
    vx2e_idx[v - 1] = vx2e_idx[v - 2] + tmp;
    tmp = tmp2;
  }
  int nv2e = vx2e_idx[nvx] - 1;
  int *vx2e = xcalloc(nv2e + 1, 4);
  for (int s = 0; s < c->nsec; s++)
    for (int e = c->esec[2 * s]; e <= c->esec[2 * s + 1]; e++)
      for (int j = 0; j < elem_nvx(c->etype[s]); j++) {
        int v = c->e2vx[8 * (e - 1) + j];
        int i1 = vx2e_idx[v - 1];
        while (vx2e[i1 - 1] != 0 && i1 < vx2e_idx[v]) i1++;
        vx2e[i1 - 1] = e;
      }
  int cap = emax * fmax;
  int *indices = xcalloc(cap + 1, 4), *f2s = xcalloc(cap + 1, 4);
  int nf = 0;
  for (int v = 1; v <= nvx; v++) {
    int n = 0;
    for (int i1 = vx2e_idx[v - 1]; i1 <= vx2e_idx[v] - 1; i1++) {
      int e = vx2e[i1 - 1];
      int s = section_of(c, e), t = c->etype[s];
      for (int f = 1; f <= elem_nface(t); f++) {
        int lst[4];
        int nl = face_vxlist(c->e2vx + 8 * (e - 1), t, f, lst);
        for (int l = 0; l < nl; l++)
          if (lst[l] == v) { indices[n] = PK(e, f); f2s[n] = s; n++; }
      }
    }
    for (int l = 0; l < n; l++)
      for (int m = l + 1; m < n; m++) {
        int t1 = c->etype[f2s[l]], t2 = c->etype[f2s[m]];
        int e1 = PK_E(indices[l]), f1 = PK_F(indices[l]);
        int e2 = PK_E(indices[m]), f2 = PK_F(indices[m]);
        if (!face_compare(e1, f1, t1, c->e2vx + 8 * (e1 - 1), e2, f2, t2, c->e2vx + 8 * (e2 - 1)))
          continue;
        if (e1 > ne && e2 <= ne) {
          int i2 = c->ef2nb_idx[e2 - 1] + f2 - 1;
          if (g1[i2 - 1] != 0) continue;
          nf++;
          g1[i2 - 1] = PK(e1, 0); g2[i2 - 1] = nf;
          c->s2g[nf - 1] = indices[m];
          c->bs[e1 - ne - 1] = PK(e2, f2);
        } else if (e2 > ne && e1 <= ne) {
          int i1 = c->ef2nb_idx[e1 - 1] + f1 - 1;
          if (g1[i1 - 1] != 0) continue;
          nf++;
          g1[i1 - 1] = PK(e2, 0); g2[i1 - 1] = nf;
          c->s2g[nf - 1] = indices[l];
          c->bs[e2 - ne - 1] = PK(e1, f1);
        } else if (e1 <= ne && e2 <= ne) {
          int i1 = c->ef2nb_idx[e1 - 1] + f1 - 1, i2 = c->ef2nb_idx[e2 - 1] + f2 - 1;
          if (g1[i1 - 1] != 0 && g1[i2 - 1] != 0) continue;
          nf++;
          g1[i1 - 1] = indices[m]; g1[i2 - 1] = indices[l];
          g2[i1 - 1] = nf; c->s2g[nf - 1] = indices[l];
          if (e1 > e2) { g2[i1 - 1] = -nf; c->s2g[nf - 1] = indices[m]; }
          g2[i2 - 1] = -g2[i1 - 1];
        }
      }
  }
  /* sign flag on halos of edge/corner cells, mod_mg_lvl_uns.f90:422-433 */
  for (int e1 = ne + 1; e1 <= nelem; e1++) {
    if (c->bs[e1 - ne - 1] < 0) continue;
    int e = PK_E(c->bs[e1 - ne - 1]), f = PK_F(c->bs[e1 - ne - 1]);
    for (int idx = c->ef2nb_idx[e - 1]; idx <= c->ef2nb_idx[e] - 1; idx++) {
      if (idx - c->ef2nb_idx[e - 1] + 1 == f) continue;
      int en = PK_E(g1[idx - 1]), fn = PK_F(g1[idx - 1]);
      if (fn == 0) {
        c->bs[e1 - ne - 1] = -abs(c->bs[e1 - ne - 1]);
        c->bs[en - ne - 1] = -c->bs[en - ne - 1];
      }
    }
  }
  free(indices); free(f2s); free(vx2e); free(vx2e_idx);
  if (nf != c->nf) {
    fprintf(stderr, "oracle: face count mismatch %d vs %d\n", nf, c->nf);
    return 1;
  }
  return 0;
}

/* calc_aip_xyzip_uns, src/setup/calc_aip_xyzip.f90:7-72 (quad faces) */
static void cross(double *a, const double *b, const double *cc) {
  a[0] = b[1] * cc[2] - b[2] * cc[1];
  a[1] = b[2] * cc[0] - b[0] * cc[2];
  a[2] = b[0] * cc[1] - b[1] * cc[0];
}

void or_geometry(or_case *c) {
  int nf = c->nf, ne = c->ne, nelem = c->nelem;
  c->aip = xcalloc(3 * (size_t)nf, 8);
  c->rip = xcalloc(3 * (size_t)nf, 8);
  c->xc = xcalloc(nelem, 8); c->yc = xcalloc(nelem, 8); c->zc = xcalloc(nelem, 8);
  c->vol = xcalloc(ne, 8);
  for (int fg = 1; fg <= nf; fg++) {
    int e = PK_E(c->s2g[fg - 1]), fl = PK_F(c->s2g[fg - 1]);
    int t = c->etype[section_of(c, e)];
    int lst[4];
    face_vxlist(c->e2vx + 8 * (e - 1), t, fl, lst);
    double r[4][3];
    for (int l = 0; l < 4; l++) {
      r[l][0] = c->x[lst[l] - 1]; r[l][1] = c->y[lst[l] - 1]; r[l][2] = c->z[lst[l] - 1];
    }
    double dr1[3], dr2[3], av[3], sub[3], sc[3][3], *aip = c->aip + 3 * (fg - 1);
    for (int i = 0; i < 3; i++) { dr1[i] = r[1][i] - r[0][i]; dr2[i] = r[2][i] - r[0][i]; }
    cross(av, dr1, dr2);
    for (int i = 0; i < 3; i++) { av[i] = 0.5 * av[i]; aip[i] = av[i]; }
    for (int i = 0; i < 3; i++) sub[i] = (r[0][i] + r[1][i] + r[2][i]) / 3.0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) sc[i][j] = sub[i] * av[j];
    for (int i = 0; i < 3; i++) { dr1[i] = r[2][i] - r[0][i]; dr2[i] = r[3][i] - r[0][i]; }
    cross(av, dr1, dr2);
    for (int i = 0; i < 3; i++) { av[i] = 0.5 * av[i]; aip[i] = aip[i] + av[i]; }
    for (int i = 0; i < 3; i++) sub[i] = (r[0][i] + r[2][i] + r[3][i]) / 3.0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) sc[i][j] = sc[i][j] + sub[i] * av[j];
    double dd = aip[0] * aip[0] + aip[1] * aip[1] + aip[2] * aip[2];
    for (int i = 0; i < 3; i++) {
      double m = sc[i][0] * aip[0] + sc[i][1] * aip[1] + sc[i][2] * aip[2];
      c->rip[3 * (fg - 1) + i] = m / dd;
    }
  }
  /* calc_vol_cv_centers_uns, src/setup/calc_vol_cv_centers.f90:4-59 */
  const int *g1 = c->ef2nb, *g2 = c->ef2nb + c->L;
  for (int s = 0; s < c->nsec; s++)
    for (int e = c->esec[2 * s]; e <= c->esec[2 * s + 1]; e++) {
      if (e > ne) continue;
      int nv = elem_nvx(c->etype[s]);
      double gc[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < nv; k++) {
        int v = c->e2vx[8 * (e - 1) + k];
        gc[0] = gc[0] + c->x[v - 1]; gc[1] = gc[1] + c->y[v - 1]; gc[2] = gc[2] + c->z[v - 1];
      }
      for (int i = 0; i < 3; i++) gc[i] = gc[i] / (double)nv;
      double sum_vol = 0.0, rc[3] = {0.0, 0.0, 0.0};
      for (int idx = c->ef2nb_idx[e - 1]; idx <= c->ef2nb_idx[e] - 1; idx++) {
        int gf = g2[idx - 1];
        double sg = gf >= 0 ? 1.0 : -1.0;
        gf = abs(gf);
        const double *cs = c->rip + 3 * (gf - 1), *a = c->aip + 3 * (gf - 1);
        int enb = PK_E(g1[idx - 1]), lf = PK_F(g1[idx - 1]);
        if (lf == 0) { c->xc[enb - 1] = cs[0]; c->yc[enb - 1] = cs[1]; c->zc[enb - 1] = cs[2]; }
        double h[3], sc3[3];
        for (int i = 0; i < 3; i++) h[i] = cs[i] - gc[i];
        double sub_vol = ((sg * a[0]) * h[0] + (sg * a[1]) * h[1] + (sg * a[2]) * h[2]) / 3.0;
        for (int i = 0; i < 3; i++) sc3[i] = 0.25 * gc[i] + 0.75 * cs[i];
        sum_vol = sum_vol + sub_vol;
        for (int i = 0; i < 3; i++) rc[i] = rc[i] + sc3[i] * sub_vol;
      }
      for (int i = 0; i < 3; i++) rc[i] = rc[i] / sum_vol;
      c->xc[e - 1] = rc[0]; c->yc[e - 1] = rc[1]; c->zc[e - 1] = rc[2];
      c->vol[e - 1] = sum_vol;
    }
}

/* ------------------------------------------------------------------------ */
/* Oct-tree bisection seeds, src/setup/mod_agglomeration.f90:331-545.        */
/* The reference keeps leaves in an in-order threaded binary tree and element */
/* lists in push order; both orders decide sums and leaf numbering, so they  */
/* are reproduced here with the same threading rules.                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  int parent_group_id, group_id;
  double rmin[3], rmax[3], group_vol;
  int *list, len;
} bb_t;

typedef struct tnode {
  int id, lthread, rthread;
  struct tnode *left, *right;
  bb_t *val;
} tnode;

typedef struct {
  int ncv, nleaf, target_ncv, last_id;
  double vol_ave;
  const double *vol, *xc, *yc, *zc;
  tnode *root, *begin, *end;
  tnode **all; int nall, capall;
} seedgen;

static tnode *new_node(seedgen *g, int id, bb_t *val) {
  tnode *n = xcalloc(1, sizeof(tnode));
  n->id = id; n->val = val; n->lthread = 1; n->rthread = 1;
  if (g->nall == g->capall) {
    g->capall = g->capall ? 2 * g->capall : 16;
    g->all = realloc(g->all, sizeof(tnode *) * g->capall);
  }
  g->all[g->nall++] = n;
  return n;
}

static int is_leaf(const tnode *n) { return n->lthread && n->rthread; }

/* iterate_next, mod_agglomeration.f90:184-198 */
static int iterate_next(seedgen *g, tnode **it) {
  if (*it == g->end) return 0;
  if (!(*it)->rthread) {
    *it = (*it)->right;
    while (!(*it)->lthread) *it = (*it)->left;
  } else {
    *it = (*it)->right;
  }
  return 1;
}

/* goto_next_leaf, mod_agglomeration.f90:216-227 */
static int goto_next_leaf(seedgen *g, tnode **it) {
  int ok = 0;
  while (iterate_next(g, it)) {
    ok = is_leaf(*it);
    if (ok) break;
  }
  return ok;
}

/* goto_1st_leaf, mod_agglomeration.f90:270-279 */
static int goto_1st_leaf(seedgen *g, tnode **it) {
  *it = g->begin;
  if (!g->root) return 0;
  if (!is_leaf(*it)) return goto_next_leaf(g, it);
  return 1;
}

/* split_node, mod_agglomeration.f90:141-172 */
static void split_node(seedgen *g, tnode *n, bb_t *lv, bb_t *rv) {
  tnode *left = n->left, *right = n->right;
  if (lv) {
    n->lthread = 0;
    tnode *L = new_node(g, ++g->last_id, lv);
    n->left = L; L->left = left; L->right = n;
    if (n == g->begin) g->begin = L;
  }
  if (rv) {
    n->rthread = 0;
    tnode *R = new_node(g, ++g->last_id, rv);
    n->right = R; R->left = n; R->right = right;
    if (n == g->end) g->end = R;
  }
}

static bb_t *new_bb(int cap) {
  bb_t *b = xcalloc(1, sizeof(bb_t));
  b->parent_group_id = 0; b->group_id = -1;
  b->list = xcalloc(cap, 4);
  return b;
}

/* split_leaf, mod_agglomeration.f90:386-495 */
static int split_leaf(seedgen *g, tnode *it, double thr, int *nchild) {
  *nchild = 0;
  if (!is_leaf(it)) return 0;
  bb_t *b = it->val;
  if (b->group_vol / g->vol_ave < thr) return 0;
  if (b->len == 1) return 0;
  double d = 0.0;
  int isplit = 0;
  for (int i = 0; i < 3; i++)
    if (fabs(b->rmin[i] - b->rmax[i]) > d) { d = fabs(b->rmin[i] - b->rmax[i]); isplit = i + 1; }
  if (isplit == 0) { fprintf(stderr, "oracle: degenerate bounding box\n"); abort(); }
  d = (b->rmin[isplit - 1] + b->rmax[isplit - 1]) / 2.0;
  const double *pos = isplit == 1 ? g->xc : (isplit == 2 ? g->yc : g->zc);
  bb_t *L = new_bb(b->len), *R = new_bb(b->len);
  double lmin[3] = {1e20, 1e20, 1e20}, lmax[3] = {-1e20, -1e20, -1e20};
  double rmin[3] = {1e20, 1e20, 1e20}, rmax[3] = {-1e20, -1e20, -1e20};
  double vl = 0.0, vr = 0.0;
  for (int k = 0; k < b->len; k++) {
    int e = b->list[k];
    double r[3] = {g->xc[e - 1], g->yc[e - 1], g->zc[e - 1]};
    if (pos[e - 1] < d) {
      L->list[L->len++] = e;
      vl = vl + g->vol[e - 1] * 1.0;
      for (int i = 0; i < 3; i++) { lmin[i] = fmin(lmin[i], r[i]); lmax[i] = fmax(lmax[i], r[i]); }
    } else {
      R->list[R->len++] = e;
      vr = vr + g->vol[e - 1] * 1.0;
      for (int i = 0; i < 3; i++) { rmin[i] = fmin(rmin[i], r[i]); rmax[i] = fmax(rmax[i], r[i]); }
    }
  }
  b->len = 0;
  L->group_vol = vl; R->group_vol = vr;
  memcpy(L->rmin, lmin, 24); memcpy(L->rmax, lmax, 24);
  memcpy(R->rmin, rmin, 24); memcpy(R->rmax, rmax, 24);
  int pg = b->group_id == -1 ? b->parent_group_id : b->group_id;
  L->parent_group_id = pg; R->parent_group_id = pg;
  if (L->len > 0 && R->len > 0) { *nchild = 2; split_node(g, it, L, R); }
  else if (L->len > 0) { *nchild = 1; split_node(g, it, L, NULL); free(R->list); free(R); }
  else { *nchild = 1; split_node(g, it, NULL, R); free(L->list); free(L); }
  return 1;
}

/* grow, mod_agglomeration.f90:497-545 */
static void grow(seedgen *g, int new_target, int *gf2g, int lvl) {
  int nleaf0 = g->nleaf, nchild;
  double thr = 2.0;
  g->vol_ave = g->vol_ave * (double)g->target_ncv / (double)new_target;
  g->target_ncv = new_target;
  tnode *it;
  if (lvl > 1) {
    while (g->nleaf < new_target) {
      if (goto_1st_leaf(g, &it)) {
        for (;;) {
          if (split_leaf(g, it, thr, &nchild)) {
            goto_next_leaf(g, &it);
            g->nleaf = g->nleaf + nchild - 1;
            if (g->nleaf == new_target) break;
          }
          if (!goto_next_leaf(g, &it)) break;
        }
      }
      if (g->nleaf == nleaf0) thr = thr * 0.75;
      nleaf0 = g->nleaf;
    }
    int nleaf = 0;
    if (goto_1st_leaf(g, &it)) {
      for (;;) {
        nleaf++;
        bb_t *leaf = it->val;
        if (leaf->group_id != -1) leaf->parent_group_id = leaf->group_id;
        leaf->group_id = nleaf;
        gf2g[nleaf - 1] = leaf->parent_group_id;
        if (!goto_next_leaf(g, &it)) break;
      }
    }
  } else {
    if (goto_1st_leaf(g, &it)) {
      for (;;) {
        bb_t *leaf = it->val;
        for (int k = 0; k < leaf->len; k++) gf2g[leaf->list[k] - 1] = leaf->group_id;
        leaf->len = 0;
        if (!goto_next_leaf(g, &it)) break;
      }
    }
  }
}

/* qsort_key_nRec, src/modules/mod_util.f90:1683-1730: the non-recursive   */
/* quicksort published by Darel Rex Finley, pivot = first key of the range. */
/* A range whose keys are all equal is left untouched by that algorithm, so */
/* it is skipped here without changing the output permutation.             */
void or_qsort_key_nrec(int *key, int *b, int n) {
  enum { MAXLV = 1000 };
  int beg[MAXLV], end[MAXLV];
  int i = 0;
  beg[0] = 0; end[0] = n;
  while (i >= 0) {
    int L = beg[i], R = end[i] - 1;
    int alleq = 1;
    for (int k = L + 1; k <= R && alleq; k++) alleq = key[k] == key[L];
    if (L < R && !alleq) {
      int piv = key[L], b0 = b[L];
      while (L < R) {
        while (key[R] >= piv && L < R) R--;
        if (L < R) { key[L] = key[R]; b[L] = b[R]; L++; }
        while (key[L] <= piv && L < R) L++;
        if (L < R) { key[R] = key[L]; b[R] = b[L]; R--; }
      }
      key[L] = piv; b[L] = b0;
      beg[i + 1] = L + 1; end[i + 1] = end[i]; end[i] = L;
      i++;
      if (end[i] - beg[i] > end[i - 1] - beg[i - 1]) {
        int t = beg[i]; beg[i] = beg[i - 1]; beg[i - 1] = t;
        t = end[i]; end[i] = end[i - 1]; end[i - 1] = t;
      }
    } else {
      i--;
    }
  }
}

/* generate_seeds + add_transformation_bt (first level),                    */
/* src/setup/mod_mg_lvl_uns.f90:95-118 and :873-895                         */
int or_partition(or_case *c, int nsubd) {
  int ne = c->ne;
  c->nsubd = nsubd;
  c->gf2g = xcalloc(ne, 4);
  c->g2gf = xcalloc(ne, 4);
  c->g2gf_idx = xcalloc(nsubd + 1, 4);
  if (nsubd <= 1) {
    for (int e = 0; e < ne; e++) { c->gf2g[e] = 1; c->g2gf[e] = e + 1; }
    c->g2gf_idx[0] = 1; c->g2gf_idx[1] = ne + 1;
    return 0;
  }
  seedgen g;
  memset(&g, 0, sizeof(g));
  g.ncv = ne; g.nleaf = 1; g.target_ncv = 1; g.last_id = 1;
  g.vol = c->vol; g.xc = c->xc; g.yc = c->yc; g.zc = c->zc;
  bb_t *rb = new_bb(ne);
  double mn[3] = {c->xc[0], c->yc[0], c->zc[0]}, mx[3] = {c->xc[0], c->yc[0], c->zc[0]};
  for (int e = 0; e < ne; e++) {
    double r[3] = {c->xc[e], c->yc[e], c->zc[e]};
    for (int i = 0; i < 3; i++) { if (r[i] < mn[i]) mn[i] = r[i]; if (r[i] > mx[i]) mx[i] = r[i]; }
  }
  memcpy(rb->rmin, mn, 24); memcpy(rb->rmax, mx, 24);
  g.vol_ave = 0.0;
  for (int e = 1; e <= ne; e++) { rb->list[rb->len++] = e; g.vol_ave = g.vol_ave + c->vol[e - 1] * 1.0; }
  rb->group_vol = 2.1 * g.vol_ave;
  rb->group_id = 1;
  g.root = new_node(&g, 1, rb);
  g.root->left = g.root; g.root->right = g.root;
  g.begin = g.root; g.end = g.root;
  int *gf2g2 = xcalloc(nsubd, 4);
  grow(&g, nsubd, gf2g2, 2);
  grow(&g, ne, c->gf2g, 1);
  free(gf2g2);
  for (int k = 0; k < g.nall; k++) {
    if (g.all[k]->val) { free(g.all[k]->val->list); free(g.all[k]->val); }
    free(g.all[k]);
  }
  free(g.all);
  int *key = xcalloc(ne, 4);
  for (int e = 0; e < ne; e++) { key[e] = c->gf2g[e]; c->g2gf[e] = e + 1; }
  or_qsort_key_nrec(key, c->g2gf, ne);
  int ngrp = key[ne - 1], gcur = 0;
  if (ngrp != nsubd) { free(key); return 1; }
  for (int idx = 1; idx <= ne; idx++)
    if (gcur != key[idx - 1]) { gcur = key[idx - 1]; c->g2gf_idx[gcur - 1] = idx; }
  c->g2gf_idx[nsubd] = ne + 1;
  free(key);
  return 0;
}

/* construct_subdomains, src/modules/mod_subdomains.f90:18-160 */
static void qsort_key_rec(int *key, int *b, int i, int f) {
  /* qsort_key, mod_util.f90:1602-1625; keys here are distinct, so any */
  /* correct sort gives the same result, the recursion is kept anyway. */
  if (i >= f) return;
  int n = i, p = f;
  while (n < p) {
    if (key[n] > key[p]) {
      int t = key[n]; key[n] = key[p]; key[p] = t;
      t = b[n]; b[n] = b[p]; b[p] = t;
      p--;
      if (p > n) {
        t = key[n]; key[n] = key[p]; key[p] = t;
        t = b[n]; b[n] = b[p]; b[p] = t;
      }
    } else {
      n++;
    }
  }
  qsort_key_rec(key, b, i, p - 1);
  qsort_key_rec(key, b, p + 1, f);
}

void or_construct_subdomains(or_case *c) {
  int ns = c->nsubd;
  const int *g1 = c->ef2nb;
  c->sd = xcalloc(ns, sizeof(or_subd));
  c->intf = xcalloc((size_t)ns * ns, sizeof(or_intf));
  int **tmpidx = xcalloc((size_t)ns * ns, sizeof(int *));
  int *ef2nb_tmp = xcalloc(c->L, 4);
#define IF(a, b) c->intf[(a) * ns + (b)]
  for (int cc = 0; cc < ns; cc++) {
    or_subd *s = &c->sd[cc];
    s->ne = c->g2gf_idx[cc + 1] - c->g2gf_idx[cc];
    int nf = 0, nbf = 0;
    for (int n = c->g2gf_idx[cc]; n <= c->g2gf_idx[cc + 1] - 1; n++) {
      int gf = c->g2gf[n - 1];
      for (int idx = c->ef2nb_idx[gf - 1]; idx <= c->ef2nb_idx[gf] - 1; idx++) {
        int gfnb = PK_E(g1[idx - 1]), lfnb = PK_F(g1[idx - 1]);
        int cnb = 0;
        if (lfnb != 0) cnb = c->gf2g[gfnb - 1];
        if (cnb == 0) { IF(cc, cc).ncs++; nbf++; }
        else if (cc + 1 != cnb) { IF(cc, cnb - 1).ncs++; nbf++; }
        nf++;
      }
    }
    nf = (nf + nbf) / 2;
    s->nbf = nbf; s->nf = nf;
    s->ap = xcalloc(s->ne, 8); s->b = xcalloc(s->ne, 8);
    s->anb = xcalloc(2 * nf - nbf, 8);
    s->phic = xcalloc(s->ne + nbf, 8);
    s->ef2nb = xcalloc(2 * nf - nbf, 4);
    s->ef2nb_idx = xcalloc(s->ne + 1, 4);
    for (int cnb = 0; cnb < ns; cnb++) {
      IF(cc, cnb).c1 = cc + 1; IF(cc, cnb).c2 = cnb + 1;
      IF(cc, cnb).index1 = xcalloc(IF(cc, cnb).ncs + 1, 4);
      if (cc != cnb) tmpidx[cc * ns + cnb] = xcalloc(IF(cc, cnb).ncs + 1, 4);
    }
  }
  for (int cc = 0; cc < ns; cc++)
    for (int cnb = 0; cnb < ns; cnb++) {
      if (cc != cnb) IF(cc, cnb).index2 = IF(cnb, cc).index1;
      IF(cc, cnb).ncs = 0;
    }
  for (int cc = 0; cc < ns; cc++) {
    or_subd *s = &c->sd[cc];
    int nbf = 0;
    s->ef2nb_idx[0] = 1;
    for (int n = c->g2gf_idx[cc]; n <= c->g2gf_idx[cc + 1] - 1; n++) {
      int g = n - c->g2gf_idx[cc] + 1, gf = c->g2gf[n - 1];
      int nl = c->ef2nb_idx[gf] - c->ef2nb_idx[gf - 1];
      s->ef2nb_idx[g] = s->ef2nb_idx[g - 1] + nl;
      for (int idx = c->ef2nb_idx[gf - 1]; idx <= c->ef2nb_idx[gf] - 1; idx++) {
        int lf = idx - c->ef2nb_idx[gf - 1] + 1;
        int gfnb = PK_E(g1[idx - 1]), lfnb = PK_F(g1[idx - 1]);
        int cnb = 0;
        if (lfnb != 0) cnb = c->gf2g[gfnb - 1];
        if (cc + 1 == cnb) {
          int idx2 = c->ef2nb_idx[gfnb - 1] + lfnb - 1;
          ef2nb_tmp[idx2 - 1] = PK(g, lf);
        } else {
          nbf++;
          ef2nb_tmp[idx - 1] = PK(s->ne + nbf, 0);
        }
      }
    }
  }
  for (int cc = 0; cc < ns; cc++) {
    or_subd *s = &c->sd[cc];
    for (int n = c->g2gf_idx[cc]; n <= c->g2gf_idx[cc + 1] - 1; n++) {
      int g = n - c->g2gf_idx[cc] + 1, gf = c->g2gf[n - 1];
      for (int idx = c->ef2nb_idx[gf - 1]; idx <= c->ef2nb_idx[gf] - 1; idx++) {
        int lf = idx - c->ef2nb_idx[gf - 1] + 1;
        int gfnb = PK_E(g1[idx - 1]), lfnb = PK_F(g1[idx - 1]);
        int cnb = 0;
        if (lfnb != 0) cnb = c->gf2g[gfnb - 1];
        int idx2 = s->ef2nb_idx[g - 1] + lf - 1;
        s->ef2nb[idx2 - 1] = ef2nb_tmp[idx - 1];
        if (cnb == 0) {
          or_intf *t = &IF(cc, cc);
          t->ncs++;
          t->index1[t->ncs - 1] = PK(g, lf);
        } else if (cc + 1 != cnb) {
          or_intf *t = &IF(cc, cnb - 1);
          t->ncs++;
          t->index1[t->ncs - 1] = PK(g, lf);
          if (cc + 1 < cnb) tmpidx[cc * ns + cnb - 1][t->ncs - 1] = PK(gfnb, lfnb);
          else tmpidx[cc * ns + cnb - 1][t->ncs - 1] = PK(gf, lf);
        }
      }
    }
  }
  for (int cc = 0; cc < ns; cc++)
    for (int cnb = 0; cnb < ns; cnb++) {
      if (cc == cnb) continue;
      qsort_key_rec(tmpidx[cc * ns + cnb], IF(cc, cnb).index1, 0, IF(cc, cnb).ncs - 1);
      free(tmpidx[cc * ns + cnb]);
    }
#undef IF
  free(tmpidx);
  free(ef2nb_tmp);
}

void or_init_physics(or_case *c, double rho, double mu, double dt, int nit) {
  int ne = c->ne, n = c->ne + c->nbf, nf = c->nf;
  c->dt = dt; c->nit = nit;
  c->alpha_u = 1.0; c->alpha_p = 1.0;
  c->inlet[0] = 1.0; c->inlet[1] = 0.0; c->inlet[2] = 1.0; c->inlet[3] = 0.0; c->inlet[4] = 1.0;
  c->rho = xcalloc(ne, 8); c->mu = xcalloc(ne, 8);
  for (int e = 0; e < ne; e++) { c->rho[e] = rho; c->mu[e] = mu; }
  double **f1[] = {&c->u, &c->v, &c->w, &c->p, &c->u0, &c->v0, &c->w0, &c->phic};
  for (int k = 0; k < 8; k++) *f1[k] = xcalloc(n, 8);
  double **f3[] = {&c->gu, &c->gv, &c->gw, &c->gp, &c->gpc};
  for (int k = 0; k < 5; k++) *f3[k] = xcalloc(3 * (size_t)n, 8);
  c->mip = xcalloc(nf, 8); c->mip0 = xcalloc(nf, 8);
  double **f0[] = {&c->bu, &c->bv, &c->bw, &c->d, &c->dc, &c->ap, &c->b};
  for (int k = 0; k < 7; k++) *f0[k] = xcalloc(ne, 8);
  c->anb = xcalloc(c->L, 8);
  /* thermal properties of init_properties, mod_properties.f90:86-89 */
  c->cp = xcalloc(ne, 8); c->tc = xcalloc(ne, 8);
  for (int e = 0; e < ne; e++) { c->cp[e] = 1000.0; c->tc[e] = 5.0; }
  c->t = xcalloc(n, 8); c->h = xcalloc(n, 8); c->h0 = xcalloc(n, 8);
  c->gt = xcalloc(3 * (size_t)n, 8); c->gh = xcalloc(3 * (size_t)n, 8);
}

void or_case_free(or_case *c) {
  if (!c) return;
  void *ptrs[] = {c->x, c->y, c->z, c->e2vx, c->etype, c->esec, c->bckind, c->ef2nb_idx,
                  c->ef2nb, c->s2g, c->bs, c->xc, c->yc, c->zc, c->vol, c->aip, c->rip,
                  c->gf2g, c->g2gf, c->g2gf_idx, c->rho, c->mu, c->u, c->v, c->w, c->p,
                  c->u0, c->v0, c->w0, c->gu, c->gv, c->gw, c->gp, c->gpc, c->mip, c->mip0,
                  c->bu, c->bv, c->bw, c->d, c->dc, c->ap, c->anb, c->b, c->phic, c->log,
                  c->cp, c->tc, c->t, c->gt, c->h, c->h0, c->gh};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
  or_set_pc_order(c, 0);
  or_set_gs_order(c, 0);
  or_mg_free(c->mg);
  if (c->sd) {
    for (int k = 0; k < c->nsubd; k++) {
      or_subd *s = &c->sd[k];
      free(s->ef2nb_idx); free(s->ef2nb); free(s->ap); free(s->b); free(s->anb); free(s->phic);
    }
    free(c->sd);
  }
  if (c->intf) {
    for (int k = 0; k < c->nsubd * c->nsubd; k++) free(c->intf[k].index1);
    free(c->intf);
  }
  free(c);
}
