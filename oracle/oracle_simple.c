/*
 * oracle_simple.c -- restatement of one SIMPLE(C) outer iteration of the
 * reference: momentum assembly, the Gauss-Seidel solves, least-squares
 * gradients, Rhie-Chow face fluxes, pressure-correction assembly, the
 * 4-subdomain block solver and the correction step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, not the product).
 *
 * Each expression keeps the reference's evaluation order (Fortran evaluates
 * a*b/c as (a*b)/c, a-b-c as (a-b)-c); compile without FMA contraction.
 * Parity: pinned bitwise against the reference compiled from its own
 * sources (oracle/_ref) and against closed forms in the CPU tests.  The
 * extension modes (config-mode relaxation, the colour order and the
 * mixed-precision pc solve) have no reference counterpart: they are pinned
 * by their reduction to the reference (relaxation 1), by the colouring rule,
 * and by the mixed mode's distance from the FP64 colour order.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PK_E(p) ((int)((unsigned)(p) >> 5))
#define PK_F(p) ((p) & 31)

static double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static void log_line(or_case *c, const char *name, int it, double ri, double rf, double rm) {
  if (!c) return;
  if (c->nlog == c->caplog) {
    c->caplog = c->caplog ? 2 * c->caplog : 64;
    c->log = realloc(c->log, sizeof(or_logline) * c->caplog);
  }
  or_logline *l = &c->log[c->nlog++];
  memset(l, 0, sizeof(*l));
  strncpy(l->name, name, sizeof(l->name) - 1);
  l->it = it; l->res_i = ri; l->res_f = rf; l->res_max = rm;
}

static int bc_kind_of_halo(const or_case *c, int enb);

void or_log_line(or_case *c, const char *name, int it, double ri, double rf, double rm) {
  log_line(c, name, it, ri, rf, rm);
}

/* vec_weight, src/modules/mod_util.f90:729-743 */
double or_vec_weight(const double *r0, const double *r1, const double *r2) {
  double ra[3], rb[3];
  for (int i = 0; i < 3; i++) { ra[i] = r1[i] - r0[i]; rb[i] = r2[i] - r0[i]; }
  double la = sqrt(ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2]);
  double lb = sqrt(rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2]);
  if (la + lb > 0.0) return la / (la + lb);
  return 0.5;
}

/* matinv3, src/modules/mod_solver.f90:8-38 (A, B column-major) */
void or_matinv3(const double *A, double *B) {
#define a(i, j) A[((j)-1) * 3 + (i)-1]
#define bb(i, j) B[((j)-1) * 3 + (i)-1]
  double det = a(1, 1) * a(2, 2) * a(3, 3) - a(1, 1) * a(2, 3) * a(3, 2) -
               a(1, 2) * a(2, 1) * a(3, 3) + a(1, 2) * a(2, 3) * a(3, 1) +
               a(1, 3) * a(2, 1) * a(3, 2) - a(1, 3) * a(2, 2) * a(3, 1);
  if (fabs(det) > 2.2250738585072014e-308) {
    double di = 1.0 / det;
    bb(1, 1) = +di * (a(2, 2) * a(3, 3) - a(2, 3) * a(3, 2));
    bb(2, 1) = -(di * (a(2, 1) * a(3, 3) - a(2, 3) * a(3, 1)));
    bb(3, 1) = +di * (a(2, 1) * a(3, 2) - a(2, 2) * a(3, 1));
    bb(1, 2) = -(di * (a(1, 2) * a(3, 3) - a(1, 3) * a(3, 2)));
    bb(2, 2) = +di * (a(1, 1) * a(3, 3) - a(1, 3) * a(3, 1));
    bb(3, 2) = -(di * (a(1, 1) * a(3, 2) - a(1, 2) * a(3, 1)));
    bb(1, 3) = +di * (a(1, 2) * a(2, 3) - a(1, 3) * a(2, 2));
    bb(2, 3) = -(di * (a(1, 1) * a(2, 3) - a(1, 3) * a(2, 1)));
    bb(3, 3) = +di * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1));
  } else {
    for (int k = 0; k < 9; k++) B[k] = 0.0;
    det = a(1, 1) + a(2, 2) + a(3, 3);
    double di = 1.0 / det;
    bb(1, 1) = di; bb(2, 2) = di; bb(3, 3) = di;
  }
#undef a
#undef bb
}

/* calc_grad, src/modules/mod_solver.f90:40-81 */
void or_calc_grad(const double *phi, double *grad, const double *xc, const double *yc,
                  const double *zc, const int *ef2nb_idx, const int *ef2nb,
                  int ne, int nf, int nbf) {
  (void)nf; (void)nbf;
  for (int e = 1; e <= ne; e++) {
    double A[9] = {0}, g[3] = {0.0, 0.0, 0.0};
    for (int idx = ef2nb_idx[e - 1]; idx <= ef2nb_idx[e] - 1; idx++) {
      int enb = PK_E(ef2nb[idx - 1]);
      double dr[3] = {xc[enb - 1] - xc[e - 1], yc[enb - 1] - yc[e - 1], zc[enb - 1] - zc[e - 1]};
      double wt = 1.0 / (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      double dphi = phi[enb - 1] - phi[e - 1];
      for (int i = 0; i < 3; i++) g[i] = g[i] + wt * dphi * dr[i];
      A[0] = A[0] + wt * dr[0] * dr[0];
      A[3] = A[3] + wt * dr[0] * dr[1];
      A[6] = A[6] + wt * dr[0] * dr[2];
      A[4] = A[4] + wt * dr[1] * dr[1];
      A[7] = A[7] + wt * dr[1] * dr[2];
      A[8] = A[8] + wt * dr[2] * dr[2];
    }
    A[1] = A[3]; A[2] = A[6]; A[5] = A[7];
    double Ai[9];
    or_matinv3(A, Ai);
    for (int i = 0; i < 3; i++)
      grad[3 * (e - 1) + i] = Ai[i] * g[0] + Ai[3 + i] * g[1] + Ai[6 + i] * g[2];
  }
}

static double gs_sum(const double *phi, const double *anb, const double *b,
                     const int *ef2nb_idx, const int *ef2nb, int e) {
  double s = b[e - 1];
  for (int idx = ef2nb_idx[e - 1]; idx <= ef2nb_idx[e] - 1; idx++)
    s = s + anb[idx - 1] * phi[PK_E(ef2nb[idx - 1]) - 1];
  return s;
}

static double sor_of(const char *name) {
  return (name[0] == 'p' && name[1] == 'c') ? 1.02 : 1.0;
}

/* smoother_gs, src/modules/mod_solver.f90:191-228 */
void or_smoother_gs(const char *name, double *phi, const double *ap, const double *anb,
                    const double *b, const int *ef2nb_idx, const int *ef2nb,
                    int ne, int nf, int nbf, int nit) {
  (void)nf; (void)nbf;
  double sor = sor_of(name);
  for (int it = 1; it <= nit; it++) {
    for (int e = 1; e <= ne; e++) {
      double s = gs_sum(phi, anb, b, ef2nb_idx, ef2nb, e);
      phi[e - 1] = (s + (sor - 1.0) * ap[e - 1] * phi[e - 1]) / ap[e - 1] / sor;
    }
    for (int e = ne; e >= 1; e--) {
      double s = gs_sum(phi, anb, b, ef2nb_idx, ef2nb, e);
      phi[e - 1] = (s + (sor - 1.0) * ap[e - 1] * phi[e - 1]) / ap[e - 1] / sor;
    }
  }
}

/* calc_residual, src/modules/mod_solver.f90:230-253 (signed max, Q2) */
void or_calc_residual(const double *phi, const double *ap, const double *anb,
                      const double *b, const int *ef2nb_idx, const int *ef2nb,
                      int ne, int nf, int nbf, double *res, double *res_max) {
  (void)nf; (void)nbf;
  double r2 = 0.0, rm = 0.0;
  for (int e = 1; e <= ne; e++) {
    double r = gs_sum(phi, anb, b, ef2nb_idx, ef2nb, e) - ap[e - 1] * phi[e - 1];
    rm = fmax(r, rm);
    r2 = r2 + r * r;
  }
  *res = sqrt(r2 / (double)ne);
  *res_max = rm;
}

/* Colour order of solve_gs (an extension, no counterpart in the reference):
   a greedy colouring of the global system in global order, a cell taking the
   least colour none of its earlier interior neighbours has; gs_perm lists
   the cells by (colour, global index). */
void or_set_gs_order(or_case *c, int order) {
  free(c->gs_perm);
  c->gs_perm = NULL;
  c->gs_order = order;
  if (order != 1) return;
  int ne = c->ne, ncol = 0;
  int *col = calloc((size_t)ne, sizeof(int));
  for (int e = 1; e <= ne; e++) {
    unsigned long long used = 0;
    for (int idx = c->ef2nb_idx[e - 1]; idx <= c->ef2nb_idx[e] - 1; idx++) {
      int d = PK_E(c->ef2nb[idx - 1]);
      if (d <= ne && d < e && col[d - 1] < 64) used |= 1ull << col[d - 1];
    }
    int k = 0;
    while (k < 64 && (used >> k & 1ull)) k++;
    col[e - 1] = k;
    if (k + 1 > ncol) ncol = k + 1;
  }
  c->gs_perm = malloc((size_t)ne * sizeof(int));
  int n = 0;
  for (int k = 0; k < ncol; k++)
    for (int e = 1; e <= ne; e++)
      if (col[e - 1] == k) c->gs_perm[n++] = e;
  free(col);
}

/* solve_gs, src/modules/mod_solver.f90:255-327 (with gs_order 1 the sweeps
   visit the cells in gs_perm order forward and reversed backward) */
void or_solve_gs(or_case *c, const char *name, double *phi, const double *ap,
                 const double *anb, const double *b, const int *ef2nb_idx,
                 const int *ef2nb, int ne, int nf, int nbf, int nit) {
  (void)nf; (void)nbf;
  const int *perm = (c && c->gs_order == 1 && ne == c->ne) ? c->gs_perm : NULL;
  double sor = sor_of(name);
  double res_i = 0.0;
  for (int e = 1; e <= ne; e++) {
    double r = gs_sum(phi, anb, b, ef2nb_idx, ef2nb, e) - ap[e - 1] * phi[e - 1];
    res_i = res_i + r * r;
  }
  res_i = sqrt(res_i / (double)ne);
  int it = 0;
  double res_f = res_i, res_max = 0.0, target = res_i / 10.0;
  while (it < nit && res_f > target) {
    it++;
    for (int k = 1; k <= ne; k++) {
      int e = perm ? perm[k - 1] : k;
      double s = gs_sum(phi, anb, b, ef2nb_idx, ef2nb, e);
      phi[e - 1] = (s + (sor - 1.0) * ap[e - 1] * phi[e - 1]) / ap[e - 1] / sor;
    }
    for (int k = ne; k >= 1; k--) {
      int e = perm ? perm[k - 1] : k;
      double s = gs_sum(phi, anb, b, ef2nb_idx, ef2nb, e);
      phi[e - 1] = (s + (sor - 1.0) * ap[e - 1] * phi[e - 1]) / ap[e - 1] / sor;
    }
    res_max = 0.0;
    res_f = 0.0;
    for (int e = 1; e <= ne; e++) {
      double r = fabs(gs_sum(phi, anb, b, ef2nb_idx, ef2nb, e) - ap[e - 1] * phi[e - 1]);
      res_max = fmax(r, res_max);
      res_f = res_f + r * r;
    }
    res_f = sqrt(res_f / (double)ne);
  }
  log_line(c, name, it, res_i, res_f, res_max);
}

/* assemble_coef, src/modules/mod_subdomains.f90:162-189 */
static void assemble_coef(or_case *c, const double *phi) {
  const int *g1 = c->ef2nb;
  for (int cc = 0; cc < c->nsubd; cc++) {
    or_subd *s = &c->sd[cc];
    for (int e = 1; e <= s->ne; e++) {
      int g = c->g2gf[c->g2gf_idx[cc] + e - 2];
      s->ap[e - 1] = c->ap[g - 1];
      s->b[e - 1] = c->b[g - 1];
      s->phic[e - 1] = phi[g - 1];
      for (int idx = s->ef2nb_idx[e - 1]; idx <= s->ef2nb_idx[e] - 1; idx++) {
        int enb = PK_E(s->ef2nb[idx - 1]), lfnb = PK_F(s->ef2nb[idx - 1]);
        int lf = idx - s->ef2nb_idx[e - 1] + 1;
        int idx2 = c->ef2nb_idx[g - 1] + lf - 1;
        s->anb[idx - 1] = c->anb[idx2 - 1];
        if (lfnb == 0) s->phic[enb - 1] = phi[PK_E(g1[idx2 - 1]) - 1];
      }
    }
  }
}

/* update_halos, src/modules/mod_subdomains.f90:191-212 */
static void update_halos(or_intf *t, or_subd *s1, or_subd *s2) {
  if (t->c1 == t->c2) return;
  for (int cs = 0; cs < t->ncs; cs++) {
    int g1 = PK_E(t->index1[cs]), lf1 = PK_F(t->index1[cs]);
    int gnb1 = PK_E(s1->ef2nb[s1->ef2nb_idx[g1 - 1] + lf1 - 2]);
    int g2 = PK_E(t->index2[cs]), lf2 = PK_F(t->index2[cs]);
    int gnb2 = PK_E(s2->ef2nb[s2->ef2nb_idx[g2 - 1] + lf2 - 2]);
    s1->phic[gnb1 - 1] = s2->phic[g2 - 1];
    s2->phic[gnb2 - 1] = s1->phic[g1 - 1];
  }
}

/* smoother_gs (src/modules/mod_solver.f90:191-228) visiting the cells in the
   order perm[0..ne-1] (1-based cells) forward and in reverse backward: the
   same update, a different sequential order (colour order, extension; the
   colours come from a greedy colouring in global cell order) */
static void smoother_gs_perm(const char *name, double *phi, const double *ap, const double *anb,
                             const double *b, const int *ef2nb_idx, const int *ef2nb, int ne,
                             const int *perm, int nit) {
  double sor = sor_of(name);
  for (int it = 1; it <= nit; it++) {
    for (int k = 0; k < ne; k++) {
      int e = perm[k];
      double s = gs_sum(phi, anb, b, ef2nb_idx, ef2nb, e);
      phi[e - 1] = (s + (sor - 1.0) * ap[e - 1] * phi[e - 1]) / ap[e - 1] / sor;
    }
    for (int k = ne - 1; k >= 0; k--) {
      int e = perm[k];
      double s = gs_sum(phi, anb, b, ef2nb_idx, ef2nb, e);
      phi[e - 1] = (s + (sor - 1.0) * ap[e - 1] * phi[e - 1]) / ap[e - 1] / sor;
    }
  }
}

/* qsort order of int pairs by their first element (distinct here) */
static int cmp_first(const void *a, const void *b) {
  int x = ((const int *)a)[0], y = ((const int *)b)[0];
  return (x > y) - (x < y);
}

void or_set_pc_order(or_case *c, int order) {
  if (c->pc_perm) {
    for (int k = 0; k < c->nsubd; k++) free(c->pc_perm[k]);
    free(c->pc_perm);
    c->pc_perm = NULL;
  }
  c->pc_order = order;
  if (order != 1 && order != 2) return;
  c->pc_perm = calloc((size_t)c->nsubd, sizeof(int *));
  for (int cc = 0; cc < c->nsubd; cc++) {
    const or_subd *s = &c->sd[cc];
    /* global index of each local cell, and the local cells in global order */
    const int *gl = c->g2gf + c->g2gf_idx[cc] - 1;
    int *col = calloc((size_t)s->ne, sizeof(int));
    int ncol = 0;
    int (*byg)[2] = malloc((size_t)s->ne * sizeof(*byg)); /* (global index, local cell) */
    for (int e = 1; e <= s->ne; e++) { byg[e - 1][0] = gl[e - 1]; byg[e - 1][1] = e; }
    qsort(byg, (size_t)s->ne, sizeof(*byg), cmp_first);
    for (int i = 0; i < s->ne; i++) {
      int e = byg[i][1];
      unsigned long long used = 0; /* a hex cell has at most 6 neighbours */
      for (int idx = s->ef2nb_idx[e - 1]; idx <= s->ef2nb_idx[e] - 1; idx++) {
        int d = PK_E(s->ef2nb[idx - 1]);
        if (d <= s->ne && gl[d - 1] < gl[e - 1] && col[d - 1] < 64) used |= 1ull << col[d - 1];
      }
      int k = 0;
      while (k < 64 && (used >> k & 1ull)) k++;
      col[e - 1] = k;
      if (k + 1 > ncol) ncol = k + 1;
    }
    free(byg);
    int *perm = malloc((size_t)s->ne * sizeof(int)), n = 0;
    for (int k = 0; k < ncol; k++)
      for (int e = 1; e <= s->ne; e++)
        if (col[e - 1] == k) perm[n++] = e;
    free(col);
    c->pc_perm[cc] = perm;
  }
}

/* Mixed-precision colour order (an extension, no counterpart in the
   reference): smoother_gs_perm's update in single precision on the
   correction system ap*d - sum(anb*d_nb) = r, cells in perm order forward
   and reversed backward, sor 1.02f, IEEE single divisions. */
static void smoother_gs_perm32(float *d, const float *a, const float *anb, const float *r,
                               const int *ef2nb_idx, const int *ef2nb, int ne, const int *perm,
                               int nit) {
  const float sor = 1.02f;
  for (int it = 1; it <= nit; it++) {
    for (int dir = 0; dir < 2; dir++) {
      for (int k = 0; k < ne; k++) {
        int e = perm[dir == 0 ? k : ne - 1 - k];
        float sum = r[e - 1];
        for (int idx = ef2nb_idx[e - 1]; idx <= ef2nb_idx[e] - 1; idx++)
          sum = sum + anb[idx - 1] * d[PK_E(ef2nb[idx - 1]) - 1];
        d[e - 1] = (sum + (sor - 1.0f) * a[e - 1] * d[e - 1]) / a[e - 1] / sor;
      }
    }
  }
}

/* update_halos (mod_subdomains.f90:191-212) on the FP32 corrections */
static void update_halos32(const or_intf *t, const or_subd *s1, const or_subd *s2, float *d1,
                           float *d2) {
  if (t->c1 == t->c2) return;
  for (int cs = 0; cs < t->ncs; cs++) {
    int g1 = PK_E(t->index1[cs]), lf1 = PK_F(t->index1[cs]);
    int gnb1 = PK_E(s1->ef2nb[s1->ef2nb_idx[g1 - 1] + lf1 - 2]);
    int g2 = PK_E(t->index2[cs]), lf2 = PK_F(t->index2[cs]);
    int gnb2 = PK_E(s2->ef2nb[s2->ef2nb_idx[g2 - 1] + lf2 - 2]);
    d1[gnb1 - 1] = d2[g2 - 1];
    d2[gnb2 - 1] = d1[g1 - 1];
  }
}

/* The FP64 residual of calc_residual (norms) together with each cell's
   r = b + sum(anb*phi_nb) - ap*phi rounded to FP32. */
static void residual_mixed(const or_subd *s, float *r32, double *res, double *rmax) {
  or_calc_residual(s->phic, s->ap, s->anb, s->b, s->ef2nb_idx, s->ef2nb, s->ne, s->nf, s->nbf,
                   res, rmax);
  for (int e = 1; e <= s->ne; e++)
    r32[e - 1] = (float)(gs_sum(s->phic, s->anb, s->b, s->ef2nb_idx, s->ef2nb, e) -
                         s->ap[e - 1] * s->phic[e - 1]);
}

/* multi_subdomain_solver (mod_solver.f90:124-189, quirks Q1-Q3 kept) with
   the mixed-precision inner solve: the two-sweep blocks, halo exchanges and
   residual checks of the reference loop, on an FP32 correction d; at every
   check phic += d on cells and halos, d = 0, and the FP64 residual of the
   updated phic gives the norms and the next r; once more after the loop. */
static void multi_subdomain_solver_mixed(or_case *c, const char *name, double *phi) {
  int ns = c->nsubd, nit = c->nit;
  assemble_coef(c, phi);
  float **a32 = calloc((size_t)ns, sizeof(float *)), **anb32 = calloc((size_t)ns, sizeof(float *));
  float **r32 = calloc((size_t)ns, sizeof(float *)), **d32 = calloc((size_t)ns, sizeof(float *));
  double res, rmax, ri = 0.0, rf = 0.0, rmt = 0.0;
  for (int cc = 0; cc < ns; cc++) {
    or_subd *s = &c->sd[cc];
    int L = s->ef2nb_idx[s->ne] - 1;
    a32[cc] = malloc((size_t)s->ne * sizeof(float));
    anb32[cc] = malloc((size_t)L * sizeof(float));
    r32[cc] = malloc((size_t)s->ne * sizeof(float));
    d32[cc] = calloc((size_t)(s->ne + s->nbf), sizeof(float));
    for (int e = 0; e < s->ne; e++) a32[cc][e] = (float)s->ap[e];
    for (int k = 0; k < L; k++) anb32[cc][k] = (float)s->anb[k];
    residual_mixed(s, r32[cc], &res, &rmax);
    ri = ri + res * res;
    rmt = fmax(rmt, rmax);
  }
  ri = sqrt(ri / (double)ns);
  int it = 0;
  rf = ri;
  double target = ri / 10.0;
  while (it < nit && rf > target) {
    for (int cc = 0; cc < ns; cc++) {
      or_subd *s = &c->sd[cc];
      smoother_gs_perm32(d32[cc], a32[cc], anb32[cc], r32[cc], s->ef2nb_idx, s->ef2nb, s->ne,
                         c->pc_perm[cc], 2);
    }
    it += 2;
    for (int i = 0; i < ns; i++)
      for (int j = i + 1; j < ns; j++)
        update_halos32(&c->intf[i * ns + j], &c->sd[i], &c->sd[j], d32[i], d32[j]);
    if (it % 10 == 0) {
      for (int cc = 0; cc < ns; cc++) {
        or_subd *s = &c->sd[cc];
        for (int i = 0; i < s->ne + s->nbf; i++) {
          s->phic[i] = s->phic[i] + (double)d32[cc][i];
          d32[cc][i] = 0.0f;
        }
        residual_mixed(s, r32[cc], &res, &rmax);
        rf = rf + res * res; /* Q1 */
        rmt = fmax(rmt, rmax);
      }
      rf = sqrt(rf / (double)ns);
    }
  }
  for (int cc = 0; cc < ns; cc++) {
    or_subd *s = &c->sd[cc];
    for (int i = 0; i < s->ne + s->nbf; i++) s->phic[i] = s->phic[i] + (double)d32[cc][i];
    free(a32[cc]);
    free(anb32[cc]);
    free(r32[cc]);
    free(d32[cc]);
  }
  free(a32);
  free(anb32);
  free(r32);
  free(d32);
  log_line(c, name, it, ri, rf, rmt);
  for (int cc = 0; cc < ns; cc++) {
    or_subd *s = &c->sd[cc];
    for (int e = 1; e <= s->ne; e++) phi[c->g2gf[c->g2gf_idx[cc] + e - 2] - 1] = s->phic[e - 1];
  }
}

/* multi_subdomain_solver, src/modules/mod_solver.f90:124-189 (quirks Q1-Q3) */
static void multi_subdomain_solver(or_case *c, const char *name, double *phi) {
  int ns = c->nsubd, nit = c->nit;
  if (c->pc_order == 2) {
    multi_subdomain_solver_mixed(c, name, phi);
    return;
  }
  assemble_coef(c, phi);
  double res, rmax, ri = 0.0, rf = 0.0, rmt = 0.0;
  for (int cc = 0; cc < ns; cc++) {
    or_subd *s = &c->sd[cc];
    or_calc_residual(s->phic, s->ap, s->anb, s->b, s->ef2nb_idx, s->ef2nb, s->ne, s->nf, s->nbf,
                     &res, &rmax);
    ri = ri + res * res;
    rmt = fmax(rmt, rmax);
  }
  ri = sqrt(ri / (double)ns);
  int it = 0;
  rf = ri;
  double target = ri / 10.0;
  while (it < nit && rf > target) {
    for (int cc = 0; cc < ns; cc++) {
      or_subd *s = &c->sd[cc];
      if (c->pc_order == 1)
        smoother_gs_perm(name, s->phic, s->ap, s->anb, s->b, s->ef2nb_idx, s->ef2nb, s->ne,
                         c->pc_perm[cc], 2);
      else
        or_smoother_gs(name, s->phic, s->ap, s->anb, s->b, s->ef2nb_idx, s->ef2nb, s->ne,
                       s->nf, s->nbf, 2);
    }
    it += 2;
    for (int i = 0; i < ns; i++)
      for (int j = i + 1; j < ns; j++)
        update_halos(&c->intf[i * ns + j], &c->sd[i], &c->sd[j]);
    if (it % 10 == 0) {
      for (int cc = 0; cc < ns; cc++) {
        or_subd *s = &c->sd[cc];
        or_calc_residual(s->phic, s->ap, s->anb, s->b, s->ef2nb_idx, s->ef2nb, s->ne, s->nf,
                         s->nbf, &res, &rmax);
        rf = rf + res * res; /* Q1: accumulates onto the previous value */
        rmt = fmax(rmt, rmax);
      }
      rf = sqrt(rf / (double)ns);
    }
  }
  log_line(c, name, it, ri, rf, rmt);
  /* update_phi, src/modules/mod_subdomains.f90:214-229 */
  for (int cc = 0; cc < ns; cc++) {
    or_subd *s = &c->sd[cc];
    for (int e = 1; e <= s->ne; e++) phi[c->g2gf[c->g2gf_idx[cc] + e - 2] - 1] = s->phic[e - 1];
  }
}

/* solve, src/modules/mod_solver.f90:329-344 */
void or_solve(or_case *c, const char *name, double *phi) {
  if (c->pc_order == 3 || c->pc_order == 4)
    or_mgcg_solve(c, name, phi, c->pc_order == 4);
  else if (c->nsubd <= 1)
    or_solve_gs(c, name, phi, c->ap, c->anb, c->b, c->ef2nb_idx, c->ef2nb, c->ne, c->nf, c->nbf,
                c->nit);
  else
    multi_subdomain_solver(c, name, phi);
}

/* BC setters lid / dirichlet0 / symmetry, src/equations/mod_uvwp.f90:493-570, */
/* called for every boundary section in order by update_boundaries            */
/* (src/modules/mod_physics.f90:38-50).                                        */
void or_update_boundaries(or_case *c) {
  const int *g2 = c->ef2nb + c->L;
  for (int s = 0; s < c->nsec; s++) {
    if (c->etype[s] >= 10) continue;
    int kind = c->bckind[s];
    for (int e = c->esec[2 * s]; e <= c->esec[2 * s + 1]; e++) {
      int pk = abs(c->bs[e - c->ne - 1]);
      int enb = PK_E(pk), lfnb = PK_F(pk);
      if (kind == OR_BC_SYMMETRY) {
        int fg = g2[c->ef2nb_idx[enb - 1] + lfnb - 2];
        const double *a = c->aip + 3 * (fg - 1);
        double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
        double vel[3] = {c->u[enb - 1], c->v[enb - 1], c->w[enb - 1]}, vn[3], vt[3];
        double t = dot3(vel, nrm);
        for (int i = 0; i < 3; i++) { vn[i] = t * nrm[i]; vt[i] = vel[i] - vn[i]; }
        c->u[e - 1] = vt[0] - 2.0 * vn[0];
        c->v[e - 1] = vt[1] - 2.0 * vn[1];
        c->w[e - 1] = vt[2] - 2.0 * vn[2];
        c->p[e - 1] = c->p[enb - 1];
        c->mip[fg - 1] = 0.0;
      } else if (kind == OR_BC_INLET || kind == OR_BC_OUTLET) {
        /* channel extension: inlet velocity along the inward normal with the
           parabolic profile, outlet velocity and pressure from the owner;
           the face mass flux rho_P (v.n) A (the outlet's is rescaled below) */
        int fg = g2[c->ef2nb_idx[enb - 1] + lfnb - 2];
        const double *a = c->aip + 3 * (fg - 1);
        double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
        double vel[3];
        if (kind == OR_BC_INLET) {
          const double *q = c->inlet;
          double sy = (c->yc[e - 1] - q[1]) / (q[2] - q[1]);
          double sz = (c->zc[e - 1] - q[3]) / (q[4] - q[3]);
          double un = q[0] * 16.0 * sy * (1.0 - sy) * sz * (1.0 - sz);
          for (int i = 0; i < 3; i++) vel[i] = un * -nrm[i];
        } else {
          vel[0] = c->u[enb - 1];
          vel[1] = c->v[enb - 1];
          vel[2] = c->w[enb - 1];
        }
        c->u[e - 1] = vel[0];
        c->v[e - 1] = vel[1];
        c->w[e - 1] = vel[2];
        c->p[e - 1] = c->p[enb - 1];
        c->mip[fg - 1] = dot3(vel, nrm) * c->rho[enb - 1] * area;
      } else {
        c->u[e - 1] = kind == OR_BC_LID ? 1.0 : 0.0;
        c->v[e - 1] = 0.0;
        c->w[e - 1] = 0.0;
        c->p[e - 1] = c->p[enb - 1];
        int fg = g2[c->ef2nb_idx[enb - 1] + lfnb - 2];
        c->mip[fg - 1] = 0.0;
      }
    }
  }
  or_scale_outlet(c);
}

/* Outlet mass flux of the channel extension: the extrapolated fluxes are
   scaled so that the outflow equals the inflow, scale = -Q_in / Q_out with
   both sums (inlet, then outlet halos in halo order) in the blocked tree
   order of or_tree_sum; if Q_out is not positive the inflow is spread by
   face area instead, -Q_in * A_f / A_out.  The pc system stays Neumann with
   zero net boundary flux. */
void or_scale_outlet(or_case *c) {
  const int *g2 = c->ef2nb + c->L;
  int nin = 0, nout = 0;
  int *fin = malloc(sizeof(int) * (size_t)(c->nbf + 1)), *fout = malloc(sizeof(int) * (size_t)(c->nbf + 1));
  for (int h = 0; h < c->nbf; h++) {
    const int kind = bc_kind_of_halo(c, c->ne + 1 + h);
    if (kind != OR_BC_INLET && kind != OR_BC_OUTLET) continue;
    const int pk = abs(c->bs[h]);
    const int fg = g2[c->ef2nb_idx[PK_E(pk) - 1] + PK_F(pk) - 2];
    if (kind == OR_BC_INLET) fin[nin++] = fg; else fout[nout++] = fg;
  }
  if (nout > 0) {
    double *v = malloc(sizeof(double) * (size_t)(nin > nout ? nin : nout));
    double *ar = malloc(sizeof(double) * (size_t)nout);
    for (int k = 0; k < nin; k++) v[k] = c->mip[fin[k] - 1];
    const double qin = or_tree_sum(v, nin);
    for (int k = 0; k < nout; k++) {
      const double *a = c->aip + 3 * (fout[k] - 1);
      v[k] = c->mip[fout[k] - 1];
      ar[k] = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    }
    const double qout = or_tree_sum(v, nout), aout = or_tree_sum(ar, nout);
    if (qout > 0.0) {
      const double sc = -qin / qout;
      for (int k = 0; k < nout; k++) c->mip[fout[k] - 1] = c->mip[fout[k] - 1] * sc;
    } else {
      for (int k = 0; k < nout; k++) c->mip[fout[k] - 1] = -qin * ar[k] / aout;
    }
    free(v);
    free(ar);
  }
  free(fin);
  free(fout);
}

static int bc_kind_of_halo(const or_case *c, int enb) {
  for (int s = 0; s < c->nsec; s++)
    if (enb >= c->esec[2 * s] && enb <= c->esec[2 * s + 1]) return c->bckind[s];
  return OR_BC_WALL;
}

/* calc_coef_uvw, src/equations/mod_uvwp.f90:161-286 */
void or_calc_coef_uvw(or_case *c) {
  int ne = c->ne;
  const int *g1 = c->ef2nb, *g2 = c->ef2nb + c->L;
  for (int e = 1; e <= ne; e++) {
    double rp[3] = {c->xc[e - 1], c->yc[e - 1], c->zc[e - 1]};
    c->ap[e - 1] = 0.0;
    double sumf = 0.0, sumss[3] = {0, 0, 0}, sumdefc[3] = {0, 0, 0};
    for (int idx = c->ef2nb_idx[e - 1]; idx <= c->ef2nb_idx[e] - 1; idx++) {
      c->anb[idx - 1] = 0.0;
      int enb = PK_E(g1[idx - 1]), lfnb = PK_F(g1[idx - 1]);
      double f = 0.0, fnb = 0.0, d = 0.0;
      if (lfnb > 0) {
        int fg = g2[idx - 1];
        double sg = fg >= 0 ? 1.0 : -1.0;
        fg = abs(fg);
        const double *a = c->aip + 3 * (fg - 1), *rip = c->rip + 3 * (fg - 1);
        double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double nrm[3], rpnb[3], dr[3], drip[3], rpp[3], rnp[3], drp[3];
        for (int i = 0; i < 3; i++) nrm[i] = sg * a[i] / area;
        rpnb[0] = c->xc[enb - 1]; rpnb[1] = c->yc[enb - 1]; rpnb[2] = c->zc[enb - 1];
        for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
        double ds = sqrt(dot3(dr, dr));
        double wt = or_vec_weight(rip, rp, rpnb);
        for (int i = 0; i < 3; i++) drip[i] = rip[i] - rp[i];
        double t = dot3(drip, nrm);
        for (int i = 0; i < 3; i++) rpp[i] = rip[i] - t * nrm[i];
        for (int i = 0; i < 3; i++) drip[i] = rip[i] - rpnb[i];
        t = dot3(drip, nrm);
        for (int i = 0; i < 3; i++) rnp[i] = rip[i] - t * nrm[i];
        for (int i = 0; i < 3; i++) drp[i] = rnp[i] - rpp[i];
        double dsp = sqrt(dot3(drp, drp));
        f = -(sg * c->mip[fg - 1]);
        fnb = fmax(f, 0.0);
        sumf = sumf + f;
        double muip = (1.0 - wt) * c->mu[e - 1] + wt * c->mu[enb - 1];
        d = muip * area / ds;
        for (int m = 0; m < 3; m++) {
          double gip[3];
          gip[0] = (1.0 - wt) * c->gu[3 * (e - 1) + m] + wt * c->gu[3 * (enb - 1) + m];
          gip[1] = (1.0 - wt) * c->gv[3 * (e - 1) + m] + wt * c->gv[3 * (enb - 1) + m];
          gip[2] = (1.0 - wt) * c->gw[3 * (e - 1) + m] + wt * c->gw[3 * (enb - 1) + m];
          sumss[m] = sumss[m] + muip * area * dot3(gip, dr) / ds;
        }
        const double *gs[3] = {c->gu, c->gv, c->gw};
        for (int m = 0; m < 3; m++) {
          double gip[3];
          for (int i = 0; i < 3; i++)
            gip[i] = (1.0 - wt) * gs[m][3 * (e - 1) + i] + wt * gs[m][3 * (enb - 1) + i];
          sumdefc[m] = sumdefc[m] + muip * area * (dot3(gip, drp) / dsp - dot3(gip, dr) / ds);
        }
      }
      c->anb[idx - 1] = d + fnb;
      c->ap[e - 1] = c->ap[e - 1] + d + fnb;
    }
    double ap0 = c->rho[e - 1] * c->vol[e - 1] / c->dt;
    c->ap[e - 1] = c->ap[e - 1] + ap0;
    c->bu[e - 1] = ap0 * c->u0[e - 1] + sumf * c->u[e - 1] - c->vol[e - 1] * c->gp[3 * (e - 1)] +
                   sumss[0] + sumdefc[0];
    c->bv[e - 1] = ap0 * c->v0[e - 1] + sumf * c->v[e - 1] -
                   c->vol[e - 1] * c->gp[3 * (e - 1) + 1] + sumss[1] + sumdefc[1];
    c->bw[e - 1] = ap0 * c->w0[e - 1] + sumf * c->w[e - 1] -
                   c->vol[e - 1] * c->gp[3 * (e - 1) + 2] + sumss[2] + sumdefc[2];
  }
  for (int enb = ne + 1; enb <= ne + c->nbf; enb++) {
    int pk = abs(c->bs[enb - ne - 1]);
    int e = PK_E(pk), lf = PK_F(pk);
    int idx = c->ef2nb_idx[e - 1] + lf - 1;
    int fg = g2[idx - 1];
    const double *a = c->aip + 3 * (fg - 1);
    double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
    double dr[3] = {c->xc[enb - 1] - c->xc[e - 1], c->yc[enb - 1] - c->yc[e - 1],
                    c->zc[enb - 1] - c->zc[e - 1]};
    double ds = sqrt(dot3(dr, dr));
    double f = 0.0, d = c->mu[e - 1] * area / ds;
    const int kind = bc_kind_of_halo(c, enb);
    if (kind == OR_BC_INLET || kind == OR_BC_OUTLET) {
      /* channel extension: the face's convective flux as on interior faces
         (anb += max(f, 0) upwind, f * u_P into the right-hand side as the
         cell loop's sumf), diffusion d at the inlet and none at the
         zero-gradient outlet */
      const double fl = -c->mip[fg - 1], fnb = fmax(fl, 0.0);
      const double dd = kind == OR_BC_INLET ? d : 0.0;
      c->bu[e - 1] = c->bu[e - 1] + fl * c->u[e - 1];
      c->bv[e - 1] = c->bv[e - 1] + fl * c->v[e - 1];
      c->bw[e - 1] = c->bw[e - 1] + fl * c->w[e - 1];
      c->ap[e - 1] = c->ap[e - 1] + dd + fnb;
      c->anb[idx - 1] = c->anb[idx - 1] + dd + fnb;
      continue;
    }
    if (kind != OR_BC_SYMMETRY) {
      double vb[3] = {c->u[enb - 1], c->v[enb - 1], c->w[enb - 1]};
      double vr[3] = {c->u[e - 1], c->v[e - 1], c->w[e - 1]};
      double t = dot3(vr, nrm);
      for (int i = 0; i < 3; i++) vr[i] = vr[i] - t * nrm[i];
      for (int i = 0; i < 3; i++) vr[i] = vb[i] - vr[i];
      c->bu[e - 1] = c->bu[e - 1] + d * vr[0] - d * c->u[e - 1];
      c->bv[e - 1] = c->bv[e - 1] + d * vr[1] - d * c->v[e - 1];
      c->bw[e - 1] = c->bw[e - 1] + d * vr[2] - d * c->w[e - 1];
    }
    c->ap[e - 1] = c->ap[e - 1] + d + f;
    c->anb[idx - 1] = c->anb[idx - 1] + d + f;
  }
  /* config mode (no counterpart in the reference): Patankar implicit
     under-relaxation, ap -> ap/alpha_u and b += (ap/alpha_u - ap)*phi, applied
     before the Rhie-Chow / SIMPLEC coefficients; alpha_u = 1 is skipped so
     that reference mode stays bit-identical. */
  if (c->alpha_u != 1.0) {
    for (int e = 1; e <= ne; e++) {
      double apr = c->ap[e - 1] / c->alpha_u, da = apr - c->ap[e - 1];
      c->bu[e - 1] = c->bu[e - 1] + da * c->u[e - 1];
      c->bv[e - 1] = c->bv[e - 1] + da * c->v[e - 1];
      c->bw[e - 1] = c->bw[e - 1] + da * c->w[e - 1];
      c->ap[e - 1] = apr;
    }
  }
  for (int e = 1; e <= ne; e++) {
    c->d[e - 1] = c->vol[e - 1] / c->ap[e - 1];
    double dc = c->ap[e - 1];
    for (int idx = c->ef2nb_idx[e - 1]; idx <= c->ef2nb_idx[e] - 1; idx++) dc = dc - c->anb[idx - 1];
    c->dc[e - 1] = c->vol[e - 1] / dc;
  }
}

/* calc_coef_p, src/equations/mod_uvwp.f90:289-368 */
void or_calc_coef_p(or_case *c) {
  int ne = c->ne;
  const int *g1 = c->ef2nb, *g2 = c->ef2nb + c->L;
  for (int e = 1; e <= ne; e++) {
    c->ap[e - 1] = 0.0;
    double sumf = 0.0, rp[3] = {c->xc[e - 1], c->yc[e - 1], c->zc[e - 1]};
    for (int idx = c->ef2nb_idx[e - 1]; idx <= c->ef2nb_idx[e] - 1; idx++) {
      c->anb[idx - 1] = 0.0;
      int enb = PK_E(g1[idx - 1]), lfnb = PK_F(g1[idx - 1]);
      double d = 0.0;
      if (lfnb > 0) {
        int fg = g2[idx - 1];
        double sg = fg >= 0 ? 1.0 : -1.0;
        fg = abs(fg);
        const double *a = c->aip + 3 * (fg - 1), *rip = c->rip + 3 * (fg - 1);
        double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double nrm[3], rpnb[3] = {c->xc[enb - 1], c->yc[enb - 1], c->zc[enb - 1]}, dr[3];
        for (int i = 0; i < 3; i++) nrm[i] = sg * a[i] / area;
        for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
        double wt = or_vec_weight(rip, rp, rpnb);
        double f = -(sg * c->mip[fg - 1]);
        sumf = sumf + f;
        double rhoip = (1.0 - wt) * c->rho[e - 1] + wt * c->rho[enb - 1];
        d = ((1.0 - wt) * c->dc[e - 1] + wt * c->dc[enb - 1]) / dot3(dr, nrm) * rhoip * area;
      }
      c->anb[idx - 1] = d;
      c->ap[e - 1] = c->ap[e - 1] + d;
    }
    c->b[e - 1] = sumf;
  }
  for (int enb = ne + 1; enb <= ne + c->nbf; enb++) {
    int pk = abs(c->bs[enb - ne - 1]);
    int e = PK_E(pk), lf = PK_F(pk);
    int idx = c->ef2nb_idx[e - 1] + lf - 1;
    int fg = g2[idx - 1];
    double d = 0.0;
    c->ap[e - 1] = c->ap[e - 1] + d;
    c->anb[idx - 1] = c->anb[idx - 1] + d;
    c->b[e - 1] = c->b[e - 1] - c->mip[fg - 1];
  }
}

/* calc_mip, src/equations/mod_uvwp.f90:438-490 */
void or_calc_mip(or_case *c, int rc) {
  const int *g1 = c->ef2nb;
  for (int fg = 1; fg <= c->nf; fg++) {
    int e = PK_E(c->s2g[fg - 1]), lf = PK_F(c->s2g[fg - 1]);
    int idx = c->ef2nb_idx[e - 1] + lf - 1;
    int enb = PK_E(g1[idx - 1]), lfnb = PK_F(g1[idx - 1]);
    if (lfnb == 0) continue;
    double rp[3] = {c->xc[e - 1], c->yc[e - 1], c->zc[e - 1]};
    double rpnb[3] = {c->xc[enb - 1], c->yc[enb - 1], c->zc[enb - 1]};
    const double *a = c->aip + 3 * (fg - 1), *rip = c->rip + 3 * (fg - 1);
    double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
    double wt = or_vec_weight(rip, rp, rpnb);
    double v1[3] = {c->u[e - 1], c->v[e - 1], c->w[e - 1]};
    double v2[3] = {c->u[enb - 1], c->v[enb - 1], c->w[enb - 1]}, vip[3];
    for (int i = 0; i < 3; i++) vip[i] = (1.0 - wt) * v1[i] + wt * v2[i];
    double rhoip = c->rho[e - 1] * (1.0 - wt) + c->rho[enb - 1] * wt;
    c->mip[fg - 1] = dot3(vip, nrm) * rhoip * area;
    if (rc) {
      double dr[3], gpip[3], vip0[3];
      for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
      for (int i = 0; i < 3; i++)
        gpip[i] = (1.0 - wt) * c->gp[3 * (e - 1) + i] + wt * c->gp[3 * (enb - 1) + i];
      double dip = (1.0 - wt) * c->d[e - 1] + wt * c->d[enb - 1];
      double w1[3] = {c->u0[e - 1], c->v0[e - 1], c->w0[e - 1]};
      double w2[3] = {c->u0[enb - 1], c->v0[enb - 1], c->w0[enb - 1]};
      for (int i = 0; i < 3; i++) vip0[i] = (1.0 - wt) * w1[i] + wt * w2[i];
      c->mip[fg - 1] = c->mip[fg - 1] -
                       rhoip * area * dip / dot3(dr, nrm) *
                           (c->p[enb - 1] - c->p[e - 1] - dot3(gpip, dr)) -
                       rhoip / c->dt * dip * (c->mip0[fg - 1] - dot3(vip0, nrm) * rhoip * area);
    }
  }
}

/* adjust_pc, src/equations/mod_uvwp.f90:136-158 (zeroth-order halos) */
void or_adjust_pc(or_case *c, double *pc, double pref) {
  for (int e = 0; e < c->ne; e++) pc[e] = pc[e] - pref;
  for (int enb = c->ne + 1; enb <= c->ne + c->nbf; enb++)
    pc[enb - 1] = pc[PK_E(abs(c->bs[enb - c->ne - 1])) - 1];
}

/* update_uvwp, src/equations/mod_uvwp.f90:370-436 (no velocity correction) */
void or_update_uvwp(or_case *c, const double *pc) {
  const int *g1 = c->ef2nb;
  /* config mode: p and gp take alpha_p * correction (alpha_p = 1 multiplies
     by one exactly, i.e. the reference's p+pc, gp+gpc) */
  for (int e = 0; e < c->ne; e++) {
    c->p[e] = c->p[e] + c->alpha_p * pc[e];
    for (int i = 0; i < 3; i++)
      c->gp[3 * e + i] = c->gp[3 * e + i] + c->alpha_p * c->gpc[3 * e + i];
  }
  for (int fg = 1; fg <= c->nf; fg++) {
    int e = PK_E(c->s2g[fg - 1]), lf = PK_F(c->s2g[fg - 1]);
    int idx = c->ef2nb_idx[e - 1] + lf - 1;
    int enb = PK_E(g1[idx - 1]), lfnb = PK_F(g1[idx - 1]);
    if (lfnb == 0) continue;
    const double *a = c->aip + 3 * (fg - 1), *rip = c->rip + 3 * (fg - 1);
    double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
    double rp[3] = {c->xc[e - 1], c->yc[e - 1], c->zc[e - 1]};
    double rpnb[3] = {c->xc[enb - 1], c->yc[enb - 1], c->zc[enb - 1]}, dr[3];
    for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
    double wt = or_vec_weight(rip, rp, rpnb);
    double dip = (1.0 - wt) * c->dc[e - 1] + wt * c->dc[enb - 1];
    double rhoip = (c->rho[e - 1] + c->rho[enb - 1]) / 2.0;
    double dmip = rhoip * area * dip * (pc[enb - 1] - pc[e - 1]) / dot3(dr, nrm);
    c->mip[fg - 1] = c->mip[fg - 1] - dmip;
  }
}

/* solve_uvwp, src/equations/mod_uvwp.f90:95-134 */
void or_solve_uvwp(or_case *c) {
  int ne = c->ne, nf = c->nf, nbf = c->nbf;
  if (c->asm32)
    or_calc_coef_uvw32(c);
  else
    or_calc_coef_uvw(c);
  or_solve_gs(c, "u", c->u, c->ap, c->anb, c->bu, c->ef2nb_idx, c->ef2nb, ne, nf, nbf, c->nit);
  or_solve_gs(c, "v", c->v, c->ap, c->anb, c->bv, c->ef2nb_idx, c->ef2nb, ne, nf, nbf, c->nit);
  or_solve_gs(c, "w", c->w, c->ap, c->anb, c->bw, c->ef2nb_idx, c->ef2nb, ne, nf, nbf, c->nit);
  or_calc_grad(c->u, c->gu, c->xc, c->yc, c->zc, c->ef2nb_idx, c->ef2nb, ne, nf, nbf);
  or_calc_grad(c->v, c->gv, c->xc, c->yc, c->zc, c->ef2nb_idx, c->ef2nb, ne, nf, nbf);
  or_calc_grad(c->w, c->gw, c->xc, c->yc, c->zc, c->ef2nb_idx, c->ef2nb, ne, nf, nbf);
  or_calc_mip(c, 1);
  if (c->asm32)
    or_calc_coef_p32(c);
  else
    or_calc_coef_p(c);
  for (int k = 0; k < ne + nbf; k++) c->phic[k] = 0.0;
  or_solve(c, "pc", c->phic);
  double pref = c->phic[0];
  or_adjust_pc(c, c->phic, pref);
  or_calc_grad(c->phic, c->gpc, c->xc, c->yc, c->zc, c->ef2nb_idx, c->ef2nb, ne, nf, nbf);
  or_update_uvwp(c, c->phic);
}

/* update_time, src/modules/mod_physics.f90:101-112 */
void or_update_time(or_case *c) {
  size_t n = (size_t)(c->ne + c->nbf);
  memcpy(c->u0, c->u, 8 * n);
  memcpy(c->v0, c->v, 8 * n);
  memcpy(c->w0, c->w, 8 * n);
  memcpy(c->mip0, c->mip, 8 * (size_t)c->nf);
}

/* construct_uvwp state initialisation, src/equations/mod_uvwp.f90:57-82 */
void or_construct_uvwp(or_case *c) {
  or_calc_mip(c, 0);
  memcpy(c->mip0, c->mip, 8 * (size_t)c->nf);
}

/* main loop nesting, src/main.f90:50-63; update_boundaries (mod_physics.f90:
   38-50) applies the uvwp and the energy BC setters */
void or_run(or_case *c, int ntstep, int ncoef) {
  for (int t = 0; t < ntstep; t++) {
    for (int k = 0; k < ncoef; k++) {
      or_update_boundaries(c);
      or_update_boundaries_energy(c);
      or_solve_uvwp(c);
    }
    or_update_time(c);
    or_update_time_energy(c);
  }
}

/* ---------------------------------------------------------------- energy */

/* construct_energy, src/equations/mod_energy.f90:14-48: t = 273 K, the
   enthalpy phi = cp*t on the cells (the halos are set by the BC setters
   before any use), phi0 = phi, zero gradients */
void or_construct_energy(or_case *c) {
  int n = c->ne + c->nbf;
  for (int e = 0; e < n; e++) c->t[e] = 273.0;
  for (int e = 0; e < c->ne; e++) c->h[e] = c->t[e] * c->cp[e];
  memcpy(c->h0, c->h, 8 * (size_t)n);
  memset(c->gt, 0, 24 * (size_t)n);
  memset(c->gh, 0, 24 * (size_t)n);
}

/* energy BC setters lid (t = 373) and dirichlet0 (t = 273), mod_energy.f90:
   182-212, phi = cp(owner)*t; construct_energy binds 'top' to lid and every
   other boundary section to dirichlet0 (mod_energy.f90:41-47) */
void or_update_boundaries_energy(or_case *c) {
  for (int s = 0; s < c->nsec; s++) {
    if (c->etype[s] >= 10) continue;
    double tb = c->bckind[s] == OR_BC_LID ? 373.0 : 273.0;
    for (int e = c->esec[2 * s]; e <= c->esec[2 * s + 1]; e++) {
      int enb = PK_E(abs(c->bs[e - c->ne - 1]));
      c->t[e - 1] = tb;
      c->h[e - 1] = c->cp[enb - 1] * c->t[e - 1];
    }
  }
}

/* calc_coef_energy, src/equations/mod_energy.f90:82-169 */
void or_calc_coef_energy(or_case *c) {
  int ne = c->ne;
  const int *g1 = c->ef2nb, *g2 = c->ef2nb + c->L;
  for (int e = 1; e <= ne; e++) {
    c->ap[e - 1] = 0.0;
    double sumf = 0.0, sumdefc = 0.0;
    double rp[3] = {c->xc[e - 1], c->yc[e - 1], c->zc[e - 1]};
    for (int idx = c->ef2nb_idx[e - 1]; idx <= c->ef2nb_idx[e] - 1; idx++) {
      c->anb[idx - 1] = 0.0;
      int enb = PK_E(g1[idx - 1]), lfnb = PK_F(g1[idx - 1]);
      if (lfnb == 0) continue;
      int fg = g2[idx - 1];
      double sg = fg >= 0 ? 1.0 : -1.0;
      fg = abs(fg);
      const double *a = c->aip + 3 * (fg - 1), *rip = c->rip + 3 * (fg - 1);
      double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      double nrm[3], rpnb[3] = {c->xc[enb - 1], c->yc[enb - 1], c->zc[enb - 1]}, dr[3];
      for (int i = 0; i < 3; i++) nrm[i] = sg * a[i] / area;
      for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
      double wt = or_vec_weight(rip, rp, rpnb);
      double f = -(sg * c->mip[fg - 1]);
      double fnb = f > 0.0 ? f : 0.0;
      sumf = sumf + f;
      double tci = (1.0 - wt) * c->tc[e - 1] + wt * c->tc[enb - 1];
      double cpi = (1.0 - wt) * c->cp[e - 1] + wt * c->cp[enb - 1];
      double d = tci / cpi / dot3(dr, nrm) * area;
      double ghi[3], gti[3];
      for (int i = 0; i < 3; i++) {
        ghi[i] = (1.0 - wt) * c->gh[3 * (e - 1) + i] + wt * c->gh[3 * (enb - 1) + i];
        gti[i] = (1.0 - wt) * c->gt[3 * (e - 1) + i] + wt * c->gt[3 * (enb - 1) + i];
      }
      sumdefc = sumdefc + tci * area * (dot3(gti, nrm) - dot3(ghi, nrm) / cpi);
      c->anb[idx - 1] = d + fnb;
      c->ap[e - 1] = c->ap[e - 1] + d + fnb;
    }
    double ap0 = c->rho[e - 1] * c->vol[e - 1] / c->dt;
    c->ap[e - 1] = c->ap[e - 1] + ap0;
    c->b[e - 1] = ap0 * c->h0[e - 1] + sumf * c->h[e - 1] + sumdefc;
  }
  /* boundary loop: interfaces (sections) in order, halos in order; both
     setters are 'dirichlet' (f = 0, d = tc*A/ds/cp) */
  for (int s = 0; s < c->nsec; s++) {
    if (c->etype[s] >= 10) continue;
    for (int enb = c->esec[2 * s]; enb <= c->esec[2 * s + 1]; enb++) {
      int pk = abs(c->bs[enb - c->ne - 1]);
      int e = PK_E(pk), lf = PK_F(pk);
      int idx = c->ef2nb_idx[e - 1] + lf - 1;
      int fg = g2[idx - 1];
      const double *a = c->aip + 3 * (fg - 1);
      double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
      double dr[3] = {c->xc[enb - 1] - c->xc[e - 1], c->yc[enb - 1] - c->yc[e - 1],
                      c->zc[enb - 1] - c->zc[e - 1]};
      double ds = dot3(dr, nrm);
      double f = 0.0;
      double d = c->tc[e - 1] * area / ds / c->cp[e - 1];
      c->ap[e - 1] = c->ap[e - 1] + d + f;
      c->anb[idx - 1] = c->anb[idx - 1] + d + f;
    }
  }
}

/* solve_energy, src/equations/mod_energy.f90:58-80 (calc_temperature,
   mod_properties.f90:214-222) */
void or_solve_energy(or_case *c) {
  int ne = c->ne, nf = c->nf, nbf = c->nbf;
  or_calc_grad(c->t, c->gt, c->xc, c->yc, c->zc, c->ef2nb_idx, c->ef2nb, ne, nf, nbf);
  or_calc_grad(c->h, c->gh, c->xc, c->yc, c->zc, c->ef2nb_idx, c->ef2nb, ne, nf, nbf);
  or_calc_coef_energy(c);
  or_solve_gs(c, "e", c->h, c->ap, c->anb, c->b, c->ef2nb_idx, c->ef2nb, ne, nf, nbf, c->nit);
  for (int e = 0; e < ne; e++) c->t[e] = c->h[e] / c->cp[e];
}

/* update_time, src/modules/mod_physics.f90:111 */
void or_update_time_energy(or_case *c) {
  memcpy(c->h0, c->h, 8 * (size_t)(c->ne + c->nbf));
}

void or_run_energy(or_case *c, int ntstep, int ncoef) {
  for (int t = 0; t < ntstep; t++) {
    for (int k = 0; k < ncoef; k++) {
      or_update_boundaries(c);
      or_update_boundaries_energy(c);
      or_solve_uvwp(c);
      or_solve_energy(c);
    }
    or_update_time(c);
    or_update_time_energy(c);
  }
}
