/*
 * oracle_asm32.c -- restatement of the FP32 coefficient assembly (an
 * extension: the reference assembles calc_coef_uvw, src/equations/
 * mod_uvwp.f90:161-286, and calc_coef_p, :289-368, in FP64 only).
 *
 * TEST INFRASTRUCTURE ONLY (the checker, not the product).
 *
 * The same operations as or_calc_coef_uvw / or_calc_coef_p (the FP64
 * restatement pinned to the reference), in single precision on the geometry,
 * properties and fields rounded to float, results stored as double; the pc
 * right-hand side (each cell's net inflow) stays double.  Order of every
 * float operation as in the GPU kernels: per row entry its terms, then the
 * cell's sums in CSR order, then its boundary faces in halo order, the
 * relaxation, d and dc.  Compile without FMA contraction.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

#define PK_E(p) ((int)((unsigned)(p) >> 5))
#define PK_F(p) ((p) & 31)

static float dot3f(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static float vec_weightf(const float *r0, const float *r1, const float *r2) {
  float ra[3], rb[3];
  for (int i = 0; i < 3; i++) {
    ra[i] = r1[i] - r0[i];
    rb[i] = r2[i] - r0[i];
  }
  const float la = sqrtf(ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2]);
  const float lb = sqrtf(rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2]);
  if (la + lb > 0.0f) return la / (la + lb);
  return 0.5f;
}

static int kind_of_halo(const or_case *c, int enb) {
  for (int s = 0; s < c->nsec; s++)
    if (enb >= c->esec[2 * s] && enb <= c->esec[2 * s + 1]) return c->bckind[s];
  return OR_BC_WALL;
}

static void centref(const or_case *c, int e, float *r) { /* e 1-based */
  r[0] = (float)c->xc[e - 1];
  r[1] = (float)c->yc[e - 1];
  r[2] = (float)c->zc[e - 1];
}

static void face3(const double *v, int fg, float *out) { /* fg 1-based */
  for (int i = 0; i < 3; i++) out[i] = (float)v[3 * (fg - 1) + i];
}

void or_calc_coef_uvw32(or_case *c) {
  const int ne = c->ne;
  const int *g1 = c->ef2nb, *g2 = c->ef2nb + c->L;
  const float dt = (float)c->dt, au = (float)c->alpha_u;
  /* boundary faces of each cell in halo order */
  int *bcnt = calloc((size_t)ne + 1, sizeof(int)), *bslot = malloc(sizeof(int) * (size_t)(c->nbf + 1));
  for (int h = 0; h < c->nbf; h++) bcnt[PK_E(abs(c->bs[h]))]++;
  for (int e = 1; e <= ne; e++) bcnt[e] += bcnt[e - 1];
  {
    int *pos = malloc(sizeof(int) * (size_t)(ne + 1));
    for (int e = 0; e <= ne; e++) pos[e] = bcnt[e];
    for (int h = 0; h < c->nbf; h++) bslot[pos[PK_E(abs(c->bs[h])) - 1]++] = h;
    free(pos);
  }
  for (int e = 1; e <= ne; e++) {
    float rp[3];
    centref(c, e, rp);
    float ap = 0.0f, sumf = 0.0f, sumss[3] = {0.0f, 0.0f, 0.0f}, sumdefc[3] = {0.0f, 0.0f, 0.0f};
    float anbk[64];
    const int r0 = c->ef2nb_idx[e - 1], r1 = c->ef2nb_idx[e];
    for (int idx = r0; idx < r1; idx++) {
      const int enb = PK_E(g1[idx - 1]), lfnb = PK_F(g1[idx - 1]);
      float fnb = 0.0f, d = 0.0f;
      if (lfnb > 0) {
        int fg = g2[idx - 1];
        const float sg = fg >= 0 ? 1.0f : -1.0f;
        fg = abs(fg);
        float a[3], rip[3], rpnb[3], nrm[3], dr[3], drip[3], rpp[3], rnp[3], drp[3];
        face3(c->aip, fg, a);
        face3(c->rip, fg, rip);
        const float area = sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        for (int i = 0; i < 3; i++) nrm[i] = sg * a[i] / area;
        centref(c, enb, rpnb);
        for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
        const float ds = sqrtf(dot3f(dr, dr));
        const float wt = vec_weightf(rip, rp, rpnb);
        for (int i = 0; i < 3; i++) drip[i] = rip[i] - rp[i];
        float t = dot3f(drip, nrm);
        for (int i = 0; i < 3; i++) rpp[i] = rip[i] - t * nrm[i];
        for (int i = 0; i < 3; i++) drip[i] = rip[i] - rpnb[i];
        t = dot3f(drip, nrm);
        for (int i = 0; i < 3; i++) rnp[i] = rip[i] - t * nrm[i];
        for (int i = 0; i < 3; i++) drp[i] = rnp[i] - rpp[i];
        const float dsp = sqrtf(dot3f(drp, drp));
        const float fl = -(sg * (float)c->mip[fg - 1]);
        fnb = fmaxf(fl, 0.0f);
        const float muip = (1.0f - wt) * (float)c->mu[e - 1] + wt * (float)c->mu[enb - 1];
        d = muip * area / ds;
        const double *gs[3] = {c->gu, c->gv, c->gw};
        float ge[3][3], gn[3][3];
        for (int q = 0; q < 3; q++)
          for (int i = 0; i < 3; i++) {
            ge[q][i] = (float)gs[q][3 * (e - 1) + i];
            gn[q][i] = (float)gs[q][3 * (enb - 1) + i];
          }
        float ss[3], dfc[3];
        for (int mm = 0; mm < 3; mm++) {
          float gip[3];
          for (int q = 0; q < 3; q++) gip[q] = (1.0f - wt) * ge[q][mm] + wt * gn[q][mm];
          ss[mm] = muip * area * dot3f(gip, dr) / ds;
        }
        for (int q = 0; q < 3; q++) {
          float gip[3];
          for (int i = 0; i < 3; i++) gip[i] = (1.0f - wt) * ge[q][i] + wt * gn[q][i];
          dfc[q] = muip * area * (dot3f(gip, drp) / dsp - dot3f(gip, dr) / ds);
        }
        sumf = sumf + fl;
        for (int i = 0; i < 3; i++) sumss[i] = sumss[i] + ss[i];
        for (int i = 0; i < 3; i++) sumdefc[i] = sumdefc[i] + dfc[i];
      }
      anbk[idx - r0] = d + fnb;
      ap = ap + d + fnb;
    }
    const float vol = (float)c->vol[e - 1];
    const float ap0 = (float)c->rho[e - 1] * vol / dt;
    ap = ap + ap0;
    const float ue = (float)c->u[e - 1], ve = (float)c->v[e - 1], we = (float)c->w[e - 1];
    float bu = ap0 * (float)c->u0[e - 1] + sumf * ue - vol * (float)c->gp[3 * (e - 1)] + sumss[0] +
               sumdefc[0];
    float bv = ap0 * (float)c->v0[e - 1] + sumf * ve - vol * (float)c->gp[3 * (e - 1) + 1] +
               sumss[1] + sumdefc[1];
    float bw = ap0 * (float)c->w0[e - 1] + sumf * we - vol * (float)c->gp[3 * (e - 1) + 2] +
               sumss[2] + sumdefc[2];
    for (int q = bcnt[e - 1]; q < bcnt[e]; q++) {
      const int h = bslot[q], H = ne + 1 + h;
      const int pk = abs(c->bs[h]), lf = PK_F(pk);
      const int idx = r0 + lf - 1;
      const int fg = g2[idx - 1];
      float a[3], nrm[3], rh[3], dr[3];
      face3(c->aip, fg, a);
      const float area = sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      for (int i = 0; i < 3; i++) nrm[i] = a[i] / area;
      centref(c, H, rh);
      for (int i = 0; i < 3; i++) dr[i] = rh[i] - rp[i];
      const float ds = sqrtf(dot3f(dr, dr));
      const float fz = 0.0f, d = (float)c->mu[e - 1] * area / ds;
      const int kind = kind_of_halo(c, H);
      if (kind == OR_BC_INLET || kind == OR_BC_OUTLET) {
        const float fl = -(float)c->mip[fg - 1], fnb = fmaxf(fl, 0.0f);
        const float dd = kind == OR_BC_INLET ? d : 0.0f;
        bu = bu + fl * ue;
        bv = bv + fl * ve;
        bw = bw + fl * we;
        ap = ap + dd + fnb;
        anbk[idx - r0] = anbk[idx - r0] + dd + fnb;
        continue;
      }
      if (kind != OR_BC_SYMMETRY) {
        float vr[3] = {ue, ve, we};
        const float t = dot3f(vr, nrm);
        for (int i = 0; i < 3; i++) vr[i] = vr[i] - t * nrm[i];
        vr[0] = (float)c->u[H - 1] - vr[0];
        vr[1] = (float)c->v[H - 1] - vr[1];
        vr[2] = (float)c->w[H - 1] - vr[2];
        bu = bu + d * vr[0] - d * ue;
        bv = bv + d * vr[1] - d * ve;
        bw = bw + d * vr[2] - d * we;
      }
      ap = ap + d + fz;
      anbk[idx - r0] = anbk[idx - r0] + d + fz;
    }
    if (au != 1.0f) {
      const float apr = ap / au, da = apr - ap;
      bu = bu + da * ue;
      bv = bv + da * ve;
      bw = bw + da * we;
      ap = apr;
    }
    for (int idx = r0; idx < r1; idx++) c->anb[idx - 1] = (double)anbk[idx - r0];
    c->ap[e - 1] = (double)ap;
    c->bu[e - 1] = (double)bu;
    c->bv[e - 1] = (double)bv;
    c->bw[e - 1] = (double)bw;
    c->d[e - 1] = (double)(vol / ap);
    float dc = ap;
    for (int idx = r0; idx < r1; idx++) dc = dc - anbk[idx - r0];
    c->dc[e - 1] = (double)(vol / dc);
  }
  free(bcnt);
  free(bslot);
}

void or_calc_coef_p32(or_case *c) {
  const int ne = c->ne;
  const int *g1 = c->ef2nb, *g2 = c->ef2nb + c->L;
  for (int e = 1; e <= ne; e++) {
    float rp[3];
    centref(c, e, rp);
    float ap = 0.0f;
    double sumf = 0.0;
    for (int idx = c->ef2nb_idx[e - 1]; idx <= c->ef2nb_idx[e] - 1; idx++) {
      const int enb = PK_E(g1[idx - 1]), lfnb = PK_F(g1[idx - 1]);
      float d = 0.0f;
      if (lfnb > 0) {
        int fg = g2[idx - 1];
        const float sg = fg >= 0 ? 1.0f : -1.0f;
        const double sgd = fg >= 0 ? 1.0 : -1.0;
        fg = abs(fg);
        float a[3], rip[3], rpnb[3], nrm[3], dr[3];
        face3(c->aip, fg, a);
        face3(c->rip, fg, rip);
        const float area = sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        for (int i = 0; i < 3; i++) nrm[i] = sg * a[i] / area;
        centref(c, enb, rpnb);
        for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
        const float wt = vec_weightf(rip, rp, rpnb);
        const double fl = -(sgd * c->mip[fg - 1]);
        sumf = sumf + fl;
        const float rhoip = (1.0f - wt) * (float)c->rho[e - 1] + wt * (float)c->rho[enb - 1];
        d = ((1.0f - wt) * (float)c->dc[e - 1] + wt * (float)c->dc[enb - 1]) / dot3f(dr, nrm) *
            rhoip * area;
      }
      c->anb[idx - 1] = (double)d;
      ap = ap + d;
    }
    c->ap[e - 1] = (double)ap;
    c->b[e - 1] = sumf;
  }
  for (int enb = ne + 1; enb <= ne + c->nbf; enb++) {
    const int pk = abs(c->bs[enb - ne - 1]);
    const int e = PK_E(pk), lf = PK_F(pk);
    const int idx = c->ef2nb_idx[e - 1] + lf - 1;
    const int fg = g2[idx - 1];
    const float d = 0.0f;
    c->ap[e - 1] = (double)((float)c->ap[e - 1] + d);
    c->anb[idx - 1] = (double)((float)c->anb[idx - 1] + d);
    c->b[e - 1] = c->b[e - 1] - c->mip[fg - 1];
  }
}
