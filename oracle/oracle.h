/*
 * oracle.h -- CPU restatement of CFD-Lite's SIMPLE pressure-velocity loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it.
 * The HIP library under cfd-lite_amd/ does not link or call anything here.
 *
 * Every routine restates the reference (Fortran, real(8), serial) operation by
 * operation, in the same evaluation order, so that results are bitwise equal
 * to the reference built without FMA contraction.  File:line citations point
 * into the reference tree (src/...).
 *
 * Conventions kept from the reference (src/modules/mod_util.f90:1428-1448):
 *   - elements are numbered 1..ne (cells) then ne+1..ne+nbf (boundary halos);
 *   - packed index = (elem << 5) | local_face, local_face 0 marks a halo;
 *   - ef2nb is the CSR neighbour table, column 1 = packed neighbour,
 *     column 2 = signed global face (+ on the owner side);
 *   - C arrays are 0-based, the integers stored inside them are 1-based.
 *
 * Parity status of each function is stated in its comment: "pinned" means a
 * test compares it with the reference compiled from its own sources
 * (oracle/_ref) and/or with closed forms.
 */
#ifndef CFDL_ORACLE_H
#define CFDL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_HEXA_8 = 17, OR_QUAD_4 = 7 };
enum { OR_BC_WALL = 0, OR_BC_LID = 1, OR_BC_SYMMETRY = 2 };
/* channel boundaries (an extension: the reference's inlet/outlet setters,
   src/equations/mod_uvwp.f90:572-606, are bound nowhere and its assembly
   cannot use them, mod_uvwp.f90:254-270): a parabolic velocity inlet and a
   zero-gradient outlet whose mass flux is scaled to the inflow */
enum { OR_BC_INLET = 3, OR_BC_OUTLET = 4 };

typedef struct {
  int ne, nf, nbf;
  int *ef2nb_idx; /* ne+1 */
  int *ef2nb;     /* L = 2nf-nbf entries, halo entries packed (ne+k)<<5 */
  double *ap, *b, *anb, *phic;
} or_subd;

typedef struct {
  int c1, c2, ncs;
  int *index1; /* owned */
  int *index2; /* alias of intf(c2,c1).index1 */
} or_intf;

typedef struct {
  int it;
  double res_i, res_f, res_max;
  char name[8];
} or_logline;

typedef struct or_case {
  /* mesh input (CGNS-like element lists, sections sorted by element range) */
  int nvx, nsec, nelem, ne, nbf, nf, L;
  double *x, *y, *z;
  int *e2vx;              /* 8 per element, 1-based vertex ids */
  int *etype, *esec;      /* nsec, 2*nsec */
  int *bckind;            /* nsec: BC kind of each boundary section (-1 for cells) */
  /* connectivity */
  int *ef2nb_idx, *ef2nb, *s2g, *bs;  /* bs has nbf entries (halo ne+1+k) */
  /* geometry */
  double *xc, *yc, *zc, *vol, *aip, *rip;
  /* partition */
  int nsubd;
  int *gf2g, *g2gf, *g2gf_idx;
  or_subd *sd;
  or_intf *intf; /* nsubd*nsubd, (c,cnb) at c*nsubd+cnb */
  /* physics */
  double dt;
  int nit;
  double alpha_u, alpha_p; /* config-mode under-relaxation (1 = reference) */
  double *rho, *mu;
  /* equation state (mod_uvwp.f90:6-16) */
  double *u, *v, *w, *p, *u0, *v0, *w0;
  double *gu, *gv, *gw, *gp, *gpc;
  double *mip, *mip0;
  double *bu, *bv, *bw, *d, *dc;
  /* shared work arrays (mod_physics.f90:58-63) */
  double *ap, *anb, *b, *phic;
  /* residual log */
  int nlog, caplog;
  or_logline *log;
  /* energy equation (energy_t, mod_energy.f90:4-10): temperature t and its
     gradient gt, enthalpy h (the equation's phi), h0, gradient gh; thermal
     properties cp, tc (properties_t, mod_properties.f90:76-90) */
  double *cp, *tc;
  double *t, *gt, *h, *h0, *gh;
  /* pressure-correction visiting order of the subdomain smoother: 0 is the
     reference's (local cell order); 1 visits the cells colour by colour
     (or_set_pc_order), an extension the reference does not have; 2 is the
     colour order with FP32 sweeps on the correction system under FP64
     defect correction at every residual check (an extension as well) */
  int pc_order;
  int **pc_perm; /* per subdomain: its ne cells (1-based) in visiting order */
  /* visiting order of solve_gs: 0 is the reference's (global cell order); 1
     visits the cells colour by colour (or_set_gs_order), an extension */
  int gs_order;
  int *gs_perm; /* ne cells (1-based) in visiting order */
  /* multigrid hierarchy of the pc solve in pc_order 3 and 4 (an extension:
     multigrid-preconditioned conjugate gradients on the global system, FP64
     or with an FP32 preconditioner), built on first use */
  struct or_mg *mg;
  /* inlet profile (OR_BC_INLET): umax, y0, y1, z0, z1; the halo velocity is
     umax*16*s*(1-s)*t*(1-t) along the inward normal, s = (y-y0)/(y1-y0),
     t = (z-z0)/(z1-z0) at the face centre */
  double inlet[5];
  /* FP32 coefficient assembly (an extension, oracle_asm32.c): 1 = ap, anb,
     sources, d, dc in single precision; the pc right-hand side stays double */
  int asm32;
} or_case;

/* ---- setup ---- */
or_case *or_case_new(int nvx, const double *x, const double *y, const double *z,
                     int nsec, const int *etype, const int *esec, const int *e2vx,
                     const int *bckind);
int or_connectivity(or_case *c);       /* returns 0 or error */
void or_geometry(or_case *c);
int or_partition(or_case *c, int nsubd);
void or_construct_subdomains(or_case *c);
void or_init_physics(or_case *c, double rho, double mu, double dt, int nit);
void or_case_free(or_case *c);

/* ---- helpers exported for tests ---- */
void or_qsort_key_nrec(int *key, int *b, int n);
double or_vec_weight(const double *r0, const double *r1, const double *r2);
void or_matinv3(const double *A, double *B); /* column-major 3x3 */

/* ---- hot path (plain-array routines mirror mod_solver.f90) ---- */
void or_calc_grad(const double *phi, double *grad, const double *xc, const double *yc,
                  const double *zc, const int *ef2nb_idx, const int *ef2nb,
                  int ne, int nf, int nbf);
void or_solve_gs(or_case *c, const char *name, double *phi, const double *ap,
                 const double *anb, const double *b, const int *ef2nb_idx,
                 const int *ef2nb, int ne, int nf, int nbf, int nit);
void or_smoother_gs(const char *name, double *phi, const double *ap, const double *anb,
                    const double *b, const int *ef2nb_idx, const int *ef2nb,
                    int ne, int nf, int nbf, int nit);
void or_calc_residual(const double *phi, const double *ap, const double *anb,
                      const double *b, const int *ef2nb_idx, const int *ef2nb,
                      int ne, int nf, int nbf, double *res, double *res_max);
void or_solve(or_case *c, const char *name, double *phi);
/* colour order of the subdomain pressure-correction smoother (extension):
   orders 1 and 2 colour every subdomain greedily in global cell order (its
   cells taken by increasing global index l2g; a cell takes the least colour
   that none of its interior neighbours with a smaller global index has) and
   visit the cells by colour, then by local index; order 0 restores the
   reference's order. */
void or_set_pc_order(or_case *c, int order);
void or_set_gs_order(or_case *c, int order);
/* pc_order 3 / 4 (extension, oracle_mgcg.c): the pc solve is multigrid-
   preconditioned conjugate gradients on the global system, with the
   preconditioner in FP64 (3) or FP32 (4) */
typedef struct or_mg or_mg;
or_mg *or_mg_build(int ne, const int *ef2nb_idx, const int *ef2nb, const double *xc,
                   const double *yc, const double *zc, int maxagg, int ncoarse);
void or_mg_free(or_mg *M);
int or_mg_nlev(const or_mg *M);
int or_mg_get(const or_mg *M, int l, const char *name, int *dst, int n);
or_mg *or_case_mg(or_case *c);
void or_mg_apply(or_case *c, const double *ap, const double *anb, const double *r, double *z,
                 int fp32);
int or_mg_operator(or_case *c, const double *ap, const double *anb, int l, double *apl,
                   double *anbl);
void or_mgcg_solve(or_case *c, const char *name, double *phi, int fp32);
double or_tree_dot(const double *a, const double *b, int n);
double or_tree_sum(const double *a, int n);
void or_tree_blocks(int n, int *nblk, int *chunk);
void or_log_line(or_case *c, const char *name, int it, double ri, double rf, double rm);

/* ---- hot path (equation routines mirror mod_uvwp.f90) ---- */
void or_update_boundaries(or_case *c);
void or_scale_outlet(or_case *c); /* channel extension, called by or_update_boundaries */
void or_calc_coef_uvw(or_case *c);
void or_calc_mip(or_case *c, int rhie_chow);
void or_calc_coef_p(or_case *c);
void or_calc_coef_uvw32(or_case *c); /* FP32 assembly, oracle_asm32.c */
void or_calc_coef_p32(or_case *c);
void or_adjust_pc(or_case *c, double *pc, double pref);
void or_update_uvwp(or_case *c, const double *pc);
void or_solve_uvwp(or_case *c);
void or_update_time(or_case *c);
void or_construct_uvwp(or_case *c); /* zero state, calc_mip(.false.), mip0=mip */

/* run ntstep x ncoef outer iterations exactly as src/main.f90:50-63 */
void or_run(or_case *c, int ntstep, int ncoef);

/* ---- energy equation (mirror mod_energy.f90; never solved by the
   reference's main loop, whose update_boundaries still applies its BCs) ---- */
void or_construct_energy(or_case *c);          /* mod_energy.f90:14-48 */
void or_update_boundaries_energy(or_case *c);  /* lid / dirichlet0, :182-212 */
void or_calc_coef_energy(or_case *c);          /* :82-169 */
void or_solve_energy(or_case *c);              /* :58-80 */
void or_update_time_energy(or_case *c);        /* mod_physics.f90:111 */
/* or_run with solve_energy after every solve_uvwp */
void or_run_energy(or_case *c, int ntstep, int ncoef);

#ifdef __cplusplus
}
#endif
#endif
