/*
 * cfdl.h -- C ABI of libcfdl_hip.so, the MI355X (gfx950) implementation of
 * CFD-Lite's SIMPLE pressure-velocity coupling loop.
 *
 * The reference is Fortran; its hot path is the module procedures of
 * src/equations/mod_uvwp.f90 and src/modules/mod_solver.f90.  Their derived
 * types (uvwp_t, geometry_t, subdomain_t) carry allocatables and type-bound
 * procedures and cannot cross bind(C), so this ABI takes the component arrays
 * (plain pointers + sizes) and keeps device copies in an opaque context.  A
 * Fortran host binds every function below with iso_c_binding (the interface
 * module is cfd-lite_amd/fortran/cfdl_c.f90).
 *
 * Conventions are the reference's: reals are IEEE double, integers int32,
 * arrays are passed as the reference stores them (Fortran column-major,
 * 1-based packed indices (elem << 5) | local_face inside ef2nb/s2g/bs,
 * src/modules/mod_util.f90:1428-1448).  Every function returns 0 on success
 * or a CFDL_E* code; cfdl_last_error() gives the message (the reference stops
 * the program instead, src/modules/mod_eqn_setup.f90:63-66).
 *
 * Threading: one context = one HIP stream = one calling thread.  Calls on a
 * context are asynchronous with respect to the host except where stated;
 * cfdl_get_field and the log read-out synchronise.
 */
#ifndef CFDL_H
#define CFDL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum cfdl_status {
  CFDL_OK = 0,
  CFDL_EINVAL = 1, /* bad argument or size */
  CFDL_EHIP = 2,   /* HIP runtime error */
  CFDL_ENOMEM = 3, /* host or device allocation failed */
  CFDL_ESTATE = 4, /* call out of order (e.g. no partition) */
  CFDL_ENODEV = 5  /* no GPU visible */
};

/* boundary-section kinds: the BC setters of src/equations/mod_uvwp.f90 */
enum cfdl_bc_kind {
  CFDL_BC_WALL = 0,     /* dirichlet0, mod_uvwp.f90:493-515 */
  CFDL_BC_LID = 1,      /* lid (u=1), mod_uvwp.f90:548-570 */
  CFDL_BC_SYMMETRY = 2, /* symmetry (zero_flux), mod_uvwp.f90:517-546 */
  /* Channel boundaries, no counterpart that runs in the reference (its inlet
     and outlet setters, mod_uvwp.f90:572-606, are bound nowhere and the BC
     loops of calc_coef_uvw / calc_coef_p handle only 'dirichlet' and
     'zero_flux', :254-270, 344-366):
     INLET: velocity along the inward face normal with the parabolic profile
       of cfdl_set_inlet_profile at the face centre, p = p(owner), mass flux
       rho(owner) (v.n) A; momentum: diffusion mu A / ds and upwind
       convection of that flux (anb += d + max(f, 0), b += f u_P);
     OUTLET: zero gradient, u, v, w, p = the owner's; mass flux rho (v.n) A
       from the owner's velocity, then all outlet fluxes scaled so that the
       outflow equals the inflow (by face area if the extrapolated outflow is
       not positive); momentum: convection only.
     The pc system keeps every boundary flux fixed (Neumann). */
  CFDL_BC_INLET = 3,
  CFDL_BC_OUTLET = 4
};

/* fields of uvwp_t (mod_uvwp.f90:6-16) and the shared work arrays of phys_t
   (mod_physics.f90:58-63).  Sizes: CELLS+HALOS = ne+nbf, CELLS = ne,
   FACES = nf, CSR = 2nf-nbf, gradients 3 per element (x,y,z interleaved). */
enum cfdl_field {
  CFDL_U = 0, CFDL_V, CFDL_W, CFDL_P, CFDL_U0, CFDL_V0, CFDL_W0, /* ne+nbf   */
  CFDL_GU, CFDL_GV, CFDL_GW, CFDL_GP, CFDL_GPC,                   /* 3(ne+nbf) */
  CFDL_MIP, CFDL_MIP0,                                            /* nf       */
  CFDL_BU, CFDL_BV, CFDL_BW, CFDL_D, CFDL_DC,                     /* ne       */
  CFDL_AP, CFDL_B,                                                /* ne       */
  CFDL_ANB,                                                       /* 2nf-nbf  */
  CFDL_PHIC,                                                      /* ne+nbf   */
  CFDL_WORK,   /* ne+nbf scratch: any element field a caller passes in */
  CFDL_GWORK,  /* 3(ne+nbf) scratch gradient */
  /* energy equation (energy_t, mod_energy.f90:4-10): temperature t, its
     gradient gt, enthalpy (the equation's phi), phi0, grad */
  CFDL_T, CFDL_GT, CFDL_ENTH, CFDL_ENTH0, CFDL_GENTH, /* ne+nbf, 3(ne+nbf), ne+nbf x2, 3(ne+nbf) */
  CFDL_NFIELDS
};

const char *cfdl_last_error(void);
int cfdl_version(void);

/* ------------------------------------------------------------------------ */
/* Host-side setup.  Replaces the reference's pre-processing on a generated   */
/* or given element list: find_element_nb (src/setup/mod_mg_lvl_uns.f90:     */
/* 283-488), calc_aip_xyzip_uns (src/setup/calc_aip_xyzip.f90:7-72),          */
/* calc_vol_cv_centers_uns (src/setup/calc_vol_cv_centers.f90:4-59) and the   */
/* subdomain partition generate_seeds + add_transformation_bt                 */
/* (src/setup/mod_mg_lvl_uns.f90:95-118, 873-895).                            */
/* ------------------------------------------------------------------------ */
typedef struct cfdl_mesh cfdl_mesh;

/* Hexahedral box, CGNS HEXA_8 cells (i fastest) followed by QUAD_4 sections
   bottom, south, east, north, west, top.  jitter > 0 moves interior vertices
   by uniform(-jitter*h, jitter*h) per coordinate (splitmix64 stream, seed). */
int cfdl_mesh_box(int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                  uint64_t seed, cfdl_mesh **out);
/* Same input the reference reads from CGNS (cell_input.f90:36-93): sections
   sorted by element range, e2vx with 8 slots per element. */
int cfdl_mesh_from_elements(int nvx, const double *x, const double *y, const double *z,
                            int nsec, const int *etype, const int *esec, const int *e2vx,
                            cfdl_mesh **out);
/* Section names, which bind the boundary conditions (the reference's
   make_bc matches them, mod_eqn_setup.f90:46-67; cfdl_ctx_create_cavity
   needs a section named "top").  cfdl_mesh_box names its sections itself. */
int cfdl_mesh_set_section_names(cfdl_mesh *m, int nsec, const char *const *names);
/* Oct-tree bisection into nsubd subdomains and the subdomain-local order
   (nsubd = 1: no partition, the pc solve uses solve_gs). */
int cfdl_mesh_partition(cfdl_mesh *m, int nsubd);
int cfdl_mesh_info(const cfdl_mesh *m, int *ne, int *nf, int *nbf, int *nsec, int *nsubd,
                   int *nvx);
/* Cell renumbering of the distorted-channel input (no counterpart in the
   reference, which keeps its file's element order): a seeded random
   permutation of the cells (Fisher-Yates, position i = ne-1 .. 1 swaps with
   splitmix64 % (i+1)), then reverse Cuthill-McKee on the interior-face graph
   (each component from its lowest-numbered least-degree cell, neighbours by
   increasing degree then number, sequence reversed).  Connectivity and
   geometry are rebuilt; boundary elements and halos keep their order.  Call
   before cfdl_mesh_partition. */
int cfdl_mesh_renumber(cfdl_mesh *m, uint64_t seed);
/* Copy out one setup array by name: "ef2nb_idx" (ne+1), "ef2nb" (2(2nf-nbf),
   column-major), "s2g" (nf), "bs" (nbf), "esec" (2nsec), "etype" (nsec),
   "gf2g" (ne), "g2gf" (ne), "g2gf_idx" (nsubd+1), "e2vx" (8(ne+nbf));
   f64: "xc","yc","zc" (ne+nbf), "vol" (ne), "aip","rip" (3nf), "x","y","z". */
int cfdl_mesh_get_i32(const cfdl_mesh *m, const char *name, int *dst, int64_t n);
int cfdl_mesh_get_f64(const cfdl_mesh *m, const char *name, double *dst, int64_t n);
void cfdl_mesh_free(cfdl_mesh *m);
/* The agglomeration multigrid hierarchy the CFDL_PC_MGCG pc solve builds
   for this mesh (no counterpart in the reference; see cfdl_set_pc_order):
   copies level `level`'s array `name` -- "n", "row", "col", "cptr", "cperm",
   and towards the next level "agg", "mptr", "mem", "iptr", "ilist", "gptr",
   "glist" -- up to cap values into dst and its length into *count (-1 when
   the level or name does not exist).  Level 0 is the mesh CSR (0-based,
   neighbours >= ne are halos). */
int cfdl_mesh_mg_get(const cfdl_mesh *m, int level, const char *name, int *dst, int64_t cap,
                     int64_t *count);
/* VTU file of cell arrays on the mesh, byte for byte the layout the
   reference's write_vtubin / vtu_data writes (src/modules/mod_vtu_output.f90:
   6-326): nvar Float64 cell arrays (names[v], ncomp[v] components, ne*ncomp
   values in cell order), then points, connectivity, offsets and types in
   one appended raw block.  For the uvwp equation the reference writes u, v,
   w, p, gpc (3 components) and the cell mass imbalance as 'mip'
   (cfdl_calc_mass_imbalance). */
int cfdl_mesh_write_vtu(const cfdl_mesh *m, const char *path, int nvar, const char *const *names,
                        const int *ncomp, const double *const *data);

/* ------------------------------------------------------------------------ */
/* Device context: the drop-in for the hot path.                             */
/* ------------------------------------------------------------------------ */
typedef struct cfdl_ctx cfdl_ctx;

/* Uploads the static mesh (geometry_t + meshds_t arrays, mod_mg_lvl_uns.f90:
   46-58, mod_meshds_uns.f90:14-26), the BC binding per section, the
   partition (g2gf(1), gf2g(1)) from which the subdomain systems of
   construct_subdomains (src/modules/mod_subdomains.f90:18-160) are built,
   rho/mu per cell, dt and the linear-solver cap nit (phys_t,
   mod_physics.f90:13-35).  ef2nb is the reference's ef2nb(L,2).  bckind
   has one entry per section (ignored for cell sections).  Fields start at
   zero and mip = mip0 = 0, as construct_uvwp leaves them
   (mod_uvwp.f90:57-82). */
int cfdl_ctx_create(int ne, int nf, int nbf, const int *ef2nb_idx, const int *ef2nb,
                    const int *s2g, const int *bs, const double *xc, const double *yc,
                    const double *zc, const double *vol, const double *aip, const double *rip,
                    int nsec, const int *esec, const int *bckind, int nsubd, const int *gf2g,
                    const int *g2gf, const int *g2gf_idx, const double *rho, const double *mu,
                    double dt, int nit, cfdl_ctx **out);
/* Convenience: a context from a cfdl_mesh with uniform rho, mu and the
   lid-driven-cavity binding ('top' -> lid, every other section -> wall,
   mod_uvwp.f90:73-78). */
int cfdl_ctx_create_cavity(const cfdl_mesh *m, double rho, double mu, double dt, int nit,
                           cfdl_ctx **out);
/* Convenience: a channel along x from a cfdl_mesh with uniform rho, mu:
   section 'west' -> CFDL_BC_INLET with u_max = umax over the y-z extent of
   the mesh's vertices, 'east' -> CFDL_BC_OUTLET, every other section a
   no-slip wall (dirichlet0). */
int cfdl_ctx_create_channel(const cfdl_mesh *m, double rho, double mu, double dt, int nit,
                            double umax, cfdl_ctx **out);
void cfdl_ctx_destroy(cfdl_ctx *c);

/* Config mode (no counterpart in the reference, which relaxes only through
   its pseudo-time step dt): Patankar implicit under-relaxation of the
   momentum equations, ap -> ap/alpha_u and b += (ap/alpha_u - ap)*phi before
   d and dc are formed, and p += alpha_p*pc, gp += alpha_p*gpc.  Both 1 (the
   default) is the reference algorithm bit for bit. */
int cfdl_set_relaxation(cfdl_ctx *c, double alpha_u, double alpha_p);

/* Visiting order of the multi-subdomain pressure-correction smoother
   (smoother_gs inside multi_subdomain_solver, mod_solver.f90:124-228).
   CFDL_PC_REFERENCE (the default) is the reference's order, bit for bit.
   CFDL_PC_COLORED has no counterpart in the reference: every subdomain is
   coloured greedily in global cell order (a cell takes the least colour
   none of its interior neighbours earlier in that order has; two colours on
   a structured box) and the sweeps visit the cells
   colour by colour (forward) and in reverse (backward) -- the same
   Gauss-Seidel update in another sequential order, with every stop rule
   unchanged, in a few fully parallel launches per sweep.
   CFDL_PC_COLORED_MIXED (no counterpart in the reference either) is the
   colour order with an FP32 inner solve under FP64 defect correction: the
   sweeps run in single precision on the correction system
   ap*d - sum(anb*d_nb) = r, and at every residual check (every ten
   iterations, as in the reference) d is added to the FP64 values and
   cleared, and the FP64 true residual r = b + sum(anb*pc_nb) - ap*pc gives
   the stop test's norms and the next right-hand side.  Assembly, residuals,
   stop rules and the fields outside the pc solve stay FP64.
   Neither colour order has an effect with one subdomain (the pc solve is
   then solve_gs on the global system).
   CFDL_PC_MGCG and CFDL_PC_MGCG_MIXED (no counterpart in the reference, whose
   agglomeration levels, src/setup/mod_agglomeration.f90:331-561, no solver
   uses) replace the subdomain smoother by conjugate gradients on the global
   system, preconditioned by one V-cycle of agglomeration multigrid (cells
   grouped by recursive longest-axis bisection of their centres, at most 8^l
   per level-l aggregate; Galerkin coarse operators; one forward colour
   Gauss-Seidel sweep before and one backward sweep after each coarse
   correction).  The stop rule is the reference's (rms residual at most a
   tenth of the initial one, or nit iterations) and the log line keeps the
   solve's name.  In CFDL_PC_MGCG_MIXED the V-cycle runs in FP32; CG, the
   residual, the norms and the stop test stay FP64.  The hierarchy is built
   on the first call that selects either mode. */
enum {
  CFDL_PC_REFERENCE = 0,
  CFDL_PC_COLORED = 1,
  CFDL_PC_COLORED_MIXED = 2,
  CFDL_PC_MGCG = 3,
  CFDL_PC_MGCG_MIXED = 4
};
int cfdl_set_pc_order(cfdl_ctx *c, int order);

/* Arithmetic of the coefficient assembly (calc_coef_uvw and calc_coef_p,
   mod_uvwp.f90:161-368): 64 (the default) is the reference's FP64 bit for
   bit; 32 (no counterpart in the reference) computes ap, anb, the momentum
   sources, d and dc in FP32 from FP32 copies of the geometry and properties
   and the fields rounded to FP32, stored as FP64.  The pc right-hand side
   (each cell's net inflow, the continuity check), the fields, the linear
   solves' residuals and norms and the update stay FP64.  Needs a mesh whose
   rows all have six entries. */
int cfdl_set_assembly_precision(cfdl_ctx *c, int bits);

/* Visiting order of solve_gs (mod_solver.f90:255-327: the momentum solves,
   the energy solve, and the pc solve with one subdomain).
   CFDL_GS_REFERENCE (the default) is the reference's global order, bit for
   bit.  CFDL_GS_COLORED has no counterpart in the reference: the global
   system is coloured greedily in global order (a cell takes the least colour
   none of its earlier interior neighbours has) and each sweep visits the
   cells colour by colour, forward, and in reverse, backward -- the same
   update and stop rule in another sequential order, one launch per colour. */
enum { CFDL_GS_REFERENCE = 0, CFDL_GS_COLORED = 1 };
int cfdl_set_gs_order(cfdl_ctx *c, int order);

/* Host-side state a Fortran drop-in keeps authoritative (properties_t and
   phys_t%dt, mod_properties.f90:9-51, mod_physics.f90:15; the BC type the
   setters assign, mod_uvwp.f90:500,525,555: 'dirichlet' -> CFDL_BC_WALL,
   'zero_flux' -> CFDL_BC_SYMMETRY) for the halos first..last (1-based
   element numbers, ne < first <= last <= ne+nbf). */
int cfdl_set_properties(cfdl_ctx *c, const double *rho, const double *mu);
/* dt = +infinity (here or at context creation) is the steady mode (no
   counterpart in the reference, whose loop is pseudo-transient, main.f90:
   50-63 with dt = 0.01, mod_physics.f90:15): rho V / dt and the Rhie-Chow
   transient term (mod_uvwp.f90:233-238, 479-484) are then exactly zero and
   the outer iterations are steady SIMPLE, which needs momentum
   under-relaxation (cfdl_set_relaxation with alpha_u < 1; without it
   cfdl_calc_coef_uvw returns CFDL_ESTATE). */
int cfdl_set_dt(cfdl_ctx *c, double dt);
int cfdl_set_bc_kind(cfdl_ctx *c, int first, int last, int kind);
/* Inlet profile of CFDL_BC_INLET halos (channel extension): normal velocity
   umax*16*s*(1-s)*t*(1-t) at the face centre, s = (y-y0)/(y1-y0),
   t = (z-z0)/(z1-z0); default umax = 1 over the unit square. */
int cfdl_set_inlet_profile(cfdl_ctx *c, double umax, double y0, double y1, double z0, double z1);

/* host <-> device copies of one field (n = element count of that field) */
int cfdl_set_field(cfdl_ctx *c, int field, const double *src, int64_t n);
int cfdl_get_field(cfdl_ctx *c, int field, double *dst, int64_t n);
int cfdl_synchronize(cfdl_ctx *c);
/* The context back to its state at creation (every field cleared and the
   energy fields initialised, the log and the outer-iteration count
   cleared), keeping its mesh, properties, BCs and settings.  No counterpart
   in the reference; serving reuses a context for the next case this way. */
int cfdl_ctx_reset(cfdl_ctx *c);

/* One entry point per reference routine, all on device-resident state.    */
int cfdl_update_boundaries(cfdl_ctx *c);   /* mod_physics.f90:38-50 (uvwp, energy BCs) */
int cfdl_calc_coef_uvw(cfdl_ctx *c);       /* mod_uvwp.f90:161-286             */
/* solve_gs (mod_solver.f90:255-327) on the shared ap/anb for phi = field
   (CFDL_U/V/W or CFDL_PHIC) with right-hand side rhs (CFDL_BU/BV/BW or
   CFDL_B); name decides the SOR factor ("pc..." -> 1.02) and labels the log. */
int cfdl_solve_gs(cfdl_ctx *c, const char *name, int field, int rhs);
/* the three momentum solves of solve_uvwp (mod_uvwp.f90:114-116) at once */
int cfdl_solve_gs_uvw(cfdl_ctx *c);
int cfdl_calc_grad(cfdl_ctx *c, int src, int dst); /* mod_solver.f90:40-81 */
int cfdl_calc_mip(cfdl_ctx *c, int rhie_chow);     /* mod_uvwp.f90:438-490 */
int cfdl_calc_coef_p(cfdl_ctx *c);                 /* mod_uvwp.f90:289-368 */
int cfdl_set_a_0(cfdl_ctx *c, int field);          /* mod_util.f90:425-435 */
/* solve (mod_solver.f90:329-344): multi_subdomain_solver (:124-189) on
   phic with ap/anb/b, or solve_gs when the context has one subdomain */
int cfdl_solve(cfdl_ctx *c, const char *name);
int cfdl_adjust_pc(cfdl_ctx *c);   /* pref = phic(1); mod_uvwp.f90:129-130,136-158 */
int cfdl_update_uvwp(cfdl_ctx *c); /* mod_uvwp.f90:370-436 */
int cfdl_solve_uvwp(cfdl_ctx *c);  /* mod_uvwp.f90:95-134, fused */
int cfdl_update_time(cfdl_ctx *c); /* mod_physics.f90:101-112 (uvwp, energy) */
/* Energy equation (src/equations/mod_energy.f90), which the reference's main
   loop never solves but whose BC setters its update_boundaries applies
   (mod_physics.f90:44-48).  cfdl_update_boundaries and cfdl_update_time act
   on it as the reference's update_boundaries / update_time do.  Thermal
   properties start as init_properties sets them (cp = 1000, tc = 5,
   mod_properties.f90:86-89) and the state as construct_energy leaves it
   (t = 273, enthalpy = cp*t, mod_energy.f90:30-38). */
int cfdl_set_thermal(cfdl_ctx *c, const double *cp, const double *tc); /* ne each */
int cfdl_construct_energy(cfdl_ctx *c);          /* mod_energy.f90:30-38 */
int cfdl_update_boundaries_energy(cfdl_ctx *c);  /* lid / dirichlet0, :182-212 */
int cfdl_calc_coef_energy(cfdl_ctx *c);          /* :82-169, into ap/anb/b */
int cfdl_solve_energy(cfdl_ctx *c);              /* solve_energy, :58-80 */
/* cfdl_run with solve_energy after every solve_uvwp */
int cfdl_run_energy(cfdl_ctx *c, int ntstep, int ncoef);

/* Net outward mass flux of each cell into field dst (ne values): the array
   the reference's VTU output computes into dc and writes as 'mip'
   (mod_vtu_output.f90:113-123); dc itself is left alone. */
int cfdl_calc_mass_imbalance(cfdl_ctx *c, int dst);
/* ntstep x ncoef outer iterations with the loop nesting of src/main.f90:50-63 */
int cfdl_run(cfdl_ctx *c, int ntstep, int ncoef);

/* Residual log: one line per linear solve, the values the reference prints
   with format (5x,A16,x,i5,x,15x,es9.3e2,...) at mod_solver.f90:6,184,325.
   Reading synchronises the context. */
int cfdl_log_count(cfdl_ctx *c, int *n);
int cfdl_log_get(cfdl_ctx *c, int k, char name[8], int *it, double *res_i, double *res_f,
                 double *res_max);
int cfdl_log_clear(cfdl_ctx *c);

/* calc_residual (src/modules/mod_solver.f90:230-253) on a system already in
   device memory, outside any context: r = b + sum(anb * phi_nb) - ap * phi
   over the cells, *res = sqrt(sum(r^2) / ne) and *res_max = max(r, 0)
   (signed, as the reference's).  Device arrays: row (ne+1, 0-based CSR),
   nbr (row[ne] neighbour indices, 0-based; halos >= ne), ap, b (ne), anb
   (row[ne]), phi (every index nbr names).  res and res_max are host
   pointers.  stream is a hipStream_t (NULL: the default stream); the call
   synchronises it.  1 <= ne <= 2^30.  The sum is a fixed-order tree. */
int cfdl_calc_residual(int ne, const int *row, const int *nbr, const double *ap,
                       const double *anb, const double *b, const double *phi, double *res,
                       double *res_max, void *stream);
/* The same, allocating nothing (no counterpart in the reference): for
   callers that cannot allocate device memory -- code under HIP graph
   capture, tight serving loops, harnesses that forbid it.  Its block
   partials live in a fixed device array of 4098 doubles (2048 blocks) that
   each GPU keeps until the process ends, so ne is limited to 2^30 as above;
   calls that overlap on one GPU share that array and race.  Results equal
   cfdl_calc_residual's bit for bit. */
int cfdl_calc_residual_noalloc(int ne, const int *row, const int *nbr, const double *ap,
                               const double *anb, const double *b, const double *phi,
                               double *res, double *res_max, void *stream);

/* Batches: continuous batching of independent cases (no counterpart in the
   reference, whose program solves one case per run, src/main.f90:25-38).
   cfdl_batch_run(b, ntstep, ncoef) does for every attached context what
   cfdl_run(c, ntstep, ncoef) does for one (the loop of src/main.f90:50-63),
   with one grouped launch per step for all cases: each case keeps its own
   fields, solver state, stop tests and log, and its results equal those of
   cfdl_run on the case alone, bit for bit.  Up to 32 cases per batch, all on
   the device current at cfdl_batch_create.  A case needs six-entry
   (hexahedral) rows, the colour order of solve_gs (cfdl_set_gs_order),
   pc order mgcg or mgcg-mixed (cfdl_set_pc_order), and the assembly and pc
   precision of the batch's first case.  While attached, a context's work
   (including cfdl_get_field, cfdl_log_*, and single steps) runs on the
   batch's stream, so cases attach and detach between cfdl_batch_run calls
   while others are part-way through their solves; cfdl_ctx_destroy detaches
   first.  cfdl_batch_set_block sets the lane count of the element-wise
   launches (a multiple of 64 in 64 .. 1024; 0, the default, lets the batch
   time the candidates at the first run after each change of composition and
   keep the fastest); results do not depend on it.  Each outer iteration
   runs as captured HIP graphs of its segments (the steps before the
   momentum solve, one momentum-solver iteration, the steps between the
   solves, one CG iteration, the steps after; captured once per composition
   and lane count, replayed while the host checks the cases' stop flags
   between solver iterations); cfdl_batch_set_graphs(b, 0) launches the
   same work directly, with the same results, and cfdl_batch_graph_stats
   counts captures and graph launches.  Not thread-safe. */
typedef struct cfdl_batch cfdl_batch;
int cfdl_batch_create(cfdl_batch **out);
int cfdl_batch_attach(cfdl_batch *b, cfdl_ctx *c);
int cfdl_batch_detach(cfdl_batch *b, cfdl_ctx *c);
int cfdl_batch_size(cfdl_batch *b, int *n);
int cfdl_batch_set_block(cfdl_batch *b, int block);
int cfdl_batch_block(cfdl_batch *b, int *block);
int cfdl_batch_set_graphs(cfdl_batch *b, int on);
int cfdl_batch_graph_stats(cfdl_batch *b, int64_t *captures, int64_t *launches);
int cfdl_batch_run(cfdl_batch *b, int ntstep, int ncoef);
int cfdl_batch_synchronize(cfdl_batch *b);
void cfdl_batch_destroy(cfdl_batch *b);

/* Launch counters and sampled kernel timing (no counterpart in the
   reference, which has no timers).  cfdl_set_profile(c, k) brackets every
   k-th pressure-correction sweep (all of its launches, back to back) with
   HIP events on the context stream (0 = off).  cfdl_stats synchronises and
   fills up to 8 values: [0] Gauss-Seidel launches (all solves),
   [1] pressure-correction sweep launches, [2] cells updated by those
   launches, [3] outer iterations, [4] launches inside sampled sweeps,
   [5] mean launch duration in sampled sweeps (ms: elapsed / launches),
   [6] mean cells per launch in sampled sweeps, [7] dependency levels per
   forward + backward pc sweep. */
int cfdl_set_profile(cfdl_ctx *c, int sample_every);
int cfdl_stats(cfdl_ctx *c, double *stats, int n);
int cfdl_stats_reset(cfdl_ctx *c);

#ifdef __cplusplus
}
#endif
#endif
