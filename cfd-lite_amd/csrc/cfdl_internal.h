// cfdl_internal.h -- shared host-side types of libcfdl_hip.so.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "cfdl.h"

namespace cfdl {

constexpr int kHexa8 = 17;  // CGNS element type ids, src/modules/mod_util.f90:16-39
constexpr int kQuad4 = 7;

// packed (element << 5) | local_face, src/modules/mod_util.f90:1428-1448
inline int pack(int e, int lf) { return (e << 5) | lf; }
inline int pk_elem(int p) { return static_cast<int>(static_cast<unsigned>(p) >> 5); }
inline int pk_face(int p) { return p & 31; }

void set_error(const std::string &msg);

}  // namespace cfdl

// The host-side setup product (arrays in the reference's own layout).
struct cfdl_mesh {
  int nvx = 0, nsec = 0, ne = 0, nbf = 0, nf = 0, L = 0, nelem = 0;
  std::vector<double> x, y, z;
  std::vector<int> etype, esec, e2vx;  // esec: 2 per section, e2vx: 8 per element
  std::vector<std::string> secname;
  std::vector<int> ef2nb_idx, ef2nb, s2g, bs;  // ef2nb: column-major L x 2
  std::vector<double> xc, yc, zc, vol, aip, rip;
  int nsubd = 0;
  std::vector<int> gf2g, g2gf, g2gf_idx;
};

namespace cfdl {
int mesh_build(cfdl_mesh *m);  // connectivity + geometry

// Agglomeration multigrid hierarchy of a CSR system (cfdl_mg.cpp).  Level 0
// is the given CSR (neighbours >= ne are boundary halos); lists named for
// the next level (agg, mptr/mem, iptr/ilist, gptr/glist) are empty on the
// coarsest level.
constexpr int kMgMaxLevels = 16;
struct MgLevel {
  int n = 0, nnz = 0, ncol = 0;
  std::vector<int> row, col;    // CSR, neighbours in increasing order (level >= 1)
  std::vector<int> cptr, cperm;  // greedy colours: cperm[cptr[c] .. cptr[c+1])
  std::vector<int> agg;          // node -> next-level aggregate
  std::vector<int> mptr, mem;    // next-level aggregate -> its nodes, increasing
  std::vector<int> iptr, ilist;  // next-level aggregate -> slots inside it
  std::vector<int> gptr, glist;  // next-level slot -> slots summed into it
};
struct MgHier {
  std::vector<MgLevel> lv;
};
int mg_build(int ne, const int *row, const int *nbr, const double *xc, const double *yc,
             const double *zc, int maxagg, int ncoarse, MgHier *out);

// one line of the device residual log (a linear solve's name, iterations,
// initial, final and maximum residual)
struct LogLine {
  char name[8];
  int it, pad;
  double res_i, res_f, res_max;
};
constexpr int kLogCap = 1 << 15;

// Multigrid-preconditioned conjugate gradients on a CSR system held in
// device memory (cfdl_mgcg.hip).  The solver owns the hierarchy's device
// arrays and its work vectors; the system's arrays belong to the caller.
struct MgCgSolver;
struct MgCgSystem {
  int ne, nbf, nit;
  const int *row, *nbr;               // device CSR, neighbours >= ne are halos
  const double *ap, *anb, *b;          // device coefficients (ne, L, ne)
  double *phi;                         // device values (ne + nbf), updated on cells
  char name[8];                        // log label
  LogLine *log;                        // device log and its count
  int *nlog;
};
int mgcg_create(const MgHier &h, int nbf, int fp32, MgCgSolver **out);
void mgcg_destroy(MgCgSolver *s);
// enqueues the solve on stream (a hipStream_t); synchronises with the host
// once per iteration to stop at the stop rule
int mgcg_solve(MgCgSolver *s, const MgCgSystem &sys, void *stream);
// iterations of the last solve and launches it made (diagnostics)
void mgcg_stats(const MgCgSolver *s, int *launches, int *levels);

// Grouped launches over independent systems (cfdl_group.h): slot ids and
// the first block of each in the launch.
constexpr int kMaxGroup = 32;
struct GroupMap {
  int n;
  int slot[kMaxGroup];
  int b0[kMaxGroup + 1];
};
// The same solve for n <= kMaxGroup systems at once, each with its own
// solver (all FP64 or all FP32 preconditioners), in grouped launches: every
// system's arithmetic and stop test are those of mgcg_solve alone.  The
// group object holds the systems' device descriptors (uploaded again when
// the composition changes) and the flag buffers; block is the lane count of
// the element-wise kernels (a multiple of 64 in 64 .. 1024; the results do
// not depend on it).
struct MgGroup;
int mgcg_group_create(MgGroup **out);
void mgcg_group_destroy(MgGroup *g);
int mgcg_solve_group(MgGroup *g, MgCgSolver *const *s, const MgCgSystem *sys, int n, void *stream,
                     int block);
// The same solve in phases, for callers that drive the iterations
// themselves (and may capture the enqueued phases in a graph): begin binds
// the systems and uploads their descriptors when they changed (*changed;
// it synchronises the stream then, so it runs outside any capture); setup,
// iteration and finish only enqueue work on the stream.  iteration copies
// the systems' active flags to host slot 0 or 1, which any_active reads
// once the stream has passed that point; nit is the largest iteration cap.
int mgcg_group_begin(MgGroup *g, MgCgSolver *const *s, const MgCgSystem *sys, int n, void *stream,
                     int block, int *changed);
int mgcg_group_setup(MgGroup *g, void *stream);
int mgcg_group_iteration(MgGroup *g, void *stream, int slot);
int mgcg_group_finish(MgGroup *g, void *stream);
int mgcg_group_nit(const MgGroup *g);
bool mgcg_group_any_active(const MgGroup *g, int slot);
}  // namespace cfdl
