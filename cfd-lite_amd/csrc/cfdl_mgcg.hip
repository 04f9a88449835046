// cfdl_mgcg.hip -- multigrid-preconditioned conjugate gradients for the
// pressure-correction system on MI355X (an extension: the reference solves
// that system with the 4-subdomain Gauss-Seidel solver of
// src/modules/mod_solver.f90:124-189; its stop rule, rms residual at most a
// tenth of the initial one or nit iterations, :271-322, is kept).
//
// Design:
//  * Conjugate gradients in FP64 on the mesh CSR as it stands (ap, anb, b of
//    calc_coef_p, no copies), vectors in the global cell order.  Every dot
//    product and norm is a blocked reduction of fixed shape: block x of 256
//    lanes owns a contiguous chunk, lane t sums its strided elements in
//    order, the lanes combine as s[t] += s[t+h] for h = 128, 64 (LDS) and
//    32 .. 1 (wavefront shuffles), and the last block to finish (a ticket
//    counter) combines the block sums in the same way and does the scalar
//    work (alpha, beta, the stop test) in the same launch.  So each CG
//    iteration is four launches besides the preconditioner, and the result
//    does not depend on scheduling.
//  * The preconditioner is one V-cycle of agglomeration multigrid (the
//    hierarchy of cfdl_mg.cpp): Galerkin coarse operators summed once per
//    solve from lists (no atomics), one forward colour Gauss-Seidel sweep
//    before and one backward after each coarse correction, so the operator
//    is symmetric.  Levels of more than kCoarseMax nodes run as grid-wide
//    launches (one per colour); the small levels below run inside one
//    workgroup of 1024 lanes, which walks them with barriers instead of
//    kernel boundaries.  In the mixed variant the preconditioner runs in
//    FP32 (FP32 copies of ap, anb and of the residual); CG, its residual
//    and the stop test stay FP64.
//  * All kernels read the solve's active flag and return at once when it is
//    clear, so the host can queue an iteration before knowing whether the
//    previous one stopped the solve.
//  * Grouped solves (mgcg_solve_group) run the same steps for up to
//    kMaxGroup independent systems, one launch per step for all of them:
//    the launch's blocks are split into per-system ranges, each system keeps
//    its own vectors, reductions and stop test, so its results are those of
//    a solve alone, bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "cfdl_group.h"
#include "cfdl_internal.h"

namespace cfdl {

namespace {

#define MGCHK(x)                                                                    \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));                    \
      return CFDL_EHIP;                                                             \
    }                                                                               \
  } while (0)

constexpr int kRed = 256;          // lanes per reduction block
constexpr int kMaxRedBlocks = 1024;
constexpr int kBlk = 256;          // lanes per block of the grid-wide V-cycle kernels
constexpr int kCoarseMax = 4096;   // levels at most this large run in one workgroup
constexpr int kCoarseThreads = 1024;
constexpr int kCoarseSweeps = 10;  // forward + backward sweep pairs on the coarsest level
// a coarsest level at most this large is swept from LDS by one wavefront
constexpr int kLdsNodes = 64, kLdsEntries = 4096;

// control block of one solve
struct CgCtl {
  int it, active, ticket, nit;
  double res_i, res_f, res_max, target, rho, rho_old, alpha, beta;
};

// blocked reduction shape of an n-element vector
struct Red {
  int n, nblk, chunk;
};
Red red_of(int n) {
  int nb = (n + 4095) / 4096;
  nb = nb < 1 ? 1 : (nb > kMaxRedBlocks ? kMaxRedBlocks : nb);
  return Red{n, nb, (n + nb - 1) / nb};
}

// s[t] += s[t+h], h = 128, 64 in LDS, then 32 .. 1 by shuffles in wave 0;
// lane 0 returns the block's sum (the same pairing as the LDS tree)
__device__ inline double tree_sum(double v, double *s) {
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  if (t < 128) s[t] = s[t] + s[t + 128];
  __syncthreads();
  double x = 0.0;
  if (t < 64) {
    x = s[t] + s[t + 64];
#pragma unroll
    for (int h = 32; h >= 1; h >>= 1) x = x + __shfl_down(x, h, 64);
  }
  __syncthreads();
  return x;
}
__device__ inline double tree_max(double v, double *s) {
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  if (t < 128) s[t] = fmax(s[t], s[t + 128]);
  __syncthreads();
  double x = 0.0;
  if (t < 64) {
    x = fmax(s[t], s[t + 64]);
#pragma unroll
    for (int h = 32; h >= 1; h >>= 1) x = fmax(x, __shfl_down(x, h, 64));
  }
  __syncthreads();
  return x;
}

// Writes this block's partial(s) and returns true in every lane of the last
// block to arrive, after which the partials of all blocks are visible.
__device__ inline bool arrive_last(CgCtl *c, int nblk) {
  __shared__ int last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(&c->ticket, 1) == nblk - 1;
  __syncthreads();
  if (last) __threadfence();
  return last;
}
// the last block's sum of the nblk partials part[k * stride]
__device__ inline double final_sum(const double *part, int nblk, int stride, double *s) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRed) acc = acc + part[i * stride];
  return tree_sum(acc, s);
}
__device__ inline double final_max(const double *part, int nblk, int stride, double *s) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRed) acc = fmax(acc, part[i * stride]);
  return tree_max(acc, s);
}

// ---------------------------------------------------------------------------
// CG kernels.  Each step is a device body run by block bx of the system's own
// grid (bx = blockIdx.x for a solve alone; the block's index inside the
// system's range in a grouped launch) and two thin kernels around it.
// ---------------------------------------------------------------------------

// r = b + sum anb phi_nb - ap phi; z = 0 (and r in FP32 for the mixed
// preconditioner); res_i, target and the active flag from rms(r)
template <typename T>
__device__ inline void d_cg_init(int bx, Red R, const int *__restrict__ row,
                                 const int *__restrict__ nbr, const double *__restrict__ ap,
                                 const double *__restrict__ anb, const double *__restrict__ b,
                                 const double *__restrict__ phi, double *__restrict__ r,
                                 T *__restrict__ rt, T *__restrict__ z, CgCtl *c, double *part,
                                 int nit) {
  __shared__ double s[kRed];
  const int lo = bx * R.chunk, hi = min(lo + R.chunk, R.n);
  double acc = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += kRed) {
    double v = b[i];
    for (int k = row[i]; k < row[i + 1]; k++) v = v + anb[k] * phi[nbr[k]];
    v = v - ap[i] * phi[i];
    r[i] = v;
    if (rt) rt[i] = static_cast<T>(v);
    z[i] = static_cast<T>(0);
    acc = acc + v * v;
  }
  const double bs = tree_sum(acc, s);
  if (threadIdx.x == 0) part[bx] = bs;
  if (!arrive_last(c, R.nblk)) return;
  const double rr = final_sum(part, R.nblk, 1, s);
  if (threadIdx.x == 0) {
    const double ri = sqrt(rr / static_cast<double>(R.n));
    c->res_i = ri;
    c->res_f = ri;
    c->target = ri / 10.0;
    c->res_max = 0.0;
    c->it = 0;
    c->nit = nit;
    c->rho_old = 0.0;
    c->active = (0 < nit && ri > c->target) ? 1 : 0;
    c->ticket = 0;
  }
}

// rho = r.z, beta = rho / rho_old (0 on the first iteration)
template <typename T>
__device__ inline void d_cg_rz(int bx, Red R, const double *__restrict__ r,
                               const T *__restrict__ z, CgCtl *c, double *part) {
  __shared__ double s[kRed];
  if (!c->active) return;
  const int lo = bx * R.chunk, hi = min(lo + R.chunk, R.n);
  double acc = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += kRed) acc = acc + r[i] * static_cast<double>(z[i]);
  const double bs = tree_sum(acc, s);
  if (threadIdx.x == 0) part[bx] = bs;
  if (!arrive_last(c, R.nblk)) return;
  const double rho = final_sum(part, R.nblk, 1, s);
  if (threadIdx.x == 0) {
    c->rho = rho;
    c->beta = c->it == 0 ? 0.0 : rho / c->rho_old;
    c->ticket = 0;
  }
}

// p = z, or z + beta p (element i)
template <typename T>
__device__ inline void d_cg_p(int i, int n, const T *__restrict__ z, double *__restrict__ p,
                              const CgCtl *c) {
  if (i >= n || !c->active) return;
  if (c->it == 0)
    p[i] = static_cast<double>(z[i]);
  else
    p[i] = static_cast<double>(z[i]) + c->beta * p[i];
}

// q = ap p - sum anb p_nb, sigma = p.q, alpha = rho / sigma (a breakdown,
// sigma or rho not positive, ends the solve)
__device__ inline void d_cg_spmv(int bx, Red R, const int *__restrict__ row,
                                 const int *__restrict__ nbr, const double *__restrict__ ap,
                                 const double *__restrict__ anb, const double *__restrict__ p,
                                 double *__restrict__ q, CgCtl *c, double *part) {
  __shared__ double s[kRed];
  if (!c->active) return;
  const int lo = bx * R.chunk, hi = min(lo + R.chunk, R.n);
  double acc = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += kRed) {
    const int r0 = row[i], r1 = row[i + 1];
    double sm = 0.0;
    if (r1 - r0 == 6) {  // hexahedral row: all loads first, the sum in CSR order
      int nb[6];
      double an[6], v[6];
#pragma unroll
      for (int k = 0; k < 6; k++) {
        nb[k] = nbr[r0 + k];
        an[k] = anb[r0 + k];
      }
#pragma unroll
      for (int k = 0; k < 6; k++) v[k] = p[nb[k]];
#pragma unroll
      for (int k = 0; k < 6; k++) sm = sm + an[k] * v[k];
    } else {
      for (int k = r0; k < r1; k++) sm = sm + anb[k] * p[nbr[k]];
    }
    const double pi = p[i], qi = ap[i] * pi - sm;
    q[i] = qi;
    acc = acc + pi * qi;
  }
  const double bs = tree_sum(acc, s);
  if (threadIdx.x == 0) part[bx] = bs;
  if (!arrive_last(c, R.nblk)) return;
  const double sigma = final_sum(part, R.nblk, 1, s);
  if (threadIdx.x == 0) {
    if (!(sigma > 0.0) || !(c->rho > 0.0))
      c->active = 0;
    else
      c->alpha = c->rho / sigma;
    c->ticket = 0;
  }
}

// phi += alpha p, r -= alpha q, z = 0 for the next V-cycle; rms and max |r|;
// the iteration count and the stop test
template <typename T>
__device__ inline void d_cg_update(int bx, Red R, double *__restrict__ phi,
                                   const double *__restrict__ p, const double *__restrict__ q,
                                   double *__restrict__ r, T *__restrict__ rt, T *__restrict__ z,
                                   CgCtl *c, double *part) {
  __shared__ double s[kRed];
  if (!c->active) return;
  const double alpha = c->alpha;
  const int lo = bx * R.chunk, hi = min(lo + R.chunk, R.n);
  double acc = 0.0, mx = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += kRed) {
    phi[i] = phi[i] + alpha * p[i];
    const double v = r[i] - alpha * q[i];
    r[i] = v;
    if (rt) rt[i] = static_cast<T>(v);
    z[i] = static_cast<T>(0);
    acc = acc + v * v;
    mx = fmax(mx, fabs(v));
  }
  const double bs = tree_sum(acc, s), bm = tree_max(mx, s);
  if (threadIdx.x == 0) {
    part[2 * bx] = bs;
    part[2 * bx + 1] = bm;
  }
  if (!arrive_last(c, R.nblk)) return;
  const double rr = final_sum(part, R.nblk, 2, s);
  const double rm = final_max(part + 1, R.nblk, 2, s);
  if (threadIdx.x == 0) {
    c->it = c->it + 1;
    c->res_f = sqrt(rr / static_cast<double>(R.n));
    c->res_max = rm;
    c->rho_old = c->rho;
    c->active = (c->it < c->nit && c->res_f > c->target) ? 1 : 0;
    c->ticket = 0;
  }
}

__device__ inline void d_cg_log(const CgCtl *c, LogLine *log, int *nlog, const LogLine &tmpl) {
  const int k = *nlog;
  if (k >= kLogCap) return;
  LogLine l = tmpl;
  l.it = c->it;
  l.res_i = c->res_i;
  l.res_f = c->res_f;
  l.res_max = c->res_max;
  log[k] = l;
  *nlog = k + 1;
}

template <typename T>
__global__ void __launch_bounds__(kRed)
k_cg_init(Red R, const int *__restrict__ row, const int *__restrict__ nbr,
          const double *__restrict__ ap, const double *__restrict__ anb,
          const double *__restrict__ b, const double *__restrict__ phi, double *__restrict__ r,
          T *__restrict__ rt, T *__restrict__ z, CgCtl *c, double *part, int nit) {
  d_cg_init<T>(blockIdx.x, R, row, nbr, ap, anb, b, phi, r, rt, z, c, part, nit);
}
template <typename T>
__global__ void __launch_bounds__(kRed)
k_cg_rz(Red R, const double *__restrict__ r, const T *__restrict__ z, CgCtl *c, double *part) {
  d_cg_rz<T>(blockIdx.x, R, r, z, c, part);
}
template <typename T>
__global__ void k_cg_p(int n, const T *__restrict__ z, double *__restrict__ p, const CgCtl *c) {
  d_cg_p<T>(blockIdx.x * blockDim.x + threadIdx.x, n, z, p, c);
}
__global__ void __launch_bounds__(kRed)
k_cg_spmv(Red R, const int *__restrict__ row, const int *__restrict__ nbr,
          const double *__restrict__ ap, const double *__restrict__ anb,
          const double *__restrict__ p, double *__restrict__ q, CgCtl *c, double *part) {
  d_cg_spmv(blockIdx.x, R, row, nbr, ap, anb, p, q, c, part);
}
template <typename T>
__global__ void __launch_bounds__(kRed)
k_cg_update(Red R, double *__restrict__ phi, const double *__restrict__ p,
            const double *__restrict__ q, double *__restrict__ r, T *__restrict__ rt,
            T *__restrict__ z, CgCtl *c, double *part) {
  d_cg_update<T>(blockIdx.x, R, phi, p, q, r, rt, z, c, part);
}
__global__ void k_cg_log(const CgCtl *c, LogLine *log, int *nlog, LogLine tmpl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) d_cg_log(c, log, nlog, tmpl);
}

// ---------------------------------------------------------------------------
// V-cycle kernels on one level (T = double or float)
// ---------------------------------------------------------------------------
// one level as the kernels see it (device pointers)
template <typename T>
struct Lev {
  int n, nnz, ncol;
  const int *row, *col, *cperm, *cptr;  // cptr on the device (coarse and grouped kernels)
  const int *agg, *mptr, *mem, *iptr, *ilist, *gptr, *glist;
  const T *ap, *anb;
  T *b, *x;
  T *apw, *anbw;  // writable views of ap / anb (coarse levels, Galerkin)
};

// acc + sum over row i of anb * x[col], in CSR order; a six-entry
// (hexahedral) row issues all its loads before the sum
template <typename T>
__device__ inline T row_sum(const Lev<T> &L, int i, T acc) {
  const int r0 = L.row[i], r1 = L.row[i + 1];
  if (r1 - r0 == 6) {
    int c[6];
    T a[6], v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
      c[k] = L.col[r0 + k];
      a[k] = L.anb[r0 + k];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = L.x[c[k]];
#pragma unroll
    for (int k = 0; k < 6; k++) acc = acc + a[k] * v[k];
    return acc;
  }
  for (int k = r0; k < r1; k++) acc = acc + L.anb[k] * L.x[L.col[k]];
  return acc;
}

// FP32 copies of the level-0 coefficients (mixed preconditioner), element i
__device__ inline void d_mg_copy0(int i, int ne, int L, const double *__restrict__ ap,
                                  const double *__restrict__ anb, float *__restrict__ ap32,
                                  float *__restrict__ anb32) {
  if (i < L) anb32[i] = static_cast<float>(anb[i]);
  if (i < ne) ap32[i] = static_cast<float>(ap[i]);
}

// Galerkin operator of the next level: thread g < nC sums aggregate g's
// diagonal (members' ap in order, minus the inside entries in order), thread
// nC + s sums coarse slot s (its list of fine entries in order)
template <typename T>
__device__ inline void d_mg_galerkin(int g, const Lev<T> &F, const Lev<T> &C) {
  if (g < C.n) {
    T acc = static_cast<T>(0);
    for (int q = F.mptr[g]; q < F.mptr[g + 1]; q++) acc = acc + F.ap[F.mem[q]];
    for (int q = F.iptr[g]; q < F.iptr[g + 1]; q++) acc = acc - F.anb[F.ilist[q]];
    C.apw[g] = acc;
  } else if (g < C.n + C.nnz) {
    const int sl = g - C.n;
    T acc = static_cast<T>(0);
    for (int q = F.gptr[sl]; q < F.gptr[sl + 1]; q++) acc = acc + F.anb[F.glist[q]];
    C.anbw[sl] = acc;
  }
}

// one colour of a Gauss-Seidel sweep: x_i = (b_i + sum anb x_nb) / ap_i
template <typename T>
__device__ inline void d_mg_sweep(int t, const Lev<T> &L, int c0, int cnt, const CgCtl *ctl) {
  if (t >= cnt || !ctl->active) return;
  const int i = L.cperm[c0 + t];
  L.x[i] = row_sum(L, i, L.b[i]) / L.ap[i];
}

// restriction of the residual: b_C(I) = sum over members of
// (b_i + sum anb x_nb) - ap_i x_i, from zero; x_C(I) = 0
template <typename T>
__device__ inline void d_mg_restrict(int I, const Lev<T> &F, const Lev<T> &C, const CgCtl *ctl) {
  if (I >= C.n || !ctl->active) return;
  T acc = static_cast<T>(0);
  for (int q = F.mptr[I]; q < F.mptr[I + 1]; q++) {
    const int i = F.mem[q];
    const T t = row_sum(F, i, F.b[i]) - F.ap[i] * F.x[i];
    acc = acc + t;
  }
  C.b[I] = acc;
  C.x[I] = static_cast<T>(0);
}

// The same restriction with NL lanes per aggregate (NL a power of two, at
// least the aggregate's member count): lane j computes member j's term,
// then every lane of the group adds the terms in member order from
// shuffles, so the sum is the loop's above.  All lanes reach the shuffles
// (the active flag is the same for the whole launch of a system).
template <typename T, int NL>
__device__ inline void d_mg_restrict_lanes(int gt, const Lev<T> &F, const Lev<T> &C,
                                           const CgCtl *ctl) {
  if (!ctl->active) return;
  const int I = gt / NL, j = gt % NL;
  T t = static_cast<T>(0);
  int cnt = 0;
  if (I < C.n) {
    const int q0 = F.mptr[I];
    cnt = F.mptr[I + 1] - q0;
    if (j < cnt) {
      const int i = F.mem[q0 + j];
      t = row_sum(F, i, F.b[i]) - F.ap[i] * F.x[i];
    }
  }
  T acc = static_cast<T>(0);
#pragma unroll
  for (int jj = 0; jj < NL; jj++) {
    const T v = __shfl(t, jj, NL);
    if (jj < cnt) acc = acc + v;
  }
  if (j == 0 && I < C.n) {
    C.b[I] = acc;
    C.x[I] = static_cast<T>(0);
  }
}

// prolongation: x_i += x_C(agg_i)
template <typename T>
__device__ inline void d_mg_prolong(int i, const Lev<T> &F, const T *__restrict__ xc,
                                    const CgCtl *ctl) {
  if (i < F.n && ctl->active) F.x[i] = F.x[i] + xc[F.agg[i]];
}

// The small levels lc .. nlev-1 of the V-cycle in one workgroup: the same
// sweeps, restrictions and prolongations as the grid-wide kernels, with a
// barrier where those have a kernel boundary.  lv[lc].b is the restricted
// residual and lv[lc].x is zero on entry.
template <typename T>
__device__ void wg_sweep(const Lev<T> &L, bool backward) {
  for (int cc = 0; cc < L.ncol; cc++) {
    const int c = backward ? L.ncol - 1 - cc : cc;
    for (int q = L.cptr[c] + threadIdx.x; q < L.cptr[c + 1]; q += blockDim.x) {
      const int i = L.cperm[q];
      L.x[i] = row_sum(L, i, L.b[i]) / L.ap[i];
    }
    __syncthreads();
  }
}
template <typename T>
__device__ inline void d_mg_coarse(const Lev<T> *__restrict__ lv, int lc, int nlev,
                                   const CgCtl *ctl) {
  if (!ctl->active) return;
  for (int l = lc; l < nlev - 1; l++) {
    const Lev<T> &F = lv[l], &C = lv[l + 1];
    wg_sweep(F, false);
    for (int I = threadIdx.x; I < C.n; I += blockDim.x) {
      T acc = static_cast<T>(0);
      for (int q = F.mptr[I]; q < F.mptr[I + 1]; q++) {
        const int i = F.mem[q];
        const T t = row_sum(F, i, F.b[i]) - F.ap[i] * F.x[i];
        acc = acc + t;
      }
      C.b[I] = acc;
      C.x[I] = static_cast<T>(0);
    }
    __syncthreads();
  }
  const Lev<T> &Z = lv[nlev - 1];
  if (Z.n <= kLdsNodes && Z.nnz <= kLdsEntries && Z.ncol <= kLdsNodes) {
    // The coarsest level in LDS, swept by wavefront 0 alone: one wavefront
    // holds every node of a colour, and its LDS accesses complete in program
    // order, so no barrier is needed between colours.  The same operations
    // in the same order as wg_sweep.
    __shared__ T xs[kLdsNodes], bs[kLdsNodes], as[kLdsNodes], anbs[kLdsEntries];
    __shared__ int rows[kLdsNodes + 1], cols[kLdsEntries], perms[kLdsNodes], cps[kLdsNodes + 1];
    for (int i = threadIdx.x; i < Z.n; i += blockDim.x) {
      xs[i] = Z.x[i];
      bs[i] = Z.b[i];
      as[i] = Z.ap[i];
      perms[i] = Z.cperm[i];
    }
    for (int i = threadIdx.x; i <= Z.n; i += blockDim.x) rows[i] = Z.row[i];
    for (int i = threadIdx.x; i <= Z.ncol; i += blockDim.x) cps[i] = Z.cptr[i];
    for (int k = threadIdx.x; k < Z.nnz; k += blockDim.x) {
      cols[k] = Z.col[k];
      anbs[k] = Z.anb[k];
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      const int t = threadIdx.x;
      auto colour = [&](int c) {
        const int q = cps[c] + t;
        if (q < cps[c + 1]) {
          const int i = perms[q];
          T acc = bs[i];
          for (int k = rows[i]; k < rows[i + 1]; k++) acc = acc + anbs[k] * xs[cols[k]];
          xs[i] = acc / as[i];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      };
      for (int k = 0; k < kCoarseSweeps; k++) {
        for (int c = 0; c < Z.ncol; c++) colour(c);
        for (int c = Z.ncol - 1; c >= 0; c--) colour(c);
      }
      for (int i = t; i < Z.n; i += 64) Z.x[i] = xs[i];
    }
    __syncthreads();
  } else {
    for (int k = 0; k < kCoarseSweeps; k++) {
      wg_sweep(Z, false);
      wg_sweep(Z, true);
    }
  }
  for (int l = nlev - 2; l >= lc; l--) {
    const Lev<T> &F = lv[l];
    const T *xc = lv[l + 1].x;
    for (int i = threadIdx.x; i < F.n; i += blockDim.x) F.x[i] = F.x[i] + xc[F.agg[i]];
    __syncthreads();
    wg_sweep(F, true);
  }
}

__global__ void k_mg_copy0(int ne, int L, const double *__restrict__ ap,
                           const double *__restrict__ anb, float *__restrict__ ap32,
                           float *__restrict__ anb32) {
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L || i < ne; i += stride)
    d_mg_copy0(i, ne, L, ap, anb, ap32, anb32);
}
template <typename T>
__global__ void k_mg_galerkin(Lev<T> F, Lev<T> C) {
  d_mg_galerkin<T>(blockIdx.x * blockDim.x + threadIdx.x, F, C);
}
template <typename T>
__global__ void k_mg_sweep(Lev<T> L, int c0, int cnt, const CgCtl *ctl) {
  d_mg_sweep<T>(blockIdx.x * blockDim.x + threadIdx.x, L, c0, cnt, ctl);
}
template <typename T>
__global__ void k_mg_restrict(Lev<T> F, Lev<T> C, const CgCtl *ctl) {
  d_mg_restrict<T>(blockIdx.x * blockDim.x + threadIdx.x, F, C, ctl);
}
template <typename T, int NL>
__global__ void k_mg_restrict_lanes(Lev<T> F, Lev<T> C, const CgCtl *ctl) {
  d_mg_restrict_lanes<T, NL>(blockIdx.x * blockDim.x + threadIdx.x, F, C, ctl);
}
template <typename T>
__global__ void k_mg_prolong(Lev<T> F, const T *__restrict__ xc, const CgCtl *ctl) {
  d_mg_prolong<T>(blockIdx.x * blockDim.x + threadIdx.x, F, xc, ctl);
}
template <typename T>
__global__ void __launch_bounds__(kCoarseThreads)
k_mg_coarse(const Lev<T> *__restrict__ lv, int lc, int nlev, const CgCtl *ctl) {
  d_mg_coarse<T>(lv, lc, nlev, ctl);
}

// ---------------------------------------------------------------------------
// Grouped kernels: one launch runs a step for every system of a group.  Each
// system's descriptor (device memory) holds its vectors, its control block
// and its levels with level 0 bound to the system.
// ---------------------------------------------------------------------------
template <typename T>
struct MgSlot {
  Red R;
  int ne, nit, nlev, lc;
  const int *row, *nbr;
  const double *ap, *anb, *b;
  double *phi, *r, *p, *q, *part;
  T *rt, *z;              // FP32 residual copy (nullptr for FP64), preconditioned residual
  float *ap32, *anb32;    // FP32 level-0 coefficients (mixed)
  CgCtl *ctl;
  const Lev<T> *lv;       // device levels
  LogLine *log;
  int *nlog;
  LogLine tmpl;
};

template <typename T>
__global__ void g_mg_copy0(GroupMap g, const MgSlot<T> *__restrict__ S) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  d_mg_copy0(bx * blockDim.x + threadIdx.x, x.ne, x.lv[0].nnz, x.ap, x.anb, x.ap32, x.anb32);
}
template <typename T>
__global__ void g_mg_galerkin(GroupMap g, const MgSlot<T> *__restrict__ S, int l) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  const Lev<T> F = x.lv[l], C = x.lv[l + 1];
  d_mg_galerkin<T>(bx * blockDim.x + threadIdx.x, F, C);
}
template <typename T>
__global__ void __launch_bounds__(kRed) g_cg_init(GroupMap g, const MgSlot<T> *__restrict__ S) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  d_cg_init<T>(bx, x.R, x.row, x.nbr, x.ap, x.anb, x.b, x.phi, x.r, x.rt, x.z, x.ctl, x.part,
               x.nit);
}
// colour cc (counted from the last colour when backward) of level l, or of
// the system's coarsest level for l < 0
template <typename T>
__global__ void g_mg_sweep(GroupMap g, const MgSlot<T> *__restrict__ S, int l, int cc,
                           int backward) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  const Lev<T> L = x.lv[l < 0 ? x.nlev - 1 : l];
  const int c = backward ? L.ncol - 1 - cc : cc;
  const int c0 = L.cptr[c];
  d_mg_sweep<T>(bx * blockDim.x + threadIdx.x, L, c0, L.cptr[c + 1] - c0, x.ctl);
}
template <typename T>
__global__ void g_mg_restrict(GroupMap g, const MgSlot<T> *__restrict__ S, int l) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  const Lev<T> F = x.lv[l], C = x.lv[l + 1];
  d_mg_restrict<T>(bx * blockDim.x + threadIdx.x, F, C, x.ctl);
}
template <typename T, int NL>
__global__ void g_mg_restrict_lanes(GroupMap g, const MgSlot<T> *__restrict__ S, int l) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  const Lev<T> F = x.lv[l], C = x.lv[l + 1];
  d_mg_restrict_lanes<T, NL>(bx * blockDim.x + threadIdx.x, F, C, x.ctl);
}
template <typename T>
__global__ void g_mg_prolong(GroupMap g, const MgSlot<T> *__restrict__ S, int l) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  const Lev<T> F = x.lv[l];
  d_mg_prolong<T>(bx * blockDim.x + threadIdx.x, F, x.lv[l + 1].x, x.ctl);
}
template <typename T>
__global__ void __launch_bounds__(kCoarseThreads)
g_mg_coarse(GroupMap g, const MgSlot<T> *__restrict__ S) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  d_mg_coarse<T>(x.lv, x.lc, x.nlev, x.ctl);
}
template <typename T>
__global__ void __launch_bounds__(kRed) g_cg_rz(GroupMap g, const MgSlot<T> *__restrict__ S) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  d_cg_rz<T>(bx, x.R, x.r, x.z, x.ctl, x.part);
}
template <typename T>
__global__ void g_cg_p(GroupMap g, const MgSlot<T> *__restrict__ S) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  d_cg_p<T>(bx * blockDim.x + threadIdx.x, x.ne, x.z, x.p, x.ctl);
}
template <typename T>
__global__ void __launch_bounds__(kRed) g_cg_spmv(GroupMap g, const MgSlot<T> *__restrict__ S) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  d_cg_spmv(bx, x.R, x.row, x.nbr, x.ap, x.anb, x.p, x.q, x.ctl, x.part);
}
template <typename T>
__global__ void __launch_bounds__(kRed) g_cg_update(GroupMap g, const MgSlot<T> *__restrict__ S) {
  int bx;
  const MgSlot<T> &x = S[group_slot(g, bx)];
  d_cg_update<T>(bx, x.R, x.phi, x.p, x.q, x.r, x.rt, x.z, x.ctl, x.part);
}
// the active flags of systems 0 .. n-1, and their log lines at the end
template <typename T>
__global__ void g_cg_flags(int n, const MgSlot<T> *__restrict__ S, int *flags) {
  const int k = threadIdx.x;
  if (k < n) flags[k] = S[k].ctl->active;
}
template <typename T>
__global__ void g_cg_log(int n, const MgSlot<T> *__restrict__ S) {
  const int k = threadIdx.x;
  if (k < n) d_cg_log(S[k].ctl, S[k].log, S[k].nlog, S[k].tmpl);
}

inline int nb(int n, int b) { return (n + b - 1) / b; }

// lanes per aggregate of the restriction from a level whose aggregates have
// at most maxm members: the least power of two >= maxm, 0 (one lane per
// aggregate, k_mg_restrict) above a wavefront
inline int restrict_lanes(int maxm) {
  int nl = 1;
  while (nl < maxm) nl <<= 1;
  return nl <= 64 ? nl : 0;
}

// the restriction from level l of a solve alone
template <typename T>
void launch_restrict(int nl, const Lev<T> &F, const Lev<T> &C, const CgCtl *ctl, hipStream_t st) {
  const int n = C.n * (nl > 0 ? nl : 1);
  const dim3 g(nb(n, kBlk)), b(kBlk);
  switch (nl) {
    case 1: hipLaunchKernelGGL((k_mg_restrict_lanes<T, 1>), g, b, 0, st, F, C, ctl); break;
    case 2: hipLaunchKernelGGL((k_mg_restrict_lanes<T, 2>), g, b, 0, st, F, C, ctl); break;
    case 4: hipLaunchKernelGGL((k_mg_restrict_lanes<T, 4>), g, b, 0, st, F, C, ctl); break;
    case 8: hipLaunchKernelGGL((k_mg_restrict_lanes<T, 8>), g, b, 0, st, F, C, ctl); break;
    case 16: hipLaunchKernelGGL((k_mg_restrict_lanes<T, 16>), g, b, 0, st, F, C, ctl); break;
    case 32: hipLaunchKernelGGL((k_mg_restrict_lanes<T, 32>), g, b, 0, st, F, C, ctl); break;
    case 64: hipLaunchKernelGGL((k_mg_restrict_lanes<T, 64>), g, b, 0, st, F, C, ctl); break;
    default: hipLaunchKernelGGL(k_mg_restrict<T>, g, b, 0, st, F, C, ctl); break;
  }
}
// the same for a grouped launch (nbk blocks of blk lanes)
template <typename T>
void launch_restrict_group(int nl, int nbk, int blk, const GroupMap &gm, const MgSlot<T> *S, int l,
                           hipStream_t st) {
  const dim3 g(nbk), b(blk);
  switch (nl) {
    case 1: hipLaunchKernelGGL((g_mg_restrict_lanes<T, 1>), g, b, 0, st, gm, S, l); break;
    case 2: hipLaunchKernelGGL((g_mg_restrict_lanes<T, 2>), g, b, 0, st, gm, S, l); break;
    case 4: hipLaunchKernelGGL((g_mg_restrict_lanes<T, 4>), g, b, 0, st, gm, S, l); break;
    case 8: hipLaunchKernelGGL((g_mg_restrict_lanes<T, 8>), g, b, 0, st, gm, S, l); break;
    case 16: hipLaunchKernelGGL((g_mg_restrict_lanes<T, 16>), g, b, 0, st, gm, S, l); break;
    case 32: hipLaunchKernelGGL((g_mg_restrict_lanes<T, 32>), g, b, 0, st, gm, S, l); break;
    case 64: hipLaunchKernelGGL((g_mg_restrict_lanes<T, 64>), g, b, 0, st, gm, S, l); break;
    default: hipLaunchKernelGGL(g_mg_restrict<T>, g, b, 0, st, gm, S, l); break;
  }
}

}  // namespace
// device copy of the hierarchy and the solve's work vectors
struct MgCgSolver {
  int fp32 = 0, ne = 0, nbf = 0, nlev = 0, lc = 0;
  std::vector<void *> mem;
  std::vector<std::vector<int>> cptr;  // host colour offsets per level
  std::vector<int> lanes;              // restriction lanes per aggregate, per level
  // per level, for each real type (index 0 double, 1 float)
  std::vector<Lev<double>> lvd;
  std::vector<Lev<float>> lvf;
  Lev<double> *d_lvd = nullptr;
  Lev<float> *d_lvf = nullptr;
  double *r = nullptr, *p = nullptr, *q = nullptr, *zd = nullptr, *part = nullptr;
  float *r32 = nullptr, *z32 = nullptr, *ap32 = nullptr, *anb32 = nullptr;
  CgCtl *ctl = nullptr;
  int *hflag = nullptr;  // pinned
  hipEvent_t ev[2] = {nullptr, nullptr};
  int launches = 0;
  // grouped solves: the arrays level 0 was last bound to on the device
  // (d_lvd[0] / d_lvf[0]); nullptr until the first grouped solve
  const int *bound_row = nullptr;
  const double *bound_ap = nullptr;

  ~MgCgSolver() {
    for (void *m : mem) (void)hipFree(m);
    if (hflag) (void)hipHostFree(hflag);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  template <typename X>
  int alloc(X **p, size_t n) {
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, sizeof(X) * (n > 0 ? n : 1));
    if (e != hipSuccess) {
      set_error(std::string("hipMalloc (multigrid): ") + hipGetErrorString(e));
      return CFDL_ENOMEM;
    }
    mem.push_back(v);
    *p = static_cast<X *>(v);
    return CFDL_OK;
  }
  int upload(const std::vector<int> &v, const int **out) {
    int *d = nullptr;
    if (int rc = alloc(&d, v.size())) return rc;
    if (!v.empty()) MGCHK(hipMemcpy(d, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice));
    *out = d;
    return CFDL_OK;
  }
};

#define MGCALL(x)          \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

namespace {
template <typename T>
int fill_levels(MgCgSolver *s, const MgHier &h, std::vector<Lev<T>> &lv) {
  lv.assign(s->nlev, Lev<T>{});
  for (int l = 0; l < s->nlev; l++) {
    const MgLevel &H = h.lv[l];
    Lev<T> &L = lv[l];
    L.n = H.n;
    L.nnz = H.nnz;
    L.ncol = H.ncol;
    if (l > 0) {
      MGCALL(s->upload(H.row, &L.row));
      MGCALL(s->upload(H.col, &L.col));
      T *a = nullptr, *an = nullptr, *b = nullptr, *x = nullptr;
      MGCALL(s->alloc(&a, H.n));
      MGCALL(s->alloc(&an, H.nnz));
      MGCALL(s->alloc(&b, H.n));
      MGCALL(s->alloc(&x, H.n));
      L.ap = L.apw = a;
      L.anb = L.anbw = an;
      L.b = b;
      L.x = x;
    }
    MGCALL(s->upload(H.cperm, &L.cperm));
    MGCALL(s->upload(H.cptr, &L.cptr));
    if (l + 1 < s->nlev) {
      MGCALL(s->upload(H.agg, &L.agg));
      MGCALL(s->upload(H.mptr, &L.mptr));
      MGCALL(s->upload(H.mem, &L.mem));
      MGCALL(s->upload(H.iptr, &L.iptr));
      MGCALL(s->upload(H.ilist, &L.ilist));
      MGCALL(s->upload(H.gptr, &L.gptr));
      MGCALL(s->upload(H.glist, &L.glist));
    }
  }
  return CFDL_OK;
}

// one V-cycle: z = M r on level 0 (lv[0].b = r, lv[0].x = z, zero on entry)
template <typename T>
int vcycle(MgCgSolver *s, const std::vector<Lev<T>> &lv, const Lev<T> *d_lv, hipStream_t st,
           const CgCtl *ctl) {
  const int nlev = s->nlev, lc = s->lc;
  auto sweep = [&](int l, bool backward) {
    const Lev<T> &L = lv[l];
    for (int cc = 0; cc < L.ncol; cc++) {
      const int c = backward ? L.ncol - 1 - cc : cc;
      const int c0 = s->cptr[l][c], cnt = s->cptr[l][c + 1] - c0;
      hipLaunchKernelGGL(k_mg_sweep<T>, dim3(nb(cnt, kBlk)), dim3(kBlk), 0, st, L, c0, cnt, ctl);
      s->launches++;
    }
  };
  // grid-wide levels 0 .. lc-1 (the coarsest among them, if no small
  // levels follow, smoothed in place)
  const int top = lc < nlev ? lc : nlev - 1;
  for (int l = 0; l < top; l++) {
    sweep(l, false);
    launch_restrict<T>(s->lanes[l], lv[l], lv[l + 1], ctl, st);
    s->launches++;
  }
  if (lc < nlev) {
    hipLaunchKernelGGL(k_mg_coarse<T>, dim3(1), dim3(kCoarseThreads), 0, st, d_lv, lc, nlev, ctl);
    s->launches++;
  } else {
    for (int k = 0; k < kCoarseSweeps; k++) {
      sweep(nlev - 1, false);
      sweep(nlev - 1, true);
    }
  }
  for (int l = top - 1; l >= 0; l--) {
    hipLaunchKernelGGL(k_mg_prolong<T>, dim3(nb(lv[l].n, kBlk)), dim3(kBlk), 0, st, lv[l],
                       static_cast<const T *>(lv[l + 1].x), ctl);
    s->launches++;
    sweep(l, true);
  }
  MGCHK(hipGetLastError());
  return CFDL_OK;
}
}  // namespace

int mgcg_create(const MgHier &h, int nbf, int fp32, MgCgSolver **out) {
  std::unique_ptr<MgCgSolver> s(new MgCgSolver());
  s->fp32 = fp32 ? 1 : 0;
  s->nlev = static_cast<int>(h.lv.size());
  s->ne = h.lv[0].n;
  s->nbf = nbf;
  if (s->nlev < 1 || s->ne < 1) {
    set_error("mgcg_create: empty hierarchy");
    return CFDL_EINVAL;
  }
  // the small levels that run in one workgroup: the trailing levels (below
  // level 0) of at most kCoarseMax nodes
  s->lc = s->nlev;
  while (s->lc - 1 >= 1 && h.lv[s->lc - 1].n <= kCoarseMax) s->lc--;
  for (int l = 0; l < s->nlev; l++) {
    s->cptr.push_back(h.lv[l].cptr);
    int maxm = 1;
    const std::vector<int> &mp = h.lv[l].mptr;
    for (size_t g = 0; g + 1 < mp.size(); g++) maxm = std::max(maxm, mp[g + 1] - mp[g]);
    s->lanes.push_back(restrict_lanes(maxm));
  }
  const int ne = s->ne, nall = ne + nbf;
  MGCALL(s->alloc(&s->r, ne));
  MGCALL(s->alloc(&s->q, ne));
  MGCALL(s->alloc(&s->p, nall));
  MGCALL(s->alloc(&s->part, 2 * kMaxRedBlocks));
  MGCALL(s->alloc(&s->ctl, 1));
  MGCHK(hipMemset(s->ctl, 0, sizeof(CgCtl)));
  MGCHK(hipMemset(s->p, 0, sizeof(double) * nall));
  if (s->fp32) {
    MGCALL(s->alloc(&s->r32, ne));
    MGCALL(s->alloc(&s->z32, nall));
    MGCALL(s->alloc(&s->ap32, ne));
    MGCALL(s->alloc(&s->anb32, h.lv[0].nnz));
    MGCHK(hipMemset(s->z32, 0, sizeof(float) * nall));
    MGCALL(fill_levels<float>(s.get(), h, s->lvf));
    MGCALL(s->alloc(&s->d_lvf, s->nlev));
    // the workgroup kernel reads levels >= 1 only, which never change
    MGCHK(hipMemcpy(s->d_lvf, s->lvf.data(), sizeof(Lev<float>) * s->nlev, hipMemcpyHostToDevice));
  } else {
    MGCALL(s->alloc(&s->zd, nall));
    MGCHK(hipMemset(s->zd, 0, sizeof(double) * nall));
    MGCALL(fill_levels<double>(s.get(), h, s->lvd));
    MGCALL(s->alloc(&s->d_lvd, s->nlev));
    MGCHK(hipMemcpy(s->d_lvd, s->lvd.data(), sizeof(Lev<double>) * s->nlev, hipMemcpyHostToDevice));
  }
  MGCHK(hipHostMalloc(reinterpret_cast<void **>(&s->hflag), 2 * sizeof(int), hipHostMallocDefault));
  MGCHK(hipEventCreateWithFlags(&s->ev[0], hipEventDisableTiming));
  MGCHK(hipEventCreateWithFlags(&s->ev[1], hipEventDisableTiming));
  *out = s.release();
  return CFDL_OK;
}

void mgcg_destroy(MgCgSolver *s) { delete s; }

void mgcg_stats(const MgCgSolver *s, int *launches, int *levels) {
  if (launches) *launches = s ? s->launches : 0;
  if (levels) *levels = s ? s->nlev : 0;
}

int mgcg_solve(MgCgSolver *s, const MgCgSystem &sys, void *stream) {
  if (!s || sys.ne != s->ne || sys.nbf != s->nbf) {
    set_error("mgcg_solve: system does not match the hierarchy");
    return CFDL_EINVAL;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int ne = s->ne;
  const Red R = red_of(ne);
  s->launches = 0;
  // level 0 is the system itself (FP32 copies for the mixed preconditioner)
  if (s->fp32) {
    Lev<float> &z = s->lvf[0];
    z.row = sys.row;
    z.col = sys.nbr;
    const int Lnz = z.nnz;
    hipLaunchKernelGGL(k_mg_copy0, dim3(std::min(nb(std::max(Lnz, ne), kBlk), 8192)), dim3(kBlk), 0,
                       st, ne, Lnz, sys.ap, sys.anb, s->ap32, s->anb32);
    z.ap = z.apw = s->ap32;
    z.anb = z.anbw = s->anb32;
    z.b = s->r32;
    z.x = s->z32;
    for (int l = 0; l + 1 < s->nlev; l++)
      hipLaunchKernelGGL(k_mg_galerkin<float>, dim3(nb(s->lvf[l + 1].n + s->lvf[l + 1].nnz, kBlk)),
                         dim3(kBlk), 0, st, s->lvf[l], s->lvf[l + 1]);
    hipLaunchKernelGGL(k_cg_init<float>, dim3(R.nblk), dim3(kRed), 0, st, R, sys.row, sys.nbr,
                       sys.ap, sys.anb, sys.b, static_cast<const double *>(sys.phi), s->r, s->r32,
                       s->z32, s->ctl, s->part, sys.nit);
  } else {
    Lev<double> &z = s->lvd[0];
    z.row = sys.row;
    z.col = sys.nbr;
    z.ap = sys.ap;
    z.anb = sys.anb;
    z.apw = nullptr;
    z.anbw = nullptr;
    z.b = s->r;
    z.x = s->zd;
    for (int l = 0; l + 1 < s->nlev; l++)
      hipLaunchKernelGGL(k_mg_galerkin<double>, dim3(nb(s->lvd[l + 1].n + s->lvd[l + 1].nnz, kBlk)),
                         dim3(kBlk), 0, st, s->lvd[l], s->lvd[l + 1]);
    hipLaunchKernelGGL(k_cg_init<double>, dim3(R.nblk), dim3(kRed), 0, st, R, sys.row, sys.nbr,
                       sys.ap, sys.anb, sys.b, static_cast<const double *>(sys.phi), s->r,
                       static_cast<double *>(nullptr), s->zd, s->ctl, s->part, sys.nit);
  }
  MGCHK(hipGetLastError());
  int *act = &s->ctl->active;
  // Iteration k is queued before the host reads the flag after iteration
  // k-1, so the GPU never waits for the host; an iteration queued after the
  // stop exits at once.
  for (int it = 0; it < sys.nit; it++) {
    if (s->fp32) {
      MGCALL(vcycle<float>(s, s->lvf, s->d_lvf, st, s->ctl));
      hipLaunchKernelGGL(k_cg_rz<float>, dim3(R.nblk), dim3(kRed), 0, st, R,
                         static_cast<const double *>(s->r), static_cast<const float *>(s->z32),
                         s->ctl, s->part);
      hipLaunchKernelGGL(k_cg_p<float>, dim3(nb(ne, kBlk)), dim3(kBlk), 0, st, ne,
                         static_cast<const float *>(s->z32), s->p,
                         static_cast<const CgCtl *>(s->ctl));
    } else {
      MGCALL(vcycle<double>(s, s->lvd, s->d_lvd, st, s->ctl));
      hipLaunchKernelGGL(k_cg_rz<double>, dim3(R.nblk), dim3(kRed), 0, st, R,
                         static_cast<const double *>(s->r), static_cast<const double *>(s->zd),
                         s->ctl, s->part);
      hipLaunchKernelGGL(k_cg_p<double>, dim3(nb(ne, kBlk)), dim3(kBlk), 0, st, ne,
                         static_cast<const double *>(s->zd), s->p,
                         static_cast<const CgCtl *>(s->ctl));
    }
    hipLaunchKernelGGL(k_cg_spmv, dim3(R.nblk), dim3(kRed), 0, st, R, sys.row, sys.nbr, sys.ap,
                       sys.anb, static_cast<const double *>(s->p), s->q, s->ctl, s->part);
    if (s->fp32)
      hipLaunchKernelGGL(k_cg_update<float>, dim3(R.nblk), dim3(kRed), 0, st, R, sys.phi,
                         static_cast<const double *>(s->p), static_cast<const double *>(s->q),
                         s->r, s->r32, s->z32, s->ctl, s->part);
    else
      hipLaunchKernelGGL(k_cg_update<double>, dim3(R.nblk), dim3(kRed), 0, st, R, sys.phi,
                         static_cast<const double *>(s->p), static_cast<const double *>(s->q),
                         s->r, static_cast<double *>(nullptr), s->zd, s->ctl, s->part);
    s->launches += 4;
    MGCHK(hipGetLastError());
    MGCHK(hipMemcpyAsync(&s->hflag[it & 1], act, sizeof(int), hipMemcpyDeviceToHost, st));
    MGCHK(hipEventRecord(s->ev[it & 1], st));
    if (it > 0) {
      MGCHK(hipEventSynchronize(s->ev[(it - 1) & 1]));
      if (!s->hflag[(it - 1) & 1]) break;
    }
  }
  LogLine tmpl{};
  std::memcpy(tmpl.name, sys.name, 8);
  hipLaunchKernelGGL(k_cg_log, dim3(1), dim3(64), 0, st, static_cast<const CgCtl *>(s->ctl),
                     sys.log, sys.nlog, tmpl);
  MGCHK(hipGetLastError());
  return CFDL_OK;
}

// ---------------------------------------------------------------------------
// Grouped solves
// ---------------------------------------------------------------------------
namespace {
// A grouped solve between its phases (host side): the systems, their
// descriptors as uploaded, and the launch shape.
template <typename T>
struct GroupRun {
  std::vector<MgCgSolver *> ss;
  std::vector<MgSlot<T>> h;
  std::vector<Lev<T>> l0;
  int n = 0, blk = 256, maxtop = 0, maxlev = 0, nitmax = 0;
};
}  // namespace

struct MgGroup {
  int fp32 = 0;            // precision of the current run
  GroupRun<double> rd;
  GroupRun<float> rf;
  void *slots = nullptr;  // device MgSlot<T>[kMaxGroup]
  int *flags = nullptr;   // device active flags
  int *hflags = nullptr;  // pinned, two sets of kMaxGroup
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<char> key;  // the composition whose descriptors are on the device
  int launches = 0;
  ~MgGroup() {
    if (slots) (void)hipFree(slots);
    if (flags) (void)hipFree(flags);
    if (hflags) (void)hipHostFree(hflags);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

int mgcg_group_create(MgGroup **out) {
  std::unique_ptr<MgGroup> g(new MgGroup());
  const size_t bytes = sizeof(MgSlot<double>) > sizeof(MgSlot<float>) ? sizeof(MgSlot<double>)
                                                                       : sizeof(MgSlot<float>);
  MGCHK(hipMalloc(&g->slots, bytes * kMaxGroup));
  MGCHK(hipMalloc(reinterpret_cast<void **>(&g->flags), sizeof(int) * kMaxGroup));
  MGCHK(hipHostMalloc(reinterpret_cast<void **>(&g->hflags), 2 * sizeof(int) * kMaxGroup,
                      hipHostMallocDefault));
  MGCHK(hipEventCreateWithFlags(&g->ev[0], hipEventDisableTiming));
  MGCHK(hipEventCreateWithFlags(&g->ev[1], hipEventDisableTiming));
  *out = g.release();
  return CFDL_OK;
}

void mgcg_group_destroy(MgGroup *g) { delete g; }

namespace {
// Level 0 of a grouped solve lives on the device (the grouped kernels read
// every level from there); it is bound to the system's arrays once, with the
// stream drained, and again only if they change.
template <typename T>
int bind_level0(MgCgSolver *s, const MgCgSystem &sys, std::vector<Lev<T>> &lv, Lev<T> *d_lv,
                hipStream_t st, Lev<T> *out) {
  Lev<T> z = lv[0];
  z.row = sys.row;
  z.col = sys.nbr;
  if (s->fp32) {
    z.ap = z.apw = reinterpret_cast<T *>(s->ap32);
    z.anb = z.anbw = reinterpret_cast<T *>(s->anb32);
    z.b = reinterpret_cast<T *>(s->r32);
    z.x = reinterpret_cast<T *>(s->z32);
  } else {
    z.ap = reinterpret_cast<const T *>(sys.ap);
    z.anb = reinterpret_cast<const T *>(sys.anb);
    z.apw = z.anbw = nullptr;
    z.b = reinterpret_cast<T *>(s->r);
    z.x = reinterpret_cast<T *>(s->zd);
  }
  if (s->bound_row != sys.row || s->bound_ap != sys.ap) {
    MGCHK(hipStreamSynchronize(st));
    MGCHK(hipMemcpy(d_lv, &z, sizeof(Lev<T>), hipMemcpyHostToDevice));
    s->bound_row = sys.row;
    s->bound_ap = sys.ap;
  }
  *out = z;
  return CFDL_OK;
}

template <typename T>
Lev<T> *dev_levels(MgCgSolver *s);
template <>
Lev<double> *dev_levels<double>(MgCgSolver *s) { return s->d_lvd; }
template <>
Lev<float> *dev_levels<float>(MgCgSolver *s) { return s->d_lvf; }
template <typename T>
std::vector<Lev<T>> &host_levels(MgCgSolver *s);
template <>
std::vector<Lev<double>> &host_levels<double>(MgCgSolver *s) { return s->lvd; }
template <>
std::vector<Lev<float>> &host_levels<float>(MgCgSolver *s) { return s->lvf; }

template <typename T>
GroupRun<T> &run_of(MgGroup *G);
template <>
GroupRun<double> &run_of<double>(MgGroup *G) { return G->rd; }
template <>
GroupRun<float> &run_of<float>(MgGroup *G) { return G->rf; }

// The solve of mgcg_solve for n systems at once, in phases: begin (level 0
// bound, descriptors uploaded when they changed; synchronises then, so it
// runs outside any graph capture), setup (FP32 copies, Galerkin operators,
// initial residual), one iteration at a time (V-cycle and CG step, then the
// systems' active flags copied to a host slot), finish (the log lines).
// Every launch covers the systems that take part in its step; each system
// runs exactly its own sequence of steps (a system with fewer levels or
// colours sits out the launches it has no part in), so its arithmetic is
// that of a solve alone.
template <typename T>
int group_begin(MgGroup *G, MgCgSolver *const *ss, const MgCgSystem *sys, int n, hipStream_t st,
                int blk, int *changed) {
  GroupRun<T> &R = run_of<T>(G);
  R.ss.assign(ss, ss + n);
  R.n = n;
  R.blk = blk;
  R.h.assign(n, MgSlot<T>{});
  std::memset(static_cast<void *>(R.h.data()), 0, sizeof(MgSlot<T>) * n);  // the key compares bytes
  R.l0.assign(n, Lev<T>{});
  R.maxtop = R.maxlev = R.nitmax = 0;
  for (int k = 0; k < n; k++) {
    MgCgSolver *s = ss[k];
    MGCALL(bind_level0<T>(s, sys[k], host_levels<T>(s), dev_levels<T>(s), st, &R.l0[k]));
    MgSlot<T> &x = R.h[k];
    x.R = red_of(s->ne);
    x.ne = s->ne;
    x.nit = sys[k].nit;
    x.nlev = s->nlev;
    x.lc = s->lc;
    x.row = sys[k].row;
    x.nbr = sys[k].nbr;
    x.ap = sys[k].ap;
    x.anb = sys[k].anb;
    x.b = sys[k].b;
    x.phi = sys[k].phi;
    x.r = s->r;
    x.p = s->p;
    x.q = s->q;
    x.part = s->part;
    x.rt = s->fp32 ? reinterpret_cast<T *>(s->r32) : nullptr;
    x.z = s->fp32 ? reinterpret_cast<T *>(s->z32) : reinterpret_cast<T *>(s->zd);
    x.ap32 = s->ap32;
    x.anb32 = s->anb32;
    x.ctl = s->ctl;
    x.lv = dev_levels<T>(s);
    x.log = sys[k].log;
    x.nlog = sys[k].nlog;
    std::memcpy(x.tmpl.name, sys[k].name, 8);
    R.maxtop = std::max(R.maxtop, s->lc < s->nlev ? s->lc : s->nlev - 1);
    R.maxlev = std::max(R.maxlev, s->nlev);
    R.nitmax = std::max(R.nitmax, sys[k].nit);
  }
  // the descriptors go to the device when the composition changes
  const char *kb = reinterpret_cast<const char *>(R.h.data());
  std::vector<char> key(kb, kb + sizeof(MgSlot<T>) * n);
  key.push_back(static_cast<char>(sizeof(T)));
  *changed = key != G->key;
  if (*changed) {
    MGCHK(hipStreamSynchronize(st));
    MGCHK(hipMemcpy(G->slots, R.h.data(), sizeof(MgSlot<T>) * n, hipMemcpyHostToDevice));
    G->key.swap(key);
  }
  G->launches = 0;
  return CFDL_OK;
}

// the launch shapes of a run
template <typename T>
struct GroupShape {
  GroupRun<T> &R;
  GroupMap g{};
  std::vector<int> ids, cnt;
  explicit GroupShape(GroupRun<T> &r) : R(r), ids(r.n), cnt(r.n) {
    for (int k = 0; k < R.n; k++) ids[k] = k;
  }
  // one grouped launch: blocks_of(k) blocks for system k (0 = sits out)
  template <typename F>
  int map(F blocks_of) {
    for (int k = 0; k < R.n; k++) cnt[k] = blocks_of(k);
    return group_map(g, ids.data(), cnt.data(), R.n);
  }
  int top_of(int k) const {
    return R.ss[k]->lc < R.ss[k]->nlev ? R.ss[k]->lc : R.ss[k]->nlev - 1;
  }
};

template <typename T>
int group_setup(MgGroup *G, hipStream_t st) {
  GroupRun<T> &R = run_of<T>(G);
  GroupShape<T> sh(R);
  const MgSlot<T> *S = static_cast<const MgSlot<T> *>(G->slots);
  const int blk = R.blk;
  int nbk = 0;
  if (std::is_same<T, float>::value) {
    nbk = sh.map([&](int k) { return nb(std::max(R.l0[k].nnz, R.ss[k]->ne), blk); });
    if (nbk > 0) {
      hipLaunchKernelGGL(g_mg_copy0<T>, dim3(nbk), dim3(blk), 0, st, sh.g, S);
      G->launches++;
    }
  }
  for (int l = 0; l + 1 < R.maxlev; l++) {
    nbk = sh.map([&](int k) {
      if (l + 1 >= R.ss[k]->nlev) return 0;
      const Lev<T> &C = host_levels<T>(R.ss[k])[l + 1];
      return nb(C.n + C.nnz, blk);
    });
    if (nbk > 0) {
      hipLaunchKernelGGL(g_mg_galerkin<T>, dim3(nbk), dim3(blk), 0, st, sh.g, S, l);
      G->launches++;
    }
  }
  nbk = sh.map([&](int k) { return R.h[k].R.nblk; });
  hipLaunchKernelGGL(g_cg_init<T>, dim3(nbk), dim3(kRed), 0, st, sh.g, S);
  G->launches++;
  MGCHK(hipGetLastError());
  return CFDL_OK;
}

template <typename T>
int group_iteration(MgGroup *G, hipStream_t st, int slot) {
  GroupRun<T> &R = run_of<T>(G);
  GroupShape<T> sh(R);
  const MgSlot<T> *S = static_cast<const MgSlot<T> *>(G->slots);
  const int n = R.n, blk = R.blk;
  auto colour_count = [&](int k, int l, int c) {
    return R.ss[k]->cptr[l][c + 1] - R.ss[k]->cptr[l][c];
  };
  // colour cc of level l (l < 0: the coarsest level of systems without small
  // levels), forward or backward
  auto sweep = [&](int l, bool backward) {
    int ncmax = 0;
    for (int k = 0; k < n; k++) {
      const int lk = l < 0 ? R.ss[k]->nlev - 1 : l;
      const bool part = l < 0 ? R.ss[k]->lc == R.ss[k]->nlev : l < sh.top_of(k);
      if (part) ncmax = std::max(ncmax, host_levels<T>(R.ss[k])[lk].ncol);
    }
    for (int cc = 0; cc < ncmax; cc++) {
      const int nbk = sh.map([&](int k) {
        const int lk = l < 0 ? R.ss[k]->nlev - 1 : l;
        const bool part = l < 0 ? R.ss[k]->lc == R.ss[k]->nlev : l < sh.top_of(k);
        const int nc = host_levels<T>(R.ss[k])[lk].ncol;
        if (!part || cc >= nc) return 0;
        const int c = backward ? nc - 1 - cc : cc;
        return nb(colour_count(k, lk, c), blk);
      });
      if (nbk > 0) {
        hipLaunchKernelGGL(g_mg_sweep<T>, dim3(nbk), dim3(blk), 0, st, sh.g, S, l, cc,
                           backward ? 1 : 0);
        G->launches++;
      }
    }
  };
  int nbk = 0;
  // V-cycle
  for (int l = 0; l < R.maxtop; l++) {
    sweep(l, false);
    // one lane count for the launch: the largest of the systems taking part
    // (0, one lane per aggregate, if any needs it); sums do not depend on it
    int nl = 1;
    for (int k = 0; k < n; k++)
      if (l < sh.top_of(k))
        nl = (nl == 0 || R.ss[k]->lanes[l] == 0) ? 0 : std::max(nl, R.ss[k]->lanes[l]);
    nbk = sh.map([&](int k) {
      return l < sh.top_of(k) ? nb(host_levels<T>(R.ss[k])[l + 1].n * (nl > 0 ? nl : 1), blk) : 0;
    });
    if (nbk > 0) {
      launch_restrict_group<T>(nl, nbk, blk, sh.g, S, l, st);
      G->launches++;
    }
  }
  nbk = sh.map([&](int k) { return R.ss[k]->lc < R.ss[k]->nlev ? 1 : 0; });
  if (nbk > 0) {
    hipLaunchKernelGGL(g_mg_coarse<T>, dim3(nbk), dim3(kCoarseThreads), 0, st, sh.g, S);
    G->launches++;
  }
  bool any_flat = false;
  for (int k = 0; k < n; k++) any_flat = any_flat || R.ss[k]->lc == R.ss[k]->nlev;
  if (any_flat)
    for (int r = 0; r < kCoarseSweeps; r++) {
      sweep(-1, false);
      sweep(-1, true);
    }
  for (int l = R.maxtop - 1; l >= 0; l--) {
    nbk = sh.map([&](int k) { return l < sh.top_of(k) ? nb(host_levels<T>(R.ss[k])[l].n, blk) : 0; });
    if (nbk > 0) {
      hipLaunchKernelGGL(g_mg_prolong<T>, dim3(nbk), dim3(blk), 0, st, sh.g, S, l);
      G->launches++;
    }
    sweep(l, true);
  }
  // CG
  nbk = sh.map([&](int k) { return R.h[k].R.nblk; });
  hipLaunchKernelGGL(g_cg_rz<T>, dim3(nbk), dim3(kRed), 0, st, sh.g, S);
  const GroupMap gR = sh.g;
  nbk = sh.map([&](int k) { return nb(R.ss[k]->ne, blk); });
  hipLaunchKernelGGL(g_cg_p<T>, dim3(nbk), dim3(blk), 0, st, sh.g, S);
  hipLaunchKernelGGL(g_cg_spmv<T>, dim3(gR.b0[gR.n]), dim3(kRed), 0, st, gR, S);
  hipLaunchKernelGGL(g_cg_update<T>, dim3(gR.b0[gR.n]), dim3(kRed), 0, st, gR, S);
  hipLaunchKernelGGL(g_cg_flags<T>, dim3(1), dim3(kMaxGroup), 0, st, n, S, G->flags);
  G->launches += 5;
  MGCHK(hipGetLastError());
  MGCHK(hipMemcpyAsync(G->hflags + slot * kMaxGroup, G->flags, sizeof(int) * n,
                       hipMemcpyDeviceToHost, st));
  return CFDL_OK;
}

template <typename T>
int group_finish(MgGroup *G, hipStream_t st) {
  GroupRun<T> &R = run_of<T>(G);
  const MgSlot<T> *S = static_cast<const MgSlot<T> *>(G->slots);
  hipLaunchKernelGGL(g_cg_log<T>, dim3(1), dim3(kMaxGroup), 0, st, R.n, S);
  G->launches++;
  MGCHK(hipGetLastError());
  for (MgCgSolver *s : R.ss) s->launches = G->launches;
  return CFDL_OK;
}
}  // namespace

int mgcg_group_begin(MgGroup *G, MgCgSolver *const *s, const MgCgSystem *sys, int n, void *stream,
                     int block, int *changed) {
  if (!G || !s || !sys || n < 1 || n > kMaxGroup) {
    set_error("mgcg_solve_group: 1 .. " + std::to_string(kMaxGroup) + " systems");
    return CFDL_EINVAL;
  }
  if (block < 64 || block > 1024 || block % 64 != 0) {
    set_error("mgcg_solve_group: block must be a multiple of 64 in 64 .. 1024");
    return CFDL_EINVAL;
  }
  for (int k = 0; k < n; k++) {
    if (!s[k] || sys[k].ne != s[k]->ne || sys[k].nbf != s[k]->nbf) {
      set_error("mgcg_solve_group: system " + std::to_string(k) + " does not match its hierarchy");
      return CFDL_EINVAL;
    }
    if (s[k]->fp32 != s[0]->fp32) {
      set_error("mgcg_solve_group: the systems of a group share one preconditioner precision");
      return CFDL_EINVAL;
    }
    for (int j = 0; j < k; j++)
      if (s[j] == s[k]) {
        set_error("mgcg_solve_group: a solver appears twice in the group");
        return CFDL_EINVAL;
      }
  }
  int dummy = 0;
  if (!changed) changed = &dummy;
  hipStream_t st = static_cast<hipStream_t>(stream);
  G->fp32 = s[0]->fp32;
  return G->fp32 ? group_begin<float>(G, s, sys, n, st, block, changed)
                 : group_begin<double>(G, s, sys, n, st, block, changed);
}

int mgcg_group_setup(MgGroup *G, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  return G->fp32 ? group_setup<float>(G, st) : group_setup<double>(G, st);
}

int mgcg_group_iteration(MgGroup *G, void *stream, int slot) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  return G->fp32 ? group_iteration<float>(G, st, slot) : group_iteration<double>(G, st, slot);
}

int mgcg_group_finish(MgGroup *G, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  return G->fp32 ? group_finish<float>(G, st) : group_finish<double>(G, st);
}

int mgcg_group_nit(const MgGroup *G) { return G->fp32 ? G->rf.nitmax : G->rd.nitmax; }

bool mgcg_group_any_active(const MgGroup *G, int slot) {
  const int n = G->fp32 ? G->rf.n : G->rd.n;
  const int *f = G->hflags + slot * kMaxGroup;
  for (int k = 0; k < n; k++)
    if (f[k]) return true;
  return false;
}

int mgcg_solve_group(MgGroup *G, MgCgSolver *const *s, const MgCgSystem *sys, int n,
                     void *stream, int block) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  MGCALL(mgcg_group_begin(G, s, sys, n, stream, block, nullptr));
  MGCALL(mgcg_group_setup(G, stream));
  // iteration it is queued before the host reads the flags after it-1
  for (int it = 0; it < mgcg_group_nit(G); it++) {
    MGCALL(mgcg_group_iteration(G, stream, it & 1));
    MGCHK(hipEventRecord(G->ev[it & 1], st));
    if (it > 0) {
      MGCHK(hipEventSynchronize(G->ev[(it - 1) & 1]));
      if (!mgcg_group_any_active(G, (it - 1) & 1)) break;
    }
  }
  return mgcg_group_finish(G, stream);
}

}  // namespace cfdl
