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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

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

// r = b + sum anb phi_nb - ap phi; z = 0 (and r in FP32 for the mixed
// preconditioner); res_i, target and the active flag from rms(r)
template <typename T>
__global__ void __launch_bounds__(kRed)
k_cg_init(Red R, const int *__restrict__ row, const int *__restrict__ nbr,
          const double *__restrict__ ap, const double *__restrict__ anb,
          const double *__restrict__ b, const double *__restrict__ phi, double *__restrict__ r,
          T *__restrict__ rt, T *__restrict__ z, CgCtl *c, double *part, int nit) {
  __shared__ double s[kRed];
  const int lo = blockIdx.x * R.chunk, hi = min(lo + R.chunk, R.n);
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
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
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
__global__ void __launch_bounds__(kRed)
k_cg_rz(Red R, const double *__restrict__ r, const T *__restrict__ z, CgCtl *c, double *part) {
  __shared__ double s[kRed];
  if (!c->active) return;
  const int lo = blockIdx.x * R.chunk, hi = min(lo + R.chunk, R.n);
  double acc = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += kRed) acc = acc + r[i] * static_cast<double>(z[i]);
  const double bs = tree_sum(acc, s);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
  if (!arrive_last(c, R.nblk)) return;
  const double rho = final_sum(part, R.nblk, 1, s);
  if (threadIdx.x == 0) {
    c->rho = rho;
    c->beta = c->it == 0 ? 0.0 : rho / c->rho_old;
    c->ticket = 0;
  }
}

// p = z, or z + beta p
template <typename T>
__global__ void k_cg_p(int n, const T *__restrict__ z, double *__restrict__ p, const CgCtl *c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !c->active) return;
  if (c->it == 0)
    p[i] = static_cast<double>(z[i]);
  else
    p[i] = static_cast<double>(z[i]) + c->beta * p[i];
}

// q = ap p - sum anb p_nb, sigma = p.q, alpha = rho / sigma (a breakdown,
// sigma or rho not positive, ends the solve)
__global__ void __launch_bounds__(kRed)
k_cg_spmv(Red R, const int *__restrict__ row, const int *__restrict__ nbr,
          const double *__restrict__ ap, const double *__restrict__ anb,
          const double *__restrict__ p, double *__restrict__ q, CgCtl *c, double *part) {
  __shared__ double s[kRed];
  if (!c->active) return;
  const int lo = blockIdx.x * R.chunk, hi = min(lo + R.chunk, R.n);
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
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
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
__global__ void __launch_bounds__(kRed)
k_cg_update(Red R, double *__restrict__ phi, const double *__restrict__ p,
            const double *__restrict__ q, double *__restrict__ r, T *__restrict__ rt,
            T *__restrict__ z, CgCtl *c, double *part) {
  __shared__ double s[kRed];
  if (!c->active) return;
  const double alpha = c->alpha;
  const int lo = blockIdx.x * R.chunk, hi = min(lo + R.chunk, R.n);
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
    part[2 * blockIdx.x] = bs;
    part[2 * blockIdx.x + 1] = bm;
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

__global__ void k_cg_log(const CgCtl *c, LogLine *log, int *nlog, LogLine tmpl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
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

// ---------------------------------------------------------------------------
// V-cycle kernels on one level (T = double or float)
// ---------------------------------------------------------------------------
// one level as the kernels see it (device pointers)
template <typename T>
struct Lev {
  int n, nnz, ncol;
  const int *row, *col, *cperm, *cptr;  // cptr on the device (coarse kernel)
  const int *agg, *mptr, *mem, *iptr, *ilist, *gptr, *glist;
  const T *ap, *anb;
  T *b, *x;
  T *apw, *anbw;  // writable views of ap / anb (coarse levels, Galerkin)
};

template <typename T>
__device__ inline T row_sum(const Lev<T> &L, int i, T acc) {
  for (int k = L.row[i]; k < L.row[i + 1]; k++) acc = acc + L.anb[k] * L.x[L.col[k]];
  return acc;
}

// FP32 copies of the level-0 coefficients (mixed preconditioner)
__global__ void k_mg_copy0(int ne, int L, const double *__restrict__ ap,
                           const double *__restrict__ anb, float *__restrict__ ap32,
                           float *__restrict__ anb32) {
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L || i < ne; i += stride) {
    if (i < L) anb32[i] = static_cast<float>(anb[i]);
    if (i < ne) ap32[i] = static_cast<float>(ap[i]);
  }
}

// Galerkin operator of the next level: thread g < nC sums aggregate g's
// diagonal (members' ap in order, minus the inside entries in order), thread
// nC + s sums coarse slot s (its list of fine entries in order)
template <typename T>
__global__ void k_mg_galerkin(Lev<T> F, Lev<T> C) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
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
__global__ void k_mg_sweep(Lev<T> L, int c0, int cnt, const CgCtl *ctl) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt || !ctl->active) return;
  const int i = L.cperm[c0 + t];
  L.x[i] = row_sum(L, i, L.b[i]) / L.ap[i];
}

// restriction of the residual: b_C(I) = sum over members of
// (b_i + sum anb x_nb) - ap_i x_i, from zero; x_C(I) = 0
template <typename T>
__global__ void k_mg_restrict(Lev<T> F, Lev<T> C, const CgCtl *ctl) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
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

// prolongation: x_i += x_C(agg_i)
template <typename T>
__global__ void k_mg_prolong(Lev<T> F, const T *__restrict__ xc, const CgCtl *ctl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
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
__global__ void __launch_bounds__(kCoarseThreads)
k_mg_coarse(const Lev<T> *__restrict__ lv, int lc, int nlev, const CgCtl *ctl) {
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
  for (int k = 0; k < kCoarseSweeps; k++) {
    wg_sweep(lv[nlev - 1], false);
    wg_sweep(lv[nlev - 1], true);
  }
  for (int l = nlev - 2; l >= lc; l--) {
    const Lev<T> &F = lv[l];
    const T *xc = lv[l + 1].x;
    for (int i = threadIdx.x; i < F.n; i += blockDim.x) F.x[i] = F.x[i] + xc[F.agg[i]];
    __syncthreads();
    wg_sweep(F, true);
  }
}

inline int nb(int n, int b) { return (n + b - 1) / b; }

}  // namespace

// device copy of the hierarchy and the solve's work vectors
struct MgCgSolver {
  int fp32 = 0, ne = 0, nbf = 0, nlev = 0, lc = 0;
  std::vector<void *> mem;
  std::vector<std::vector<int>> cptr;  // host colour offsets per level
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
    hipLaunchKernelGGL(k_mg_restrict<T>, dim3(nb(lv[l + 1].n, kBlk)), dim3(kBlk), 0, st, lv[l],
                       lv[l + 1], ctl);
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
  for (int l = 0; l < s->nlev; l++) s->cptr.push_back(h.lv[l].cptr);
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

}  // namespace cfdl
