// cfdl_device.hip -- device context, FP64 kernels and the C ABI of the hot
// path (CFD-Lite solve_uvwp, src/equations/mod_uvwp.f90:95-134, and the
// solvers of src/modules/mod_solver.f90) for MI355X / gfx950.
//
// Numerics: every kernel evaluates the reference's expressions in the
// reference's order (Fortran a*b/c = (a*b)/c, a-b-c = (a-b)-c) and the file
// is compiled with -ffp-contract=off, so each per-cell and per-face value is
// bit-identical to the reference's.  Independent cell / face loops become one
// thread per cell / face, reading neighbours through the CSR (gathers, no
// atomics).  The Gauss-Seidel sweeps keep the reference's exact order: cells
// are grouped by dependency level (a cell's level is one more than the
// highest level among its already-updated neighbours) and each level is one
// launch, so every cell sees exactly the neighbour values it sees in the
// sequential sweep.  Norms are fixed-order tree reductions (reproducible;
// equal to the sequential sums up to rounding in the last bits).
//
// Device layout (HBM, 0-based; cells 0..ne-1, boundary halos ne..ne+nbf-1):
//   row[ne+1], nbr[L], fsg[L]   CSR of ef2nb: neighbour, signed 1-based face
//   own[nf], fcell[nf]          owner-side CSR slot / owner cell of a face
//   bcrow[ne+1], bcslot[nbf]    per cell, its boundary slots in halo order
//   hcell[nbf], hslot[nbf], hkind[nbf]   per halo: owner cell, slot, BC kind
//   fields: element arrays of ne+nbf, gradients 3 per element (x,y,z).
#include <execinfo.h>
#include <hip/hip_runtime.h>
#include <signal.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <type_traits>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "cfdl_group.h"
#include "cfdl_internal.h"

namespace cfdl {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// CFDL_TRACE=1 prints set-up stages to stderr and a native backtrace on a
// segmentation fault (diagnostics for host-side faults).
static bool trace_on() {
  static const bool on = [] {
    const char *v = std::getenv("CFDL_TRACE");
    return v && v[0] == '1';
  }();
  return on;
}
static void segv_handler(int sig) {
  void *frames[64];
  const int n = backtrace(frames, 64);
  const char msg[] = "cfdl: fatal signal, native backtrace:\n";
  (void)!write(2, msg, sizeof(msg) - 1);
  backtrace_symbols_fd(frames, n, 2);
  signal(sig, SIG_DFL);
  raise(sig);
}
static void trace(const char *stage) {
  if (!trace_on()) return;
  static bool installed = false;
  if (!installed) {
    signal(SIGSEGV, segv_handler);
    installed = true;
  }
  std::fprintf(stderr, "cfdl: %s\n", stage);
  std::fflush(stderr);
}

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));                    \
      return CFDL_EHIP;                                                             \
    }                                                                               \
  } while (0)

#define CALL(x)              \
  do {                       \
    int rc_ = (x);           \
    if (rc_) return rc_;     \
  } while (0)

constexpr int kBlock = 256;
constexpr int kMaxSys = 16;
// multigrid hierarchy of the MGCG pc solve: at most kMgAgg^l cells per
// level-l aggregate; levels are added while more than kMgCoarse remain
constexpr int kMgAgg = 8, kMgCoarse = 64;

struct MeshDev {
  int ne, nf, nbf, L;
  const int *row, *nbr, *fsg, *own, *fcell, *bcrow, *bcslot, *hcell, *hslot, *hkind;
  const double *xc, *yc, *zc, *vol, *aip, *rip, *rho, *mu;
  const double *cp, *tc;  // thermal properties of the energy equation
  // FP32 copies of the geometry and properties for the FP32 assembly (null
  // until cfdl_set_assembly_precision selects it)
  const float *xc32, *yc32, *zc32, *vol32, *aip32, *rip32, *rho32, *mu32;
};

// the geometry of the arithmetic type T: the mesh's FP64 arrays, or their
// FP32 copies
struct Geo64 {
  const double *xc, *yc, *zc, *vol, *aip, *rip, *rho, *mu;
};
struct Geo32 {
  const float *xc, *yc, *zc, *vol, *aip, *rip, *rho, *mu;
};
template <typename T>
__device__ inline auto geo_of(const MeshDev &m) {
  if constexpr (sizeof(T) == 8)
    return Geo64{m.xc, m.yc, m.zc, m.vol, m.aip, m.rip, m.rho, m.mu};
  else
    return Geo32{m.xc32, m.yc32, m.zc32, m.vol32, m.aip32, m.rip32, m.rho32, m.mu32};
}

// energy equation state (energy_t, mod_energy.f90:4-10): temperature t and
// its gradient gt, the enthalpy h it solves for (phi), h0, gradient gh
struct EnergyDev {
  double *t, *gt, *h, *h0, *gh;
};

struct FieldsDev {
  double *u, *v, *w, *p, *u0, *v0, *w0;
  double *gu, *gv, *gw, *gp, *gpc;
  double *mip, *mip0;
  double *bu, *bv, *bw, *d, *dc;
  double *ap, *b, *anb, *phic;
};

// state of a solve_gs call (up to 3 right-hand sides sharing ap/anb)
struct GsCtl {
  int it[3], active[3];
  double res_i[3], res_f[3], res_max[3], target[3];
};

// state of multi_subdomain_solver, src/modules/mod_solver.f90:124-189
struct PcCtl {
  int it, active;
  double ri, rf, rmt, target;
};


// A linear system for the residual / sweep kernels: ny right-hand sides or
// subdomains, each a contiguous cell range sharing row/nbr/ap/anb.
struct SysDesc {
  int ny;
  int beg[kMaxSys], end[kMaxSys];
  double *phi[kMaxSys];
  const double *b[kMaxSys];
  const int *row, *nbr;
  const double *ap, *anb;
  const double *rap;  // 1/ap per cell (tiled LDS sweeps), correctly rounded
};

// A single right-hand side in FP32: the correction system of the
// mixed-precision pressure-correction solve (ap, anb, residual, correction)
struct SysDesc32 {
  const int *row, *nbr;
  const float *ap, *anb, *b;
  float *phi;
};

// (x / a) / s, correctly rounded at each step, from y = 1/a and ys = 1/s
// correctly rounded.  A quotient estimate q0 = x*y corrected once by the
// exact remainder, q0 + (x - a*q0)*y, is the correctly rounded x/a
// (Markstein's theorem) while q0, the remainder and r*y stay normal: |q0|
// in [2^-960, 2^960], or q0 = 0.  The same step then divides by s.  Only
// the sign of a zero result can come out wrong, and for s > 0 the final
// sign is that of x/a, i.e. of q0 = x*y: one copysign at the end fixes it
// (no intermediate is a zero of the wrong sign unless x/a is zero).  This
// takes the divisions off the Gauss-Seidel critical path, since y and ys
// are known before the sweep.  Callers keep the operands in range (rcp_of,
// kRcpExp, s in [2^-64, 2^64]) and fall back to the hardware division
// outside it.
__device__ inline double div_rcp2(double x, double a, double y, double s, double ys) {
  const double q0 = x * y;
  const double q1 = fma(fma(-a, q0, x), y, q0);
  const double p0 = q1 * ys;
  return __builtin_copysign(fma(fma(-s, p0, q1), ys, p0), q0);
}

// Operand ranges of the reciprocal division: divisors (ap, sor) with
// |a| in [2^-64, 2^64] and dividends x = 0 or with frexp exponent in
// [-kRcpExp, kRcpExp] put every quotient estimate of two chained divisions
// (x/ap, then /sor) within [2^-880, 2^880] or at zero.
constexpr int kRcpExp = 750;
constexpr double kRcpDivMin = 0x1p-64, kRcpDivMax = 0x1p64;

// 1/a for div_rcp2, NaN for a divisor outside [2^-64, 2^64] (zero, huge, NaN
// included): every quotient that uses it is then NaN, which the sweep
// detects and recomputes with the hardware division
__device__ inline double rcp_of(double a) {
  const double m = fabs(a);
  return (m >= kRcpDivMin && m <= kRcpDivMax) ? 1.0 / a : __builtin_nan("");
}

// The helpers are templates on the arithmetic type: FP64 everywhere, FP32 in
// the FP32 assembly (cfdl_set_assembly_precision), with the same operations
// in the same order.
template <typename T>
__device__ inline T dot3(const T *a, const T *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

template <typename T>
__device__ inline T area_of(const T *a) {
  return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

// vec_weight, src/modules/mod_util.f90:729-743 (r0 = face centre)
template <typename T>
__device__ inline T vec_weight(const T *r0, const T *r1, const T *r2) {
  T ra[3], rb[3];
  for (int i = 0; i < 3; i++) {
    ra[i] = r1[i] - r0[i];
    rb[i] = r2[i] - r0[i];
  }
  const T la = sqrt(ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2]);
  const T lb = sqrt(rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2]);
  if (la + lb > static_cast<T>(0)) return la / (la + lb);
  return static_cast<T>(0.5);
}

__device__ inline void centre(const MeshDev &m, int e, double *r) {
  r[0] = m.xc[e];
  r[1] = m.yc[e];
  r[2] = m.zc[e];
}

// ---------------------------------------------------------------------------
// BC setters lid / dirichlet0 / symmetry (mod_uvwp.f90:493-570), per halo.
// ---------------------------------------------------------------------------
// Inlet profile of the channel extension: umax, y0, y1, z0, z1
struct Inlet {
  double q[5];
};

__device__ inline void d_bc_update(int bx_, MeshDev m, FieldsDev f, Inlet in) {
  const int h = bx_ * blockDim.x + threadIdx.x;
  if (h >= m.nbf) return;
  const int e = m.hcell[h], H = m.ne + h, kind = m.hkind[h];
  const int fg = abs(m.fsg[m.hslot[h]]) - 1;
  if (kind == CFDL_BC_INLET || kind == CFDL_BC_OUTLET) {
    // channel extension: inlet velocity along the inward normal with the
    // parabolic profile at the face centre, outlet velocity and pressure
    // from the owner; the face mass flux rho_P (v.n) A, the outlet's later
    // scaled to the inflow (k_outlet_flux)
    const double *a = m.aip + 3 * fg;
    const double area = area_of(a);
    const double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
    double vel[3];
    if (kind == CFDL_BC_INLET) {
      const double sy = (m.yc[H] - in.q[1]) / (in.q[2] - in.q[1]);
      const double sz = (m.zc[H] - in.q[3]) / (in.q[4] - in.q[3]);
      const double un = in.q[0] * 16.0 * sy * (1.0 - sy) * sz * (1.0 - sz);
      for (int i = 0; i < 3; i++) vel[i] = un * -nrm[i];
    } else {
      vel[0] = f.u[e];
      vel[1] = f.v[e];
      vel[2] = f.w[e];
    }
    f.u[H] = vel[0];
    f.v[H] = vel[1];
    f.w[H] = vel[2];
    f.p[H] = f.p[e];
    f.mip[fg] = dot3(vel, nrm) * m.rho[e] * area;
    return;
  }
  if (kind == CFDL_BC_SYMMETRY) {
    const double *a = m.aip + 3 * fg;
    const double area = area_of(a);
    const double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
    const double vel[3] = {f.u[e], f.v[e], f.w[e]};
    const double t = dot3(vel, nrm);
    double vn[3], vt[3];
    for (int i = 0; i < 3; i++) {
      vn[i] = t * nrm[i];
      vt[i] = vel[i] - vn[i];
    }
    f.u[H] = vt[0] - 2.0 * vn[0];
    f.v[H] = vt[1] - 2.0 * vn[1];
    f.w[H] = vt[2] - 2.0 * vn[2];
  } else {
    f.u[H] = kind == CFDL_BC_LID ? 1.0 : 0.0;
    f.v[H] = 0.0;
    f.w[H] = 0.0;
  }
  f.p[H] = f.p[e];
  f.mip[fg] = 0.0;
}
__global__ void k_bc_update(MeshDev m, FieldsDev f, Inlet in) {
  d_bc_update(blockIdx.x, m, f, in);
}

// Outlet mass flux of the channel extension: Q_in (inlet faces) and the
// extrapolated Q_out (outlet faces), each summed in the blocked tree order
// (blocks of ceil(n / nblk) faces, nblk = min(1024, ceil(n / 4096)); lane t
// of 256 sums its strided faces, lanes combined pairwise s[t] += s[t + h],
// h = 128 .. 1, block sums combined the same way), here all in one
// workgroup that walks the blocks in turn; then the outlet fluxes are
// scaled by -Q_in / Q_out, or, when Q_out is not positive, replaced by the
// inflow spread by face area.  Face lists are in halo order.
constexpr int kOutThreads = 256;
__device__ double wg_tree(double v, double *s) {
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int h = kOutThreads / 2; h >= 1; h >>= 1) {
    if (t < h) s[t] = s[t] + s[t + h];
    __syncthreads();
  }
  const double r = s[0];
  __syncthreads();
  return r;
}
// sum of v(k) for k < n in the blocked order; part holds 1024 block sums
template <typename F>
__device__ double wg_blocked_sum(int n, F v, double *s, double *part) {
  int nblk = (n + 4095) / 4096;
  nblk = nblk < 1 ? 1 : (nblk > 1024 ? 1024 : nblk);
  const int chunk = (n + nblk - 1) / nblk;
  for (int x = 0; x < nblk; x++) {
    const int lo = x * chunk, hi = min(lo + chunk, n);
    double acc = 0.0;
    for (int k = lo + threadIdx.x; k < hi; k += kOutThreads) acc = acc + v(k);
    const double bs = wg_tree(acc, s);
    if (threadIdx.x == 0) part[x] = bs;
  }
  __syncthreads();
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kOutThreads) acc = acc + part[i];
  return wg_tree(acc, s);
}
__device__ inline void d_outlet_flux(int bx_, MeshDev m, double *mip, const int *fin, int nin, const int *fout, int nout) {
  __shared__ double s[kOutThreads], part[1024];
  if (nout == 0) return;
  const double qin = wg_blocked_sum(nin, [&](int k) { return mip[fin[k]]; }, s, part);
  const double qout = wg_blocked_sum(nout, [&](int k) { return mip[fout[k]]; }, s, part);
  const double aout = wg_blocked_sum(nout, [&](int k) { return area_of(m.aip + 3 * fout[k]); }, s, part);
  if (qout > 0.0) {
    const double sc = -qin / qout;
    for (int k = threadIdx.x; k < nout; k += kOutThreads) mip[fout[k]] = mip[fout[k]] * sc;
  } else {
    for (int k = threadIdx.x; k < nout; k += kOutThreads)
      mip[fout[k]] = -qin * area_of(m.aip + 3 * fout[k]) / aout;
  }
}
__global__ void __launch_bounds__(kOutThreads)
k_outlet_flux(MeshDev m, double *mip, const int *fin, int nin, const int *fout, int nout) {
  d_outlet_flux(blockIdx.x, m, mip, fin, nin, fout, nout);
}

// ---------------------------------------------------------------------------
// calc_coef_uvw, mod_uvwp.f90:161-286: one thread per cell; the boundary
// loop (:243-273) runs for the cell's own boundary faces in halo order, then
// the Rhie-Chow / SIMPLEC coefficients (:276-284).
// ---------------------------------------------------------------------------
__global__ void k_coef_uvw(MeshDev m, FieldsDev f, double dt, double alpha_u) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  double rp[3];
  centre(m, e, rp);
  double ap = 0.0, sumf = 0.0, sumss[3] = {0.0, 0.0, 0.0}, sumdefc[3] = {0.0, 0.0, 0.0};
  const int r0 = m.row[e], r1 = m.row[e + 1];
  for (int k = r0; k < r1; k++) {
    const int nb = m.nbr[k];
    double fnb = 0.0, d = 0.0;
    if (nb < m.ne) {
      const int fs = m.fsg[k];
      const double sg = fs >= 0 ? 1.0 : -1.0;
      const int fg = abs(fs) - 1;
      const double *a = m.aip + 3 * fg, *rip = m.rip + 3 * fg;
      const double area = area_of(a);
      double nrm[3], rpnb[3], dr[3], drip[3], rpp[3], rnp[3], drp[3];
      for (int i = 0; i < 3; i++) nrm[i] = sg * a[i] / area;
      centre(m, nb, rpnb);
      for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
      const double ds = sqrt(dot3(dr, dr));
      const double wt = vec_weight(rip, rp, rpnb);
      for (int i = 0; i < 3; i++) drip[i] = rip[i] - rp[i];
      double t = dot3(drip, nrm);
      for (int i = 0; i < 3; i++) rpp[i] = rip[i] - t * nrm[i];
      for (int i = 0; i < 3; i++) drip[i] = rip[i] - rpnb[i];
      t = dot3(drip, nrm);
      for (int i = 0; i < 3; i++) rnp[i] = rip[i] - t * nrm[i];
      for (int i = 0; i < 3; i++) drp[i] = rnp[i] - rpp[i];
      const double dsp = sqrt(dot3(drp, drp));
      const double fl = -(sg * f.mip[fg]);
      fnb = fmax(fl, 0.0);
      sumf = sumf + fl;
      const double muip = (1.0 - wt) * m.mu[e] + wt * m.mu[nb];
      d = muip * area / ds;
      const double ge[3][3] = {{f.gu[3 * e], f.gu[3 * e + 1], f.gu[3 * e + 2]},
                               {f.gv[3 * e], f.gv[3 * e + 1], f.gv[3 * e + 2]},
                               {f.gw[3 * e], f.gw[3 * e + 1], f.gw[3 * e + 2]}};
      const double gn[3][3] = {{f.gu[3 * nb], f.gu[3 * nb + 1], f.gu[3 * nb + 2]},
                               {f.gv[3 * nb], f.gv[3 * nb + 1], f.gv[3 * nb + 2]},
                               {f.gw[3 * nb], f.gw[3 * nb + 1], f.gw[3 * nb + 2]}};
      // secondary stress, :213-218 (gip = m-th column of grad(u,v,w))
      for (int mm = 0; mm < 3; mm++) {
        double gip[3];
        for (int q = 0; q < 3; q++) gip[q] = (1.0 - wt) * ge[q][mm] + wt * gn[q][mm];
        sumss[mm] = sumss[mm] + muip * area * dot3(gip, dr) / ds;
      }
      // deferred non-orthogonal correction, :220-225
      for (int q = 0; q < 3; q++) {
        double gip[3];
        for (int i = 0; i < 3; i++) gip[i] = (1.0 - wt) * ge[q][i] + wt * gn[q][i];
        sumdefc[q] = sumdefc[q] + muip * area * (dot3(gip, drp) / dsp - dot3(gip, dr) / ds);
      }
    }
    f.anb[k] = d + fnb;
    ap = ap + d + fnb;
  }
  const double vol = m.vol[e];
  const double ap0 = m.rho[e] * vol / dt;
  ap = ap + ap0;
  double bu = ap0 * f.u0[e] + sumf * f.u[e] - vol * f.gp[3 * e] + sumss[0] + sumdefc[0];
  double bv = ap0 * f.v0[e] + sumf * f.v[e] - vol * f.gp[3 * e + 1] + sumss[1] + sumdefc[1];
  double bw = ap0 * f.w0[e] + sumf * f.w[e] - vol * f.gp[3 * e + 2] + sumss[2] + sumdefc[2];
  for (int q = m.bcrow[e]; q < m.bcrow[e + 1]; q++) {
    const int k = m.bcslot[q], H = m.nbr[k];
    const int fg = m.fsg[k] - 1;  // boundary faces are always owner side
    const double *a = m.aip + 3 * fg;
    const double area = area_of(a);
    const double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
    const double dr[3] = {m.xc[H] - rp[0], m.yc[H] - rp[1], m.zc[H] - rp[2]};
    const double ds = sqrt(dot3(dr, dr));
    const double fz = 0.0, d = m.mu[e] * area / ds;
    const int kind = m.hkind[H - m.ne];
    if (kind == CFDL_BC_INLET || kind == CFDL_BC_OUTLET) {
      // channel extension: the face's convective flux as on interior faces
      // (anb += max(f, 0), f * u_P into the right-hand side as sumf),
      // diffusion at the inlet only (zero-gradient outlet)
      const double fl = -f.mip[fg], fnb = fmax(fl, 0.0);
      const double dd = kind == CFDL_BC_INLET ? d : 0.0;
      bu = bu + fl * f.u[e];
      bv = bv + fl * f.v[e];
      bw = bw + fl * f.w[e];
      ap = ap + dd + fnb;
      f.anb[k] = f.anb[k] + dd + fnb;
      continue;
    }
    if (kind != CFDL_BC_SYMMETRY) {  // 'dirichlet', :255-265
      const double ue = f.u[e], ve = f.v[e], we = f.w[e];
      double vr[3] = {ue, ve, we};
      const double t = dot3(vr, nrm);
      for (int i = 0; i < 3; i++) vr[i] = vr[i] - t * nrm[i];
      vr[0] = f.u[H] - vr[0];
      vr[1] = f.v[H] - vr[1];
      vr[2] = f.w[H] - vr[2];
      bu = bu + d * vr[0] - d * ue;
      bv = bv + d * vr[1] - d * ve;
      bw = bw + d * vr[2] - d * we;
    }
    ap = ap + d + fz;
    f.anb[k] = f.anb[k] + d + fz;
  }
  // config mode (no counterpart in the reference): Patankar implicit
  // under-relaxation before the Rhie-Chow / SIMPLEC coefficients; skipped at
  // alpha_u = 1 so that reference mode stays bit-identical.
  if (alpha_u != 1.0) {
    const double apr = ap / alpha_u, da = apr - ap;
    bu = bu + da * f.u[e];
    bv = bv + da * f.v[e];
    bw = bw + da * f.w[e];
    ap = apr;
  }
  f.ap[e] = ap;
  f.bu[e] = bu;
  f.bv[e] = bv;
  f.bw[e] = bw;
  f.d[e] = vol / ap;
  double dc = ap;
  for (int k = r0; k < r1; k++) dc = dc - f.anb[k];
  f.dc[e] = vol / dc;
}

// The same assembly for meshes whose rows all have six entries (row e at 6e),
// six threads per cell: thread j of a cell computes the terms of its row
// entry j (the loop body above) into LDS, then the cell's first thread adds
// them in CSR order exactly as the loop does (interior entries only for the
// flux, secondary-stress and deferred-correction sums) and finishes the cell.
// Interior anb are written by their entry's thread, boundary anb by the cell.
constexpr int kFaceThreads = 256, kFaceCells = kFaceThreads / 6;
template <typename T>
__device__ inline void d_coef_uvw6(int bx_, MeshDev m, FieldsDev f, double dt_, double alpha_u_) {
  // T = double: the reference's FP64 assembly bit for bit.  T = float: the
  // same operations in FP32 on the FP32 geometry copies and on the fields
  // rounded to FP32 (the FP32 assembly; results stored as FP64).
  const auto G = geo_of<T>(m);
  const T one = static_cast<T>(1), zero = static_cast<T>(0);
  const T dt = static_cast<T>(dt_), alpha_u = static_cast<T>(alpha_u_);
  auto ld = [](double x) { return static_cast<T>(x); };
  __shared__ T st[9][kFaceThreads];  // d, fnb, fl, secondary stress, deferred correction
  __shared__ int sin_[kFaceThreads];  // the entry is interior
  const int lc = threadIdx.x / 6, j = threadIdx.x - 6 * lc;
  const int e = bx_ * kFaceCells + lc;
  const bool valid = lc < kFaceCells && e < m.ne;
  if (valid) {
    const int k = 6 * e + j;
    const int nb = m.nbr[k];
    T fnb = zero, d = zero, fl = zero, ss[3] = {zero, zero, zero}, dfc[3] = {zero, zero, zero};
    if (nb < m.ne) {
      const T rp[3] = {G.xc[e], G.yc[e], G.zc[e]};
      const int fs = m.fsg[k];
      const T sg = fs >= 0 ? one : -one;
      const int fg = abs(fs) - 1;
      const T *a = G.aip + 3 * fg, *rip = G.rip + 3 * fg;
      const T area = area_of(a);
      T nrm[3], dr[3], drip[3], rpp[3], rnp[3], drp[3];
      for (int i = 0; i < 3; i++) nrm[i] = sg * a[i] / area;
      const T rpnb[3] = {G.xc[nb], G.yc[nb], G.zc[nb]};
      for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
      const T ds = sqrt(dot3(dr, dr));
      const T wt = vec_weight(rip, rp, rpnb);
      for (int i = 0; i < 3; i++) drip[i] = rip[i] - rp[i];
      T t = dot3(drip, nrm);
      for (int i = 0; i < 3; i++) rpp[i] = rip[i] - t * nrm[i];
      for (int i = 0; i < 3; i++) drip[i] = rip[i] - rpnb[i];
      t = dot3(drip, nrm);
      for (int i = 0; i < 3; i++) rnp[i] = rip[i] - t * nrm[i];
      for (int i = 0; i < 3; i++) drp[i] = rnp[i] - rpp[i];
      const T dsp = sqrt(dot3(drp, drp));
      fl = -(sg * ld(f.mip[fg]));
      fnb = fmax(fl, zero);
      const T muip = (one - wt) * G.mu[e] + wt * G.mu[nb];
      d = muip * area / ds;
      const T ge[3][3] = {{ld(f.gu[3 * e]), ld(f.gu[3 * e + 1]), ld(f.gu[3 * e + 2])},
                          {ld(f.gv[3 * e]), ld(f.gv[3 * e + 1]), ld(f.gv[3 * e + 2])},
                          {ld(f.gw[3 * e]), ld(f.gw[3 * e + 1]), ld(f.gw[3 * e + 2])}};
      const T gn[3][3] = {{ld(f.gu[3 * nb]), ld(f.gu[3 * nb + 1]), ld(f.gu[3 * nb + 2])},
                          {ld(f.gv[3 * nb]), ld(f.gv[3 * nb + 1]), ld(f.gv[3 * nb + 2])},
                          {ld(f.gw[3 * nb]), ld(f.gw[3 * nb + 1]), ld(f.gw[3 * nb + 2])}};
      for (int mm = 0; mm < 3; mm++) {
        T gip[3];
        for (int q = 0; q < 3; q++) gip[q] = (one - wt) * ge[q][mm] + wt * gn[q][mm];
        ss[mm] = muip * area * dot3(gip, dr) / ds;
      }
      for (int q = 0; q < 3; q++) {
        T gip[3];
        for (int i = 0; i < 3; i++) gip[i] = (one - wt) * ge[q][i] + wt * gn[q][i];
        dfc[q] = muip * area * (dot3(gip, drp) / dsp - dot3(gip, dr) / ds);
      }
      f.anb[k] = static_cast<double>(d + fnb);
    }
    st[0][threadIdx.x] = d;
    st[1][threadIdx.x] = fnb;
    st[2][threadIdx.x] = fl;
    for (int i = 0; i < 3; i++) {
      st[3 + i][threadIdx.x] = ss[i];
      st[6 + i][threadIdx.x] = dfc[i];
    }
    sin_[threadIdx.x] = nb < m.ne;
  }
  __syncthreads();
  if (!valid || j != 0) return;
  const T rp[3] = {G.xc[e], G.yc[e], G.zc[e]};
  T ap = zero, sumf = zero, sumss[3] = {zero, zero, zero}, sumdefc[3] = {zero, zero, zero};
  T anbk[6];
  for (int q = 0; q < 6; q++) {
    const int t = threadIdx.x + q;
    const T d = st[0][t], fnb = st[1][t];
    if (sin_[t]) {
      sumf = sumf + st[2][t];
      for (int i = 0; i < 3; i++) sumss[i] = sumss[i] + st[3 + i][t];
      for (int i = 0; i < 3; i++) sumdefc[i] = sumdefc[i] + st[6 + i][t];
    }
    anbk[q] = d + fnb;
    ap = ap + d + fnb;
  }
  const T vol = G.vol[e];
  const T ap0 = G.rho[e] * vol / dt;
  ap = ap + ap0;
  const T ue = ld(f.u[e]), ve = ld(f.v[e]), we = ld(f.w[e]);
  T bu = ap0 * ld(f.u0[e]) + sumf * ue - vol * ld(f.gp[3 * e]) + sumss[0] + sumdefc[0];
  T bv = ap0 * ld(f.v0[e]) + sumf * ve - vol * ld(f.gp[3 * e + 1]) + sumss[1] + sumdefc[1];
  T bw = ap0 * ld(f.w0[e]) + sumf * we - vol * ld(f.gp[3 * e + 2]) + sumss[2] + sumdefc[2];
  for (int q = 0; q < 6; q++)
    if (!sin_[threadIdx.x + q]) f.anb[6 * e + q] = static_cast<double>(anbk[q]);
  for (int q = m.bcrow[e]; q < m.bcrow[e + 1]; q++) {
    const int k = m.bcslot[q], H = m.nbr[k];
    const int fg = m.fsg[k] - 1;  // boundary faces are always owner side
    const T *a = G.aip + 3 * fg;
    const T area = area_of(a);
    const T nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
    const T dr[3] = {G.xc[H] - rp[0], G.yc[H] - rp[1], G.zc[H] - rp[2]};
    const T ds = sqrt(dot3(dr, dr));
    const T fz = zero, d = G.mu[e] * area / ds;
    const int kind = m.hkind[H - m.ne];
    if (kind == CFDL_BC_INLET || kind == CFDL_BC_OUTLET) {
      // channel extension: the face's convective flux as on interior faces
      // (anb += max(f, 0), f * u_P into the right-hand side as sumf),
      // diffusion at the inlet only (zero-gradient outlet)
      const T fl = -ld(f.mip[fg]), fnb = fmax(fl, zero);
      const T dd = kind == CFDL_BC_INLET ? d : zero;
      bu = bu + fl * ue;
      bv = bv + fl * ve;
      bw = bw + fl * we;
      ap = ap + dd + fnb;
      anbk[k - 6 * e] = anbk[k - 6 * e] + dd + fnb;
      f.anb[k] = static_cast<double>(anbk[k - 6 * e]);
      continue;
    }
    if (kind != CFDL_BC_SYMMETRY) {  // 'dirichlet', :255-265
      T vr[3] = {ue, ve, we};
      const T t = dot3(vr, nrm);
      for (int i = 0; i < 3; i++) vr[i] = vr[i] - t * nrm[i];
      vr[0] = ld(f.u[H]) - vr[0];
      vr[1] = ld(f.v[H]) - vr[1];
      vr[2] = ld(f.w[H]) - vr[2];
      bu = bu + d * vr[0] - d * ue;
      bv = bv + d * vr[1] - d * ve;
      bw = bw + d * vr[2] - d * we;
    }
    ap = ap + d + fz;
    anbk[k - 6 * e] = anbk[k - 6 * e] + d + fz;
    f.anb[k] = static_cast<double>(anbk[k - 6 * e]);
  }
  if (alpha_u != one) {
    const T apr = ap / alpha_u, da = apr - ap;
    bu = bu + da * ue;
    bv = bv + da * ve;
    bw = bw + da * we;
    ap = apr;
  }
  f.ap[e] = static_cast<double>(ap);
  f.bu[e] = static_cast<double>(bu);
  f.bv[e] = static_cast<double>(bv);
  f.bw[e] = static_cast<double>(bw);
  f.d[e] = static_cast<double>(vol / ap);
  T dc = ap;
  for (int q = 0; q < 6; q++) dc = dc - anbk[q];
  f.dc[e] = static_cast<double>(vol / dc);
}
template <typename T>
__global__ void __launch_bounds__(kFaceThreads)
k_coef_uvw6(MeshDev m, FieldsDev f, double dt_, double alpha_u_) {
  d_coef_uvw6<T>(blockIdx.x, m, f, dt_, alpha_u_);
}

// ---------------------------------------------------------------------------
// Energy equation (src/equations/mod_energy.f90; never solved by the
// reference's main loop, whose update_boundaries still applies its BCs).
// ---------------------------------------------------------------------------
// construct_energy, :30-38: t = 273 K, enthalpy h = cp*t on the cells, h0 = h,
// zero gradients (the halos of h are set by the BC setters before any use)
__global__ void k_energy_init(MeshDev m, EnergyDev en) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = m.ne + m.nbf;
  if (i >= n) return;
  en.t[i] = 273.0;
  en.h[i] = i < m.ne ? 273.0 * m.cp[i] : 0.0;
  en.h0[i] = en.h[i];
  for (int q = 0; q < 3; q++) {
    en.gt[3 * i + q] = 0.0;
    en.gh[3 * i + q] = 0.0;
  }
}

// BC setters lid (t = 373) and dirichlet0 (t = 273), :182-212, per halo;
// construct_energy binds 'top' to lid and every other section to
// dirichlet0 (:41-47), i.e. the lid-driven cavity's uvwp binding
__device__ inline void d_bc_energy(int bx_, MeshDev m, EnergyDev en) {
  const int h = bx_ * blockDim.x + threadIdx.x;
  if (h >= m.nbf) return;
  const int H = m.ne + h, e = m.hcell[h];
  en.t[H] = m.hkind[h] == CFDL_BC_LID ? 373.0 : 273.0;
  en.h[H] = m.cp[e] * en.t[H];
}
__global__ void k_bc_energy(MeshDev m, EnergyDev en) {
  d_bc_energy(blockIdx.x, m, en);
}

// calc_coef_energy, :82-169: upwind convection with the face mass fluxes of
// uvwp, diffusion tc/cp with a deferred non-orthogonal correction from the
// gradients of t and h, then the Dirichlet boundary faces of the cell in
// halo order (f = 0, d = tc*A/ds/cp)
__global__ void k_coef_energy(MeshDev m, FieldsDev f, EnergyDev en, double dt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  double rp[3];
  centre(m, e, rp);
  double ap = 0.0, sumf = 0.0, sumdefc = 0.0;
  for (int k = m.row[e]; k < m.row[e + 1]; k++) {
    f.anb[k] = 0.0;
    const int nb = m.nbr[k];
    if (nb >= m.ne) continue;
    const int fs = m.fsg[k];
    const double sg = fs >= 0 ? 1.0 : -1.0;
    const int fg = abs(fs) - 1;
    const double *a = m.aip + 3 * fg, *rip = m.rip + 3 * fg;
    const double area = area_of(a);
    double nrm[3], rpnb[3], dr[3];
    for (int i = 0; i < 3; i++) nrm[i] = sg * a[i] / area;
    centre(m, nb, rpnb);
    for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
    const double wt = vec_weight(rip, rp, rpnb);
    const double fl = -(sg * f.mip[fg]);
    const double fnb = fmax(fl, 0.0);
    sumf = sumf + fl;
    const double tci = (1.0 - wt) * m.tc[e] + wt * m.tc[nb];
    const double cpi = (1.0 - wt) * m.cp[e] + wt * m.cp[nb];
    const double d = tci / cpi / dot3(dr, nrm) * area;
    double ghi[3], gti[3];
    for (int i = 0; i < 3; i++) {
      ghi[i] = (1.0 - wt) * en.gh[3 * e + i] + wt * en.gh[3 * nb + i];
      gti[i] = (1.0 - wt) * en.gt[3 * e + i] + wt * en.gt[3 * nb + i];
    }
    sumdefc = sumdefc + tci * area * (dot3(gti, nrm) - dot3(ghi, nrm) / cpi);
    f.anb[k] = d + fnb;
    ap = ap + d + fnb;
  }
  const double ap0 = m.rho[e] * m.vol[e] / dt;
  ap = ap + ap0;
  f.b[e] = ap0 * en.h0[e] + sumf * en.h[e] + sumdefc;
  for (int q = m.bcrow[e]; q < m.bcrow[e + 1]; q++) {
    const int k = m.bcslot[q], H = m.nbr[k];
    const int fg = m.fsg[k] - 1;  // boundary faces are always owner side
    const double *a = m.aip + 3 * fg;
    const double area = area_of(a);
    const double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
    const double dr[3] = {m.xc[H] - rp[0], m.yc[H] - rp[1], m.zc[H] - rp[2]};
    const double ds = dot3(dr, nrm);
    const double fz = 0.0, d = m.tc[e] * area / ds / m.cp[e];
    ap = ap + d + fz;
    f.anb[k] = f.anb[k] + d + fz;
  }
  f.ap[e] = ap;
}

// calc_temperature, mod_properties.f90:214-222
__global__ void k_temperature(MeshDev m, EnergyDev en) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m.ne) en.t[e] = en.h[e] / m.cp[e];
}

// ---------------------------------------------------------------------------
// calc_grad + matinv3, mod_solver.f90:8-81: weighted least squares over all
// CSR neighbours (boundary halos included), one thread per cell.
// ---------------------------------------------------------------------------
__device__ inline void d_grad(int bx_, MeshDev m, const double *__restrict__ phi, double *__restrict__ grad) {
  const int e = bx_ * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  double a11 = 0.0, a12 = 0.0, a13 = 0.0, a22 = 0.0, a23 = 0.0, a33 = 0.0;
  double g[3] = {0.0, 0.0, 0.0};
  const double xe = m.xc[e], ye = m.yc[e], ze = m.zc[e], pe = phi[e];
  for (int k = m.row[e]; k < m.row[e + 1]; k++) {
    const int nb = m.nbr[k];
    const double dr[3] = {m.xc[nb] - xe, m.yc[nb] - ye, m.zc[nb] - ze};
    const double wt = 1.0 / (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    const double dphi = phi[nb] - pe;
    for (int i = 0; i < 3; i++) g[i] = g[i] + wt * dphi * dr[i];
    a11 = a11 + wt * dr[0] * dr[0];
    a12 = a12 + wt * dr[0] * dr[1];
    a13 = a13 + wt * dr[0] * dr[2];
    a22 = a22 + wt * dr[1] * dr[1];
    a23 = a23 + wt * dr[1] * dr[2];
    a33 = a33 + wt * dr[2] * dr[2];
  }
  const double a21 = a12, a31 = a13, a32 = a23;
  double b11, b12, b13, b21, b22, b23, b31, b32, b33;
  double det = a11 * a22 * a33 - a11 * a23 * a32 - a12 * a21 * a33 + a12 * a23 * a31 +
               a13 * a21 * a32 - a13 * a22 * a31;
  if (fabs(det) > 2.2250738585072014e-308) {
    const double di = 1.0 / det;
    b11 = di * (a22 * a33 - a23 * a32);
    b21 = -(di * (a21 * a33 - a23 * a31));
    b31 = di * (a21 * a32 - a22 * a31);
    b12 = -(di * (a12 * a33 - a13 * a32));
    b22 = di * (a11 * a33 - a13 * a31);
    b32 = -(di * (a11 * a32 - a12 * a31));
    b13 = di * (a12 * a23 - a13 * a22);
    b23 = -(di * (a11 * a23 - a13 * a21));
    b33 = di * (a11 * a22 - a12 * a21);
  } else {
    det = a11 + a22 + a33;
    const double di = 1.0 / det;
    b11 = di; b22 = di; b33 = di;
    b12 = b13 = b21 = b23 = b31 = b32 = 0.0;
  }
  grad[3 * e] = b11 * g[0] + b12 * g[1] + b13 * g[2];
  grad[3 * e + 1] = b21 * g[0] + b22 * g[1] + b23 * g[2];
  grad[3 * e + 2] = b31 * g[0] + b32 * g[1] + b33 * g[2];
}
__global__ void k_grad(MeshDev m, const double *__restrict__ phi, double *__restrict__ grad) {
  d_grad(blockIdx.x, m, phi, grad);
}

// ---------------------------------------------------------------------------
// calc_mip, mod_uvwp.f90:438-490: one thread per face, owner side, interior
// faces only (boundary mip is set by the BC setters).
// ---------------------------------------------------------------------------
__device__ inline void d_mip(int bx_, MeshDev m, FieldsDev f, int rc, double dt) {
  const int fg = bx_ * blockDim.x + threadIdx.x;
  if (fg >= m.nf) return;
  const int k = m.own[fg], e = m.fcell[fg], nb = m.nbr[k];
  if (nb >= m.ne) return;
  double rp[3], rpnb[3];
  centre(m, e, rp);
  centre(m, nb, rpnb);
  const double *a = m.aip + 3 * fg, *rip = m.rip + 3 * fg;
  const double area = area_of(a);
  const double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
  const double wt = vec_weight(rip, rp, rpnb);
  const double v1[3] = {f.u[e], f.v[e], f.w[e]}, v2[3] = {f.u[nb], f.v[nb], f.w[nb]};
  double vip[3];
  for (int i = 0; i < 3; i++) vip[i] = (1.0 - wt) * v1[i] + wt * v2[i];
  const double rhoip = m.rho[e] * (1.0 - wt) + m.rho[nb] * wt;
  double mip = dot3(vip, nrm) * rhoip * area;
  if (rc) {
    double dr[3], gpip[3], vip0[3];
    for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
    for (int i = 0; i < 3; i++) gpip[i] = (1.0 - wt) * f.gp[3 * e + i] + wt * f.gp[3 * nb + i];
    const double dip = (1.0 - wt) * f.d[e] + wt * f.d[nb];
    const double w1[3] = {f.u0[e], f.v0[e], f.w0[e]}, w2[3] = {f.u0[nb], f.v0[nb], f.w0[nb]};
    for (int i = 0; i < 3; i++) vip0[i] = (1.0 - wt) * w1[i] + wt * w2[i];
    mip = mip - rhoip * area * dip / dot3(dr, nrm) * (f.p[nb] - f.p[e] - dot3(gpip, dr)) -
          rhoip / dt * dip * (f.mip0[fg] - dot3(vip0, nrm) * rhoip * area);
  }
  f.mip[fg] = mip;
}
__global__ void k_mip(MeshDev m, FieldsDev f, int rc, double dt) {
  d_mip(blockIdx.x, m, f, rc, dt);
}

// ---------------------------------------------------------------------------
// calc_coef_p, mod_uvwp.f90:289-368: Neumann pressure-correction matrix and
// net inflow, one thread per cell; boundary loop (:344-366) in halo order.
// ---------------------------------------------------------------------------
__global__ void k_coef_p(MeshDev m, FieldsDev f) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  double rp[3];
  centre(m, e, rp);
  double ap = 0.0, sumf = 0.0;
  for (int k = m.row[e]; k < m.row[e + 1]; k++) {
    const int nb = m.nbr[k];
    double d = 0.0;
    if (nb < m.ne) {
      const int fs = m.fsg[k];
      const double sg = fs >= 0 ? 1.0 : -1.0;
      const int fg = abs(fs) - 1;
      const double *a = m.aip + 3 * fg, *rip = m.rip + 3 * fg;
      const double area = area_of(a);
      double nrm[3], rpnb[3], dr[3];
      for (int i = 0; i < 3; i++) nrm[i] = sg * a[i] / area;
      centre(m, nb, rpnb);
      for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
      const double wt = vec_weight(rip, rp, rpnb);
      const double fl = -(sg * f.mip[fg]);
      sumf = sumf + fl;
      const double rhoip = (1.0 - wt) * m.rho[e] + wt * m.rho[nb];
      d = ((1.0 - wt) * f.dc[e] + wt * f.dc[nb]) / dot3(dr, nrm) * rhoip * area;
    }
    f.anb[k] = d;
    ap = ap + d;
  }
  double b = sumf;
  for (int q = m.bcrow[e]; q < m.bcrow[e + 1]; q++) {
    const int k = m.bcslot[q];
    const double d = 0.0;  // 'dirichlet' and 'zero_flux' both add nothing
    ap = ap + d;
    f.anb[k] = f.anb[k] + d;
    b = b - f.mip[m.fsg[k] - 1];
  }
  f.ap[e] = ap;
  f.b[e] = b;
}

// calc_coef_p for meshes whose rows all have six entries, six threads per
// cell as in k_coef_uvw6: entry terms in parallel, then the cell's sums in
// CSR order.  T = float (the FP32 assembly) computes the coefficients in
// FP32; the right-hand side, the cell's net inflow (the continuity check),
// stays FP64 in both.
template <typename T>
__device__ inline void d_coef_p6(int bx_, MeshDev m, FieldsDev f) {
  const auto G = geo_of<T>(m);
  const T one = static_cast<T>(1), zero = static_cast<T>(0);
  __shared__ T sd[kFaceThreads];
  __shared__ double sfl[kFaceThreads];
  __shared__ int sin_[kFaceThreads];
  const int lc = threadIdx.x / 6, j = threadIdx.x - 6 * lc;
  const int e = bx_ * kFaceCells + lc;
  const bool valid = lc < kFaceCells && e < m.ne;
  if (valid) {
    const int k = 6 * e + j;
    const int nb = m.nbr[k];
    T d = zero;
    double fl = 0.0;
    if (nb < m.ne) {
      const T rp[3] = {G.xc[e], G.yc[e], G.zc[e]};
      const int fs = m.fsg[k];
      const T sg = fs >= 0 ? one : -one;
      const int fg = abs(fs) - 1;
      const T *a = G.aip + 3 * fg, *rip = G.rip + 3 * fg;
      const T area = area_of(a);
      T nrm[3], dr[3];
      for (int i = 0; i < 3; i++) nrm[i] = sg * a[i] / area;
      const T rpnb[3] = {G.xc[nb], G.yc[nb], G.zc[nb]};
      for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
      const T wt = vec_weight(rip, rp, rpnb);
      fl = -((fs >= 0 ? 1.0 : -1.0) * f.mip[fg]);
      const T rhoip = (one - wt) * G.rho[e] + wt * G.rho[nb];
      d = ((one - wt) * static_cast<T>(f.dc[e]) + wt * static_cast<T>(f.dc[nb])) / dot3(dr, nrm) *
          rhoip * area;
      f.anb[k] = static_cast<double>(d);
    }
    sd[threadIdx.x] = d;
    sfl[threadIdx.x] = fl;
    sin_[threadIdx.x] = nb < m.ne;
  }
  __syncthreads();
  if (!valid || j != 0) return;
  T ap = zero;
  double sumf = 0.0;
  for (int q = 0; q < 6; q++) {
    const int t = threadIdx.x + q;
    if (sin_[t]) {
      sumf = sumf + sfl[t];
    } else {
      f.anb[6 * e + q] = static_cast<double>(sd[t]);
    }
    ap = ap + sd[t];
  }
  double b = sumf;
  for (int q = m.bcrow[e]; q < m.bcrow[e + 1]; q++) {
    const int k = m.bcslot[q];
    const T d = zero;  // 'dirichlet' and 'zero_flux' both add nothing
    ap = ap + d;
    f.anb[k] = static_cast<double>(sd[threadIdx.x + (k - 6 * e)] + d);
    b = b - f.mip[m.fsg[k] - 1];
  }
  f.ap[e] = static_cast<double>(ap);
  f.b[e] = b;
}
template <typename T>
__global__ void __launch_bounds__(kFaceThreads) k_coef_p6(MeshDev m, FieldsDev f) {
  d_coef_p6<T>(blockIdx.x, m, f);
}

// update_uvwp cell part, mod_uvwp.f90:383-392 (no velocity correction)
// (config mode scales the correction by alpha_p; alpha_p = 1 is exact)
__device__ inline void d_update_cells(int bx_, int ne, FieldsDev f, double alpha_p) {
  const int e = bx_ * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  f.p[e] = f.p[e] + alpha_p * f.phic[e];
  for (int i = 0; i < 3; i++) f.gp[3 * e + i] = f.gp[3 * e + i] + alpha_p * f.gpc[3 * e + i];
}
__global__ void k_update_cells(int ne, FieldsDev f, double alpha_p) {
  d_update_cells(blockIdx.x, ne, f, alpha_p);
}

// update_uvwp face part, mod_uvwp.f90:394-415 (arithmetic-mean rho)
__device__ inline void d_update_faces(int bx_, MeshDev m, FieldsDev f) {
  const int fg = bx_ * blockDim.x + threadIdx.x;
  if (fg >= m.nf) return;
  const int k = m.own[fg], e = m.fcell[fg], nb = m.nbr[k];
  if (nb >= m.ne) return;
  const double *a = m.aip + 3 * fg, *rip = m.rip + 3 * fg;
  const double area = area_of(a);
  const double nrm[3] = {a[0] / area, a[1] / area, a[2] / area};
  double rp[3], rpnb[3], dr[3];
  centre(m, e, rp);
  centre(m, nb, rpnb);
  for (int i = 0; i < 3; i++) dr[i] = rpnb[i] - rp[i];
  const double wt = vec_weight(rip, rp, rpnb);
  const double dip = (1.0 - wt) * f.dc[e] + wt * f.dc[nb];
  const double rhoip = (m.rho[e] + m.rho[nb]) / 2.0;
  const double dmip = rhoip * area * dip * (f.phic[nb] - f.phic[e]) / dot3(dr, nrm);
  f.mip[fg] = f.mip[fg] - dmip;
}
__global__ void k_update_faces(MeshDev m, FieldsDev f) {
  d_update_faces(blockIdx.x, m, f);
}

// adjust_pc, mod_uvwp.f90:136-158 with pref = phic(1) (:129)
__device__ inline void d_pref(int bx_, const double *phic, double *scal) {
  if (bx_ == 0 && threadIdx.x == 0) scal[0] = phic[0];
}
__global__ void k_pref(const double *phic, double *scal) {
  d_pref(blockIdx.x, phic, scal);
}
__device__ inline void d_sub_pref(int bx_, int ne, double *phic, const double *scal) {
  const int e = bx_ * blockDim.x + threadIdx.x;
  if (e < ne) phic[e] = phic[e] - scal[0];
}
__global__ void k_sub_pref(int ne, double *phic, const double *scal) {
  d_sub_pref(blockIdx.x, ne, phic, scal);
}
__device__ inline void d_halo_from_owner(int bx_, MeshDev m, double *phi) {
  const int h = bx_ * blockDim.x + threadIdx.x;
  if (h < m.nbf) phi[m.ne + h] = phi[m.hcell[h]];
}
__global__ void k_halo_from_owner(MeshDev m, double *phi) {
  d_halo_from_owner(blockIdx.x, m, phi);
}

__device__ inline void d_fill(int bx_, double *p, int64_t n, double v) {
  const int64_t i = static_cast<int64_t>(bx_) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill(double *p, int64_t n, double v) {
  d_fill(blockIdx.x, p, n, v);
}
__global__ void k_to_f32(const double *src, float *dst, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = static_cast<float>(src[i]);
}
__device__ inline void d_copy(int bx_, double *dst, const double *src, int64_t n) {
  const int64_t i = static_cast<int64_t>(bx_) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
__global__ void k_copy(double *dst, const double *src, int64_t n) {
  d_copy(blockIdx.x, dst, src, n);
}

// Net outward mass flux of each cell, sum over its CSR entries of
// mip(face) * sign in row order: the quantity the reference's VTU output
// stores in dc and writes as 'mip' (src/modules/mod_vtu_output.f90:113-123).
__global__ void k_mass_imbalance(MeshDev m, const double *__restrict__ mip, double *out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  double s = 0.0;
  for (int k = m.row[e]; k < m.row[e + 1]; k++) {
    const int f = m.fsg[k];
    s = s + mip[(f > 0 ? f : -f) - 1] * (f > 0 ? 1.0 : -1.0);
  }
  out[e] = s;
}

// ---------------------------------------------------------------------------
// Gauss-Seidel: one dependency level of a sweep, up to NC right-hand sides
// sharing ap/anb (smoother_gs / solve_gs inner loop, mod_solver.f90:205-224,
// 287-307).  phi(e) = (sumnb + (sor-1)*ap*phi(e))/ap/sor.
// ---------------------------------------------------------------------------
// cells == nullptr: the cells are base .. base + n - 1 (a colour range).
template <int NC>
__device__ inline void d_gs_level(int bx_, SysDesc s, const int *__restrict__ cells, int base, int n, double sor,
                           const int *__restrict__ active) {
  const int t = bx_ * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int e = cells ? cells[t] : base + t;
  const int r0 = s.row[e], r1 = s.row[e + 1];
  const double a = s.ap[e];
  if (r1 - r0 == 6) {
    // hexahedral row: every index, coefficient and value is loaded before
    // the row sum, which then runs in CSR order as below
    int nb[6];
    double an[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
      nb[k] = s.nbr[r0 + k];
      an[k] = s.anb[r0 + k];
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
      if (!active[c]) continue;
      double *phi = s.phi[c];
      double v[6];
#pragma unroll
      for (int k = 0; k < 6; k++) v[k] = phi[nb[k]];
      double sum = s.b[c][e];
#pragma unroll
      for (int k = 0; k < 6; k++) sum = sum + an[k] * v[k];
      phi[e] = (sum + (sor - 1.0) * a * phi[e]) / a / sor;
    }
    return;
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    if (!active[c]) continue;
    double *phi = s.phi[c];
    double sum = s.b[c][e];
    for (int k = r0; k < r1; k++) sum = sum + s.anb[k] * phi[s.nbr[k]];
    phi[e] = (sum + (sor - 1.0) * a * phi[e]) / a / sor;
  }
}
template <int NC>
__global__ void k_gs_level(SysDesc s, const int *__restrict__ cells, int base, int n, double sor,
                           const int *__restrict__ active) {
  d_gs_level<NC>(blockIdx.x, s, cells, base, n, sor, active);
}

// The same update in FP32 on a correction system: cells of one colour (all
// independent), row sum in CSR order, then (sum + (sor-1)*ap*phi)/ap/sor with
// IEEE single-precision divisions.
// One colour is a contiguous cell range in each subdomain of the colour
// numbering: cells beg[y] .. beg[y] + off[y+1] - off[y] - 1 are launch
// positions off[y] .. off[y+1] - 1.  HEX: every row has six entries, so row
// e starts at 6e.
struct ColRanges {
  int n;
  int beg[kMaxSys], off[kMaxSys + 1];
};
template <bool HEX>
__global__ void k_gs_level32(SysDesc32 s, ColRanges cr, float sor,
                             const int *__restrict__ active) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cr.off[cr.n] || !active[0]) return;
  int y = 0;
  while (t >= cr.off[y + 1]) y++;
  const int e = cr.beg[y] + (t - cr.off[y]);
  const int r0 = HEX ? 6 * e : s.row[e], r1 = HEX ? r0 + 6 : s.row[e + 1];
  const float a = s.ap[e];
  float sum = s.b[e];
  if (r1 - r0 == 6) {
    int nb[6];
    float an[6], v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
      nb[k] = s.nbr[r0 + k];
      an[k] = s.anb[r0 + k];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = s.phi[nb[k]];
#pragma unroll
    for (int k = 0; k < 6; k++) sum = sum + an[k] * v[k];
  } else {
    for (int k = r0; k < r1; k++) sum = sum + s.anb[k] * s.phi[s.nbr[k]];
  }
  s.phi[e] = (sum + (sor - 1.0f) * a * s.phi[e]) / a / sor;
}

// Mixed-precision pc solve: FP32 copies of ap and anb, and a zero correction
// (cells and halos).  Then, at every residual check, the correction is added
// to the FP64 values and cleared (once the solve has stopped there is
// nothing to add, so the check-time update is skipped).
__global__ void k_mixed_prep(int ne, int L, int nall, const double *__restrict__ ap,
                             const double *__restrict__ anb, float *__restrict__ a32,
                             float *__restrict__ anb32, float *__restrict__ d32) {
  const int stride = gridDim.x * blockDim.x, n = max(L, max(ne, nall));
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (i < L) anb32[i] = static_cast<float>(anb[i]);
    if (i < ne) a32[i] = static_cast<float>(ap[i]);
    if (i < nall) d32[i] = 0.0f;
  }
}
__global__ void k_mixed_update(int nall, double *__restrict__ phic, float *__restrict__ d32,
                               const int *__restrict__ active) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nall || (active && !active[0])) return;
  phic[i] = phic[i] + static_cast<double>(d32[i]);
  d32[i] = 0.0f;
}

// Space-time tiled sweeps of the pressure-correction system.  The host
// numbers that system so that every (launch, tile) batch is one contiguous
// cell range whose dependency-level segments are contiguous sub-ranges:
// seg[bseg[b] .. bseg[b+1]) are the nq+1 segment bounds of batch b.  A
// dependency between tiles always spans launches, so a workgroup only waits
// for its own earlier segments (a barrier between segments).  The forward
// sweep walks the launches and each batch's segments in order; the
// backward sweep walks both in reverse, which visits every cell after all
// of its backward predecessors (its neighbours on higher forward levels).
constexpr int kTileThreads = 64;
__global__ void __launch_bounds__(kTileThreads)
k_gs_tiles(SysDesc s, const int *__restrict__ bseg, const int *__restrict__ seg, int batch0,
           int backward, double sor, const int *__restrict__ active) {
  if (!active[0]) return;
  const int b = batch0 + blockIdx.x;
  const int s0 = bseg[b], nq = bseg[b + 1] - s0 - 1;
  double *phi = s.phi[0];
  const double *rhs = s.b[0];
  for (int j = 0; j < nq; j++) {
    const int q = backward ? nq - 1 - j : j;
    const int c1 = seg[s0 + q + 1];
    for (int e = seg[s0 + q] + threadIdx.x; e < c1; e += kTileThreads) {
      const int r0 = s.row[e], r1 = s.row[e + 1];
      const double a = s.ap[e];
      double sum = rhs[e];
      for (int k = r0; k < r1; k++) sum = sum + s.anb[k] * phi[s.nbr[k]];
      phi[e] = (sum + (sor - 1.0) * a * phi[e]) / a / sor;
    }
    __syncthreads();
  }
}

// The same sweep with the batch resident in LDS.  NW wavefronts load the
// batch's values, coefficients and local slot table (contiguous ranges
// starting at its first cell c0 and first CSR entry r0) plus the values of
// its outside neighbours (xslot); the segments are walked on LDS only (by
// wavefront 0 alone in the pipelined walk) and the NW wavefronts write the
// batch back.  Outside neighbours are constant during the launch: a cell of
// another tile is never updated in the same launch.  lnbr gives, per CSR
// entry, the LDS slot of the neighbour: < nc a batch cell, else outside
// value nc + j; for the pipelined walk the host stores the LDS byte address
// of that slot's values instead (8*NC*slot).  Every load of a
// gather pass is issued before the first LDS
// store, so the gather costs two memory latencies (the outside indices, then
// their values) plus the issue of its loads, which NW wavefronts share.  The
// segment count is at most 63 (host).
constexpr int kLdsThreads = 64;  // one wavefront: the pipelined walk's width
constexpr int kXcds = 8;         // XCDs of an MI355X; each has its own L2
// The pipelined walk's values start at LDS byte address 0: the kernel has
// no static LDS, so its dynamic block starts there.

// Orders the LDS accesses of the wavefront that walks an LDS sweep alone: a
// wavefront's LDS instructions execute in issue order, so what one lane
// stores is seen by every later load of any lane without s_barrier or a
// wait for the store.  The wavefront-scope fences keep the compiler from
// moving LDS accesses across this point.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Brings batch b's read-only data and its own cells' values into the L2 of
// the XCD that runs this workgroup: one byte of every 128-byte line of its
// descriptor-given ranges (segment bounds, ap, 1/ap, right-hand sides,
// values, anb, slot table, outside-neighbour indices), nt threads, at most
// kTouch lines per thread.  The loaded bytes are discarded.  A batch's own
// cells are written only by the batch itself, so their values stay what the
// next launch reads.
constexpr int kTouch = 4;

// The arrays a tiled sweep reads and writes, as a compact kernel argument:
// the argument block is a cost of each of the sweep's thousands of launches.
// The kernel's first three arguments (batch0, bdesc, active: all its first
// loads need) are scalars and pointers, which the build preloads into
// scalar registers at wave launch (-amdgpu-kernarg-preload-count=3).
template <int NC>
struct TileSys {
  double *phi[NC];
  const double *b[NC];
  const double *ap, *anb, *rap;
  const int *row;
};

template <int NC>
__device__ void touch_batch(const TileSys<NC> &s, const int4 *__restrict__ bdesc,
                            const int *__restrict__ seg, const unsigned short *__restrict__ lnbr,
                            const int *__restrict__ xslot, int b, int t, int nt) {
  const int4 d0 = bdesc[2 * b], d1 = bdesc[2 * b + 1];
  constexpr int NR = 5 + 2 * NC;
  const unsigned char *base[NR];
  int len[NR];  // bytes
  base[0] = reinterpret_cast<const unsigned char *>(seg + d0.x);
  len[0] = 4 * (d0.y + 1);
  base[1] = reinterpret_cast<const unsigned char *>(s.ap + d0.z);
  base[2] = reinterpret_cast<const unsigned char *>(s.rap + d0.z);
  len[1] = len[2] = 8 * d0.w;
  base[3] = reinterpret_cast<const unsigned char *>(s.anb + d1.x);
  len[3] = 8 * d1.y;
  base[4] = reinterpret_cast<const unsigned char *>(lnbr + d1.x);
  len[4] = 2 * d1.y;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    base[5 + 2 * c] = reinterpret_cast<const unsigned char *>(s.b[c] + d0.z);
    base[6 + 2 * c] = reinterpret_cast<const unsigned char *>(s.phi[c] + d0.z);
    len[5 + 2 * c] = len[6 + 2 * c] = 8 * d0.w;
  }
  // the outside-neighbour indices come last
  const unsigned char *xb = reinterpret_cast<const unsigned char *>(xslot + d1.z);
  const int xl = 4 * d1.w;
  unsigned v[kTouch];
#pragma unroll
  for (int k = 0; k < kTouch; k++) {
    int line = t + k * nt;
    const unsigned char *p = nullptr;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int nl = (len[r] + 127) >> 7;
      if (!p && line < nl) p = base[r] + min(line << 7, len[r] - 1);
      if (!p) line -= nl;
    }
    if (!p && line < ((xl + 127) >> 7)) p = xb + min(line << 7, xl - 1);
    v[k] = p ? *p : 0;
  }
  // the loads complete here; their bytes are not used
#pragma unroll
  for (int k = 0; k < kTouch; k++) asm volatile("" ::"v"(v[k]));
}

template <bool HEX, bool PIPE, int NC, int NW>
__global__ void __launch_bounds__(kLdsThreads * NW)
k_gs_tiles_lds(int batch0, const int4 *__restrict__ bdesc, const int *__restrict__ active,
               TileSys<NC> s, const int *__restrict__ seg,
               const unsigned short *__restrict__ lnbr, const int *__restrict__ xslot,
               int backward, double sor, double rsor, int next0, int nb_next, int rstride) {
  extern __shared__ double lds[];
  constexpr int T = kLdsThreads * NW;  // threads of the gather and write-back
  // the batch's descriptor (segment, cell, CSR and outside-neighbour ranges)
  // comes in one load, issued with the active flags
  const int b = batch0 + blockIdx.x;
  const int4 d0 = bdesc[2 * b], d1 = bdesc[2 * b + 1];
  __builtin_amdgcn_sched_barrier(0);  // the descriptor load goes out first
  // NC right-hand sides share ap/anb (u, v, w of a momentum solve); a
  // right-hand side whose solve has stopped is neither updated nor written
  bool act[NC];
  bool any = false;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    act[c] = active[c] != 0;
    any = any || act[c];
  }
  // The descriptor's second half and the kernel arguments of the gather
  // arrive with the first half and the flags: without these uses the
  // compiler loads them after the branch below, one more scalar round trip
  // before the gather's first load.
  asm volatile("" ::"s"(d1.x), "s"(d1.y), "s"(d1.z), "s"(d1.w), "s"(s.ap), "s"(s.rap),
               "s"(s.anb), "s"(lnbr), "s"(seg), "s"(xslot));
#pragma unroll
  for (int c = 0; c < NC; c++) asm volatile("" ::"s"(s.phi[c]), "s"(s.b[c]));
  // (no batch is empty: testing nc here makes both loads precede the branch)
  if (!any | (d0.w == 0)) return;
  const int s0 = d0.x, nq = d0.y, c0 = d0.z, nc = d0.w;
  const int r0 = d1.x, nent = d1.y, x0 = d1.z, nx = d1.w;
  const int np = nc + nx;
  // LDS layout of the plain walk: component-major values, then per-array
  // coefficient ranges
  double *lphi = lds;               // NC x (nc + nx) values
  double *lap = lphi + NC * np;     // nc
  double *lra = lap + nc;           // nc reciprocals 1/ap
  double *lb = lra + nc;            // NC x nc
  double *lanb = lb + NC * nc;      // nent
  int *lseg = reinterpret_cast<int *>(lanb + nent);  // nq + 1 bounds relative to c0
  int *lrow = lseg + nq + 1;                          // nc + 1 row starts (general rows)
  unsigned short *lnb = reinterpret_cast<unsigned short *>(lrow + (HEX ? 0 : nc + 1));
  // LDS layout of the pipelined walk: values interleaved by component (slot
  // s, component c at pphi[NC*s + c]); per cell (ap, 1/ap) as one 16-byte
  // record, b, its six anb, and the LDS byte addresses of its six
  // neighbours' values as 16-bit numbers, which the host computes (pphi is
  // at address 0 and the host keeps 8*NC*np within 16 bits for this walk),
  // in a 16-byte record (the fourth word unused); then 32 bytes that take
  // the stores of lanes without a cell.  Records are 16-byte aligned, so
  // each is one aligned read.
  typedef unsigned U4 __attribute__((ext_vector_type(4)));
  typedef double D2 __attribute__((ext_vector_type(2)));
  double *pphi = lds;
  const unsigned o_ar = (8u * NC * np + 15u) & ~15u;  // byte offset of (ap, 1/ap)
  D2 *par = reinterpret_cast<D2 *>(reinterpret_cast<char *>(lds) + o_ar);
  double *pb = reinterpret_cast<double *>(par + nc);  // NC x nc
  double *panb = pb + NC * nc;                          // 6 per cell
  U4 *psl = reinterpret_cast<U4 *>(reinterpret_cast<char *>(lds) +
                                   ((o_ar + 8u * (2 + NC + 6) * nc + 15u) & ~15u));
  const double *pa = s.ap + c0, *pr = s.rap + c0, *pn = s.anb + r0;
  const unsigned short *pl = lnbr + r0;
  // segment bound of lane q of wavefront 0 (lanes past nq reload the last
  // one); used after the gather, so its load overlaps the gather's
  const int sg_raw = seg[s0 + min(static_cast<int>(threadIdx.x), nq)];
  // gather: each pass covers U*T cells (values, coefficients, slots) and
  // X*T outside neighbours; every load of a pass is issued before its LDS
  // stores, so a pass costs two memory latencies (the outside indices, then
  // their values).  With four wavefronts one pass takes 768 cells and 768
  // outside neighbours, more than the 16 x 8 x 4 tiles' batches have.
  // Loads under a branch make the compiler's wait for a later use count
  // conservatively, so: the index and outside-value loads are unconditional
  // (lanes past the end reload an entry in range); the rows of the pass's
  // later thirds (when they exist) go out before the indices and the first
  // third's rows after them, so the wait for the indices leaves exactly the
  // first third's loads in flight (loads return in issue order).
  constexpr int U = NW == 1 ? (NC == 1 ? 8 : 4) : 3;
  constexpr int X = NW == 1 ? 8 : 3;
  const int nrc = (nc + U * T - 1) / (U * T);
  const int nrx = (nx + X * T - 1) / (X * T);
  const int nr = nrc > nrx ? nrc : nrx;
  auto gather = [&](int r) {
    const int cb = r * U * T, xb = r * X * T;
    int xs[X];
    double va[U], vr[U], vn[U][6], vp[NC][U], vb[NC][U], vx[NC][X];
    unsigned sl[U][3];  // six 16-bit slots or LDS addresses, two per word
    auto load_rows = [&](int u) {
      const int i = cb + u * T + threadIdx.x;
      const int ic = i < nc ? i : 0;  // lanes past the end reload row 0
      va[u] = pa[ic];
      vr[u] = pr[ic];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        vp[c][u] = s.phi[c][c0 + ic];
        vb[c][u] = s.b[c][c0 + ic];
      }
      if (HEX) {
#pragma unroll
        for (int k = 0; k < 6; k++) {
          vn[u][k] = pn[6 * ic + k];
          if (k < 3)  // pl is 4-byte aligned: a hexahedral batch's rows start at 6 * c0
            sl[u][k] = reinterpret_cast<const unsigned *>(pl)[3 * ic + k];
        }
      }
      // keeps the compiler from moving the loads into the stores' branch,
      // after the outside-value loads and their wait
      asm volatile("" ::: "memory");
    };
#pragma unroll
    for (int u = U - 1; u > 0; u--)
      if (cb + u * T < nc) load_rows(u);  // workgroup-uniform: rows of this pass exist
#pragma unroll
    for (int u = 0; u < X; u++) {
      const int j = min(xb + u * T + static_cast<int>(threadIdx.x), nx - 1);
      xs[u] = xslot[max(x0 + j, 0)];
    }
    __builtin_amdgcn_sched_barrier(0);
    if (r == 0 || cb < nc) load_rows(0);  // the first pass always has rows
    __builtin_amdgcn_sched_barrier(0);   // the indices' first use comes after all row loads
#pragma unroll
    for (int u = 0; u < X; u++)
#pragma unroll
      for (int c = 0; c < NC; c++) vx[c][u] = s.phi[c][static_cast<unsigned>(xs[u])];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = cb + u * T + threadIdx.x;
      if (i < nc) {
        if constexpr (PIPE) {
          par[i] = D2{va[u], vr[u]};
#pragma unroll
          for (int c = 0; c < NC; c++) {
            pphi[NC * i + c] = vp[c][u];
            pb[c * nc + i] = vb[c][u];
          }
#pragma unroll
          for (int k = 0; k < 6; k++) {
            panb[6 * i + k] = vn[u][k];
            if (k == 0) psl[i] = U4{sl[u][0], sl[u][1], sl[u][2], 0u};  // LDS addresses (host)
          }
        } else {
          lap[i] = va[u];
          lra[i] = vr[u];
#pragma unroll
          for (int c = 0; c < NC; c++) {
            lphi[c * np + i] = vp[c][u];
            lb[c * nc + i] = vb[c][u];
          }
          if (HEX) {
#pragma unroll
            for (int k = 0; k < 6; k++) {
              lanb[6 * i + k] = vn[u][k];
              lnb[6 * i + k] = static_cast<unsigned short>(sl[u][k >> 1] >> (k & 1 ? 16 : 0));
            }
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < X; u++) {
      const int j = xb + u * T + threadIdx.x;
      if (j < nx)
#pragma unroll
        for (int c = 0; c < NC; c++) {
          if constexpr (PIPE)
            pphi[NC * (nc + j) + c] = vx[c][u];
          else
            lphi[c * np + nc + j] = vx[c][u];
        }
    }
  };
  // The first pass, which is the only one for batches of up to U*T cells and
  // X*T outside neighbours, is outside the loop: in the loop the compiler
  // waits for the previous pass's loads before reusing their registers, which
  // serialises the index loads and the coefficient loads of the first pass.
  gather(0);
  for (int r = 1; r < nr; r++) gather(r);
  if (!HEX) {
    for (int k = threadIdx.x; k < nent; k += T) {
      lanb[k] = pn[k];
      lnb[k] = pl[k];
    }
    for (int i = threadIdx.x; i <= nc; i += T) lrow[i] = s.row[c0 + i] - r0;
  }
  const int sg = sg_raw - c0;
  if (!PIPE && threadIdx.x <= nq) lseg[threadIdx.x] = sg;
  __syncthreads();
  if constexpr (PIPE) {
    // wavefront 0 walks; the others wait at the barrier before the write-back
    if (NW == 1 || threadIdx.x < kLdsThreads) {
      // Every segment has at most 64 cells: lane l updates cell sg[q] + l of
      // segment q (bounds from lane q of sg).  Two operand sets alternate: while
      // segment j computes from one, the other receives segment j+1's ap, 1/ap,
      // b, anb, neighbour addresses and own old value (its cells are not
      // written at level j); right after segment j's stores it receives its
      // neighbour values.  The walk is one wavefront's, whose LDS accesses
      // execute in issue order, so those reads see the stores without waiting
      // for them.  Lanes without a cell read a later cell's operands and store
      // to the 32 bytes after the table.
      // RCP: the divisions by ap and sor use div_rcp2.  Without a branch in the
      // walk, each lane keeps the least and greatest frexp exponent of its
      // dividends x and a NaN detector (0*nv is NaN for an infinite or NaN
      // result, which an out-of-range 1/ap or an overflow produces); if any
      // lane saw a dividend outside [-kRcpExp, kRcpExp] or a non-finite
      // result, the batch is walked again with hardware division from its
      // original values, which global memory still holds.
      const int lane = threadIdx.x;
      const double sm1 = sor - 1.0;
      // The next segment's neighbour addresses are read first and unpacked
      // before this level's arithmetic (they are next in the LDS queue after
      // this level's values); its other operands are used only after this
      // level's store, so the own values stay raw (p).
      struct Ops {
        int st, en;   // the segment's cells [st, en) (uniform)
        unsigned pv;  // LDS byte address of the lane's cell's values (a later cell without a cell)
        bool on;
        double a, ya, bv[NC], p[NC], an[6], v[NC][6];
        unsigned sl[6];
      };
      // LDS byte addresses of the walk's arrays (the dynamic block starts at
      // 0), so that an operand's address is one add or 24-bit multiply-add
      typedef const __attribute__((address_space(3))) double *LdsD;
      const unsigned b_ar = o_ar, b_b = b_ar + 16u * nc, b_an = b_b + 8u * NC * nc,
                     b_sl = (b_an + 48u * nc + 15u) & ~15u, b_dum = b_sl + 16u * nc;
      auto lds_f64 = [](unsigned a) { return *reinterpret_cast<LdsD>(static_cast<size_t>(a)); };
      // operands and own values of the segment of cells [st, en)
      auto load_ops = [&](int st, int en, Ops &o) {
        o.st = st;
        o.en = en;
        const int ii = st + lane;
        o.on = ii < en;
        // lanes past the segment read a later cell of the batch (its last
        // past the end): real operands, so the division guard stays valid
        const unsigned ic = min(ii, nc - 1);
        // 12 of the record's 16 bytes (a fourth, dead register would be
        // reused at once and make the compiler wait for this read)
        typedef unsigned U3 __attribute__((ext_vector_type(3)));
        const U3 w = *reinterpret_cast<const __attribute__((address_space(3))) U3 *>(
            static_cast<size_t>(b_sl + 16u * ic));
        __builtin_amdgcn_sched_barrier(0);  // first in the LDS queue after the values
#pragma unroll
        for (int k = 0; k < 6; k++) o.sl[k] = (k & 1 ? w[k >> 1] >> 16 : w[k >> 1] & 0xffffu);
        const unsigned c8 = 8u * ic;
        o.pv = NC == 1 ? c8 : __umul24(ic, 8u * NC);
        const D2 ar = *reinterpret_cast<const __attribute__((address_space(3))) D2 *>(
            static_cast<size_t>(b_ar + 2u * c8));
        o.a = ar.x;
        o.ya = ar.y;
#pragma unroll
        for (int c = 0; c < NC; c++) {
          o.bv[c] = lds_f64(b_b + 8u * c * nc + c8);
          o.p[c] = lds_f64(o.pv + 8u * c);
        }
        const unsigned an = __umul24(ic, 48u) + b_an;
#pragma unroll
        for (int k = 0; k < 6; k++) o.an[k] = lds_f64(an + 8u * k);
      };
      // neighbour values at the LDS byte addresses the host stored
      auto load_v = [&](Ops &o) {
#pragma unroll
        for (int k = 0; k < 6; k++)
#pragma unroll
          for (int c = 0; c < NC; c++)
            o.v[c][k] = *reinterpret_cast<const __attribute__((address_space(3))) double *>(
                static_cast<size_t>(o.sl[k] + 8 * c));
      };
      // The visit direction is a compile-time argument (bwd is a
      // std::integral_constant): the next segment shares one bound with the
      // current one, so a level reads one segment bound (lane q of sg).
      auto walk = [&](bool rcp, auto bwd) __attribute__((always_inline)) -> bool {
        constexpr bool BWD = decltype(bwd)::value;
        int emin = 0, emax = 0;
        double nan_seen = 0.0;
        // level j of the walk: the j-th segment, then (more) the operands of
        // the next
        auto step = [&](int j, Ops &cur, Ops &nxt, bool more) {
          if (more) {
            if (BWD)
              load_ops(__builtin_amdgcn_readlane(sg, nq - 2 - j), cur.st, nxt);
            else
              load_ops(cur.en, __builtin_amdgcn_readlane(sg, j + 2), nxt);
          }
          // the operand loads go out before this level's arithmetic
          __builtin_amdgcn_sched_barrier(0);
          double xg[NC], ng[NC];
          // with several right-hand sides every one is computed, and one whose
          // solve has stopped (not act) stores past the table, like a lane
          // without a cell (fewer instructions than a select per use)
#pragma unroll
          for (int c = 0; c < NC; c++) {
            if (NC == 1 && !act[c]) continue;
            double sum = cur.bv[c];
#pragma unroll
            for (int k = 0; k < 6; k++) sum = sum + cur.an[k] * cur.v[c][k];
            const double x = sum + sm1 * cur.a * cur.p[c];
            double nv;
            if (rcp)
              nv = div_rcp2(x, cur.a, cur.ya, sor, rsor);
            else
              nv = x / cur.a / sor;
            xg[c] = x;
            ng[c] = nv;
            // unconditional store, so no branch separates it from the next
            // segment's loads; lanes without a cell store past the table
            *reinterpret_cast<__attribute__((address_space(3))) double *>(
                static_cast<size_t>(((NC == 1 || act[c]) && cur.on ? cur.pv : b_dum) + 8u * c)) = nv;
          }
          wave_sync();  // the walk is one wavefront (kLdsThreads = 64)
          if (more) load_v(nxt);
          // the guard's bookkeeping waits with the next level's values, off
          // the chain from this level's dividend to its store
          __builtin_amdgcn_sched_barrier(0);
          if (rcp) {
#pragma unroll
            for (int c = 0; c < NC; c++) {
              if (NC == 1 && !act[c]) continue;
              const int ex = __builtin_amdgcn_frexp_exp(xg[c]);  // 0 for x = 0
              emin = min(emin, ex);
              emax = max(emax, ex);
              nan_seen = fma(0.0, ng[c], nan_seen);
            }
          }
        };
        Ops A, B;
        if (BWD)
          load_ops(__builtin_amdgcn_readlane(sg, nq - 1), __builtin_amdgcn_readlane(sg, nq), A);
        else
          load_ops(__builtin_amdgcn_readlane(sg, 0), __builtin_amdgcn_readlane(sg, 1), A);
        load_v(A);
        // the loop's levels all have a next segment; the last one or two
        // levels follow it
        int j = 0;
        for (; j + 2 < nq; j += 2) {
          step(j, A, B, true);
          step(j + 1, B, A, true);
        }
        if (j + 1 < nq) {
          step(j, A, B, true);
          step(j + 1, B, A, false);
        } else {
          step(j, A, B, false);
        }
        return emin < -kRcpExp || emax > kRcpExp || nan_seen != nan_seen;
      };
      auto walk_dir = [&](bool rcp) __attribute__((always_inline)) -> bool {
        return backward ? walk(rcp, std::true_type{}) : walk(rcp, std::false_type{});
      };
      if (rsor > 0.0) {
        if (__builtin_amdgcn_ballot_w64(walk_dir(true))) {
          for (int r = threadIdx.x; r < nc; r += kLdsThreads)
#pragma unroll
            for (int c = 0; c < NC; c++) pphi[NC * r + c] = s.phi[c][c0 + r];
          wave_sync();
          walk_dir(false);
        }
      } else {
        walk_dir(false);
      }
    } else {
      // wavefronts 1.. meanwhile bring the next launch's batches that will
      // run in this workgroup's slot (same block index modulo the XCD count,
      // so the same XCD) into L2: blocks blockIdx.x + k * rstride
      for (int j = blockIdx.x; j < nb_next; j += rstride)
        touch_batch<NC>(s, bdesc, seg, lnbr, xslot, next0 + j,
                        static_cast<int>(threadIdx.x) - kLdsThreads, T - kLdsThreads);
    }
  } else {
    for (int j = 0; j < nq; j++) {
      const int q = backward ? nq - 1 - j : j;
      const int i1 = lseg[q + 1];
      for (int i = lseg[q] + threadIdx.x; i < i1; i += T) {
        const double a = lap[i];
#pragma unroll
        for (int c = 0; c < NC; c++) {
          if (!act[c]) continue;
          double *lp = lphi + c * np;
          double sum = lb[c * nc + i];
          if (HEX) {  // six faces: gather all slots first, then sum in row order
            const int k0 = 6 * i;
            int sl[6];
            double v[6];
#pragma unroll
            for (int k = 0; k < 6; k++) sl[k] = lnb[k0 + k];
#pragma unroll
            for (int k = 0; k < 6; k++) v[k] = lp[sl[k]];
#pragma unroll
            for (int k = 0; k < 6; k++) sum = sum + lanb[k0 + k] * v[k];
          } else {
            for (int k = lrow[i]; k < lrow[i + 1]; k++) sum = sum + lanb[k] * lp[lnb[k]];
          }
          lp[i] = (sum + (sor - 1.0) * a * lp[i]) / a / sor;
        }
      }
      __syncthreads();
    }
  }
  // the pipelined walk's stores were made by wavefront 0 alone
  if constexpr (PIPE && NW > 1) __syncthreads();
  // System-scope stores write through the XCD's L2, so the end of the
  // kernel has none of these lines to write back before the next launch,
  // whose batches on other XCDs read them (about 1% of the sweep's time).
#pragma unroll
  for (int c = 0; c < NC; c++) {
    if (!act[c]) continue;
    for (int r = threadIdx.x; r < nc; r += T)
      __hip_atomic_store(&s.phi[c][c0 + r], PIPE ? pphi[NC * r + c] : lphi[c * np + r],
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Residual partial sums: block (x, y) covers cells [beg_y + x*chunk, ...).
// absmax: |r| and max|r| (solve_gs, :309-321); else signed r and max(r,0)
// (calc_residual, :239-251).  Fixed-order tree, so reproducible.
// rout, when given, receives each cell's r rounded to FP32 (the right-hand
// side of the mixed-precision correction system).
__device__ inline void d_resid_partial(int bx_, SysDesc s, int chunk, int nblk, int absmax, double *part,
                                float *rout) {
  __shared__ double ssum[kBlock], smax[kBlock];
  const int y = blockIdx.y, x = bx_, tid = threadIdx.x;
  const int lo = s.beg[y] + x * chunk;
  const int hi = min(lo + chunk, s.end[y]);
  const double *phi = s.phi[y], *b = s.b[y];
  double sum = 0.0, mx = 0.0;
  for (int e = lo + tid; e < hi; e += blockDim.x) {
    double r = b[e];
    const int r0 = s.row[e], r1 = s.row[e + 1];
    if (r1 - r0 == 6) {  // hexahedral row: all loads first, the sum in CSR order
      int nb[6];
      double an[6], v[6];
#pragma unroll
      for (int k = 0; k < 6; k++) {
        nb[k] = s.nbr[r0 + k];
        an[k] = s.anb[r0 + k];
      }
#pragma unroll
      for (int k = 0; k < 6; k++) v[k] = phi[nb[k]];
#pragma unroll
      for (int k = 0; k < 6; k++) r = r + an[k] * v[k];
    } else {
      for (int k = r0; k < r1; k++) r = r + s.anb[k] * phi[s.nbr[k]];
    }
    r = r - s.ap[e] * phi[e];
    if (rout) rout[e] = static_cast<float>(r);
    if (absmax) r = fabs(r);
    mx = fmax(r, mx);
    sum = sum + r * r;
  }
  ssum[tid] = sum;
  smax[tid] = mx;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (tid < off) {
      ssum[tid] = ssum[tid] + ssum[tid + off];
      smax[tid] = fmax(smax[tid], smax[tid + off]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    part[2 * (y * nblk + x)] = ssum[0];
    part[2 * (y * nblk + x) + 1] = smax[0];
  }
}
__global__ void k_resid_partial(SysDesc s, int chunk, int nblk, int absmax, double *part,
                                float *rout) {
  d_resid_partial(blockIdx.x, s, chunk, nblk, absmax, part, rout);
}

__device__ inline void d_resid_final(int bx_, const double *part, int nblk, double *red) {
  __shared__ double ssum[kBlock], smax[kBlock];
  const int y = bx_, tid = threadIdx.x;
  double sum = 0.0, mx = 0.0;
  for (int i = tid; i < nblk; i += blockDim.x) {
    sum = sum + part[2 * (y * nblk + i)];
    mx = fmax(mx, part[2 * (y * nblk + i) + 1]);
  }
  ssum[tid] = sum;
  smax[tid] = mx;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (tid < off) {
      ssum[tid] = ssum[tid] + ssum[tid + off];
      smax[tid] = fmax(smax[tid], smax[tid + off]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    red[2 * y] = ssum[0];
    red[2 * y + 1] = smax[0];
  }
}
__global__ void k_resid_final(const double *part, int nblk, double *red) {
  d_resid_final(blockIdx.x, part, nblk, red);
}

// solve_gs bookkeeping, mod_solver.f90:271-287 and :309-322
__device__ inline void d_gs_init(int bx_, GsCtl *c, const double *red, int ny, double ne, int nit) {
  if (threadIdx.x != 0) return;
  for (int y = 0; y < ny; y++) {
    const double ri = sqrt(red[2 * y] / ne);
    c->res_i[y] = ri;
    c->res_f[y] = ri;
    c->target[y] = ri / 10.0;
    c->res_max[y] = 0.0;
    c->it[y] = 0;
    c->active[y] = (0 < nit && ri > c->target[y]) ? 1 : 0;
  }
}
__global__ void k_gs_init(GsCtl *c, const double *red, int ny, double ne, int nit) {
  d_gs_init(blockIdx.x, c, red, ny, ne, nit);
}
__device__ inline void d_gs_iter_end(int bx_, GsCtl *c, const double *red, int ny, double ne, int nit) {
  if (threadIdx.x != 0) return;
  for (int y = 0; y < ny; y++) {
    if (!c->active[y]) continue;
    c->it[y] = c->it[y] + 1;
    c->res_max[y] = red[2 * y + 1];
    c->res_f[y] = sqrt(red[2 * y] / ne);
    c->active[y] = (c->it[y] < nit && c->res_f[y] > c->target[y]) ? 1 : 0;
  }
}
__global__ void k_gs_iter_end(GsCtl *c, const double *red, int ny, double ne, int nit) {
  d_gs_iter_end(blockIdx.x, c, red, ny, ne, nit);
}
struct Names {
  char n[3][8];
};
__device__ inline void d_log_gs(int bx_, const GsCtl *c, int ny, Names nm, LogLine *log, int *nlog) {
  if (threadIdx.x != 0) return;
  for (int y = 0; y < ny; y++) {
    const int k = *nlog;
    if (k >= kLogCap) return;
    LogLine &l = log[k];
    for (int i = 0; i < 8; i++) l.name[i] = nm.n[y][i];
    l.it = c->it[y];
    l.res_i = c->res_i[y];
    l.res_f = c->res_f[y];
    l.res_max = c->res_max[y];
    *nlog = k + 1;
  }
}
__global__ void k_log_gs(const GsCtl *c, int ny, Names nm, LogLine *log, int *nlog) {
  d_log_gs(blockIdx.x, c, ny, nm, log, nlog);
}

// multi_subdomain_solver bookkeeping, mod_solver.f90:140-181 (Q1: res_f_tot
// accumulates onto its previous value; Q2: signed max)
struct SubCounts {
  int n;
  double ne[kMaxSys];
};
__global__ void k_pc_init(PcCtl *c, const double *red, SubCounts sc, int nit) {
  if (threadIdx.x != 0) return;
  double ri = 0.0, rmt = 0.0;
  for (int y = 0; y < sc.n; y++) {
    const double res = sqrt(red[2 * y] / sc.ne[y]);
    ri = ri + res * res;
    rmt = fmax(rmt, red[2 * y + 1]);
  }
  ri = sqrt(ri / static_cast<double>(sc.n));
  c->ri = ri;
  c->rf = ri;
  c->rmt = rmt;
  c->target = ri / 10.0;
  c->it = 0;
  c->active = (0 < nit && c->rf > c->target) ? 1 : 0;
}
__device__ void pc_block_end(PcCtl *c, const double *red, SubCounts sc, int nit, int has_res) {
  if (!c->active) return;
  c->it = c->it + 2;
  if (has_res) {
    double rf = c->rf, rmt = c->rmt;
    for (int y = 0; y < sc.n; y++) {
      const double res = sqrt(red[2 * y] / sc.ne[y]);
      rf = rf + res * res;
      rmt = fmax(rmt, red[2 * y + 1]);
    }
    c->rf = sqrt(rf / static_cast<double>(sc.n));
    c->rmt = rmt;
  }
  c->active = (c->it < nit && c->rf > c->target) ? 1 : 0;
}
__global__ void k_pc_block_end(PcCtl *c, const double *red, SubCounts sc, int nit, int has_res) {
  if (threadIdx.x == 0) pc_block_end(c, red, sc, nit, has_res);
}
// A block without a residual check: the halo exchange and the block's stop
// test in one launch.  The exchange does not read the active flag the stop
// test writes: once a solve has stopped its values no longer change, so the
// copies it makes are the ones already made (or, stopped on entry, the
// gathered halo values themselves).
template <typename T>
__global__ void k_exchange_end(int n, const int *dst, const int *src, T *phic, PcCtl *c,
                               const double *red, SubCounts sc, int nit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) phic[dst[i]] = phic[src[i]];
  if (i == 0) pc_block_end(c, red, sc, nit, 0);
}
__global__ void k_log_pc(const PcCtl *c, LogLine *log, int *nlog) {
  if (threadIdx.x != 0) return;
  const int k = *nlog;
  if (k >= kLogCap) return;
  LogLine &l = log[k];
  for (int i = 0; i < 8; i++) l.name[i] = 0;
  l.name[0] = 'p';
  l.name[1] = 'c';
  l.it = c->it;
  l.res_i = c->ri;
  l.res_f = c->rf;
  l.res_max = c->rmt;
  *nlog = k + 1;
}

// assemble_coef / update_halos / update_phi, mod_subdomains.f90:162-229
__global__ void k_pc_assemble(int ne, const int *l2g_cell, const int *rowp, const int *rowg,
                              const double *ap, const double *b, const double *anb,
                              const double *phi, double *ap_p, double *b_p, double *anb_p,
                              double *phic_p, double *rap_p) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ne) return;
  const int g = l2g_cell[c];
  ap_p[c] = ap[g];
  rap_p[c] = rcp_of(ap[g]);
  b_p[c] = b[g];
  phic_p[c] = phi[g];
  const int kg = rowg[g];
  for (int k = rowp[c]; k < rowp[c + 1]; k++) anb_p[k] = anb[kg + (k - rowp[c])];
}
__global__ void k_gather(int n, const int *idx, const double *src, double *dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
template <typename T>
__global__ void k_exchange(int n, const int *dst, const int *src, T *phic, const int *active) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && active[0]) phic[dst[i]] = phic[src[i]];
}
__global__ void k_scatter(int n, const int *idx, const double *src, double *dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = src[i];
}

// The global system in tile numbering for a tiled solve_gs: gather the
// shared coefficients, the right-hand sides and the values (boundary halos
// keep their slots ne..ne+nbf-1; rows keep their entry order) before the
// solve, scatter the values back after it.
struct TileMap {
  int ne, nbf, ny;
  const int *l2g, *rowm, *rowg;
  const double *ap, *anb;
  double *ap_m, *anb_m, *rap_m;
  const double *b[3];
  double *b_m[3];
  double *phi[3];
  double *phi_m[3];
};
__device__ inline void d_tile_gather(int bx_, TileMap t) {
  const int i = bx_ * blockDim.x + threadIdx.x;
  if (i < t.ne) {
    const int g = t.l2g[i];
    t.ap_m[i] = t.ap[g];
    t.rap_m[i] = rcp_of(t.ap[g]);
    const int r = t.rowm[i], rg = t.rowg[g], n = t.rowg[g + 1] - rg;
    for (int k = 0; k < n; k++) t.anb_m[r + k] = t.anb[rg + k];
    for (int y = 0; y < t.ny; y++) {
      t.b_m[y][i] = t.b[y][g];
      t.phi_m[y][i] = t.phi[y][g];
    }
  } else if (i < t.ne + t.nbf) {
    for (int y = 0; y < t.ny; y++) t.phi_m[y][i] = t.phi[y][i];
  }
}
__global__ void k_tile_gather(TileMap t) {
  d_tile_gather(blockIdx.x, t);
}
__device__ inline void d_tile_scatter(int bx_, TileMap t) {
  const int i = bx_ * blockDim.x + threadIdx.x;
  if (i >= t.ne) return;
  const int g = t.l2g[i];
  for (int y = 0; y < t.ny; y++) t.phi[y][g] = t.phi_m[y][i];
}
__global__ void k_tile_scatter(TileMap t) {
  d_tile_scatter(blockIdx.x, t);
}


// ---------------------------------------------------------------------------
// Batched iterations (continuous batching of independent cases): one launch
// per step of the SIMPLE iteration covers every case of a batch.  Case k
// owns the blocks [b0[k], b0[k+1]) of a launch (cfdl_group.h) and runs them
// as blocks 0 .. of its own grid with its own descriptor, so its arithmetic
// is that of the same step on the case alone, bit for bit.
// ---------------------------------------------------------------------------
constexpr int kMaxColours = 16;  // colours of a case's momentum system
struct BSlot {
  MeshDev md;
  FieldsDev fd;
  EnergyDev en;
  Inlet in;
  double dt, alpha_u, alpha_p;
  const int *fin, *fout;
  int nin, nout;
  int nblk_g, chunk_g, nit, nq;
  int gq[kMaxColours + 1];  // colour ranges of the colour numbering
  TileMap tmap;             // u, v, w and their system in the colour numbering
  SysDesc sys;              // the same system as the sweep and residual kernels see it
  GsCtl *gctl;
  double *part, *red, *scal;
  LogLine *log;
  int *nlog;
};

#define BSLOT                                                                       \
  int bx;                                                                           \
  const BSlot &x = S[group_slot(g, bx)]

__global__ void gb_bc_update(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_bc_update(bx, x.md, x.fd, x.in);
}
__global__ void __launch_bounds__(kOutThreads)
gb_outlet_flux(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_outlet_flux(bx, x.md, x.fd.mip, x.fin, x.nin, x.fout, x.nout);
}
__global__ void gb_bc_energy(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_bc_energy(bx, x.md, x.en);
}
template <typename T>
__global__ void __launch_bounds__(kFaceThreads)
gb_coef_uvw6(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_coef_uvw6<T>(bx, x.md, x.fd, x.dt, x.alpha_u);
}
__global__ void gb_tile_gather(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_tile_gather(bx, x.tmap);
}
__global__ void gb_tile_scatter(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_tile_scatter(bx, x.tmap);
}
// grid (blocks, 3): the three right-hand sides in y, as k_resid_partial
__global__ void __launch_bounds__(kBlock) gb_resid_partial(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_resid_partial(bx, x.sys, x.chunk_g, x.nblk_g, 1, x.part, nullptr);
}
// grid (cases, 3): one block per case and right-hand side
__global__ void __launch_bounds__(kBlock) gb_resid_final(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_resid_final(blockIdx.y, x.part, x.nblk_g, x.red);
}
__global__ void gb_gs_init(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_gs_init(bx, x.gctl, x.red, 3, static_cast<double>(x.md.ne), x.nit);
}
__global__ void gb_gs_iter_end(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_gs_iter_end(bx, x.gctl, x.red, 3, static_cast<double>(x.md.ne), x.nit);
}
// colour i of the case's forward order, or of its reverse (backward)
__global__ void gb_gs_colour(GroupMap g, const BSlot *__restrict__ S, int i, int backward,
                             double sor) {
  BSLOT;
  const int q = backward ? x.nq - 1 - i : i;
  const int b0 = x.gq[q];
  d_gs_level<3>(bx, x.sys, nullptr, b0, x.gq[q + 1] - b0, sor, x.gctl->active);
}
// flags[k]: case k's momentum solve is still running
__global__ void gb_gs_flags(int n, const BSlot *__restrict__ S, int *flags) {
  const int k = threadIdx.x;
  if (k < n) {
    const int *a = S[k].gctl->active;
    flags[k] = a[0] | a[1] | a[2];
  }
}
__global__ void gb_log_gs(GroupMap g, const BSlot *__restrict__ S, Names nm) {
  BSLOT;
  d_log_gs(bx, x.gctl, 3, nm, x.log, x.nlog);
}
// which: 0 u -> gu, 1 v -> gv, 2 w -> gw, 3 phic -> gpc
__global__ void gb_grad(GroupMap g, const BSlot *__restrict__ S, int which) {
  BSLOT;
  const FieldsDev &f = x.fd;
  const double *phi = which == 0 ? f.u : which == 1 ? f.v : which == 2 ? f.w : f.phic;
  double *gr = which == 0 ? f.gu : which == 1 ? f.gv : which == 2 ? f.gw : f.gpc;
  d_grad(bx, x.md, phi, gr);
}
__global__ void gb_mip(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_mip(bx, x.md, x.fd, 1, x.dt);
}
template <typename T>
__global__ void __launch_bounds__(kFaceThreads) gb_coef_p6(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_coef_p6<T>(bx, x.md, x.fd);
}
__global__ void gb_zero_phic(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_fill(bx, x.fd.phic, int64_t{x.md.ne} + x.md.nbf, 0.0);
}
__global__ void gb_pref(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_pref(bx, x.fd.phic, x.scal);
}
__global__ void gb_sub_pref(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_sub_pref(bx, x.md.ne, x.fd.phic, x.scal);
}
__global__ void gb_halo_from_owner(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_halo_from_owner(bx, x.md, x.fd.phic);
}
__global__ void gb_update_cells(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_update_cells(bx, x.md.ne, x.fd, x.alpha_p);
}
__global__ void gb_update_faces(GroupMap g, const BSlot *__restrict__ S) {
  BSLOT;
  d_update_faces(bx, x.md, x.fd);
}
// update_time: which 0 u0 <- u, 1 v0 <- v, 2 w0 <- w, 3 mip0 <- mip, 4 h0 <- h
__global__ void gb_update_time(GroupMap g, const BSlot *__restrict__ S, int which) {
  BSLOT;
  const FieldsDev &f = x.fd;
  const int64_t n = which == 3 ? int64_t{x.md.nf} : int64_t{x.md.ne} + x.md.nbf;
  double *dst = which == 0 ? f.u0 : which == 1 ? f.v0 : which == 2 ? f.w0 : which == 3 ? f.mip0
                                                                                         : x.en.h0;
  const double *src = which == 0 ? f.u : which == 1 ? f.v : which == 2 ? f.w : which == 3 ? f.mip
                                                                                         : x.en.h;
  d_copy(bx, dst, src, n);
}
#undef BSLOT

}  // namespace cfdl

using namespace cfdl;

// ===========================================================================
// Host side
// ===========================================================================
namespace {

struct Sched {
  std::vector<int> ptr;  // level starts (host)
  int *cells = nullptr;  // device, cells sorted by level
  int nlev = 0;
};

struct DevBuf {
  std::vector<void *> ptrs;
  ~DevBuf() {
    for (void *p : ptrs) (void)hipFree(p);
  }
};

}  // namespace

struct cfdl_ctx {
  int dev = 0;
  hipStream_t st = nullptr;
  // while the case is attached to a batch (cfdl_batch_attach) its work runs
  // on the batch's stream; own_st is the context's own stream meanwhile
  cfdl_batch *batch = nullptr;
  hipStream_t own_st = nullptr;
  int ne = 0, nf = 0, nbf = 0, L = 0, nsubd = 1, nit = 100;
  double dt = 0.01;
  double alpha_u = 1.0, alpha_p = 1.0;
  DevBuf mem;
  MeshDev md{};
  FieldsDev fd{};
  EnergyDev en{};
  double *fields[CFDL_NFIELDS] = {};
  int64_t fsize[CFDL_NFIELDS] = {};
  Sched gf, gb;  // global system, forward / backward sweeps
  // pressure-correction subdomain systems (concatenated)
  int Lp = 0, nh = 0, nx = 0;
  int *rowp = nullptr, *nbrp = nullptr, *l2g_cell = nullptr, *halo_src = nullptr;
  // the pc system numbered by (subdomain, colour, local order) for the
  // colour order: each colour of a subdomain is a contiguous cell range
  int *rowp_c = nullptr, *nbrp_c = nullptr, *l2g_c = nullptr, *xsrc_c = nullptr;
  int *xdst = nullptr, *xsrc = nullptr, *rowg = nullptr;
  double *ap_p = nullptr, *b_p = nullptr, *anb_p = nullptr, *phic_p = nullptr, *rap_p = nullptr;
  std::vector<int> co;  // subdomain cell offsets (nsubd+1)
  Sched pcl;            // pc system by forward level (CFDL_GS_MODE=level)
  // colour order of the pc smoother (cfdl_set_pc_order, an extension): the
  // colours of a greedy colouring in each subdomain's local order, as levels
  Sched pcc;
  std::vector<ColRanges> pcc_rng;  // per colour, its range in each subdomain
  bool pc_hex = false;             // every pc row has six entries (row e at 6e)
  int pc_order = CFDL_PC_REFERENCE;
  // FP32 correction system of CFDL_PC_COLORED_MIXED (colour numbering):
  // ap, anb, residual and correction (cells and halos)
  float *a32 = nullptr, *anb32 = nullptr, *r32 = nullptr, *d32 = nullptr;
  int pc_nlev = 0;
  // FP32 assembly of the momentum and pc coefficients (an extension; its
  // FP32 copies of the geometry live in md.*32)
  bool asm32 = false;
  // channel extension: inlet profile, the host copy of the halo BC kinds and
  // the inlet / outlet faces (halo order) whose fluxes k_outlet_flux balances
  Inlet inlet{{1.0, 0.0, 1.0, 0.0, 1.0}};
  std::vector<int> hkind_h, hslot_h, fsg_h;
  int *fin = nullptr, *fout = nullptr;
  int nin = 0, nout = 0;
  // multigrid-preconditioned CG pc solve (CFDL_PC_MGCG[_MIXED], an
  // extension): the hierarchy's device copy, FP64 and FP32 preconditioner
  cfdl::MgCgSolver *mg[2] = {nullptr, nullptr};
  // space-time tiled schedule of the pc sweeps (the backward sweep runs it
  // in reverse)
  struct TileSched {
    std::vector<int> launch_ptr;   // batches of launch m: [launch_ptr[m], launch_ptr[m+1])
    std::vector<int> launch_cells;  // cells updated by launch m
    int *bseg = nullptr, *seg = nullptr, *xptr = nullptr, *xslot = nullptr;
    // per batch: {first segment, segments, first cell, cells, first CSR entry,
    // entries, first outside neighbour, outside neighbours}
    int4 *bdesc = nullptr;
    // per CSR entry: the neighbour's LDS slot, or with pipe its values' LDS
    // byte address for the table's right-hand-side count (lnbr) and for one
    // right-hand side (lnbr1, tables of several: the energy solve_gs)
    unsigned short *lnbr = nullptr, *lnbr1 = nullptr;
    bool hex = false;   // every row has six entries
    bool pipe = false;  // hex, and no level segment has more than 64 cells
    int nw = 4;         // wavefronts per batch in the gather and write-back (1 or 4)
    bool prefetch = true;  // idle wavefronts bring the next launch's batches into L2
    int nbatch = 0, nseg = 0;
    size_t lds_bytes = 0;  // 0: batches too large for LDS, global-memory kernel
  } ts, tm;
  // the global system in tile numbering (tiled solve_gs): new -> old cell,
  // CSR, and copies of the shared coefficients, right-hand sides and values
  bool mtiled = false;
  int *l2g_m = nullptr, *rowm = nullptr, *nbrm = nullptr;
  double *ap_m = nullptr, *anb_m = nullptr, *rap_m = nullptr, *b_m[3] = {}, *phi_m[3] = {};
  bool tiled = true;
  // colour order of solve_gs (cfdl_set_gs_order, an extension): the global
  // system numbered by (colour, global order), each colour one range
  int gs_order = CFDL_GS_REFERENCE;
  int *l2g_gc = nullptr, *rowgc = nullptr, *nbrgc = nullptr;
  std::vector<int> gcol_ptr;
  bool mhex = false;  // every row of the mesh CSR has six entries (row e at 6e)
  bool rcp = true;  // reciprocal-based division in the pipelined walk
  // control
  GsCtl *gctl = nullptr;
  PcCtl *pctl = nullptr;
  double *part = nullptr, *red = nullptr, *scal = nullptr;
  int nblk_g = 1, chunk_g = 1, nblk_p = 1, chunk_p = 1;
  int *host_flags = nullptr;  // pinned
  LogLine *dlog = nullptr;
  int *dlog_n = nullptr;
  std::vector<LogLine> hlog;
  // launch counters and sampled timing of the pressure-correction sweep
  // kernel: events around every sample_every-th sweep (all its launches,
  // back to back on the stream)
  int64_t gs_launches = 0, pc_launches = 0, pc_cells = 0, outer = 0, pc_sweeps = 0;
  int sample_every = 0;
  struct Sample {
    hipEvent_t a, b;
    int64_t cells;
    int launches;
  };
  std::vector<Sample> samples;
  size_t nsamples = 0;
};

template <typename T>
static int dalloc(cfdl_ctx *c, T **p, int64_t n) {
  void *q = nullptr;
  const size_t bytes = sizeof(T) * static_cast<size_t>(n > 0 ? n : 1);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc of ") + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    return CFDL_ENOMEM;
  }
  c->mem.ptrs.push_back(q);
  *p = static_cast<T *>(q);
  return CFDL_OK;
}

template <typename T>
static int upload(cfdl_ctx *c, T **p, const std::vector<T> &v) {
  CALL(dalloc(c, p, static_cast<int64_t>(v.size())));
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return CFDL_OK;
}

static inline int nblocks(int64_t n) { return static_cast<int>((n + kBlock - 1) / kBlock); }

// Dependency levels of a sweep over cells [beg, end) in increasing (fwd) or
// decreasing (bwd) order: a cell waits for neighbours in the same range that
// the sweep visits before it.  Cells of one level are independent.
static void levels(const std::vector<int> &row, const std::vector<int> &nbr, int beg, int end,
                   bool fwd, std::vector<int> &lev) {
  if (fwd) {
    for (int e = beg; e < end; e++) {
      int l = 0;
      for (int k = row[e]; k < row[e + 1]; k++) {
        const int n = nbr[k];
        if (n >= beg && n < e) l = std::max(l, lev[n]);
      }
      lev[e] = l + 1;
    }
  } else {
    for (int e = end - 1; e >= beg; e--) {
      int l = 0;
      for (int k = row[e]; k < row[e + 1]; k++) {
        const int n = nbr[k];
        if (n > e && n < end) l = std::max(l, lev[n]);
      }
      lev[e] = l + 1;
    }
  }
}

static int make_sched(cfdl_ctx *c, const std::vector<int> &lev, int n, Sched *s) {
  int nlev = 0;
  for (int e = 0; e < n; e++) nlev = std::max(nlev, lev[e]);
  std::vector<int> cnt(nlev + 1, 0);
  for (int e = 0; e < n; e++) cnt[lev[e]]++;
  s->ptr.assign(nlev + 1, 0);
  for (int l = 1; l <= nlev; l++) s->ptr[l] = s->ptr[l - 1] + cnt[l];
  std::vector<int> pos(s->ptr.begin(), s->ptr.end() - 1), cells(n);
  for (int e = 0; e < n; e++) cells[pos[lev[e] - 1]++] = e;
  s->nlev = nlev;
  return upload(c, &s->cells, cells);
}

static int sweep(cfdl_ctx *c, const Sched &s, const SysDesc &sys, int nc, double sor,
                 const int *active) {
  for (int l = 0; l < s.nlev; l++) {
    const int n = s.ptr[l + 1] - s.ptr[l];
    if (nc == 3)
      hipLaunchKernelGGL(k_gs_level<3>, dim3(nblocks(n)), dim3(kBlock), 0, c->st, sys,
                         s.cells + s.ptr[l], 0, n, sor, active);
    else
      hipLaunchKernelGGL(k_gs_level<1>, dim3(nblocks(n)), dim3(kBlock), 0, c->st, sys,
                         s.cells + s.ptr[l], 0, n, sor, active);
  }
  c->gs_launches += s.nlev;
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

// Sampled timing of the pressure-correction sweeps: HIP events on the
// context stream before the first and after the last launch of every
// sample_every-th sweep.  The launches run back to back in between, so
// elapsed / launches is the mean launch duration without the dispatch
// latency that an event pair around a single launch adds.
static int sample_begin(cfdl_ctx *c, cfdl_ctx::Sample **smp) {
  *smp = nullptr;
  const int64_t k = c->pc_sweeps++;
  if (c->sample_every <= 0 || k % c->sample_every != 0) return CFDL_OK;
  if (c->nsamples == c->samples.size()) {
    cfdl_ctx::Sample x{};
    HIPCHK(hipEventCreate(&x.a));
    HIPCHK(hipEventCreate(&x.b));
    c->samples.push_back(x);
  }
  *smp = &c->samples[c->nsamples++];
  HIPCHK(hipEventRecord((*smp)->a, c->st));
  return CFDL_OK;
}

static int sample_end(cfdl_ctx *c, cfdl_ctx::Sample *smp, int launches, int64_t cells) {
  if (!smp) return CFDL_OK;
  smp->launches = launches;
  smp->cells = cells;
  HIPCHK(hipEventRecord(smp->b, c->st));
  return CFDL_OK;
}

// A pressure-correction sweep, one launch per dependency level: forward
// visits the levels in order, backward in reverse (a cell's neighbours on
// lower forward levels come earlier in the sequential order, so reversing
// the level order is a valid backward schedule).
static int pc_sweep(cfdl_ctx *c, const Sched &s, const SysDesc &sys, bool forward,
                    const int *active) {
  cfdl_ctx::Sample *smp = nullptr;
  CALL(sample_begin(c, &smp));
  for (int i = 0; i < s.nlev; i++) {
    const int l = forward ? i : s.nlev - 1 - i;
    const int n = s.ptr[l + 1] - s.ptr[l];
    hipLaunchKernelGGL(k_gs_level<1>, dim3(nblocks(n)), dim3(kBlock), 0, c->st, sys,
                       s.cells + s.ptr[l], 0, n, 1.02, active);
  }
  CALL(sample_end(c, smp, s.nlev, c->ne));
  c->pc_launches += s.nlev;
  c->pc_cells += c->ne;
  c->gs_launches += s.nlev;
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

// An FP32 colour sweep of the mixed-precision correction system (sor 1.02 in
// single precision), colours in order forward and in reverse backward.
static int pc_sweep32(cfdl_ctx *c, const SysDesc32 &sys, bool forward, const int *active) {
  cfdl_ctx::Sample *smp = nullptr;
  CALL(sample_begin(c, &smp));
  const int nl = static_cast<int>(c->pcc_rng.size());
  for (int i = 0; i < nl; i++) {
    const ColRanges &cr = c->pcc_rng[forward ? i : nl - 1 - i];
    if (c->pc_hex)
      hipLaunchKernelGGL(k_gs_level32<true>, dim3(nblocks(cr.off[cr.n])), dim3(kBlock), 0, c->st,
                         sys, cr, 1.02f, active);
    else
      hipLaunchKernelGGL(k_gs_level32<false>, dim3(nblocks(cr.off[cr.n])), dim3(kBlock), 0,
                         c->st, sys, cr, 1.02f, active);
  }
  CALL(sample_end(c, smp, nl, c->ne));
  c->pc_launches += nl;
  c->pc_cells += c->ne;
  c->gs_launches += nl;
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

// One launch of the tiled sweep: batches [b0, b0 + nb), NW wavefronts each;
// the next launch in sweep order has nb1 batches from b1 (nb1 = 0: none, or
// no prefetch).  Workgroups go to the XCDs round robin by block index,
// starting at the same XCD in every launch, so block j + k * 8 * ceil(nb/8)
// of the next launch runs on block j's XCD.
template <int NC, int NW>
static void tiled_launch(cfdl_ctx *c, const cfdl_ctx::TileSched &t, const SysDesc &sys, int b0,
                         int nb, int b1, int nb1, int bwd, double sor, double rsor,
                         const int *active) {
  const dim3 blk(kLdsThreads * NW);
  const int rs = kXcds * ((nb + kXcds - 1) / kXcds);
  TileSys<NC> ts;
  for (int c = 0; c < NC; c++) {
    ts.phi[c] = sys.phi[c];
    ts.b[c] = sys.b[c];
  }
  ts.ap = sys.ap;
  ts.anb = sys.anb;
  ts.rap = sys.rap;
  ts.row = sys.row;
  const unsigned short *ln = (NC == 1 && t.lnbr1) ? t.lnbr1 : t.lnbr;
  if (t.pipe)
    hipLaunchKernelGGL((k_gs_tiles_lds<true, true, NC, NW>), dim3(nb), blk, t.lds_bytes, c->st,
                       b0, t.bdesc, active, ts, t.seg, ln, t.xslot, bwd, sor, rsor, b1, nb1,
                       rs);
  else if (t.hex)
    hipLaunchKernelGGL((k_gs_tiles_lds<true, false, NC, NW>), dim3(nb), blk, t.lds_bytes, c->st,
                       b0, t.bdesc, active, ts, t.seg, ln, t.xslot, bwd, sor, rsor, b1, nb1,
                       rs);
  else
    hipLaunchKernelGGL((k_gs_tiles_lds<false, false, NC, NW>), dim3(nb), blk, t.lds_bytes, c->st,
                       b0, t.bdesc, active, ts, t.seg, ln, t.xslot, bwd, sor, rsor, b1, nb1,
                       rs);
}

// One tiled sweep of a system in its tile numbering: the launches of the
// schedule in order (forward) or in reverse (backward), NC right-hand sides.
template <int NC>
static void tiled_launches(cfdl_ctx *c, const cfdl_ctx::TileSched &t, const SysDesc &sys,
                           bool forward, double sor, const int *active) {
  const int nl = static_cast<int>(t.launch_ptr.size()) - 1;
  const int bwd = forward ? 0 : 1;
  // 1/sor correctly rounded, as div_rcp2 requires; 0 selects the hardware
  // division in the kernel (CFDL_TILE_RCP=0, or sor outside div_rcp2's
  // divisor range)
  const bool rcp_ok = c->rcp && sor >= kRcpDivMin && sor <= kRcpDivMax;
  const double rsor = rcp_ok ? 1.0 / sor : 0.0;
  for (int i = 0; i < nl; i++) {
    const int m = forward ? i : nl - 1 - i;
    const int b0 = t.launch_ptr[m], nb = t.launch_ptr[m + 1] - b0;
    const int m1 = forward ? m + 1 : m - 1;
    const bool has_next = t.prefetch && m1 >= 0 && m1 < nl;
    const int b1 = has_next ? t.launch_ptr[m1] : 0;
    const int nb1 = has_next ? t.launch_ptr[m1 + 1] - b1 : 0;
    if (t.nw == 4)
      tiled_launch<NC, 4>(c, t, sys, b0, nb, b1, nb1, bwd, sor, rsor, active);
    else
      tiled_launch<NC, 1>(c, t, sys, b0, nb, b1, nb1, bwd, sor, rsor, active);
  }
}

static int pc_sweep_tiled(cfdl_ctx *c, const SysDesc &sys, bool forward, const int *active) {
  const cfdl_ctx::TileSched &t = c->ts;
  const int nl = static_cast<int>(t.launch_ptr.size()) - 1;
  cfdl_ctx::Sample *smp = nullptr;
  CALL(sample_begin(c, &smp));
  if (t.lds_bytes > 0) {
    tiled_launches<1>(c, t, sys, forward, 1.02, active);
  } else {
    for (int i = 0; i < nl; i++) {
      const int m = forward ? i : nl - 1 - i;
      const int b0 = t.launch_ptr[m], nb = t.launch_ptr[m + 1] - b0;
      hipLaunchKernelGGL(k_gs_tiles, dim3(nb), dim3(kTileThreads), 0, c->st, sys, t.bseg, t.seg,
                         b0, forward ? 0 : 1, 1.02, active);
    }
  }
  CALL(sample_end(c, smp, nl, c->ne));
  c->pc_launches += nl;
  c->pc_cells += c->ne;
  c->gs_launches += nl;
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

static int residual(cfdl_ctx *c, const SysDesc &sys, int nblk, int chunk, int absmax,
                    float *rout = nullptr) {
  hipLaunchKernelGGL(k_resid_partial, dim3(nblk, sys.ny), dim3(kBlock), 0, c->st, sys, chunk,
                     nblk, absmax, c->part, rout);
  hipLaunchKernelGGL(k_resid_final, dim3(sys.ny), dim3(kBlock), 0, c->st, c->part, nblk, c->red);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

static double sor_of(const char *name) {
  return (name && name[0] == 'p' && name[1] == 'c') ? 1.02 : 1.0;
}

// solve_gs, mod_solver.f90:255-327, for ny right-hand sides on the global
// system; the host checks the per-RHS stop flags after every iteration.
static int *gs_active(cfdl_ctx *c) {
  return reinterpret_cast<int *>(reinterpret_cast<char *>(c->gctl) + offsetof(GsCtl, active));
}
static int *pc_active(cfdl_ctx *c) {
  return reinterpret_cast<int *>(reinterpret_cast<char *>(c->pctl) + offsetof(PcCtl, active));
}

static int solve_gs_n(cfdl_ctx *c, int ny, double **phi, const double **b,
                      const char *const *names, double sor) {
  Names nm{};
  for (int y = 0; y < ny; y++) std::strncpy(nm.n[y], names[y], 7);
  // tiled: the system and its right-hand sides in tile numbering (one or
  // three right-hand sides; solve_gs_uvw passes three); colour order: in the
  // colour numbering of the same system, through the same copies
  const bool gcol = c->gs_order == CFDL_GS_COLORED && (ny == 1 || ny == 3);
  const bool tiled = !gcol && c->mtiled && (ny == 1 || ny == 3);
  const bool mapped = tiled || gcol;
  TileMap tmap{};
  if (mapped) {
    tmap.ne = c->ne;
    tmap.nbf = c->nbf;
    tmap.ny = ny;
    tmap.l2g = gcol ? c->l2g_gc : c->l2g_m;
    tmap.rowm = gcol ? c->rowgc : c->rowm;
    tmap.rowg = c->md.row;
    tmap.ap = c->fd.ap;
    tmap.anb = c->fd.anb;
    tmap.ap_m = c->ap_m;
    tmap.anb_m = c->anb_m;
    tmap.rap_m = c->rap_m;
    for (int y = 0; y < ny; y++) {
      tmap.b[y] = b[y];
      tmap.b_m[y] = c->b_m[y];
      tmap.phi[y] = phi[y];
      tmap.phi_m[y] = c->phi_m[y];
    }
    hipLaunchKernelGGL(k_tile_gather, dim3(nblocks(int64_t{c->ne} + c->nbf)), dim3(kBlock), 0,
                       c->st, tmap);
  }
  SysDesc s{};
  s.ny = ny;
  for (int y = 0; y < ny; y++) {
    s.beg[y] = 0;
    s.end[y] = c->ne;
    s.phi[y] = mapped ? c->phi_m[y] : phi[y];
    s.b[y] = mapped ? c->b_m[y] : b[y];
  }
  s.row = gcol ? c->rowgc : tiled ? c->rowm : c->md.row;
  s.nbr = gcol ? c->nbrgc : tiled ? c->nbrm : c->md.nbr;
  s.ap = mapped ? c->ap_m : c->fd.ap;
  s.anb = mapped ? c->anb_m : c->fd.anb;
  s.rap = mapped ? c->rap_m : nullptr;
  CALL(residual(c, s, c->nblk_g, c->chunk_g, 1));
  hipLaunchKernelGGL(k_gs_init, dim3(1), dim3(64), 0, c->st, c->gctl, c->red, ny,
                     static_cast<double>(c->ne), c->nit);
  // one SysDesc per sweep kernel with its RHS subset in slots 0..ny-1
  for (int it = 0; it < c->nit; it++) {
    if (gcol) {
      // colours in order, then in reverse; one launch per colour
      const int nq = static_cast<int>(c->gcol_ptr.size()) - 1;
      for (int d = 0; d < 2; d++)
        for (int i = 0; i < nq; i++) {
          const int q = d == 0 ? i : nq - 1 - i;
          const int b0 = c->gcol_ptr[q], n = c->gcol_ptr[q + 1] - b0;
          if (ny == 3)
            hipLaunchKernelGGL(k_gs_level<3>, dim3(nblocks(n)), dim3(kBlock), 0, c->st, s,
                               static_cast<const int *>(nullptr), b0, n, sor, gs_active(c));
          else
            hipLaunchKernelGGL(k_gs_level<1>, dim3(nblocks(n)), dim3(kBlock), 0, c->st, s,
                               static_cast<const int *>(nullptr), b0, n, sor, gs_active(c));
        }
      c->gs_launches += 2 * nq;
      HIPCHK(hipGetLastError());
    } else if (tiled) {
      const int nl = static_cast<int>(c->tm.launch_ptr.size()) - 1;
      if (ny == 3) {
        tiled_launches<3>(c, c->tm, s, true, sor, gs_active(c));
        tiled_launches<3>(c, c->tm, s, false, sor, gs_active(c));
      } else {
        tiled_launches<1>(c, c->tm, s, true, sor, gs_active(c));
        tiled_launches<1>(c, c->tm, s, false, sor, gs_active(c));
      }
      c->gs_launches += 2 * nl;
      HIPCHK(hipGetLastError());
    } else {
      CALL(sweep(c, c->gf, s, ny, sor, gs_active(c)));
      CALL(sweep(c, c->gb, s, ny, sor, gs_active(c)));
    }
    CALL(residual(c, s, c->nblk_g, c->chunk_g, 1));
    hipLaunchKernelGGL(k_gs_iter_end, dim3(1), dim3(64), 0, c->st, c->gctl, c->red, ny,
                       static_cast<double>(c->ne), c->nit);
    HIPCHK(hipMemcpyAsync(c->host_flags, gs_active(c), sizeof(int) * 3, hipMemcpyDeviceToHost,
                          c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    bool any = false;
    for (int y = 0; y < ny; y++) any = any || c->host_flags[y];
    if (!any) break;
  }
  if (mapped)
    hipLaunchKernelGGL(k_tile_scatter, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, tmap);
  hipLaunchKernelGGL(k_log_gs, dim3(1), dim3(64), 0, c->st, c->gctl, ny, nm, c->dlog, c->dlog_n);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

static int multi_subdomain_solver(cfdl_ctx *c) {
  const int ne = c->ne, ns = c->nsubd;
  // the colour orders run on the colour numbering of the same system
  const bool mixed = c->pc_order == CFDL_PC_COLORED_MIXED;
  const bool col = c->pc_order == CFDL_PC_COLORED || mixed;
  const int *l2g = col ? c->l2g_c : c->l2g_cell, *rowp = col ? c->rowp_c : c->rowp;
  const int *nbrp = col ? c->nbrp_c : c->nbrp, *xsrc = col ? c->xsrc_c : c->xsrc;
  // assemble_coef, mod_subdomains.f90:162-189
  hipLaunchKernelGGL(k_pc_assemble, dim3(nblocks(ne)), dim3(kBlock), 0, c->st, ne, l2g,
                     rowp, c->md.row, c->fd.ap, c->fd.b, c->fd.anb, c->fd.phic, c->ap_p,
                     c->b_p, c->anb_p, c->phic_p, c->rap_p);
  if (c->nh > 0)
    hipLaunchKernelGGL(k_gather, dim3(nblocks(c->nh)), dim3(kBlock), 0, c->st, c->nh, c->halo_src,
                       c->fd.phic, c->phic_p + ne);
  SysDesc s{};
  s.ny = ns;
  SubCounts sc{};
  sc.n = ns;
  for (int y = 0; y < ns; y++) {
    s.beg[y] = c->co[y];
    s.end[y] = c->co[y + 1];
    s.phi[y] = c->phic_p;
    s.b[y] = c->b_p;
    sc.ne[y] = static_cast<double>(c->co[y + 1] - c->co[y]);
  }
  s.row = rowp;
  s.nbr = nbrp;
  s.ap = c->ap_p;
  s.anb = c->anb_p;
  s.rap = c->rap_p;
  // mixed precision (an extension): FP64 defect correction around FP32
  // sweeps.  The sweeps solve ap*d - sum(anb*d_nb) = r for a correction d
  // (halos exchanged like the values); at every residual check the
  // correction is added to the FP64 values and cleared, and the FP64 true
  // residual gives both the stop test's norms and the next r.
  const int nall = ne + c->nh;
  SysDesc32 s32{rowp, nbrp, c->a32, c->anb32, c->r32, c->d32};
  if (mixed)
    hipLaunchKernelGGL(k_mixed_prep, dim3(std::min(nblocks(std::max(c->Lp, nall)), 4096)),
                       dim3(kBlock), 0, c->st,
                       ne, c->Lp, nall, c->ap_p, c->anb_p, c->a32, c->anb32, c->d32);
  CALL(residual(c, s, c->nblk_p, c->chunk_p, 0, mixed ? c->r32 : nullptr));
  hipLaunchKernelGGL(k_pc_init, dim3(1), dim3(64), 0, c->st, c->pctl, c->red, sc, c->nit);
  SysDesc s1 = s;  // the sweeps see the whole concatenated system as one RHS
  s1.ny = 1;
  const int *act = pc_active(c);
  const int nblock = (c->nit + 1) / 2;
  for (int blk = 1; blk <= nblock; blk++) {
    for (int rep = 0; rep < 2; rep++) {  // smoother_gs(..., 2): two symmetric iterations
      if (mixed) {
        CALL(pc_sweep32(c, s32, true, act));
        CALL(pc_sweep32(c, s32, false, act));
      } else if (c->pc_order == CFDL_PC_COLORED) {
        CALL(pc_sweep(c, c->pcc, s1, true, act));
        CALL(pc_sweep(c, c->pcc, s1, false, act));
      } else if (c->tiled) {
        CALL(pc_sweep_tiled(c, s1, true, act));
        CALL(pc_sweep_tiled(c, s1, false, act));
      } else {
        CALL(pc_sweep(c, c->pcl, s1, true, act));
        CALL(pc_sweep(c, c->pcl, s1, false, act));
      }
    }
    const int has_res = ((2 * blk) % 10 == 0) ? 1 : 0;
    if (!has_res && c->nx > 0) {  // exchange and stop test in one launch
      if (mixed)
        hipLaunchKernelGGL(k_exchange_end<float>, dim3(nblocks(c->nx)), dim3(kBlock), 0, c->st,
                           c->nx, c->xdst, xsrc, c->d32, c->pctl, c->red, sc, c->nit);
      else
        hipLaunchKernelGGL(k_exchange_end<double>, dim3(nblocks(c->nx)), dim3(kBlock), 0, c->st,
                           c->nx, c->xdst, xsrc, c->phic_p, c->pctl, c->red, sc, c->nit);
      continue;
    }
    if (c->nx > 0 && mixed)
      hipLaunchKernelGGL(k_exchange<float>, dim3(nblocks(c->nx)), dim3(kBlock), 0, c->st, c->nx,
                         c->xdst, xsrc, c->d32, act);
    else if (c->nx > 0)
      hipLaunchKernelGGL(k_exchange<double>, dim3(nblocks(c->nx)), dim3(kBlock), 0, c->st, c->nx,
                         c->xdst, xsrc, c->phic_p, act);
    if (has_res && mixed)
      hipLaunchKernelGGL(k_mixed_update, dim3(nblocks(nall)), dim3(kBlock), 0, c->st, nall,
                         c->phic_p, c->d32, act);
    if (has_res) CALL(residual(c, s, c->nblk_p, c->chunk_p, 0, mixed ? c->r32 : nullptr));
    hipLaunchKernelGGL(k_pc_block_end, dim3(1), dim3(64), 0, c->st, c->pctl, c->red, sc, c->nit,
                       has_res);
  }
  // the correction left since the last check (a cap of nit iterations that
  // is not a multiple of ten), added whether or not the solve is still active
  if (mixed)
    hipLaunchKernelGGL(k_mixed_update, dim3(nblocks(nall)), dim3(kBlock), 0, c->st, nall,
                       c->phic_p, c->d32, static_cast<const int *>(nullptr));
  // update_phi, mod_subdomains.f90:214-229
  hipLaunchKernelGGL(k_scatter, dim3(nblocks(ne)), dim3(kBlock), 0, c->st, ne, l2g,
                     c->phic_p, c->fd.phic);
  hipLaunchKernelGGL(k_log_pc, dim3(1), dim3(64), 0, c->st, c->pctl, c->dlog, c->dlog_n);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

// Dense ids of spatial tiles of about edge[0] x edge[1] x edge[2] cells
// (x, y, z): bins of the cell centres in the bounding box of each group of
// positions [grp[s], grp[s+1]) (the subdomains of the pc system, or all
// cells); tiles never cross groups.  cell[q] is the mesh cell at position q.
static void bin_tiles(const std::vector<int> &grp, const std::vector<int> &cell, const double *xc,
                      const double *yc, const double *zc, const double *vol, const int *edge,
                      std::vector<int> *tid) {
  std::unordered_map<int64_t, int> tid_of;
  tid->assign(cell.size(), 0);
  for (size_t s = 0; s + 1 < grp.size(); s++) {
    double lo[3] = {1e300, 1e300, 1e300}, vsum = 0.0;
    for (int q = grp[s]; q < grp[s + 1]; q++) {
      const int g = cell[q];
      const double r[3] = {xc[g], yc[g], zc[g]};
      for (int i = 0; i < 3; i++) lo[i] = std::min(lo[i], r[i]);
      vsum += vol[g];
    }
    const double h = std::cbrt(vsum / (grp[s + 1] - grp[s]));
    const double hx = h * edge[0], hy = h * edge[1], hz = h * edge[2];
    for (int q = grp[s]; q < grp[s + 1]; q++) {
      const int g = cell[q];
      const int64_t bx = static_cast<int64_t>((xc[g] - lo[0]) / hx);
      const int64_t by = static_cast<int64_t>((yc[g] - lo[1]) / hy);
      const int64_t bz = static_cast<int64_t>((zc[g] - lo[2]) / hz);
      const int64_t key = ((static_cast<int64_t>(s) * 4096 + bx) * 4096 + by) * 4096 + bz;
      auto it = tid_of.find(key);
      if (it == tid_of.end()) it = tid_of.emplace(key, static_cast<int>(tid_of.size())).first;
      (*tid)[q] = it->second;
    }
  }
}

// Space-time tiles of the pressure-correction sweeps, on the pc system in
// its set-up order (cells 0..ne-1, halo slots >= ne).  lev is each cell's
// forward dependency level inside its subdomain and tid its spatial tile
// (dense ids; a tile never crosses subdomains).  In the forward sweep's
// processing order (increasing level) each cell gets launch = max over the
// neighbours it depends on of (their launch, +1 when the neighbour lies in
// another tile).  So a dependency between two tiles always spans launches,
// and inside one launch a tile's cells only depend on earlier cells of the
// same tile.  A (launch, tile) batch spans fewer than `span` levels: a cell
// that would stretch its batch further moves to the tile's batch in a later
// launch.
static void tile_launches(int ne, const std::vector<int> &row, const std::vector<int> &nbr,
                          const std::vector<int> &lev, const std::vector<int> &tid, int span,
                          std::vector<int> *launch) {
  std::vector<int> order(ne);
  for (int e = 0; e < ne; e++) order[e] = e;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lev[a] < lev[b]; });
  int64_t ntile = 1;
  for (int e = 0; e < ne; e++) ntile = std::max<int64_t>(ntile, tid[e] + 1);
  std::unordered_map<int64_t, int> first_level;  // (launch, tile) -> its first level
  launch->assign(ne, 0);
  for (int e : order) {
    int L = 0;
    for (int k = row[e]; k < row[e + 1]; k++) {
      const int d = nbr[k];
      if (d >= ne || lev[d] >= lev[e]) continue;  // halo, or visited after e
      L = std::max(L, (*launch)[d] + (tid[d] != tid[e] ? 1 : 0));
    }
    for (;;) {
      const int64_t key = static_cast<int64_t>(L) * ntile + tid[e];
      auto it = first_level.find(key);
      if (it == first_level.end()) {
        first_level.emplace(key, lev[e]);
        break;
      }
      if (lev[e] - it->second < span) break;
      L++;
    }
    (*launch)[e] = L;
  }
}

// Batch tables of the tiled sweeps on the renumbered pc system, whose cells
// are ordered by (subdomain, launch, tile, level): each batch is the run of
// cells with equal (subdomain, launch, tile), each segment the run of equal
// level inside it.  Batches are listed launch by launch.  Per batch the LDS
// kernel also needs the distinct outside neighbours it reads (xslot) and,
// per CSR entry, the neighbour's LDS slot (lnbr).
static int build_tile_tables(cfdl_ctx *c, const std::vector<int> &rowp,
                             const std::vector<int> &nbrp, const std::vector<int> &sd,
                             const std::vector<int> &launch, const std::vector<int> &tid,
                             const std::vector<int> &lev, int nrhs, cfdl_ctx::TileSched *tp) {
  const int ne = c->ne;
  cfdl_ctx::TileSched &t = *tp;
  struct Batch {
    int launch, c0, c1, s0, s1;  // cells [c0, c1), level-segment starts sb[s0..s1)
  };
  std::vector<Batch> bl;
  std::vector<int> sb;  // segment starts, batch by batch in cell order
  for (int e = 0; e < ne; e++) {
    const bool nb = e == 0 || sd[e] != sd[e - 1] || launch[e] != launch[e - 1] ||
                    tid[e] != tid[e - 1];
    if (nb) {
      if (!bl.empty()) {
        bl.back().c1 = e;
        bl.back().s1 = static_cast<int>(sb.size());
      }
      bl.push_back(Batch{launch[e], e, e, static_cast<int>(sb.size()), 0});
    }
    if (nb || lev[e] != lev[e - 1]) sb.push_back(e);
  }
  bl.back().c1 = ne;
  bl.back().s1 = static_cast<int>(sb.size());
  std::stable_sort(bl.begin(), bl.end(),
                   [](const Batch &x, const Batch &y) { return x.launch < y.launch; });
  const int nl = bl.back().launch + 1;
  t.nbatch = static_cast<int>(bl.size());
  t.launch_ptr.assign(nl + 1, 0);
  t.launch_cells.assign(nl, 0);
  std::vector<int> bseg(t.nbatch + 1, 0), seg, xptr(t.nbatch + 1, 0), xslot;
  std::vector<unsigned short> lnbr(rowp[ne]);
  std::unordered_map<int, int> ext;
  size_t lds_max = 0, lds_pipe_max = 0;
  int nc_max = 0, nx_max = 0;
  int seg_max = 0;
  bool addr16 = true;
  bool fits = true, hex = true;
  for (int e = 0; e < ne; e++) hex = hex && rowp[e + 1] - rowp[e] == 6;
  for (int i = 0; i < t.nbatch; i++) {
    const Batch &B = bl[i];
    t.launch_ptr[B.launch + 1]++;
    t.launch_cells[B.launch] += B.c1 - B.c0;
    bseg[i] = static_cast<int>(seg.size());
    for (int q = B.s0; q < B.s1; q++) {
      seg.push_back(sb[q]);
      seg_max = std::max(seg_max, (q + 1 < B.s1 ? sb[q + 1] : B.c1) - sb[q]);
    }
    seg.push_back(B.c1);
    const int nc = B.c1 - B.c0, nq = B.s1 - B.s0;
    if (nq > 63) fits = false;  // the LDS kernel loads at most 64 bounds
    ext.clear();
    for (int k = rowp[B.c0]; k < rowp[B.c1]; k++) {
      const int d = nbrp[k];
      int slot;
      if (d >= B.c0 && d < B.c1) {
        slot = d - B.c0;
      } else {
        auto it = ext.find(d);
        if (it == ext.end()) {
          slot = nc + static_cast<int>(ext.size());
          ext.emplace(d, slot);
          xslot.push_back(d);
        } else {
          slot = it->second;
        }
      }
      if (slot > 65535) fits = false;
      lnbr[k] = static_cast<unsigned short>(slot & 0xffff);
    }
    xptr[i + 1] = static_cast<int>(xslot.size());
    const size_t nent = rowp[B.c1] - rowp[B.c0];
    // the two LDS layouts of k_gs_tiles_lds: plain walk, pipelined walk
    const size_t np = static_cast<size_t>(nc) + ext.size();
    const size_t bytes = 8 * ((2 * nrhs + 2) * static_cast<size_t>(nc) + nrhs * ext.size() + nent) +
                         4 * (nq + 1 + (hex ? 0 : nc + 1)) + 2 * nent;
    const size_t r16 = ~static_cast<size_t>(15);
    const size_t bytes_pipe =
        ((((8 * nrhs * np + 15) & r16) + 8 * (2 + nrhs + 6) * static_cast<size_t>(nc) + 15) & r16) +
        16 * static_cast<size_t>(nc) + 32;
    // the pipelined walk addresses values with 16-bit LDS byte addresses
    if (8 * nrhs * np > 0xffff) addr16 = false;
    lds_max = std::max(lds_max, bytes);
    lds_pipe_max = std::max(lds_pipe_max, bytes_pipe);
    nc_max = std::max(nc_max, nc);
    nx_max = std::max(nx_max, static_cast<int>(ext.size()));
  }
  bseg[t.nbatch] = static_cast<int>(seg.size());
  for (int m = 0; m < nl; m++) t.launch_ptr[m + 1] += t.launch_ptr[m];
  t.nseg = static_cast<int>(seg.size()) - t.nbatch;
  if (xslot.empty()) xslot.push_back(0);
  const char *pp = std::getenv("CFDL_TILE_PIPE");
  t.pipe = hex && addr16 && seg_max <= kLdsThreads && !(pp && pp[0] == '0');
  // the pipelined walk reads its neighbours' values at these LDS addresses
  if (t.pipe) {
    if (nrhs != 1) {
      std::vector<unsigned short> a1(lnbr.size());
      for (size_t k = 0; k < lnbr.size(); k++)
        a1[k] = static_cast<unsigned short>(8u * lnbr[k]);
      CALL(upload(c, &t.lnbr1, a1));
    }
    for (auto &v : lnbr) v = static_cast<unsigned short>(8u * nrhs * v);
  }
  const char *tw = std::getenv("CFDL_TILE_WAVES");
  t.nw = (tw && tw[0] == '1') ? 1 : 4;
  const char *pf = std::getenv("CFDL_TILE_PREFETCH");
  t.prefetch = t.nw > 1 && !(pf && pf[0] == '0');
  if (t.pipe) lds_max = lds_pipe_max;
  if (trace_on()) {
    std::fprintf(stderr, "cfdl: tiled sweep: %d launches, %d batches, %d segments (largest "
                 "%d cells), largest batch LDS %zu bytes, at most %d cells and %d outside "
                 "neighbours per batch\n", nl, t.nbatch, t.nseg, seg_max, lds_max, nc_max,
                 nx_max);
  }
  const char *no_lds = std::getenv("CFDL_TILE_LDS");
  const bool lds_ok = !(no_lds && no_lds[0] == '0');
  // a workgroup may use the device's whole per-block LDS (160 KB on gfx950)
  // once the kernel's dynamic-LDS limit is raised to it
  int lds_dev = 0;
  HIPCHK(hipDeviceGetAttribute(&lds_dev, hipDeviceAttributeMaxSharedMemoryPerBlock, c->dev));
  size_t lds_cap = 64 * 1024;
  if (static_cast<size_t>(lds_dev) > lds_cap && lds_max > lds_cap) {
    const void *kern[] = {reinterpret_cast<const void *>(&k_gs_tiles_lds<true, true, 1, 1>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<true, false, 1, 1>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<false, false, 1, 1>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<true, true, 3, 1>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<true, false, 3, 1>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<false, false, 3, 1>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<true, true, 1, 4>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<true, false, 1, 4>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<false, false, 1, 4>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<true, true, 3, 4>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<true, false, 3, 4>),
                          reinterpret_cast<const void *>(&k_gs_tiles_lds<false, false, 3, 4>)};
    bool ok = true;
    for (const void *k : kern)
      ok = ok && hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lds_dev) ==
                     hipSuccess;
    if (ok) lds_cap = static_cast<size_t>(lds_dev);
  }
  (void)hipGetLastError();
  t.lds_bytes = (lds_ok && fits && lds_max <= lds_cap) ? lds_max : 0;
  t.hex = hex;
  std::vector<int4> bdesc(2 * static_cast<size_t>(t.nbatch));
  for (int i = 0; i < t.nbatch; i++) {
    const int s0 = bseg[i], nq = bseg[i + 1] - s0 - 1, c0 = seg[s0], nc = seg[s0 + nq] - c0;
    bdesc[2 * i] = make_int4(s0, nq, c0, nc);
    bdesc[2 * i + 1] = make_int4(rowp[c0], rowp[c0 + nc] - rowp[c0], xptr[i], xptr[i + 1] - xptr[i]);
  }
  CALL(upload(c, &t.bdesc, bdesc));
  CALL(upload(c, &t.bseg, bseg));
  CALL(upload(c, &t.seg, seg));
  CALL(upload(c, &t.xptr, xptr));
  CALL(upload(c, &t.xslot, xslot));
  CALL(upload(c, &t.lnbr, lnbr));
  return CFDL_OK;
}

static double *field_ptr(cfdl_ctx *c, int f) {
  if (f < 0 || f >= CFDL_NFIELDS) return nullptr;
  return c->fields[f];
}

static int check_ctx(cfdl_ctx *c) {
  if (!c) {
    set_error("null context");
    return CFDL_EINVAL;
  }
  return CFDL_OK;
}

// Device scratch of cfdl_calc_residual_noalloc: the block partials (2 per
// block, at most 2048 blocks) and the two reduced values.  Each GPU holds
// its own copy until the process ends; host code uses the address that
// hipGetSymbolAddress gives for the current device, never the array's name.
// Calls that overlap on one GPU share it and race.
constexpr int kResidualMaxBlocks = 2048;
static __device__ __align__(256) double scratch[2 * kResidualMaxBlocks + 2];

static int resid_scratch(double **out) {
  static std::mutex mtx;
  static std::vector<double *> addr;  // per device id, resolved once
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mtx);
  if (static_cast<int>(addr.size()) <= dev) addr.resize(dev + 1, nullptr);
  if (!addr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    addr[dev] = static_cast<double *>(p);
  }
  *out = addr[dev];
  return CFDL_OK;
}

extern "C" {

const char *cfdl_last_error(void) { return g_err.c_str(); }
int cfdl_version(void) { return 100; }

// the inlet and outlet faces (0-based, halo order) of the channel extension
static int channel_lists(cfdl_ctx *c) {
  std::vector<int> fin, fout;
  for (int h = 0; h < c->nbf; h++) {
    const int fg = std::abs(c->fsg_h[c->hslot_h[h]]) - 1;
    if (c->hkind_h[h] == CFDL_BC_INLET) fin.push_back(fg);
    if (c->hkind_h[h] == CFDL_BC_OUTLET) fout.push_back(fg);
  }
  c->nin = static_cast<int>(fin.size());
  c->nout = static_cast<int>(fout.size());
  if (!fin.empty())
    HIPCHK(hipMemcpy(c->fin, fin.data(), sizeof(int) * fin.size(), hipMemcpyHostToDevice));
  if (!fout.empty())
    HIPCHK(hipMemcpy(c->fout, fout.data(), sizeof(int) * fout.size(), hipMemcpyHostToDevice));
  return CFDL_OK;
}

int cfdl_ctx_create(int ne, int nf, int nbf, const int *ef2nb_idx, const int *ef2nb,
                    const int *s2g, const int *bs, const double *xc, const double *yc,
                    const double *zc, const double *vol, const double *aip, const double *rip,
                    int nsec, const int *esec, const int *bckind, int nsubd, const int *gf2g,
                    const int *g2gf, const int *g2gf_idx, const double *rho, const double *mu,
                    double dt, int nit, cfdl_ctx **out) {
  if (!out || ne <= 0 || nf <= 0 || nbf < 0 || !ef2nb_idx || !ef2nb || !s2g || !xc || !yc ||
      !zc || !vol || !aip || !rip || !rho || !mu || nsec <= 0 || !esec || !bckind ||
      nsubd < 1 || nsubd > kMaxSys || nit < 0 || !(dt > 0.0) || (nbf > 0 && !bs) ||
      (nsubd > 1 && (!gf2g || !g2gf || !g2gf_idx))) {
    set_error("cfdl_ctx_create: bad arguments");
    return CFDL_EINVAL;
  }
  trace("ctx_create: begin");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("cfdl_ctx_create: no HIP device visible");
    return CFDL_ENODEV;
  }
  std::unique_ptr<cfdl_ctx> c(new cfdl_ctx());
  HIPCHK(hipGetDevice(&c->dev));
  trace("ctx_create: device");
  HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  trace("ctx_create: stream");
  c->ne = ne;
  c->nf = nf;
  c->nbf = nbf;
  c->L = 2 * nf - nbf;
  c->nsubd = nsubd;
  c->nit = nit;
  c->dt = dt;
  const int L = c->L, n = ne + nbf;
  if (ef2nb_idx[0] != 1 || ef2nb_idx[ne] != L + 1) {
    set_error("ef2nb_idx does not describe 2nf-nbf entries");
    return CFDL_EINVAL;
  }
  // --- CSR, faces, halos (0-based) ---
  std::vector<int> row(ne + 1), nbr(L), fsg(L), own(nf, -1), fcell(nf, -1);
  for (int e = 0; e <= ne; e++) row[e] = ef2nb_idx[e] - 1;
  for (int k = 0; k < L; k++) {
    const int enb = pk_elem(ef2nb[k]) - 1, lf = pk_face(ef2nb[k]);
    if (enb < 0 || enb >= n || (lf == 0) != (enb >= ne)) {
      set_error("ef2nb entry " + std::to_string(k + 1) + " is inconsistent");
      return CFDL_EINVAL;
    }
    nbr[k] = enb;
    fsg[k] = ef2nb[L + k];
    if (fsg[k] == 0 || std::abs(fsg[k]) > nf) {
      set_error("ef2nb(:,2) entry " + std::to_string(k + 1) + " is not a face");
      return CFDL_EINVAL;
    }
  }
  for (int f = 0; f < nf; f++) {
    const int e = pk_elem(s2g[f]) - 1, lf = pk_face(s2g[f]);
    if (e < 0 || e >= ne || lf < 1 || row[e] + lf - 1 >= row[e + 1]) {
      set_error("s2g entry " + std::to_string(f + 1) + " is inconsistent");
      return CFDL_EINVAL;
    }
    own[f] = row[e] + lf - 1;
    fcell[f] = e;
  }
  std::vector<int> hcell(nbf), hslot(nbf), hkind(nbf, CFDL_BC_WALL), bcrow(ne + 1, 0), bcslot(nbf);
  for (int h = 0; h < nbf; h++) {
    const int p = std::abs(bs[h]);
    const int e = pk_elem(p) - 1, lf = pk_face(p);
    if (e < 0 || e >= ne || lf < 1 || row[e] + lf - 1 >= row[e + 1] ||
        nbr[row[e] + lf - 1] != ne + h) {
      set_error("bs entry " + std::to_string(h + 1) + " is inconsistent");
      return CFDL_EINVAL;
    }
    hcell[h] = e;
    hslot[h] = row[e] + lf - 1;
    bcrow[e + 1]++;
  }
  for (int s = 0; s < nsec; s++)
    for (int el = esec[2 * s]; el <= esec[2 * s + 1]; el++)
      if (el > ne && el <= n) {
        if (bckind[s] < CFDL_BC_WALL || bckind[s] > CFDL_BC_OUTLET) {
          set_error("cfdl_ctx_create: bckind of section " + std::to_string(s + 1) +
                    " is not a CFDL_BC_* kind");
          return CFDL_EINVAL;
        }
        hkind[el - ne - 1] = bckind[s];
      }
  for (int e = 0; e < ne; e++) bcrow[e + 1] += bcrow[e];
  {
    std::vector<int> pos(bcrow.begin(), bcrow.end() - 1);
    for (int h = 0; h < nbf; h++) bcslot[pos[hcell[h]]++] = hslot[h];  // halo order per cell
  }
  trace("ctx_create: host tables");
  cfdl_ctx *C = c.get();
  int *d_row, *d_nbr, *d_fsg, *d_own, *d_fcell, *d_bcrow, *d_bcslot, *d_hcell, *d_hslot, *d_hkind;
  CALL(upload(C, &d_row, row));
  CALL(upload(C, &d_nbr, nbr));
  CALL(upload(C, &d_fsg, fsg));
  CALL(upload(C, &d_own, own));
  CALL(upload(C, &d_fcell, fcell));
  CALL(upload(C, &d_bcrow, bcrow));
  CALL(upload(C, &d_bcslot, bcslot));
  CALL(upload(C, &d_hcell, hcell));
  CALL(upload(C, &d_hslot, hslot));
  CALL(upload(C, &d_hkind, hkind));
  C->hkind_h = hkind;
  C->hslot_h = hslot;
  C->fsg_h = fsg;
  CALL(dalloc(C, &C->fin, nbf));
  CALL(dalloc(C, &C->fout, nbf));
  CALL(channel_lists(C));
  auto up64 = [&](double **p, const double *src, int64_t cnt) -> int {
    CALL(dalloc(C, p, cnt));
    HIPCHK(hipMemcpy(*p, src, 8 * cnt, hipMemcpyHostToDevice));
    return CFDL_OK;
  };
  double *d_xc, *d_yc, *d_zc, *d_vol, *d_aip, *d_rip, *d_rho, *d_mu, *d_cp, *d_tc;
  CALL(up64(&d_xc, xc, n));
  CALL(up64(&d_yc, yc, n));
  CALL(up64(&d_zc, zc, n));
  CALL(up64(&d_vol, vol, ne));
  CALL(up64(&d_aip, aip, 3 * int64_t{nf}));
  CALL(up64(&d_rip, rip, 3 * int64_t{nf}));
  CALL(up64(&d_rho, rho, ne));
  CALL(up64(&d_mu, mu, ne));
  // thermal properties as init_properties sets them, mod_properties.f90:88-89
  CALL(up64(&d_cp, std::vector<double>(ne, 1000.0).data(), ne));
  CALL(up64(&d_tc, std::vector<double>(ne, 5.0).data(), ne));
  // CFDL_COEF_SERIAL=1 keeps the one-thread-per-cell assembly kernels, which
  // also serve meshes with other row lengths
  const char *cs = std::getenv("CFDL_COEF_SERIAL");
  C->mhex = row[ne] == 6 * ne && !(cs && cs[0] == '1');
  for (int e = 0; e < ne && C->mhex; e++) C->mhex = row[e] == 6 * e;
  C->md = MeshDev{ne, nf, nbf, L, d_row, d_nbr, d_fsg, d_own, d_fcell, d_bcrow, d_bcslot,
                  d_hcell, d_hslot, d_hkind, d_xc, d_yc, d_zc, d_vol, d_aip, d_rip, d_rho, d_mu,
                  d_cp, d_tc};
  trace("ctx_create: mesh uploaded");
  // --- fields ---
  const int64_t sz[CFDL_NFIELDS] = {n, n, n, n, n, n, n, 3 * int64_t{n}, 3 * int64_t{n},
                                    3 * int64_t{n}, 3 * int64_t{n}, 3 * int64_t{n}, nf, nf, ne,
                                    ne, ne, ne, ne, ne, ne, L, n, n, 3 * int64_t{n},
                                    n, 3 * int64_t{n}, n, n, 3 * int64_t{n}};
  for (int f = 0; f < CFDL_NFIELDS; f++) {
    C->fsize[f] = sz[f];
    CALL(dalloc(C, &C->fields[f], sz[f]));
    hipLaunchKernelGGL(k_fill, dim3(nblocks(sz[f])), dim3(kBlock), 0, C->st, C->fields[f], sz[f], 0.0);
  }
  double **F = C->fields;
  C->fd = FieldsDev{F[CFDL_U], F[CFDL_V], F[CFDL_W], F[CFDL_P], F[CFDL_U0], F[CFDL_V0],
                    F[CFDL_W0], F[CFDL_GU], F[CFDL_GV], F[CFDL_GW], F[CFDL_GP], F[CFDL_GPC],
                    F[CFDL_MIP], F[CFDL_MIP0], F[CFDL_BU], F[CFDL_BV], F[CFDL_BW], F[CFDL_D],
                    F[CFDL_DC], F[CFDL_AP], F[CFDL_B], F[CFDL_ANB], F[CFDL_PHIC]};
  C->en = EnergyDev{F[CFDL_T], F[CFDL_GT], F[CFDL_ENTH], F[CFDL_ENTH0], F[CFDL_GENTH]};
  hipLaunchKernelGGL(k_energy_init, dim3(nblocks(n)), dim3(kBlock), 0, C->st, C->md, C->en);
  trace("ctx_create: fields");
  // sweep schedules: spatial tiles of edge[0] x edge[1] x edge[2] cells
  // (CFDL_TILE = "A" for A^3 or "AxBxC"), batches spanning fewer than `span`
  // levels; CFDL_GS_MODE=level keeps one launch per level.  16 x 8 x 4 cells
  // and 32 levels measured fastest on the 128^3 cavity (fewest launches
  // whose batches keep level segments within 64 cells and LDS within half
  // a CU's, so two batches fit per CU).
  int edge[3] = {16, 8, 4};
  if (const char *te = std::getenv("CFDL_TILE")) {
    int a = 0, b = 0, d = 0;
    const int n = std::sscanf(te, "%dx%dx%d", &a, &b, &d);
    if (n == 3) {
      edge[0] = a, edge[1] = b, edge[2] = d;
    } else if (n >= 1) {
      edge[0] = edge[1] = edge[2] = a;
    }
    for (int i = 0; i < 3; i++) edge[i] = std::max(1, edge[i]);
  }
  const char *sp = std::getenv("CFDL_TILE_SPAN");
  const int span = sp ? std::min(63, std::max(1, std::atoi(sp))) : 32;
  const char *mode = std::getenv("CFDL_GS_MODE");
  C->tiled = !(mode && std::string(mode) == "level");
  const char *rc = std::getenv("CFDL_TILE_RCP");
  C->rcp = !(rc && rc[0] == '0');
  // --- global Gauss-Seidel schedules ---
  {
    std::vector<int> lev(ne);
    levels(row, nbr, 0, ne, true, lev);
    CALL(make_sched(C, lev, ne, &C->gf));
    if (C->tiled) {
      // the global system in tile numbering, for the tiled solve_gs
      std::vector<int> all(ne), grp{0, ne}, tid, launch;
      for (int e = 0; e < ne; e++) all[e] = e;
      bin_tiles(grp, all, xc, yc, zc, vol, edge, &tid);
      tile_launches(ne, row, nbr, lev, tid, span, &launch);
      std::vector<int> n2o(all), o2n(ne);
      std::stable_sort(n2o.begin(), n2o.end(), [&](int x, int y) {
        if (launch[x] != launch[y]) return launch[x] < launch[y];
        if (tid[x] != tid[y]) return tid[x] < tid[y];
        return lev[x] < lev[y];
      });
      for (int i = 0; i < ne; i++) o2n[n2o[i]] = i;
      std::vector<int> rowm(ne + 1, 0), nbrm, zero(ne, 0), launchm(ne), tidm(ne), levm(ne);
      nbrm.reserve(L);
      for (int i = 0; i < ne; i++) {
        const int g = n2o[i];
        launchm[i] = launch[g];
        tidm[i] = tid[g];
        levm[i] = lev[g];
        rowm[i + 1] = rowm[i] + (row[g + 1] - row[g]);
        for (int k = row[g]; k < row[g + 1]; k++) nbrm.push_back(nbr[k] < ne ? o2n[nbr[k]] : nbr[k]);
      }
      CALL(build_tile_tables(C, rowm, nbrm, zero, launchm, tidm, levm, 3, &C->tm));
      C->mtiled = C->tm.lds_bytes > 0;
      if (C->mtiled) {
        CALL(upload(C, &C->l2g_m, n2o));
        CALL(upload(C, &C->rowm, rowm));
        CALL(upload(C, &C->nbrm, nbrm));
        CALL(dalloc(C, &C->ap_m, ne));
        CALL(dalloc(C, &C->rap_m, ne));
        CALL(dalloc(C, &C->anb_m, L));
        for (int y = 0; y < 3; y++) {
          CALL(dalloc(C, &C->b_m[y], ne));
          CALL(dalloc(C, &C->phi_m[y], int64_t{ne} + nbf));
        }
      }
    }
    levels(row, nbr, 0, ne, false, lev);
    CALL(make_sched(C, lev, ne, &C->gb));
    // colour order of solve_gs: greedy colouring in global order (a cell
    // takes the least colour none of its earlier interior neighbours has),
    // cells renumbered by (colour, global order)
    {
      std::vector<int> col(ne, 0);
      int ncol = 1;
      for (int e = 0; e < ne; e++) {
        uint64_t used = 0;
        for (int k = row[e]; k < row[e + 1]; k++) {
          const int d = nbr[k];
          if (d < e && col[d] < 64) used |= uint64_t{1} << col[d];
        }
        int q = 0;
        while (q < 64 && ((used >> q) & 1)) q++;
        col[e] = q;
        ncol = std::max(ncol, q + 1);
      }
      std::vector<int> n2o(ne), o2n(ne), rowc(ne + 1, 0), nbrc;
      for (int e = 0; e < ne; e++) n2o[e] = e;
      std::stable_sort(n2o.begin(), n2o.end(), [&](int x, int y) { return col[x] < col[y]; });
      for (int i = 0; i < ne; i++) o2n[n2o[i]] = i;
      C->gcol_ptr.assign(ncol + 1, 0);
      for (int e = 0; e < ne; e++) C->gcol_ptr[col[e] + 1]++;
      for (int q = 0; q < ncol; q++) C->gcol_ptr[q + 1] += C->gcol_ptr[q];
      nbrc.reserve(L);
      for (int i = 0; i < ne; i++) {
        const int g = n2o[i];
        rowc[i + 1] = rowc[i] + (row[g + 1] - row[g]);
        for (int k = row[g]; k < row[g + 1]; k++) nbrc.push_back(nbr[k] < ne ? o2n[nbr[k]] : nbr[k]);
      }
      CALL(upload(C, &C->l2g_gc, n2o));
      CALL(upload(C, &C->rowgc, rowc));
      CALL(upload(C, &C->nbrgc, nbrc));
    }
  }
  // --- subdomain systems (construct_subdomains, mod_subdomains.f90:18-160):
  // local cells in g2gf order, CSR rows in the global row order, halo slots
  // numbered in encounter order, each interface face a copy pair ---
  trace("ctx_create: global schedules");
  if (nsubd > 1) {
    C->co.assign(nsubd + 1, 0);
    std::vector<int> cid_of(ne, -1), sd_of(ne, -1), l2g(ne);
    for (int s = 0; s < nsubd; s++) {
      C->co[s] = g2gf_idx[s] - 1;
      if (g2gf_idx[s + 1] <= g2gf_idx[s]) {
        set_error("g2gf_idx describes an empty subdomain");
        return CFDL_EINVAL;
      }
    }
    C->co[nsubd] = g2gf_idx[nsubd] - 1;
    if (C->co[0] != 0 || C->co[nsubd] != ne) {
      set_error("g2gf_idx does not cover the cells");
      return CFDL_EINVAL;
    }
    for (int s = 0; s < nsubd; s++)
      for (int q = C->co[s]; q < C->co[s + 1]; q++) {
        const int g = g2gf[q] - 1;
        if (g < 0 || g >= ne || cid_of[g] >= 0 || gf2g[g] != s + 1) {
          set_error("g2gf / gf2g are not a consistent partition");
          return CFDL_EINVAL;
        }
        cid_of[g] = q;
        sd_of[g] = s;
        l2g[q] = g;
      }
    // subdomain systems in g2gf order first (old index q = position in
    // g2gf), halo slots numbered in encounter order; then each subdomain's
    // cells are renumbered by forward dependency level
    std::vector<int> rowo(ne + 1, 0), nbro, halo_src, xdst, xsrc;
    nbro.reserve(L);
    for (int q = 0; q < ne; q++) rowo[q + 1] = rowo[q] + (row[l2g[q] + 1] - row[l2g[q]]);
    int nh = 0;
    for (int q = 0; q < ne; q++) {
      const int g = l2g[q], s = sd_of[g];
      for (int k = row[g]; k < row[g + 1]; k++) {
        const int nb = nbr[k];
        if (nb < ne && sd_of[nb] == s) {
          nbro.push_back(cid_of[nb]);
        } else {
          const int slot = ne + nh++;
          nbro.push_back(slot);
          halo_src.push_back(nb);
          if (nb < ne) {
            xdst.push_back(slot);
            xsrc.push_back(cid_of[nb]);
          }
        }
      }
    }
    // forward dependency level of every cell inside its subdomain
    std::vector<int> lev(ne);
    C->pc_nlev = 0;
    for (int s = 0; s < nsubd; s++) levels(rowo, nbro, C->co[s], C->co[s + 1], true, lev);
    for (int q = 0; q < ne; q++) C->pc_nlev = std::max(C->pc_nlev, lev[q]);
    std::vector<int> tid;
    bin_tiles(C->co, l2g, xc, yc, zc, vol, edge, &tid);
    std::vector<int> launch;
    tile_launches(ne, rowo, nbro, lev, tid, span, &launch);
    // renumber the pc system by (subdomain, launch, tile, level): every
    // (launch, tile) batch of the tiled sweeps becomes a contiguous range
    std::vector<int> n2o(ne), o2n(ne);
    for (int q = 0; q < ne; q++) n2o[q] = q;
    std::stable_sort(n2o.begin(), n2o.end(), [&](int x, int y) {
      if (sd_of[l2g[x]] != sd_of[l2g[y]]) return sd_of[l2g[x]] < sd_of[l2g[y]];
      if (launch[x] != launch[y]) return launch[x] < launch[y];
      if (tid[x] != tid[y]) return tid[x] < tid[y];
      return lev[x] < lev[y];
    });
    for (int nq = 0; nq < ne; nq++) o2n[n2o[nq]] = nq;
    std::vector<int> rowp(ne + 1, 0), nbrp, l2gn(ne), sdn(ne), launchn(ne), tidn(ne), levn(ne);
    nbrp.reserve(nbro.size());
    for (int nq = 0; nq < ne; nq++) {
      const int q = n2o[nq];
      l2gn[nq] = l2g[q];
      sdn[nq] = sd_of[l2g[q]];
      launchn[nq] = launch[q];
      tidn[nq] = tid[q];
      levn[nq] = lev[q];
      rowp[nq + 1] = rowp[nq] + (rowo[q + 1] - rowo[q]);
      for (int k = rowo[q]; k < rowo[q + 1]; k++) {
        const int v = nbro[k];
        nbrp.push_back(v < ne ? o2n[v] : v);
      }
    }
    const std::vector<int> xsrc_old = xsrc;
    for (auto &v : xsrc) v = o2n[v];
    // colour order (extension): greedy colouring of each subdomain in global
    // cell order, a cell taking the least colour that none of its interior
    // neighbours earlier in that order has (two colours on a structured hex
    // box, where the subdomain-local quicksort order needs four); colour + 1
    // serves as the cell's level
    std::vector<int> colq(ne, 0), coln(ne), byg(ne);
    for (int q = 0; q < ne; q++) byg[q] = q;
    for (int s = 0; s < nsubd; s++)
      std::sort(byg.begin() + C->co[s], byg.begin() + C->co[s + 1],
                [&](int x, int y) { return l2g[x] < l2g[y]; });
    for (int s = 0; s < nsubd; s++)
      for (int i = C->co[s]; i < C->co[s + 1]; i++) {
        const int q = byg[i];
        uint64_t used = 0;
        for (int k = rowo[q]; k < rowo[q + 1]; k++) {
          const int d = nbro[k];
          if (d < ne && l2g[d] < l2g[q] && colq[d] < 64) used |= uint64_t{1} << colq[d];
        }
        int col = 0;
        while (col < 64 && ((used >> col) & 1)) col++;
        colq[q] = col;
      }
    // the colour numbering: (subdomain, colour, global cell index).  Cells
    // of one colour are independent, so their order inside the colour does
    // not change a bit.  The global (mesh) order keeps neighbouring threads'
    // neighbours closer in memory than the partition's quicksort order; on
    // the 128^3 cavity the sweeps measured about 1% faster with it.
    std::vector<int> c2o(ne), o2c(ne);
    for (int q = 0; q < ne; q++) c2o[q] = q;
    std::sort(c2o.begin(), c2o.end(), [&](int x, int y) {
      if (sd_of[l2g[x]] != sd_of[l2g[y]]) return sd_of[l2g[x]] < sd_of[l2g[y]];
      if (colq[x] != colq[y]) return colq[x] < colq[y];
      return l2g[x] < l2g[y];
    });
    for (int i = 0; i < ne; i++) o2c[c2o[i]] = i;
    std::vector<int> rowc(ne + 1, 0), nbrc, l2gc(ne), xsrcc(xsrc_old.size());
    nbrc.reserve(nbro.size());
    for (int i = 0; i < ne; i++) {
      const int q = c2o[i];
      l2gc[i] = l2g[q];
      coln[i] = colq[q] + 1;
      rowc[i + 1] = rowc[i] + (rowo[q + 1] - rowo[q]);
      for (int k = rowo[q]; k < rowo[q + 1]; k++)
        nbrc.push_back(nbro[k] < ne ? o2c[nbro[k]] : nbro[k]);
    }
    for (size_t j = 0; j < xsrc_old.size(); j++) xsrcc[j] = o2c[xsrc_old[j]];
    CALL(build_tile_tables(C, rowp, nbrp, sdn, launchn, tidn, levn, 1, &C->ts));
    CALL(make_sched(C, levn, ne, &C->pcl));
    CALL(make_sched(C, coln, ne, &C->pcc));
    // the same colours as ranges per subdomain (colour numbering is
    // subdomain-major, colour-minor)
    C->pcc_rng.assign(C->pcc.nlev, ColRanges{});
    for (int l = 0; l < C->pcc.nlev; l++) {
      ColRanges &cr = C->pcc_rng[l];
      cr.n = nsubd;
      cr.off[0] = 0;
      for (int y = 0; y < nsubd; y++) {
        int b = C->co[y + 1], e1 = C->co[y];
        for (int i = C->co[y]; i < C->co[y + 1]; i++)
          if (coln[i] == l + 1) b = std::min(b, i), e1 = i + 1;
        cr.beg[y] = b;
        cr.off[y + 1] = cr.off[y] + std::max(0, e1 - b);
      }
    }
    C->pc_hex = true;
    for (int i = 0; i < ne && C->pc_hex; i++) C->pc_hex = rowc[i] == 6 * i;
    C->pc_hex = C->pc_hex && rowc[ne] == 6 * ne;
    C->Lp = static_cast<int>(nbrp.size());
    C->nh = nh;
    C->nx = static_cast<int>(xdst.size());
    CALL(upload(C, &C->rowp, rowp));
    CALL(upload(C, &C->nbrp, nbrp));
    CALL(upload(C, &C->l2g_cell, l2gn));
    CALL(upload(C, &C->halo_src, halo_src));
    CALL(upload(C, &C->xdst, xdst));
    CALL(upload(C, &C->xsrc, xsrc));
    CALL(upload(C, &C->rowp_c, rowc));
    CALL(upload(C, &C->nbrp_c, nbrc));
    CALL(upload(C, &C->l2g_c, l2gc));
    CALL(upload(C, &C->xsrc_c, xsrcc));
    CALL(dalloc(C, &C->ap_p, ne));
    CALL(dalloc(C, &C->rap_p, ne));
    CALL(dalloc(C, &C->b_p, ne));
    CALL(dalloc(C, &C->anb_p, C->Lp));
    CALL(dalloc(C, &C->phic_p, int64_t{ne} + nh));
    int maxsd = 0;
    for (int s = 0; s < nsubd; s++) maxsd = std::max(maxsd, C->co[s + 1] - C->co[s]);
    C->chunk_p = std::max(kBlock * 4, (maxsd + 2047) / 2048);
    C->nblk_p = (maxsd + C->chunk_p - 1) / C->chunk_p;
  }
  trace("ctx_create: subdomains");
  C->chunk_g = std::max(kBlock * 4, (ne + 2047) / 2048);
  C->nblk_g = (ne + C->chunk_g - 1) / C->chunk_g;
  const int nblk = std::max(C->nblk_g * 3, C->nblk_p * nsubd);
  CALL(dalloc(C, &C->part, 2 * int64_t{nblk}));
  CALL(dalloc(C, &C->red, 2 * kMaxSys));
  CALL(dalloc(C, &C->scal, 8));
  CALL(dalloc(C, &C->gctl, 1));
  CALL(dalloc(C, &C->pctl, 1));
  CALL(dalloc(C, &C->dlog, kLogCap));
  CALL(dalloc(C, &C->dlog_n, 1));
  {
    const int zero = 0;
    HIPCHK(hipMemcpy(C->dlog_n, &zero, sizeof(int), hipMemcpyHostToDevice));
  }
  HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&C->host_flags), 64, hipHostMallocDefault));
  HIPCHK(hipStreamSynchronize(C->st));
  trace("ctx_create: done");
  *out = c.release();
  return CFDL_OK;
}

void cfdl_ctx_destroy(cfdl_ctx *c) {
  if (!c) return;
  if (c->batch) (void)cfdl_batch_detach(c->batch, c);
  (void)hipStreamSynchronize(c->st);
  for (auto &m : c->mg) cfdl::mgcg_destroy(m);
  if (c->host_flags) (void)hipHostFree(c->host_flags);
  for (auto &x : c->samples) {
    (void)hipEventDestroy(x.a);
    (void)hipEventDestroy(x.b);
  }
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

int cfdl_ctx_create_cavity(const cfdl_mesh *m, double rho, double mu, double dt, int nit,
                           cfdl_ctx **out) {
  if (!m || m->nsubd < 1) {
    set_error("cfdl_ctx_create_cavity: mesh is not partitioned (call cfdl_mesh_partition)");
    return CFDL_ESTATE;
  }
  // BCs bind by section name, as make_bc does (mod_eqn_setup.f90:46-67): a
  // mesh without a 'top' section has no lid, which the reference stops on
  std::vector<int> kind(m->nsec, CFDL_BC_WALL);
  bool lid = false;
  for (int s = 0; s < m->nsec; s++)
    if (s < static_cast<int>(m->secname.size()) && m->secname[s] == "top") {
      kind[s] = CFDL_BC_LID;
      lid = true;
    }
  if (!lid) {
    set_error("cfdl_ctx_create_cavity: BC top not found (name the sections with "
              "cfdl_mesh_set_section_names)");
    return CFDL_EINVAL;
  }
  std::vector<double> r(m->ne, rho), u(m->ne, mu);
  return cfdl_ctx_create(m->ne, m->nf, m->nbf, m->ef2nb_idx.data(), m->ef2nb.data(),
                         m->s2g.data(), m->bs.data(), m->xc.data(), m->yc.data(), m->zc.data(),
                         m->vol.data(), m->aip.data(), m->rip.data(), m->nsec, m->esec.data(),
                         kind.data(), m->nsubd, m->gf2g.data(), m->g2gf.data(),
                         m->g2gf_idx.data(), r.data(), u.data(), dt, nit, out);
}

int cfdl_ctx_create_channel(const cfdl_mesh *m, double rho, double mu, double dt, int nit,
                            double umax, cfdl_ctx **out) {
  if (!m || m->nsubd < 1) {
    set_error("cfdl_ctx_create_channel: mesh is not partitioned (call cfdl_mesh_partition)");
    return CFDL_ESTATE;
  }
  std::vector<int> kind(m->nsec, CFDL_BC_WALL);
  int found = 0;
  for (int s = 0; s < m->nsec && s < static_cast<int>(m->secname.size()); s++) {
    if (m->secname[s] == "west") kind[s] = CFDL_BC_INLET, found |= 1;
    if (m->secname[s] == "east") kind[s] = CFDL_BC_OUTLET, found |= 2;
  }
  if (found != 3) {
    set_error("cfdl_ctx_create_channel: sections 'west' (inlet) and 'east' (outlet) needed");
    return CFDL_EINVAL;
  }
  // the cross-section from the vertex bounding box
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int v = 0; v < m->nvx; v++) {
    const double p[3] = {m->x[v], m->y[v], m->z[v]};
    for (int i = 0; i < 3; i++) {
      lo[i] = std::min(lo[i], p[i]);
      hi[i] = std::max(hi[i], p[i]);
    }
  }
  std::vector<double> r(m->ne, rho), u(m->ne, mu);
  CALL(cfdl_ctx_create(m->ne, m->nf, m->nbf, m->ef2nb_idx.data(), m->ef2nb.data(), m->s2g.data(),
                       m->bs.data(), m->xc.data(), m->yc.data(), m->zc.data(), m->vol.data(),
                       m->aip.data(), m->rip.data(), m->nsec, m->esec.data(), kind.data(),
                       m->nsubd, m->gf2g.data(), m->g2gf.data(), m->g2gf_idx.data(), r.data(),
                       u.data(), dt, nit, out));
  return cfdl_set_inlet_profile(*out, umax, lo[1], hi[1], lo[2], hi[2]);
}

int cfdl_set_properties(cfdl_ctx *c, const double *rho, const double *mu) {
  CALL(check_ctx(c));
  if (!rho || !mu) {
    set_error("cfdl_set_properties: null array");
    return CFDL_EINVAL;
  }
  HIPCHK(hipMemcpyAsync(const_cast<double *>(c->md.rho), rho, 8 * size_t(c->ne), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(const_cast<double *>(c->md.mu), mu, 8 * size_t(c->ne), hipMemcpyHostToDevice, c->st));
  if (c->md.rho32) {
    hipLaunchKernelGGL(k_to_f32, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->md.rho,
                       const_cast<float *>(c->md.rho32), int64_t{c->ne});
    hipLaunchKernelGGL(k_to_f32, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->md.mu,
                       const_cast<float *>(c->md.mu32), int64_t{c->ne});
  }
  HIPCHK(hipStreamSynchronize(c->st));
  return CFDL_OK;
}

int cfdl_set_dt(cfdl_ctx *c, double dt) {
  CALL(check_ctx(c));
  if (!(dt > 0.0)) {
    set_error("cfdl_set_dt: dt must be positive");
    return CFDL_EINVAL;
  }
  c->dt = dt;
  return CFDL_OK;
}

int cfdl_set_bc_kind(cfdl_ctx *c, int first, int last, int kind) {
  CALL(check_ctx(c));
  if (first <= c->ne || last > c->ne + c->nbf || first > last || kind < CFDL_BC_WALL ||
      kind > CFDL_BC_OUTLET) {
    set_error("cfdl_set_bc_kind: halo range or kind out of range");
    return CFDL_EINVAL;
  }
  std::vector<int> k(last - first + 1, kind);
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(const_cast<int *>(c->md.hkind) + (first - c->ne - 1), k.data(),
                   sizeof(int) * k.size(), hipMemcpyHostToDevice));
  for (int el = first; el <= last; el++) c->hkind_h[el - c->ne - 1] = kind;
  return channel_lists(c);
}

int cfdl_set_inlet_profile(cfdl_ctx *c, double umax, double y0, double y1, double z0, double z1) {
  CALL(check_ctx(c));
  if (!(y1 > y0) || !(z1 > z0) || !std::isfinite(umax)) {
    set_error("cfdl_set_inlet_profile: need y1 > y0, z1 > z0 and a finite umax");
    return CFDL_EINVAL;
  }
  c->inlet = Inlet{{umax, y0, y1, z0, z1}};
  return CFDL_OK;
}

int cfdl_set_relaxation(cfdl_ctx *c, double alpha_u, double alpha_p) {
  CALL(check_ctx(c));
  if (!(alpha_u > 0.0 && alpha_u <= 1.0 && alpha_p > 0.0 && alpha_p <= 1.0)) {
    set_error("cfdl_set_relaxation: factors must lie in (0, 1]");
    return CFDL_EINVAL;
  }
  c->alpha_u = alpha_u;
  c->alpha_p = alpha_p;
  return CFDL_OK;
}

// the multigrid hierarchy of the mesh CSR and its device copy for the
// FP64 (fp32 = 0) or FP32 (fp32 = 1) preconditioner, built once
static int mg_ensure(cfdl_ctx *c, int fp32) {
  if (c->mg[fp32]) return CFDL_OK;
  const int ne = c->ne, n = ne + c->nbf;
  std::vector<int> row(ne + 1), nbr(c->L);
  std::vector<double> x(n), y(n), z(n);
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(row.data(), c->md.row, sizeof(int) * row.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(nbr.data(), c->md.nbr, sizeof(int) * nbr.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(x.data(), c->md.xc, sizeof(double) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(y.data(), c->md.yc, sizeof(double) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(z.data(), c->md.zc, sizeof(double) * n, hipMemcpyDeviceToHost));
  cfdl::MgHier h;
  CALL(cfdl::mg_build(ne, row.data(), nbr.data(), x.data(), y.data(), z.data(), kMgAgg,
                      kMgCoarse, &h));
  return cfdl::mgcg_create(h, c->nbf, fp32, &c->mg[fp32]);
}

int cfdl_set_assembly_precision(cfdl_ctx *c, int bits) {
  CALL(check_ctx(c));
  if (bits != 32 && bits != 64) {
    set_error("cfdl_set_assembly_precision: bits must be 32 or 64");
    return CFDL_EINVAL;
  }
  if (bits == 32 && !c->mhex) {
    set_error("cfdl_set_assembly_precision: FP32 assembly needs a mesh whose rows all have "
              "six entries");
    return CFDL_EINVAL;
  }
  if (bits == 32 && !c->md.xc32) {
    // FP32 copies of the geometry and properties (properties are re-copied
    // by cfdl_set_properties)
    const int n = c->ne + c->nbf;
    float *p[8];
    const double *src[8] = {c->md.xc, c->md.yc, c->md.zc, c->md.vol, c->md.aip, c->md.rip,
                            c->md.rho, c->md.mu};
    const int64_t cnt[8] = {n, n, n, c->ne, 3 * int64_t{c->nf}, 3 * int64_t{c->nf}, c->ne, c->ne};
    for (int i = 0; i < 8; i++) {
      CALL(dalloc(c, &p[i], cnt[i]));
      hipLaunchKernelGGL(k_to_f32, dim3(std::min<int64_t>((cnt[i] + kBlock - 1) / kBlock, 65535)),
                         dim3(kBlock), 0, c->st, src[i], p[i], cnt[i]);
    }
    HIPCHK(hipGetLastError());
    c->md.xc32 = p[0], c->md.yc32 = p[1], c->md.zc32 = p[2], c->md.vol32 = p[3];
    c->md.aip32 = p[4], c->md.rip32 = p[5], c->md.rho32 = p[6], c->md.mu32 = p[7];
  }
  c->asm32 = bits == 32;
  return CFDL_OK;
}

int cfdl_set_pc_order(cfdl_ctx *c, int order) {
  CALL(check_ctx(c));
  if (order < CFDL_PC_REFERENCE || order > CFDL_PC_MGCG_MIXED) {
    set_error("cfdl_set_pc_order: order must be CFDL_PC_REFERENCE, CFDL_PC_COLORED, "
              "CFDL_PC_COLORED_MIXED, CFDL_PC_MGCG or CFDL_PC_MGCG_MIXED");
    return CFDL_EINVAL;
  }
  if (order == CFDL_PC_MGCG || order == CFDL_PC_MGCG_MIXED)
    CALL(mg_ensure(c, order == CFDL_PC_MGCG_MIXED ? 1 : 0));
  if (order == CFDL_PC_COLORED_MIXED && !c->d32) {
    CALL(dalloc(c, &c->a32, c->ne));
    CALL(dalloc(c, &c->anb32, c->Lp));
    CALL(dalloc(c, &c->r32, c->ne));
    CALL(dalloc(c, &c->d32, static_cast<int64_t>(c->ne) + c->nh));
  }
  c->pc_order = order;
  return CFDL_OK;
}

int cfdl_set_gs_order(cfdl_ctx *c, int order) {
  CALL(check_ctx(c));
  if (order != CFDL_GS_REFERENCE && order != CFDL_GS_COLORED) {
    set_error("cfdl_set_gs_order: order must be CFDL_GS_REFERENCE or CFDL_GS_COLORED");
    return CFDL_EINVAL;
  }
  if (order == CFDL_GS_COLORED && !c->ap_m) {  // the copies the tiled path would own
    CALL(dalloc(c, &c->ap_m, c->ne));
    CALL(dalloc(c, &c->rap_m, c->ne));
    CALL(dalloc(c, &c->anb_m, c->L));
    for (int y = 0; y < 3; y++) {
      CALL(dalloc(c, &c->b_m[y], c->ne));
      CALL(dalloc(c, &c->phi_m[y], int64_t{c->ne} + c->nbf));
    }
  }
  c->gs_order = order;
  return CFDL_OK;
}

int cfdl_set_field(cfdl_ctx *c, int field, const double *src, int64_t n) {
  CALL(check_ctx(c));
  double *p = field_ptr(c, field);
  if (!p || !src || n != c->fsize[field]) {
    set_error("cfdl_set_field: bad field id or size");
    return CFDL_EINVAL;
  }
  HIPCHK(hipMemcpyAsync(p, src, 8 * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return CFDL_OK;
}

int cfdl_get_field(cfdl_ctx *c, int field, double *dst, int64_t n) {
  CALL(check_ctx(c));
  double *p = field_ptr(c, field);
  if (!p || !dst || n != c->fsize[field]) {
    set_error("cfdl_get_field: bad field id or size");
    return CFDL_EINVAL;
  }
  HIPCHK(hipMemcpyAsync(dst, p, 8 * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return CFDL_OK;
}

int cfdl_synchronize(cfdl_ctx *c) {
  CALL(check_ctx(c));
  HIPCHK(hipStreamSynchronize(c->st));
  return CFDL_OK;
}

int cfdl_ctx_reset(cfdl_ctx *c) {
  CALL(check_ctx(c));
  for (int f = 0; f < CFDL_NFIELDS; f++)
    hipLaunchKernelGGL(k_fill, dim3(nblocks(c->fsize[f])), dim3(kBlock), 0, c->st, c->fields[f],
                       c->fsize[f], 0.0);
  hipLaunchKernelGGL(k_energy_init, dim3(nblocks(int64_t{c->ne} + c->nbf)), dim3(kBlock), 0, c->st,
                     c->md, c->en);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  const int zero = 0;
  HIPCHK(hipMemcpy(c->dlog_n, &zero, sizeof(int), hipMemcpyHostToDevice));
  c->hlog.clear();
  c->outer = 0;
  return CFDL_OK;
}

int cfdl_update_boundaries(cfdl_ctx *c) {
  CALL(check_ctx(c));
  if (c->nbf > 0) {
    hipLaunchKernelGGL(k_bc_update, dim3(nblocks(c->nbf)), dim3(kBlock), 0, c->st, c->md, c->fd,
                       c->inlet);
    if (c->nout > 0)
      hipLaunchKernelGGL(k_outlet_flux, dim3(1), dim3(kOutThreads), 0, c->st, c->md, c->fd.mip,
                         static_cast<const int *>(c->fin), c->nin,
                         static_cast<const int *>(c->fout), c->nout);
    hipLaunchKernelGGL(k_bc_energy, dim3(nblocks(c->nbf)), dim3(kBlock), 0, c->st, c->md, c->en);
  }
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_update_boundaries_energy(cfdl_ctx *c) {
  CALL(check_ctx(c));
  if (c->nbf > 0)
    hipLaunchKernelGGL(k_bc_energy, dim3(nblocks(c->nbf)), dim3(kBlock), 0, c->st, c->md, c->en);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_set_thermal(cfdl_ctx *c, const double *cp, const double *tc) {
  CALL(check_ctx(c));
  if (!cp || !tc) {
    set_error("cfdl_set_thermal: null array");
    return CFDL_EINVAL;
  }
  HIPCHK(hipMemcpyAsync(const_cast<double *>(c->md.cp), cp, 8 * size_t(c->ne), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(const_cast<double *>(c->md.tc), tc, 8 * size_t(c->ne), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return CFDL_OK;
}

int cfdl_construct_energy(cfdl_ctx *c) {
  CALL(check_ctx(c));
  hipLaunchKernelGGL(k_energy_init, dim3(nblocks(int64_t{c->ne} + c->nbf)), dim3(kBlock), 0, c->st,
                     c->md, c->en);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_calc_coef_energy(cfdl_ctx *c) {
  CALL(check_ctx(c));
  hipLaunchKernelGGL(k_coef_energy, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->md, c->fd,
                     c->en, c->dt);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

// solve_energy, mod_energy.f90:58-80: gradients of t and h, coefficients,
// solve_gs('e') on the enthalpy (SOR factor 1), temperature from enthalpy
int cfdl_solve_energy(cfdl_ctx *c) {
  CALL(check_ctx(c));
  hipLaunchKernelGGL(k_grad, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->md, c->en.t, c->en.gt);
  hipLaunchKernelGGL(k_grad, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->md, c->en.h, c->en.gh);
  CALL(cfdl_calc_coef_energy(c));
  double *phi = c->en.h;
  const double *b = c->fd.b;
  const char *name = "e";
  CALL(solve_gs_n(c, 1, &phi, &b, &name, sor_of(name)));
  hipLaunchKernelGGL(k_temperature, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->md, c->en);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_calc_coef_uvw(cfdl_ctx *c) {
  CALL(check_ctx(c));
  // steady mode (dt = infinity): the transient terms vanish and, without
  // under-relaxation, ap - sum(anb) = 0 on interior cells, so dc = V / 0
  if (std::isinf(c->dt) && c->alpha_u == 1.0) {
    set_error("cfdl_calc_coef_uvw: dt = infinity (steady) needs momentum under-relaxation "
              "alpha_u < 1 (cfdl_set_relaxation)");
    return CFDL_ESTATE;
  }
  if (c->mhex && c->asm32)
    hipLaunchKernelGGL(k_coef_uvw6<float>, dim3((c->ne + kFaceCells - 1) / kFaceCells),
                       dim3(kFaceThreads), 0, c->st, c->md, c->fd, c->dt, c->alpha_u);
  else if (c->mhex)
    hipLaunchKernelGGL(k_coef_uvw6<double>, dim3((c->ne + kFaceCells - 1) / kFaceCells),
                       dim3(kFaceThreads), 0, c->st, c->md, c->fd, c->dt, c->alpha_u);
  else
    hipLaunchKernelGGL(k_coef_uvw, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->md, c->fd,
                       c->dt, c->alpha_u);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_solve_gs(cfdl_ctx *c, const char *name, int field, int rhs) {
  CALL(check_ctx(c));
  double *phi = field_ptr(c, field);
  double *b = field_ptr(c, rhs);
  if (!phi || !b || c->fsize[field] != c->ne + c->nbf || c->fsize[rhs] != c->ne || !name) {
    set_error("cfdl_solve_gs: phi must be an element field and rhs a cell field");
    return CFDL_EINVAL;
  }
  const double *bb = b;
  return solve_gs_n(c, 1, &phi, &bb, &name, sor_of(name));
}

int cfdl_solve_gs_uvw(cfdl_ctx *c) {
  CALL(check_ctx(c));
  double *phi[3] = {c->fd.u, c->fd.v, c->fd.w};
  const double *b[3] = {c->fd.bu, c->fd.bv, c->fd.bw};
  const char *names[3] = {"u", "v", "w"};
  return solve_gs_n(c, 3, phi, b, names, 1.0);
}

int cfdl_calc_grad(cfdl_ctx *c, int src, int dst) {
  CALL(check_ctx(c));
  double *p = field_ptr(c, src), *g = field_ptr(c, dst);
  if (!p || !g || c->fsize[src] != c->ne + c->nbf || c->fsize[dst] != 3 * int64_t{c->ne + c->nbf}) {
    set_error("cfdl_calc_grad: src must be an element field, dst a gradient field");
    return CFDL_EINVAL;
  }
  hipLaunchKernelGGL(k_grad, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->md, p, g);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_calc_mip(cfdl_ctx *c, int rhie_chow) {
  CALL(check_ctx(c));
  hipLaunchKernelGGL(k_mip, dim3(nblocks(c->nf)), dim3(kBlock), 0, c->st, c->md, c->fd,
                     rhie_chow ? 1 : 0, c->dt);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_calc_coef_p(cfdl_ctx *c) {
  CALL(check_ctx(c));
  if (c->mhex && c->asm32)
    hipLaunchKernelGGL(k_coef_p6<float>, dim3((c->ne + kFaceCells - 1) / kFaceCells),
                       dim3(kFaceThreads), 0, c->st, c->md, c->fd);
  else if (c->mhex)
    hipLaunchKernelGGL(k_coef_p6<double>, dim3((c->ne + kFaceCells - 1) / kFaceCells),
                       dim3(kFaceThreads), 0, c->st, c->md, c->fd);
  else
    hipLaunchKernelGGL(k_coef_p, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->md, c->fd);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_set_a_0(cfdl_ctx *c, int field) {
  CALL(check_ctx(c));
  double *p = field_ptr(c, field);
  if (!p) {
    set_error("cfdl_set_a_0: bad field id");
    return CFDL_EINVAL;
  }
  hipLaunchKernelGGL(k_fill, dim3(nblocks(c->fsize[field])), dim3(kBlock), 0, c->st, p,
                     c->fsize[field], 0.0);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_solve(cfdl_ctx *c, const char *name) {
  CALL(check_ctx(c));
  if (c->pc_order == CFDL_PC_MGCG || c->pc_order == CFDL_PC_MGCG_MIXED) {
    cfdl::MgCgSystem sys{};
    sys.ne = c->ne;
    sys.nbf = c->nbf;
    sys.nit = c->nit;
    sys.row = c->md.row;
    sys.nbr = c->md.nbr;
    sys.ap = c->fd.ap;
    sys.anb = c->fd.anb;
    sys.b = c->fd.b;
    sys.phi = c->fd.phic;
    std::strncpy(sys.name, name ? name : "pc", 7);
    sys.log = c->dlog;
    sys.nlog = c->dlog_n;
    return cfdl::mgcg_solve(c->mg[c->pc_order == CFDL_PC_MGCG_MIXED ? 1 : 0], sys, c->st);
  }
  if (c->nsubd <= 1) {
    double *phi = c->fd.phic;
    const double *b = c->fd.b;
    const char *nm = name ? name : "pc";
    return solve_gs_n(c, 1, &phi, &b, &nm, sor_of(nm));
  }
  return multi_subdomain_solver(c);
}

int cfdl_adjust_pc(cfdl_ctx *c) {
  CALL(check_ctx(c));
  hipLaunchKernelGGL(k_pref, dim3(1), dim3(64), 0, c->st, c->fd.phic, c->scal);
  hipLaunchKernelGGL(k_sub_pref, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->ne, c->fd.phic,
                     c->scal);
  if (c->nbf > 0)
    hipLaunchKernelGGL(k_halo_from_owner, dim3(nblocks(c->nbf)), dim3(kBlock), 0, c->st, c->md,
                       c->fd.phic);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_update_uvwp(cfdl_ctx *c) {
  CALL(check_ctx(c));
  hipLaunchKernelGGL(k_update_cells, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->ne, c->fd,
                     c->alpha_p);
  hipLaunchKernelGGL(k_update_faces, dim3(nblocks(c->nf)), dim3(kBlock), 0, c->st, c->md, c->fd);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_solve_uvwp(cfdl_ctx *c) {
  CALL(check_ctx(c));
  CALL(cfdl_calc_coef_uvw(c));
  CALL(cfdl_solve_gs_uvw(c));
  CALL(cfdl_calc_grad(c, CFDL_U, CFDL_GU));
  CALL(cfdl_calc_grad(c, CFDL_V, CFDL_GV));
  CALL(cfdl_calc_grad(c, CFDL_W, CFDL_GW));
  CALL(cfdl_calc_mip(c, 1));
  CALL(cfdl_calc_coef_p(c));
  CALL(cfdl_set_a_0(c, CFDL_PHIC));
  CALL(cfdl_solve(c, "pc"));
  CALL(cfdl_adjust_pc(c));
  CALL(cfdl_calc_grad(c, CFDL_PHIC, CFDL_GPC));
  CALL(cfdl_update_uvwp(c));
  c->outer++;
  return CFDL_OK;
}

int cfdl_calc_mass_imbalance(cfdl_ctx *c, int dst) {
  CALL(check_ctx(c));
  double *out = field_ptr(c, dst);
  if (!out || c->fsize[dst] < c->ne) {
    set_error("cfdl_calc_mass_imbalance: destination must be a field of at least ne values");
    return CFDL_EINVAL;
  }
  hipLaunchKernelGGL(k_mass_imbalance, dim3(nblocks(c->ne)), dim3(kBlock), 0, c->st, c->md,
                     c->fd.mip, out);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int cfdl_update_time(cfdl_ctx *c) {
  CALL(check_ctx(c));
  const int64_t n = c->ne + c->nbf;
  hipLaunchKernelGGL(k_copy, dim3(nblocks(n)), dim3(kBlock), 0, c->st, c->fd.u0, c->fd.u, n);
  hipLaunchKernelGGL(k_copy, dim3(nblocks(n)), dim3(kBlock), 0, c->st, c->fd.v0, c->fd.v, n);
  hipLaunchKernelGGL(k_copy, dim3(nblocks(n)), dim3(kBlock), 0, c->st, c->fd.w0, c->fd.w, n);
  hipLaunchKernelGGL(k_copy, dim3(nblocks(c->nf)), dim3(kBlock), 0, c->st, c->fd.mip0, c->fd.mip,
                     int64_t{c->nf});
  hipLaunchKernelGGL(k_copy, dim3(nblocks(n)), dim3(kBlock), 0, c->st, c->en.h0, c->en.h, n);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

static int run_loop(cfdl_ctx *c, int ntstep, int ncoef, bool energy) {
  CALL(check_ctx(c));
  for (int t = 0; t < ntstep; t++) {
    for (int k = 0; k < ncoef; k++) {
      CALL(cfdl_update_boundaries(c));
      CALL(cfdl_solve_uvwp(c));
      if (energy) CALL(cfdl_solve_energy(c));
    }
    CALL(cfdl_update_time(c));
  }
  return CFDL_OK;
}

int cfdl_run(cfdl_ctx *c, int ntstep, int ncoef) { return run_loop(c, ntstep, ncoef, false); }

int cfdl_run_energy(cfdl_ctx *c, int ntstep, int ncoef) { return run_loop(c, ntstep, ncoef, true); }

static int drain_log(cfdl_ctx *c) {
  HIPCHK(hipStreamSynchronize(c->st));
  int n = 0;
  HIPCHK(hipMemcpy(&n, c->dlog_n, sizeof(int), hipMemcpyDeviceToHost));
  if (n > 0) {
    std::vector<LogLine> tmp(n);
    HIPCHK(hipMemcpy(tmp.data(), c->dlog, sizeof(LogLine) * n, hipMemcpyDeviceToHost));
    c->hlog.insert(c->hlog.end(), tmp.begin(), tmp.end());
    const int zero = 0;
    HIPCHK(hipMemcpy(c->dlog_n, &zero, sizeof(int), hipMemcpyHostToDevice));
  }
  return CFDL_OK;
}

int cfdl_log_count(cfdl_ctx *c, int *n) {
  CALL(check_ctx(c));
  CALL(drain_log(c));
  if (n) *n = static_cast<int>(c->hlog.size());
  return CFDL_OK;
}

int cfdl_log_get(cfdl_ctx *c, int k, char name[8], int *it, double *res_i, double *res_f,
                 double *res_max) {
  CALL(check_ctx(c));
  if (k < 0 || k >= static_cast<int>(c->hlog.size())) {
    set_error("cfdl_log_get: index out of range (call cfdl_log_count first)");
    return CFDL_EINVAL;
  }
  const LogLine &l = c->hlog[k];
  if (name) std::memcpy(name, l.name, 8);
  if (it) *it = l.it;
  if (res_i) *res_i = l.res_i;
  if (res_f) *res_f = l.res_f;
  if (res_max) *res_max = l.res_max;
  return CFDL_OK;
}

int cfdl_log_clear(cfdl_ctx *c) {
  CALL(check_ctx(c));
  CALL(drain_log(c));
  c->hlog.clear();
  return CFDL_OK;
}

// calc_residual (src/modules/mod_solver.f90:230-253) on a system in device
// memory, outside any context: the residual kernels of the solves
// (k_resid_partial + k_resid_final, signed max as in the reference) with
// block partials in a buffer of 2 * nblk + 2 doubles (partials, then the sum
// of squares and the max).
constexpr int kResidualMaxCells = 1 << 30;
static void resid_shape(int ne, int *nblk, int *chunk) {
  *chunk = std::max(kBlock * 4, (ne + 2047) / 2048);
  *nblk = (ne + *chunk - 1) / *chunk;
}
static int resid_args(int ne, const int *row, const int *nbr, const double *ap, const double *anb,
                      const double *b, const double *phi, double *res, double *res_max,
                      const char *who) {
  if (ne < 1 || ne > kResidualMaxCells || !row || !nbr || !ap || !anb || !b || !phi || !res ||
      !res_max) {
    set_error(std::string(who) + ": need 1 <= ne <= 2^30 and non-null arrays");
    return CFDL_EINVAL;
  }
  return CFDL_OK;
}
static int resid_run(int ne, const int *row, const int *nbr, const double *ap, const double *anb,
                     const double *b, const double *phi, double *buf, hipStream_t st,
                     double *res, double *res_max) {
  int nblk = 0, chunk = 0;
  resid_shape(ne, &nblk, &chunk);
  SysDesc s{};
  s.ny = 1;
  s.beg[0] = 0;
  s.end[0] = ne;
  s.phi[0] = const_cast<double *>(phi);
  s.b[0] = b;
  s.row = row;
  s.nbr = nbr;
  s.ap = ap;
  s.anb = anb;
  double *part = buf, *red = buf + 2 * nblk;
  hipLaunchKernelGGL(k_resid_partial, dim3(nblk, 1), dim3(kBlock), 0, st, s, chunk, nblk, 0, part,
                     static_cast<float *>(nullptr));
  hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(kBlock), 0, st, static_cast<const double *>(part),
                     nblk, red);
  HIPCHK(hipGetLastError());
  double h[2];
  HIPCHK(hipMemcpyAsync(h, red, sizeof(h), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *res = std::sqrt(h[0] / static_cast<double>(ne));
  *res_max = h[1];
  return CFDL_OK;
}

int cfdl_calc_residual(int ne, const int *row, const int *nbr, const double *ap, const double *anb,
                       const double *b, const double *phi, double *res, double *res_max,
                       void *stream) {
  CALL(resid_args(ne, row, nbr, ap, anb, b, phi, res, res_max, "cfdl_calc_residual"));
  int nblk = 0, chunk = 0;
  resid_shape(ne, &nblk, &chunk);
  double *buf = nullptr;
  HIPCHK(hipMalloc(&buf, sizeof(double) * (2 * static_cast<size_t>(nblk) + 2)));
  const int rc = resid_run(ne, row, nbr, ap, anb, b, phi, buf, static_cast<hipStream_t>(stream),
                           res, res_max);
  HIPCHK(hipFree(buf));
  return rc;
}

int cfdl_calc_residual_noalloc(int ne, const int *row, const int *nbr, const double *ap,
                               const double *anb, const double *b, const double *phi, double *res,
                               double *res_max, void *stream) {
  CALL(resid_args(ne, row, nbr, ap, anb, b, phi, res, res_max, "cfdl_calc_residual_noalloc"));
  double *buf = nullptr;
  CALL(resid_scratch(&buf));
  return resid_run(ne, row, nbr, ap, anb, b, phi, buf, static_cast<hipStream_t>(stream), res,
                   res_max);
}

int cfdl_set_profile(cfdl_ctx *c, int sample_every) {
  CALL(check_ctx(c));
  if (sample_every < 0) {
    set_error("cfdl_set_profile: sample_every must be >= 0");
    return CFDL_EINVAL;
  }
  c->sample_every = sample_every;
  return CFDL_OK;
}

int cfdl_stats(cfdl_ctx *c, double *stats, int n) {
  CALL(check_ctx(c));
  HIPCHK(hipStreamSynchronize(c->st));
  double ms = 0.0, cells = 0.0, nl = 0.0;
  for (size_t k = 0; k < c->nsamples; k++) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->samples[k].a, c->samples[k].b));
    ms += t;
    cells += static_cast<double>(c->samples[k].cells);
    nl += c->samples[k].launches;
  }
  const double v[8] = {static_cast<double>(c->gs_launches), static_cast<double>(c->pc_launches),
                       static_cast<double>(c->pc_cells), static_cast<double>(c->outer), nl,
                       nl > 0 ? ms / nl : 0.0, nl > 0 ? cells / nl : 0.0,
                       static_cast<double>(2 * c->pc_nlev)};
  for (int i = 0; i < n && i < 8; i++) stats[i] = v[i];
  return CFDL_OK;
}

int cfdl_stats_reset(cfdl_ctx *c) {
  CALL(check_ctx(c));
  HIPCHK(hipStreamSynchronize(c->st));
  c->gs_launches = c->pc_launches = c->pc_cells = c->outer = c->pc_sweeps = 0;
  c->nsamples = 0;
  return CFDL_OK;
}

}  // extern "C"

// ===========================================================================
// Batches (continuous batching of independent cases, an extension with no
// counterpart in the reference, whose program solves one case per run)
// ===========================================================================
// A batch runs the SIMPLE loop of cfdl_run for every attached context at
// once: each step is one grouped launch over all cases (BSlot descriptors in
// device memory, rebuilt when a case's settings or the composition change),
// the momentum and pressure-correction solves stop per case, and the host
// waits once per solver iteration for all cases together.  Cases attach and
// detach between calls, so a serving loop admits new cases while others are
// part-way through their solves.  Requirements of a case: hexahedral rows,
// the colour order of solve_gs, a multigrid-preconditioned pc solve, and the
// batch's assembly and preconditioner precision (set by its first case).
struct cfdl_batch {
  int dev = 0;
  hipStream_t st = nullptr;
  std::vector<cfdl_ctx *> ctx;
  int asm32 = -1, fp32 = -1;
  cfdl::MgGroup *mg = nullptr;
  BSlot *d_slots = nullptr;
  int *d_flags = nullptr;
  int *h_flags = nullptr;  // pinned, two sets of kMaxGroup
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<char> key;   // the descriptors on the device
  int block = 0;           // lane count of the element-wise launches; 0 = tuned
  int tuned = 0;           // the tuner's choice for the current composition
  bool tuning = false;     // tuner events recorded, not yet read
  hipEvent_t tev[6] = {};
  int64_t runs = 0;
  // Captured graphs of the iteration's segments (prefix, one momentum-solver
  // iteration per flag slot, middle, one CG iteration per flag slot, suffix,
  // update_time), valid for one composition and lane count
  struct Seg {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
  };
  bool use_graphs = true;
  bool graphs_valid = false;
  int graph_blk = 0;
  Seg seg[8];
  int64_t graph_launches = 0, captures = 0;
};

namespace {
constexpr int kTuneBlocks[5] = {64, 128, 256, 512, 1024};
inline int nbk(int64_t n, int blk) { return static_cast<int>((n + blk - 1) / blk); }

// the descriptor of a case (its device pointers and settings)
int batch_slot(cfdl_ctx *c, BSlot *s) {
  std::memset(static_cast<void *>(s), 0, sizeof(BSlot));
  s->md = c->md;
  s->fd = c->fd;
  s->en = c->en;
  s->in = c->inlet;
  s->dt = c->dt;
  s->alpha_u = c->alpha_u;
  s->alpha_p = c->alpha_p;
  s->fin = c->fin;
  s->fout = c->fout;
  s->nin = c->nin;
  s->nout = c->nout;
  s->nblk_g = c->nblk_g;
  s->chunk_g = c->chunk_g;
  s->nit = c->nit;
  s->nq = static_cast<int>(c->gcol_ptr.size()) - 1;
  if (s->nq < 1 || s->nq > kMaxColours) {
    set_error("cfdl_batch: a case's momentum system has " + std::to_string(s->nq) +
              " colours (1 .. " + std::to_string(kMaxColours) + " allowed)");
    return CFDL_ESTATE;
  }
  for (int q = 0; q <= s->nq; q++) s->gq[q] = c->gcol_ptr[q];
  TileMap &tm = s->tmap;
  tm.ne = c->ne;
  tm.nbf = c->nbf;
  tm.ny = 3;
  tm.l2g = c->l2g_gc;
  tm.rowm = c->rowgc;
  tm.rowg = c->md.row;
  tm.ap = c->fd.ap;
  tm.anb = c->fd.anb;
  tm.ap_m = c->ap_m;
  tm.anb_m = c->anb_m;
  tm.rap_m = c->rap_m;
  double *phi[3] = {c->fd.u, c->fd.v, c->fd.w};
  const double *b[3] = {c->fd.bu, c->fd.bv, c->fd.bw};
  SysDesc &sy = s->sys;
  sy.ny = 3;
  for (int y = 0; y < 3; y++) {
    tm.b[y] = b[y];
    tm.b_m[y] = c->b_m[y];
    tm.phi[y] = phi[y];
    tm.phi_m[y] = c->phi_m[y];
    sy.beg[y] = 0;
    sy.end[y] = c->ne;
    sy.phi[y] = c->phi_m[y];
    sy.b[y] = c->b_m[y];
  }
  sy.row = c->rowgc;
  sy.nbr = c->nbrgc;
  sy.ap = c->ap_m;
  sy.anb = c->anb_m;
  sy.rap = c->rap_m;
  s->gctl = c->gctl;
  s->part = c->part;
  s->red = c->red;
  s->scal = c->scal;
  s->log = c->dlog;
  s->nlog = c->dlog_n;
  return CFDL_OK;
}

int batch_check(cfdl_batch *b, cfdl_ctx *c) {
  if (c->dev != b->dev) {
    set_error("cfdl_batch: the case is on another device");
    return CFDL_EINVAL;
  }
  if (!c->mhex) {
    set_error("cfdl_batch: cases need hexahedral rows (six entries each)");
    return CFDL_ESTATE;
  }
  if (c->gs_order != CFDL_GS_COLORED) {
    set_error("cfdl_batch: cases need the colour order of solve_gs (cfdl_set_gs_order)");
    return CFDL_ESTATE;
  }
  if (c->pc_order != CFDL_PC_MGCG && c->pc_order != CFDL_PC_MGCG_MIXED) {
    set_error("cfdl_batch: cases need a multigrid-preconditioned pc solve (cfdl_set_pc_order)");
    return CFDL_ESTATE;
  }
  const int fp32 = c->pc_order == CFDL_PC_MGCG_MIXED ? 1 : 0, asm32 = c->asm32 ? 1 : 0;
  if (!c->mg[fp32]) {
    set_error("cfdl_batch: the case has no multigrid hierarchy");
    return CFDL_ESTATE;
  }
  // the precision of the batch is its first case's
  const cfdl_ctx *c0 = b->ctx.empty() ? c : b->ctx[0];
  if (fp32 != (c0->pc_order == CFDL_PC_MGCG_MIXED ? 1 : 0) || asm32 != (c0->asm32 ? 1 : 0)) {
    set_error("cfdl_batch: the cases of a batch share the assembly and pc precision");
    return CFDL_ESTATE;
  }
  if (std::isinf(c->dt) && c->alpha_u == 1.0) {
    set_error("cfdl_batch: dt = infinity (steady) needs momentum under-relaxation alpha_u < 1");
    return CFDL_ESTATE;
  }
  return CFDL_OK;
}

// descriptors of every case, uploaded when they differ from the device copy
int batch_sync_slots(cfdl_batch *b) {
  const int n = static_cast<int>(b->ctx.size());
  std::vector<BSlot> h(n);
  for (int k = 0; k < n; k++) {
    CALL(batch_check(b, b->ctx[k]));
    CALL(batch_slot(b->ctx[k], &h[k]));
  }
  if (n > 0) {
    b->fp32 = b->ctx[0]->pc_order == CFDL_PC_MGCG_MIXED ? 1 : 0;
    b->asm32 = b->ctx[0]->asm32 ? 1 : 0;
  }
  const char *kb = reinterpret_cast<const char *>(h.data());
  std::vector<char> key(kb, kb + sizeof(BSlot) * n);
  if (key != b->key) {
    HIPCHK(hipStreamSynchronize(b->st));
    if (n > 0) HIPCHK(hipMemcpy(b->d_slots, h.data(), sizeof(BSlot) * n, hipMemcpyHostToDevice));
    b->key.swap(key);
    b->graphs_valid = false;
  }
  return CFDL_OK;
}

// The batch's grouped launches: one GroupMap per launch from the per-case
// block counts.
struct BatchLaunch {
  cfdl_batch *b;
  int n;
  GroupMap g;
  std::vector<int> ids, cnt;
  explicit BatchLaunch(cfdl_batch *b_) : b(b_), n(static_cast<int>(b_->ctx.size())), g{} {
    ids.resize(n);
    cnt.resize(n);
    for (int k = 0; k < n; k++) ids[k] = k;
  }
  template <typename F>
  int map(F blocks_of) {
    for (int k = 0; k < n; k++) cnt[k] = blocks_of(b->ctx[k]);
    return group_map(g, ids.data(), cnt.data(), n);
  }
};

// The iteration in segments, each only enqueueing work on the batch's
// stream (so it can be captured once and replayed as a graph):
//   prefix: update_boundaries, calc_coef_uvw, the momentum system gathered
//           into the colour numbering, its initial residuals;
//   gs:     one iteration of solve_gs_uvw for every case still running, its
//           stop test, the cases' flags copied to host slot 0 or 1;
//   middle: the system scattered back, the momentum log lines, calc_grad of
//           u, v, w, calc_mip, calc_coef_p, set_a_0(phic), the pc solve's
//           setup;
//   cg:     one CG iteration with its V-cycle, flags to host slot 0 or 1;
//   suffix: the pc log lines, adjust_pc, calc_grad(phic), update_uvwp;
//   time:   update_time.
// The host runs the gs and cg segments until no case's flag is set (it
// queues iteration it+1 before reading the flags after iteration it; an
// iteration queued after a case stopped changes nothing of that case: its
// kernels skip, and a recomputed residual of unchanged values is not used).
int seg_prefix(cfdl_batch *b, int blk) {
  BatchLaunch L(b);
  const BSlot *S = b->d_slots;
  hipStream_t st = b->st;
  int nb = L.map([&](cfdl_ctx *c) { return nbk(c->nbf, blk); });
  if (nb > 0) hipLaunchKernelGGL(gb_bc_update, dim3(nb), dim3(blk), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *c) { return c->nbf > 0 && c->nout > 0 ? 1 : 0; });
  if (nb > 0) hipLaunchKernelGGL(gb_outlet_flux, dim3(nb), dim3(kOutThreads), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *c) { return nbk(c->nbf, blk); });
  if (nb > 0) hipLaunchKernelGGL(gb_bc_energy, dim3(nb), dim3(blk), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *c) { return (c->ne + kFaceCells - 1) / kFaceCells; });
  if (b->asm32)
    hipLaunchKernelGGL(gb_coef_uvw6<float>, dim3(nb), dim3(kFaceThreads), 0, st, L.g, S);
  else
    hipLaunchKernelGGL(gb_coef_uvw6<double>, dim3(nb), dim3(kFaceThreads), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *c) { return nbk(int64_t{c->ne} + c->nbf, blk); });
  hipLaunchKernelGGL(gb_tile_gather, dim3(nb), dim3(blk), 0, st, L.g, S);
  const int np = L.map([&](cfdl_ctx *c) { return c->nblk_g; });
  hipLaunchKernelGGL(gb_resid_partial, dim3(np, 3), dim3(kBlock), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *) { return 1; });
  hipLaunchKernelGGL(gb_resid_final, dim3(nb, 3), dim3(kBlock), 0, st, L.g, S);
  hipLaunchKernelGGL(gb_gs_init, dim3(nb), dim3(64), 0, st, L.g, S);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int seg_gs(cfdl_batch *b, int blk, int slot) {
  BatchLaunch L(b);
  const BSlot *S = b->d_slots;
  hipStream_t st = b->st;
  int nq = 0;
  for (cfdl_ctx *c : b->ctx) nq = std::max(nq, static_cast<int>(c->gcol_ptr.size()) - 1);
  for (int d = 0; d < 2; d++)
    for (int i = 0; i < nq; i++) {
      const int nb = L.map([&](cfdl_ctx *c) {
        const int q_n = static_cast<int>(c->gcol_ptr.size()) - 1;
        if (i >= q_n) return 0;
        const int q = d == 0 ? i : q_n - 1 - i;
        return nbk(c->gcol_ptr[q + 1] - c->gcol_ptr[q], blk);
      });
      if (nb > 0) hipLaunchKernelGGL(gb_gs_colour, dim3(nb), dim3(blk), 0, st, L.g, S, i, d, 1.0);
    }
  const int np = L.map([&](cfdl_ctx *c) { return c->nblk_g; });
  hipLaunchKernelGGL(gb_resid_partial, dim3(np, 3), dim3(kBlock), 0, st, L.g, S);
  const int nb = L.map([&](cfdl_ctx *) { return 1; });
  hipLaunchKernelGGL(gb_resid_final, dim3(nb, 3), dim3(kBlock), 0, st, L.g, S);
  hipLaunchKernelGGL(gb_gs_iter_end, dim3(nb), dim3(64), 0, st, L.g, S);
  hipLaunchKernelGGL(gb_gs_flags, dim3(1), dim3(kMaxGroup), 0, st, L.n, S, b->d_flags);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->h_flags + slot * kMaxGroup, b->d_flags, sizeof(int) * L.n,
                        hipMemcpyDeviceToHost, st));
  return CFDL_OK;
}

// The tuner: at the first gradient step after the composition changed,
// the gradient of u is launched at each candidate lane count (three times
// each; the launches recompute the same values) between events; the
// fastest is read at the end of the run and used for every element-wise
// launch of the batch from then on.
int batch_grad_u(cfdl_batch *b, int blk) {
  BatchLaunch L(b);
  const BSlot *S = b->d_slots;
  if (b->block == 0 && b->tuned == 0 && !b->tuning) {
    for (int j = 0; j < 5; j++) {
      const int nb = L.map([&](cfdl_ctx *c) { return nbk(c->ne, kTuneBlocks[j]); });
      HIPCHK(hipEventRecord(b->tev[j], b->st));
      for (int r = 0; r < 3; r++)
        hipLaunchKernelGGL(gb_grad, dim3(nb), dim3(kTuneBlocks[j]), 0, b->st, L.g, S, 0);
    }
    HIPCHK(hipEventRecord(b->tev[5], b->st));
    b->tuning = true;
  }
  const int nb = L.map([&](cfdl_ctx *c) { return nbk(c->ne, blk); });
  hipLaunchKernelGGL(gb_grad, dim3(nb), dim3(blk), 0, b->st, L.g, S, 0);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}
int batch_tune_read(cfdl_batch *b) {
  if (!b->tuning) return CFDL_OK;
  HIPCHK(hipEventSynchronize(b->tev[5]));
  float best = 0.f;
  int bj = 2;
  for (int j = 0; j < 5; j++) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->tev[j], b->tev[j + 1]));
    if (j == 0 || ms < best) {
      best = ms;
      bj = j;
    }
  }
  b->tuned = kTuneBlocks[bj];
  b->tuning = false;
  return CFDL_OK;
}

int seg_middle(cfdl_batch *b, int blk, bool tune) {
  BatchLaunch L(b);
  const BSlot *S = b->d_slots;
  hipStream_t st = b->st;
  int nb = L.map([&](cfdl_ctx *c) { return nbk(c->ne, blk); });
  hipLaunchKernelGGL(gb_tile_scatter, dim3(nb), dim3(blk), 0, st, L.g, S);
  Names nm{};
  std::strncpy(nm.n[0], "u", 7);
  std::strncpy(nm.n[1], "v", 7);
  std::strncpy(nm.n[2], "w", 7);
  nb = L.map([&](cfdl_ctx *) { return 1; });
  hipLaunchKernelGGL(gb_log_gs, dim3(nb), dim3(64), 0, st, L.g, S, nm);
  HIPCHK(hipGetLastError());
  if (tune) {
    CALL(batch_grad_u(b, blk));
  } else {
    nb = L.map([&](cfdl_ctx *c) { return nbk(c->ne, blk); });
    hipLaunchKernelGGL(gb_grad, dim3(nb), dim3(blk), 0, st, L.g, S, 0);
  }
  nb = L.map([&](cfdl_ctx *c) { return nbk(c->ne, blk); });
  hipLaunchKernelGGL(gb_grad, dim3(nb), dim3(blk), 0, st, L.g, S, 1);
  hipLaunchKernelGGL(gb_grad, dim3(nb), dim3(blk), 0, st, L.g, S, 2);
  nb = L.map([&](cfdl_ctx *c) { return nbk(c->nf, blk); });
  hipLaunchKernelGGL(gb_mip, dim3(nb), dim3(blk), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *c) { return (c->ne + kFaceCells - 1) / kFaceCells; });
  if (b->asm32)
    hipLaunchKernelGGL(gb_coef_p6<float>, dim3(nb), dim3(kFaceThreads), 0, st, L.g, S);
  else
    hipLaunchKernelGGL(gb_coef_p6<double>, dim3(nb), dim3(kFaceThreads), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *c) { return nbk(int64_t{c->ne} + c->nbf, blk); });
  hipLaunchKernelGGL(gb_zero_phic, dim3(nb), dim3(blk), 0, st, L.g, S);
  HIPCHK(hipGetLastError());
  return cfdl::mgcg_group_setup(b->mg, st);
}

int seg_cg(cfdl_batch *b, int slot) { return cfdl::mgcg_group_iteration(b->mg, b->st, slot); }

int seg_suffix(cfdl_batch *b, int blk) {
  BatchLaunch L(b);
  const BSlot *S = b->d_slots;
  hipStream_t st = b->st;
  CALL(cfdl::mgcg_group_finish(b->mg, st));
  int nb = L.map([&](cfdl_ctx *) { return 1; });
  hipLaunchKernelGGL(gb_pref, dim3(nb), dim3(64), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *c) { return nbk(c->ne, blk); });
  hipLaunchKernelGGL(gb_sub_pref, dim3(nb), dim3(blk), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *c) { return nbk(c->nbf, blk); });
  if (nb > 0) hipLaunchKernelGGL(gb_halo_from_owner, dim3(nb), dim3(blk), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *c) { return nbk(c->ne, blk); });
  hipLaunchKernelGGL(gb_grad, dim3(nb), dim3(blk), 0, st, L.g, S, 3);
  hipLaunchKernelGGL(gb_update_cells, dim3(nb), dim3(blk), 0, st, L.g, S);
  nb = L.map([&](cfdl_ctx *c) { return nbk(c->nf, blk); });
  hipLaunchKernelGGL(gb_update_faces, dim3(nb), dim3(blk), 0, st, L.g, S);
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

int seg_time(cfdl_batch *b, int blk) {
  BatchLaunch L(b);
  for (int w = 0; w < 5; w++) {
    const int nb = L.map([&](cfdl_ctx *c) {
      return nbk(w == 3 ? int64_t{c->nf} : int64_t{c->ne} + c->nbf, blk);
    });
    hipLaunchKernelGGL(gb_update_time, dim3(nb), dim3(blk), 0, b->st, L.g, b->d_slots, w);
  }
  HIPCHK(hipGetLastError());
  return CFDL_OK;
}

void graphs_drop(cfdl_batch *b) {
  for (auto &s : b->seg) {
    if (s.x) (void)hipGraphExecDestroy(s.x);
    if (s.g) (void)hipGraphDestroy(s.g);
    s.x = nullptr;
    s.g = nullptr;
  }
  b->graphs_valid = false;
}

// Runs segment k: eagerly, or as its graph (captured on first use for the
// current composition and lane count)
template <typename F>
int run_seg(cfdl_batch *b, bool graph, int k, F fn) {
  if (!graph) return fn();
  cfdl_batch::Seg &s = b->seg[k];
  if (!s.x) {
    HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed));
    const int rc = fn();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->st, &g);
    if (rc != CFDL_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIPCHK(e);
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      (void)hipGraphDestroy(g);
      HIPCHK(ei);
    }
    s.g = g;
    s.x = x;
    b->captures++;
  }
  HIPCHK(hipGraphLaunch(s.x, b->st));
  b->graph_launches++;
  return CFDL_OK;
}

int batch_iteration(cfdl_batch *b) {
  hipStream_t st = b->st;
  const bool tune = b->block == 0 && b->tuned == 0 && !b->tuning;
  const int blk = b->block > 0 ? b->block : b->tuned > 0 ? b->tuned : kBlock;
  // the pc solve's descriptors (outside any capture: this may synchronise)
  const int n = static_cast<int>(b->ctx.size());
  std::vector<cfdl::MgCgSolver *> sv(n);
  std::vector<cfdl::MgCgSystem> sys(n);
  for (int k = 0; k < n; k++) {
    cfdl_ctx *c = b->ctx[k];
    sv[k] = c->mg[b->fp32];
    cfdl::MgCgSystem &y = sys[k];
    y = cfdl::MgCgSystem{};
    y.ne = c->ne;
    y.nbf = c->nbf;
    y.nit = c->nit;
    y.row = c->md.row;
    y.nbr = c->md.nbr;
    y.ap = c->fd.ap;
    y.anb = c->fd.anb;
    y.b = c->fd.b;
    y.phi = c->fd.phic;
    std::strncpy(y.name, "pc", 7);
    y.log = c->dlog;
    y.nlog = c->dlog_n;
  }
  int changed = 0;
  CALL(cfdl::mgcg_group_begin(b->mg, sv.data(), sys.data(), n, st, blk, &changed));
  if (changed || blk != b->graph_blk || !b->graphs_valid) {
    graphs_drop(b);
    b->graph_blk = blk;
    b->graphs_valid = true;
  }
  // the tuner's first iteration runs eagerly (its timed launches are not
  // captured)
  const bool graph = b->use_graphs && !tune;
  CALL(run_seg(b, graph, 0, [&] { return seg_prefix(b, blk); }));
  int nit = 0;
  for (cfdl_ctx *c : b->ctx) nit = std::max(nit, c->nit);
  for (int it = 0; it < nit; it++) {
    CALL(run_seg(b, graph, 1 + (it & 1), [&] { return seg_gs(b, blk, it & 1); }));
    HIPCHK(hipEventRecord(b->ev[it & 1], st));
    if (it > 0) {
      HIPCHK(hipEventSynchronize(b->ev[(it - 1) & 1]));
      const int *pf = b->h_flags + ((it - 1) & 1) * kMaxGroup;
      bool any = false;
      for (int k = 0; k < n; k++) any = any || pf[k] != 0;
      if (!any) break;
    }
  }
  CALL(run_seg(b, graph, 3, [&] { return seg_middle(b, blk, tune); }));
  const int cit = cfdl::mgcg_group_nit(b->mg);
  for (int it = 0; it < cit; it++) {
    CALL(run_seg(b, graph, 4 + (it & 1), [&] { return seg_cg(b, it & 1); }));
    HIPCHK(hipEventRecord(b->ev[it & 1], st));
    if (it > 0) {
      HIPCHK(hipEventSynchronize(b->ev[(it - 1) & 1]));
      if (!cfdl::mgcg_group_any_active(b->mg, (it - 1) & 1)) break;
    }
  }
  CALL(run_seg(b, graph, 6, [&] { return seg_suffix(b, blk); }));
  for (cfdl_ctx *c : b->ctx) c->outer++;
  return CFDL_OK;
}

int batch_update_time(cfdl_batch *b) {
  const int blk = b->block > 0 ? b->block : b->tuned > 0 ? b->tuned : kBlock;
  const bool graph = b->use_graphs && blk == b->graph_blk && b->graphs_valid;
  return run_seg(b, graph, 7, [&] { return seg_time(b, blk); });
}

void batch_free(cfdl_batch *b) {
  graphs_drop(b);
  if (b->mg) cfdl::mgcg_group_destroy(b->mg);
  if (b->d_slots) (void)hipFree(b->d_slots);
  if (b->d_flags) (void)hipFree(b->d_flags);
  if (b->h_flags) (void)hipHostFree(b->h_flags);
  for (auto &e : b->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : b->tev)
    if (e) (void)hipEventDestroy(e);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}
}  // namespace

extern "C" {

int cfdl_batch_create(cfdl_batch **out) {
  if (!out) {
    set_error("cfdl_batch_create: null output");
    return CFDL_EINVAL;
  }
  *out = nullptr;
  cfdl_batch *b = new cfdl_batch();
  auto fail = [&](hipError_t e, const char *what) {
    set_error(std::string("cfdl_batch_create: ") + what + ": " + hipGetErrorString(e));
    batch_free(b);
    return CFDL_EHIP;
  };
  hipError_t e = hipGetDevice(&b->dev);
  if (e != hipSuccess) return fail(e, "hipGetDevice");
  if ((e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking)) != hipSuccess)
    return fail(e, "hipStreamCreateWithFlags");
  if ((e = hipMalloc(reinterpret_cast<void **>(&b->d_slots), sizeof(BSlot) * kMaxGroup)) !=
      hipSuccess)
    return fail(e, "hipMalloc");
  if ((e = hipMalloc(reinterpret_cast<void **>(&b->d_flags), sizeof(int) * kMaxGroup)) != hipSuccess)
    return fail(e, "hipMalloc");
  if ((e = hipHostMalloc(reinterpret_cast<void **>(&b->h_flags), 2 * sizeof(int) * kMaxGroup,
                         hipHostMallocDefault)) != hipSuccess)
    return fail(e, "hipHostMalloc");
  for (auto &v : b->ev)
    if ((e = hipEventCreateWithFlags(&v, hipEventDisableTiming)) != hipSuccess)
      return fail(e, "hipEventCreate");
  for (auto &v : b->tev)
    if ((e = hipEventCreate(&v)) != hipSuccess) return fail(e, "hipEventCreate");
  if (cfdl::mgcg_group_create(&b->mg) != CFDL_OK) {
    batch_free(b);
    return CFDL_EHIP;
  }
  *out = b;
  return CFDL_OK;
}

int cfdl_batch_attach(cfdl_batch *b, cfdl_ctx *c) {
  if (!b || !c) {
    set_error("cfdl_batch_attach: null batch or context");
    return CFDL_EINVAL;
  }
  if (c->batch) {
    set_error("cfdl_batch_attach: the context is already in a batch");
    return CFDL_ESTATE;
  }
  if (static_cast<int>(b->ctx.size()) >= kMaxGroup) {
    set_error("cfdl_batch_attach: the batch is full (" + std::to_string(kMaxGroup) + " cases)");
    return CFDL_ESTATE;
  }
  CALL(batch_check(b, c));
  // the case's queued work completes on its own stream; from here on its
  // work runs on the batch's
  HIPCHK(hipStreamSynchronize(c->st));
  c->own_st = c->st;
  c->st = b->st;
  c->batch = b;
  b->ctx.push_back(c);
  b->tuned = 0;
  return CFDL_OK;
}

int cfdl_batch_detach(cfdl_batch *b, cfdl_ctx *c) {
  if (!b || !c || c->batch != b) {
    set_error("cfdl_batch_detach: the context is not in this batch");
    return CFDL_EINVAL;
  }
  HIPCHK(hipStreamSynchronize(b->st));
  c->st = c->own_st;
  c->own_st = nullptr;
  c->batch = nullptr;
  b->ctx.erase(std::find(b->ctx.begin(), b->ctx.end(), c));
  b->tuned = 0;
  return CFDL_OK;
}

int cfdl_batch_size(cfdl_batch *b, int *n) {
  if (!b || !n) {
    set_error("cfdl_batch_size: null argument");
    return CFDL_EINVAL;
  }
  *n = static_cast<int>(b->ctx.size());
  return CFDL_OK;
}

int cfdl_batch_set_block(cfdl_batch *b, int block) {
  if (!b || (block != 0 && (block < 64 || block > 1024 || block % 64 != 0))) {
    set_error("cfdl_batch_set_block: 0 (tuned) or a multiple of 64 in 64 .. 1024");
    return CFDL_EINVAL;
  }
  b->block = block;
  return CFDL_OK;
}

int cfdl_batch_block(cfdl_batch *b, int *block) {
  if (!b || !block) {
    set_error("cfdl_batch_block: null argument");
    return CFDL_EINVAL;
  }
  *block = b->block > 0 ? b->block : b->tuned > 0 ? b->tuned : kBlock;
  return CFDL_OK;
}

int cfdl_batch_set_graphs(cfdl_batch *b, int on) {
  if (!b) {
    set_error("cfdl_batch_set_graphs: null batch");
    return CFDL_EINVAL;
  }
  b->use_graphs = on != 0;
  return CFDL_OK;
}

int cfdl_batch_graph_stats(cfdl_batch *b, int64_t *captures, int64_t *launches) {
  if (!b) {
    set_error("cfdl_batch_graph_stats: null batch");
    return CFDL_EINVAL;
  }
  if (captures) *captures = b->captures;
  if (launches) *launches = b->graph_launches;
  return CFDL_OK;
}

int cfdl_batch_run(cfdl_batch *b, int ntstep, int ncoef) {
  if (!b || ntstep < 0 || ncoef < 0) {
    set_error("cfdl_batch_run: null batch or negative counts");
    return CFDL_EINVAL;
  }
  if (b->ctx.empty()) return CFDL_OK;
  CALL(batch_sync_slots(b));
  for (int t = 0; t < ntstep; t++) {
    for (int k = 0; k < ncoef; k++) CALL(batch_iteration(b));
    CALL(batch_update_time(b));
  }
  CALL(batch_tune_read(b));
  b->runs++;
  return CFDL_OK;
}

int cfdl_batch_synchronize(cfdl_batch *b) {
  if (!b) {
    set_error("cfdl_batch_synchronize: null batch");
    return CFDL_EINVAL;
  }
  HIPCHK(hipStreamSynchronize(b->st));
  return CFDL_OK;
}

void cfdl_batch_destroy(cfdl_batch *b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->st);
  while (!b->ctx.empty()) (void)cfdl_batch_detach(b, b->ctx.back());
  batch_free(b);
}

}  // extern "C"
