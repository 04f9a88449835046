// cfdl_mg.cpp -- host set-up of the agglomeration multigrid hierarchy used
// by the multigrid-preconditioned conjugate-gradient pressure-correction
// solve (an extension; the reference solves the pc system with the
// 4-subdomain Gauss-Seidel solver of src/modules/mod_solver.f90:124-189 and
// builds agglomeration levels, src/setup/mod_agglomeration.f90:331-561 and
// src/setup/mod_mg_lvl_uns.f90:490-1001, that no solver uses).
//
// Agglomeration follows the reference's seed splitting rule
// (mod_agglomeration.f90:380-410): a node of more than `maxagg` cells is
// split on the axis of greatest extent of its cell centres (the first on
// ties) at (min + max) / 2, centres below the midpoint going left, both sides
// keeping increasing cell order (halves if a side would be empty).  One
// bisection tree gives every level at once: a level-l aggregate is a maximal
// node of at most maxagg^l cells, numbered in pre-order, so every aggregate is
// a union of whole aggregates of the level below.  The coarse graphs, greedy
// colourings and Galerkin sum lists are what the device kernels walk; their
// order fixes the order of every floating-point sum.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "cfdl_internal.h"

namespace cfdl {

namespace {

struct Bisect {
  const double *xyz[3];
  std::vector<long> cap;            // cap[l] = maxagg^l (l >= 1)
  int nl = 0;                       // levels considered
  std::vector<int> count;           // aggregates numbered so far, per level
  std::vector<std::vector<int>> up;  // up[l]: level-(l-1) node -> level-l aggregate
  std::vector<int> tmp;
};

void bisect_node(Bisect &B, int *ids, int n, long parent) {
  // level-(l-1) aggregates numbered inside this node start here
  int start[32];
  for (int l = 1; l <= B.nl; l++) start[l] = B.count[l - 1];
  int form[32];
  for (int l = 1; l <= B.nl; l++) {
    form[l] = -1;
    if (n <= B.cap[l] && parent > B.cap[l]) form[l] = B.count[l]++;
  }
  if (form[1] >= 0)
    for (int i = 0; i < n; i++) B.up[1][ids[i]] = form[1];
  if (n > B.cap[1]) {
    double mn[3], mx[3];
    for (int a = 0; a < 3; a++) mn[a] = mx[a] = B.xyz[a][ids[0]];
    for (int i = 1; i < n; i++)
      for (int a = 0; a < 3; a++) {
        const double v = B.xyz[a][ids[i]];
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
    int left = 0;
    if (ax >= 0) {
      const double mid = (mn[ax] + mx[ax]) / 2.0;
      int right = 0;
      for (int i = 0; i < n; i++) {
        if (B.xyz[ax][ids[i]] < mid)
          ids[left++] = ids[i];
        else
          B.tmp[right++] = ids[i];
      }
      std::memcpy(ids + left, B.tmp.data(), sizeof(int) * static_cast<size_t>(right));
    }
    if (left == 0 || left == n) left = n / 2;
    bisect_node(B, ids, left, n);
    bisect_node(B, ids + left, n - left, n);
  }
  // every level-(l-1) aggregate of this subtree belongs to the level-l
  // aggregate this node forms (l >= 2; level 1 is mapped cell by cell above)
  for (int l = 2; l <= B.nl; l++)
    if (form[l] >= 0) {
      if (static_cast<int>(B.up[l].size()) < B.count[l - 1]) B.up[l].resize(2 * B.count[l - 1]);
      for (int a = start[l]; a < B.count[l - 1]; a++) B.up[l][a] = form[l];
    }
}

// Colours of a level: two, by the parity of the breadth-first distance from
// the least node of each connected part (neighbours taken in CSR order),
// when that is a proper colouring -- the cell graph of a hex box and the
// coarse graphs of its box-shaped aggregates are bipartite; otherwise the
// greedy colouring in index order.  Nodes >= nint are halos (level 0).
void colour(MgLevel &L, int nint) {
  const int n = L.n;
  std::vector<int> col(n, -1), queue;
  queue.reserve(n);
  bool two = true;
  for (int s = 0; s < n && two; s++) {
    if (col[s] >= 0) continue;
    col[s] = 0;
    queue.assign(1, s);
    for (size_t h = 0; h < queue.size() && two; h++) {
      const int e = queue[h];
      for (int k = L.row[e]; k < L.row[e + 1]; k++) {
        const int j = L.col[k];
        if (j >= nint || j == e) continue;
        if (col[j] < 0) {
          col[j] = 1 - col[e];
          queue.push_back(j);
        } else if (col[j] == col[e]) {
          two = false;
          break;
        }
      }
    }
  }
  int ncol = 0;
  if (two) {
    for (int e = 0; e < n; e++) ncol = std::max(ncol, col[e] + 1);
  } else {
    std::fill(col.begin(), col.end(), 0);
  }
  for (int e = 0; e < n && !two; e++) {
    unsigned long long used = 0;
    for (int k = L.row[e]; k < L.row[e + 1]; k++) {
      const int j = L.col[k];
      if (j < nint && j < e && col[j] < 64) used |= 1ull << col[j];
    }
    int c = 0;
    while (c < 64 && ((used >> c) & 1ull)) c++;
    col[e] = c;
    ncol = std::max(ncol, c + 1);
  }
  L.ncol = ncol;
  L.cptr.assign(ncol + 1, 0);
  for (int e = 0; e < n; e++) L.cptr[col[e] + 1]++;
  for (int c = 0; c < ncol; c++) L.cptr[c + 1] += L.cptr[c];
  std::vector<int> pos(L.cptr.begin(), L.cptr.end());
  L.cperm.assign(n, 0);
  for (int e = 0; e < n; e++) L.cperm[pos[col[e]]++] = e;
}

// the next level's graph and the Galerkin sum lists of aggregate map F.agg
void coarsen(MgLevel &F, MgLevel &C, int nagg, int nint) {
  const int n = F.n;
  C.n = nagg;
  F.mptr.assign(nagg + 1, 0);
  for (int i = 0; i < n; i++) F.mptr[F.agg[i] + 1]++;
  for (int a = 0; a < nagg; a++) F.mptr[a + 1] += F.mptr[a];
  F.mem.assign(n, 0);
  {
    std::vector<int> pos(F.mptr.begin(), F.mptr.end());
    for (int i = 0; i < n; i++) F.mem[pos[F.agg[i]]++] = i;
  }
  C.row.assign(nagg + 1, 0);
  C.col.clear();
  std::vector<int> buf;
  for (int I = 0; I < nagg; I++) {
    buf.clear();
    for (int q = F.mptr[I]; q < F.mptr[I + 1]; q++) {
      const int i = F.mem[q];
      for (int k = F.row[i]; k < F.row[i + 1]; k++) {
        const int j = F.col[k];
        if (j >= nint) continue;
        const int J = F.agg[j];
        if (J != I) buf.push_back(J);
      }
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    C.col.insert(C.col.end(), buf.begin(), buf.end());
    C.row[I + 1] = static_cast<int>(C.col.size());
  }
  C.nnz = static_cast<int>(C.col.size());
  // slot of (I, J) in the coarse row I
  auto slot = [&C](int I, int J) {
    const int *b = C.col.data() + C.row[I], *e = C.col.data() + C.row[I + 1];
    return static_cast<int>(std::lower_bound(b, e, J) - C.col.data());
  };
  F.iptr.assign(nagg + 1, 0);
  F.gptr.assign(C.nnz + 1, 0);
  for (int I = 0; I < nagg; I++)
    for (int q = F.mptr[I]; q < F.mptr[I + 1]; q++) {
      const int i = F.mem[q];
      for (int k = F.row[i]; k < F.row[i + 1]; k++) {
        const int j = F.col[k];
        if (j >= nint) continue;
        const int J = F.agg[j];
        if (J == I)
          F.iptr[I + 1]++;
        else
          F.gptr[slot(I, J) + 1]++;
      }
    }
  for (int I = 0; I < nagg; I++) F.iptr[I + 1] += F.iptr[I];
  for (int s = 0; s < C.nnz; s++) F.gptr[s + 1] += F.gptr[s];
  F.ilist.assign(F.iptr[nagg], 0);
  F.glist.assign(F.gptr[C.nnz], 0);
  std::vector<int> ip(F.iptr.begin(), F.iptr.end()), gp(F.gptr.begin(), F.gptr.end());
  for (int I = 0; I < nagg; I++)
    for (int q = F.mptr[I]; q < F.mptr[I + 1]; q++) {
      const int i = F.mem[q];
      for (int k = F.row[i]; k < F.row[i + 1]; k++) {
        const int j = F.col[k];
        if (j >= nint) continue;
        const int J = F.agg[j];
        if (J == I)
          F.ilist[ip[I]++] = k;
        else
          F.glist[gp[slot(I, J)]++] = k;
      }
    }
}

}  // namespace

int mg_build(int ne, const int *row, const int *nbr, const double *xc, const double *yc,
             const double *zc, int maxagg, int ncoarse, MgHier *out) {
  out->lv.clear();
  out->lv.emplace_back();
  MgLevel &L0 = out->lv[0];
  L0.n = ne;
  L0.row.assign(row, row + ne + 1);
  L0.nnz = row[ne];
  L0.col.assign(nbr, nbr + L0.nnz);
  colour(L0, ne);
  Bisect B;
  B.xyz[0] = xc;
  B.xyz[1] = yc;
  B.xyz[2] = zc;
  B.cap.assign(1, 1);
  long capl = 1;
  for (int l = 1; l < kMgMaxLevels; l++) {
    capl *= maxagg;
    if (capl > static_cast<long>(ne) * maxagg || capl > (1L << 30)) break;
    B.cap.push_back(capl);
    B.nl = l;
  }
  B.count.assign(B.nl + 1, 0);
  B.up.assign(B.nl + 1, std::vector<int>());
  if (B.nl >= 1) B.up[1].assign(ne, 0);
  B.tmp.assign(ne, 0);
  std::vector<int> ids(ne);
  for (int e = 0; e < ne; e++) ids[e] = e;
  if (ne > 0) bisect_node(B, ids.data(), ne, 2L * (1L << 30));
  for (int l = 1; l <= B.nl; l++) {
    const int nagg = B.count[l];
    MgLevel &F = out->lv[l - 1];
    if (F.n <= ncoarse || nagg <= 1 || nagg >= F.n) break;
    F.agg.assign(B.up[l].begin(), B.up[l].begin() + F.n);
    out->lv.emplace_back();
    MgLevel &F2 = out->lv[l - 1];  // emplace_back may move the levels
    MgLevel &C = out->lv[l];
    coarsen(F2, C, nagg, l == 1 ? ne : F2.n);
    colour(C, nagg);
  }
  return CFDL_OK;
}

}  // namespace cfdl

// Copy out one array of the hierarchy the MGCG pc solve builds for this
// mesh: "n" (1), "row", "col", "cptr", "cperm", and towards the next level
// "agg", "mptr", "mem", "iptr", "ilist", "gptr", "glist"; level 0's col is
// the mesh CSR (neighbours >= ne are halos).  *count receives the array's
// length (-1 for an unknown name or level).
extern "C" int cfdl_mesh_mg_get(const cfdl_mesh *m, int level, const char *name, int *dst,
                                int64_t cap, int64_t *count) {
  if (!m || !name || !count || m->ne <= 0) {
    cfdl::set_error("cfdl_mesh_mg_get: bad arguments");
    return CFDL_EINVAL;
  }
  const int ne = m->ne;
  std::vector<int> row(ne + 1), nbr(m->L);
  for (int e = 0; e <= ne; e++) row[e] = m->ef2nb_idx[e] - 1;
  for (int k = 0; k < m->L; k++) nbr[k] = cfdl::pk_elem(m->ef2nb[k]) - 1;
  cfdl::MgHier h;
  if (int rc = cfdl::mg_build(ne, row.data(), nbr.data(), m->xc.data(), m->yc.data(),
                              m->zc.data(), 8, 64, &h))
    return rc;
  *count = -1;
  if (level < 0 || level >= static_cast<int>(h.lv.size())) return CFDL_OK;
  const cfdl::MgLevel &L = h.lv[level];
  std::vector<int> one(1, L.n);
  const std::vector<int> *v = nullptr;
  const std::string nm(name);
  if (nm == "n") v = &one;
  else if (nm == "row") v = &L.row;
  else if (nm == "col") v = &L.col;
  else if (nm == "cptr") v = &L.cptr;
  else if (nm == "cperm") v = &L.cperm;
  else if (level + 1 < static_cast<int>(h.lv.size())) {
    if (nm == "agg") v = &L.agg;
    else if (nm == "mptr") v = &L.mptr;
    else if (nm == "mem") v = &L.mem;
    else if (nm == "iptr") v = &L.iptr;
    else if (nm == "ilist") v = &L.ilist;
    else if (nm == "gptr") v = &L.gptr;
    else if (nm == "glist") v = &L.glist;
  }
  if (!v) return CFDL_OK;
  *count = static_cast<int64_t>(v->size());
  if (dst) std::memcpy(dst, v->data(), sizeof(int) * static_cast<size_t>(std::min<int64_t>(cap, *count)));
  return CFDL_OK;
}

