// cfdl_setup.cpp -- host-side pre-processing of libcfdl_hip.so.
//
// Builds, from CGNS-style element lists, the arrays the reference's setup
// produces (src/setup/*): packed face connectivity, face and cell geometry,
// and the subdomain partition with its local order.  Not on the hot path,
// but its outputs fix the hot path's arithmetic (owner side of each face,
// CSR row order, Gauss-Seidel order of the subdomains), so each rule below
// follows the reference exactly; the algorithms are near-linear where the
// reference's are quadratic (vertex-sweep matching, quicksort on runs).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>

#include "cfdl_internal.h"

#define CALL_RC(x)          \
  do {                      \
    int rc_ = (x);          \
    if (rc_) return rc_;    \
  } while (0)

namespace cfdl {

// CGNS hexahedron faces (vertex slots), src/modules/mod_util.f90:78-85
static const int kHexFaces[6][4] = {{1, 4, 3, 2}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                    {3, 4, 8, 7}, {1, 5, 8, 4}, {5, 6, 7, 8}};

static int nface_of(int t) { return t == kHexa8 ? 6 : (t == kQuad4 ? 1 : -1); }

static void face_vertices(const int *ev, int t, int f, int *lst) {
  if (t == kHexa8)
    for (int l = 0; l < 4; l++) lst[l] = ev[kHexFaces[f - 1][l] - 1];
  else
    for (int l = 0; l < 4; l++) lst[l] = ev[l];
}

// Face numbering and ownership (find_element_nb, mod_mg_lvl_uns.f90:283-488).
// The reference sweeps vertices in ascending order and, at each vertex, pairs
// the incident (element, face) entries in ascending (element, face) order; a
// face gets its number when first paired, i.e. at its smallest vertex, and
// the lower-numbered element owns it.  Bucketing element faces by their
// smallest vertex and ordering each bucket's matches by (owner element,
// owner face) reproduces that numbering in O(n).
static int build_connectivity(cfdl_mesh *m) {
  const int ne = m->ne, nelem = m->nelem, nvx = m->nvx;
  m->ef2nb_idx.assign(ne + 1, 0);
  m->ef2nb_idx[0] = 1;
  std::vector<int> bstart(static_cast<size_t>(nvx) + 2, 0);
  auto for_faces = [&](auto fn) -> int {
    for (int s = 0; s < m->nsec; s++) {
      const int t = m->etype[s], nfc = nface_of(t);
      if (nfc < 0) {
        set_error("unsupported element type " + std::to_string(t) + " (HEXA_8/QUAD_4 only)");
        return CFDL_EINVAL;
      }
      for (int e = m->esec[2 * s]; e <= m->esec[2 * s + 1]; e++)
        for (int f = 1; f <= nfc; f++) {
          int lst[4];
          face_vertices(&m->e2vx[8 * static_cast<size_t>(e - 1)], t, f, lst);
          fn(e, f, std::min(std::min(lst[0], lst[1]), std::min(lst[2], lst[3])));
        }
    }
    return CFDL_OK;
  };
  for (int s = 0; s < m->nsec; s++)
    for (int e = m->esec[2 * s]; e <= m->esec[2 * s + 1] && e <= ne; e++)
      m->ef2nb_idx[e] = m->ef2nb_idx[e - 1] + nface_of(m->etype[s]);
  CALL_RC(for_faces([&](int, int, int v) { bstart[v + 1]++; }));
  for (int v = 1; v <= nvx + 1; v++) bstart[v] += bstart[v - 1];
  std::vector<int64_t> inc(bstart[nvx + 1]);
  {
    std::vector<int> fill(bstart.begin(), bstart.end() - 1);
    CALL_RC(for_faces([&](int e, int f, int v) {
      inc[fill[v]++] = (static_cast<int64_t>(e) << 3) | f;
    }));
  }
  const int L = m->L;
  m->ef2nb.assign(2 * static_cast<size_t>(L), 0);
  m->s2g.assign(m->nf, 0);
  m->bs.assign(m->nbf, 0);
  int *g1 = m->ef2nb.data(), *g2 = m->ef2nb.data() + L;
  struct Ent {
    std::array<int, 4> key;
    int e, f;
  };
  struct Pair {
    int e1, f1, e2, f2;
  };
  std::vector<Ent> ent;
  std::vector<Pair> pairs;
  int nf = 0;
  for (int v = 1; v <= nvx; v++) {
    ent.clear();
    pairs.clear();
    for (int q = bstart[v]; q < bstart[v + 1]; q++) {
      Ent a;
      a.e = static_cast<int>(inc[q] >> 3);
      a.f = static_cast<int>(inc[q] & 7);
      const int t = a.e <= ne ? kHexa8 : kQuad4;
      face_vertices(&m->e2vx[8 * static_cast<size_t>(a.e - 1)], t, a.f, a.key.data());
      std::sort(a.key.begin(), a.key.end());
      ent.push_back(a);
    }
    std::sort(ent.begin(), ent.end(), [](const Ent &a, const Ent &b) {
      if (a.key != b.key) return a.key < b.key;
      if (a.e != b.e) return a.e < b.e;
      return a.f < b.f;
    });
    for (size_t k = 0; k < ent.size();) {
      size_t j = k + 1;
      while (j < ent.size() && ent[j].key == ent[k].key) j++;
      if (j - k != 2) {
        set_error("face shared by " + std::to_string(j - k) + " elements (element " +
                  std::to_string(ent[k].e) + ")");
        return CFDL_EINVAL;
      }
      if (ent[k].e > ne) {
        set_error("boundary element matched another boundary element");
        return CFDL_EINVAL;
      }
      pairs.push_back({ent[k].e, ent[k].f, ent[k + 1].e, ent[k + 1].f});
      k = j;
    }
    std::sort(pairs.begin(), pairs.end(), [](const Pair &a, const Pair &b) {
      return a.e1 != b.e1 ? a.e1 < b.e1 : a.f1 < b.f1;
    });
    for (const Pair &p : pairs) {
      const int fg = ++nf;
      if (fg > m->nf) break;
      const int i1 = m->ef2nb_idx[p.e1 - 1] + p.f1 - 2;
      m->s2g[fg - 1] = pack(p.e1, p.f1);
      if (p.e2 > ne) {
        g1[i1] = pack(p.e2, 0);
        g2[i1] = fg;
        m->bs[p.e2 - ne - 1] = pack(p.e1, p.f1);
      } else {
        const int i2 = m->ef2nb_idx[p.e2 - 1] + p.f2 - 2;
        g1[i1] = pack(p.e2, p.f2);
        g1[i2] = pack(p.e1, p.f1);
        g2[i1] = fg;
        g2[i2] = -fg;
      }
    }
  }
  if (nf != m->nf) {
    set_error("face count " + std::to_string(nf) + " differs from " + std::to_string(m->nf));
    return CFDL_EINVAL;
  }
  // edge/corner flag: negative bs on halos of cells with several boundary
  // faces (mod_mg_lvl_uns.f90:422-433); only |bs| is used downstream.
  for (int e1 = ne + 1; e1 <= nelem; e1++) {
    int &b1 = m->bs[e1 - ne - 1];
    if (b1 < 0) continue;
    const int e = pk_elem(b1), f = pk_face(b1);
    for (int idx = m->ef2nb_idx[e - 1]; idx <= m->ef2nb_idx[e] - 1; idx++) {
      if (idx - m->ef2nb_idx[e - 1] + 1 == f) continue;
      const int en = pk_elem(g1[idx - 1]), fn = pk_face(g1[idx - 1]);
      if (fn == 0) {
        b1 = -std::abs(b1);
        m->bs[en - ne - 1] = -m->bs[en - ne - 1];
      }
    }
  }
  return CFDL_OK;
}

static int section_of(const cfdl_mesh *m, int e) {
  for (int s = 0; s < m->nsec; s++)
    if (e <= m->esec[2 * s + 1]) return s;
  return m->nsec - 1;
}

// Face area vectors and centroids (calc_aip_xyzip_uns, calc_aip_xyzip.f90:
// 25-72) and cell volumes / centroids with halo centres at face centroids
// (calc_vol_cv_centers_uns, calc_vol_cv_centers.f90:22-59).  The expression
// order is the reference's so that the values are bit-identical.
static void build_geometry(cfdl_mesh *m) {
  const int nf = m->nf, ne = m->ne;
  m->aip.assign(3 * static_cast<size_t>(nf), 0.0);
  m->rip.assign(3 * static_cast<size_t>(nf), 0.0);
  m->xc.assign(m->nelem, 0.0);
  m->yc.assign(m->nelem, 0.0);
  m->zc.assign(m->nelem, 0.0);
  m->vol.assign(ne, 0.0);
  for (int k = 0; k < nf; k++) {
    const int e = pk_elem(m->s2g[k]), fl = pk_face(m->s2g[k]);
    int lst[4];
    face_vertices(&m->e2vx[8 * static_cast<size_t>(e - 1)], m->etype[section_of(m, e)], fl, lst);
    double r[4][3];
    for (int l = 0; l < 4; l++) {
      r[l][0] = m->x[lst[l] - 1];
      r[l][1] = m->y[lst[l] - 1];
      r[l][2] = m->z[lst[l] - 1];
    }
    double *aip = &m->aip[3 * static_cast<size_t>(k)];
    double sc[3][3];
    for (int tri = 0; tri < 2; tri++) {
      const double *p1 = r[0], *p2 = tri == 0 ? r[1] : r[2], *p3 = tri == 0 ? r[2] : r[3];
      double d1[3], d2[3], av[3], sub[3];
      for (int i = 0; i < 3; i++) {
        d1[i] = p2[i] - p1[i];
        d2[i] = p3[i] - p1[i];
      }
      av[0] = d1[1] * d2[2] - d1[2] * d2[1];
      av[1] = d1[2] * d2[0] - d1[0] * d2[2];
      av[2] = d1[0] * d2[1] - d1[1] * d2[0];
      for (int i = 0; i < 3; i++) av[i] = 0.5 * av[i];
      for (int i = 0; i < 3; i++) sub[i] = (p1[i] + p2[i] + p3[i]) / 3.0;
      for (int i = 0; i < 3; i++) {
        aip[i] = tri == 0 ? av[i] : aip[i] + av[i];
        for (int j = 0; j < 3; j++) sc[i][j] = tri == 0 ? sub[i] * av[j] : sc[i][j] + sub[i] * av[j];
      }
    }
    const double dd = aip[0] * aip[0] + aip[1] * aip[1] + aip[2] * aip[2];
    for (int i = 0; i < 3; i++)
      m->rip[3 * static_cast<size_t>(k) + i] = (sc[i][0] * aip[0] + sc[i][1] * aip[1] + sc[i][2] * aip[2]) / dd;
  }
  const int *g1 = m->ef2nb.data(), *g2 = m->ef2nb.data() + m->L;
  for (int e = 1; e <= ne; e++) {
    double gc[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 8; k++) {
      const int v = m->e2vx[8 * static_cast<size_t>(e - 1) + k];
      gc[0] = gc[0] + m->x[v - 1];
      gc[1] = gc[1] + m->y[v - 1];
      gc[2] = gc[2] + m->z[v - 1];
    }
    for (int i = 0; i < 3; i++) gc[i] = gc[i] / 8.0;
    double sum_vol = 0.0, rc[3] = {0.0, 0.0, 0.0};
    for (int idx = m->ef2nb_idx[e - 1]; idx <= m->ef2nb_idx[e] - 1; idx++) {
      int gf = g2[idx - 1];
      const double sg = gf >= 0 ? 1.0 : -1.0;
      gf = std::abs(gf);
      const double *cs = &m->rip[3 * static_cast<size_t>(gf - 1)];
      const double *a = &m->aip[3 * static_cast<size_t>(gf - 1)];
      const int enb = pk_elem(g1[idx - 1]);
      if (pk_face(g1[idx - 1]) == 0) {
        m->xc[enb - 1] = cs[0];
        m->yc[enb - 1] = cs[1];
        m->zc[enb - 1] = cs[2];
      }
      double h[3];
      for (int i = 0; i < 3; i++) h[i] = cs[i] - gc[i];
      const double sub_vol = ((sg * a[0]) * h[0] + (sg * a[1]) * h[1] + (sg * a[2]) * h[2]) / 3.0;
      sum_vol = sum_vol + sub_vol;
      for (int i = 0; i < 3; i++) rc[i] = rc[i] + (0.25 * gc[i] + 0.75 * cs[i]) * sub_vol;
    }
    for (int i = 0; i < 3; i++) rc[i] = rc[i] / sum_vol;
    m->xc[e - 1] = rc[0];
    m->yc[e - 1] = rc[1];
    m->zc[e - 1] = rc[2];
    m->vol[e - 1] = sum_vol;
  }
}

int mesh_build(cfdl_mesh *m) {
  int ne = 0, nbf = 0, nfc = 0;
  for (int s = 0; s < m->nsec; s++) {
    const int cnt = m->esec[2 * s + 1] - m->esec[2 * s] + 1;
    if (s > 0 && m->esec[2 * s] != m->esec[2 * s - 1] + 1) {
      set_error("sections must be sorted and contiguous");
      return CFDL_EINVAL;
    }
    if (m->etype[s] == kHexa8) {
      ne += cnt;
      nfc += 6 * cnt;
    } else if (m->etype[s] == kQuad4) {
      nbf += cnt;
    } else {
      set_error("unsupported element type " + std::to_string(m->etype[s]));
      return CFDL_EINVAL;
    }
  }
  m->ne = ne;
  m->nbf = nbf;
  m->nelem = ne + nbf;
  m->nf = (nfc + nbf) / 2;
  m->L = 2 * m->nf - nbf;
  if (static_cast<int64_t>(m->nelem) >= (int64_t{1} << 26)) {
    set_error("mesh too large for 5-bit packed face indices");
    return CFDL_EINVAL;
  }
  int rc = build_connectivity(m);
  if (rc) return rc;
  build_geometry(m);
  return CFDL_OK;
}

// ---------------------------------------------------------------------------
// Partition: longest-axis midpoint bisection of the cell-centre bounding box
// (seed_gen_t%grow/split_leaf, src/setup/mod_agglomeration.f90:331-545).  The
// reference keeps leaves in an in-order threaded binary tree and cell lists
// in insertion order; which leaves split in which pass, the leaf numbering
// and the group volumes (summed in list order, compared with a threshold)
// all depend on that, so the same threading is kept.
// ---------------------------------------------------------------------------
namespace {
struct Box {
  int parent_group_id = 0, group_id = -1;
  double rmin[3], rmax[3], group_vol = 0.0;
  std::vector<int> cells;
};
struct Node {
  int id;
  bool lthread = true, rthread = true;
  Node *left = nullptr, *right = nullptr;
  std::unique_ptr<Box> box;
};
struct Tree {
  std::vector<std::unique_ptr<Node>> nodes;
  Node *root = nullptr, *begin = nullptr, *end = nullptr;
  int last_id = 1, nleaf = 1, target = 1;
  double vol_ave = 0.0;
  const double *vol, *c[3];

  Node *make(std::unique_ptr<Box> b) {
    nodes.emplace_back(new Node());
    Node *n = nodes.back().get();
    n->id = ++last_id;
    n->box = std::move(b);
    return n;
  }
  static bool leaf(const Node *n) { return n->lthread && n->rthread; }
  bool next(Node *&it) const {  // iterate_next
    if (it == end) return false;
    if (!it->rthread) {
      it = it->right;
      while (!it->lthread) it = it->left;
    } else {
      it = it->right;
    }
    return true;
  }
  bool next_leaf(Node *&it) const {  // goto_next_leaf
    bool ok = false;
    while (next(it)) {
      ok = leaf(it);
      if (ok) break;
    }
    return ok;
  }
  bool first_leaf(Node *&it) const {  // goto_1st_leaf
    it = begin;
    if (!leaf(it)) return next_leaf(it);
    return true;
  }
  void split_node(Node *n, std::unique_ptr<Box> lb, std::unique_ptr<Box> rb) {
    Node *oldl = n->left, *oldr = n->right;
    if (lb) {
      n->lthread = false;
      Node *L = make(std::move(lb));
      n->left = L;
      L->left = oldl;
      L->right = n;
      if (n == begin) begin = L;
    }
    if (rb) {
      n->rthread = false;
      Node *R = make(std::move(rb));
      n->right = R;
      R->left = n;
      R->right = oldr;
      if (n == end) end = R;
    }
  }
  int split(Node *it, double thr, int *nchild) {  // split_leaf
    *nchild = 0;
    if (!leaf(it)) return 0;
    Box *b = it->box.get();
    if (b->group_vol / vol_ave < thr) return 0;
    if (b->cells.size() == 1) return 0;
    double d = 0.0;
    int ax = -1;
    for (int i = 0; i < 3; i++)
      if (std::fabs(b->rmin[i] - b->rmax[i]) > d) {
        d = std::fabs(b->rmin[i] - b->rmax[i]);
        ax = i;
      }
    if (ax < 0) return -1;
    d = (b->rmin[ax] + b->rmax[ax]) / 2.0;
    std::unique_ptr<Box> L(new Box()), R(new Box());
    for (int i = 0; i < 3; i++) {
      L->rmin[i] = R->rmin[i] = 1e20;
      L->rmax[i] = R->rmax[i] = -1e20;
    }
    for (int e : b->cells) {
      Box *t = c[ax][e - 1] < d ? L.get() : R.get();
      t->cells.push_back(e);
      t->group_vol = t->group_vol + vol[e - 1] * 1.0;
      for (int i = 0; i < 3; i++) {
        t->rmin[i] = std::fmin(t->rmin[i], c[i][e - 1]);
        t->rmax[i] = std::fmax(t->rmax[i], c[i][e - 1]);
      }
    }
    b->cells.clear();
    const int pg = b->group_id == -1 ? b->parent_group_id : b->group_id;
    L->parent_group_id = R->parent_group_id = pg;
    const bool hl = !L->cells.empty(), hr = !R->cells.empty();
    *nchild = (hl ? 1 : 0) + (hr ? 1 : 0);
    split_node(it, hl ? std::move(L) : nullptr, hr ? std::move(R) : nullptr);
    return 1;
  }
};

// Positions of one key value, as a 64-ary bit tree: set/clear and the
// nearest set position at or below / at or above a query point in a few
// word operations.
class PosSet {
 public:
  explicit PosSet(int64_t n) : n_(n) {
    int64_t m = n;
    do {
      m = (m + 63) / 64;
      lv_.emplace_back(static_cast<size_t>(m), 0ull);
    } while (m > 1);
  }
  void set(int64_t p) {
    for (auto &w : lv_) {
      const uint64_t bit = 1ull << (p & 63);
      const bool was_empty = w[p >> 6] == 0;
      w[p >> 6] |= bit;
      if (!was_empty) break;
      p >>= 6;
    }
  }
  void clear(int64_t p) {
    for (auto &w : lv_) {
      w[p >> 6] &= ~(1ull << (p & 63));
      if (w[p >> 6] != 0) break;
      p >>= 6;
    }
  }
  // largest set position <= p, or -1
  int64_t prev(int64_t p) const {
    if (p < 0) return -1;
    if (p >= n_) p = n_ - 1;
    size_t l = 0;
    for (;;) {
      const uint64_t w = lv_[l][p >> 6] & (~0ull >> (63 - (p & 63)));
      if (w) {
        p = (p & ~int64_t{63}) | (63 - __builtin_clzll(w));
        break;
      }
      p = (p >> 6) - 1;
      if (p < 0) return -1;
      if (++l == lv_.size()) return -1;
    }
    while (l > 0) {
      --l;
      p = p * 64 + (63 - __builtin_clzll(lv_[l][p]));
    }
    return p;
  }
  // smallest set position >= p, or n
  int64_t next(int64_t p) const {
    if (p >= n_) return n_;
    if (p < 0) p = 0;
    size_t l = 0;
    for (;;) {
      const size_t wi = static_cast<size_t>(p >> 6);
      if (wi >= lv_[l].size()) return n_;
      const uint64_t w = lv_[l][wi] & (~0ull << (p & 63));
      if (w) {
        p = (p & ~int64_t{63}) | __builtin_ctzll(w);
        break;
      }
      p = (p >> 6) + 1;
      if (++l == lv_.size()) return n_;
    }
    while (l > 0) {
      --l;
      p = p * 64 + __builtin_ctzll(lv_[l][p]);
    }
    return p;
  }

 private:
  int64_t n_;
  std::vector<std::vector<uint64_t>> lv_;
};

// Darel Rex Finley's non-recursive quicksort as used by qsort_key_nRec
// (src/modules/mod_util.f90:1683-1730), pivot = first key of the range.
// The moves and the range stack are the algorithm's own, so the output
// permutation is identical; only its two scan loops ("skip keys >= pivot
// from the right", "skip keys <= pivot from the left") jump straight to
// their stopping position through one PosSet per key value.  On the few
// distinct keys sorted here the plain scans cost O(m) per partition and
// O(m^2) per range; the jumps make the whole sort O(n * nkeys).
void finley_sort(int *key, int *b, int n, int kmin, int kmax) {
  const int nk = kmax - kmin + 1;
  std::vector<PosSet> pos;
  pos.reserve(nk);
  for (int v = 0; v < nk; v++) pos.emplace_back(n);
  for (int k = 0; k < n; k++) pos[key[k] - kmin].set(k);
  auto write = [&](int p, int val, int bv) {
    pos[key[p] - kmin].clear(p);
    key[p] = val;
    b[p] = bv;
    pos[val - kmin].set(p);
  };
  // last position in (lo, hi] whose key < piv, else lo
  auto scan_right = [&](int lo, int hi, int piv) {
    int64_t best = lo;
    for (int v = kmin; v < piv; v++) {
      const int64_t q = pos[v - kmin].prev(hi);
      if (q > lo && q > best) best = q;
    }
    return static_cast<int>(best);
  };
  // first position in [lo, hi) whose key > piv, else hi
  auto scan_left = [&](int lo, int hi, int piv) {
    int64_t best = hi;
    for (int v = piv + 1; v <= kmax; v++) {
      const int64_t q = pos[v - kmin].next(lo);
      if (q < hi && q < best) best = q;
    }
    return static_cast<int>(best);
  };
  std::vector<int> beg(1, 0), end(1, n);
  int i = 0;
  while (i >= 0) {
    int L = beg[i], R = end[i] - 1;
    if (L < R) {
      const int piv = key[L], b0 = b[L];
      while (L < R) {
        R = scan_right(L, R, piv);
        if (L < R) {
          write(L, key[R], b[R]);
          L++;
        }
        L = scan_left(L, R, piv);
        if (L < R) {
          write(R, key[L], b[L]);
          R--;
        }
      }
      write(L, piv, b0);
      if (static_cast<int>(beg.size()) <= i + 1) {
        beg.resize(i + 2);
        end.resize(i + 2);
      }
      beg[i + 1] = L + 1;
      end[i + 1] = end[i];
      end[i] = L;
      i++;
      if (end[i] - beg[i] > end[i - 1] - beg[i - 1]) {
        std::swap(beg[i], beg[i - 1]);
        std::swap(end[i], end[i - 1]);
      }
    } else {
      i--;
    }
  }
}
}  // namespace

}  // namespace cfdl

using namespace cfdl;

extern "C" int cfdl_mesh_partition(cfdl_mesh *m, int nsubd) {
  if (!m || nsubd < 1 || nsubd > m->ne) {
    set_error("cfdl_mesh_partition: bad arguments");
    return CFDL_EINVAL;
  }
  const int ne = m->ne;
  m->nsubd = nsubd;
  m->gf2g.assign(ne, 1);
  m->g2gf.resize(ne);
  std::iota(m->g2gf.begin(), m->g2gf.end(), 1);
  m->g2gf_idx.assign(nsubd + 1, 0);
  if (nsubd == 1) {
    m->g2gf_idx[0] = 1;
    m->g2gf_idx[1] = ne + 1;
    return CFDL_OK;
  }
  // generate_seeds([nsubd,1], 2, ...): mod_mg_lvl_uns.f90:95-118
  Tree t;
  t.vol = m->vol.data();
  t.c[0] = m->xc.data();
  t.c[1] = m->yc.data();
  t.c[2] = m->zc.data();
  std::unique_ptr<Box> rb(new Box());
  for (int i = 0; i < 3; i++) {
    rb->rmin[i] = *std::min_element(t.c[i], t.c[i] + ne);
    rb->rmax[i] = *std::max_element(t.c[i], t.c[i] + ne);
  }
  rb->cells.resize(ne);
  std::iota(rb->cells.begin(), rb->cells.end(), 1);
  for (int e = 0; e < ne; e++) t.vol_ave = t.vol_ave + m->vol[e] * 1.0;
  rb->group_vol = 2.1 * t.vol_ave;
  rb->group_id = 1;
  t.nodes.emplace_back(new Node());
  t.root = t.nodes.back().get();
  t.root->id = 1;
  t.root->box = std::move(rb);
  t.root->left = t.root->right = t.root;
  t.begin = t.end = t.root;
  // grow(nsubd, gf2g(2), 2), mod_agglomeration.f90:497-528
  t.vol_ave = t.vol_ave * static_cast<double>(t.target) / static_cast<double>(nsubd);
  t.target = nsubd;
  int nleaf0 = t.nleaf, nchild;
  double thr = 2.0;
  Node *it;
  int guard = 0;
  while (t.nleaf < nsubd) {
    if (t.first_leaf(it)) {
      for (;;) {
        const int s = t.split(it, thr, &nchild);
        if (s < 0) {
          set_error("partition: degenerate bounding box");
          return CFDL_EINVAL;
        }
        if (s) {
          t.next_leaf(it);
          t.nleaf = t.nleaf + nchild - 1;
          if (t.nleaf == nsubd) break;
        }
        if (!t.next_leaf(it)) break;
      }
    }
    if (t.nleaf == nleaf0) thr = thr * 0.75;
    nleaf0 = t.nleaf;
    if (++guard > 100000) {
      set_error("partition did not reach the requested number of subdomains");
      return CFDL_EINVAL;
    }
  }
  int nleaf = 0;
  if (t.first_leaf(it)) {
    for (;;) {
      Box *b = it->box.get();
      ++nleaf;
      if (b->group_id != -1) b->parent_group_id = b->group_id;
      b->group_id = nleaf;
      if (!t.next_leaf(it)) break;
    }
  }
  // grow(ne, gf2g(1), 1): every cell takes its leaf's number
  if (t.first_leaf(it)) {
    for (;;) {
      for (int e : it->box->cells) m->gf2g[e - 1] = it->box->group_id;
      if (!t.next_leaf(it)) break;
    }
  }
  // add_transformation_bt first level, mod_mg_lvl_uns.f90:873-895
  std::vector<int> key(m->gf2g);
  finley_sort(key.data(), m->g2gf.data(), ne, 1, nsubd);
  if (key[ne - 1] != nsubd) {
    set_error("partition produced a different number of groups");
    return CFDL_EINVAL;
  }
  int g = 0;
  for (int idx = 1; idx <= ne; idx++)
    if (g != key[idx - 1]) {
      g = key[idx - 1];
      m->g2gf_idx[g - 1] = idx;
    }
  m->g2gf_idx[nsubd] = ne + 1;
  for (int c = 0; c < nsubd; c++)
    if (m->g2gf_idx[c] == 0) {
      set_error("partition produced an empty subdomain");
      return CFDL_EINVAL;
    }
  return CFDL_OK;
}

extern "C" int cfdl_mesh_from_elements(int nvx, const double *x, const double *y,
                                       const double *z, int nsec, const int *etype,
                                       const int *esec, const int *e2vx, cfdl_mesh **out) {
  if (!out || nvx <= 0 || nsec <= 0 || !x || !y || !z || !etype || !esec || !e2vx) {
    set_error("cfdl_mesh_from_elements: bad arguments");
    return CFDL_EINVAL;
  }
  std::unique_ptr<cfdl_mesh> m(new cfdl_mesh());
  m->nvx = nvx;
  m->nsec = nsec;
  m->x.assign(x, x + nvx);
  m->y.assign(y, y + nvx);
  m->z.assign(z, z + nvx);
  m->etype.assign(etype, etype + nsec);
  m->esec.assign(esec, esec + 2 * nsec);
  const int nelem = esec[2 * nsec - 1];
  m->e2vx.assign(e2vx, e2vx + 8 * static_cast<size_t>(nelem));
  for (int v : m->e2vx)
    if (v < 0 || v > nvx) {
      set_error("e2vx holds a vertex id out of range");
      return CFDL_EINVAL;
    }
  int rc = mesh_build(m.get());
  if (rc) return rc;
  *out = m.release();
  return CFDL_OK;
}

// splitmix64 (Steele, Lea, Flood 2014): the jitter stream of the generator
static uint64_t splitmix64(uint64_t &s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

extern "C" int cfdl_mesh_box(int nx, int ny, int nz, double lx, double ly, double lz,
                             double jitter, uint64_t seed, cfdl_mesh **out) {
  if (!out || nx < 1 || ny < 1 || nz < 1 || jitter < 0.0 || jitter >= 0.5) {
    set_error("cfdl_mesh_box: bad arguments");
    return CFDL_EINVAL;
  }
  const int64_t ne64 = int64_t{nx} * ny * nz;
  if (ne64 + 2 * (int64_t{nx} * ny + int64_t{nx} * nz + int64_t{ny} * nz) >= (int64_t{1} << 26)) {
    set_error("cfdl_mesh_box: mesh too large for 5-bit packed face indices");
    return CFDL_EINVAL;
  }
  std::unique_ptr<cfdl_mesh> m(new cfdl_mesh());
  const int NX = nx + 1, NY = ny + 1, NZ = nz + 1;
  m->nvx = NX * NY * NZ;
  m->x.resize(m->nvx);
  m->y.resize(m->nvx);
  m->z.resize(m->nvx);
  for (int k = 0; k < NZ; k++)
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++) {
        const size_t v = i + static_cast<size_t>(NX) * (j + static_cast<size_t>(NY) * k);
        m->x[v] = lx * i / nx;
        m->y[v] = ly * j / ny;
        m->z[v] = lz * k / nz;
      }
  if (jitter > 0.0) {
    uint64_t s = seed;
    const double hx = lx / nx, hy = ly / ny, hz = lz / nz;
    for (int k = 1; k < nz; k++)
      for (int j = 1; j < ny; j++)
        for (int i = 1; i < nx; i++) {
          const size_t v = i + static_cast<size_t>(NX) * (j + static_cast<size_t>(NY) * k);
          double u[3];
          for (int q = 0; q < 3; q++)
            u[q] = static_cast<double>(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
          m->x[v] = m->x[v] + (2.0 * u[0] - 1.0) * (jitter * hx);
          m->y[v] = m->y[v] + (2.0 * u[1] - 1.0) * (jitter * hy);
          m->z[v] = m->z[v] + (2.0 * u[2] - 1.0) * (jitter * hz);
        }
  }
  auto vid = [&](int a, int b, int c) { return 1 + a + NX * (b + NY * c); };
  const int ne = static_cast<int>(ne64);
  const int nbf = 2 * (nx * ny + nx * nz + ny * nz);
  m->e2vx.assign(8 * static_cast<size_t>(ne + nbf), 0);
  int *ev = m->e2vx.data();
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int *q = ev + 8 * (i + static_cast<size_t>(nx) * (j + static_cast<size_t>(ny) * k));
        q[0] = vid(i, j, k); q[1] = vid(i + 1, j, k); q[2] = vid(i + 1, j + 1, k);
        q[3] = vid(i, j + 1, k); q[4] = vid(i, j, k + 1); q[5] = vid(i + 1, j, k + 1);
        q[6] = vid(i + 1, j + 1, k + 1); q[7] = vid(i, j + 1, k + 1);
      }
  size_t e = ne;
  m->etype.push_back(kHexa8);
  m->esec.push_back(1);
  m->esec.push_back(ne);
  m->secname.push_back("cells");
  auto section = [&](const char *name, int n1, int n2, auto quad) {
    const int first = static_cast<int>(e) + 1;
    for (int b = 0; b < n2; b++)
      for (int a = 0; a < n1; a++) {
        quad(a, b, ev + 8 * e);
        e++;
      }
    m->etype.push_back(kQuad4);
    m->esec.push_back(first);
    m->esec.push_back(static_cast<int>(e));
    m->secname.push_back(name);
  };
  section("bottom", nx, ny, [&](int a, int b, int *q) {
    q[0] = vid(a, b, 0); q[1] = vid(a, b + 1, 0); q[2] = vid(a + 1, b + 1, 0); q[3] = vid(a + 1, b, 0);
  });
  section("south", nx, nz, [&](int a, int b, int *q) {
    q[0] = vid(a, 0, b); q[1] = vid(a + 1, 0, b); q[2] = vid(a + 1, 0, b + 1); q[3] = vid(a, 0, b + 1);
  });
  section("east", ny, nz, [&](int a, int b, int *q) {
    q[0] = vid(nx, a, b); q[1] = vid(nx, a + 1, b); q[2] = vid(nx, a + 1, b + 1); q[3] = vid(nx, a, b + 1);
  });
  section("north", nx, nz, [&](int a, int b, int *q) {
    q[0] = vid(a + 1, ny, b); q[1] = vid(a, ny, b); q[2] = vid(a, ny, b + 1); q[3] = vid(a + 1, ny, b + 1);
  });
  section("west", ny, nz, [&](int a, int b, int *q) {
    q[0] = vid(0, a + 1, b); q[1] = vid(0, a, b); q[2] = vid(0, a, b + 1); q[3] = vid(0, a + 1, b + 1);
  });
  section("top", nx, ny, [&](int a, int b, int *q) {
    q[0] = vid(a, b, nz); q[1] = vid(a + 1, b, nz); q[2] = vid(a + 1, b + 1, nz); q[3] = vid(a, b + 1, nz);
  });
  m->nsec = static_cast<int>(m->etype.size());
  int rc = mesh_build(m.get());
  if (rc) return rc;
  *out = m.release();
  return CFDL_OK;
}

// Cell renumbering for the distorted-channel input (no counterpart in the
// reference, which keeps the element order of its CGNS file): a seeded
// uniform permutation of the cells (Fisher-Yates from the last position:
// position i swaps with j = splitmix64 % (i + 1)), then the reverse
// Cuthill-McKee order of the interior-face graph of the shuffled mesh.  RCM
// here: every component starts at its lowest-numbered cell of least degree
// (interior neighbours); a breadth-first walk appends each visited cell's
// unvisited neighbours by increasing (degree, number); the whole sequence is
// then reversed.  Boundary elements keep their order, so the halos keep
// theirs.
namespace {
void rcm_order(int ne, const std::vector<int> &idx, const std::vector<int> &ef2nb,
               std::vector<int> &order) {
  std::vector<int> deg(ne, 0);
  for (int e = 0; e < ne; e++)
    for (int k = idx[e] - 1; k < idx[e + 1] - 1; k++)
      if (pk_elem(ef2nb[k]) <= ne) deg[e]++;
  std::vector<int> start(ne);
  for (int e = 0; e < ne; e++) start[e] = e;
  std::stable_sort(start.begin(), start.end(), [&](int a, int b) { return deg[a] < deg[b]; });
  std::vector<char> seen(ne, 0);
  order.clear();
  order.reserve(ne);
  std::vector<int> nb;
  for (int s0 : start) {
    if (seen[s0]) continue;
    seen[s0] = 1;
    size_t head = order.size();
    order.push_back(s0);
    while (head < order.size()) {
      const int v = order[head++];
      nb.clear();
      for (int k = idx[v] - 1; k < idx[v + 1] - 1; k++) {
        const int j = pk_elem(ef2nb[k]) - 1;
        if (j < ne && !seen[j]) nb.push_back(j);
      }
      std::sort(nb.begin(), nb.end(),
                [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
      for (int j : nb) {
        seen[j] = 1;
        order.push_back(j);
      }
    }
  }
  std::reverse(order.begin(), order.end());
}

// cell k of the new numbering is cell from[k] of the current one
int permute_cells(cfdl_mesh *m, const std::vector<int> &from) {
  const int ne = m->ne;
  std::vector<int> ev(m->e2vx.begin(), m->e2vx.begin() + 8 * static_cast<size_t>(ne));
  for (int k = 0; k < ne; k++)
    std::memcpy(m->e2vx.data() + 8 * static_cast<size_t>(k), ev.data() + 8 * static_cast<size_t>(from[k]),
                8 * sizeof(int));
  return mesh_build(m);
}
}  // namespace

extern "C" int cfdl_mesh_renumber(cfdl_mesh *m, uint64_t seed) {
  if (!m || m->ne <= 0) {
    set_error("cfdl_mesh_renumber: null mesh");
    return CFDL_EINVAL;
  }
  if (m->nsubd != 0) {
    set_error("cfdl_mesh_renumber: renumber before cfdl_mesh_partition");
    return CFDL_ESTATE;
  }
  if (m->nsec < 1 || m->etype[0] != kHexa8 || m->esec[0] != 1 || m->esec[1] != m->ne) {
    set_error("cfdl_mesh_renumber: the cells must be the first section");
    return CFDL_EINVAL;
  }
  const int ne = m->ne;
  std::vector<int> from(ne);
  for (int k = 0; k < ne; k++) from[k] = k;
  uint64_t st = seed;
  for (int i = ne - 1; i >= 1; i--) {
    const int j = static_cast<int>(splitmix64(st) % static_cast<uint64_t>(i + 1));
    std::swap(from[i], from[j]);
  }
  int rc = permute_cells(m, from);
  if (rc) return rc;
  std::vector<int> order;
  rcm_order(ne, m->ef2nb_idx, m->ef2nb, order);
  return permute_cells(m, order);
}

extern "C" int cfdl_mesh_info(const cfdl_mesh *m, int *ne, int *nf, int *nbf, int *nsec,
                              int *nsubd, int *nvx) {
  if (!m) {
    set_error("cfdl_mesh_info: null mesh");
    return CFDL_EINVAL;
  }
  if (ne) *ne = m->ne;
  if (nf) *nf = m->nf;
  if (nbf) *nbf = m->nbf;
  if (nsec) *nsec = m->nsec;
  if (nsubd) *nsubd = m->nsubd;
  if (nvx) *nvx = m->nvx;
  return CFDL_OK;
}

template <typename T>
static int copy_out(const std::vector<T> &v, T *dst, int64_t n, const char *name) {
  if (!dst || n != static_cast<int64_t>(v.size())) {
    set_error(std::string("mesh array '") + name + "' has " + std::to_string(v.size()) +
              " entries, caller gave " + std::to_string(n));
    return CFDL_EINVAL;
  }
  std::memcpy(dst, v.data(), sizeof(T) * v.size());
  return CFDL_OK;
}

extern "C" int cfdl_mesh_get_i32(const cfdl_mesh *m, const char *name, int *dst, int64_t n) {
  if (!m || !name) {
    set_error("cfdl_mesh_get_i32: null argument");
    return CFDL_EINVAL;
  }
  const std::string s(name);
  if (s == "ef2nb_idx") return copy_out(m->ef2nb_idx, dst, n, name);
  if (s == "ef2nb") return copy_out(m->ef2nb, dst, n, name);
  if (s == "s2g") return copy_out(m->s2g, dst, n, name);
  if (s == "bs") return copy_out(m->bs, dst, n, name);
  if (s == "esec") return copy_out(m->esec, dst, n, name);
  if (s == "etype") return copy_out(m->etype, dst, n, name);
  if (s == "e2vx") return copy_out(m->e2vx, dst, n, name);
  if (s == "gf2g") return copy_out(m->gf2g, dst, n, name);
  if (s == "g2gf") return copy_out(m->g2gf, dst, n, name);
  if (s == "g2gf_idx") return copy_out(m->g2gf_idx, dst, n, name);
  set_error("unknown mesh array " + s);
  return CFDL_EINVAL;
}

extern "C" int cfdl_mesh_get_f64(const cfdl_mesh *m, const char *name, double *dst, int64_t n) {
  if (!m || !name) {
    set_error("cfdl_mesh_get_f64: null argument");
    return CFDL_EINVAL;
  }
  const std::string s(name);
  if (s == "xc") return copy_out(m->xc, dst, n, name);
  if (s == "yc") return copy_out(m->yc, dst, n, name);
  if (s == "zc") return copy_out(m->zc, dst, n, name);
  if (s == "vol") return copy_out(m->vol, dst, n, name);
  if (s == "aip") return copy_out(m->aip, dst, n, name);
  if (s == "rip") return copy_out(m->rip, dst, n, name);
  if (s == "x") return copy_out(m->x, dst, n, name);
  if (s == "y") return copy_out(m->y, dst, n, name);
  if (s == "z") return copy_out(m->z, dst, n, name);
  set_error("unknown mesh array " + s);
  return CFDL_EINVAL;
}

extern "C" void cfdl_mesh_free(cfdl_mesh *m) { delete m; }

extern "C" int cfdl_mesh_set_section_names(cfdl_mesh *m, int nsec, const char *const *names) {
  if (!m || !names || nsec != m->nsec) {
    cfdl::set_error("cfdl_mesh_set_section_names: one name per section is needed");
    return CFDL_EINVAL;
  }
  m->secname.clear();
  for (int s = 0; s < nsec; s++) {
    if (!names[s]) {
      cfdl::set_error("cfdl_mesh_set_section_names: null name");
      return CFDL_EINVAL;
    }
    m->secname.emplace_back(names[s]);
  }
  return CFDL_OK;
}

// VTU output of cell fields, in the file layout of the reference's
// write_vtubin / vtu_data (src/modules/mod_vtu_output.f90:6-326): an XML
// header that lists every array with its byte offset into the appended
// block, then '_' and the raw little-endian data, each array preceded by its
// byte count as a 64-bit integer -- the cell arrays (values of the 3D
// sections, ncomp per cell), the vertex coordinates (x,y,z interleaved), the
// 0-based connectivity, the running vertex offsets and the VTK cell types
// (12 = hexahedron), all Int32.
extern "C" int cfdl_mesh_write_vtu(const cfdl_mesh *m, const char *path, int nvar,
                                   const char *const *names, const int *ncomp,
                                   const double *const *data) {
  using namespace cfdl;
  if (!m || !path || nvar < 0 || (nvar > 0 && (!names || !ncomp || !data))) {
    set_error("cfdl_mesh_write_vtu: bad arguments");
    return CFDL_EINVAL;
  }
  std::vector<int> cells;  // 3D sections
  for (int s = 0; s < m->nsec; s++) {
    if (m->etype[s] == kQuad4) continue;
    if (m->etype[s] != kHexa8) {
      set_error("cfdl_mesh_write_vtu: only HEXA_8 cell sections are supported");
      return CFDL_EINVAL;
    }
    cells.push_back(s);
  }
  for (int v = 0; v < nvar; v++)
    if (!names[v] || !data[v] || ncomp[v] < 1 || ncomp[v] > 9) {
      set_error("cfdl_mesh_write_vtu: bad array " + std::to_string(v));
      return CFDL_EINVAL;
    }
  int64_t ncell = 0;
  for (int s : cells) ncell += m->esec[2 * s + 1] - m->esec[2 * s] + 1;
  // offsets are printed as the reference's I12 edit descriptor prints them:
  // right-justified in 12 characters, leading blanks kept
  auto i12 = [](int64_t v) {
    char buf[32];
    std::snprintf(buf, sizeof buf, "%12lld", static_cast<long long>(v));
    return std::string(buf);
  };
  std::string h;
  const char lf = '\n';
  h += std::string("<?xml version=\"1.0\"?>") + lf;
  h += std::string("<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
                   "header_type=\"UInt64\">") + lf;
  h += std::string(" <UnstructuredGrid>") + lf;
  h += "  <Piece NumberOfPoints=\"" + std::to_string(m->nvx) + "\" NumberOfCells=\"" +
       std::to_string(m->ne) + "\">" + lf;
  h += std::string("   <CellData Scalars=\"scalars\">") + lf;
  int64_t off = 0;
  std::vector<int64_t> nbytes(nvar);
  for (int v = 0; v < nvar; v++) {
    h += "    <DataArray type=\"Float64\" Name=\"" + std::string(names[v]) +
         "\" NumberOfComponents=\"" + std::to_string(ncomp[v]) +
         "\" format=\"appended\" offset=\"" + i12(off) + "\" />" + lf;
    nbytes[v] = static_cast<int64_t>(m->ne) * ncomp[v] * 8;
    off += 8 + nbytes[v];
  }
  h += std::string("   </CellData>") + lf;
  h += std::string("   <Points>") + lf;
  h += "    <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" +
       i12(off) + "\" />" + lf;
  const int64_t nb_pts = 3 * static_cast<int64_t>(m->nvx) * 8;
  off += 8 + nb_pts;
  h += std::string("   </Points>") + lf;
  h += std::string("   <Cells>") + lf;
  h += "    <DataArray type=\"Int32\" Name=\"connectivity\" format=\"appended\" offset=\"" +
       i12(off) + "\" />" + lf;
  const int64_t nb_conn = ncell * 8 * 4;
  off += 8 + nb_conn;
  h += "    <DataArray type=\"Int32\" Name=\"offsets\" format=\"appended\" offset=\"" +
       i12(off) + "\" />" + lf;
  const int64_t nb_offs = ncell * 4;
  off += 8 + nb_offs;
  h += "    <DataArray type=\"Int32\" Name=\"types\" format=\"appended\" offset=\"" +
       i12(off) + "\" />" + lf;
  const int64_t nb_types = static_cast<int64_t>(m->ne) * 4;
  h += std::string("   </Cells>") + lf;
  h += std::string("  </Piece>") + lf;
  h += std::string(" </UnstructuredGrid>") + lf;
  h += std::string(" <AppendedData encoding=\"raw\">") + lf;
  h += "_";
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    set_error(std::string("cfdl_mesh_write_vtu: cannot open ") + path);
    return CFDL_EINVAL;
  }
  bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
  auto put = [&](const void *p, size_t n) { ok = ok && std::fwrite(p, 1, n, f) == n; };
  for (int v = 0; v < nvar; v++) {
    put(&nbytes[v], 8);
    for (int s : cells) {
      const int64_t e0 = m->esec[2 * s] - 1, e1 = m->esec[2 * s + 1];
      put(data[v] + e0 * ncomp[v], static_cast<size_t>((e1 - e0) * ncomp[v]) * 8);
    }
  }
  put(&nb_pts, 8);
  for (int j = 0; j < m->nvx; j++) {
    const double xyz[3] = {m->x[j], m->y[j], m->z[j]};
    put(xyz, 24);
  }
  put(&nb_conn, 8);
  for (int s : cells)
    for (int e = m->esec[2 * s]; e <= m->esec[2 * s + 1]; e++) {
      int32_t v[8];
      for (int k = 0; k < 8; k++) v[k] = m->e2vx[8 * static_cast<size_t>(e - 1) + k] - 1;
      put(v, 32);
    }
  put(&nb_offs, 8);
  int32_t running = 0;
  for (int s : cells)
    for (int e = m->esec[2 * s]; e <= m->esec[2 * s + 1]; e++) {
      running += 8;
      put(&running, 4);
    }
  put(&nb_types, 8);
  for (int s : cells)
    for (int e = m->esec[2 * s]; e <= m->esec[2 * s + 1]; e++) {
      const int32_t t = 12;
      put(&t, 4);
    }
  const std::string tail = std::string(1, lf) + " </AppendedData>" + lf + "</VTKFile>" + lf;
  put(tail.data(), tail.size());
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    set_error(std::string("cfdl_mesh_write_vtu: write failed: ") + path);
    return CFDL_EINVAL;
  }
  return CFDL_OK;
}
