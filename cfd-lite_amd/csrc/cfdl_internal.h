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
}
