// Host-side structure builders for HEX8 and TET4 (vec = 3) and QUAD4
// (vec = 2) meshes.
//
// fx_hex8_pattern produces the CSR pattern that feax builds in
// Problem._build_csr_assembly_structure (feax/problem.py:316-484): the
// unique (row, col) pairs of the element DOF blocks, row-major, columns
// ascending.  For a vec = 3 problem row 3n+i holds columns 3m+j for every
// node m sharing an element with node n (boundary faces re-use their cell's
// DOF block, feax/problem.py:195-205, so they add no new pairs).  Instead of
// feax's argsort over 576(E+F) int64 keys we build it from node adjacency,
// O(E) memory; the result is the same integer arrays.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "common.h"

namespace {

// feax local face vertex lists (basix hexahedron facets through the
// re-order [0,1,3,2,4,5,7,6], feax/basis.py:65-66, 299-322).
const int kFaceVerts[6][4] = {{0, 1, 3, 2}, {0, 1, 4, 5}, {0, 3, 4, 7},
                              {1, 2, 5, 6}, {3, 2, 7, 6}, {4, 5, 7, 6}};

// basix tetrahedron facets: facet i is the triangle opposite vertex i
// (feax keeps basix's TET4 vertex order, feax/basis.py:78-83).
const int kTetFaceVerts[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

// The CSR pattern, element ranks and node incidence of a mesh of NPE-node
// cells (the incidence entry of node a of element e is (e << shift) | a:
// shift 3 for up to 8 nodes per cell, 5 for HEX20).
int cell_pattern(const char* name, int npe, int vec, int32_t num_nodes, int32_t num_cells, const int32_t* cells,
                 int64_t* nse_out, int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                 int32_t* inc_ptr, int32_t* inc, int shift = 3) {
  if (num_nodes <= 0 || num_cells < 0 || !cells || !nse_out || !indptr)
    return fx::fail(FX_EINVAL, "%s: bad arguments", name);
  if (npe > (1 << shift)) return fx::fail(FX_EINVAL, "%s: %d nodes per cell need a wider incidence", name, npe);
  if (static_cast<int64_t>(num_nodes) * vec >= INT32_MAX || (static_cast<int64_t>(num_cells) << shift) >= INT32_MAX)
    return fx::fail(FX_ELIMIT, "%s: mesh too large for int32 indices", name);
  const int64_t E = num_cells, N = num_nodes;
  // node -> (element, local) incidence, ascending element id per node.
  std::vector<int32_t> iptr(N + 1, 0);
  for (int64_t k = 0; k < E * npe; ++k) {
    int32_t nd = cells[k];
    if (nd < 0 || nd >= num_nodes) return fx::fail(FX_EINVAL, "%s: node id %d out of range", name, nd);
    iptr[nd + 1]++;
  }
  for (int64_t i = 0; i < N; ++i) iptr[i + 1] += iptr[i];
  std::vector<int32_t> ilist(E * npe);
  {
    std::vector<int32_t> pos(iptr.begin(), iptr.end() - 1);
    for (int64_t e = 0; e < E; ++e)
      for (int a = 0; a < npe; ++a) ilist[pos[cells[e * npe + a]]++] = static_cast<int32_t>((e << shift) | a);
  }
  // neighbour lists
  std::vector<int32_t> nptr(N + 1, 0);
  std::vector<int32_t> nbr;
  nbr.reserve(N * 27);
  std::vector<int32_t> tmp;
  int32_t maxdeg = 0;
  for (int64_t nd = 0; nd < N; ++nd) {
    tmp.clear();
    for (int32_t k = iptr[nd]; k < iptr[nd + 1]; ++k) {
      int32_t e = ilist[k] >> shift;
      for (int b = 0; b < npe; ++b) tmp.push_back(cells[static_cast<int64_t>(e) * npe + b]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    nbr.insert(nbr.end(), tmp.begin(), tmp.end());
    nptr[nd + 1] = static_cast<int32_t>(nbr.size());
    maxdeg = std::max<int32_t>(maxdeg, static_cast<int32_t>(tmp.size()));
  }
  if (maxdeg > 255) return fx::fail(FX_ELIMIT, "%s: node degree %d > 255", name, maxdeg);
  int64_t nse = 0;
  indptr[0] = 0;
  for (int64_t nd = 0; nd < N; ++nd) {
    int64_t deg = nptr[nd + 1] - nptr[nd];
    for (int i = 0; i < vec; ++i) {
      nse += vec * deg;
      if (nse >= INT32_MAX) return fx::fail(FX_ELIMIT, "%s: nse exceeds int32", name);
      indptr[nd * vec + i + 1] = static_cast<int32_t>(nse);
    }
  }
  *nse_out = nse;
  if (max_degree) *max_degree = maxdeg;
  if (!indices) return FX_OK;
  for (int64_t nd = 0; nd < N; ++nd) {
    for (int i = 0; i < vec; ++i) {
      int64_t s = indptr[nd * vec + i];
      for (int32_t k = nptr[nd]; k < nptr[nd + 1]; ++k)
        for (int j = 0; j < vec; ++j) indices[s++] = nbr[k] * vec + j;
    }
  }
  if (elem_rank) {
    for (int64_t e = 0; e < E; ++e)
      for (int a = 0; a < npe; ++a) {
        int32_t nd = cells[e * npe + a];
        const int32_t* lo = nbr.data() + nptr[nd];
        const int32_t* hi = nbr.data() + nptr[nd + 1];
        for (int b = 0; b < npe; ++b) {
          const int32_t* it = std::lower_bound(lo, hi, cells[e * npe + b]);
          elem_rank[(e * npe + a) * npe + b] = static_cast<uint8_t>(it - lo);
        }
      }
  }
  if (inc_ptr) std::copy(iptr.begin(), iptr.end(), inc_ptr);
  if (inc) std::copy(ilist.begin(), ilist.end(), inc);
  return FX_OK;
}

}  // namespace

extern "C" int fx_hex8_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse_out,
                               int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                               int32_t* inc_ptr, int32_t* inc) {
  return cell_pattern("fx_hex8_pattern", 8, 3, num_nodes, num_cells, cells, nse_out, max_degree, indptr, indices,
                      elem_rank, inc_ptr, inc);
}

extern "C" int fx_tet4_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse_out,
                               int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                               int32_t* inc_ptr, int32_t* inc) {
  return cell_pattern("fx_tet4_pattern", 4, 3, num_nodes, num_cells, cells, nse_out, max_degree, indptr, indices,
                      elem_rank, inc_ptr, inc);
}

extern "C" int fx_quad4_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse_out,
                                int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                                int32_t* inc_ptr, int32_t* inc) {
  return cell_pattern("fx_quad4_pattern", 4, 2, num_nodes, num_cells, cells, nse_out, max_degree, indptr, indices,
                      elem_rank, inc_ptr, inc);
}

extern "C" int fx_hex8s_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse_out,
                                int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                                int32_t* inc_ptr, int32_t* inc) {
  return cell_pattern("fx_hex8s_pattern", 8, 1, num_nodes, num_cells, cells, nse_out, max_degree, indptr, indices,
                      elem_rank, inc_ptr, inc);
}

extern "C" int fx_hex8_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells,
                                      int32_t num_faces, const int32_t* faces, int32_t* finc_ptr,
                                      int32_t* finc) {
  if (num_nodes <= 0 || num_faces < 0 || !cells || !finc_ptr || (num_faces > 0 && (!faces || !finc)))
    return fx::fail(FX_EINVAL, "fx_hex8_face_incidence: bad arguments");
  std::vector<int32_t> cnt(static_cast<size_t>(num_nodes) + 1, 0);
  for (int32_t f = 0; f < num_faces; ++f) {
    int32_t e = faces[2 * f], lf = faces[2 * f + 1];
    if (e < 0 || e >= num_cells || lf < 0 || lf > 5)
      return fx::fail(FX_EINVAL, "fx_hex8_face_incidence: bad face (%d, %d)", e, lf);
    for (int v = 0; v < 4; ++v) cnt[cells[static_cast<int64_t>(e) * 8 + kFaceVerts[lf][v]] + 1]++;
  }
  for (int32_t i = 0; i < num_nodes; ++i) cnt[i + 1] += cnt[i];
  std::copy(cnt.begin(), cnt.end(), finc_ptr);
  std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
  for (int32_t f = 0; f < num_faces; ++f) {
    int32_t e = faces[2 * f], lf = faces[2 * f + 1];
    for (int v = 0; v < 4; ++v) {
      int a = kFaceVerts[lf][v];
      finc[pos[cells[static_cast<int64_t>(e) * 8 + a]]++] = f * 8 + a;
    }
  }
  return FX_OK;
}

extern "C" int fx_tet4_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells,
                                      int32_t num_faces, const int32_t* faces, int32_t* finc_ptr,
                                      int32_t* finc) {
  if (num_nodes <= 0 || num_faces < 0 || !cells || !finc_ptr || (num_faces > 0 && (!faces || !finc)))
    return fx::fail(FX_EINVAL, "fx_tet4_face_incidence: bad arguments");
  std::vector<int32_t> cnt(static_cast<size_t>(num_nodes) + 1, 0);
  for (int32_t f = 0; f < num_faces; ++f) {
    int32_t e = faces[2 * f], lf = faces[2 * f + 1];
    if (e < 0 || e >= num_cells || lf < 0 || lf > 3)
      return fx::fail(FX_EINVAL, "fx_tet4_face_incidence: bad face (%d, %d)", e, lf);
    for (int v = 0; v < 3; ++v) cnt[cells[static_cast<int64_t>(e) * 4 + kTetFaceVerts[lf][v]] + 1]++;
  }
  for (int32_t i = 0; i < num_nodes; ++i) cnt[i + 1] += cnt[i];
  std::copy(cnt.begin(), cnt.end(), finc_ptr);
  std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
  for (int32_t f = 0; f < num_faces; ++f) {
    int32_t e = faces[2 * f], lf = faces[2 * f + 1];
    for (int v = 0; v < 3; ++v) {
      int a = kTetFaceVerts[lf][v];
      finc[pos[cells[static_cast<int64_t>(e) * 4 + a]]++] = f * 8 + a;
    }
  }
  return FX_OK;
}

// QUAD4 edges in feax numbering: basix quadrilateral facets through the
// re-order [0, 1, 3, 2] (feax/basis.py:90-95): bottom, left, right, top.
extern "C" int fx_quad4_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells,
                                       int32_t num_faces, const int32_t* faces, int32_t* finc_ptr,
                                       int32_t* finc) {
  static const int kQuadEdges[4][2] = {{0, 1}, {0, 3}, {1, 2}, {3, 2}};
  if (num_nodes <= 0 || num_faces < 0 || !cells || !finc_ptr || (num_faces > 0 && (!faces || !finc)))
    return fx::fail(FX_EINVAL, "fx_quad4_face_incidence: bad arguments");
  std::vector<int32_t> cnt(static_cast<size_t>(num_nodes) + 1, 0);
  for (int32_t f = 0; f < num_faces; ++f) {
    int32_t e = faces[2 * f], lf = faces[2 * f + 1];
    if (e < 0 || e >= num_cells || lf < 0 || lf > 3)
      return fx::fail(FX_EINVAL, "fx_quad4_face_incidence: bad face (%d, %d)", e, lf);
    for (int v = 0; v < 2; ++v) cnt[cells[static_cast<int64_t>(e) * 4 + kQuadEdges[lf][v]] + 1]++;
  }
  for (int32_t i = 0; i < num_nodes; ++i) cnt[i + 1] += cnt[i];
  std::copy(cnt.begin(), cnt.end(), finc_ptr);
  std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
  for (int32_t f = 0; f < num_faces; ++f) {
    int32_t e = faces[2 * f], lf = faces[2 * f + 1];
    for (int v = 0; v < 2; ++v) {
      int a = kQuadEdges[lf][v];
      finc[pos[cells[static_cast<int64_t>(e) * 4 + a]]++] = f * 8 + a;
    }
  }
  return FX_OK;
}

// HEX20 (feax/basis.py:71-77): 20 nodes, meshio order; incidence entries
// (e << 5) | a.
extern "C" int fx_hex20_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse_out,
                                int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                                int32_t* inc_ptr, int32_t* inc) {
  return cell_pattern("fx_hex20_pattern", 20, 3, num_nodes, num_cells, cells, nse_out, max_degree, indptr, indices,
                      elem_rank, inc_ptr, inc, 5);
}

// The 8 nodes of each HEX20 facet: its 4 corners (the HEX8 facet, in the
// same order) and the 4 edge nodes between them.
extern "C" int fx_hex20_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells,
                                       int32_t num_faces, const int32_t* faces, int32_t* finc_ptr,
                                       int32_t* finc) {
  static const int kH20FaceNodes[6][8] = {{0, 1, 3, 2, 8, 9, 10, 11},   {0, 1, 4, 5, 8, 12, 16, 17},
                                          {0, 3, 4, 7, 11, 15, 16, 19}, {1, 2, 5, 6, 9, 13, 17, 18},
                                          {3, 2, 7, 6, 10, 14, 18, 19}, {4, 5, 7, 6, 12, 13, 14, 15}};
  if (num_nodes <= 0 || num_faces < 0 || !cells || !finc_ptr || (num_faces > 0 && (!faces || !finc)))
    return fx::fail(FX_EINVAL, "fx_hex20_face_incidence: bad arguments");
  if ((static_cast<int64_t>(num_faces) << 5) >= INT32_MAX)
    return fx::fail(FX_ELIMIT, "fx_hex20_face_incidence: too many faces");
  std::vector<int32_t> cnt(static_cast<size_t>(num_nodes) + 1, 0);
  for (int32_t f = 0; f < num_faces; ++f) {
    int32_t e = faces[2 * f], lf = faces[2 * f + 1];
    if (e < 0 || e >= num_cells || lf < 0 || lf > 5)
      return fx::fail(FX_EINVAL, "fx_hex20_face_incidence: bad face (%d, %d)", e, lf);
    for (int v = 0; v < 8; ++v) cnt[cells[static_cast<int64_t>(e) * 20 + kH20FaceNodes[lf][v]] + 1]++;
  }
  for (int32_t i = 0; i < num_nodes; ++i) cnt[i + 1] += cnt[i];
  std::copy(cnt.begin(), cnt.end(), finc_ptr);
  std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
  for (int32_t f = 0; f < num_faces; ++f) {
    int32_t e = faces[2 * f], lf = faces[2 * f + 1];
    for (int v = 0; v < 8; ++v) {
      int a = kH20FaceNodes[lf][v];
      finc[pos[cells[static_cast<int64_t>(e) * 20 + a]]++] = (f << 5) | a;
    }
  }
  return FX_OK;
}

// The monolithic (u, T) HEX8 problem (vec = [3, 1], DOFs u: 3 n + i,
// T: 3 N + n): every row of node n holds the u columns 3 m + j of its
// neighbours m, then their T columns 3 N + m.  elem_rank and the node
// incidence are those of the HEX8 pattern.
extern "C" int fx_hex8ut_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse_out,
                                 int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                                 int32_t* inc_ptr, int32_t* inc) {
  if (num_nodes <= 0 || !nse_out || !indptr) return fx::fail(FX_EINVAL, "fx_hex8ut_pattern: bad arguments");
  if (static_cast<int64_t>(num_nodes) * 4 >= INT32_MAX) return fx::fail(FX_ELIMIT, "fx_hex8ut_pattern: too many nodes");
  const int64_t N = num_nodes;
  std::vector<int32_t> nptr(N + 1);
  int64_t nnb = 0;
  int rc = cell_pattern("fx_hex8ut_pattern", 8, 1, num_nodes, num_cells, cells, &nnb, max_degree, nptr.data(),
                        nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  // rows 3 n + i (u) then 3 N + n (T), 4 deg entries each
  int64_t nse = 0;
  indptr[0] = 0;
  for (int64_t r = 0; r < 4 * N; ++r) {
    const int64_t nd = r < 3 * N ? r / 3 : r - 3 * N;
    nse += 4 * static_cast<int64_t>(nptr[nd + 1] - nptr[nd]);
    if (nse >= INT32_MAX) return fx::fail(FX_ELIMIT, "fx_hex8ut_pattern: nse exceeds int32");
    indptr[r + 1] = static_cast<int32_t>(nse);
  }
  *nse_out = nse;
  if (!indices) return FX_OK;
  std::vector<int32_t> nbr(nnb);
  rc = cell_pattern("fx_hex8ut_pattern", 8, 1, num_nodes, num_cells, cells, &nnb, nullptr, nptr.data(), nbr.data(),
                    elem_rank, inc_ptr, inc);
  if (rc) return rc;
  for (int64_t r = 0; r < 4 * N; ++r) {
    const int64_t nd = r < 3 * N ? r / 3 : r - 3 * N;
    int64_t s = indptr[r];
    for (int32_t k = nptr[nd]; k < nptr[nd + 1]; ++k)
      for (int j = 0; j < 3; ++j) indices[s++] = 3 * nbr[k] + j;
    for (int32_t k = nptr[nd]; k < nptr[nd + 1]; ++k) indices[s++] = static_cast<int32_t>(3 * N + nbr[k]);
  }
  return FX_OK;
}

extern "C" int fx_hex8ut_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells,
                                        int32_t num_faces, const int32_t* faces, int32_t* finc_ptr,
                                        int32_t* finc) {
  return fx_hex8_face_incidence(num_nodes, num_cells, cells, num_faces, faces, finc_ptr, finc);
}

extern "C" int fx_csr_is_block3(int32_t n, const int32_t* indptr, const int32_t* indices,
                                int32_t* is_block3) {
  if (n < 0 || !indptr || !is_block3) return fx::fail(FX_EINVAL, "fx_csr_is_block3: bad arguments");
  *is_block3 = 0;
  if (n % 3) return FX_OK;
  for (int32_t m = 0; m < n / 3; ++m) {
    int32_t r0 = indptr[3 * m], len = indptr[3 * m + 1] - r0;
    if (len % 3 || indptr[3 * m + 2] - indptr[3 * m + 1] != len || indptr[3 * m + 3] - indptr[3 * m + 2] != len)
      return FX_OK;
    for (int32_t k = 0; k < len; ++k) {
      int32_t c = indices[r0 + k];
      if (indices[r0 + len + k] != c || indices[r0 + 2 * len + k] != c) return FX_OK;
      if (k % 3 == 0 ? (c % 3 != 0) : (c != indices[r0 + k - 1] + 1)) return FX_OK;
    }
  }
  *is_block3 = 1;
  return FX_OK;
}

extern "C" int fx_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return FX_EINVAL;
  const std::string& m = fx::last_error();
  size_t k = std::min(len - 1, m.size());
  std::memcpy(buf, m.data(), k);
  buf[k] = 0;
  return FX_OK;
}

extern "C" int fx_version(char* buf, size_t len) {
  if (!buf || len == 0) return FX_EINVAL;
  snprintf(buf, len, "feaxhip 0.1 gfx950");
  return FX_OK;
}

// The face incidence depends on the nodes only: the scalar problem shares it.
extern "C" int fx_hex8s_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells,
                                       int32_t num_faces, const int32_t* faces, int32_t* finc_ptr,
                                       int32_t* finc) {
  return fx_hex8_face_incidence(num_nodes, num_cells, cells, num_faces, faces, finc_ptr, finc);
}

extern "C" int fx_quad4s_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse_out,
                                 int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                                 int32_t* inc_ptr, int32_t* inc) {
  return cell_pattern("fx_quad4s_pattern", 4, 1, num_nodes, num_cells, cells, nse_out, max_degree, indptr, indices,
                      elem_rank, inc_ptr, inc);
}

extern "C" int fx_quad4s_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells,
                                        int32_t num_faces, const int32_t* faces, int32_t* finc_ptr,
                                        int32_t* finc) {
  return fx_quad4_face_incidence(num_nodes, num_cells, cells, num_faces, faces, finc_ptr, finc);
}

extern "C" int fx_tet4s_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse_out,
                                int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                                int32_t* inc_ptr, int32_t* inc) {
  return cell_pattern("fx_tet4s_pattern", 4, 1, num_nodes, num_cells, cells, nse_out, max_degree, indptr, indices,
                      elem_rank, inc_ptr, inc);
}

extern "C" int fx_tet4s_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells,
                                       int32_t num_faces, const int32_t* faces, int32_t* finc_ptr,
                                       int32_t* finc) {
  return fx_tet4_face_incidence(num_nodes, num_cells, cells, num_faces, faces, finc_ptr, finc);
}
