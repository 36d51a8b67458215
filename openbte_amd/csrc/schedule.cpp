// Host-side sweep scheduling for the upwind BTE operator.
//
// For every direction s_l the upwind operator A_ml = I + Lambda_m (D_l + G^-_l)
// (openbte/bte.py:181-185) couples a cell only to the neighbours across its
// inflow faces (s_l . n_f < 0).  Without the periodic faces that dependency
// graph is acyclic for convex 2D cells, so cutting the periodic inflow faces
// (openbte/mesh.py:268-296 keeps them inside `internal`) leaves a DAG that a
// level-ordered sweep inverts exactly.  The cut couplings become interface
// unknowns solved by the Krylov layer.  Any cycle left after that cut (possible
// on non-Delaunay tetrahedra) is broken greedily and its faces join the
// interface as well.
//
// Output per direction, in sweep ("position") order:
//   order[p]        cell at position p (cells grouped by level)
//   pos[c]          inverse permutation
//   lvl_ptr[k]      first position of level k (nlev+1 entries)
//   in_ptr[p]       CSR row pointer into the inflow entries of position p
//   in_src[e]       upstream position (>= 0) or ~cut_index (< 0)
//   in_face[e]      directed face id of the entry (for coefficient assembly)
//   cut_exit[f]     position of the upstream cell of cut face f
//   cut_face[f]     directed face id of cut face f
// Cut faces are numbered in ascending directed-face id, which does not depend
// on the layering: the interface vectors (and the order of every sum over
// them) are the same for any level width.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include <omp.h>

#include "capi_common.h"

namespace {

struct DirSchedule {
  std::vector<int32_t> order, pos, lvl_ptr, in_ptr, in_src, in_face, cut_exit, cut_face;
};

struct Schedule {
  int32_t n = 0, n_dirs = 0;
  std::vector<DirSchedule> dirs;
};

// Directed-face inflow test: row cell i_mat[f] receives from j_mat[f] when
// s . k_f < 0 (G^- of openbte/bte.py:58-60).
inline double face_dot(const double* s, const double* k, int dim) {
  double g = 0.0;
  for (int d = 0; d < dim; ++d) g += s[d] * k[d];
  return g;
}

void build_direction(int32_t n, int64_t n_faces, int dim, const int32_t* i_mat, const int32_t* j_mat,
                     const double* k_mat, const uint8_t* cut_candidate, const double* s, int32_t max_width,
                     DirSchedule& out) {
  // one pass: inflow and cut flags, inflow faces per row cell (cnt), kept
  // (uncut) inflow faces per upstream cell (dcnt) and per row cell (indeg)
  std::vector<int32_t> cnt(n + 1, 0), dcnt(n + 1, 0), indeg(n, 0);
  std::vector<uint8_t> flag(n_faces, 0);  // 1 inflow, 3 inflow and cut
  for (int64_t f = 0; f < n_faces; ++f) {
    if (face_dot(s, k_mat + f * dim, dim) < 0.0) {
      const int32_t i = i_mat[f];
      cnt[i + 1]++;
      if (cut_candidate[f]) {
        flag[f] = 3;
      } else {
        flag[f] = 1;
        dcnt[j_mat[f] + 1]++;
        indeg[i]++;
      }
    }
  }
  for (int32_t c = 0; c < n; ++c) {
    cnt[c + 1] += cnt[c];
    dcnt[c + 1] += dcnt[c];
  }
  // second pass: the faces of each row cell and each upstream cell
  std::vector<int32_t> row_faces(cnt[n]), down_face(dcnt[n]);
  std::vector<uint8_t> is_cut(n_faces, 0);
  {
    std::vector<int32_t> fr(cnt.begin(), cnt.end() - 1), fd(dcnt.begin(), dcnt.end() - 1);
    for (int64_t f = 0; f < n_faces; ++f) {
      if (!flag[f]) continue;
      row_faces[fr[i_mat[f]]++] = (int32_t)f;
      if (flag[f] == 3) is_cut[f] = 1;
      else down_face[fd[j_mat[f]]++] = (int32_t)f;
    }
  }

  // Kahn's algorithm by levels (longest-path level of each cell).  With
  // max_width > 0 a level takes at most max_width ready cells; the rest wait
  // for the next level ahead of the newly ready ones.  Any such layering is a
  // valid sweep order and gives the same values.
  std::vector<int32_t> level(n, -1);
  std::vector<int32_t> front, next, carry, cur;
  for (int32_t c = 0; c < n; ++c)
    if (indeg[c] == 0) front.push_back(c);
  int32_t done = 0, lev = 0;
  out.order.clear();
  out.order.reserve(n);
  out.lvl_ptr.clear();
  while (done < n) {
    if (front.empty() && carry.empty()) {
      // cycle: make the pending cell with the fewest remaining inflows a
      // source by cutting its remaining inflow faces
      int32_t best = -1, bestdeg = 1 << 30;
      for (int32_t c = 0; c < n; ++c)
        if (level[c] < 0 && indeg[c] < bestdeg) { best = c; bestdeg = indeg[c]; }
      for (int32_t e = cnt[best]; e < cnt[best + 1]; ++e) {
        int32_t f = row_faces[e];
        if (!is_cut[f] && level[j_mat[f]] < 0) {
          is_cut[f] = 1;
          // drop the face from its upstream cell's downstream list
          int32_t u = j_mat[f];
          for (int32_t q = dcnt[u]; q < dcnt[u + 1]; ++q)
            if (down_face[q] == f) down_face[q] = -1;
        }
      }
      indeg[best] = 0;
      front.push_back(best);
    }
    if (max_width <= 0 || carry.size() + front.size() <= (size_t)max_width) {
      // nothing is deferred: the level is carry and front whole (one sort,
      // same result)
      cur.assign(carry.begin(), carry.end());
      cur.insert(cur.end(), front.begin(), front.end());
      carry.clear();
    } else {
      std::sort(front.begin(), front.end());
      cur.assign(carry.begin(), carry.end());
      cur.insert(cur.end(), front.begin(), front.end());
      size_t take = cur.size();
      if (max_width > 0 && take > (size_t)max_width) take = (size_t)max_width;
      carry.assign(cur.begin() + take, cur.end());
      cur.resize(take);
    }
    std::sort(cur.begin(), cur.end());
    out.lvl_ptr.push_back(done);
    next.clear();
    for (int32_t c : cur) {
      level[c] = lev;
      out.order.push_back(c);
    }
    done += (int32_t)cur.size();
    for (int32_t c : cur) {
      for (int32_t q = dcnt[c]; q < dcnt[c + 1]; ++q) {
        int32_t f = down_face[q];
        if (f < 0) continue;
        int32_t d = i_mat[f];
        if (--indeg[d] == 0) next.push_back(d);
      }
    }
    front.swap(next);
    ++lev;
  }
  out.lvl_ptr.push_back(n);

  out.pos.assign(n, 0);
  for (int32_t p = 0; p < n; ++p) out.pos[out.order[p]] = p;

  // cut faces numbered in ascending face id (is_cut is only set on inflow
  // faces of this direction)
  std::vector<int32_t> cut_index(n_faces, -1);
  int32_t ncut = 0;
  for (int64_t f = 0; f < n_faces; ++f)
    if (is_cut[f]) cut_index[f] = ncut++;
  out.cut_exit.resize(ncut);
  out.cut_face.resize(ncut);
  // inflow CSR in position order
  out.in_ptr.assign(n + 1, 0);
  out.in_src.resize(cnt[n]);
  out.in_face.resize(cnt[n]);
  int32_t w = 0;
  for (int32_t p = 0; p < n; ++p) {
    const int32_t c = out.order[p];
    for (int32_t e = cnt[c]; e < cnt[c + 1]; ++e, ++w) {
      const int32_t f = row_faces[e];
      const int32_t up = out.pos[j_mat[f]];
      if (is_cut[f]) {
        const int32_t q = cut_index[f];
        out.cut_face[q] = f;
        out.cut_exit[q] = up;
        out.in_src[w] = ~q;
      } else {
        out.in_src[w] = up;
      }
      out.in_face[w] = f;
    }
    out.in_ptr[p + 1] = w;
  }
}

}  // namespace

extern "C" {

int obte_sched_create(int32_t n, int64_t n_faces, int32_t dim, const int32_t* i_mat, const int32_t* j_mat,
                      const double* k_mat, const uint8_t* cut_candidate, int32_t n_dirs,
                      const double* dirs, int32_t max_width, void** handle) {
  try {
    if (!handle || n <= 0 || n_faces < 0 || (dim != 2 && dim != 3) || n_dirs <= 0 || max_width < 0)
      return obte_fail(OBTE_EINVAL, "obte_sched_create: invalid arguments");
    for (int64_t f = 0; f < n_faces; ++f)
      if (i_mat[f] < 0 || i_mat[f] >= n || j_mat[f] < 0 || j_mat[f] >= n)
        return obte_fail(OBTE_EINVAL, "obte_sched_create: face cell index out of range");
    std::unique_ptr<Schedule> s(new Schedule());
    s->n = n;
    s->n_dirs = n_dirs;
    s->dirs.resize(n_dirs);
    // an exception may not leave an OpenMP region: each direction records
    // its own failure and the first one is reported after the region
    std::vector<int> err(n_dirs, OBTE_OK);
#pragma omp parallel for schedule(dynamic)
    for (int32_t l = 0; l < n_dirs; ++l) {
      try {
        build_direction(n, n_faces, dim, i_mat, j_mat, k_mat, cut_candidate, dirs + (int64_t)l * dim,
                        max_width, s->dirs[l]);
      } catch (const std::bad_alloc&) {
        err[l] = OBTE_ENOMEM;
      } catch (...) {
        err[l] = OBTE_EINTERNAL;
      }
    }
    for (int32_t l = 0; l < n_dirs; ++l) {
      if (err[l] == OBTE_ENOMEM) return obte_fail(OBTE_ENOMEM, "obte_sched_create: out of host memory");
      if (err[l] != OBTE_OK) return obte_fail(OBTE_EINTERNAL, "obte_sched_create: internal error");
    }
    *handle = s.release();
    return OBTE_OK;
  } catch (const std::bad_alloc&) {
    return obte_fail(OBTE_ENOMEM, "obte_sched_create: out of host memory");
  } catch (...) {
    return obte_fail(OBTE_EINTERNAL, "obte_sched_create: internal error");
  }
}

// sizes[l*3 + 0] = number of levels, [1] = inflow entries, [2] = cut faces
int obte_sched_sizes(void* handle, int64_t* sizes) {
  if (!handle || !sizes) return obte_fail(OBTE_EINVAL, "obte_sched_sizes: null argument");
  auto* s = static_cast<Schedule*>(handle);
  for (int32_t l = 0; l < s->n_dirs; ++l) {
    const DirSchedule& d = s->dirs[l];
    sizes[3 * l + 0] = (int64_t)d.lvl_ptr.size() - 1;
    sizes[3 * l + 1] = (int64_t)d.in_src.size();
    sizes[3 * l + 2] = (int64_t)d.cut_face.size();
  }
  return OBTE_OK;
}

// Concatenate all directions.  Array lengths: order/pos n*L; lvl_ptr
// sum(nlev_l+1); in_ptr L*(n+1) (local entry indices); in_src/in_face
// sum(entries_l); cut_exit/cut_face sum(cuts_l).
int obte_sched_export(void* handle, int32_t* order, int32_t* pos, int32_t* lvl_ptr, int32_t* in_ptr,
                      int32_t* in_src, int32_t* in_face, int32_t* cut_exit, int32_t* cut_face) {
  if (!handle) return obte_fail(OBTE_EINVAL, "obte_sched_export: null handle");
  auto* s = static_cast<Schedule*>(handle);
  const int64_t n = s->n;
  const int32_t L = s->n_dirs;
  // offsets of every direction's blocks, then the copies in parallel
  std::vector<int64_t> lo_off(L + 1, 0), eo_off(L + 1, 0), co_off(L + 1, 0);
  for (int32_t l = 0; l < L; ++l) {
    lo_off[l + 1] = lo_off[l] + (int64_t)s->dirs[l].lvl_ptr.size();
    eo_off[l + 1] = eo_off[l] + (int64_t)s->dirs[l].in_src.size();
    co_off[l + 1] = co_off[l] + (int64_t)s->dirs[l].cut_face.size();
  }
#pragma omp parallel for schedule(static)
  for (int32_t l = 0; l < L; ++l) {
    const DirSchedule& d = s->dirs[l];
    const int64_t lo = lo_off[l], eo = eo_off[l], co = co_off[l];
    std::memcpy(order + l * n, d.order.data(), n * sizeof(int32_t));
    std::memcpy(pos + l * n, d.pos.data(), n * sizeof(int32_t));
    std::memcpy(lvl_ptr + lo, d.lvl_ptr.data(), d.lvl_ptr.size() * sizeof(int32_t));
    std::memcpy(in_ptr + l * (n + 1), d.in_ptr.data(), (n + 1) * sizeof(int32_t));
    std::memcpy(in_src + eo, d.in_src.data(), d.in_src.size() * sizeof(int32_t));
    std::memcpy(in_face + eo, d.in_face.data(), d.in_face.size() * sizeof(int32_t));
    if (!d.cut_face.empty()) {
      std::memcpy(cut_exit + co, d.cut_exit.data(), d.cut_exit.size() * sizeof(int32_t));
      std::memcpy(cut_face + co, d.cut_face.data(), d.cut_face.size() * sizeof(int32_t));
    }
  }
  return OBTE_OK;
}

// Fixed-width inflow slots in position order, [L][kin][n] each: slot k of
// position p holds p's k-th inflow entry as an upstream row (a position, or
// n + cut index for a cut face) and its directed face; unused slots point at
// p itself with face -1.  Directions are filled in parallel.
int obte_sched_export_slots(void* handle, int32_t kin, int32_t* nb_src, int32_t* nb_face) {
  if (!handle || !nb_src || !nb_face || kin <= 0)
    return obte_fail(OBTE_EINVAL, "obte_sched_export_slots: invalid arguments");
  auto* s = static_cast<Schedule*>(handle);
  const int64_t n = s->n;
  for (int32_t l = 0; l < s->n_dirs; ++l) {
    const DirSchedule& d = s->dirs[l];
    for (int64_t p = 0; p < n; ++p)
      if (d.in_ptr[p + 1] - d.in_ptr[p] > kin)
        return obte_fail(OBTE_EINVAL, "obte_sched_export_slots: more inflow faces than kin slots");
  }
#pragma omp parallel for schedule(static)
  for (int32_t l = 0; l < s->n_dirs; ++l) {
    const DirSchedule& d = s->dirs[l];
    int32_t* src = nb_src + (int64_t)l * kin * n;
    int32_t* face = nb_face + (int64_t)l * kin * n;
    for (int64_t p = 0; p < n; ++p) {
      const int32_t e0 = d.in_ptr[p], e1 = d.in_ptr[p + 1];
      for (int32_t k = 0; k < kin; ++k) {
        const int32_t e = e0 + k;
        if (e < e1) {
          const int32_t v = d.in_src[e];
          src[k * n + p] = v >= 0 ? v : (int32_t)(n + ~v);
          face[k * n + p] = d.in_face[e];
        } else {
          src[k * n + p] = (int32_t)p;
          face[k * n + p] = -1;
        }
      }
    }
  }
  return OBTE_OK;
}

int obte_sched_destroy(void* handle) {
  delete static_cast<Schedule*>(handle);
  return OBTE_OK;
}

int obte_set_host_threads(int32_t n) {
  if (n < 1) return obte_fail(OBTE_EINVAL, "obte_set_host_threads: n must be at least 1");
  omp_set_num_threads(n);  // the calling thread's parallel regions from now on
  return OBTE_OK;
}

}  // extern "C"
