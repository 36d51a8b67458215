// Host-side setup of the smoothed-aggregation multigrid hierarchy that
// preconditions the Fourier CG solve.
//
// The reference factorises F = V^-1 Lap with SuperLU and back-solves it at
// every correction step (openbte/fourier.py:109-118,152-156).  The device
// solves the symmetric form A T = V b by CG instead; each CG iteration applies
// one V-cycle of the hierarchy built here.  A row of a level is stored as a
// diagonal plus off-diagonal weights w (A x)_i = d_i x_i - sum_k w_k x_col[k],
// the layout of obte_fourier_plan_t.
//
// Coarsening: the cells arrive renumbered along a Morton curve of their
// centroids, so every `agg` consecutive rows form a compact aggregate, and the
// aggregates, numbered in the same order, are again Morton-ordered for the next
// level.  The tentative prolongation Pt is piecewise constant on aggregates;
// on levels whose operator has at most `smooth_max_row` off-diagonal entries
// per row (on average) it is smoothed once by weighted Jacobi,
//   P = (I - omega D^-1 A) Pt,
// elsewhere P = Pt (each smoothing widens the coarse stencil: on tetrahedra a
// second smoothed level would hold ~520 entries per row), and the coarse
// operator is the Galerkin product A_c = P^T A P.  Levels are added until a
// level has at most `coarse_max` rows; that level is exported dense (its
// inverse is formed by the caller).
//
// Every sum runs in a fixed order (rows in order, entries of a row in
// ascending column order), so the hierarchy is a deterministic function of the
// input.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "capi_common.h"

namespace {

struct Csr {
  std::vector<int32_t> ptr, col;
  std::vector<double> val;
};

struct Level {
  int32_t n = 0;
  Csr A;                    // off-diagonal weights (A x)_i = diag_i x_i - sum val x[col]
  std::vector<double> diag;
  Csr P, R;                 // P: n x n_next, R = P^T: n_next x n
};

struct Hierarchy {
  std::vector<Level> levels;  // the last level is the coarsest (dense export, no P / R)
};

// Sparse accumulator over the columns of one output row: values summed in the
// order the contributions arrive, columns reported in ascending order.
struct Accumulator {
  std::vector<double> acc;
  std::vector<int32_t> mark, cols;
  explicit Accumulator(int32_t n) : acc(n, 0.0), mark(n, -1) {}
  void add(int32_t row, int32_t c, double v) {
    if (mark[c] != row) {
      mark[c] = row;
      acc[c] = 0.0;
      cols.push_back(c);
    }
    acc[c] += v;
  }
  // append the row to out (sorted columns) and reset
  void flush(Csr& out) {
    std::sort(cols.begin(), cols.end());
    for (int32_t c : cols) {
      out.col.push_back(c);
      out.val.push_back(acc[c]);
    }
    cols.clear();
    out.ptr.push_back((int32_t)out.col.size());
  }
};

// P = (I - omega D^-1 A) Pt for aggregates of `agg` consecutive rows:
//   P[i, J] = [agg(i) = J] (1 - omega) + (omega / d_i) sum_k w_ik [agg(col_k) = J]
// (omega = 0: P = Pt)
void smoothed_prolongation(const Level& L, int32_t agg, int32_t nc, double omega, Csr& P) {
  Accumulator a(nc);
  P.ptr.assign(1, 0);
  P.col.reserve((size_t)L.n * 3);
  P.val.reserve((size_t)L.n * 3);
  for (int32_t i = 0; i < L.n; ++i) {
    a.add(i, i / agg, 1.0 - omega);
    if (omega != 0.0) {
      const double s = omega / L.diag[i];
      for (int32_t k = L.A.ptr[i]; k < L.A.ptr[i + 1]; ++k) a.add(i, L.A.col[k] / agg, s * L.A.val[k]);
    }
    a.flush(P);
  }
}

// R = P^T (rows of R in order of the fine rows within each coarse row)
void transpose(const Csr& P, int32_t n, int32_t nc, Csr& R) {
  R.ptr.assign(nc + 1, 0);
  for (int32_t c : P.col) ++R.ptr[c + 1];
  for (int32_t I = 0; I < nc; ++I) R.ptr[I + 1] += R.ptr[I];
  R.col.resize(P.col.size());
  R.val.resize(P.val.size());
  std::vector<int32_t> fill(R.ptr.begin(), R.ptr.end() - 1);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t k = P.ptr[i]; k < P.ptr[i + 1]; ++k) {
      const int32_t J = P.col[k];
      R.col[fill[J]] = i;
      R.val[fill[J]] = P.val[k];
      ++fill[J];
    }
}

// Galerkin product A_c = R A P, split into diagonal and off-diagonal weights.
void galerkin(const Level& L, const Csr& P, const Csr& R, int32_t nc, Level& C) {
  // AP row i = d_i P[i, :] - sum_k w_ik P[col_k, :]
  Csr AP;
  AP.ptr.assign(1, 0);
  {
    Accumulator a(nc);
    for (int32_t i = 0; i < L.n; ++i) {
      for (int32_t k = P.ptr[i]; k < P.ptr[i + 1]; ++k) a.add(i, P.col[k], L.diag[i] * P.val[k]);
      for (int32_t e = L.A.ptr[i]; e < L.A.ptr[i + 1]; ++e) {
        const int32_t j = L.A.col[e];
        const double w = L.A.val[e];
        for (int32_t k = P.ptr[j]; k < P.ptr[j + 1]; ++k) a.add(i, P.col[k], -w * P.val[k]);
      }
      a.flush(AP);
    }
  }
  C.n = nc;
  C.diag.assign(nc, 0.0);
  C.A.ptr.assign(1, 0);
  Accumulator a(nc);
  Csr row;
  for (int32_t I = 0; I < nc; ++I) {
    for (int32_t k = R.ptr[I]; k < R.ptr[I + 1]; ++k) {
      const int32_t i = R.col[k];
      const double r = R.val[k];
      for (int32_t e = AP.ptr[i]; e < AP.ptr[i + 1]; ++e) a.add(I, AP.col[e], r * AP.val[e]);
    }
    row.ptr.assign(1, 0);
    row.col.clear();
    row.val.clear();
    a.flush(row);
    for (size_t e = 0; e < row.col.size(); ++e) {
      if (row.col[e] == I) {
        C.diag[I] = row.val[e];
      } else {
        C.A.col.push_back(row.col[e]);
        C.A.val.push_back(-row.val[e]);
      }
    }
    C.A.ptr.push_back((int32_t)C.A.col.size());
  }
}

Hierarchy* handle_of(void* h) { return static_cast<Hierarchy*>(h); }

}  // namespace

extern "C" {

int obte_amg_create(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* w, const double* diag,
                    int32_t agg, int32_t coarse_max, double omega, int32_t smooth_max_row, void** handle) {
  if (!handle) return obte_fail(OBTE_EINVAL, "obte_amg_create: null handle");
  *handle = nullptr;
  if (n <= 0 || !row_ptr || !diag || (row_ptr[n] > 0 && (!col || !w)))
    return obte_fail(OBTE_EINVAL, "obte_amg_create: invalid operator");
  if (agg < 2 || coarse_max < 1 || !(omega > 0.0 && omega < 1.0))
    return obte_fail(OBTE_EINVAL, "obte_amg_create: agg >= 2, coarse_max >= 1 and 0 < omega < 1 required");
  for (int32_t i = 0; i < n; ++i) {
    if (row_ptr[i + 1] < row_ptr[i]) return obte_fail(OBTE_EINVAL, "obte_amg_create: row_ptr not monotone");
    if (!(diag[i] > 0.0)) return obte_fail(OBTE_EINVAL, "obte_amg_create: non-positive diagonal");
  }
  for (int32_t k = 0; k < row_ptr[n]; ++k)
    if (col[k] < 0 || col[k] >= n) return obte_fail(OBTE_EINVAL, "obte_amg_create: column out of range");
  try {
    std::unique_ptr<Hierarchy> H(new Hierarchy());
    H->levels.emplace_back();
    Level& L0 = H->levels.back();
    L0.n = n;
    L0.A.ptr.assign(row_ptr, row_ptr + n + 1);
    L0.A.col.assign(col, col + row_ptr[n]);
    L0.A.val.assign(w, w + row_ptr[n]);
    L0.diag.assign(diag, diag + n);
    while (H->levels.back().n > coarse_max) {
      Level& F = H->levels.back();
      const int32_t nc = (F.n + agg - 1) / agg;
      const bool smooth = (int64_t)F.A.col.size() <= (int64_t)smooth_max_row * F.n;
      smoothed_prolongation(F, agg, nc, smooth ? omega : 0.0, F.P);
      transpose(F.P, F.n, nc, F.R);
      Level C;
      galerkin(F, F.P, F.R, nc, C);
      for (int32_t I = 0; I < nc; ++I)
        if (!(C.diag[I] > 0.0)) return obte_fail(OBTE_EINVAL, "obte_amg_create: non-positive coarse diagonal");
      H->levels.push_back(std::move(C));
    }
    *handle = H.release();
  } catch (const std::bad_alloc&) {
    return obte_fail(OBTE_ENOMEM, "obte_amg_create: out of host memory");
  } catch (...) {
    return obte_fail(OBTE_EINVAL, "obte_amg_create: failed");
  }
  return OBTE_OK;
}

int obte_amg_sizes(void* handle, int32_t* n_levels, int64_t* sizes) {
  Hierarchy* H = handle_of(handle);
  if (!H || !n_levels) return obte_fail(OBTE_EINVAL, "obte_amg_sizes: null argument");
  *n_levels = (int32_t)H->levels.size();
  if (sizes) {
    for (size_t l = 0; l < H->levels.size(); ++l) {
      const Level& L = H->levels[l];
      sizes[4 * l + 0] = L.n;
      sizes[4 * l + 1] = (int64_t)L.A.col.size();
      sizes[4 * l + 2] = (int64_t)L.P.col.size();
      sizes[4 * l + 3] = (int64_t)L.R.col.size();
    }
  }
  return OBTE_OK;
}

int obte_amg_export(void* handle, int32_t level, int32_t* a_ptr, int32_t* a_col, double* a_w, double* a_diag,
                    int32_t* p_ptr, int32_t* p_col, double* p_val, int32_t* r_ptr, int32_t* r_col, double* r_val) {
  Hierarchy* H = handle_of(handle);
  if (!H || level < 0 || level >= (int32_t)H->levels.size())
    return obte_fail(OBTE_EINVAL, "obte_amg_export: null handle or level out of range");
  const Level& L = H->levels[level];
  auto put = [](const auto& v, auto* out) {
    if (out && !v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(v[0]));
  };
  put(L.A.ptr, a_ptr);
  put(L.A.col, a_col);
  put(L.A.val, a_w);
  put(L.diag, a_diag);
  put(L.P.ptr, p_ptr);
  put(L.P.col, p_col);
  put(L.P.val, p_val);
  put(L.R.ptr, r_ptr);
  put(L.R.col, r_col);
  put(L.R.val, r_val);
  return OBTE_OK;
}

int obte_amg_export_dense(void* handle, double* dense) {
  Hierarchy* H = handle_of(handle);
  if (!H || !dense) return obte_fail(OBTE_EINVAL, "obte_amg_export_dense: null argument");
  const Level& L = H->levels.back();
  const int64_t n = L.n;
  std::fill(dense, dense + n * n, 0.0);
  for (int32_t i = 0; i < L.n; ++i) {
    dense[i * n + i] = L.diag[i];
    for (int32_t k = L.A.ptr[i]; k < L.A.ptr[i + 1]; ++k) dense[i * n + L.A.col[k]] -= L.A.val[k];
  }
  return OBTE_OK;
}

int obte_amg_destroy(void* handle) {
  delete handle_of(handle);
  return OBTE_OK;
}

}  // extern "C"
