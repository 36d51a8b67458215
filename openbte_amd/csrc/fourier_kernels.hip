// FP64 HIP kernels of the Fourier initial guess (openbte/fourier.py:15-185).
//
// The reference factorises F = V^-1 Lap (one row pinned) with SuperLU and
// back-solves up to 40 times.  Here F T = b is solved in its symmetric form
// A T = V b with A = Lap (+ isothermal / modified diagonal), by
// Jacobi-preconditioned CG: every iteration is one CSR pass (~4 nnz per row)
// plus vector streams, i.e. HBM-bound sparse work.  The pinned row is
// replaced by the compatibility condition of the singular Laplacian; since the
// reference centres T by (max + min) / 2 right after the solve, the pin's
// constant drops out.
//
// Reductions are deterministic: per-block partial sums in a fixed tree, then
// the partials summed in order.  A CG iteration is two launches: the kernel
// after a dot product's partials sums them itself (every block, same order,
// same bits), so no launch is spent on a one-block reduction.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "capi_common.h"

namespace {

constexpr int kB = 256;

enum { S_RZ0 = 0, S_RZ1, S_ALPHA, S_BETA, S_RR, S_BB, S_TOT, S_MAX, S_MIN, S_ITER, S_TGT, S_DONE };

__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double apply_row(const obte_fourier_plan_t& P, const double* x, int i) {
  double v = P.diag[i] * x[i];
  for (int k = P.row_ptr[i]; k < P.row_ptr[i + 1]; ++k) v -= P.w[k] * x[P.col[k]];
  return v;
}

// bh = V (b + c); the pinned entry is set later to minus the sum of the rest.
__global__ void k_rhs(obte_fourier_plan_t P, const double* b, const double* c, double* bh, double* part) {
  __shared__ double red[kB];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < P.n) {
    v = P.vol[i] * (b[i] + (c ? c[i] : 0.0));
    if (i == P.pin) v = 0.0;
    bh[i] = v;
  }
  const double s = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_rhs_fix(obte_fourier_plan_t P, double* bh, const double* part, int nb) {
  __shared__ double red[1024];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0 && P.pin >= 0) bh[P.pin] = -s;
}

// Sum of v over the workgroup in a fixed order (xor tree inside each wave,
// then the wave totals in wave order); every thread gets the same bits.
// red: 16 doubles, free on entry (a barrier precedes the first write).
__device__ __forceinline__ double wg_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int j = 0; j < (int)(blockDim.x >> 6); ++j) s += red[j];
  return s;
}

// sum of the nb block partials, computed the same way by every block
__device__ __forceinline__ double sum_partials(const double* part, int nb, double* red) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
  return wg_sum(s, red);
}

// the sums of two partial arrays with one pair of barriers (each with the
// bits sum_partials gives it); red: 32 doubles
__device__ __forceinline__ void sum_partials2(const double* a, const double* b, int nb, double* red, double& sa,
                                              double& sb) {
  double u = 0.0, v = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    u += a[i];
    v += b[i];
  }
  for (int o = 32; o > 0; o >>= 1) {
    u += __shfl_xor(u, o, 64);
    v += __shfl_xor(v, o, 64);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = u;
    red[16 + (threadIdx.x >> 6)] = v;
  }
  __syncthreads();
  sa = 0.0;
  sb = 0.0;
  for (int j = 0; j < (int)(blockDim.x >> 6); ++j) {
    sa += red[j];
    sb += red[16 + j];
  }
}

// Partial sums of the multi-launch CG live after the five work vectors:
// [pq | rz | rr | bb] x nb doubles (4 nb <= n + 64 for every n).
// With the multigrid preconditioner the V-cycle's last kernel forms the
// preconditioned residual z and the r.z partials; without it z = r / diag
// is formed by the kernels that update r.

// r = bh - A x; partials of r.r and bh.bh, and (Jacobi) z = r / diag with the
// partials of r.z
__global__ void k_cg_init(obte_fourier_plan_t P, const double* bh, const double* x, double* r, double* z,
                          double* part, int nb) {
  __shared__ double red[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool jac = P.n_levels == 0;
  double rz = 0.0, rr = 0.0, bb = 0.0;
  if (i < P.n) {
    const double ri = bh[i] - apply_row(P, x, i);
    r[i] = ri;
    if (jac) {
      const double zi = ri / P.diag[i];
      z[i] = zi;
      rz = ri * zi;
    }
    rr = ri * ri;
    bb = bh[i] * bh[i];
  }
  if (jac) rz = wg_sum(rz, red);
  rr = wg_sum(rr, red);
  bb = wg_sum(bb, red);
  if (threadIdx.x == 0) {
    if (jac) part[nb + blockIdx.x] = rz;
    part[2 * nb + blockIdx.x] = rr;
    part[3 * nb + blockIdx.x] = bb;
  }
}

// scal[S_RR], scal[S_BB] from the partials (host convergence test); with
// with_bb (the initial residual) also the converged threshold
// scal[S_TGT] = (rtol |b|)^2 and scal[S_DONE] = 0
__global__ void k_cg_norms(const double* part, int nb, double* scal, int with_bb, double rtol) {
  __shared__ double red[16];
  const double rr = sum_partials(part + 2 * nb, nb, red);
  const double bb = with_bb ? sum_partials(part + 3 * nb, nb, red) : 0.0;
  if (threadIdx.x == 0) {
    scal[S_RR] = rr;
    if (with_bb) {
      scal[S_BB] = bb;
      const double t = rtol * (bb > 0.0 ? sqrt(bb) : 1.0);
      scal[S_TGT] = t * t;
      scal[S_DONE] = 0.0;
    }
  }
}

// Iteration k, first half: rz_k = r.z (every block sums the partials of the
// previous update), beta = rz_k / rz_(k-1), p_k = z + beta p_(k-1) formed on
// the fly for the row and its neighbours (p_(k-1) is read from pold, p_k
// written to pnew: no block reads what another writes), q = A p_k, partial
// p.q.  rz_k goes to scal slot k & 1 (slot (k - 1) & 1 holds rz_(k-1)).
// The row's loads (up to kRowHoist entries) are issued before the partial
// sums, whose barriers do not wait for them.
constexpr int kRowHoist = 4;

__global__ void k_cg_pq(obte_fourier_plan_t P, const double* z, const double* pold, double* pnew, double* q,
                        double* part, double* scal, int nb, int k) {
  __shared__ double red[32];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool row = i < P.n;
  const bool lag = k > 0;
  double di = 0.0, zi = 0.0, pi0 = 0.0;
  double we[kRowHoist], ze[kRowHoist], pe[kRowHoist];
  int e0 = 0, len = 0;
  if (row) {
    e0 = P.row_ptr[i];
    len = P.row_ptr[i + 1] - e0;
    di = P.diag[i];
    zi = z[i];
    pi0 = lag ? pold[i] : 0.0;
  }
#pragma unroll
  for (int e = 0; e < kRowHoist; ++e) {
    we[e] = 0.0;
    ze[e] = 0.0;
    pe[e] = 0.0;
    if (e < len) {
      const int j = P.col[e0 + e];
      we[e] = P.w[e0 + e];
      ze[e] = z[j];
      pe[e] = lag ? pold[j] : 0.0;
    }
  }
  // converged (the residual of the previous update below the host's
  // threshold): this and every later iteration leave x and r unchanged, so
  // the iterations run between two host tests cannot disturb a converged
  // solution.  Every block forms the same sums, so all decide alike.
  double rz, rr;
  sum_partials2(part + nb, part + 2 * nb, nb, red, rz, rr);
  const bool done = scal[S_DONE] != 0.0 || rr <= scal[S_TGT];
  double beta = 0.0;
  if (lag && !done) {
    const double rzo = scal[S_RZ0 + ((k - 1) & 1)];
    beta = rzo != 0.0 ? rz / rzo : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[S_RZ0 + (k & 1)] = rz;
    if (done) scal[S_DONE] = 1.0;
  }
  double pq = 0.0;
  if (row) {
    const double pi = lag ? zi + beta * pi0 : zi;
    double v = di * pi;
#pragma unroll
    for (int e = 0; e < kRowHoist; ++e)
      if (e < len) v -= we[e] * (lag ? ze[e] + beta * pe[e] : ze[e]);
    for (int e = kRowHoist; e < len; ++e) {  // rows longer than the hoisted part
      const int j = P.col[e0 + e];
      v -= P.w[e0 + e] * (lag ? z[j] + beta * pold[j] : z[j]);
    }
    pnew[i] = pi;
    q[i] = v;
    pq = pi * v;
  }
  pq = wg_sum(pq, red);
  if (threadIdx.x == 0) part[blockIdx.x] = pq;
}

// Iteration k, second half: alpha = rz_k / (p.q) (every block sums the p.q
// partials); x += alpha p, r -= alpha q; partials r.r and (Jacobi) z = r /
// diag with the partials of r.z.  The row's loads go out before the partial
// sums.
__global__ void k_cg_xr(obte_fourier_plan_t P, double* x, double* r, double* z, const double* p, const double* q,
                        double* part, const double* scal, int nb, int k) {
  __shared__ double red[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool row = i < P.n;
  const bool jac = P.n_levels == 0;
  double xi = 0.0, pi = 0.0, ri = 0.0, qi = 0.0, di = 1.0;
  if (row) {
    xi = x[i];
    pi = p[i];
    ri = r[i];
    qi = q[i];
    di = P.diag[i];
  }
  const double pq = sum_partials(part, nb, red);
  const double rz = scal[S_RZ0 + (k & 1)];
  const double alpha = pq != 0.0 && scal[S_DONE] == 0.0 ? rz / pq : 0.0;
  double rzn = 0.0, rr = 0.0;
  if (row) {
    x[i] = xi + alpha * pi;
    const double rn = ri - alpha * qi;
    r[i] = rn;
    if (jac) {
      const double zn = rn / di;
      z[i] = zn;
      rzn = rn * zn;
    }
    rr = rn * rn;
  }
  if (jac) rzn = wg_sum(rzn, red);
  rr = wg_sum(rr, red);
  if (threadIdx.x == 0) {
    if (jac) part[nb + blockIdx.x] = rzn;
    part[2 * nb + blockIdx.x] = rr;
  }
}

// ---------------------------------------------------------------------------
// Multigrid V-cycle z = M^-1 b (one per CG iteration, b = the CG residual on
// the finest level).  Per level: pre-smoothing from zero x = w b / d and its
// residual res = b - A x (k_mg_pre), restriction b_next = R res
// (k_mg_restrict); the dense coarsest level x = coarse_inv b (k_mg_coarse);
// on the way up y = x + P z_next and the post-smoothing z = y + w (b - A y) /
// d (k_mg_post; y of the neighbours is formed on the fly: the hierarchy keeps
// the prolongations of wide levels unsmoothed, so a row's neighbours and
// their prolongation entries stay few).  Every row sum runs in a fixed order.
// Per CG iteration: three launches per level and one for the coarsest.
// ---------------------------------------------------------------------------
constexpr double kMgOmega = 2.0 / 3.0;   // Jacobi weight of the smoother
constexpr int kMgCoarseMax = 4096;       // rows of the dense coarsest level at most

// one level's operator as kernel arguments (level 0: the plan's rows)
struct MgOp {
  int n;
  const int32_t* ptr;
  const int32_t* col;
  const double* w;
  const double* d;
};

__device__ __forceinline__ MgOp mg_op(const obte_amg_level_t& L) { return MgOp{L.n, L.a_ptr, L.a_col, L.a_w, L.a_diag}; }

// A row of a level is handled by G consecutive lanes (G = 1, 4 or 16 by the
// level's entries per row: the row loops are chains of dependent gathers, so
// wide rows are split across lanes to keep more loads in flight).  Lane g of
// the group takes entries g, g + G, ...; the partial sums are combined by an
// xor tree over the group (fixed order); every lane of the group gets the
// sum.  Out-of-range rows (i >= n) run with empty rows so that every lane
// reaches the shuffles.
template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// pre-smoothing of row i: x_i = w b_i / d_i, res_i = b_i - (A x)_i
template <int G>
__device__ __forceinline__ void mg_pre_row(const MgOp& A, const double* b, double* x, double* res, int i, int g) {
  const bool row = i < A.n;
  double acc = 0.0;
  const int k1 = row ? A.ptr[i + 1] : 0;
  for (int k = (row ? A.ptr[i] : 0) + g; k < k1; k += G) {
    const int j = A.col[k];
    acc += A.w[k] * (kMgOmega * b[j] / A.d[j]);
  }
  acc = group_sum<G>(acc);
  if (row && g == 0) {
    const double di = A.d[i];
    const double xi = kMgOmega * b[i] / di;
    x[i] = xi;
    res[i] = b[i] - (di * xi - acc);
  }
}

// restriction of row I: b_next_I = sum_k R_Ik res_k
template <int G>
__device__ __forceinline__ void mg_restrict_row(const obte_amg_level_t& L, int nn, const double* res, double* bn,
                                                int I, int g) {
  const bool row = I < nn;
  double s = 0.0;
  const int k1 = row ? L.r_ptr[I + 1] : 0;
  for (int k = (row ? L.r_ptr[I] : 0) + g; k < k1; k += G) s += L.r_val[k] * res[L.r_col[k]];
  s = group_sum<G>(s);
  if (row && g == 0) bn[I] = s;
}

// y_i = x_i + (P z_next)_i
__device__ __forceinline__ double mg_prolong(const obte_amg_level_t& L, const double* x, const double* zn, int i) {
  double y = x[i];
  for (int k = L.p_ptr[i]; k < L.p_ptr[i + 1]; ++k) y += L.p_val[k] * zn[L.p_col[k]];
  return y;
}

// post-smoothing of row i: z_i = y_i + w (b_i - (A y)_i) / d_i; returns z_i
// on lane 0 of the group (0 elsewhere and for rows past n)
template <int G>
__device__ __forceinline__ double mg_post_row(const MgOp& A, const obte_amg_level_t& L, const double* b,
                                              const double* x, const double* zn, int i, int g) {
  const bool row = i < A.n;
  double acc = 0.0;
  const int k1 = row ? A.ptr[i + 1] : 0;
  for (int k = (row ? A.ptr[i] : 0) + g; k < k1; k += G) acc += A.w[k] * mg_prolong(L, x, zn, A.col[k]);
  acc = group_sum<G>(acc);
  if (!row || g != 0) return 0.0;
  const double yi = mg_prolong(L, x, zn, i);
  const double di = A.d[i];
  return yi + kMgOmega * (b[i] - (di * yi - acc)) / di;
}

template <int G>
__global__ void k_mg_pre(MgOp A, const double* b, double* x, double* res) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  mg_pre_row<G>(A, b, x, res, t / G, t % G);
}

template <int G>
__global__ void k_mg_restrict(obte_amg_level_t L, int n_next, const double* res, double* bn) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  mg_restrict_row<G>(L, n_next, res, bn, t / G, t % G);
}

// rz: with `part` (the finest level, G = 1: b is the CG residual, one row
// per thread), the r.z partial of each block goes to part[nb + block]
template <int G>
__global__ void k_mg_post(MgOp A, obte_amg_level_t L, const double* b, const double* x, const double* zn, double* z,
                          double* part, int nb) {
  __shared__ double red[16];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / G, g = t % G;
  const double zi = mg_post_row<G>(A, L, b, x, zn, i, g);
  double rz = 0.0;
  if (i < A.n && g == 0) {
    z[i] = zi;
    rz = b[i] * zi;
  }
  if (part != nullptr) {
    rz = wg_sum(rz, red);
    if (threadIdx.x == 0) part[nb + blockIdx.x] = rz;
  }
}

// coarsest level: x = coarse_inv b, one wave per row (lanes stride over the
// row: coalesced loads), each row summed in a fixed order (lane partial sums,
// then an xor tree); four rows per block
constexpr int kMgCoarseRows = 4;

__global__ void k_mg_coarse(int nc, const double* cinv, const double* b, double* x) {
  const int lane = threadIdx.x & 63;
  const int I = blockIdx.x * kMgCoarseRows + (threadIdx.x >> 6);  // uniform over the wave
  if (I >= nc) return;
  const double* row = cinv + (int64_t)I * nc;
  double s = 0.0;
  for (int J = lane; J < nc; J += 64) s += row[J] * b[J];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) x[I] = s;
}

// ---------------------------------------------------------------------------
// Small systems: the whole solve (right-hand side, CG iterations with the
// residual test every iteration, centring) in one workgroup of 1024 threads,
// one launch.  A multi-launch iteration costs two launches (~11.5 us from 5k
// to 200k cells); here an iteration costs a few workgroup barriers and
// dependent L2 round trips, which grow with the rows per thread: measured
// ~9 us per iteration at 2k cells, ~12 us at 3k and ~14.5 us at 4k, so the
// one-workgroup solve takes systems up to 3072 cells.
// ---------------------------------------------------------------------------
constexpr int kCgBlock = 1024;
constexpr int kCgBlockMaxN = 3072;

// sums of NV values over the workgroup in a fixed order (xor tree inside each
// wave, then the wave totals in wave order); red: NV * 16 doubles
template <int NV>
__device__ __forceinline__ void wg_sums(double (&v)[NV], double* red) {
#pragma unroll
  for (int k = 0; k < NV; ++k)
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  const int w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  __syncthreads();  // the previous call's readers are done with red
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k * 16 + w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int j = 0; j < nw; ++j) s += red[k * 16 + j];
    v[k] = s;
  }
}

__global__ __launch_bounds__(kCgBlock) void k_cg_block(obte_fourier_plan_t P, const double* b, const double* c,
                                                        double* x, double* ws, double* scal, double rtol,
                                                        int maxit) {
  __shared__ double red[3 * 16];
  const int n = P.n;
  const int tid = threadIdx.x, nt = blockDim.x;
  double* bh = ws;
  double* r = ws + n;
  double* z = ws + 2 * (int64_t)n;
  double* p = ws + 3 * (int64_t)n;
  double* q = ws + 4 * (int64_t)n;
  // bh = V (b + c); the pinned entry is minus the sum of the rest
  double s1[1] = {0.0};
  for (int i = tid; i < n; i += nt) {
    const double v = i == P.pin ? 0.0 : P.vol[i] * (b[i] + (c ? c[i] : 0.0));
    bh[i] = v;
    s1[0] += v;
  }
  wg_sums<1>(s1, red);
  if (tid == 0 && P.pin >= 0) bh[P.pin] = -s1[0];
  __syncthreads();
  // r = bh - A x, z = r / diag, p = z
  double s3[3] = {0.0, 0.0, 0.0};  // r.z, r.r, bh.bh
  for (int i = tid; i < n; i += nt) {
    const double ri = bh[i] - apply_row(P, x, i);
    const double zi = ri / P.diag[i];
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    s3[0] += ri * zi;
    s3[1] += ri * ri;
    s3[2] += bh[i] * bh[i];
  }
  wg_sums<3>(s3, red);
  double rz = s3[0], rr = s3[1];
  const double bnorm = sqrt(s3[2]);
  const double target = rtol * (bnorm > 0.0 ? bnorm : 1.0);
  int it = 0;
  while (sqrt(rr) > target && it < maxit) {
    double pq[1] = {0.0};
    for (int i = tid; i < n; i += nt) {
      const double qi = apply_row(P, p, i);
      q[i] = qi;
      pq[0] += p[i] * qi;
    }
    wg_sums<1>(pq, red);
    const double alpha = pq[0] != 0.0 ? rz / pq[0] : 0.0;
    double s2[2] = {0.0, 0.0};
    for (int i = tid; i < n; i += nt) {
      x[i] += alpha * p[i];
      const double ri = r[i] - alpha * q[i];
      const double zi = ri / P.diag[i];
      r[i] = ri;
      z[i] = zi;
      s2[0] += ri * zi;
      s2[1] += ri * ri;
    }
    wg_sums<2>(s2, red);
    const double beta = rz != 0.0 ? s2[0] / rz : 0.0;
    rz = s2[0];
    rr = s2[1];
    for (int i = tid; i < n; i += nt) p[i] = z[i] + beta * p[i];
    __syncthreads();  // p of every row before the next product
    ++it;
  }
  if (P.pin >= 0) {  // (max + min) / 2 centring (openbte/fourier.py:155-156)
    double mx = -INFINITY, mn = INFINITY;
    for (int i = tid; i < n; i += nt) {
      mx = fmax(mx, x[i]);
      mn = fmin(mn, x[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
      mx = fmax(mx, __shfl_xor(mx, o, 64));
      mn = fmin(mn, __shfl_xor(mn, o, 64));
    }
    __syncthreads();
    if ((tid & 63) == 0) {
      red[tid >> 6] = mx;
      red[16 + (tid >> 6)] = mn;
    }
    __syncthreads();
    mx = -INFINITY;
    mn = INFINITY;
    for (int j = 0; j < (nt >> 6); ++j) {
      mx = fmax(mx, red[j]);
      mn = fmin(mn, red[16 + j]);
    }
    const double shift = (mx + mn) * 0.5;
    for (int i = tid; i < n; i += nt) x[i] -= shift;
  }
  if (tid == 0) {
    scal[S_RR] = rr;
    scal[S_BB] = s3[2];
    scal[S_ITER] = (double)it;
  }
}

// (max + min) / 2 centring (openbte/fourier.py:155-156)
__global__ void k_minmax(int n, const double* t, double* part, int nb) {
  __shared__ double mx[kB], mn[kB];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  mx[threadIdx.x] = i < n ? t[i] : -INFINITY;
  mn[threadIdx.x] = i < n ? t[i] : INFINITY;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + o]);
      mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = mx[0];
    part[nb + blockIdx.x] = mn[0];
  }
}

__global__ void k_minmax_fin(const double* part, int nb, double* scal) {
  __shared__ double mx[1024], mn[1024];
  double a = -INFINITY, b = INFINITY;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    a = fmax(a, part[i]);
    b = fmin(b, part[nb + i]);
  }
  mx[threadIdx.x] = a;
  mn[threadIdx.x] = b;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + o]);
      mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    scal[S_MAX] = mx[0];
    scal[S_MIN] = mn[0];
  }
}

__global__ void k_shift(int n, double* t, const double* scal) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] -= (scal[S_MAX] + scal[S_MIN]) * 0.5;
}

__global__ void k_fill(int64_t n, double* a, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Workspace of obte_fourier_solve_noalloc: the solve's 6 n + 64 doubles plus
// its 16 scalars, for n up to OBTE_FOURIER_NOALLOC_MAX_CELLS (12 MiB).  Host
// code never names this array in a launch: it reaches it through the address
// hipGetSymbolAddress returns on each device (noalloc_scratch()).
constexpr int64_t kNoAllocWs = 6 * (int64_t)OBTE_FOURIER_NOALLOC_MAX_CELLS + 64;
constexpr int64_t kNoAllocDoubles = kNoAllocWs + 16;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

__global__ void k_gradient(obte_fourier_plan_t P, const double* t, double* grad) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.n) return;
  const int nv = P.nv;
  const int dim = P.dim;
  double dd[4];
  const double tc = t[c];
  for (int s = 0; s < nv; ++s) {
    const int m = P.slot_mode[c * nv + s];
    double v = 0.0;
    if (m == 1) v = t[P.slot_cell[c * nv + s]] - tc + P.slot_val[c * nv + s];
    else if (m == 2) v = P.slot_val[c * nv + s] - tc;
    dd[s] = v;
  }
  for (int d = 0; d < dim; ++d) {
    const double* w = P.gweights + ((int64_t)c * dim + d) * nv;
    double g = 0.0;
    for (int s = 0; s < nv; ++s) g += w[s] * dd[s];
    grad[(int64_t)c * dim + d] = g;
  }
}

__global__ void k_side_flux(obte_fourier_plan_t P, const double* grad, double* tmp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.n_int) return;
  const int dim = P.dim;
  const int a = P.side_e[2 * s], b = P.side_e[2 * s + 1];
  const double w = P.side_w[s];
  double v = 0.0;
  for (int d = 0; d < dim; ++d)
    v += (w * grad[(int64_t)a * dim + d] + (1.0 - w) * grad[(int64_t)b * dim + d]) * P.side_vn[(int64_t)s * dim + d];
  tmp[s] = v;
}

__global__ void k_correction(obte_fourier_plan_t P, const double* tmp, double* c, double* part) {
  __shared__ double red[kB];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < P.n) {
    for (int s = 0; s < P.nv; ++s) {
      const int sd = P.slot_side[i * P.nv + s];
      if (sd >= 0) v += P.slot_sgn[i * P.nv + s] * tmp[sd];
    }
    c[i] = v;
  }
  const double s2 = block_sum(v * v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s2;
}

__global__ void k_sum_fin(const double* part, int nb, double* out) {
  __shared__ double red[1024];
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) a += part[i];
  a = block_sum(a, red);
  if (threadIdx.x == 0) *out = a;
}

__global__ void k_kappa(obte_fourier_plan_t P, const double* t, double* out) {
  __shared__ double red[1024];
  double a = 0.0;
  for (int s = threadIdx.x; s < P.n_contact; s += blockDim.x) {
    const double t1 = t[P.ce1[s]];
    const double t2 = P.contact_kind == 1 ? t[P.ce2[s]] + P.c_value : P.c_value;
    a += (t1 - t2) * P.c_coef[s];
  }
  a = block_sum(a, red);
  if (threadIdx.x == 0) *out = a * P.normalization;
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

int status(const char* who) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return obte_fail_hip(who, e);
  return OBTE_OK;
}

int check(const obte_fourier_plan_t* P) {
  if (!P || P->n <= 0 || (P->dim != 2 && P->dim != 3) || P->nv <= 0 || P->nv > 4)
    return obte_fail(OBTE_EINVAL, "fourier plan: invalid sizes");
  if (P->pin >= P->n) return obte_fail(OBTE_EINVAL, "fourier plan: pin out of range");
  if (P->n_levels < 0) return obte_fail(OBTE_EINVAL, "fourier plan: n_levels < 0");
  return OBTE_OK;
}

// The multigrid hierarchy fits the kernels: level 0 is the plan's operator,
// every level has its operator (above level 0), transfer operators and work
// vectors, and the coarsest level its inverse.
int check_mg(const obte_fourier_plan_t& P) {
  const obte_amg_level_t* L = P.levels;
  if (!L || L[0].n != P.n) return obte_fail(OBTE_EINVAL, "fourier plan: levels[0].n must equal n");
  if (P.n_coarse <= 0 || !P.coarse_inv || !P.coarse_b || !P.coarse_x)
    return obte_fail(OBTE_EINVAL, "fourier plan: multigrid needs the coarsest level and its inverse");
  if (P.n_coarse > kMgCoarseMax) return obte_fail(OBTE_ELIMIT, "fourier plan: coarsest level over 4096 rows");
  for (int l = 0; l < P.n_levels; ++l) {
    if (L[l].n <= 0 || !L[l].p_ptr || !L[l].p_col || !L[l].p_val || !L[l].r_ptr || !L[l].r_col || !L[l].r_val ||
        !L[l].x || !L[l].res)
      return obte_fail(OBTE_EINVAL, "fourier plan: incomplete multigrid level");
    if (l > 0 && (!L[l].a_ptr || !L[l].a_col || !L[l].a_w || !L[l].a_diag || !L[l].b || !L[l].z))
      return obte_fail(OBTE_EINVAL, "fourier plan: incomplete multigrid level");
  }
  return OBTE_OK;
}

// z = M^-1 r: one V-cycle; the r.z partials go to part[nb ..].
void vcycle(const obte_fourier_plan_t& P, const double* r, double* z, double* part, int nb, hipStream_t st) {
  const obte_amg_level_t* L = P.levels;
  const int nl = P.n_levels;
  auto op = [&](int l) {
    return l == 0 ? MgOp{P.n, P.row_ptr, P.col, P.w, P.diag} : MgOp{L[l].n, L[l].a_ptr, L[l].a_col, L[l].a_w, L[l].a_diag};
  };
  // lanes per row from a level's entries per row (level 0 stays at one row
  // per thread: its post kernel forms the CG's r.z partials per block of rows)
  auto lanes = [](int64_t entries, int rows, bool finest) {
    if (finest || entries <= 6 * (int64_t)rows) return 1;
    return entries <= 24 * (int64_t)rows ? 4 : 16;
  };
  auto grid = [](int rows, int g) { return dim3((unsigned)(((int64_t)rows * g + kB - 1) / kB)); };
  for (int l = 0; l < nl; ++l) {
    const double* b = l == 0 ? r : L[l].b;
    const int nn = l + 1 < nl ? L[l + 1].n : P.n_coarse;
    double* bn = l + 1 < nl ? L[l + 1].b : P.coarse_b;
    const MgOp A = op(l);
    const int g_a = lanes(L[l].a_entries, A.n, l == 0);
    const int g_r = lanes(L[l].r_entries, nn, false);
    if (g_a == 1) hipLaunchKernelGGL(k_mg_pre<1>, grid(A.n, 1), dim3(kB), 0, st, A, b, L[l].x, L[l].res);
    else if (g_a == 4) hipLaunchKernelGGL(k_mg_pre<4>, grid(A.n, 4), dim3(kB), 0, st, A, b, L[l].x, L[l].res);
    else hipLaunchKernelGGL(k_mg_pre<16>, grid(A.n, 16), dim3(kB), 0, st, A, b, L[l].x, L[l].res);
    const double* res = L[l].res;
    if (g_r == 1) hipLaunchKernelGGL(k_mg_restrict<1>, grid(nn, 1), dim3(kB), 0, st, L[l], nn, res, bn);
    else if (g_r == 4) hipLaunchKernelGGL(k_mg_restrict<4>, grid(nn, 4), dim3(kB), 0, st, L[l], nn, res, bn);
    else hipLaunchKernelGGL(k_mg_restrict<16>, grid(nn, 16), dim3(kB), 0, st, L[l], nn, res, bn);
  }
  hipLaunchKernelGGL(k_mg_coarse, dim3((P.n_coarse + kMgCoarseRows - 1) / kMgCoarseRows), dim3(64 * kMgCoarseRows),
                     0, st, P.n_coarse, P.coarse_inv, (const double*)P.coarse_b, P.coarse_x);
  for (int l = nl - 1; l >= 0; --l) {
    const double* b = l == 0 ? r : L[l].b;
    const double* zn = l + 1 < nl ? L[l + 1].z : P.coarse_x;
    double* zl = l == 0 ? z : L[l].z;
    const MgOp A = op(l);
    const double* x = L[l].x;
    const int g_a = lanes(L[l].a_entries, A.n, l == 0);
    if (g_a == 1)
      hipLaunchKernelGGL(k_mg_post<1>, grid(A.n, 1), dim3(kB), 0, st, A, L[l], b, x, zn, zl, l == 0 ? part : nullptr,
                         nb);
    else if (g_a == 4)
      hipLaunchKernelGGL(k_mg_post<4>, grid(A.n, 4), dim3(kB), 0, st, A, L[l], b, x, zn, zl, nullptr, nb);
    else
      hipLaunchKernelGGL(k_mg_post<16>, grid(A.n, 16), dim3(kB), 0, st, A, L[l], b, x, zn, zl, nullptr, nb);
  }
}

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array at its own address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return obte_fail_hip("obte_fourier_solve_noalloc: hipGetDevice", e);
  if (dev < 0 || dev >= kMaxDevices) return obte_fail(OBTE_ELIMIT, "obte_fourier_solve_noalloc: device index >= 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return obte_fail_hip("obte_fourier_solve_noalloc: hipGetSymbolAddress", e);
    if (!p) return obte_fail(OBTE_EHIP, "obte_fourier_solve_noalloc: null scratch address");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return OBTE_OK;
}

}  // namespace

extern "C" {

int obte_fourier_solve(const obte_fourier_plan_t* plan, const double* b, const double* c, double* t, double* ws,
                       double* scal, double rtol, int32_t maxit, int32_t check_every, int32_t* iters_out,
                       double* relres_out, obte_stream_t stream) {
  int rc = check(plan);
  if (rc) return rc;
  if (!b || !t || !ws || !scal) return obte_fail(OBTE_EINVAL, "obte_fourier_solve: null pointer");
  const obte_fourier_plan_t& P = *plan;
  hipStream_t st = (hipStream_t)stream;
  const int n = P.n;
  const unsigned nb = nblocks(n);
  double* bh = ws;
  double* r = ws + n;
  double* z = ws + 2 * (int64_t)n;
  double* p = ws + 3 * (int64_t)n;
  double* q = ws + 4 * (int64_t)n;
  double* part = ws + 5 * (int64_t)n;  // n + 64 doubles >= 6 * nb
  if (check_every <= 0) check_every = 16;
  double host[16];

  if (n <= kCgBlockMaxN) {
    hipLaunchKernelGGL(k_cg_block, dim3(1), dim3(kCgBlock), 0, st, P, b, c, t, ws, scal, rtol, (int)maxit);
    rc = status("obte_fourier_solve: launch failed");
    if (rc) return rc;
    if (hipMemcpyAsync(host, scal, 16 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return obte_fail(OBTE_EHIP, "obte_fourier_solve: readback failed");
    const double bn = sqrt(host[S_BB]);
    if (iters_out) *iters_out = (int32_t)host[S_ITER];
    if (relres_out) *relres_out = sqrt(host[S_RR]) / (bn > 0.0 ? bn : 1.0);
    return OBTE_OK;
  }

  if (P.n_levels > 0) {
    rc = check_mg(P);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_rhs, dim3(nb), dim3(kB), 0, st, P, b, c, bh, part);
  if (P.pin >= 0) hipLaunchKernelGGL(k_rhs_fix, dim3(1), dim3(1024), 0, st, P, bh, part, (int)nb);
  // bh is dead after the initial residual: it holds the second direction buffer
  double* pbuf[2] = {p, bh};
  hipLaunchKernelGGL(k_cg_init, dim3(nb), dim3(kB), 0, st, P, bh, t, r, z, part, (int)nb);
  if (P.n_levels > 0) vcycle(P, r, z, part, (int)nb, st);
  hipLaunchKernelGGL(k_cg_norms, dim3(1), dim3(1024), 0, st, part, (int)nb, scal, 1, rtol);
  rc = status("obte_fourier_solve: init launch failed");
  if (rc) return rc;
  if (hipMemcpyAsync(host, scal, 8 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return obte_fail(OBTE_EHIP, "obte_fourier_solve: readback failed");
  const double bnorm = sqrt(host[S_BB]);
  double rnorm = sqrt(host[S_RR]);
  int it = 0;
  const double target = rtol * (bnorm > 0.0 ? bnorm : 1.0);
  while (rnorm > target && it < maxit) {
    for (int k = 0; k < check_every && it < maxit; ++k, ++it) {
      double* pk = pbuf[it & 1];
      hipLaunchKernelGGL(k_cg_pq, dim3(nb), dim3(kB), 0, st, P, z, pbuf[(it + 1) & 1], pk, q, part, scal, (int)nb,
                         it);
      hipLaunchKernelGGL(k_cg_xr, dim3(nb), dim3(kB), 0, st, P, t, r, z, pk, q, part, scal, (int)nb, it);
      if (P.n_levels > 0) vcycle(P, r, z, part, (int)nb, st);
    }
    hipLaunchKernelGGL(k_cg_norms, dim3(1), dim3(1024), 0, st, part, (int)nb, scal, 0, rtol);
    rc = status("obte_fourier_solve: iteration launch failed");
    if (rc) return rc;
    if (hipMemcpyAsync(host, scal, 8 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return obte_fail(OBTE_EHIP, "obte_fourier_solve: readback failed");
    rnorm = sqrt(host[S_RR]);
  }
  if (P.pin >= 0) {
    hipLaunchKernelGGL(k_minmax, dim3(nb), dim3(kB), 0, st, n, t, part, (int)nb);
    hipLaunchKernelGGL(k_minmax_fin, dim3(1), dim3(1024), 0, st, part, (int)nb, scal);
    hipLaunchKernelGGL(k_shift, dim3(nb), dim3(kB), 0, st, n, t, scal);
  }
  if (iters_out) *iters_out = it;
  if (relres_out) *relres_out = rnorm / (bnorm > 0.0 ? bnorm : 1.0);
  return status("obte_fourier_solve: finish launch failed");
}

int obte_fourier_solve_noalloc(const obte_fourier_plan_t* plan, const double* b, const double* c, double* t,
                               double rtol, int32_t maxit, int32_t check_every, int32_t* iters_out,
                               double* relres_out, obte_stream_t stream) {
  int rc = check(plan);
  if (rc) return rc;
  if (plan->n > OBTE_FOURIER_NOALLOC_MAX_CELLS)
    return obte_fail(OBTE_ELIMIT, "obte_fourier_solve_noalloc: more cells than OBTE_FOURIER_NOALLOC_MAX_CELLS "
                                  "(262144); use obte_fourier_solve with a caller workspace");
  double* ws = nullptr;
  rc = noalloc_scratch(&ws);
  if (rc) return rc;
  double* scal = ws + kNoAllocWs;
  // the scalars of the previous call on this device are stale: start from zero
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, (hipStream_t)stream, (int64_t)16, scal, 0.0);
  rc = status("obte_fourier_solve_noalloc: fill launch failed");
  if (rc) return rc;
  return obte_fourier_solve(plan, b, c, t, ws, scal, rtol, maxit, check_every, iters_out, relres_out, stream);
}

int obte_fourier_gradient(const obte_fourier_plan_t* plan, const double* t, double* grad, obte_stream_t stream) {
  int rc = check(plan);
  if (rc) return rc;
  if (!t || !grad) return obte_fail(OBTE_EINVAL, "obte_fourier_gradient: null pointer");
  hipLaunchKernelGGL(k_gradient, dim3(nblocks(plan->n)), dim3(kB), 0, (hipStream_t)stream, *plan, t, grad);
  return status("obte_fourier_gradient: launch failed");
}

int obte_fourier_correction(const obte_fourier_plan_t* plan, const double* grad, double* c, double* side_tmp,
                            double* ws, double* norm2, obte_stream_t stream) {
  int rc = check(plan);
  if (rc) return rc;
  if (!grad || !c || !side_tmp || !ws || !norm2) return obte_fail(OBTE_EINVAL, "obte_fourier_correction: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (plan->n_int > 0)
    hipLaunchKernelGGL(k_side_flux, dim3(nblocks(plan->n_int)), dim3(kB), 0, st, *plan, grad, side_tmp);
  const unsigned nb = nblocks(plan->n);
  hipLaunchKernelGGL(k_correction, dim3(nb), dim3(kB), 0, st, *plan, side_tmp, c, ws);
  hipLaunchKernelGGL(k_sum_fin, dim3(1), dim3(1024), 0, st, ws, (int)nb, norm2);
  return status("obte_fourier_correction: launch failed");
}

int obte_fourier_kappa(const obte_fourier_plan_t* plan, const double* t, double* out, obte_stream_t stream) {
  int rc = check(plan);
  if (rc) return rc;
  if (!t || !out) return obte_fail(OBTE_EINVAL, "obte_fourier_kappa: null pointer");
  hipLaunchKernelGGL(k_kappa, dim3(1), dim3(1024), 0, (hipStream_t)stream, *plan, t, out);
  return status("obte_fourier_kappa: launch failed");
}

}  // extern "C"
