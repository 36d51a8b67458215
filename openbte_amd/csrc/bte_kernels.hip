// FP64 HIP kernels of the steady-state mode-resolved BTE solve on gfx950.
//
// Data layout in HBM (one geometry, L directions, M MFP bins, chunks of MC
// bins):
//   * every per-direction array is in sweep ("position") order, so the cells
//     of one level are contiguous: order/diag/pert/q [L][n], inflow CSR
//     in_ptr [L][n+1] + entries in_src/in_coef;
//   * the solution workspace T is [L][chunk][n][MC]: one workgroup owns one
//     (direction, chunk) slab and each cell's MC bins are one 8*MC-byte run, so
//     the upwind gathers are coalesced over the MFP lanes;
//   * interface vectors (cut-face inflows / exit values) are [chunk][cut][MC].
//
// The sweep is one workgroup per (direction, chunk): levels are separated by
// __syncthreads only, there is no inter-workgroup waiting of any kind.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "capi_common.h"

namespace {

constexpr int kMaxBlock = 1024;

__device__ __forceinline__ double dotv(const double* a, const double* b, int dim) {
  double g = a[0] * b[0] + a[1] * b[1];
  if (dim == 3) g += a[2] * b[2];
  return g;
}

// Sum of v over the MC consecutive lanes of one cell (fixed xor-tree order).
template <int MC>
__device__ __forceinline__ double lane_group_sum(double v) {
#pragma unroll
  for (int o = MC / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, MC);
  return v;
}

// ---------------------------------------------------------------------------
// (a) face-flux assembly: D_l, inflow coefficients, perturbation
//     (openbte/bte.py:47-124), one thread per (direction, position).
// ---------------------------------------------------------------------------
__global__ void k_assemble(obte_bte_plan_t P) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)P.n_dirs * P.n;
  if (gid >= total) return;
  const int l = (int)(gid / P.n);
  const int p = (int)(gid - (int64_t)l * P.n);
  const int dim = P.dim;
  const double* s = P.dirs + (int64_t)l * dim;
  const int c = P.order[gid];
  double D = 0.0, pert = 0.0;
  // internal directed faces of row c: outflow part into D, periodic jump term
  for (int q = P.cf_ptr[c]; q < P.cf_ptr[c + 1]; ++q) {
    const int f = P.cf_face[q];
    const double g = dotv(s, P.k_face + (int64_t)f * dim, dim);
    if (g > 0.0) D += g;
    else pert += P.jump_face[f] * g;
  }
  // diffuse and isothermal walls: outflow into D, isothermal inflow into pert
  for (int w = P.wall_ptr[c]; w < P.wall_ptr[c + 1]; ++w) {
    const double g = dotv(s, P.wall_db + (int64_t)w * dim, dim);
    if (g > 0.0) D += g;
    else if (P.wall_side[w] < 0) pert -= g * P.wall_temp[w];
  }
  if (P.nano != 0.0 && dim == 3) pert -= s[2];
  P.diag[gid] = D;
  P.pert[gid] = pert;
  const int32_t* inptr = P.in_ptr + (int64_t)l * (P.n + 1);
  const int64_t eb = P.in_off[l];
  for (int e = inptr[p]; e < inptr[p + 1]; ++e) {
    const int f = P.in_face[eb + e];
    const double g = dotv(s, P.k_face + (int64_t)f * dim, dim);
    P.in_coef[eb + e] = g < 0.0 ? -g : 0.0;
  }
}

// q = pert + sum over diffuse walls of (-TB_s) min(s . db, 0)   (openbte/bte.py:199-200,214)
__global__ void k_source(obte_bte_plan_t P) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)P.n_dirs * P.n;
  if (gid >= total) return;
  const int l = (int)(gid / P.n);
  const int dim = P.dim;
  const double* s = P.dirs + (int64_t)l * dim;
  const int c = P.order[gid];
  double q = P.pert[gid];
  for (int w = P.wall_ptr[c]; w < P.wall_ptr[c + 1]; ++w) {
    const int sd = P.wall_side[w];
    if (sd < 0) continue;
    const double g = dotv(s, P.wall_db + (int64_t)w * dim, dim);
    if (g < 0.0) q += -P.tb[sd] * g;
  }
  P.q[gid] = q;
}

// ---------------------------------------------------------------------------
// (b) level-scheduled upwind sweep, (c) fused mode sums
// ---------------------------------------------------------------------------
template <int MC, int MODE>
__global__ __launch_bounds__(kMaxBlock) void k_sweep(obte_bte_plan_t P, const double* __restrict__ y,
                                                     double* __restrict__ x, const int32_t* __restrict__ active) {
  __shared__ double red[kMaxBlock];
  const int task = blockIdx.x;
  if (active != nullptr && active[task] == 0) return;  // uniform over the block
  const int L = P.n_dirs;
  const int l = task / P.n_chunks;
  const int chunk = task - l * P.n_chunks;
  const int n = P.n;
  const int dim = P.dim;
  const int lane = threadIdx.x % MC;
  const int slot = threadIdx.x / MC;
  const int cpp = blockDim.x / MC;
  const int mg = chunk * MC + lane;
  const bool valid = mg < P.n_mfp;
  const int mm = valid ? mg : P.n_mfp - 1;
  const double lam = valid ? P.mfps[mm] : 0.0;

  double* __restrict__ T = P.t + (int64_t)task * n * MC;
  const int32_t* __restrict__ order = P.order + (int64_t)l * n;
  const int32_t* __restrict__ inptr = P.in_ptr + (int64_t)l * (n + 1);
  const int32_t* __restrict__ src_e = P.in_src + P.in_off[l];
  const double* __restrict__ coef_e = P.in_coef + P.in_off[l];
  const int32_t* __restrict__ lvl = P.lvl_ptr + P.lvl_off[l];
  const int nlev = P.lvl_off[l + 1] - P.lvl_off[l] - 1;
  const int cbase = P.cut_off[l];
  const int ncut = P.cut_off[l + 1] - cbase;
  const double* __restrict__ yv = y + ((int64_t)chunk * P.n_cut + cbase) * MC;
  const double* __restrict__ diag = P.diag + (int64_t)l * n;
  const double* __restrict__ qv = P.q + (int64_t)l * n;

  double tco = 0.0, sg0 = 0.0, sg1 = 0.0, sg2 = 0.0;
  if (MODE == 2 && valid) {
    tco = P.t_coeff[(int64_t)mm * L + l];
    const double* sg = P.sigma + ((int64_t)mm * L + l) * dim;
    sg0 = sg[0];
    sg1 = sg[1];
    if (dim == 3) sg2 = sg[2];
  }

  for (int k = 0; k < nlev; ++k) {
    const int p1 = lvl[k + 1];
    for (int p = lvl[k] + slot; p < p1; p += cpp) {
      const int c = order[p];
      double b = 0.0;
      double dtc = 0.0;
      if (MODE != 1) {
        dtc = P.delta_t[c];
        b = dtc + P.heat[c] + lam * qv[p];
      }
      double acc = 0.0;
      const int e1 = inptr[p + 1];
      for (int e = inptr[p]; e < e1; ++e) {
        const int sidx = src_e[e];
        const double v = sidx >= 0 ? T[(int64_t)sidx * MC + lane] : yv[(int64_t)(~sidx) * MC + lane];
        acc += coef_e[e] * v;
      }
      const double t = (b + lam * acc) / (1.0 + lam * diag[p]);
      T[(int64_t)p * MC + lane] = t;
      if (MODE == 2) {
        const double a0 = lane_group_sum<MC>(tco * t);
        const double dj = t - dtc;
        const double j0 = lane_group_sum<MC>(dj * sg0);
        const double j1 = lane_group_sum<MC>(dj * sg1);
        double j2 = 0.0;
        if (dim == 3) j2 = lane_group_sum<MC>(dj * sg2);
        if (lane == 0) {
          const int64_t o = (int64_t)task * n + c;
          P.part_t[o] = a0;
          P.part_j[o * dim + 0] = j0;
          P.part_j[o * dim + 1] = j1;
          if (dim == 3) P.part_j[o * dim + 2] = j2;
        }
      }
    }
    __syncthreads();
  }

  // exit values of the cut faces (upstream cells across the cut)
  if (x != nullptr) {
    double* __restrict__ xv = x + ((int64_t)chunk * P.n_cut + cbase) * MC;
    const int32_t* __restrict__ ex = P.cut_exit + cbase;
    for (int f = slot; f < ncut; f += cpp) xv[(int64_t)f * MC + lane] = T[(int64_t)ex[f] * MC + lane];
  }

  if (MODE == 2) {
    const int32_t* __restrict__ pos = P.pos + (int64_t)l * n;
    // wall temperatures: TB_s = -sum_ml T_ml[i_s] GG_ml,s  (openbte/bte.py:73-78,227)
    for (int sd = slot; sd < P.n_wall; sd += cpp) {
      const double* db = P.wall_dbs + (int64_t)sd * dim;
      double gb = sg0 * db[0] + sg1 * db[1];
      if (dim == 3) gb += sg2 * db[2];
      const double gg = (gb > 0.0 ? gb : 0.0) * P.wall_rinv[sd];
      const double v = T[(int64_t)pos[P.wall_cell[sd]] * MC + lane];
      const double tb = lane_group_sum<MC>(-v * gg);
      if (lane == 0) P.part_tb[(int64_t)task * P.n_wall + sd] = tb;
    }
    // kappa and suppression of each mode (openbte/bte.py:232-243)
    if (P.contact_kind != 0) {
      double kap = 0.0, sup = 0.0;
      for (int sc = slot; sc < P.n_contact; sc += cpp) {
        const double* na = P.c_na + (int64_t)sc * dim;
        double sn = sg0 * na[0] + sg1 * na[1];
        double area2 = na[0] * na[0] + na[1] * na[1];
        if (dim == 3) {
          sn += sg2 * na[2];
          area2 += na[2] * na[2];
        }
        const double t1 = T[(int64_t)pos[P.ce1[sc]] * MC + lane];
        const double snp = sn > 0.0 ? sn : 0.0;
        const double snm = sn < 0.0 ? sn : 0.0;
        kap += t1 * snp + P.contact_jump * snm;
        if (P.contact_kind == 1) kap += T[(int64_t)pos[P.ce2[sc]] * MC + lane] * snm;
        sup += t1 * sqrt(area2);
      }
      red[threadIdx.x] = kap;
      __syncthreads();
      double ksum = 0.0;
      if (slot == 0) {
        for (int s2 = 0; s2 < cpp; ++s2) ksum += red[s2 * MC + lane];
      }
      __syncthreads();
      red[threadIdx.x] = sup;
      __syncthreads();
      if (slot == 0) {
        double ssum = 0.0;
        for (int s2 = 0; s2 < cpp; ++s2) ssum += red[s2 * MC + lane];
        if (valid) {
          P.kappa[(int64_t)mm * L + l] = ksum * P.normalization;
          P.supp[(int64_t)mm * L + l] = ssum * P.normalization;
        }
      }
    }
  }
}

// Block-wide per-lane sum over the slots (fixed order); result broadcast.
template <int MC>
__device__ double block_lane_sum(double v, double* red, double* bc) {
  const int lane = threadIdx.x % MC;
  const int slot = threadIdx.x / MC;
  const int cpp = blockDim.x / MC;
  red[threadIdx.x] = v;
  __syncthreads();
  if (slot == 0) {
    double s = 0.0;
    for (int q = 0; q < cpp; ++q) s += red[q * MC + lane];
    bc[lane] = s;
  }
  __syncthreads();
  const double r = bc[lane];
  __syncthreads();
  return r;
}

template <int MC>
__device__ double block_lane_max(double v, double* red, double* bc) {
  const int lane = threadIdx.x % MC;
  const int slot = threadIdx.x / MC;
  const int cpp = blockDim.x / MC;
  red[threadIdx.x] = v;
  __syncthreads();
  if (slot == 0) {
    double s = 0.0;
    for (int q = 0; q < cpp; ++q) s = fmax(s, red[q * MC + lane]);
    bc[lane] = s;
  }
  __syncthreads();
  const double r = bc[lane];
  __syncthreads();
  return r;
}

struct GcrView {
  int64_t stride;  // elements of one interface vector set = n_chunks * n_cut * MC
  int64_t base;    // offset of this (l, chunk) block
  int ncut;
};

__device__ __forceinline__ GcrView gcr_view(const obte_bte_plan_t& P, int l, int chunk, int MC) {
  GcrView v;
  v.stride = (int64_t)P.n_chunks * P.n_cut * MC;
  v.base = ((int64_t)chunk * P.n_cut + P.cut_off[l]) * MC;
  v.ncut = P.cut_off[l + 1] - P.cut_off[l];
  return v;
}

// Convergence test of one lane; updates flags.  Returns 1 if this lane stays active.
template <int MC>
__device__ int gcr_flags(const obte_bte_plan_t& P, const obte_gcr_t& G, int task, int l, int chunk,
                         const GcrView& v, double rn2, double ymax, double* red, double* bc) {
  const int lane = threadIdx.x % MC;
  const int slot = threadIdx.x / MC;
  const int mg = chunk * MC + lane;
  const bool valid = mg < P.n_mfp;
  const double rn = sqrt(rn2);
  const double tol = G.rtol * fmax(1.0, ymax) * sqrt((double)(v.ncut > 0 ? v.ncut : 1));
  const int act = (valid && v.ncut > 0 && rn > tol) ? 1 : 0;
  // any active lane in the block?
  const double any = block_lane_max<MC>(act ? 1.0 : 0.0, red, bc);
  __shared__ int blk_any;
  if (threadIdx.x == 0) blk_any = 0;
  __syncthreads();
  if (slot == 0) {
    G.lane_active[(int64_t)task * MC + lane] = act;
    if (valid) G.rnorm[(int64_t)mg * P.n_dirs + l] = rn;
    if (any > 0.0) atomicOr(&blk_any, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    G.active[task] = blk_any;
    if (blk_any) atomicAdd(G.n_active, 1);
  }
  return act;
}

// r = x - y; MGS projection of r on the recycled basis; y += U a.
template <int MC>
__global__ __launch_bounds__(kMaxBlock) void k_gcr_project(obte_bte_plan_t P, obte_gcr_t G) {
  __shared__ double red[kMaxBlock];
  __shared__ double bc[64];
  const int task = blockIdx.x;
  const int l = task / P.n_chunks;
  const int chunk = task - l * P.n_chunks;
  const int lane = threadIdx.x % MC;
  const int slot = threadIdx.x / MC;
  const int cpp = blockDim.x / MC;
  const GcrView v = gcr_view(P, l, chunk, MC);
  double* y = G.y + v.base;
  double* r = G.r + v.base;
  const double* x = G.x + v.base;
  for (int f = slot; f < v.ncut; f += cpp) {
    const int64_t o = (int64_t)f * MC + lane;
    r[o] = x[o] - y[o];
  }
  const int kb = G.k[task];
  for (int j = 0; j < kb; ++j) {
    const double* cj = G.c + j * v.stride + v.base;
    const double* uj = G.u + j * v.stride + v.base;
    double part = 0.0;
    for (int f = slot; f < v.ncut; f += cpp) {
      const int64_t o = (int64_t)f * MC + lane;
      part += cj[o] * r[o];
    }
    const double a = block_lane_sum<MC>(part, red, bc);
    for (int f = slot; f < v.ncut; f += cpp) {
      const int64_t o = (int64_t)f * MC + lane;
      r[o] -= a * cj[o];
      y[o] += a * uj[o];
    }
  }
  double rp = 0.0, ym = 0.0;
  for (int f = slot; f < v.ncut; f += cpp) {
    const int64_t o = (int64_t)f * MC + lane;
    rp += r[o] * r[o];
    ym = fmax(ym, fabs(y[o]));
  }
  const double rn2 = block_lane_sum<MC>(rp, red, bc);
  const double ymax = block_lane_max<MC>(ym, red, bc);
  gcr_flags<MC>(P, G, task, l, chunk, v, rn2, ymax, red, bc);
}

// w = r on unconverged lanes, 0 elsewhere (inflow of the transfer sweep).
template <int MC>
__global__ __launch_bounds__(kMaxBlock) void k_gcr_begin(obte_bte_plan_t P, obte_gcr_t G) {
  const int task = blockIdx.x;
  if (G.active[task] == 0) return;
  const int l = task / P.n_chunks;
  const int chunk = task - l * P.n_chunks;
  const int lane = threadIdx.x % MC;
  const int slot = threadIdx.x / MC;
  const int cpp = blockDim.x / MC;
  const GcrView v = gcr_view(P, l, chunk, MC);
  const int act = G.lane_active[(int64_t)task * MC + lane];
  const double* r = G.r + v.base;
  double* w = G.w + v.base;
  for (int f = slot; f < v.ncut; f += cpp) {
    const int64_t o = (int64_t)f * MC + lane;
    w[o] = act ? r[o] : 0.0;
  }
}

// c = w - x (= (I - T_op) w); orthogonalise against the basis, normalise,
// append, then y += a u, r -= a c with a = c . r.
template <int MC>
__global__ __launch_bounds__(kMaxBlock) void k_gcr_end(obte_bte_plan_t P, obte_gcr_t G) {
  __shared__ double red[kMaxBlock];
  __shared__ double bc[64];
  const int task = blockIdx.x;
  if (G.active[task] == 0) return;
  const int l = task / P.n_chunks;
  const int chunk = task - l * P.n_chunks;
  const int lane = threadIdx.x % MC;
  const int slot = threadIdx.x / MC;
  const int cpp = blockDim.x / MC;
  const GcrView v = gcr_view(P, l, chunk, MC);
  int kb = G.k[task];
  if (kb >= G.kmax) kb = 0;  // restart: drop the recycled basis
  double* un = G.u + kb * v.stride + v.base;
  double* cn = G.c + kb * v.stride + v.base;
  const double* w = G.w + v.base;
  const double* x = G.x + v.base;
  for (int f = slot; f < v.ncut; f += cpp) {
    const int64_t o = (int64_t)f * MC + lane;
    un[o] = w[o];
    cn[o] = w[o] - x[o];
  }
  for (int pass = 0; pass < 2; ++pass) {  // two MGS passes for orthogonality
    for (int j = 0; j < kb; ++j) {
      const double* cj = G.c + j * v.stride + v.base;
      const double* uj = G.u + j * v.stride + v.base;
      double part = 0.0;
      for (int f = slot; f < v.ncut; f += cpp) {
        const int64_t o = (int64_t)f * MC + lane;
        part += cj[o] * cn[o];
      }
      const double b = block_lane_sum<MC>(part, red, bc);
      for (int f = slot; f < v.ncut; f += cpp) {
        const int64_t o = (int64_t)f * MC + lane;
        cn[o] -= b * cj[o];
        un[o] -= b * uj[o];
      }
    }
  }
  double part = 0.0;
  for (int f = slot; f < v.ncut; f += cpp) {
    const int64_t o = (int64_t)f * MC + lane;
    part += cn[o] * cn[o];
  }
  const double nu = sqrt(block_lane_sum<MC>(part, red, bc));
  const double inv = nu > 0.0 ? 1.0 / nu : 0.0;
  double* y = G.y + v.base;
  double* r = G.r + v.base;
  part = 0.0;
  for (int f = slot; f < v.ncut; f += cpp) {
    const int64_t o = (int64_t)f * MC + lane;
    cn[o] *= inv;
    un[o] *= inv;
    part += cn[o] * r[o];
  }
  const double a = block_lane_sum<MC>(part, red, bc);
  double rp = 0.0, ym = 0.0;
  for (int f = slot; f < v.ncut; f += cpp) {
    const int64_t o = (int64_t)f * MC + lane;
    y[o] += a * un[o];
    r[o] -= a * cn[o];
    rp += r[o] * r[o];
    ym = fmax(ym, fabs(y[o]));
  }
  const double rn2 = block_lane_sum<MC>(rp, red, bc);
  const double ymax = block_lane_max<MC>(ym, red, bc);
  if (threadIdx.x == 0) G.k[task] = kb + 1;
  gcr_flags<MC>(P, G, task, l, chunk, v, rn2, ymax, red, bc);
}

__global__ void k_zero_int(int32_t* p) { *p = 0; }

__global__ void k_gcr_reset(obte_bte_plan_t P, obte_gcr_t G, int64_t ny) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gid < ny) G.y[gid] = 0.0;
  if (gid < (int64_t)P.n_dirs * P.n_chunks) G.k[gid] = 0;
}

// Fixed-order reduction over (direction, chunk) partials; one thread per cell.
__global__ void k_reduce_cells(obte_bte_plan_t P, double* dt, double* jn, double* blk) {
  __shared__ double red[256];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int ntask = P.n_dirs * P.n_chunks;
  const int dim = P.dim;
  double s = 0.0;
  if (c < P.n) {
    double j0 = 0.0, j1 = 0.0, j2 = 0.0;
    for (int t = 0; t < ntask; ++t) {
      const int64_t o = (int64_t)t * P.n + c;
      s += P.part_t[o];
      j0 += P.part_j[o * dim + 0];
      j1 += P.part_j[o * dim + 1];
      if (dim == 3) j2 += P.part_j[o * dim + 2];
    }
    dt[c] = s;
    jn[(int64_t)c * dim + 0] = j0;
    jn[(int64_t)c * dim + 1] = j1;
    if (dim == 3) jn[(int64_t)c * dim + 2] = j2;
  }
  red[threadIdx.x] = s * s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[blockIdx.x] = red[0];
}

__global__ void k_reduce_walls(obte_bte_plan_t P, double* tb) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.n_wall) return;
  const int ntask = P.n_dirs * P.n_chunks;
  double v = 0.0;
  for (int t = 0; t < ntask; ++t) v += P.part_tb[(int64_t)t * P.n_wall + s];
  tb[s] = v;
}

__global__ void k_reduce_final(const double* blk, int nblk, double* out) {
  __shared__ double red[1024];
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += blk[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

int check_plan(const obte_bte_plan_t* P, const char* who) {
  if (!P) return obte_fail(OBTE_EINVAL, who);
  if (P->n <= 0 || P->n_dirs <= 0 || P->n_mfp <= 0 || (P->dim != 2 && P->dim != 3))
    return obte_fail(OBTE_EINVAL, "plan: invalid sizes");
  if (P->mc != 8 && P->mc != 16 && P->mc != 32 && P->mc != 64)
    return obte_fail(OBTE_EINVAL, "plan: mc must be 8, 16, 32 or 64");
  if (P->n_chunks != (P->n_mfp + P->mc - 1) / P->mc) return obte_fail(OBTE_EINVAL, "plan: n_chunks != ceil(M/mc)");
  if (P->block < 64 || P->block > kMaxBlock || P->block % 64 != 0)
    return obte_fail(OBTE_EINVAL, "plan: block must be a multiple of 64 in [64, 1024]");
  return OBTE_OK;
}

int launch_status(const char* who) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return obte_fail(OBTE_EHIP, who);
  return OBTE_OK;
}

template <int MC>
void launch_sweep(const obte_bte_plan_t& P, int mode, const double* y, double* x, const int32_t* active,
                  hipStream_t st) {
  dim3 grid(P.n_dirs * P.n_chunks), blk(P.block);
  if (mode == 0) hipLaunchKernelGGL((k_sweep<MC, 0>), grid, blk, 0, st, P, y, x, active);
  else if (mode == 1) hipLaunchKernelGGL((k_sweep<MC, 1>), grid, blk, 0, st, P, y, x, active);
  else hipLaunchKernelGGL((k_sweep<MC, 2>), grid, blk, 0, st, P, y, x, active);
}

template <template <int> class F, typename... A>
void dispatch_mc(int mc, A&&... a) {
  switch (mc) {
    case 8: F<8>::run(a...); break;
    case 16: F<16>::run(a...); break;
    case 32: F<32>::run(a...); break;
    default: F<64>::run(a...); break;
  }
}

template <int MC>
struct SweepL { static void run(const obte_bte_plan_t& P, int mode, const double* y, double* x, const int32_t* a, hipStream_t st) { launch_sweep<MC>(P, mode, y, x, a, st); } };
template <int MC>
struct ProjL { static void run(const obte_bte_plan_t& P, const obte_gcr_t& G, hipStream_t st) { hipLaunchKernelGGL((k_gcr_project<MC>), dim3(P.n_dirs * P.n_chunks), dim3(P.block), 0, st, P, G); } };
template <int MC>
struct BeginL { static void run(const obte_bte_plan_t& P, const obte_gcr_t& G, hipStream_t st) { hipLaunchKernelGGL((k_gcr_begin<MC>), dim3(P.n_dirs * P.n_chunks), dim3(P.block), 0, st, P, G); } };
template <int MC>
struct EndL { static void run(const obte_bte_plan_t& P, const obte_gcr_t& G, hipStream_t st) { hipLaunchKernelGGL((k_gcr_end<MC>), dim3(P.n_dirs * P.n_chunks), dim3(P.block), 0, st, P, G); } };

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

int check_gcr(const obte_gcr_t* G) {
  if (!G || !G->y || !G->r || !G->x || !G->u || !G->c || !G->k || !G->active || !G->lane_active || !G->n_active ||
      !G->rnorm || !G->w || G->kmax <= 0)
    return obte_fail(OBTE_EINVAL, "gcr: null pointer or kmax <= 0");
  return OBTE_OK;
}

}  // namespace

extern "C" {

int obte_bte_assemble(const obte_bte_plan_t* plan, obte_stream_t stream) {
  int rc = check_plan(plan, "obte_bte_assemble: null plan");
  if (rc) return rc;
  const int64_t total = (int64_t)plan->n_dirs * plan->n;
  hipLaunchKernelGGL(k_assemble, dim3(blocks_for(total, 256)), dim3(256), 0, (hipStream_t)stream, *plan);
  return launch_status("obte_bte_assemble: launch failed");
}

int obte_bte_source(const obte_bte_plan_t* plan, obte_stream_t stream) {
  int rc = check_plan(plan, "obte_bte_source: null plan");
  if (rc) return rc;
  const int64_t total = (int64_t)plan->n_dirs * plan->n;
  hipLaunchKernelGGL(k_source, dim3(blocks_for(total, 256)), dim3(256), 0, (hipStream_t)stream, *plan);
  return launch_status("obte_bte_source: launch failed");
}

int obte_bte_sweep(const obte_bte_plan_t* plan, int32_t mode, const double* y, double* x, const int32_t* active,
                   obte_stream_t stream) {
  int rc = check_plan(plan, "obte_bte_sweep: null plan");
  if (rc) return rc;
  if (mode < 0 || mode > 2) return obte_fail(OBTE_EINVAL, "obte_bte_sweep: mode must be 0, 1 or 2");
  if (plan->n_cut > 0 && !y) return obte_fail(OBTE_EINVAL, "obte_bte_sweep: y is required when cuts exist");
  dispatch_mc<SweepL>(plan->mc, *plan, (int)mode, y, x, active, (hipStream_t)stream);
  return launch_status("obte_bte_sweep: launch failed");
}

int obte_gcr_project(const obte_bte_plan_t* plan, const obte_gcr_t* g, obte_stream_t stream) {
  int rc = check_plan(plan, "obte_gcr_project: null plan");
  if (rc || (rc = check_gcr(g))) return rc;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, (hipStream_t)stream, g->n_active);
  dispatch_mc<ProjL>(plan->mc, *plan, *g, (hipStream_t)stream);
  return launch_status("obte_gcr_project: launch failed");
}

int obte_gcr_begin_extend(const obte_bte_plan_t* plan, const obte_gcr_t* g, obte_stream_t stream) {
  int rc = check_plan(plan, "obte_gcr_begin_extend: null plan");
  if (rc || (rc = check_gcr(g))) return rc;
  dispatch_mc<BeginL>(plan->mc, *plan, *g, (hipStream_t)stream);
  return launch_status("obte_gcr_begin_extend: launch failed");
}

int obte_gcr_end_extend(const obte_bte_plan_t* plan, const obte_gcr_t* g, obte_stream_t stream) {
  int rc = check_plan(plan, "obte_gcr_end_extend: null plan");
  if (rc || (rc = check_gcr(g))) return rc;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, (hipStream_t)stream, g->n_active);
  dispatch_mc<EndL>(plan->mc, *plan, *g, (hipStream_t)stream);
  return launch_status("obte_gcr_end_extend: launch failed");
}

int obte_gcr_reset(const obte_bte_plan_t* plan, const obte_gcr_t* g, obte_stream_t stream) {
  int rc = check_plan(plan, "obte_gcr_reset: null plan");
  if (rc || (rc = check_gcr(g))) return rc;
  const int64_t ny = (int64_t)plan->n_chunks * plan->n_cut * plan->mc;
  const int64_t nt = ny > (int64_t)plan->n_dirs * plan->n_chunks ? ny : (int64_t)plan->n_dirs * plan->n_chunks;
  hipLaunchKernelGGL(k_gcr_reset, dim3(blocks_for(nt, 256)), dim3(256), 0, (hipStream_t)stream, *plan, *g, ny);
  return launch_status("obte_gcr_reset: launch failed");
}

int obte_bte_reduce(const obte_bte_plan_t* plan, double* delta_t_new, double* j_new, double* tb_new, double* norm2,
                    obte_stream_t stream) {
  int rc = check_plan(plan, "obte_bte_reduce: null plan");
  if (rc) return rc;
  if (!delta_t_new || !j_new || !norm2 || !plan->scratch) return obte_fail(OBTE_EINVAL, "obte_bte_reduce: null output");
  const unsigned nb = blocks_for(plan->n, 256);
  if (nb > 4096) return obte_fail(OBTE_ELIMIT, "obte_bte_reduce: more than 4096*256 cells");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_reduce_cells, dim3(nb), dim3(256), 0, st, *plan, delta_t_new, j_new, plan->scratch);
  if (plan->n_wall > 0) {
    if (!tb_new) return obte_fail(OBTE_EINVAL, "obte_bte_reduce: null tb output");
    hipLaunchKernelGGL(k_reduce_walls, dim3(blocks_for(plan->n_wall, 256)), dim3(256), 0, st, *plan, tb_new);
  }
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, st, plan->scratch, (int)nb, norm2);
  return launch_status("obte_bte_reduce: launch failed");
}

}  // extern "C"
