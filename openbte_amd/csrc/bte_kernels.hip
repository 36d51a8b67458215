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

// Packed per-position record of the sweep operator, kRec doubles (64 B):
//   [0] D_l  [1] b0 = DeltaT + heat  [2] q  [3..5] inflow coefficients (unused = 0)
//   [6] (src0, src1)  [7] (src2, cell) as int32 pairs
constexpr int kRec = 8;
// faces (or values) per thread and step of the latency-bound per-face loops:
// their loads go out together
constexpr int kFaces = 4;

// doubles per direction of the per-direction sums part_tj and per (direction,
// chunk) of part_chunk: n (nf + 1) rounded up to even (16-byte rows), nf =
// flux sums per position (n_flux)
__host__ __device__ __forceinline__ int64_t part_stride(int n, int nf) {
  return ((int64_t)n * (nf + 1) + 1) & ~(int64_t)1;
}

// flux mode sums per position: one when the group velocities are separable
// (sigma_ml = sig_a_ml sig_dir_l), else one per component
template <class PT>
__host__ __device__ __forceinline__ int n_flux(const PT& P) {
  return P.sig_a != nullptr ? 1 : P.dim;
}

__device__ __forceinline__ double pack_ints(int lo, int hi) { return __hiloint2double(hi, lo); }

// s_waitcnt immediate (gfx9 encoding) that waits for vmcnt <= n only
constexpr int vmcnt_only(int n) { return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14); }
constexpr int kSweepBlock = 256;  // sweep kernels: at most 256 threads (register budget)

__device__ __forceinline__ double dotv(const double* a, const double* b, int dim) {
  double g = a[0] * b[0] + a[1] * b[1];
  if (dim == 3) g += a[2] * b[2];
  return g;
}

// ---------------------------------------------------------------------------
// Problem sources.  Every kernel is written once against an accessor that
// yields the plan / Krylov state / IO pointers of problem `pid`:
//   One  -- a single problem whose structs are kernel arguments (the
//           obte_bte_* / obte_gcr_* entry points);
//   Many -- a batch of problems whose structs live in device arrays (the
//           obte_batch_* entry points); `live` freezes finished problems.
// ---------------------------------------------------------------------------
struct One {
  obte_bte_plan_t P;
  obte_gcr_t G;
  obte_sweep_io_t io;
  obte_reduce_io_t rio;
  __device__ const obte_bte_plan_t& plan(int) const { return P; }
  __device__ const obte_gcr_t& gcr(int) const { return G; }
  __device__ const obte_sweep_io_t& sio(int) const { return io; }
  __device__ const obte_reduce_io_t& red(int) const { return rio; }
  __device__ bool live(int) const { return true; }
  __device__ int pid_of(int) const { return 0; }
  __device__ int local_of(int t) const { return t; }
};

// The batch's struct arrays and task table are read through the constant
// address space: nothing writes them while a kernel runs, so the loads of a
// uniform problem index are scalar loads and the fields stay in scalar
// registers (through generic pointers every field was a vector load).
#define OBTE_CONST __attribute__((address_space(4)))

struct Many {
  const OBTE_CONST obte_bte_plan_t* P;
  const OBTE_CONST obte_gcr_t* G;
  const OBTE_CONST obte_sweep_io_t* io;
  const OBTE_CONST obte_reduce_io_t* rio;
  const OBTE_CONST int32_t* livef;
  const OBTE_CONST int32_t* tpid;
  const OBTE_CONST int32_t* tloc;
  __device__ const OBTE_CONST obte_bte_plan_t& plan(int p) const { return P[p]; }
  __device__ const OBTE_CONST obte_gcr_t& gcr(int p) const { return G[p]; }
  __device__ const OBTE_CONST obte_sweep_io_t& sio(int p) const { return io[p]; }
  __device__ const OBTE_CONST obte_reduce_io_t& red(int p) const { return rio[p]; }
  __device__ bool live(int p) const { return livef[p] != 0; }
  __device__ int pid_of(int t) const { return tpid[t]; }
  __device__ int local_of(int t) const { return tloc[t]; }
};

// ---------------------------------------------------------------------------
// (a) face-flux assembly: D_l, inflow coefficients, perturbation
//     (openbte/bte.py:47-124), one thread per (direction, position);
//     blockIdx.y = problem.
// ---------------------------------------------------------------------------
template <class S>
__global__ void k_assemble(S src) {
  const int pid = blockIdx.y;
  const auto& P = src.plan(pid);
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
  // (interface reflecting sides, wall_side <= -2, add neither: their outflow is
  // the interface face's, their inflow is a source term)
  for (int w = P.wall_ptr[c]; w < P.wall_ptr[c + 1]; ++w) {
    const int sd = P.wall_side[w];
    if (sd <= -2) continue;
    const double g = dotv(s, P.wall_db + (int64_t)w * dim, dim);
    if (g > 0.0) D += g;
    else if (sd == -1) pert -= g * P.wall_temp[w];
  }
  if (P.nano != 0.0 && dim == 3) pert -= s[2];
  P.pert[gid] = pert;
  const int64_t base = (int64_t)l * P.kin * P.n + p;
  double cf[3] = {0.0, 0.0, 0.0};
  int sr[3] = {0, 0, 0};
  for (int kk = 0; kk < P.kin; ++kk) {
    const int f = P.nb_face[base + (int64_t)kk * P.n];
    sr[kk] = P.nb_src[base + (int64_t)kk * P.n];
    if (f >= 0) {
      const double g = dotv(s, P.k_face + (int64_t)f * dim, dim);
      cf[kk] = g < 0.0 ? -g : 0.0;
      if (P.t_face != nullptr) cf[kk] *= P.t_face[f];
    }
  }
  double2* r = reinterpret_cast<double2*>(P.rec + gid * kRec);
  r[0] = make_double2(D, 0.0);
  r[1] = make_double2(0.0, cf[0]);
  r[2] = make_double2(cf[1], cf[2]);
  r[3] = make_double2(pack_ints(sr[0], sr[1]), pack_ints(sr[2], 0));  // keep flags: k_mark_keep
}

// Smallest LDS window (positions) any sweep of a plan with these level sizes
// uses; 0 when the sizes are unknown (no windowed sweep then).
__host__ __device__ inline int min_window(int pair_max) {
  if (pair_max <= 0) return 0;
  int cap = 64;
  while (cap < pair_max) cap <<= 1;
  return cap;
}

// keep flag (high word of rec[7]) of every position whose value a later
// position reads from t rather than from the smallest LDS window: the
// upwind source s of position p is outside the window when s < h - cap, h
// the end of p's level.  One thread per (direction, position), blockIdx.y =
// problem; runs after k_assemble.
template <class S>
__global__ void k_mark_keep(S src) {
  const int pid = blockIdx.y;
  const auto& P = src.plan(pid);
  const int cap = min_window(P.lvl_pair_max);
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (cap == 0 || gid >= (int64_t)P.n_dirs * P.n) return;
  const int n = P.n;
  const int l = (int)(gid / n);
  const int p = (int)(gid - (int64_t)l * n);
  const int32_t* lp = P.lvl_ptr + P.lvl_off[l];
  int lo = 1, hi = P.lvl_off[l + 1] - P.lvl_off[l] - 1;  // first level start > p
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (lp[mid] > p) hi = mid;
    else lo = mid + 1;
  }
  const int h = lp[lo];
  int32_t* words = reinterpret_cast<int32_t*>(P.rec);
  for (int kk = 0; kk < P.kin; ++kk) {
    const int64_t e = ((int64_t)l * P.kin + kk) * n + p;
    const int s = P.nb_src[e];
    if (P.nb_face[e] >= 0 && s < n && s < h - cap) atomicOr(&words[(((int64_t)l * n + s) * kRec + 7) * 2 + 1], 1);
  }
}

// keep flag of the exit position of every cut face (its value leaves the
// sweep as x); grid-stride over the cuts, blockIdx.y = problem
template <class S>
__global__ void k_mark_exits(S src) {
  const int pid = blockIdx.y;
  const auto& P = src.plan(pid);
  if (min_window(P.lvl_pair_max) == 0) return;
  int32_t* words = reinterpret_cast<int32_t*>(P.rec);
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < P.n_cut; f += gridDim.x * blockDim.x) {
    int lo = 0, hi = P.n_dirs;  // direction l with cut_off[l] <= f < cut_off[l + 1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (P.cut_off[mid] <= f) lo = mid;
      else hi = mid;
    }
    atomicOr(&words[(((int64_t)lo * P.n + P.cut_exit[f]) * kRec + 7) * 2 + 1], 1);
  }
}

// keep flag of the diffuse-wall and kappa-contact cells in every direction
// (the final sweep's wall-temperature and kappa sums read their values from
// t); grid-stride over (direction, entry), blockIdx.y = problem
template <class S>
__global__ void k_mark_cells(S src) {
  const int pid = blockIdx.y;
  const auto& P = src.plan(pid);
  if (min_window(P.lvl_pair_max) == 0) return;
  const int nc = P.contact_kind == 0 ? 0 : P.n_contact;
  const int per = P.n_wall + (P.contact_kind == 1 ? 2 * nc : nc);
  const int64_t total = (int64_t)per * P.n_dirs;
  int32_t* words = reinterpret_cast<int32_t*>(P.rec);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(i / per);
    const int k = (int)(i - (int64_t)l * per);
    const int c = k < P.n_wall ? P.wall_cell[k] : (k - P.n_wall < nc ? P.ce1[k - P.n_wall] : P.ce2[k - P.n_wall - nc]);
    const int p = P.pos[(int64_t)l * P.n + c];
    atomicOr(&words[(((int64_t)l * P.n + p) * kRec + 7) * 2 + 1], 1);
  }
}

// q = pert + sum over diffuse walls of (-TB_s) min(s . db, 0)   (openbte/bte.py:199-200,214)
//     + sum over interface reflecting sides of (1 - t) (-TB_k) min(s . db, 0)
// b0 = DeltaT + heat, both in position order
template <class S>
__global__ void k_source(S src) {
  const int pid = blockIdx.y;
  if (!src.live(pid)) return;
  const auto& P = src.plan(pid);
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
    if (sd == -1) continue;
    const double g = dotv(s, P.wall_db + (int64_t)w * dim, dim);
    if (g < 0.0) {
      if (sd >= 0) q += -P.tb[sd] * g;
      else q += -P.wall_temp[w] * P.tb[-2 - sd] * g;
    }
  }
  P.rec[gid * kRec + 1] = P.delta_t[c] + P.heat[c];
  P.rec[gid * kRec + 2] = q;
}

// ---------------------------------------------------------------------------
// (b) level-scheduled upwind sweep, (c) fused mode sums
//
// One workgroup per task = (problem, direction, chunk of MC MFP bins).  A
// cell's MC values are one contiguous row of T; G = MC/2 consecutive lanes
// share a cell and each owns a pair of bins (one 16-byte access), so a wave
// reads or writes 64/G whole rows with one cache line per row.  Each thread
// holds U cells of the current level (cells slot, slot + CPS, ...), issues all
// their loads first, then the divisions and stores.
//
// Software pipelining across levels: at the start of level k the per-cell
// data of level k+1 (sources, diagonal, inflow slots and coefficients) is
// loaded, so after the barrier only the upwind gathers are on the critical
// path.  Level pointers live in LDS (a scalar load in the loop would make
// every LDS access wait on it).  Levels wider than U*CPS cells take extra
// passes.
// ---------------------------------------------------------------------------
// Lanes of one cell that share a quad (2 for MC = 4, else 4) split the 64-byte
// record: lane jq of the quad loads NP = 4 / GQ 16-byte pieces, starting at
// piece jq * NP, and the quad exchanges them with DPP (no LDS, no memory).
template <int U, int NP>
struct LevelRegs {
  int p[U];            // position, -1 for an idle slot
  double2 pc[U][NP];   // this lane's pieces of the record
};

template <int U, int NP>
__device__ __forceinline__ void load_level(LevelRegs<U, NP>& R, int base, int p1, int cps, int jq,
                                           const double* __restrict__ rec) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int p = base + u * cps;
    const bool ok = p < p1;
    R.p[u] = ok ? p : -1;
    const int pc = ok ? p : 0;  // idle slots read position 0 (always present) and are discarded
    const double2* r = reinterpret_cast<const double2*>(rec + (int64_t)pc * kRec) + jq * NP;
#pragma unroll
    for (int i = 0; i < NP; ++i) R.pc[u][i] = r[i];
  }
}

// num / den for den = 1 + Lambda D >= 1 (no scaling or special cases):
// reciprocal from v_rcp_f64 (relative error < 5e-8 on gfx950) and one Newton
// step (< 3e-15), then one residual correction of the quotient, which rounds
// it as IEEE division does (measured: no difference from x / d over 4M
// random quotients, the same as with a second Newton step).  Every sweep uses
// it, so their values agree bit for bit.
__device__ __forceinline__ double upwind_value(double num, double den) {
  double r = __builtin_amdgcn_rcp(den);
  r = fma(fma(-den, r, 1.0), r, r);
  const double q = num * r;
  return fma(fma(-den, q, num), r, q);
}

// v of lane S of this lane's group of GQ lanes inside its quad
template <int GQ, int S>
__device__ __forceinline__ double quad_bcast(double v) {
  constexpr int ctrl = GQ == 4 ? (S | S << 2 | S << 4 | S << 6) : (S | S << 2 | (2 + S) << 4 | (2 + S) << 6);
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// double D of the record held in pieces pc (all lanes of the wave active)
template <int GQ, int NP, int D>
__device__ __forceinline__ double rec_get(const double2 (&pc)[NP]) {
  constexpr int piece = D / 2;
  const double2 w = pc[piece % NP];
  return quad_bcast<GQ, piece / NP>(D % 2 ? w.y : w.x);
}

constexpr int kLvlStage = 2048;  // level pointers staged in LDS per refill

template <int MC>
struct alignas(16) SweepShared {
  static constexpr int kStage = kLvlStage;
  double sig[3 * MC];      // sigma of the chunk's bins, [bin][3] (final sweeps)
  double wk[4][MC];        // mode-sum weights t, sigma_x, sigma_y, sigma_z per bin (final sweeps)
  int lvl[kLvlStage + 2];  // staged level starts of the running direction
};

// Shared state of a task run by one wave (a batched sweep runs up to four of
// them per workgroup side by side: problems whose levels are narrow): the
// same fields, fewer level starts staged per refill.
template <int MC>
struct alignas(16) WaveShared {
  static constexpr int kStage = 510;
  double sig[3 * MC];
  double wk[4][MC];
  int lvl[kStage + 2];
};

template <int MC>
union BatchShared {
  SweepShared<MC> blk;
  WaveShared<MC> wav[4];
};

// LDS ordering point of a task run by one wave: the wave's LDS operations
// execute in order, and no other wave takes part (no s_barrier)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-chunk mode sums of one cell from its row of MC values (G = MC / 2
// double2 pairs): e[0] = sum over the chunk's bins of t_m T_m, e[1 + d] =
// sum of sigma_m[d] T_m, one FMA chain per component in bin order.  wk(k, j)
// gives the weights of component k for the bin pair j.  The windowed final
// sweep (row in its LDS window) and k_modesum (row in t) both use it, so
// their sums have the same bits.
// kSerial: keep each bin pair's loads next to their use (the sweep's register
// budget), instead of letting the scheduler hoist them all.
template <int G, bool kSerial, class Row, class W>
__device__ __forceinline__ void chunk_sums(const Row& row, const W& wk, int dim, double (&e)[4]) {
#pragma unroll
  for (int j = 0; j < G; ++j) {
    if constexpr (kSerial) __builtin_amdgcn_sched_barrier(0);
    const double2 v = row(j);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k <= dim) {
        const double2 w = wk(k, j);
        e[k] = j == 0 ? w.x * v.x : fma(w.x, v.x, e[k]);
        e[k] = fma(w.y, v.y, e[k]);
      }
    }
  }
}

// cells per thread and level pass; the windowed sweep also holds the far
// upwind values of two levels, so it takes fewer (one 256-thread pass covers
// 192 cells at MC = 8 and 256 at MC = 4)
template <int MC, bool WIN>
constexpr int units_of() { return WIN ? (MC == 8 ? 3 : 2) : (MC >= 16 ? 2 : 4); }
// the persistent batch kernel carries its queue state as well: two cells per
// thread in the windowed sweep keep it within three waves per SIMD
template <int MC, bool WIN>
constexpr int units_batch() { return WIN ? 2 : units_of<MC, WIN>(); }

// ---------------------------------------------------------------------------
// Windowed sweep.  The values of the last `cap` positions (cap a power of two
// holding any two adjacent levels) are also kept in LDS, slot
// (position & (cap - 1)) * (MC / 2) + lane.  An upwind value inside that
// window is read from LDS; one outside it belongs to a level at least two back
// and is loaded from T one level ahead (its store went out before the barrier
// that precedes the load; the waves of a workgroup share the CU's vector L1,
// so a store issued before s_barrier is seen by loads issued after it).  Per
// level the critical path is then LDS reads, the arithmetic and the barrier:
// the global loads of a level are issued a level (records: two) ahead and the
// only wait is for operations of the previous level.
//
// Every global load and store of the level loop is unconditional, so the
// number of vector-memory operations a wave has in flight is static: slots
// past the end of a level recompute the level's last cell (same inputs, same
// value, same address), and near or converged lanes load row 0 instead of a
// far row.  The wait at the top of a level, vmcnt <= U, leaves exactly the U
// stores of the previous level in flight (vector-memory operations complete
// in issue order).
// ---------------------------------------------------------------------------
extern __shared__ double2 win_lds[];

// Raw buffer access (gfx9 descriptor word 3): 32-bit offsets, and lanes past
// the buffer's size load zeros and drop their stores without a memory access.
constexpr int kBufWord3 = 0x00020000;
constexpr int kBufOob = 0x7FFFFFF0;  // an offset past every buffer used here
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kBufWord3);
}
__device__ __forceinline__ double2 buf_load2(__amdgpu_buffer_rsrc_t rs, int off) {
  const f64x2 v = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
  return make_double2(v.x, v.y);
}
// a pointer every lane holds the same value of, moved to scalar registers (a
// buffer descriptor built from it then needs no per-lane handling)
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double buf_load1(__amdgpu_buffer_rsrc_t rs, int off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
}
__device__ __forceinline__ void buf_store1(__amdgpu_buffer_rsrc_t rs, int off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rs, off, 0, 0);
}
__device__ __forceinline__ void buf_store2(__amdgpu_buffer_rsrc_t rs, int off, double2 t) {
  const f64x2 v = {t.x, t.y};
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), rs, off, 0, 0);
}

template <int U, int NP, int KIN>
struct WinRegs {
  int p[U];               // position (slots past the level: its last cell)
  double2 pc[U][NP];      // this lane's pieces of the record
  double2 far[U][KIN];    // upwind value loaded from T (0 when it is in the window)
  int woff[U][KIN];       // window slot of the upwind value (the zero slot when not)
};

template <int U, int NP, int KIN>
__device__ __forceinline__ void load_win(WinRegs<U, NP, KIN>& R, int base, int p1, int cps, int jq,
                                         __amdgpu_buffer_rsrc_t rec_rs) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    int p = base + u * cps;
    p = p < p1 ? p : p1 - 1;
    R.p[u] = p;
#pragma unroll
    for (int i = 0; i < NP; ++i) R.pc[u][i] = buf_load2(rec_rs, p * (kRec * 8) + (jq * NP + i) * 16);
  }
}

// One (problem pid, task lt = l * n_chunks + chunk) sweep by the whole block,
// or (WAVE) by one wave of it: then tid and nthr count the wave's lanes, the
// wave synchronises alone and owns window slot range wave * (cap * G + 1).
// WIN: with an LDS window of cap positions (dynamic shared memory).
// MODE: 0 source sweep, 1 transfer sweep, 2 final sweep, 3 final sweep that
// also forms the per-chunk mode sums from its LDS window (windowed only).
template <class S, int MC, int MODE, int KIN, bool WIN, int U, bool WAVE = false, class SH = SweepShared<MC>>
__device__ __forceinline__ void sweep_task(const S& src, int pid, int lt, SH& sh, int cap) {
  constexpr bool kFinal = MODE >= 2;
  constexpr bool kFuse = MODE == 3;
  static_assert(WIN || !kFuse, "the fused mode sums read the LDS window");
  constexpr int G = MC / 2;       // lanes per cell
  constexpr int GQ = G < 4 ? G : 4;  // lanes of a cell inside one quad
  constexpr int NP = 4 / GQ;         // record pieces per lane
  const auto& P = src.plan(pid);
  const auto& io = src.sio(pid);
  const int L = P.n_dirs;
  const int l = lt / P.n_chunks;
  const int chunk = lt - l * P.n_chunks;
  const int n = P.n;
  const int dim = P.dim;
  const int tid = WAVE ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
  const int nthr = WAVE ? 64 : (int)blockDim.x;
  auto sync = [] {
    if constexpr (WAVE) wave_sync();
    else __syncthreads();
  };
  double2* const win = win_lds + (WAVE ? (int)(threadIdx.x >> 6) * (cap * G + 1) : 0);
  const int jl = tid % G;          // bin pair of this lane
  const int cslot = tid / G;       // cell slot of this lane
  const int cps = nthr / G;        // cells per pass
  const int jq = jl % GQ;          // record part of this lane

  // per-lane constants of its two bins
  const int m0 = chunk * MC + 2 * jl;
  const bool v0 = m0 < P.n_mfp, v1 = m0 + 1 < P.n_mfp;
  const double lam0 = v0 ? P.mfps[m0] : 0.0, lam1 = v1 ? P.mfps[m0 + 1] : 0.0;
  if (kFinal && cslot == 0) {
    // group velocities and weights of the chunk's bins for the wall, contact
    // and mode sums
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int m = m0 + h;
      const double* sg = P.sigma + ((int64_t)(m < P.n_mfp ? m : 0) * L + l) * dim;
      const bool ok = m < P.n_mfp;
      sh.sig[3 * (2 * jl + h) + 0] = ok ? sg[0] : 0.0;
      sh.sig[3 * (2 * jl + h) + 1] = ok ? sg[1] : 0.0;
      sh.sig[3 * (2 * jl + h) + 2] = ok && dim == 3 ? sg[2] : 0.0;
      sh.wk[0][2 * jl + h] = ok ? P.t_coeff[(int64_t)m * L + l] : 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) sh.wk[1 + d][2 * jl + h] = sh.sig[3 * (2 * jl + h) + d];
      if (P.sig_a != nullptr) sh.wk[1][2 * jl + h] = ok ? P.sig_a[(int64_t)m * L + l] : 0.0;
    }
  }
  bool act = true;  // transfer sweeps skip converged bin pairs
  if (MODE == 1 && io.lane_active != nullptr)
    act = io.lane_active[(int64_t)lt * MC + 2 * jl] != 0 || io.lane_active[(int64_t)lt * MC + 2 * jl + 1] != 0;

  double* __restrict__ T = P.t + (int64_t)lt * ((int64_t)n + P.ncut_max) * MC;
  const int32_t* __restrict__ lvl_g = P.lvl_ptr + P.lvl_off[l];
  const int nlev = P.lvl_off[l + 1] - P.lvl_off[l] - 1;
  const int cbase = P.cut_off[l];
  const int ncut = P.cut_off[l + 1] - cbase;
  const int64_t lo = (int64_t)l * n;
  const double* __restrict__ rec = P.rec + lo * kRec;

  // interface inflows of the cut faces -> rows n .. n + ncut of T
  if (ncut > 0) {
    const double* __restrict__ yv = io.y + ((int64_t)chunk * P.n_cut + cbase) * MC;
    double* __restrict__ ti = T + (int64_t)n * MC;
    for (int i = tid; i < ncut * MC; i += nthr) ti[i] = yv[i];
  }
  int* const lvl = sh.lvl;
  // level starts from `from` on; past the last level every level is empty
  auto stage = [&](int from) {
    for (int i = tid; i < SH::kStage + 2; i += nthr) lvl[i] = lvl_g[from + i <= nlev ? from + i : nlev];
  };
  if constexpr (WIN) {
    // unused inflow slots read the window with a zero coefficient: no
    // uninitialised (possibly NaN) patterns in it
    for (int i = tid; i <= cap * G; i += nthr) win[i] = make_double2(0.0, 0.0);
  }
  stage(0);
  sync();
  // Three register sets rotate (current, next, prefetch target) without
  // copies; the level count is padded to a multiple of 3 with empty levels so
  // the unrolled loop has no exits.  Level starts are staged kSeg at a time.
  constexpr int kSeg = (SH::kStage - 2) / 3 * 3;
  const int nlev3 = (nlev + 2) / 3 * 3;

  if constexpr (WIN) {
    using WR = WinRegs<U, NP, KIN>;
    const int cmask = cap - 1;
    const int zero_slot = cap * G;  // a zero double2 past the window, never written
    const __amdgpu_buffer_rsrc_t t_rs = make_rsrc(T, ((int64_t)n + P.ncut_max) * MC * 8);
    const __amdgpu_buffer_rsrc_t rec_rs = make_rsrc(rec, (int64_t)n * kRec * 8);
    // final sweeps store every value unless they form the mode sums here
    // (then only values read later, unless the caller asked for all of them)
    const bool store_all = MODE == 2 || (kFuse && P.keep_t != 0);
    const int nf = n_flux(P);
    const __amdgpu_buffer_rsrc_t ps_rs = make_rsrc(kFuse ? P.part_chunk + (int64_t)lt * part_stride(n, nf) : T,
                                                   kFuse ? (int64_t)n * (nf + 1) * 8 : 0);
    // the upwind values of N (a level ending at h): window slot, or a load
    // from T when outside the window (lanes with nothing to load read past
    // the buffer: no memory access, a zero value)
    auto issue_far = [&](WR& N, int h) {
      const int lo_w = h - cap;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const double s01 = rec_get<GQ, NP, 6>(N.pc[u]);
        const double s2c = rec_get<GQ, NP, 7>(N.pc[u]);
        const int sr[3] = {__double2loint(s01), __double2hiint(s01), __double2loint(s2c)};
#pragma unroll
        for (int kk = 0; kk < KIN; ++kk) {
          const bool inwin = (unsigned)(sr[kk] - lo_w) < (unsigned)cap;
          const bool dead = MODE == 1 && !act;
          N.woff[u][kk] = (inwin && !dead) ? (sr[kk] & cmask) * G + jl : zero_slot;
          // only lanes with a value outside the window load (a wave with none
          // skips the instruction)
          N.far[u][kk] = make_double2(0.0, 0.0);
          if (!(inwin || dead)) N.far[u][kk] = buf_load2(t_rs, sr[kk] * (MC * 8) + jl * 16);
        }
      }
    };
    // the cells of C (a level ending at h): upwind values from the window plus
    // C.far (one of the two is zero), then the store to T and to the window
    auto finish_win = [&](const WR& C) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const double dd = rec_get<GQ, NP, 0>(C.pc[u]);
        const double bb = MODE == 1 ? 0.0 : rec_get<GQ, NP, 1>(C.pc[u]);
        const double qq = MODE == 1 ? 0.0 : rec_get<GQ, NP, 2>(C.pc[u]);
        const double cf[3] = {rec_get<GQ, NP, 3>(C.pc[u]), rec_get<GQ, NP, 4>(C.pc[u]),
                              KIN == 3 ? rec_get<GQ, NP, 5>(C.pc[u]) : 0.0};
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int kk = 0; kk < KIN; ++kk) {
          const double2 w = win[C.woff[u][kk]];
          a0 += cf[kk] * (w.x + C.far[u][kk].x);
          a1 += cf[kk] * (w.y + C.far[u][kk].y);
        }
        const double2 tv = make_double2(upwind_value(bb + lam0 * (qq + a0), 1.0 + lam0 * dd),
                                        upwind_value(bb + lam1 * (qq + a1), 1.0 + lam1 * dd));
        // only values read later from t (exits, wall and contact cells included)
        // go to HBM, unless the final sweep stores them all
        const bool keep = store_all || __double2hiint(rec_get<GQ, NP, 7>(C.pc[u])) != 0;
        buf_store2(t_rs, keep ? C.p[u] * (MC * 8) + jl * 16 : kBufOob, tv);
        win[(C.p[u] & cmask) * G + jl] = tv;
      }
    };
    // mode sums of a finished level (positions q0 .. q1), read back from the
    // window while the next level is computed: the level's slots are not
    // overwritten before the level after next (cap holds two adjacent levels).
    // One thread per cell reads its row and stores its nf + 1 sums (the
    // level's sums are one contiguous range: consecutive positions).
    auto reduce_level = [&](int q0, int q1) {
      for (int p = q0 + tid; p < q1; p += nthr) {
        const double2* row = win + (p & cmask) * G;
        double e[4] = {0.0, 0.0, 0.0, 0.0};
        chunk_sums<G, true>([&](int j) { return row[j]; },
                            [&](int k, int j) { return *reinterpret_cast<const double2*>(&sh.wk[k][2 * j]); }, nf,
                            e);
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (k <= nf)
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, e[k]), ps_rs, (p * (nf + 1) + k) * 8, 0,
                                                  0);
      }
    };
    int rq0 = 0, rq1 = 0;  // the previous level (its mode sums are formed during this one)
    // level kk (cells in C); issues the far loads of level kk + 1 (N) and
    // the records of level kk + 2 (into N2)
    auto body = [&](int kk, WR& C, WR& N, WR& N2) {
      const int p0 = lvl[kk];
      const int p1 = lvl[kk + 1];
      // the U stores of the previous level (and its mode-sum stores) stay in
      // flight
      __builtin_amdgcn_s_waitcnt(vmcnt_only(kFuse ? U + 4 : U));
      __builtin_amdgcn_sched_barrier(0);
      issue_far(N, lvl[kk + 2]);
      __builtin_amdgcn_sched_barrier(0);
      load_win<U, NP, KIN>(N2, lvl[kk + 2] + cslot, lvl[kk + 3], cps, jq, rec_rs);
      __builtin_amdgcn_sched_barrier(0);
      finish_win(C);
      // levels wider than one pass (not produced by the default schedule)
      for (int base = p0 + U * cps; base < p1; base += U * cps) {
        load_win<U, NP, KIN>(C, base + cslot, p1, cps, jq, rec_rs);
        issue_far(C, p1);
        finish_win(C);
      }
      if constexpr (kFuse) {
        reduce_level(rq0, rq1);
        rq0 = p0;
        rq1 = p1;
      }
      sync();
    };
    WR A, B, C;
    load_win<U, NP, KIN>(A, lvl[0] + cslot, lvl[1], cps, jq, rec_rs);
    load_win<U, NP, KIN>(B, lvl[1] + cslot, lvl[2], cps, jq, rec_rs);
    __builtin_amdgcn_s_waitcnt(0);
    issue_far(A, lvl[1]);
    __builtin_amdgcn_s_waitcnt(0);
    for (int seg = 0; seg < nlev3; seg += kSeg) {
      if (seg > 0) {
        stage(seg);
        sync();
      }
      const int kend = min(kSeg, nlev3 - seg);
      for (int kk = 0; kk < kend; kk += 3) {
        body(kk, A, B, C);
        body(kk + 1, B, C, A);
        body(kk + 2, C, A, B);
      }
    }
    if constexpr (kFuse) reduce_level(rq0, rq1);  // the last level
  } else {

  using Regs = LevelRegs<U, NP>;
  // upwind gathers of the U cells of C (idle slots skip them)
  auto gather = [&](const Regs& C, double2 (&g)[U][KIN]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double s01 = rec_get<GQ, NP, 6>(C.pc[u]);
      const double s2c = rec_get<GQ, NP, 7>(C.pc[u]);
      int sr[3] = {__double2loint(s01), __double2hiint(s01), __double2loint(s2c)};
#pragma unroll
      for (int kk = 0; kk < KIN; ++kk) {
        g[u][kk] = make_double2(0.0, 0.0);
        if (C.p[u] >= 0 && (MODE != 1 || act))
          g[u][kk] = *reinterpret_cast<const double2*>(T + (int64_t)sr[kk] * MC + 2 * jl);
      }
    }
  };
  auto finish = [&](const Regs& C, const double2 (&g)[U][KIN]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double dd = rec_get<GQ, NP, 0>(C.pc[u]);
      const double bb = MODE == 1 ? 0.0 : rec_get<GQ, NP, 1>(C.pc[u]);
      const double qq = MODE == 1 ? 0.0 : rec_get<GQ, NP, 2>(C.pc[u]);
      const double cf[3] = {rec_get<GQ, NP, 3>(C.pc[u]), rec_get<GQ, NP, 4>(C.pc[u]),
                            KIN == 3 ? rec_get<GQ, NP, 5>(C.pc[u]) : 0.0};
      double a0 = 0.0, a1 = 0.0;
#pragma unroll
      for (int kk = 0; kk < KIN; ++kk) {
        a0 += cf[kk] * g[u][kk].x;
        a1 += cf[kk] * g[u][kk].y;
      }
      const double t0 = upwind_value(bb + lam0 * (qq + a0), 1.0 + lam0 * dd);
      const double t1 = upwind_value(bb + lam1 * (qq + a1), 1.0 + lam1 * dd);
      if (C.p[u] >= 0 && (MODE != 1 || act))
        *reinterpret_cast<double2*>(T + (int64_t)C.p[u] * MC + 2 * jl) = make_double2(t0, t1);
    }
  };

  // One level (kk levels after the staged base) with the cells of C; the
  // records of the level two further go into N2.  Issue order matters:
  // vector-memory counters retire in issue order, so the gathers go out
  // before the prefetch and the arithmetic waits for the gathers only.  The
  // records of the next level were issued before the gathers of this one, so
  // no wait on a prefetch is ever exposed.
  auto body = [&](int kk, Regs& C, Regs& N2) {
    const int p0 = lvl[kk];
    const int p1 = lvl[kk + 1];
    double2 g[U][KIN];
    gather(C, g);
    __builtin_amdgcn_sched_barrier(0);
    load_level<U, NP>(N2, lvl[kk + 2] + cslot, lvl[kk + 3], cps, jq, rec);
    __builtin_amdgcn_sched_barrier(0);
    // everything but the U * NP prefetch loads has landed (an unconditional
    // wait: the waits the compiler places sit in the per-cell branches)
    __builtin_amdgcn_s_waitcnt(vmcnt_only(U * NP));
    __builtin_amdgcn_sched_barrier(0);
    finish(C, g);
    // levels wider than one pass: C is spent, reload it per extra pass
    for (int base = p0 + U * cps; base < p1; base += U * cps) {
      load_level<U, NP>(C, base + cslot, p1, cps, jq, rec);
      gather(C, g);
      finish(C, g);
    }
    sync();
  };

  Regs A, B, C;
  load_level<U, NP>(A, lvl[0] + cslot, lvl[1], cps, jq, rec);
  load_level<U, NP>(B, lvl[1] + cslot, lvl[2], cps, jq, rec);
  // A and B arrive before the loop, so the loop body never waits on them
  __builtin_amdgcn_s_waitcnt(0);
  for (int seg = 0; seg < nlev3; seg += kSeg) {
    if (seg > 0) {
      stage(seg);
      sync();
    }
    const int kend = min(kSeg, nlev3 - seg);
    for (int kk = 0; kk < kend; kk += 3) {
      body(kk, A, C);
      body(kk + 1, B, A);
      body(kk + 2, C, B);
    }
  }
  }

  // exit values of the cut faces (upstream cells across the cut)
  // (kFaces values per thread and step, every index and value load issued
  // before the stores)
  if (io.x != nullptr) {
    double* __restrict__ xv = io.x + ((int64_t)chunk * P.n_cut + cbase) * MC;
    const int32_t* __restrict__ ex = P.cut_exit + cbase;
    const int total = ncut * MC;
    for (int i0 = tid; i0 < total; i0 += kFaces * nthr) {
      int src[kFaces];
#pragma unroll
      for (int q = 0; q < kFaces; ++q) {
        const int i = min(i0 + q * nthr, total - 1);
        const int f = i / MC;
        src[q] = ex[f] * MC + (i - f * MC);
      }
      double t[kFaces];
#pragma unroll
      for (int q = 0; q < kFaces; ++q) t[q] = T[(int64_t)src[q]];
#pragma unroll
      for (int q = 0; q < kFaces; ++q)
        if (i0 + q * nthr < total) xv[i0 + q * nthr] = t[q];
    }
  }

  if (kFinal) {
    const int32_t* __restrict__ pos = P.pos + lo;
    // wall temperatures: TB_s = -sum_ml T_ml[i_s] GG_ml,s  (openbte/bte.py:73-78,227);
    // a side's row is read KB values at a time, all loads in flight together
    constexpr int KB = 4;
    for (int sd = tid; sd < P.n_wall; sd += nthr) {
      const double* db = P.wall_dbs + (int64_t)sd * dim;
      const double rinv = P.wall_rinv[sd];
      const double2* row = reinterpret_cast<const double2*>(T + (int64_t)pos[P.wall_cell[sd]] * MC);
      double acc = 0.0;
#pragma unroll 1
      for (int m0 = 0; m0 < MC; m0 += KB) {
        double v[KB];
#pragma unroll
        for (int j = 0; j < KB / 2; ++j) {
          const double2 t = row[m0 / 2 + j];
          v[2 * j] = t.x;
          v[2 * j + 1] = t.y;
        }
#pragma unroll
        for (int j = 0; j < KB; ++j) {
          const int m = m0 + j;
          double gb = sh.sig[3 * m] * db[0] + sh.sig[3 * m + 1] * db[1];
          if (dim == 3) gb += sh.sig[3 * m + 2] * db[2];
          acc += -v[j] * ((gb > 0.0 ? gb : 0.0) * rinv);
        }
      }
      P.part_tb[(int64_t)lt * P.n_wall + sd] = acc;
    }
    // kappa and suppression of each mode (openbte/bte.py:232-243), summed
    // by the task's first wave in a fixed order (lane = contact side mod 64,
    // then an xor tree), so a task run by one wave or by a workgroup gives
    // the same bits.  KB bins at a time: a contact side's indices and normal
    // are read once per KB bins and its rows KB values at a time.
    if (P.contact_kind != 0 && tid < 64) {
#pragma unroll 1
      for (int m0 = 0; m0 < MC; m0 += KB) {
        double kap[KB], sup[KB];
#pragma unroll
        for (int j = 0; j < KB; ++j) kap[j] = sup[j] = 0.0;
        for (int sc = tid; sc < P.n_contact; sc += 64) {
          const double* na = P.c_na + (int64_t)sc * dim;
          const double2* r1 = reinterpret_cast<const double2*>(T + (int64_t)pos[P.ce1[sc]] * MC + m0);
          const double2* r2 =
              reinterpret_cast<const double2*>(T + (int64_t)pos[P.contact_kind == 1 ? P.ce2[sc] : P.ce1[sc]] * MC + m0);
          double v1[KB], v2[KB];
#pragma unroll
          for (int j = 0; j < KB / 2; ++j) {
            const double2 a = r1[j];
            v1[2 * j] = a.x;
            v1[2 * j + 1] = a.y;
            const double2 b = P.contact_kind == 1 ? r2[j] : make_double2(0.0, 0.0);
            v2[2 * j] = b.x;
            v2[2 * j + 1] = b.y;
          }
          const double ct = P.c_t != nullptr ? P.c_t[sc] : 1.0;
#pragma unroll
          for (int j = 0; j < KB; ++j) {
            const int m = m0 + j;
            const double s0 = sh.sig[3 * m], s1 = sh.sig[3 * m + 1], s2 = sh.sig[3 * m + 2];
            double sn = s0 * na[0] + s1 * na[1];
            double area2 = na[0] * na[0] + na[1] * na[1];
            if (dim == 3) {
              sn += s2 * na[2];
              area2 += na[2] * na[2];
            }
            const double t1 = v1[j];
            const double snp = sn > 0.0 ? sn : 0.0;
            const double snm = sn < 0.0 ? sn : 0.0;
            double kc = t1 * snp + P.contact_jump * snm;
            if (P.contact_kind == 1) kc += v2[j] * snm;
            kap[j] += P.c_t != nullptr ? ct * kc : kc;
            sup[j] += t1 * sqrt(area2);
          }
        }
#pragma unroll
        for (int j = 0; j < KB; ++j) {
          double ksum = kap[j], ssum = sup[j];
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) {
            ksum += __shfl_xor(ksum, o, 64);
            ssum += __shfl_xor(ssum, o, 64);
          }
          const int mg = chunk * MC + m0 + j;
          if (tid == 0 && mg < P.n_mfp) {
            P.kappa[(int64_t)mg * L + l] = ksum * P.normalization;
            P.supp[(int64_t)mg * L + l] = ssum * P.normalization;
          }
        }
      }
    }
  }
  sync();  // shared state is reused by the next task of this block
}

// Blocks are dealt to the 8 XCDs round-robin; give all chunks of one
// direction the same XCD so they share its L2 for the direction's operator.
__device__ __forceinline__ int task_of_block(int b, int ntask) {
  constexpr int kXcd = 8;
  const int per = (ntask + kXcd - 1) / kXcd;
  const int xcd = b % kXcd;
  const int idx = b / kXcd;
  const int t = xcd * per + idx;
  return t < ntask ? t : -1;
}

// single problem: one task per block
// minimum waves per SIMD asked of the compiler: the fused final sweep on
// tetrahedra at MC = 4 lands two registers over the three-wave budget
// (168 VGPRs) unasked; every other variant keeps the compiler's choice
template <int MC, int MODE, int KIN, bool WIN>
constexpr int sweep_min_waves() { return (MC == 4 && KIN == 3 && MODE == 3 && WIN) ? 3 : 1; }

template <int MC, int MODE, int KIN, bool WIN>
__global__ __launch_bounds__(kSweepBlock) __attribute__((amdgpu_waves_per_eu(sweep_min_waves<MC, MODE, KIN, WIN>())))
void k_sweep(One src, int ntask, int cap) {
  __shared__ SweepShared<MC> sh;
  const int t = task_of_block(blockIdx.x, ntask);
  if (t < 0) return;
  const obte_sweep_io_t& io = src.io;
  if (io.active != nullptr && io.active[t] == 0) return;  // uniform over the block
  sweep_task<One, MC, MODE, KIN, WIN, units_of<MC, WIN>()>(src, 0, t, sh, cap);
}

// batch: persistent workgroups draining 8 XCD-local task queues (tasks of one
// (problem, direction) share a queue; a drained queue steals from the next).
// Workgroup tasks come first; then every wave of the workgroup drains the 8
// queues of wave tasks on its own (tasks of problems whose levels are narrow:
// one wave covers a level in a pass or two, where a workgroup would leave most
// of its lanes idle), with its own slice of the shared state and window.
// The batch kernel's fused final sweep on tetrahedra at MC = 8 lands seven
// registers over the three-wave budget unasked (three waves: -2% per outer
// iteration of a batch of tetrahedral slabs)
template <int MC, int MODE, int KIN, bool WIN>
constexpr int batch_min_waves() { return (MC == 8 && KIN == 3 && MODE != 2 && WIN) ? 3 : 1; }

// next task of an XCD's queue set, stealing from the other queues once its own
// is drained (*q: queues already found empty); -1 when all are drained
__device__ __forceinline__ int take_task(const int32_t* __restrict__ off, int32_t* __restrict__ cnt, int xcd, int* q) {
  for (; *q < 8; ++*q) {
    const int qq = (xcd + *q) & 7;
    const int i = atomicAdd(&cnt[qq], 1) + off[qq];
    if (i < off[qq + 1]) return i;
  }
  return -1;
}

template <int MC, int MODE, int KIN, bool WIN>
__global__ __launch_bounds__(kSweepBlock) __attribute__((amdgpu_waves_per_eu(batch_min_waves<MC, MODE, KIN, WIN>())))
void k_sweep_batch(Many src, const int32_t* __restrict__ qoff, int32_t* __restrict__ qcnt, int cap,
                   const int32_t* __restrict__ woff, int32_t* __restrict__ wcnt, int wcap) {
  __shared__ BatchShared<MC> sh;
  __shared__ int s_task;
  const int xcd = blockIdx.x % 8;
  int q = 0;
  for (;;) {
    if (threadIdx.x == 0) s_task = take_task(qoff, qcnt, xcd, &q);
    __syncthreads();
    const int t = s_task;
    __syncthreads();
    if (t < 0) break;
    const int pid = src.pid_of(t);
    const int lt = src.local_of(t);
    if (!src.live(pid)) continue;
    const auto& io = src.sio(pid);
    if (io.active != nullptr && io.active[lt] == 0) continue;
    sweep_task<Many, MC, MODE, KIN, WIN, units_batch<MC, WIN>()>(src, pid, lt, sh.blk, cap);
  }
  if (woff == nullptr) return;
  // wave tasks: no workgroup barrier from here on (the last workgroup task
  // ended with one, so the shared state is free)
  const int w = threadIdx.x >> 6;
  q = 0;
  for (;;) {
    int t = -1;
    if ((threadIdx.x & 63) == 0) t = take_task(woff, wcnt, xcd, &q);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t < 0) return;
    const int pid = src.pid_of(t);
    const int lt = src.local_of(t);
    if (!src.live(pid)) continue;
    const auto& io = src.sio(pid);
    if (io.active != nullptr && io.active[lt] == 0) continue;
    sweep_task<Many, MC, MODE, KIN, WIN, units_batch<MC, WIN>(), true, WaveShared<MC>>(src, pid, lt, sh.wav[w], wcap);
  }
}

__global__ void k_zero_queues(int32_t* q, int32_t* wq) {
  if (threadIdx.x < 8) q[threadIdx.x] = 0;
  else if (threadIdx.x < 16 && wq != nullptr) wq[threadIdx.x - 8] = 0;
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

// Block-wide per-lane sums of JB values at once (one reduction for JB dot
// products: 2 barriers instead of 3 per value); v[j] is replaced by the sum
// over the slots of the block, in a fixed order.  red: JB * kSweepBlock
// doubles (value j of every thread at j * kSweepBlock: immediate LDS offsets,
// no per-value address registers; blockDim <= kSweepBlock).
template <int MC, int JB>
__device__ __forceinline__ void block_lane_sums(double (&v)[JB], double* red, double* bc) {
  const int lane = threadIdx.x % MC;
  const int cpp = blockDim.x / MC;
#pragma unroll
  for (int j = 0; j < JB; ++j) red[j * kSweepBlock + threadIdx.x] = v[j];
  __syncthreads();
  for (int idx = threadIdx.x; idx < JB * MC; idx += blockDim.x) {
    const int j = idx / MC, ln = idx - j * MC;
    double acc = 0.0;
    for (int q = 0; q < cpp; ++q) acc += red[j * kSweepBlock + q * MC + ln];
    bc[idx] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < JB; ++j) v[j] = bc[j * MC + lane];
  __syncthreads();
}

constexpr int kGsBlock = 8;  // basis vectors per block of the Gram-Schmidt passes

// cn -= sum_j (c_j . cn) c_j and un -= the same combination of u_j over the
// first kb basis vectors (kAddU: un += it), JB vectors at a time (dot products
// of a block together, blocks in sequence).  Each thread walks its faces
// f = slot, slot + cpp, ... in order; the loads of R consecutive faces are
// issued together before their arithmetic (the loop is bound by memory
// latency: a task's interface is a few hundred faces per thread group), which
// keeps the per-thread operation order and so the results of a face-at-a-time
// loop.  Every vector is read through a buffer resource of the task's slice:
// faces past the end read zero (adding +0 to the sums) and their stores are
// dropped, so the last group of faces needs no separate, load-by-load loop.
template <int MC, bool kAddU = false, class GT>
__device__ void gs_pass(const GT& G, int64_t stride_, int64_t base, int ncut_, int kb_, double* cn,
                        double* un, double* red, double* bc) {
  // uniform over the block (a batch reads them from memory per task)
  const int64_t stride = uniform_i64(stride_);
  const int ncut = __builtin_amdgcn_readfirstlane(ncut_);
  const int kb = __builtin_amdgcn_readfirstlane(kb_);
  constexpr int JB = kGsBlock;
  constexpr int RD = 4;  // faces in flight per thread, dot products
  constexpr int RU = 2;  // faces in flight per thread, update
  const int lane = threadIdx.x % MC;
  const int slot = threadIdx.x / MC;
  const int cpp = blockDim.x / MC;
  const int64_t vbytes = (int64_t)ncut * MC * 8;  // this task's slice of one basis vector
  const double* const gcb = uniform_ptr(G.c + base);
  const double* const gub = uniform_ptr(G.u + base);
  const __amdgpu_buffer_rsrc_t cns = make_rsrc(uniform_ptr(cn), vbytes);
  const __amdgpu_buffer_rsrc_t uns = make_rsrc(uniform_ptr(un), vbytes);
  for (int j0 = 0; j0 < kb; j0 += JB) {
    // one buffer per basis vector of the block (vectors past kb: empty
    // buffers, which read as zero and are not used)
    __amdgpu_buffer_rsrc_t crs[JB], urs[JB];
#pragma unroll
    for (int jj = 0; jj < JB; ++jj) {
      const int64_t vo = (int64_t)(j0 + jj) * stride;
      crs[jj] = make_rsrc(gcb + vo, j0 + jj < kb ? vbytes : 0);
      urs[jj] = make_rsrc(gub + vo, j0 + jj < kb ? vbytes : 0);
    }
    double d[JB];
#pragma unroll
    for (int jj = 0; jj < JB; ++jj) d[jj] = 0.0;
    for (int f = slot; f < ncut; f += RD * cpp) {
      double c0[RD], cc[RD][JB];
#pragma unroll
      for (int r = 0; r < RD; ++r) {
        const int e = (f + r * cpp) * MC + lane;
        c0[r] = buf_load1(cns, e * 8);
#pragma unroll
        for (int jj = 0; jj < JB; ++jj) cc[r][jj] = buf_load1(crs[jj], e * 8);
      }
#pragma unroll
      for (int r = 0; r < RD; ++r)
#pragma unroll
        for (int jj = 0; jj < JB; ++jj)
          if (j0 + jj < kb) d[jj] += cc[r][jj] * c0[r];
    }
    block_lane_sums<MC, JB>(d, red, bc);
    for (int f = slot; f < ncut; f += RU * cpp) {
      double cv[RU], uv[RU], cc[RU][JB], uu[RU][JB];
#pragma unroll
      for (int r = 0; r < RU; ++r) {
        const int e = (f + r * cpp) * MC + lane;
        cv[r] = buf_load1(cns, e * 8);
        uv[r] = buf_load1(uns, e * 8);
#pragma unroll
        for (int jj = 0; jj < JB; ++jj) {
          cc[r][jj] = buf_load1(crs[jj], e * 8);
          uu[r][jj] = buf_load1(urs[jj], e * 8);
        }
      }
#pragma unroll
      for (int r = 0; r < RU; ++r) {
#pragma unroll
        for (int jj = 0; jj < JB; ++jj)
          if (j0 + jj < kb) {
            cv[r] -= d[jj] * cc[r][jj];
            if constexpr (kAddU) uv[r] += d[jj] * uu[r][jj];
            else uv[r] -= d[jj] * uu[r][jj];
          }
        const int e = (f + r * cpp) * MC + lane;
        buf_store1(cns, e * 8, cv[r]);
        buf_store1(uns, e * 8, uv[r]);
      }
    }
  }
}

// true on every thread if pred holds on any thread of the block
__device__ __forceinline__ bool block_any(bool pred) {
  return __syncthreads_or(pred ? 1 : 0) != 0;
}

struct GcrView {
  int64_t stride;  // elements of one interface vector set = n_chunks * n_cut * MC
  int64_t base;    // offset of this (l, chunk) block
  int ncut;
};

template <class PT>
__device__ __forceinline__ GcrView gcr_view(const PT& P, int l, int chunk, int MC) {
  GcrView v;
  v.stride = (int64_t)P.n_chunks * P.n_cut * MC;
  v.base = ((int64_t)chunk * P.n_cut + P.cut_off[l]) * MC;
  v.ncut = P.cut_off[l + 1] - P.cut_off[l];
  return v;
}

// Convergence test of one lane; updates flags.
// count: add the task to *G.n_active when it stays active (the fused
// extension loop counts once, at its end).
template <int MC, class PT, class GT>
__device__ void gcr_flags(const PT& P, const GT& G, int task, int l, int chunk,
                          const GcrView& v, double rn2, double ymax, double* red, double* bc, bool count) {
  __shared__ int blk_any;
  const int lane = threadIdx.x % MC;
  const int slot = threadIdx.x / MC;
  const int mg = chunk * MC + lane;
  const bool valid = mg < P.n_mfp;
  const double rn = sqrt(rn2);
  const double tol = G.rtol * fmax(1.0, ymax) * sqrt((double)(v.ncut > 0 ? v.ncut : 1));
  const int act = (valid && v.ncut > 0 && rn > tol) ? 1 : 0;
  const double any = block_lane_max<MC>(act ? 1.0 : 0.0, red, bc);
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
    if (blk_any && count) atomicAdd(G.n_active, 1);
  }
}

// Per-face loops over a task's interface slice: each thread takes its faces
// f = slot, slot + cpp, ... in order, kFaces of them per step with every load
// issued before the arithmetic; vectors go through buffer resources of the
// slice, so faces past the end read zero and drop their stores.

struct Slice {
  int ncut, slot, cpp, lane, mc;
  __device__ __forceinline__ int off(int f, int q) const { return ((f + q * cpp) * mc + lane) * 8; }
};

template <int MC>
__device__ __forceinline__ Slice slice_of(const GcrView& v) {
  return Slice{v.ncut, (int)threadIdx.x / MC, (int)blockDim.x / MC, (int)threadIdx.x % MC, MC};
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t slice_rsrc(const double* p, const Slice& S) {
  return make_rsrc(uniform_ptr(p), (int64_t)S.ncut * S.mc * 8);
}

// r = x - y; MGS projection of r on the recycled basis; y += U a.  One task
// by the whole block; red, bc: kGsBlock * kSweepBlock and kGsBlock * 64 doubles.
template <int MC, class PT, class GT>
__device__ __forceinline__ void gcr_project_task(const PT& P, const GT& G, int task, double* red, double* bc) {
  const int l = task / P.n_chunks;
  const int chunk = task - l * P.n_chunks;
  const GcrView v = gcr_view(P, l, chunk, MC);
  const Slice S = slice_of<MC>(v);
  double* y = G.y + v.base;
  double* r = G.r + v.base;
  const __amdgpu_buffer_rsrc_t ys = slice_rsrc(y, S), rs = slice_rsrc(r, S), xs = slice_rsrc(G.x + v.base, S);
  for (int f = S.slot; f < S.ncut; f += kFaces * S.cpp) {
    double xv[kFaces], yv[kFaces];
#pragma unroll
    for (int q = 0; q < kFaces; ++q) {
      xv[q] = buf_load1(xs, S.off(f, q));
      yv[q] = buf_load1(ys, S.off(f, q));
    }
#pragma unroll
    for (int q = 0; q < kFaces; ++q) buf_store1(rs, S.off(f, q), xv[q] - yv[q]);
  }
  const int kb = G.k[task];
  // r -= C a, y += U a with a = C^T r
  gs_pass<MC, true>(G, v.stride, v.base, v.ncut, kb, r, y, red, bc);
  double rp = 0.0, ym = 0.0;
  for (int f = S.slot; f < S.ncut; f += kFaces * S.cpp) {
    double rv[kFaces], yv[kFaces];
#pragma unroll
    for (int q = 0; q < kFaces; ++q) {
      rv[q] = buf_load1(rs, S.off(f, q));
      yv[q] = buf_load1(ys, S.off(f, q));
    }
#pragma unroll
    for (int q = 0; q < kFaces; ++q) {
      rp += rv[q] * rv[q];
      ym = fmax(ym, fabs(yv[q]));
    }
  }
  const double rn2 = block_lane_sum<MC>(rp, red, bc);
  const double ymax = block_lane_max<MC>(ym, red, bc);
  gcr_flags<MC>(P, G, task, l, chunk, v, rn2, ymax, red, bc, true);
}

template <class S, int MC>
__global__ __launch_bounds__(kSweepBlock) void k_gcr_project(S src) {
  __shared__ double red[kGsBlock * kSweepBlock];
  __shared__ double bc[kGsBlock * 64];
  const int pid = src.pid_of(blockIdx.x);
  if (!src.live(pid)) return;
  gcr_project_task<MC>(src.plan(pid), src.gcr(pid), src.local_of(blockIdx.x), red, bc);
}

// w = r on unconverged lanes, 0 elsewhere (inflow of the transfer sweep).
template <int MC, class PT, class GT>
__device__ __forceinline__ void gcr_begin_task(const PT& P, const GT& G, int task) {
  const int l = task / P.n_chunks;
  const int chunk = task - l * P.n_chunks;
  const GcrView v = gcr_view(P, l, chunk, MC);
  const Slice S = slice_of<MC>(v);
  const int act = G.lane_active[(int64_t)task * MC + S.lane];
  const __amdgpu_buffer_rsrc_t rs = slice_rsrc(G.r + v.base, S), ws = slice_rsrc(G.w + v.base, S);
  for (int f = S.slot; f < S.ncut; f += kFaces * S.cpp) {
    double rv[kFaces];
#pragma unroll
    for (int q = 0; q < kFaces; ++q) rv[q] = buf_load1(rs, S.off(f, q));
#pragma unroll
    for (int q = 0; q < kFaces; ++q) buf_store1(ws, S.off(f, q), act ? rv[q] : 0.0);
  }
}

template <class S, int MC>
__global__ __launch_bounds__(kSweepBlock) void k_gcr_begin(S src) {
  const int pid = src.pid_of(blockIdx.x);
  if (!src.live(pid)) return;
  const int task = src.local_of(blockIdx.x);
  const auto& G = src.gcr(pid);
  if (G.active[task] == 0) return;
  gcr_begin_task<MC>(src.plan(pid), G, task);
}

// sum over this thread's faces of v^2 (faces in order)
__device__ __forceinline__ double slice_norm2(__amdgpu_buffer_rsrc_t vs, const Slice& S) {
  double part = 0.0;
  for (int f = S.slot; f < S.ncut; f += kFaces * S.cpp) {
    double t[kFaces];
#pragma unroll
    for (int q = 0; q < kFaces; ++q) t[q] = buf_load1(vs, S.off(f, q));
#pragma unroll
    for (int q = 0; q < kFaces; ++q) part += t[q] * t[q];
  }
  return part;
}

// c = w - x (= (I - T_op) w); orthogonalise against the basis, normalise,
// append, then y += a u, r -= a c with a = c . r.  One active task by the
// whole block (red, bc as for gcr_project_task).
template <int MC, class PT, class GT>
__device__ __forceinline__ void gcr_end_task(const PT& P, const GT& G, int task, double* red, double* bc,
                             bool count) {
  const int l = task / P.n_chunks;
  const int chunk = task - l * P.n_chunks;
  const GcrView v = gcr_view(P, l, chunk, MC);
  const Slice S = slice_of<MC>(v);
  int kb = G.k[task];
  if (kb >= G.kmax) kb = 0;  // restart: drop the recycled basis
  double* un = G.u + kb * v.stride + v.base;
  double* cn = G.c + kb * v.stride + v.base;
  const __amdgpu_buffer_rsrc_t us = slice_rsrc(un, S), cs = slice_rsrc(cn, S);
  const __amdgpu_buffer_rsrc_t ws = slice_rsrc(G.w + v.base, S), xs = slice_rsrc(G.x + v.base, S);
  // converged lanes carry (u, c) = (0, 0): the transfer sweep may have skipped
  // their bins, so their exit values are not T_op(0)
  const bool act = G.lane_active[(int64_t)task * MC + S.lane] != 0;
  double part = 0.0;
  for (int f = S.slot; f < S.ncut; f += kFaces * S.cpp) {
    double wv[kFaces], xv[kFaces];
#pragma unroll
    for (int q = 0; q < kFaces; ++q) {
      wv[q] = buf_load1(ws, S.off(f, q));
      xv[q] = buf_load1(xs, S.off(f, q));
    }
#pragma unroll
    for (int q = 0; q < kFaces; ++q) {
      const double c = act ? wv[q] - xv[q] : 0.0;
      buf_store1(us, S.off(f, q), act ? wv[q] : 0.0);
      buf_store1(cs, S.off(f, q), c);
      part += c * c;
    }
  }
  // block Gram-Schmidt; a second pass only where the first one cancelled
  // more than half of the vector's norm ("twice is enough", Kahan-Parlett)
  const double n_before = block_lane_sum<MC>(part, red, bc);
  gs_pass<MC>(G, v.stride, v.base, v.ncut, kb, cn, un, red, bc);
  double n_after = block_lane_sum<MC>(slice_norm2(cs, S), red, bc);
  if (block_any(n_after < 0.25 * n_before)) {  // uniform over the block (it holds barriers)
    gs_pass<MC>(G, v.stride, v.base, v.ncut, kb, cn, un, red, bc);
    n_after = block_lane_sum<MC>(slice_norm2(cs, S), red, bc);
  }
  const double nu = sqrt(n_after);
  const double inv = nu > 0.0 ? 1.0 / nu : 0.0;
  const __amdgpu_buffer_rsrc_t ys = slice_rsrc(G.y + v.base, S), rs = slice_rsrc(G.r + v.base, S);
  part = 0.0;
  for (int f = S.slot; f < S.ncut; f += kFaces * S.cpp) {
    double cv[kFaces], uv[kFaces], rv[kFaces];
#pragma unroll
    for (int q = 0; q < kFaces; ++q) {
      cv[q] = buf_load1(cs, S.off(f, q));
      uv[q] = buf_load1(us, S.off(f, q));
      rv[q] = buf_load1(rs, S.off(f, q));
    }
#pragma unroll
    for (int q = 0; q < kFaces; ++q) {
      cv[q] *= inv;
      uv[q] *= inv;
      buf_store1(cs, S.off(f, q), cv[q]);
      buf_store1(us, S.off(f, q), uv[q]);
      part += cv[q] * rv[q];
    }
  }
  const double a = block_lane_sum<MC>(part, red, bc);
  double rp = 0.0, ym = 0.0;
  for (int f = S.slot; f < S.ncut; f += kFaces * S.cpp) {
    double cv[kFaces], uv[kFaces], rv[kFaces], yv[kFaces];
#pragma unroll
    for (int q = 0; q < kFaces; ++q) {
      cv[q] = buf_load1(cs, S.off(f, q));
      uv[q] = buf_load1(us, S.off(f, q));
      rv[q] = buf_load1(rs, S.off(f, q));
      yv[q] = buf_load1(ys, S.off(f, q));
    }
#pragma unroll
    for (int q = 0; q < kFaces; ++q) {
      yv[q] += a * uv[q];
      rv[q] -= a * cv[q];
      buf_store1(ys, S.off(f, q), yv[q]);
      buf_store1(rs, S.off(f, q), rv[q]);
      rp += rv[q] * rv[q];
      ym = fmax(ym, fabs(yv[q]));
    }
  }
  const double rn2 = block_lane_sum<MC>(rp, red, bc);
  const double ymax = block_lane_max<MC>(ym, red, bc);
  if (threadIdx.x == 0) G.k[task] = kb + 1;
  gcr_flags<MC>(P, G, task, l, chunk, v, rn2, ymax, red, bc, count);
}

template <class S, int MC>
__global__ __launch_bounds__(kSweepBlock) void k_gcr_end(S src) {
  __shared__ double red[kGsBlock * kSweepBlock];
  __shared__ double bc[kGsBlock * 64];
  const int pid = src.pid_of(blockIdx.x);
  if (!src.live(pid)) return;
  const int task = src.local_of(blockIdx.x);
  const auto& G = src.gcr(pid);
  if (G.active[task] == 0) return;
  gcr_end_task<MC>(src.plan(pid), G, task, red, bc, true);
}

// Bytes of dynamic LDS the fused extension loop takes: the sweep window, or
// the Gram-Schmidt reduction buffers, whichever is larger (they take turns).
constexpr size_t kGsLdsBytes = (size_t)(kGsBlock * kSweepBlock + kGsBlock * 64) * sizeof(double);

// The interface Krylov extension loop of one problem in one launch: each
// active (direction, chunk) task runs begin, transfer sweep and end in turn
// until it converges or has made max_extend extensions, without waiting for
// the other tasks (the same steps, in the same order per task, as separate
// begin / sweep / end launches).  At the end *G.n_active counts the tasks
// still active and *ext_max is raised to this task's extension count.
template <int MC, int KIN, bool WIN>
__global__ __launch_bounds__(kSweepBlock) __attribute__((amdgpu_waves_per_eu(3)))
void k_extend(One src, int ntask, int cap, int max_extend, int32_t* ext_max) {
  __shared__ SweepShared<MC> sh;
  const int t = task_of_block(blockIdx.x, ntask);
  if (t < 0) return;
  const obte_bte_plan_t& P = src.P;
  const obte_gcr_t& G = src.G;
  if (G.active[t] == 0) return;  // uniform over the block
  double* const red = reinterpret_cast<double*>(win_lds);
  double* const bc = red + kGsBlock * kSweepBlock;
  int e = 0;
  while (e < max_extend) {
    // the task index passes through an empty asm statement, so nothing the
    // steps derive from it is hoisted out of the loop and kept live across
    // the other steps (register pressure of the sweep)
    int tt = t;
    asm volatile("" : "+s"(tt));
    gcr_begin_task<MC>(P, G, tt);
    __syncthreads();  // w (and the flags) seen by every thread of the sweep
    sweep_task<One, MC, 1, KIN, WIN, units_of<MC, WIN>()>(src, 0, tt, sh, cap);  // ends with a barrier
    asm volatile("" : "+s"(tt));
    gcr_end_task<MC>(P, G, tt, red, bc, false);
    ++e;
    __syncthreads();  // the task's flag, written by thread 0
    if (G.active[t] == 0) break;
  }
  if (threadIdx.x == 0) {
    if (G.active[t] != 0) atomicAdd(G.n_active, 1);
    atomicMax(ext_max, e);
  }
}

// The extension loop of every live problem of a batch in one persistent
// launch: workgroups drain the workgroup task queues, then the wave task
// queues (here each of those tasks is run by the whole workgroup as well), and
// run every active task's begin, transfer sweep and end steps as k_extend
// does, so no task waits for the tasks of other problems or directions.
// *ext_max is raised to the largest extension count of a task.
template <int MC, int KIN, bool WIN>
__global__ __launch_bounds__(kSweepBlock) __attribute__((amdgpu_waves_per_eu(3)))
void k_extend_batch(Many src, const int32_t* __restrict__ qoff, int32_t* __restrict__ qcnt,
                    const int32_t* __restrict__ woff, int32_t* __restrict__ wcnt, int cap, int max_extend,
                    int32_t* ext_max) {
  __shared__ SweepShared<MC> sh;
  __shared__ int s_task;
  double* const red = reinterpret_cast<double*>(win_lds);
  double* const bc = red + kGsBlock * kSweepBlock;
  const int xcd = blockIdx.x % 8;
  for (int set = 0; set < 2; ++set) {
    const int32_t* off = set == 0 ? qoff : woff;
    int32_t* cnt = set == 0 ? qcnt : wcnt;
    if (off == nullptr) break;
    int q = 0;
    for (;;) {
      if (threadIdx.x == 0) s_task = take_task(off, cnt, xcd, &q);
      __syncthreads();
      const int t = __builtin_amdgcn_readfirstlane(s_task);  // uniform: scalar registers
      __syncthreads();
      if (t < 0) break;
      const int pid = __builtin_amdgcn_readfirstlane(src.pid_of(t));
      const int lt = __builtin_amdgcn_readfirstlane(src.local_of(t));
      if (!src.live(pid)) continue;
      const auto& G = src.gcr(pid);
      if (G.active[lt] == 0) continue;  // uniform over the block
      int e = 0;
      while (e < max_extend) {
        // the task passes through an empty asm statement, so nothing derived
        // from it is kept live across the steps (register pressure)
        int tt = lt, pp = pid;
        asm volatile("" : "+s"(tt), "+s"(pp));
        gcr_begin_task<MC>(src.plan(pp), src.gcr(pp), tt);
        __syncthreads();  // w (and the flags) seen by every thread of the sweep
        sweep_task<Many, MC, 1, KIN, WIN, units_batch<MC, WIN>()>(src, pp, tt, sh, cap);  // ends with a barrier
        asm volatile("" : "+s"(tt), "+s"(pp));
        gcr_end_task<MC>(src.plan(pp), src.gcr(pp), tt, red, bc, false);
        ++e;
        __syncthreads();  // the task's flag, written by thread 0
        if (src.gcr(pid).active[lt] == 0) break;
      }
      if (threadIdx.x == 0) {
        if (src.gcr(pid).active[lt] != 0) atomicAdd(src.gcr(pid).n_active, 1);
        atomicMax(ext_max, e);
      }
    }
  }
}

// queue counters (workgroup and wave sets) and the extension count of a batched extension loop
__global__ void k_zero_batch_extend(int32_t* q, int32_t* wq, int32_t* ext_max) {
  if (threadIdx.x < 8) q[threadIdx.x] = 0;
  else if (threadIdx.x < 16 && wq != nullptr) wq[threadIdx.x - 8] = 0;
  else if (threadIdx.x == 16) *ext_max = 0;
}

__global__ void k_zero_extend(int32_t* n_active, int32_t* ext_max) {
  if (threadIdx.x == 0) *n_active = 0;
  else if (threadIdx.x == 1) *ext_max = 0;
}

// zero the active-task counters of every problem (gridDim.x = problems)
template <class S>
__global__ void k_zero_active(S src) {
  if (threadIdx.x == 0) *src.gcr(blockIdx.x).n_active = 0;
}

// y = initial lattice temperature at each cut's exit cell (or 0); basis
// cleared.  One block per (direction, problem).
template <class S>
__global__ void k_gcr_reset(S src, int use_dt) {
  const int pid = blockIdx.y;
  if (!src.live(pid)) return;
  const auto& P = src.plan(pid);
  const auto& G = src.gcr(pid);
  const int l = blockIdx.x;
  if (l >= P.n_dirs) return;
  const int cb = P.cut_off[l];
  const int ncut = P.cut_off[l + 1] - cb;
  const int mc = P.mc;
  const int64_t per = (int64_t)ncut * mc;
  for (int64_t i = threadIdx.x; i < per * P.n_chunks; i += blockDim.x) {
    const int chunk = (int)(i / per);
    const int64_t r = i - chunk * per;
    const int f = (int)(r / mc);
    const int m = (int)(r - (int64_t)f * mc);
    double v = 0.0;
    if (use_dt) v = P.delta_t[P.order[(int64_t)l * P.n + P.cut_exit[cb + f]]];
    G.y[((int64_t)chunk * P.n_cut + cb + f) * mc + m] = v;
  }
  for (int c = threadIdx.x; c < P.n_chunks; c += blockDim.x) G.k[(int64_t)l * P.n_chunks + c] = 0;
}

// Stage 1 (openbte/bte.py:222-230): mode sums of each direction in sweep
// order, interleaved per position (part_tj[l][p (nf + 1) + k]):
//   k = 0:   sum_m t_ml T_ml,
//   k = 1..: sum_m sigma_ml T_ml - DeltaT_old sum_m sigma_ml
//            (separable velocities: sig_a_ml in place of sigma_ml, one sum),
// summed chunk by chunk in order, each chunk's sum formed by chunk_sums.
// k_modesum reads T from the solution workspace; k_modesum_part sums the
// per-chunk sums a windowed final sweep wrote to part_chunk.  Both give the
// same bits.  Grid: y = direction, z = problem.
constexpr int kModesumMaxM = 1024;

// sum over the bins of sigma_ml (sig_a_ml when separable; fixed order), by
// one thread
template <class PT>
__device__ __forceinline__ void sigma_total(const PT& P, int l, double* s) {
  const int L = P.n_dirs, dim = P.dim;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  if (P.sig_a != nullptr) {
    for (int m = 0; m < P.n_mfp; ++m) a0 += P.sig_a[(int64_t)m * L + l];
  } else {
    for (int m = 0; m < P.n_mfp; ++m) {
      const double* sg = P.sigma + ((int64_t)m * L + l) * dim;
      a0 += sg[0];
      a1 += sg[1];
      if (dim == 3) a2 += sg[2];
    }
  }
  s[0] = a0;
  s[1] = a1;
  s[2] = a2;
}

template <class PT>
__device__ __forceinline__ void modesum_store(const PT& P, int l, int p, const double (&acc)[4],
                                              const double* stot) {
  const double dt = P.delta_t[P.order[(int64_t)l * P.n + p]];
  const int nf = n_flux(P);
  double* out = P.part_tj + (int64_t)l * part_stride(P.n, nf) + (int64_t)p * (nf + 1);
  out[0] = acc[0];
  for (int d = 0; d < nf; ++d) out[1 + d] = fma(-dt, stot[d], acc[1 + d]);
}

// From t: one thread per position, the bin pairs' weights staged in LDS.
template <class S, int MC>
__global__ __launch_bounds__(256) void k_modesum(S src) {
  constexpr int G = MC / 2;
  __shared__ __align__(16) double w[8 * (kModesumMaxM / 2)];  // per bin pair: t0 t1 s0x s1x s0y s1y s0z s1z
  __shared__ double stot[4];
  const int pid = blockIdx.z;
  if (!src.live(pid)) return;
  const auto& P = src.plan(pid);
  const int l = blockIdx.y;
  const int L = P.n_dirs, M = P.n_mfp, dim = P.dim, nch = P.n_chunks, n = P.n;
  if (l >= L || (int64_t)blockIdx.x * blockDim.x >= n) return;  // uniform over the block
  for (int q = threadIdx.x; q < nch * G; q += blockDim.x) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int m = 2 * q + h;
      const bool ok = m < M;
      const double* sg = P.sigma + ((int64_t)(ok ? m : 0) * L + l) * dim;
      w[8 * q + h] = ok ? P.t_coeff[(int64_t)m * L + l] : 0.0;
      w[8 * q + 2 + h] = ok ? (P.sig_a != nullptr ? P.sig_a[(int64_t)m * L + l] : sg[0]) : 0.0;
      w[8 * q + 4 + h] = ok ? sg[1] : 0.0;
      w[8 * q + 6 + h] = ok && dim == 3 ? sg[2] : 0.0;
    }
  }
  const int nf = n_flux(P);
  if (threadIdx.x == 0) sigma_total(P, l, stot);
  __syncthreads();
  const int64_t rows = (int64_t)n + P.ncut_max;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = 0; c < nch; ++c) {
      const double2* row = reinterpret_cast<const double2*>(P.t + (((int64_t)l * nch + c) * rows + p) * MC);
      const double2* wc = reinterpret_cast<const double2*>(w + 8 * c * G);
      double e[4] = {0.0, 0.0, 0.0, 0.0};
      chunk_sums<G, false>([&](int j) { return row[j]; }, [&](int k, int j) { return wc[4 * j + k]; }, nf, e);
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[k] += e[k];
    }
    modesum_store(P, l, p, acc, stot);
  }
}

// From the per-chunk sums of a windowed final sweep: one thread per position,
// the chunks added in order.
template <class S>
__global__ __launch_bounds__(256) void k_modesum_part(S src) {
  __shared__ double stot[4];
  const int pid = blockIdx.z;
  if (!src.live(pid)) return;
  const auto& P = src.plan(pid);
  const int l = blockIdx.y;
  const int n = P.n, nf = n_flux(P), nch = P.n_chunks;
  if (l >= P.n_dirs || (int64_t)blockIdx.x * blockDim.x >= n) return;  // uniform over the block
  if (threadIdx.x == 0) sigma_total(P, l, stot);
  __syncthreads();
  const int64_t ps = part_stride(n, nf);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const double* e = P.part_chunk + (int64_t)l * nch * ps + (int64_t)p * (nf + 1);
    // kFaces chunks' sums loaded together, added in chunk order
    for (int c0 = 0; c0 < nch; c0 += kFaces) {
      double v[kFaces][4];
#pragma unroll
      for (int q = 0; q < kFaces; ++q) {
        const double* eq = e + (int64_t)(c0 + q < nch ? c0 + q : c0) * ps;
#pragma unroll
        for (int k = 0; k < 4; ++k) v[q][k] = k <= nf ? eq[k] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < kFaces; ++q)
        if (c0 + q < nch) {
          acc[0] += v[q][0];
          acc[1] += v[q][1];
          if (nf >= 2) acc[2] += v[q][2];
          if (nf == 3) acc[3] += v[q][3];
        }
    }
    modesum_store(P, l, p, acc, stot);
  }
}

// Stage 2: fixed-order sum over directions, gathering each direction's
// partial at the cell's position; one thread per cell.  With separable
// velocities the flux of direction l is sig_dir_l times its one sum.
template <class S>
__global__ void k_reduce_cells(S src) {
  __shared__ double red[256];
  const int pid = blockIdx.y;
  if (!src.live(pid)) return;
  const auto& P = src.plan(pid);
  const auto& R = src.red(pid);
  if ((int64_t)blockIdx.x * blockDim.x >= P.n) return;  // uniform
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int dim = P.dim;
  double s = 0.0;
  if (c < P.n) {
    double j0 = 0.0, j1 = 0.0, j2 = 0.0;
    const int nf = n_flux(P);
    const int64_t ps = part_stride(P.n, nf);
    const int L = P.n_dirs;
    if (P.sig_a != nullptr) {
      for (int l = 0; l < L; ++l) {
        const double* e = P.part_tj + l * ps + (int64_t)P.pos[(int64_t)l * P.n + c] * 2;
        const double* dl = P.sig_dir + (int64_t)l * dim;
        s += e[0];
        j0 = fma(dl[0], e[1], j0);
        j1 = fma(dl[1], e[1], j1);
        if (dim == 3) j2 = fma(dl[2], e[1], j2);
      }
    } else {
      for (int l = 0; l < L; ++l) {
        const double* e = P.part_tj + l * ps + (int64_t)P.pos[(int64_t)l * P.n + c] * (dim + 1);
        s += e[0];
        j0 += e[1];
        j1 += e[2];
        if (dim == 3) j2 += e[3];
      }
    }
    R.dt_new[c] = s;
    R.j_new[(int64_t)c * dim + 0] = j0;
    R.j_new[(int64_t)c * dim + 1] = j1;
    if (dim == 3) R.j_new[(int64_t)c * dim + 2] = j2;
  }
  red[threadIdx.x] = s * s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) P.scratch[blockIdx.x] = red[0];
}

// TB_s = sum over the tasks of part_tb[task][s]: 16 wall sides per block,
// 16 lanes per side each summing every 16th task, the 16 partial sums then
// added in lane order (a fixed order).
constexpr int kWallsPerBlock = 16;

template <class S>
__global__ __launch_bounds__(256) void k_reduce_walls(S src) {
  __shared__ double part[16][kWallsPerBlock + 1];
  const int pid = blockIdx.y;
  if (!src.live(pid)) return;
  const auto& P = src.plan(pid);
  const auto& R = src.red(pid);
  if ((int64_t)blockIdx.x * kWallsPerBlock >= P.n_wall) return;  // uniform
  const int w = threadIdx.x % kWallsPerBlock;
  const int g = threadIdx.x / kWallsPerBlock;
  const int s = blockIdx.x * kWallsPerBlock + w;
  const int ntask = P.n_dirs * P.n_chunks;
  double v = 0.0;
  if (s < P.n_wall)
    for (int t = g; t < ntask; t += 16) v += P.part_tb[(int64_t)t * P.n_wall + s];
  part[g][w] = v;
  __syncthreads();
  if (g == 0 && s < P.n_wall) {
    double acc = 0.0;
    for (int k = 0; k < 16; ++k) acc += part[k][w];
    R.tb_new[s] = acc;
  }
}

// norm2 = sum of the block partials in order; one block per problem
template <class S>
__global__ void k_reduce_final(S src) {
  __shared__ double red[1024];
  const int pid = blockIdx.x;
  if (!src.live(pid)) return;
  const auto& P = src.plan(pid);
  const int nblk = (P.n + 255) / 256;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += P.scratch[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *src.red(pid).norm2 = red[0];
}

// total active tasks of all problems -> *out (records a round of a captured iteration)
__global__ void k_log_active(Many src, int n_problems, int32_t* out) {
  __shared__ int tot;
  if (threadIdx.x == 0) tot = 0;
  __syncthreads();
  for (int p = threadIdx.x; p < n_problems; p += blockDim.x) atomicAdd(&tot, *src.gcr(p).n_active);
  __syncthreads();
  if (threadIdx.x == 0) *out = tot;
}

// batch: DeltaT, TB <- the reduced values (pointers of the plans stay fixed)
__global__ void k_commit(Many src) {
  const int pid = blockIdx.y;
  if (!src.live(pid)) return;
  const auto& P = src.plan(pid);
  const auto& R = src.red(pid);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P.n) const_cast<double*>(P.delta_t)[i] = R.dt_new[i];
  if (i < P.n_wall) const_cast<double*>(P.tb)[i] = R.tb_new[i];
}

int check_plan(const obte_bte_plan_t* P, const char* who) {
  if (!P) return obte_fail(OBTE_EINVAL, who);
  if (P->n <= 0 || P->n_dirs <= 0 || P->n_mfp <= 0 || (P->dim != 2 && P->dim != 3))
    return obte_fail(OBTE_EINVAL, "plan: invalid sizes");
  if (P->mc != 4 && P->mc != 8 && P->mc != 16 && P->mc != 32 && P->mc != 64)
    return obte_fail(OBTE_EINVAL, "plan: mc must be 4, 8, 16, 32 or 64");
  if (P->n_chunks != (P->n_mfp + P->mc - 1) / P->mc) return obte_fail(OBTE_EINVAL, "plan: n_chunks != ceil(M/mc)");
  if (P->block < 64 || P->block > kSweepBlock || (P->block & (P->block - 1)) != 0)
    return obte_fail(OBTE_EINVAL, "plan: block must be 64, 128 or 256");
  if (P->kin != 2 && P->kin != 3) return obte_fail(OBTE_EINVAL, "plan: kin must be 2 or 3");
  if (P->n_mfp > kModesumMaxM) return obte_fail(OBTE_ELIMIT, "plan: more than 1024 MFP bins");
  if (P->ncut_max < 0) return obte_fail(OBTE_EINVAL, "plan: ncut_max < 0");
  if (P->sig_a && !P->sig_dir) return obte_fail(OBTE_EINVAL, "plan: sig_a without sig_dir");
  if ((P->n + 255) / 256 > 4096) return obte_fail(OBTE_ELIMIT, "plan: more than 4096*256 cells (scratch size)");
  return OBTE_OK;
}

int launch_status(const char* who) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return obte_fail_hip(who, e);
  return OBTE_OK;
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <template <int, int> class F, typename... A>
void dispatch(int mc, int kin, A&&... a) {
  const bool k3 = kin == 3;
  switch (mc) {
    case 4: k3 ? F<4, 3>::run(a...) : F<4, 2>::run(a...); break;
    case 8: k3 ? F<8, 3>::run(a...) : F<8, 2>::run(a...); break;
    case 16: k3 ? F<16, 3>::run(a...) : F<16, 2>::run(a...); break;
    case 32: k3 ? F<32, 3>::run(a...) : F<32, 2>::run(a...); break;
    default: k3 ? F<64, 3>::run(a...) : F<64, 2>::run(a...); break;
  }
}

// Positions of the LDS window of a sweep with MC bins whose adjacent levels
// hold at most pair_max cells: the smallest power of two >= pair_max (at least
// 64), or 0 (no window) when disabled, unknown or over kWindowMaxBytes.
constexpr int64_t kWindowMaxBytes = 48 * 1024;

inline int window_cap(int enable, int pair_max, int mc) {
  if (!enable || pair_max <= 0 || pair_max > (1 << 24) || mc > 16) return 0;
  const int64_t cap = min_window(pair_max);
  return cap * mc * (int64_t)sizeof(double) <= kWindowMaxBytes ? (int)cap : 0;
}

template <int MC, int KIN>
struct SweepOne {
  template <bool WIN>
  static void go(const One& src, int mode, int ntask, int block, int cap, hipStream_t st) {
    dim3 grid(((ntask + 7) / 8) * 8), blk(block);
    const size_t lds = WIN ? ((size_t)cap * MC / 2 + 1) * sizeof(double2) : 0;  // window + zero slot
    if (mode == 0) hipLaunchKernelGGL((k_sweep<MC, 0, KIN, WIN>), grid, blk, lds, st, src, ntask, cap);
    else if (mode == 1) hipLaunchKernelGGL((k_sweep<MC, 1, KIN, WIN>), grid, blk, lds, st, src, ntask, cap);
    else if constexpr (WIN) {
      if (mode == 3) hipLaunchKernelGGL((k_sweep<MC, 3, KIN, true>), grid, blk, lds, st, src, ntask, cap);
      else hipLaunchKernelGGL((k_sweep<MC, 2, KIN, true>), grid, blk, lds, st, src, ntask, cap);
    } else {
      hipLaunchKernelGGL((k_sweep<MC, 2, KIN, false>), grid, blk, lds, st, src, ntask, cap);
    }
  }
  // mode 3 = mode 2 forming the per-chunk mode sums (windowed sweeps only)
  static void run(const One& src, int mode, int ntask, int block, int cap, hipStream_t st) {
    if constexpr (MC <= 16) {
      if (cap > 0) return go<true>(src, mode, ntask, block, cap, st);
    }
    go<false>(src, mode == 3 ? 2 : mode, ntask, block, 0, st);
  }
};

// Where the batch kernel finds its task queues and windows.
struct BatchQueues {
  const int32_t* qoff;  // workgroup tasks
  int32_t* qcnt;
  int cap;              // their window positions (0: no window)
  const int32_t* woff;  // wave tasks (null: none)
  int32_t* wcnt;
  int wcap;             // window positions of each wave
};

template <int MC, int KIN>
struct ExtendOne {
  static void run(const One& src, int ntask, int block, int cap, int max_extend, int32_t* ext_max, hipStream_t st) {
    dim3 grid(((ntask + 7) / 8) * 8), blk(block);
    const size_t win = cap > 0 ? ((size_t)cap * MC / 2 + 1) * sizeof(double2) : 0;  // window + zero slot
    const size_t lds = win > kGsLdsBytes ? win : kGsLdsBytes;
    if constexpr (MC <= 16) {
      if (cap > 0) {
        hipLaunchKernelGGL((k_extend<MC, KIN, true>), grid, blk, lds, st, src, ntask, cap, max_extend, ext_max);
        return;
      }
    }
    hipLaunchKernelGGL((k_extend<MC, KIN, false>), grid, blk, lds, st, src, ntask, 0, max_extend, ext_max);
  }
};

template <int MC, int KIN>
struct ExtendMany {
  static void run(const Many& src, int groups, int block, const BatchQueues& Q, int max_extend, int32_t* ext_max,
                  hipStream_t st) {
    dim3 grid(groups), blk(block);
    const size_t win = Q.cap > 0 ? ((size_t)Q.cap * MC / 2 + 1) * sizeof(double2) : 0;  // window + zero slot
    const size_t lds = win > kGsLdsBytes ? win : kGsLdsBytes;
    if constexpr (MC <= 16) {
      if (Q.cap > 0) {
        hipLaunchKernelGGL((k_extend_batch<MC, KIN, true>), grid, blk, lds, st, src, Q.qoff, Q.qcnt, Q.woff, Q.wcnt,
                           Q.cap, max_extend, ext_max);
        return;
      }
    }
    hipLaunchKernelGGL((k_extend_batch<MC, KIN, false>), grid, blk, lds, st, src, Q.qoff, Q.qcnt, Q.woff, Q.wcnt, 0,
                       max_extend, ext_max);
  }
};

template <int MC, int KIN>
struct SweepMany {
  template <bool WIN>
  static void go(const Many& src, int mode, int groups, int block, const BatchQueues& Q, hipStream_t st) {
    dim3 grid(groups), blk(block);
    // the workgroup's window or the waves' windows (each with its zero slot)
    const size_t blk_win = (size_t)Q.cap * MC / 2 + 1;
    const size_t wav_win = Q.woff ? (size_t)(block / 64) * ((size_t)Q.wcap * MC / 2 + 1) : 0;
    const size_t lds = WIN ? (blk_win > wav_win ? blk_win : wav_win) * sizeof(double2) : 0;
    const int cap = WIN ? Q.cap : 0, wcap = WIN ? Q.wcap : 0;
#define OBTE_BATCH_LAUNCH(M, W) \
  hipLaunchKernelGGL((k_sweep_batch<MC, M, KIN, W>), grid, blk, lds, st, src, Q.qoff, Q.qcnt, cap, Q.woff, Q.wcnt, wcap)
    if (mode == 0) OBTE_BATCH_LAUNCH(0, WIN);
    else if (mode == 1) OBTE_BATCH_LAUNCH(1, WIN);
    else if constexpr (WIN) {
      if (mode == 3) OBTE_BATCH_LAUNCH(3, true);
      else OBTE_BATCH_LAUNCH(2, true);
    } else {
      OBTE_BATCH_LAUNCH(2, false);
    }
#undef OBTE_BATCH_LAUNCH
  }
  static void run(const Many& src, int mode, int groups, int block, const BatchQueues& Q, hipStream_t st) {
    if constexpr (MC <= 16) {
      if (Q.cap > 0) return go<true>(src, mode, groups, block, Q, st);
    }
    go<false>(src, mode == 3 ? 2 : mode, groups, block, Q, st);
  }
};

// parts: sum the per-chunk sums of a windowed final sweep; else read t.
// Grid-stride kernels with a bounded number of blocks per direction, so the
// per-block set-up (weights, sum of sigma) is paid a few times per direction.
constexpr int kModesumBlocksPerDir = 24;

template <int MC, int KIN>
struct ModesumL {
  template <class S>
  static void run(const S& src, bool parts, int max_cells, int dirs, int problems, hipStream_t st) {
    if (parts) {
      const unsigned bx = min(blocks_for(max_cells, 256), (unsigned)kModesumBlocksPerDir);
      hipLaunchKernelGGL((k_modesum_part<S>), dim3(bx, dirs, problems), dim3(256), 0, st, src);
    } else {
      const unsigned bx = min(blocks_for(max_cells, 256), (unsigned)kModesumBlocksPerDir);
      hipLaunchKernelGGL((k_modesum<S, MC>), dim3(bx, dirs, problems), dim3(256), 0, st, src);
    }
  }
};

// GCR kernels: op 0 project, 1 begin, 2 end
template <int MC, int KIN>
struct GcrL {
  template <class S>
  static void run(const S& src, int op, int nblocks, int block, hipStream_t st) {
    if (op == 0) hipLaunchKernelGGL((k_gcr_project<S, MC>), dim3(nblocks), dim3(block), 0, st, src);
    else if (op == 1) hipLaunchKernelGGL((k_gcr_begin<S, MC>), dim3(nblocks), dim3(block), 0, st, src);
    else hipLaunchKernelGGL((k_gcr_end<S, MC>), dim3(nblocks), dim3(block), 0, st, src);
  }
};

int check_gcr(const obte_gcr_t* G) {
  if (!G || !G->y || !G->r || !G->x || !G->u || !G->c || !G->k || !G->active || !G->lane_active || !G->n_active ||
      !G->rnorm || !G->w || G->kmax <= 0)
    return obte_fail(OBTE_EINVAL, "gcr: null pointer or kmax <= 0");
  return OBTE_OK;
}

One make_one(const obte_bte_plan_t* plan, const obte_gcr_t* g) {
  One o{};
  o.P = *plan;
  if (g) o.G = *g;
  return o;
}

int check_batch(const obte_bte_batch_t* B, const char* who) {
  if (!B || !B->plans || !B->gcrs || !B->io || !B->rio || !B->live || !B->task_pid || !B->task_local ||
      !B->queue_off || !B->queue_cnt || B->n_problems <= 0 || B->n_tasks <= 0)
    return obte_fail(OBTE_EINVAL, who);
  if (B->mc != 4 && B->mc != 8 && B->mc != 16 && B->mc != 32 && B->mc != 64)
    return obte_fail(OBTE_EINVAL, "batch: mc must be 4, 8, 16, 32 or 64");
  if (B->block < 64 || B->block > kSweepBlock || (B->block & (B->block - 1)) != 0)
    return obte_fail(OBTE_EINVAL, "batch: block must be 64, 128 or 256");
  if (B->kin != 2 && B->kin != 3) return obte_fail(OBTE_EINVAL, "batch: kin must be 2 or 3");
  if (B->groups <= 0 || B->max_elems <= 0 || B->max_dirs <= 0 || B->max_cells <= 0)
    return obte_fail(OBTE_EINVAL, "batch: invalid launch sizes");
  if (B->wave_off != nullptr) {
    if (!B->wave_cnt || B->wave_pair_max <= 0 || B->wave_pair_max > B->lvl_pair_max)
      return obte_fail(OBTE_EINVAL, "batch: wave tasks need wave_cnt and 0 < wave_pair_max <= lvl_pair_max");
    const int64_t wav = (int64_t)(B->block / 64) * ((int64_t)min_window(B->wave_pair_max) * B->mc / 2 + 1) * 16;
    if (wav > kWindowMaxBytes) return obte_fail(OBTE_EINVAL, "batch: the waves' windows exceed the LDS budget");
  }
  return OBTE_OK;
}

Many make_many(const obte_bte_batch_t* B, int mode) {
  Many m{};
  m.P = (const OBTE_CONST obte_bte_plan_t*)B->plans;
  m.G = (const OBTE_CONST obte_gcr_t*)B->gcrs;
  m.io = (const OBTE_CONST obte_sweep_io_t*)(B->io + (int64_t)mode * B->n_problems);
  m.rio = (const OBTE_CONST obte_reduce_io_t*)B->rio;
  m.livef = (const OBTE_CONST int32_t*)B->live;
  m.tpid = (const OBTE_CONST int32_t*)B->task_pid;
  m.tloc = (const OBTE_CONST int32_t*)B->task_local;
  return m;
}

}  // namespace

extern "C" {

// ---- single problem -------------------------------------------------------

int obte_bte_assemble(const obte_bte_plan_t* plan, obte_stream_t stream) {
  int rc = check_plan(plan, "obte_bte_assemble: null plan");
  if (rc) return rc;
  const int64_t total = (int64_t)plan->n_dirs * plan->n;
  hipLaunchKernelGGL((k_assemble<One>), dim3(blocks_for(total, 256), 1), dim3(256), 0, (hipStream_t)stream,
                     make_one(plan, nullptr));
  hipLaunchKernelGGL((k_mark_keep<One>), dim3(blocks_for(total, 256), 1), dim3(256), 0, (hipStream_t)stream,
                     make_one(plan, nullptr));
  hipLaunchKernelGGL((k_mark_exits<One>), dim3(64, 1), dim3(256), 0, (hipStream_t)stream, make_one(plan, nullptr));
  hipLaunchKernelGGL((k_mark_cells<One>), dim3(64, 1), dim3(256), 0, (hipStream_t)stream, make_one(plan, nullptr));
  return launch_status("obte_bte_assemble: launch failed");
}

int obte_bte_source(const obte_bte_plan_t* plan, obte_stream_t stream) {
  int rc = check_plan(plan, "obte_bte_source: null plan");
  if (rc) return rc;
  const int64_t total = (int64_t)plan->n_dirs * plan->n;
  hipLaunchKernelGGL((k_source<One>), dim3(blocks_for(total, 256), 1), dim3(256), 0, (hipStream_t)stream,
                     make_one(plan, nullptr));
  return launch_status("obte_bte_source: launch failed");
}

int32_t obte_sweep_window(int32_t sweep_window, int32_t lvl_pair_max, int32_t mc) {
  return window_cap(sweep_window, lvl_pair_max, mc);
}

int obte_bte_sweep(const obte_bte_plan_t* plan, int32_t mode, const double* y, double* x, const int32_t* active,
                   const int32_t* lane_active, obte_stream_t stream) {
  int rc = check_plan(plan, "obte_bte_sweep: null plan");
  if (rc) return rc;
  if (mode < 0 || mode > 2) return obte_fail(OBTE_EINVAL, "obte_bte_sweep: mode must be 0, 1 or 2");
  if (plan->n_cut > 0 && !y) return obte_fail(OBTE_EINVAL, "obte_bte_sweep: y is required when cuts exist");
  One src = make_one(plan, nullptr);
  src.io.y = y;
  src.io.x = x;
  src.io.active = active;
  src.io.lane_active = lane_active;
  const int cap = window_cap(plan->sweep_window, plan->lvl_pair_max, plan->mc);
  const int m = mode == 2 && cap > 0 && plan->part_chunk ? 3 : (int)mode;
  dispatch<SweepOne>(plan->mc, plan->kin, src, m, plan->n_dirs * plan->n_chunks, plan->block, cap,
                     (hipStream_t)stream);
  return launch_status("obte_bte_sweep: launch failed");
}

static int gcr_one(const obte_bte_plan_t* plan, const obte_gcr_t* g, int op, obte_stream_t stream, const char* who) {
  int rc = check_plan(plan, who);
  if (rc || (rc = check_gcr(g))) return rc;
  One src = make_one(plan, g);
  hipStream_t st = (hipStream_t)stream;
  if (op != 1) hipLaunchKernelGGL((k_zero_active<One>), dim3(1), dim3(64), 0, st, src);
  dispatch<GcrL>(plan->mc, 2, src, op, plan->n_dirs * plan->n_chunks, plan->block, st);
  return launch_status(who);
}

int obte_gcr_project(const obte_bte_plan_t* plan, const obte_gcr_t* g, obte_stream_t stream) {
  return gcr_one(plan, g, 0, stream, "obte_gcr_project");
}

int obte_gcr_begin_extend(const obte_bte_plan_t* plan, const obte_gcr_t* g, obte_stream_t stream) {
  return gcr_one(plan, g, 1, stream, "obte_gcr_begin_extend");
}

int obte_gcr_end_extend(const obte_bte_plan_t* plan, const obte_gcr_t* g, obte_stream_t stream) {
  return gcr_one(plan, g, 2, stream, "obte_gcr_end_extend");
}

int obte_gcr_extend(const obte_bte_plan_t* plan, const obte_gcr_t* g, int32_t max_extend, int32_t* ext_max,
                    obte_stream_t stream) {
  int rc = check_plan(plan, "obte_gcr_extend: null plan");
  if (rc || (rc = check_gcr(g))) return rc;
  if (max_extend < 0 || !ext_max) return obte_fail(OBTE_EINVAL, "obte_gcr_extend: max_extend < 0 or null ext_max");
  One src = make_one(plan, g);
  src.io.y = g->w;
  src.io.x = g->x;
  src.io.active = g->active;
  src.io.lane_active = g->lane_active;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_extend, dim3(1), dim3(64), 0, st, g->n_active, ext_max);
  if (max_extend > 0) {
    const int cap = window_cap(plan->sweep_window, plan->lvl_pair_max, plan->mc);
    dispatch<ExtendOne>(plan->mc, plan->kin, src, plan->n_dirs * plan->n_chunks, plan->block, cap, (int)max_extend,
                        ext_max, st);
  }
  return launch_status("obte_gcr_extend: launch failed");
}

int obte_gcr_reset(const obte_bte_plan_t* plan, const obte_gcr_t* g, int32_t use_delta_t, obte_stream_t stream) {
  int rc = check_plan(plan, "obte_gcr_reset: null plan");
  if (rc || (rc = check_gcr(g))) return rc;
  hipLaunchKernelGGL((k_gcr_reset<One>), dim3(plan->n_dirs, 1), dim3(256), 0, (hipStream_t)stream,
                     make_one(plan, g), (int)use_delta_t);
  return launch_status("obte_gcr_reset: launch failed");
}

int obte_bte_reduce(const obte_bte_plan_t* plan, double* delta_t_new, double* j_new, double* tb_new, double* norm2,
                    obte_stream_t stream) {
  int rc = check_plan(plan, "obte_bte_reduce: null plan");
  if (rc) return rc;
  if (!delta_t_new || !j_new || !norm2 || !plan->scratch) return obte_fail(OBTE_EINVAL, "obte_bte_reduce: null output");
  if (plan->n_wall > 0 && !tb_new) return obte_fail(OBTE_EINVAL, "obte_bte_reduce: null tb output");
  hipStream_t st = (hipStream_t)stream;
  One src = make_one(plan, nullptr);
  src.rio.dt_new = delta_t_new;
  src.rio.j_new = j_new;
  src.rio.tb_new = tb_new;
  src.rio.norm2 = norm2;
  const bool parts = window_cap(plan->sweep_window, plan->lvl_pair_max, plan->mc) > 0 && plan->part_chunk;
  dispatch<ModesumL>(plan->mc, 2, src, parts, plan->n, plan->n_dirs, 1, st);
  hipLaunchKernelGGL((k_reduce_cells<One>), dim3(blocks_for(plan->n, 256), 1), dim3(256), 0, st, src);
  if (plan->n_wall > 0)
    hipLaunchKernelGGL((k_reduce_walls<One>), dim3(blocks_for(plan->n_wall, kWallsPerBlock), 1), dim3(256), 0, st,
                       src);
  hipLaunchKernelGGL((k_reduce_final<One>), dim3(1), dim3(1024), 0, st, src);
  return launch_status("obte_bte_reduce: launch failed");
}

// ---- batch of problems ----------------------------------------------------

int obte_batch_assemble(const obte_bte_batch_t* b, obte_stream_t stream) {
  int rc = check_batch(b, "obte_batch_assemble: invalid batch");
  if (rc) return rc;
  hipLaunchKernelGGL((k_assemble<Many>), dim3(blocks_for(b->max_elems, 256), b->n_problems), dim3(256), 0,
                     (hipStream_t)stream, make_many(b, 0));
  hipLaunchKernelGGL((k_mark_keep<Many>), dim3(blocks_for(b->max_elems, 256), b->n_problems), dim3(256), 0,
                     (hipStream_t)stream, make_many(b, 0));
  hipLaunchKernelGGL((k_mark_exits<Many>), dim3(64, b->n_problems), dim3(256), 0, (hipStream_t)stream,
                     make_many(b, 0));
  hipLaunchKernelGGL((k_mark_cells<Many>), dim3(64, b->n_problems), dim3(256), 0, (hipStream_t)stream,
                     make_many(b, 0));
  return launch_status("obte_batch_assemble: launch failed");
}

int obte_batch_source(const obte_bte_batch_t* b, obte_stream_t stream) {
  int rc = check_batch(b, "obte_batch_source: invalid batch");
  if (rc) return rc;
  hipLaunchKernelGGL((k_source<Many>), dim3(blocks_for(b->max_elems, 256), b->n_problems), dim3(256), 0,
                     (hipStream_t)stream, make_many(b, 0));
  return launch_status("obte_batch_source: launch failed");
}

int obte_batch_sweep(const obte_bte_batch_t* b, int32_t mode, obte_stream_t stream) {
  int rc = check_batch(b, "obte_batch_sweep: invalid batch");
  if (rc) return rc;
  if (mode < 0 || mode > 2) return obte_fail(OBTE_EINVAL, "obte_batch_sweep: mode must be 0, 1 or 2");
  hipStream_t st = (hipStream_t)stream;
  BatchQueues Q{b->queue_off, b->queue_cnt, window_cap(b->sweep_window, b->lvl_pair_max, b->mc), nullptr, nullptr, 0};
  if (b->wave_off != nullptr) {
    Q.woff = b->wave_off;
    Q.wcnt = b->wave_cnt;
    Q.wcap = min_window(b->wave_pair_max);
  }
  hipLaunchKernelGGL(k_zero_queues, dim3(1), dim3(64), 0, st, b->queue_cnt, b->wave_cnt);
  const int m = mode == 2 && Q.cap > 0 && b->part_sums ? 3 : (int)mode;
  dispatch<SweepMany>(b->mc, b->kin, make_many(b, mode), m, b->groups, b->block, Q, st);
  return launch_status("obte_batch_sweep: launch failed");
}

int obte_batch_gcr(const obte_bte_batch_t* b, int32_t op, obte_stream_t stream) {
  int rc = check_batch(b, "obte_batch_gcr: invalid batch");
  if (rc) return rc;
  if (op < 0 || op > 2) return obte_fail(OBTE_EINVAL, "obte_batch_gcr: op must be 0, 1 or 2");
  hipStream_t st = (hipStream_t)stream;
  Many src = make_many(b, 0);
  if (op != 1) hipLaunchKernelGGL((k_zero_active<Many>), dim3(b->n_problems), dim3(64), 0, st, src);
  dispatch<GcrL>(b->mc, 2, src, (int)op, b->n_tasks, b->block, st);
  return launch_status("obte_batch_gcr: launch failed");
}

int obte_batch_extend(const obte_bte_batch_t* b, int32_t max_extend, int32_t* ext_max, obte_stream_t stream) {
  int rc = check_batch(b, "obte_batch_extend: invalid batch");
  if (rc) return rc;
  if (max_extend < 0 || !ext_max) return obte_fail(OBTE_EINVAL, "obte_batch_extend: max_extend < 0 or null ext_max");
  hipStream_t st = (hipStream_t)stream;
  Many src = make_many(b, 1);
  hipLaunchKernelGGL((k_zero_active<Many>), dim3(b->n_problems), dim3(64), 0, st, src);
  hipLaunchKernelGGL(k_zero_batch_extend, dim3(1), dim3(64), 0, st, b->queue_cnt, b->wave_cnt, ext_max);
  if (max_extend > 0) {
    // every task by a whole workgroup: the window holds the widest plan's level pairs
    BatchQueues Q{b->queue_off, b->queue_cnt, window_cap(b->sweep_window, b->lvl_pair_max, b->mc), b->wave_off,
                  b->wave_cnt, 0};
    dispatch<ExtendMany>(b->mc, b->kin, src, b->groups, b->block, Q, (int)max_extend, ext_max, st);
  }
  return launch_status("obte_batch_extend: launch failed");
}

int obte_batch_gcr_reset(const obte_bte_batch_t* b, int32_t use_delta_t, obte_stream_t stream) {
  int rc = check_batch(b, "obte_batch_gcr_reset: invalid batch");
  if (rc) return rc;
  hipLaunchKernelGGL((k_gcr_reset<Many>), dim3(b->max_dirs, b->n_problems), dim3(256), 0, (hipStream_t)stream,
                     make_many(b, 0), (int)use_delta_t);
  return launch_status("obte_batch_gcr_reset: launch failed");
}

int obte_batch_log_active(const obte_bte_batch_t* b, int32_t* out, obte_stream_t stream) {
  int rc = check_batch(b, "obte_batch_log_active: invalid batch");
  if (rc) return rc;
  if (!out) return obte_fail(OBTE_EINVAL, "obte_batch_log_active: null output");
  hipLaunchKernelGGL(k_log_active, dim3(1), dim3(256), 0, (hipStream_t)stream, make_many(b, 0), b->n_problems, out);
  return launch_status("obte_batch_log_active: launch failed");
}

int obte_batch_reduce(const obte_bte_batch_t* b, obte_stream_t stream) {
  int rc = check_batch(b, "obte_batch_reduce: invalid batch");
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  Many src = make_many(b, 0);
  const bool parts = window_cap(b->sweep_window, b->lvl_pair_max, b->mc) > 0 && b->part_sums;
  dispatch<ModesumL>(b->mc, 2, src, parts, b->max_cells, b->max_dirs, b->n_problems, st);
  hipLaunchKernelGGL((k_reduce_cells<Many>), dim3(blocks_for(b->max_cells, 256), b->n_problems), dim3(256), 0, st,
                     src);
  if (b->max_walls > 0)
    hipLaunchKernelGGL((k_reduce_walls<Many>), dim3(blocks_for(b->max_walls, kWallsPerBlock), b->n_problems),
                       dim3(256), 0, st, src);
  hipLaunchKernelGGL((k_reduce_final<Many>), dim3(b->n_problems), dim3(1024), 0, st, src);
  hipLaunchKernelGGL(k_commit, dim3(blocks_for(b->max_cells > b->max_walls ? b->max_cells : b->max_walls, 256),
                                    b->n_problems), dim3(256), 0, st, src);
  return launch_status("obte_batch_reduce: launch failed");
}

}  // extern "C"
