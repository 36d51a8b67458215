// Adjoint face gradient of the interfacial-transmission model (FP64, gfx950).
//
// With the converged forward solutions T_ml and the scaled adjoint solutions
// mu'_ml (the adjoint of direction l is solved as the forward sweep of the
// reversed direction lbar(l), see openbte_amd/tim.py), the derivative of
// kappa_eff with respect to the transmission t of a directed face f (row cell
// i, upwind cell j, reflecting side k) is
//
//   dk/dt_f = -V_i sum_l min(s_l . k_f, 0) sum_m t_ml Lambda_m mu'_{m,lbar(l)}[i]
//                                              (T_ml[j] - J_f - TB_k)
//
// (J_f: the face's periodic jump without transmission, TB_k: converged
// temperature of the reflecting side).  One wavefront per directed face; lane
// l sums its direction over the MFP bins, then a fixed xor tree adds the
// directions, so the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "capi_common.h"

namespace {

constexpr int kWave = 64;

__global__ __launch_bounds__(256) void k_tim_face_grad(obte_bte_plan_t P, const double* __restrict__ t_fwd,
                                                       const double* __restrict__ t_adj,
                                                       const int32_t* __restrict__ lbar,
                                                       const int32_t* __restrict__ i_face,
                                                       const int32_t* __restrict__ j_face,
                                                       const int32_t* __restrict__ side_of_face,
                                                       const double* __restrict__ tb,
                                                       const double* __restrict__ jump0,
                                                       const double* __restrict__ vol, int n_faces,
                                                       int l0, int l1, int accumulate,
                                                       double* __restrict__ grad) {
  const int lane = threadIdx.x % kWave;
  const int f = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kWave);
  if (f >= n_faces) return;  // uniform over the wavefront
  const int n = P.n, L = P.n_dirs, M = P.n_mfp, dim = P.dim, mc = P.mc, nch = P.n_chunks;
  const int i = i_face[f], j = j_face[f];
  const int k = side_of_face[f];
  const double shift = jump0[f] + (k >= 0 ? tb[k] : 0.0);
  const int64_t rows = (int64_t)n + P.ncut_max;
  const double* kf = P.k_face + (int64_t)f * dim;
  double acc = 0.0;
  for (int l = l0 + lane; l < l1; l += kWave) {
    const double* s = P.dirs + (int64_t)l * dim;
    double g = s[0] * kf[0] + s[1] * kf[1];
    if (dim == 3) g += s[2] * kf[2];
    if (g >= 0.0) continue;
    const int d = lbar[l];
    const int pj = P.pos[(int64_t)l * n + j];
    const int pi = P.pos[(int64_t)d * n + i];
    double sl = 0.0;
    for (int c = 0; c < nch; ++c) {
      const double* tf = t_fwd + (((int64_t)(l - l0) * nch + c) * rows + pj) * mc;
      const double* ta = t_adj + (((int64_t)d * nch + c) * rows + pi) * mc;
      for (int b = 0; b < mc; ++b) {
        const int m = c * mc + b;
        if (m >= M) break;
        sl += P.t_coeff[(int64_t)m * L + l] * P.mfps[m] * ta[b] * (tf[b] - shift);
      }
    }
    acc += g * sl;
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
  if (lane == 0) grad[f] = (accumulate ? grad[f] : 0.0) - vol[i] * acc;
}

int face_grad(const char* who, const obte_bte_plan_t* plan, const double* t_fwd, int32_t l0, int32_t l1,
              int32_t accumulate, const double* t_adj, const int32_t* lbar, const int32_t* i_face,
              const int32_t* j_face, const int32_t* side_of_face, const double* tb, const double* jump0,
              const double* vol, int32_t n_faces, double* grad, obte_stream_t stream) {
  if (!plan || !t_fwd || !t_adj || !lbar || !i_face || !j_face || !side_of_face || !jump0 || !vol || !grad)
    return obte_fail(OBTE_EINVAL, who);
  if (n_faces < 0 || n_faces > plan->n_faces) return obte_fail(OBTE_EINVAL, "obte_tim_face_grad: bad n_faces");
  if (plan->n <= 0 || plan->n_dirs <= 0 || plan->n_mfp <= 0 || (plan->dim != 2 && plan->dim != 3))
    return obte_fail(OBTE_EINVAL, "obte_tim_face_grad: invalid plan sizes");
  if (l0 < 0 || l1 > plan->n_dirs || l0 >= l1) return obte_fail(OBTE_EINVAL, "obte_tim_face_grad: bad direction range");
  if (n_faces == 0) return OBTE_OK;
  const int per_block = 256 / kWave;
  hipLaunchKernelGGL(k_tim_face_grad, dim3((n_faces + per_block - 1) / per_block), dim3(256), 0,
                     (hipStream_t)stream, *plan, t_fwd, t_adj, lbar, i_face, j_face, side_of_face, tb, jump0, vol,
                     (int)n_faces, (int)l0, (int)l1, (int)accumulate, grad);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return obte_fail_hip("obte_tim_face_grad: launch failed", e);
  return OBTE_OK;
}

}  // namespace

extern "C" {

int obte_tim_face_grad(const obte_bte_plan_t* plan, const double* t_fwd, const double* t_adj, const int32_t* lbar,
                       const int32_t* i_face, const int32_t* j_face, const int32_t* side_of_face, const double* tb,
                       const double* jump0, const double* vol, int32_t n_faces, double* grad,
                       obte_stream_t stream) {
  return face_grad("obte_tim_face_grad: null argument", plan, t_fwd, 0, plan ? plan->n_dirs : 0, 0, t_adj, lbar,
                   i_face, j_face, side_of_face, tb, jump0, vol, n_faces, grad, stream);
}

int obte_tim_face_grad_dirs(const obte_bte_plan_t* plan, const double* t_fwd_dirs, int32_t l0, int32_t l1,
                            int32_t accumulate, const double* t_adj, const int32_t* lbar, const int32_t* i_face,
                            const int32_t* j_face, const int32_t* side_of_face, const double* tb,
                            const double* jump0, const double* vol, int32_t n_faces, double* grad,
                            obte_stream_t stream) {
  return face_grad("obte_tim_face_grad_dirs: null argument", plan, t_fwd_dirs, l0, l1, accumulate, t_adj, lbar,
                   i_face, j_face, side_of_face, tb, jump0, vol, n_faces, grad, stream);
}

}  // extern "C"
