// Interpolation of phonon modes onto (MFP bin, direction) grids (FP64, gfx950).
//
// Each mode k (weight W_k = C_k / tau_k, flux sigma_k = C_k v_k, vectorial MFP
// v_k tau_k) is split linearly over the two nearest MFP bins and, per angle,
// over the two nearest grid angles (periodic in phi, extrapolated at the ends
// in theta and MFP), and added into W[m][l] and sigma[m][l][:]
// (openbte/material.py:7-102 RTA3D, 105-189 RTA2DSym, with the index rule of
// openbte/utils.py:281-320 fast_interpolation).  One thread per mode, FP64
// atomic adds in L2: the sums are exact up to the order of the additions.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "capi_common.h"

namespace {

// fast_interpolation on an increasing grid c[0..n-1]: the first index with
// c[m2] >= x, extrapolating linearly past either end ("extent").
__device__ void interp_extent(double x, const double* c, int n, int& m1, int& m2, double& a1, double& a2) {
  int lo = 0, hi = n;  // first index with c >= x, n if none
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (c[mid] >= x) hi = mid;
    else lo = mid + 1;
  }
  if (x < c[0]) {
    m1 = 0;
    m2 = 1;
  } else if (lo >= n) {  // x > c[n-1]
    m1 = n - 2;
    m2 = n - 1;
  } else if (lo == 0) {  // x == c[0]: the reference's argmax gives m2 = 0, m1 = -1
    m1 = n - 1;  // c[-1] in numpy; a2 = 1 puts all weight on m2
    m2 = 0;
    a2 = 1.0;
    a1 = 0.0;
    return;
  } else {
    m1 = lo - 1;
    m2 = lo;
  }
  a2 = (x - c[m1]) / (c[m2] - c[m1]);
  a1 = 1.0 - a2;
}

// periodic variant (phi in [0, 2 pi))
__device__ void interp_periodic(double x, const double* c, int n, int& m1, int& m2, double& a1, double& a2) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (c[mid] >= x) hi = mid;
    else lo = mid + 1;
  }
  const double delta = c[n - 1] - c[n - 2];
  if (lo == 0 || lo >= n) {  // before the first or past the last grid angle: wrap
    m1 = n - 1;
    m2 = 0;
    double xx = x;
    if (xx < 0.5 * delta) xx += 2.0 * M_PI;
    a2 = (xx - c[n - 1]) / delta;
  } else {
    m1 = lo - 1;
    m2 = lo;
    a2 = (x - c[m1]) / (c[m2] - c[m1]);
  }
  a1 = 1.0 - a2;
}

__global__ void k_rta_scatter(int n_modes, int dim, const double* __restrict__ weight,
                              const double* __restrict__ sigma, const double* __restrict__ mfp, int n_mfp,
                              const double* __restrict__ mfp_grid, int n_phi, const double* __restrict__ phi_grid,
                              int n_theta, const double* __restrict__ theta_grid, double* __restrict__ w_out,
                              double* __restrict__ s_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_modes) return;
  const double mx = mfp[3 * k], my = mfp[3 * k + 1], mz = mfp[3 * k + 2];
  const double r = dim == 3 ? sqrt(mx * mx + my * my + mz * mz) : sqrt(mx * mx + my * my);
  double phi = atan2(mx, my);
  if (phi < 0.0) phi += 2.0 * M_PI;
  int m1, m2, p1, p2, t1 = 0, t2 = 0;
  double a1, a2, b1, b2, c1 = 1.0, c2 = 0.0;
  interp_extent(r, mfp_grid, n_mfp, m1, m2, a1, a2);
  interp_periodic(phi, phi_grid, n_phi, p1, p2, b1, b2);
  if (dim == 3) interp_extent(acos(mz / r), theta_grid, n_theta, t1, t2, c1, c2);
  const double w = weight[k];
  const double s0 = sigma[3 * k], s1 = sigma[3 * k + 1], s2 = sigma[3 * k + 2];
  const int mm[2] = {m1, m2};
  const double am[2] = {a1, a2};
  const int pp[2] = {p1, p2};
  const double bp[2] = {b1, b2};
  const int tt[2] = {t1, t2};
  const double ct[2] = {c1, c2};
  const int n_ang = dim == 3 ? n_theta * n_phi : n_phi;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int q = 0; q < (dim == 3 ? 2 : 1); ++q) {
        const double f = am[i] * bp[j] * ct[q];
        const int l = (dim == 3 ? tt[q] * n_phi : 0) + pp[j];
        const int64_t o = (int64_t)mm[i] * n_ang + l;
        atomicAdd(w_out + o, f * w);
        atomicAdd(s_out + o * dim, f * s0);
        atomicAdd(s_out + o * dim + 1, f * s1);
        if (dim == 3) atomicAdd(s_out + o * dim + 2, f * s2);
      }
}

}  // namespace

extern "C" {

int obte_rta_scatter(int32_t n_modes, int32_t dim, const double* weight, const double* sigma, const double* mfp,
                     int32_t n_mfp, const double* mfp_grid, int32_t n_phi, const double* phi_grid, int32_t n_theta,
                     const double* theta_grid, double* w_out, double* sigma_out, obte_stream_t stream) {
  if (n_modes < 0 || (dim != 2 && dim != 3) || n_mfp < 2 || n_phi < 2 || (dim == 3 && n_theta < 2))
    return obte_fail(OBTE_EINVAL, "obte_rta_scatter: invalid sizes");
  if (!weight || !sigma || !mfp || !mfp_grid || !phi_grid || !w_out || !sigma_out || (dim == 3 && !theta_grid))
    return obte_fail(OBTE_EINVAL, "obte_rta_scatter: null argument");
  if (n_modes == 0) return OBTE_OK;
  hipLaunchKernelGGL(k_rta_scatter, dim3((n_modes + 255) / 256), dim3(256), 0, (hipStream_t)stream, (int)n_modes,
                     (int)dim, weight, sigma, mfp, (int)n_mfp, mfp_grid, (int)n_phi, phi_grid, (int)n_theta,
                     theta_grid, w_out, sigma_out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return obte_fail_hip("obte_rta_scatter: launch failed", e);
  return OBTE_OK;
}

}  // extern "C"
