// Element-level device code of the orbital rotation and the ADAM-family step, shared by
// the single-system kernels (dn_rot.hip) and the grouped lock-step kernels
// (dn_serve.hip), so that both compute every element with the same arithmetic.
#pragma once
#include <hip/hip_runtime.h>

#include "donof_hip.h"

namespace dn {

// packed index of the rotation variable (i, j), i < nbf5, j > i (row-wise order of
// utils.jl:503-510)
__device__ __forceinline__ long pair_index(int i, int j, int nbf) {
  return (long)i * nbf - (long)i * (i + 1) / 2 + (j - i - 1);
}

// One AMSGrad-family update (minimization.jl:406-411 ADAM, 503-508 ADABelief, 598-603
// YOGI, 307-316 Demon) of packed element e = i * nbf + j (j > i, i < nbf5); writes
// Y[i,j] = y, Y[j,i] = -y.
__device__ __forceinline__ void adam_update(long e, const double* __restrict__ g,
                                            double* __restrict__ m, double* __restrict__ v,
                                            double* __restrict__ vmax, double* __restrict__ y,
                                            int method, double beta1, double beta2, double bc1,
                                            double bc2, double alpha, int nbf,
                                            double* __restrict__ Y) {
  const int j = e % nbf, i = e / nbf;
  if (j <= i) return;
  const long k = pair_index(i, j, nbf);
  const double gk = g[k];
  const double mk = beta1 * m[k] + (1.0 - beta1) * gk;
  double vk;
  if (method == DN_ADABELIEF) {
    const double d = gk - mk;
    vk = beta2 * v[k] + (1.0 - beta2) * (d * d);
  } else if (method == DN_YOGI) {
    const double g2 = gk * gk;
    const double dv = v[k] - g2;
    const double sg = (dv > 0.0) ? 1.0 : ((dv < 0.0) ? -1.0 : 0.0);
    vk = v[k] - (1.0 - beta2) * sg * g2;
  } else {
    vk = beta2 * v[k] + (1.0 - beta2) * (gk * gk);
  }
  const double mhat = mk / bc1;
  const double vhat = vk / bc2;
  const double vm = fmax(vmax[k], vhat);
  const double yk = -alpha * mhat / sqrt(vm + 1e-16);
  m[k] = mk;
  v[k] = vk;
  vmax[k] = vm;
  y[k] = yk;
  Y[i + (long)j * nbf] = yk;
  Y[j + (long)i * nbf] = -yk;
}

// sum_i |Y[i,j]| of one column by one wavefront (lanes stride the column, then a
// butterfly); every lane returns the sum.
__device__ __forceinline__ double column_abs_sum(const double* __restrict__ col, int nbf,
                                                 int lane) {
  double s = 0.0;
  for (int i = lane; i < nbf; i += 64) s += fabs(col[i]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  return s;
}

// Element e of c0 I + c1 X + c2 X2 + c3 X3 (+ c4 X4) for an nbf x nbf block.
__device__ __forceinline__ double poly_element(long e, int nbf, const double* __restrict__ X,
                                               const double* __restrict__ X2,
                                               const double* __restrict__ X3,
                                               const double* __restrict__ X4, double c0,
                                               double c1, double c2, double c3, double c4) {
  const int a = e % nbf, b = e / nbf;
  double v = c1 * X[e] + c2 * X2[e] + c3 * X3[e];
  if (X4) v += c4 * X4[e];
  if (a == b) v += c0;
  return v;
}

// One Newton-Schulz step towards the orthogonal factor of U: U' = U (3 - U^T U) / 2 =
// 1.5 U - 0.5 U (U^T U), element-wise from U and x = (U (U^T U))[e]; the error of
// U^T U = 1 + E becomes O(E^2).  An explicit fma, so that every kernel rounds alike.
__device__ __forceinline__ double ns_polish(double u, double x) { return fma(-0.5, x, 1.5 * u); }

}  // namespace dn
