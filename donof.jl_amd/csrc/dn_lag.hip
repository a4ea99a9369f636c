// Orbital Lagrangian of the RI functional (utils.jl:1-59) and the electronic energy
// (utils.jl:216-227).
//
// With bt[q,a,Q] = b_MO[a,q,Q] (q < nbf5, a < nbf) the three two-electron terms of
// elag[a,b] (b < nbf5) are one contraction over (q, Q):
//   elag[a,b] += sum_Q sum_q bt[q,a,Q] * bmat[q,b,Q]
//   bmat[q,b,Q] = -ck12[b,q] bt[q,b,Q] + [q == b] w[b,Q]
//   w[b,Q]      = sum_q cj12[b,q] X[q,Q] + coul(b) n[b] X[b,Q],  X[q,Q] = bt[q,q,Q]
// where coul(b) marks the Coulomb self-terms of utils.jl:42-47 (nbeta > 0 and b < nbeta
// or b >= nalpha).  The first line is the C^K term (54-55), w carries C^J (50-51) and the
// Coulomb terms; the one-electron term n[b] Hmat[a,b] (39) is added afterwards.
#include "dn_internal.h"

namespace dn {

// w[b,Q] for all b < nbf5; one block per Q chunk, X read from bt's diagonal.
__global__ void lag_w_kernel(const double* __restrict__ bt, long slice, int nbf5, int naux,
                             const double* __restrict__ cj12, const double* __restrict__ n,
                             int nalpha, int nbeta, double* __restrict__ X,
                             double* __restrict__ w) {
  extern __shared__ double xs[];  // nbf5 diagonal entries of slice Q
  const int Q = blockIdx.x;
  if (Q >= naux) return;
  const double* s = bt + (long)Q * slice;
  for (int q = threadIdx.x; q < nbf5; q += blockDim.x) {
    const double v = s[(long)q * (nbf5 + 1)];
    xs[q] = v;
    X[q + (long)Q * nbf5] = v;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbf5; b += blockDim.x) {
    double acc = 0.0;
    for (int q = 0; q < nbf5; ++q) acc = fma(cj12[b + (long)q * nbf5], xs[q], acc);
    const bool coul = (nbeta > 0) && (b < nbeta || b >= nalpha);
    if (coul) acc = fma(n[b], xs[b], acc);
    w[b + (long)Q * nbf5] = acc;
  }
}

// bmat[q,b,Q] = -ck12[b,q] bt[q,b,Q] + [q==b] w[b,Q]
// Block (tq, tb, zq): the 32 x 32 (q, b) tile at (32 tq, 32 tb) for Q in
// [BMAT_Q zq, BMAT_Q (zq + 1)).  The ck12 tile is read once, along b, and kept transposed
// in LDS, so that bt and bmat are read and written along q (coalesced).
constexpr int BMAT_Q = 16;
__global__ __launch_bounds__(256) void lag_bmat_kernel(const double* __restrict__ bt, long slice,
                                                       int nbf5, int naux,
                                                       const double* __restrict__ ck12,
                                                       const double* __restrict__ w,
                                                       double* __restrict__ bmat) {
  __shared__ double ct[32][33];  // ct[q][b] = ck12[b, q]
  const int q0 = 32 * blockIdx.x, b0 = 32 * blockIdx.y;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int i = ty; i < 32; i += 8) {  // row q0 + i of the tile: ck12[b0 + tx, q0 + i]
    const int q = q0 + i, b = b0 + tx;
    ct[i][tx] = (q < nbf5 && b < nbf5) ? ck12[b + (long)q * nbf5] : 0.0;
  }
  __syncthreads();
  const long nn = (long)nbf5 * nbf5;
  const int q = q0 + tx;
  if (q >= nbf5) return;
  const int Q1 = min(naux, BMAT_Q * ((int)blockIdx.z + 1));
  for (int Q = BMAT_Q * blockIdx.z; Q < Q1; ++Q) {
    const double* s = bt + (long)Q * slice;
    double* o = bmat + (long)Q * nn;
    for (int i = ty; i < 32; i += 8) {
      const int b = b0 + i;
      if (b >= nbf5) break;
      double v = -ct[tx][i] * s[q + (long)b * nbf5];
      if (q == b) v += w[b + (long)Q * nbf5];
      o[q + (long)b * nbf5] = v;
    }
  }
}

// elag[a,b] = n[b] Hmat[a,b] for b < nbf5, 0 for b >= nbf5 (utils.jl:22,39)
__global__ void lag_init_kernel(const double* __restrict__ Hmat, const double* __restrict__ n,
                                int nbf, int nbf5, double* __restrict__ elag) {
  const long total = (long)nbf * nbf;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int a = e % nbf, b = e / nbf;
    elag[e] = (b < nbf5) ? n[b] * Hmat[a + (long)b * nbf] : 0.0;
  }
}

// E = sum_{i<nbf5} elag_ii + sum_i n_i Hmat_ii  (utils.jl:216-227); single block.
__global__ void energy_kernel(const double* __restrict__ elag, const double* __restrict__ Hmat,
                              const double* __restrict__ n, int nbf, int nbf5,
                              double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nbf5; i += blockDim.x)
    s += elag[i + (long)i * nbf] + n[i] * Hmat[i + (long)i * nbf];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

int lag_w(hipStream_t s, const double* bt, long slice, int nbf5, int naux, const double* cj12,
          const double* n, int nalpha, int nbeta, double* X, double* w) {
  hipLaunchKernelGGL(lag_w_kernel, dim3(naux), dim3(128), nbf5 * sizeof(double), s, bt, slice,
                     nbf5, naux, cj12, n, nalpha, nbeta, X, w);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int lag_bmat(hipStream_t s, const double* bt, long slice, int nbf5, int naux,
             const double* ck12, const double* w, double* bmat) {
  const unsigned t = (nbf5 + 31) / 32;
  hipLaunchKernelGGL(lag_bmat_kernel, dim3(t, t, (naux + BMAT_Q - 1) / BMAT_Q), dim3(256), 0, s,
                     bt, slice, nbf5, naux, ck12, w, bmat);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int lag_init(hipStream_t s, const double* Hmat, const double* n, int nbf, int nbf5,
             double* elag) {
  hipLaunchKernelGGL(lag_init_kernel, dim3(1024), dim3(256), 0, s, Hmat, n, nbf, nbf5, elag);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int energy_elec(hipStream_t s, const double* elag, const double* Hmat, const double* n,
                int nbf, int nbf5, double* out) {
  hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(256), 0, s, elag, Hmat, n, nbf, nbf5, out);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

}  // namespace dn
