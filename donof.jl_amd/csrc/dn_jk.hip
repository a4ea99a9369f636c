// J/K/H_core epilogues of the RI J/K-MO build (integrals.jl:144-172) and the device
// generator of the synthetic three-index tensor.
#include "dn_internal.h"

namespace dn {

// K[p,q] = sum_Q b_pqQ^2 over the Q range of split z (integrals.jl:166).  bpq is
// [p,q,Q] with leading dimension nbf5 and slice stride `slice` (nbf5^2 for the J/K
// build, nbf5*nbf when read out of the Lagrangian's b_MO^T); each thread owns one (p,q)
// and consecutive threads read consecutive addresses of a slice.
template <typename T>
__global__ void k_partial_kernel(const T* __restrict__ bpq, long slice, int nbf5, int naux,
                                 int q_per_split, double* __restrict__ Kp) {
  const long nn = (long)nbf5 * nbf5;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int z = blockIdx.y;
  if (e >= nn) return;
  const int Q0 = z * q_per_split;
  const int Q1 = min(Q0 + q_per_split, naux);
  double k = 0.0;
  for (int Q = Q0; Q < Q1; ++Q) {
    const double v = bpq[(long)Q * slice + e];
    k = fma(v, v, k);
  }
  Kp[z * nn + e] = k;
}

// X[p,Q] = b_ppQ (the diagonal that J = X X^T contracts, integrals.jl:163)
template <typename T>
__global__ void diag_kernel(const T* __restrict__ bpq, long slice, int nbf5, int naux,
                            double* __restrict__ X) {
  const long total = (long)nbf5 * naux;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long Q = e / nbf5;
    const int p = e - Q * nbf5;
    X[e] = bpq[Q * slice + (long)p * (nbf5 + 1)];
  }
}

template <typename T>
int jk_from_bpq_t(hipStream_t s, const T* bpq, long slice, int nbf5, int naux, double* X,
                  double* part, size_t part_elems, double* part2, size_t part2_elems, double* J,
                  double* K) {
  const long nn = (long)nbf5 * nbf5;
  const int threads = 256;
  const int bx = (int)((nn + threads - 1) / threads);
  int splits = (2048 + bx - 1) / bx;
  if (splits > naux) splits = naux;
  while (splits > 1 && (size_t)splits * nn > part_elems) --splits;
  const int qps = (naux + splits - 1) / splits;
  splits = (naux + qps - 1) / qps;
  hipLaunchKernelGGL(k_partial_kernel<T>, dim3(bx, splits), dim3(threads), 0, s, bpq, slice, nbf5,
                     naux, qps, part);
  DN_HIP(hipGetLastError());
  DN_TRY(reduce_splits(s, part, nn, splits, nbf5, nbf5, nbf5, K, nbf5, 0.0));
  const long nx = (long)nbf5 * naux;
  hipLaunchKernelGGL(diag_kernel<T>, dim3((unsigned)std::min<long>((nx + 255) / 256, 4096)),
                     dim3(256), 0, s, bpq, slice, nbf5, naux, X);
  DN_HIP(hipGetLastError());
  // J = X X^T: A(p,Q) = X[p + nbf5 Q], B(Q,q) = X[q + nbf5 Q]
  Gemm g;
  g.M = nbf5; g.N = nbf5; g.K = naux;
  g.A = X; g.lda = nbf5; g.a_kcontig = false;
  g.B = X; g.ldb = nbf5; g.b_kcontig = false;
  g.C = J; g.ldc = nbf5;
  return gemm_f64_splitk(s, g, part2, part2_elems);
}

int jk_from_bpq(hipStream_t s, const double* bpq, long slice, int nbf5, int naux, double* X,
                double* part, size_t part_elems, double* part2, size_t part2_elems, double* J,
                double* K) {
  return jk_from_bpq_t<double>(s, bpq, slice, nbf5, naux, X, part, part_elems, part2,
                               part2_elems, J, K);
}

int jk_from_bpq_f32(hipStream_t s, const float* bpq, long slice, int nbf5, int naux, double* X,
                    double* part, size_t part_elems, double* part2, size_t part2_elems,
                    double* J, double* K) {
  return jk_from_bpq_t<float>(s, bpq, slice, nbf5, naux, X, part, part_elems, part2,
                              part2_elems, J, K);
}

// H_core[i] = sum_m C5[m,i] (H C5)[m,i]   (integrals.jl:156,169)
__global__ void hcore_kernel(const double* __restrict__ C, const double* __restrict__ HC,
                             int nbf, int nbf5, double* __restrict__ Hcore) {
  const int i = blockIdx.x;
  double s = 0.0;
  for (int m = threadIdx.x; m < nbf; m += blockDim.x) s = fma(C[m + (long)i * nbf], HC[m + (long)i * nbf], s);
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) Hcore[i] = red[0];
}

int hcore_diag(hipStream_t s, const double* C, const double* HC, int nbf, int nbf5,
               double* Hcore) {
  hipLaunchKernelGGL(hcore_kernel, dim3(nbf5), dim3(256), 0, s, C, HC, nbf, nbf5, Hcore);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// ---- Philox4x32-10 normal stream (same definition as the host generator) -----------
__device__ __forceinline__ double philox_normal(uint64_t e, uint32_t stream, uint64_t seed) {
  const uint64_t ctr = e >> 1;
  uint32_t c0 = (uint32_t)ctr, c1 = (uint32_t)(ctr >> 32), c2 = stream, c3 = 0x0D0Fu;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  const double s53 = 1.1102230246251565e-16;  // 2^-53
  const double u = ((double)(c0 >> 5) * 67108864.0 + (double)(c1 >> 6) + 0.5) * s53;
  const double v = ((double)(c2 >> 5) * 67108864.0 + (double)(c3 >> 6) + 0.5) * s53;
  const double r = sqrt(-2.0 * log(u));
  const double ang = 6.283185307179586 * v;
  return (e & 1) ? r * sin(ang) : r * cos(ang);
}

// b[m,n,Q] = (A_Q[m,n] + A_Q[n,m]) exp(-(Q+1)/400) / (2 sqrt(nbf)),
// A_Q[m,n] = z(m + nbf n + nbf^2 Q) of stream 0.
__global__ void synth_b_kernel(double* __restrict__ b, int nbf, int naux, uint64_t seed) {
  const uint64_t nb2 = (uint64_t)nbf * nbf;
  const uint64_t total = nb2 * (uint64_t)naux;
  const double inv = 0.5 / sqrt((double)nbf);
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t Q = e / nb2;
    const uint64_t r = e - Q * nb2;
    const uint64_t m = r % nbf, n = r / nbf;
    const double a1 = philox_normal(m + nbf * n + nb2 * Q, 0u, seed);
    const double a2 = philox_normal(n + nbf * m + nb2 * Q, 0u, seed);
    b[e] = (a1 + a2) * (exp(-(double)(Q + 1) / 400.0) * inv);
  }
}

int synth_b(hipStream_t s, double* b, int nbf, int naux, uint64_t seed) {
  hipLaunchKernelGGL(synth_b_kernel, dim3(16384), dim3(256), 0, s, b, nbf, naux, seed);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

}  // namespace dn
