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
//
// Only bt's first nbf5 columns enter bmat and w: b_pq[q,p,Q] = (C5^T b_Q C5)[q,p], stored at
// p + nbf5 (q + nbf5 Q) by the flattened second transform; bmat[q,b,Q] is kept at the same
// index (an element-wise map, read and written along b).  The contraction over (q, Q) then
// runs against the first transform's output instead of the full b_MO (dn_api.hip):
//   Z[n,b] = sum_{q,Q} bpn[n,q,Q] bmat[q,b,Q],   elag[a,b] += sum_n C[n,a] Z[n,b].
#include "dn_internal.h"
#include "dn_rot_dev.h"

namespace dn {

// ---- per-block / per-element bodies (single-system and grouped kernels alike) -------------

// bt[q,a,Q] (a < nbf5 here): at a + nbf q + slice Q, or, TR (the split second transform's
// FP32 output, b_pq stored transposed), at q + nbf5 a + slice Q
template <typename T, bool TR>
__device__ __forceinline__ double bt_at(const T* __restrict__ bt, long slice, int nbf, int nbf5,
                                        int a, int q, long Q) {
  return TR ? (double)bt[q + (long)nbf5 * a + slice * Q] : (double)bt[a + (long)nbf * q + slice * Q];
}

// X[:,Q] and w[:,Q] for all b < nbf5 (one workgroup per Q); xs: nbf5 doubles of LDS.
// bt[q,q,Q] is at q (nbf + 1) + slice Q (q (nbf5 + 1) + slice Q for TR).
template <typename T = double, bool TR = false>
__device__ __forceinline__ void lag_w_block(int Q, const T* __restrict__ bt, long slice,
                                            int nbf, int nbf5, const double* __restrict__ cj12,
                                            const double* __restrict__ n, int nalpha, int nbeta,
                                            double* __restrict__ X, double* __restrict__ w,
                                            double* xs) {
  for (int q = threadIdx.x; q < nbf5; q += blockDim.x) {
    const double v = bt_at<T, TR>(bt, slice, nbf, nbf5, q, q, Q);
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

// bmat[q,b,Q] = -ck12[b,q] bt[q,b,Q] + [q==b] w[b,Q] for element e = b + nbf5 q + nbf5^2 Q
template <typename T = double, bool TR = false>
__device__ __forceinline__ double lag_bmat_element(long e, const T* __restrict__ bt,
                                                   long slice, int nbf, int nbf5,
                                                   const double* __restrict__ ck12,
                                                   const double* __restrict__ w) {
  const long nn = (long)nbf5 * nbf5;
  const long Q = e / nn;
  const int r = (int)(e - Q * nn);
  const int b = r % nbf5, q = r / nbf5;
  double v = -ck12[r] * bt_at<T, TR>(bt, slice, nbf, nbf5, b, q, Q);
  if (q == b) v += w[b + (long)nbf5 * Q];
  return v;
}

// elag[a,b] = n[b] Hmat[a,b] for b < nbf5, 0 for b >= nbf5 (utils.jl:22,39)
__device__ __forceinline__ double lag_init_element(long e, const double* __restrict__ Hmat,
                                                   const double* __restrict__ n, int nbf,
                                                   int nbf5) {
  const int a = e % nbf, b = e / nbf;
  return (b < nbf5) ? n[b] * Hmat[a + (long)b * nbf] : 0.0;
}

// E = sum_{i<nbf5} elag_ii + sum_i n_i Hmat_ii (utils.jl:216-227) by one 256-thread
// workgroup; red: 256 doubles of LDS.
__device__ __forceinline__ void energy_block(const double* __restrict__ elag,
                                             const double* __restrict__ Hmat,
                                             const double* __restrict__ n, int nbf, int nbf5,
                                             double* __restrict__ out, double* red) {
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

// ---- single-system kernels --------------------------------------------------------------

template <typename T = double, bool TR = false>
__global__ void lag_w_kernel(const T* __restrict__ bt, long slice, int nbf, int nbf5,
                             int naux, const double* __restrict__ cj12,
                             const double* __restrict__ n, int nalpha, int nbeta,
                             double* __restrict__ X, double* __restrict__ w) {
  extern __shared__ double xs[];  // nbf5 diagonal entries of slice Q
  const int Q = blockIdx.x;
  if (Q >= naux) return;
  lag_w_block<T, TR>(Q, bt, slice, nbf, nbf5, cj12, n, nalpha, nbeta, X, w, xs);
}

__global__ __launch_bounds__(256) void lag_bmat_kernel(const double* __restrict__ bt, long slice,
                                                       int nbf, int nbf5, int naux,
                                                       const double* __restrict__ ck12,
                                                       const double* __restrict__ w,
                                                       double* __restrict__ bmat) {
  const long total = (long)nbf5 * nbf5 * naux;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x)
    bmat[e] = lag_bmat_element(e, bt, slice, nbf, nbf5, ck12, w);
}

__global__ void lag_init_kernel(const double* __restrict__ Hmat, const double* __restrict__ n,
                                int nbf, int nbf5, double* __restrict__ elag) {
  const long total = (long)nbf * nbf;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x)
    elag[e] = lag_init_element(e, Hmat, n, nbf, nbf5);
}

__global__ void energy_kernel(const double* __restrict__ elag, const double* __restrict__ Hmat,
                              const double* __restrict__ n, int nbf, int nbf5,
                              double* __restrict__ out) {
  __shared__ double red[256];
  energy_block(elag, Hmat, n, nbf, nbf5, out, red);
}

// ---- grouped kernels: every system of a lock-step group in one launch ------------------

__device__ __forceinline__ int find_lag(const LagSys* __restrict__ d, int n, long b,
                                        long LagSys::*begin) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].*begin <= b) lo = mid;
    else hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

__global__ __launch_bounds__(128) void group_lag_w_kernel(const LagSys* __restrict__ d, int n) {
  extern __shared__ double xs[];
  const LagSys& r = d[find_lag(d, n, blockIdx.x, &LagSys::w_begin)];
  lag_w_block((int)(blockIdx.x - r.w_begin), r.bt, r.slice, r.ldt, r.nbf5, r.cj12, r.n,
              r.nalpha, r.nbeta, r.X, r.w, xs);
}

__global__ __launch_bounds__(256) void group_lag_bmat_kernel(const LagSys* __restrict__ d, int n) {
  const LagSys& r = d[find_lag(d, n, blockIdx.x, &LagSys::bmat_begin)];
  const long e = (blockIdx.x - r.bmat_begin) * 256 + threadIdx.x;
  if (e < (long)r.nbf5 * r.nbf5 * r.naux)
    r.bmat[e] = lag_bmat_element(e, r.bt, r.slice, r.ldt, r.nbf5, r.ck12, r.w);
}

__global__ __launch_bounds__(256) void group_lag_init_kernel(const LagSys* __restrict__ d, int n) {
  const LagSys& r = d[find_lag(d, n, blockIdx.x, &LagSys::sq_begin)];
  const long e = (blockIdx.x - r.sq_begin) * 256 + threadIdx.x;
  if (e < (long)r.nbf * r.nbf) r.elag[e] = lag_init_element(e, r.Hmat, r.n, r.nbf, r.nbf5);
}

// Z = the split partials of the C^J/C^K contraction, summed in split order (the
// arithmetic of reduce_splits with scale 1 and beta 0)
__global__ __launch_bounds__(256) void group_lag_reduce_kernel(const LagSys* __restrict__ d,
                                                               int n) {
  const LagSys& r = d[find_lag(d, n, blockIdx.x, &LagSys::rect_begin)];
  const long e = (blockIdx.x - r.rect_begin) * 256 + threadIdx.x;
  const long sP = (long)r.nbf * r.nbf5;
  if (e >= sP) return;
  double s = 0.0;
  for (int z = 0; z < r.nsplit; ++z) s += r.part[z * sP + e];
  s *= 1.0;
  r.Z[e] = s;
}

__global__ __launch_bounds__(256) void group_energy_kernel(const LagSys* __restrict__ d) {
  __shared__ double red[256];
  const LagSys& r = d[blockIdx.x];
  energy_block(r.elag, r.Hmat, r.n, r.nbf, r.nbf5, r.scal, red);
}

__global__ void group_zero_gmax_kernel(const LagSys* __restrict__ d, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) *reinterpret_cast<unsigned long long*>(d[i].scal + 1) = 0ull;
}

// g = pack(4 elag - 4 elag^T) and max|g| (the element map of pack_grad_kernel)
__global__ __launch_bounds__(256) void group_pack_grad_kernel(const LagSys* __restrict__ d, int n) {
  const LagSys& r = d[find_lag(d, n, blockIdx.x, &LagSys::rect_begin)];
  const long e = (blockIdx.x - r.rect_begin) * 256 + threadIdx.x;
  double mx = 0.0;
  if (e < (long)r.nbf5 * r.nbf) {
    const int j = e % r.nbf, i = e / r.nbf;
    if (j > i) {
      const double v = 4.0 * r.elag[i + (long)j * r.nbf] - 4.0 * r.elag[j + (long)i * r.nbf];
      r.g[pair_index(i, j, r.nbf)] = v;
      mx = fmax(mx, fabs(v));
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0)
    atomicMax(reinterpret_cast<unsigned long long*>(r.scal + 1),
              (unsigned long long)__double_as_longlong(mx));
}

int lag_w(hipStream_t s, const double* bt, long slice, int nbf, int nbf5, int naux,
          const double* cj12, const double* n, int nalpha, int nbeta, double* X, double* w) {
  hipLaunchKernelGGL((lag_w_kernel<double, false>), dim3(naux), dim3(128), nbf5 * sizeof(double),
                     s, bt, slice, nbf, nbf5, naux, cj12, n, nalpha, nbeta, X, w);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int lag_bmat(hipStream_t s, const double* bt, long slice, int nbf, int nbf5, int naux,
             const double* ck12, const double* w, double* bmat) {
  const long blocks = ((long)nbf5 * nbf5 * naux + 255) / 256;
  hipLaunchKernelGGL(lag_bmat_kernel, dim3((unsigned)std::min(blocks, 8192L)), dim3(256), 0, s,
                     bt, slice, nbf, nbf5, naux, ck12, w, bmat);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// X and w of the split path straight from the second transform's FP32 output (b_pq
// stored transposed, slice nbf5^2)
int lag_w_split(hipStream_t s, const float* bpq, int nbf5, int naux, const double* cj12,
                const double* n, int nalpha, int nbeta, double* X, double* w) {
  hipLaunchKernelGGL((lag_w_kernel<float, true>), dim3(naux), dim3(128), nbf5 * sizeof(double), s,
                     bpq, (long)nbf5 * nbf5, nbf5, nbf5, naux, cj12, n, nalpha, nbeta, X, w);
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

void lag_group_plan(LagSys* d, int n, LagBlocks* b) {
  long w = 0, bm = 0, sq = 0, rect = 0;
  int maxq = 1;
  for (int i = 0; i < n; ++i) {
    LagSys& r = d[i];
    r.w_begin = w;
    r.bmat_begin = bm;
    r.sq_begin = sq;
    r.rect_begin = rect;
    w += r.naux;
    bm += ((long)r.nbf5 * r.nbf5 * r.naux + 255) / 256;
    sq += ((long)r.nbf * r.nbf + 255) / 256;
    rect += ((long)r.nbf * r.nbf5 + 255) / 256;
    maxq = std::max(maxq, r.nbf5);
  }
  b->w = w;
  b->bmat = bm;
  b->sq = sq;
  b->rect = rect;
  b->lds_w = (size_t)maxq * sizeof(double);
}

int lag_group_before(hipStream_t s, const LagSys* d, int n, const LagBlocks& b) {
  hipLaunchKernelGGL(group_lag_w_kernel, dim3((unsigned)b.w), dim3(128), b.lds_w, s, d, n);
  hipLaunchKernelGGL(group_lag_bmat_kernel, dim3((unsigned)b.bmat), dim3(256), 0, s, d, n);
  hipLaunchKernelGGL(group_lag_init_kernel, dim3((unsigned)b.sq), dim3(256), 0, s, d, n);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int lag_group_reduce(hipStream_t s, const LagSys* d, int n, const LagBlocks& b) {
  hipLaunchKernelGGL(group_lag_reduce_kernel, dim3((unsigned)b.rect), dim3(256), 0, s, d, n);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int lag_group_after(hipStream_t s, const LagSys* d, int n, const LagBlocks& b) {
  hipLaunchKernelGGL(group_energy_kernel, dim3(n), dim3(256), 0, s, d);
  hipLaunchKernelGGL(group_zero_gmax_kernel, dim3((n + 63) / 64), dim3(64), 0, s, d, n);
  hipLaunchKernelGGL(group_pack_grad_kernel, dim3((unsigned)b.rect), dim3(256), 0, s, d, n);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

}  // namespace dn
