// J/K/H_core epilogues of the RI J/K-MO build (integrals.jl:144-172) and the device
// generator of the synthetic three-index tensor.
#include <cstring>
#include <type_traits>

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

int j_from_x(hipStream_t s, const double* X, int nbf5, int naux, double* part2,
             size_t part2_elems, double* J);

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
  return j_from_x(s, X, nbf5, naux, part2, part2_elems, J);
}

// J = X X^T (integrals.jl:163): A(p,Q) = X[p + nbf5 Q], B(Q,q) = X[q + nbf5 Q]
int j_from_x(hipStream_t s, const double* X, int nbf5, int naux, double* part2,
             size_t part2_elems, double* J) {
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

// J/K of the symmetric-b build straight from W (b_pq = W + W^T is never stored):
//   K[p,q] = sum_Q (W[p,q,Q] + W[q,p,Q])^2,   X[p,Q] = b_ppQ = 2 W[p,p,Q].
// Pass 1: workgroup (system, 32 x 32 tile pair (I, J) with I <= J, range z of Q) stages
// the tiles (I, J) and (J, I) of each slice in LDS and accumulates the squares of tile
// (I, J) in registers; every element of W is read once, coalesced.  Diagonal tiles also
// write X.  Pass 2 sums the ranges and mirrors the lower tiles (K is symmetric).
__device__ __forceinline__ int find_symk(const SymKDesc* __restrict__ d, int n, long blk,
                                         bool red) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((red ? d[mid].red_begin : d[mid].blk_begin) <= blk) lo = mid;
    else hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);  // uniform: scalar descriptor loads
}

// VEC (FP32 W with nbf5 % 4 == 0): each thread moves four consecutive rows p of one column
// q of each tile with one 16-byte load, 32 x 32 tiles in 256 loads, instead of four 4-byte
// loads per thread and tile.
template <typename T, bool VEC = false>
__global__ __launch_bounds__(256) void symk_partial_kernel(const SymKDesc* __restrict__ desc,
                                                           int n) {
  __shared__ double a[32][33], b[32][33];
  const SymKDesc d = desc[find_symk(desc, n, blockIdx.x, false)];
  const T* Wall = std::is_same<T, float>::value ? reinterpret_cast<const T*>(d.Wf)
                                                : reinterpret_cast<const T*>(d.W);
  const int n5 = d.nbf5;
  const long nn = (long)n5 * n5;
  const int ntile = (n5 + 31) / 32;
  const int pairs = ntile * (ntile + 1) / 2;
  const long local = blockIdx.x - d.blk_begin;
  int pidx = (int)(local % pairs), I = 0;
  while (pidx >= ntile - I) {
    pidx -= ntile - I;
    ++I;
  }
  const int J = I + pidx;
  const int z = (int)(local / pairs);
  const int Q0 = z * d.qps, Q1 = min(Q0 + d.qps, d.naux);
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int pI = I * 32 + tx, pJ = J * 32 + tx;
  // a diagonal tile pair is one tile: (J, I) = (I, I) is read once and mirrored in LDS
  const bool diag = I == J;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // the next slice's elements are loaded into registers while this slice is reduced
  double ra[4], rb[4];
  // VEC: rows p4 .. p4 + 3 (local) of local column c4
  const int p4 = 4 * (threadIdx.x & 7), c4 = threadIdx.x >> 3;
  auto load = [&](int Q) {
    const T* W = Wall + (long)Q * nn;
    if constexpr (VEC) {
      typedef float v4 __attribute__((ext_vector_type(4)));
      const int pa = I * 32 + p4, qa = J * 32 + c4;  // tile (I, J)
      const int pb = J * 32 + p4, qb = I * 32 + c4;  // tile (J, I)
      v4 va = {0.f, 0.f, 0.f, 0.f}, vb = {0.f, 0.f, 0.f, 0.f};
      if (pa < n5 && qa < n5) va = *reinterpret_cast<const v4*>(W + pa + (long)qa * n5);
      if (!diag && pb < n5 && qb < n5) vb = *reinterpret_cast<const v4*>(W + pb + (long)qb * n5);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        ra[e] = (double)va[e];
        rb[e] = (double)vb[e];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int r = ty + 8 * k, qJ = J * 32 + r, qI = I * 32 + r;
        ra[k] = (pI < n5 && qJ < n5) ? (double)W[pI + (long)qJ * n5] : 0.0;  // tile (I, J)
        rb[k] = (!diag && pJ < n5 && qI < n5) ? (double)W[pJ + (long)qI * n5] : 0.0;  // (J, I)
      }
    }
  };
  if (Q0 < Q1) load(Q0);
  for (int Q = Q0; Q < Q1; ++Q) {
    if constexpr (VEC) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        a[p4 + e][c4] = ra[e];
        if (!diag) b[p4 + e][c4] = rb[e];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        a[tx][ty + 8 * k] = ra[k];
        if (!diag) b[tx][ty + 8 * k] = rb[k];
      }
    }
    if (Q + 1 < Q1) load(Q + 1);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int r = ty + 8 * k;
      const double v = a[tx][r] + (diag ? a[r][tx] : b[r][tx]);  // b_pq, p = I*32+tx, q = J*32+r
      acc[k] = fma(v, v, acc[k]);
      if (diag && r == tx && pI < n5) d.X[pI + (long)Q * n5] = v;
    }
    __syncthreads();
  }
  double* part = d.part + (long)z * nn;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int qJ = J * 32 + ty + 8 * k;
    if (pI < n5 && qJ < n5) part[pI + (long)qJ * n5] = acc[k];
  }
}

// 64 elements per workgroup, the eight wavefronts summing interleaved Q ranges (two
// accumulators each, so that two loads are in flight per lane).
__global__ __launch_bounds__(512) void symk_reduce_kernel(const SymKDesc* __restrict__ desc,
                                                          int n) {
  __shared__ double red[8][64];
  const SymKDesc d = desc[find_symk(desc, n, blockIdx.x, true)];
  const int n5 = d.nbf5;
  const long nn = (long)n5 * n5;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long e = (blockIdx.x - d.red_begin) * 64 + lane;
  double k0 = 0.0, k1 = 0.0;
  if (e < nn) {
    const int p = (int)(e % n5), q = (int)(e / n5);
    // partials exist for tile(p) <= tile(q); K is symmetric
    const long src = (p / 32 <= q / 32) ? e : q + (long)p * n5;
    int z = w;
    for (; z + 8 < d.splits; z += 16) {
      k0 += d.part[z * nn + src];
      k1 += d.part[(z + 8) * nn + src];
    }
    if (z < d.splits) k0 += d.part[z * nn + src];
  }
  red[w][lane] = k0 + k1;
  __syncthreads();
  if (w == 0 && e < nn)
    d.K[e] = ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane])) +
             ((red[4][lane] + red[5][lane]) + (red[6][lane] + red[7][lane]));
}

void symk_plan(SymKDesc* desc, int n, long* blocks, long* red_blocks, bool f32) {
  long b = 0, r = 0;
  for (int i = 0; i < n; ++i) {
    SymKDesc& d = desc[i];
    const long nn = (long)d.nbf5 * d.nbf5;
    const int ntile = (d.nbf5 + 31) / 32;
    const int pairs = ntile * (ntile + 1) / 2;
    // Q ranges of 8 slabs or more.  FP64 W: the longest ranges that leave >= 360
    // workgroups (each loads the next slice while it reduces this one): 120/1600 41.6 us
    // with 32 slabs per range against 55.8 us with 8, 240/3200 293 against 478 us.  FP32
    // W: doubled only until <= 2048 workgroups (240/3200 with 4-byte loads: 317 us with
    // 64 slabs, 528 us with 128; with 16-byte loads 4096 workgroups doubled the partials'
    // reduction for 11 us less in the partial pass).
    d.qps = 8;
    d.splits = (d.naux + d.qps - 1) / d.qps;
    if (f32) {
      while ((long)pairs * d.splits > 2048 && d.qps < d.naux) {
        d.qps *= 2;
        d.splits = (d.naux + d.qps - 1) / d.qps;
      }
    } else {
      while (d.qps < d.naux && (long)pairs * ((d.naux + 2 * d.qps - 1) / (2 * d.qps)) >= 360) {
        d.qps *= 2;
        d.splits = (d.naux + d.qps - 1) / d.qps;
      }
    }
    while ((size_t)d.splits * nn > d.part_elems) {
      d.qps *= 2;
      d.splits = (d.naux + d.qps - 1) / d.qps;
    }
    d.blk_begin = b;
    d.red_begin = r;
    b += (long)pairs * d.splits;
    r += (nn + 63) / 64;
  }
  *blocks = b;
  *red_blocks = r;
}

int symk(hipStream_t s, const SymKDesc* dev, int n, long blocks, long red_blocks, bool f32,
         bool vec4) {
  if (n <= 0) return DN_OK;
  if (f32 && vec4)
    hipLaunchKernelGGL((symk_partial_kernel<float, true>), dim3((unsigned)blocks), dim3(256), 0, s,
                       dev, n);
  else if (f32)
    hipLaunchKernelGGL(symk_partial_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, s, dev, n);
  else
    hipLaunchKernelGGL(symk_partial_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, s, dev, n);
  hipLaunchKernelGGL(symk_reduce_kernel, dim3((unsigned)red_blocks), dim3(512), 0, s, dev, n);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// out[0] = max |b[m,n,Q] - b[n,m,Q]|, out[1] = max |b| (as the bit patterns of
// non-negative doubles, which order like the values).
__global__ __launch_bounds__(256) void sym_check_kernel(const double* __restrict__ b, int nbf,
                                                        int ld, int naux, int ntile,
                                                        int q_per_block,
                                                        unsigned long long* __restrict__ out) {
  // workgroup: 32 x 32 tile pair (I, J), I <= J, of a range of slabs; both tiles staged in
  // LDS so that b[m,n] and b[n,m] are each read once, coalesced
  __shared__ double a[32][33], t[32][33];
  int pidx = blockIdx.x, I = 0;
  while (pidx >= ntile - I) {
    pidx -= ntile - I;
    ++I;
  }
  const int J = I + pidx;
  const long nn = (long)ld * nbf;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int mI = I * 32 + tx, mJ = J * 32 + tx;
  double dmax = 0.0, bmax = 0.0;
  const int Q0 = blockIdx.y * q_per_block, Q1 = min(Q0 + q_per_block, naux);
  for (int Q = Q0; Q < Q1; ++Q) {
    const double* s = b + (long)Q * nn;
    for (int r = ty; r < 32; r += 8) {
      const int nJ = J * 32 + r, nI = I * 32 + r;
      a[tx][r] = (mI < nbf && nJ < nbf) ? s[mI + (long)nJ * ld] : 0.0;  // tile (I, J)
      t[tx][r] = (mJ < nbf && nI < nbf) ? s[mJ + (long)nI * ld] : 0.0;  // tile (J, I)
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
      dmax = fmax(dmax, fabs(a[tx][r] - t[r][tx]));
      bmax = fmax(bmax, fmax(fabs(a[tx][r]), fabs(t[tx][r])));
    }
    __syncthreads();
  }
  for (int off = 32; off > 0; off >>= 1) {
    dmax = fmax(dmax, __shfl_down(dmax, off, 64));
    bmax = fmax(bmax, __shfl_down(bmax, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(out, (unsigned long long)__double_as_longlong(dmax));
    atomicMax(out + 1, (unsigned long long)__double_as_longlong(bmax));
  }
}

namespace {
__global__ void clear_u64_kernel(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}
}  // namespace

int b_symmetry(hipStream_t s, const double* b, int nbf, int ld, int naux,
               unsigned long long* dev2, double* dmax, double* bmax) {
  hipLaunchKernelGGL(clear_u64_kernel, dim3(1), dim3(64), 0, s, dev2, 2);
  const int ntile = (nbf + 31) / 32;
  const int pairs = ntile * (ntile + 1) / 2;
  const int qpb = (int)std::max<long>(1, (long)naux * pairs / 8192);
  const int nq = (naux + qpb - 1) / qpb;
  hipLaunchKernelGGL(sym_check_kernel, dim3(pairs, nq), dim3(256), 0, s, b, nbf, ld, naux, ntile,
                     qpb, dev2);
  DN_HIP(hipGetLastError());
  unsigned long long h[2] = {0, 0};
  DN_HIP(hipMemcpyAsync(h, dev2, sizeof(h), hipMemcpyDeviceToHost, s));
  DN_HIP(hipStreamSynchronize(s));
  std::memcpy(dmax, &h[0], sizeof(double));
  std::memcpy(bmax, &h[1], sizeof(double));
  return DN_OK;
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

// H_core[p] = Hmat[p,p] (Hmat = C^T H C5, nbf x nbf5): the diagonal of its leading
// nbf5 block is (C5^T H C5)_pp (integrals.jl:156,169)
__global__ void hcore_from_hmat_kernel(const double* __restrict__ Hmat, int nbf, int nbf5,
                                       double* __restrict__ Hcore) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nbf5) Hcore[p] = Hmat[(long)p * nbf + p];
}

int hcore_from_hmat(hipStream_t s, const double* Hmat, int nbf, int nbf5, double* Hcore) {
  hipLaunchKernelGGL(hcore_from_hmat_kernel, dim3((nbf5 + 255) / 256), dim3(256), 0, s, Hmat,
                     nbf, nbf5, Hcore);
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
// A_Q[m,n] = z(m + nbf n + nbf^2 Q) of stream 0; stored at m + ld (n + nbf Q).
__global__ void synth_b_kernel(double* __restrict__ b, int nbf, int ld, int naux, uint64_t seed) {
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
    b[m + (uint64_t)ld * (n + (uint64_t)nbf * Q)] =
        (a1 + a2) * (exp(-(double)(Q + 1) / 400.0) * inv);
  }
}

int synth_b(hipStream_t s, double* b, int nbf, int ld, int naux, uint64_t seed) {
  hipLaunchKernelGGL(synth_b_kernel, dim3(16384), dim3(256), 0, s, b, nbf, ld, naux, seed);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// Cp[m,n] = C[m,n] for m < nbf, 0 for the pad row m = nbf (ld = nbf + 1)
__device__ __forceinline__ double pad_element(long e, const double* __restrict__ C, int nbf,
                                              int ld) {
  const int m = (int)(e % ld);
  const long n = e / ld;
  return m < nbf ? C[m + (long)nbf * n] : 0.0;
}

__global__ void pad_c_kernel(const double* __restrict__ C, double* __restrict__ Cp, int nbf,
                             int ld) {
  const long total = (long)ld * nbf;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x)
    Cp[e] = pad_element(e, C, nbf, ld);
}

__global__ __launch_bounds__(256) void pad_c_group_kernel(const PadDesc* __restrict__ d, int n) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].begin <= (long)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const PadDesc& r = d[__builtin_amdgcn_readfirstlane(lo)];
  const long e = ((long)blockIdx.x - r.begin) * 256 + threadIdx.x;
  if (e < (long)r.ld * r.nbf) r.Cp[e] = pad_element(e, r.C, r.nbf, r.ld);
}

int pad_c_launch(hipStream_t s, const double* C, double* Cp, int nbf, int ld) {
  const long blocks = ((long)ld * nbf + 255) / 256;
  hipLaunchKernelGGL(pad_c_kernel, dim3((unsigned)std::min(blocks, 4096L)), dim3(256), 0, s, C,
                     Cp, nbf, ld);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

long pad_plan(PadDesc* d, int n) {
  long blocks = 0;
  for (int i = 0; i < n; ++i) {
    d[i].begin = blocks;
    blocks += ((long)d[i].ld * d[i].nbf + 255) / 256;
  }
  return blocks;
}

int pad_c_group(hipStream_t s, const PadDesc* d, int n, long blocks) {
  if (n <= 0) return DN_OK;
  hipLaunchKernelGGL(pad_c_group_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d, n);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

}  // namespace dn
