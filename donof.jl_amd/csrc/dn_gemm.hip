// FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64), LDS double-buffered,
// with strided batches, batch-reduce splits and an XCD-aware tile order.
//
// Every dense contraction of the hot path is one of these calls: the first RI
// transform b_pn = C5^T b (one GEMM with N = nbf*naux), the per-Q second transform,
// the Lagrangian's C^K/C^J contraction (batch-reduce over Q), H C5, C^T H C5 and the
// matrix products of the orbital-rotation exponential.
#include "dn_internal.h"

namespace dn {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int kThreads = 256;  // 4 waves in a 2 x 2 wave grid

// Logical tile id for the hardware workgroup id `hw` so that consecutive logical ids
// run on the same XCD (the dispatcher deals workgroups round-robin over 8 XCDs).
__device__ __forceinline__ unsigned xcd_remap(unsigned hw, unsigned total) {
  const unsigned q = total / 8u, r = total % 8u;
  const unsigned x = hw % 8u, i = hw / 8u;
  return x * q + (x < r ? x : r) + i;
}

template <int BM, int BN, int BK, bool AK, bool BKC>
__global__ __launch_bounds__(kThreads, 2) void dgemm_mfma_kernel(Gemm g, int tiles_m,
                                                                int tiles_n, int nsplit) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int PA = 4, PB = 4;  // LDS row padding (doubles)
  constexpr int A_PER = BM * BK / kThreads;
  constexpr int B_PER = BN * BK / kThreads;
  static_assert(BM * BK % kThreads == 0 && BN * BK % kThreads == 0, "tile/threads");
  __shared__ double As[2][BK][BM + PA];
  __shared__ double Bs[2][BK][BN + PB];

  const unsigned total = (unsigned)tiles_m * tiles_n * nsplit;
  const unsigned hw = blockIdx.x;
  if (hw >= total) return;
  const unsigned lid = xcd_remap(hw, total);
  const int mt = lid % tiles_m;
  const int nt = (lid / tiles_m) % tiles_n;
  const int z = lid / (tiles_m * tiles_n);

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = mt * BM, n0 = nt * BN;
  const int M = g.M, N = g.N, K = g.K;

  const int r0 = z * g.bps;
  const int r1 = min(r0 + g.bps, g.nbatch);
  const int nkt = (K + BK - 1) / BK;
  const int T = (r1 - r0) * nkt;

  v4d acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  double ra[A_PER], rb[B_PER];

  auto load_tile = [&](int t) {
    const int r = r0 + t / nkt;
    const int k0 = (t % nkt) * BK;
    const double* A = g.A + (long)r * g.sA;
    const double* B = g.B + (long)r * g.sB;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int idx = tid + i * kThreads;
      int m, k;
      if (AK) { k = idx % BK; m = idx / BK; } else { m = idx % BM; k = idx / BM; }
      const int gm = m0 + m, gk = k0 + k;
      double v = 0.0;
      if (gm < M && gk < K) v = AK ? A[(long)gm * g.lda + gk] : A[gm + (long)gk * g.lda];
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int idx = tid + i * kThreads;
      int n, k;
      if (BKC) { k = idx % BK; n = idx / BK; } else { n = idx % BN; k = idx / BN; }
      const int gn = n0 + n, gk = k0 + k;
      double v = 0.0;
      if (gn < N && gk < K) v = BKC ? B[gk + (long)gn * g.ldb] : B[(long)gk * g.ldb + gn];
      rb[i] = v;
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int idx = tid + i * kThreads;
      int m, k;
      if (AK) { k = idx % BK; m = idx / BK; } else { m = idx % BM; k = idx / BM; }
      As[buf][k][m] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int idx = tid + i * kThreads;
      int n, k;
      if (BKC) { k = idx % BK; n = idx / BK; } else { n = idx % BN; k = idx / BN; }
      Bs[buf][k][n] = rb[i];
    }
  };

  if (T > 0) {
    load_tile(0);
    store_tile(0);
  }
  __syncthreads();
  const int fr = lane & 15, fk = lane >> 4;
  for (int t = 0; t < T; ++t) {
    const int buf = t & 1;
    if (t + 1 < T) load_tile(t + 1);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = As[buf][kk + fk][wm * WM + i * 16 + fr];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = Bs[buf][kk + fk][wn * WN + j * 16 + fr];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (t + 1 < T) store_tile(buf ^ 1);
    __syncthreads();
  }

  // D(16x16) layout of v_mfma_f64_16x16x4f64: lane l holds column l%16, rows
  // 4*(l/16) .. 4*(l/16)+3 in its four accumulator registers.
  double* Cz = g.C + (long)z * g.sC;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn * WN + j * 16 + fr;
      if (col >= N) continue;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int row = m0 + wm * WM + i * 16 + 4 * fk + rr;
        if (row < M) {
          double* c = Cz + row + (long)col * g.ldc;
          const double v = g.alpha * acc[i][j][rr];
          *c = (g.beta == 0.0) ? v : v + g.beta * (*c);
        }
      }
    }
}

__global__ void reduce_splits_kernel(const double* __restrict__ P, long sP, int nsplit, int M,
                                     int N, long ld, double* __restrict__ out, long ldo,
                                     double beta) {
  const long total = (long)M * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int m = e % M, n = e / M;
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += P[z * sP + m + n * ld];
    double* o = out + m + n * ldo;
    *o = (beta == 0.0) ? s : s + beta * (*o);
  }
}

template <int BM, int BN, bool AK, bool BKC>
int launch(hipStream_t s, const Gemm& g) {
  constexpr int BK = 16;
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  const int nsplit = (g.nbatch + g.bps - 1) / g.bps;
  const long total = (long)tm * tn * nsplit;
  if (total <= 0) return DN_OK;
  if (total > 0x7fffffffL) return fail(DN_ERR_ARG, "gemm grid too large");
  hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, BK, AK, BKC>), dim3((unsigned)total),
                     dim3(kThreads), 0, s, g, tm, tn, nsplit);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

template <int BM, int BN>
int dispatch_layout(hipStream_t s, const Gemm& g) {
  if (g.a_kcontig) {
    return g.b_kcontig ? launch<BM, BN, true, true>(s, g) : launch<BM, BN, true, false>(s, g);
  }
  return g.b_kcontig ? launch<BM, BN, false, true>(s, g) : launch<BM, BN, false, false>(s, g);
}

}  // namespace

int gemm_f64(hipStream_t s, const Gemm& g) {
  if (g.M <= 0 || g.N <= 0) return DN_OK;
  if (g.K <= 0 || g.nbatch <= 0 || g.bps <= 0)
    return fail(DN_ERR_ARG, "gemm_f64: bad K/nbatch/bps");
  // Tall-M operands get the 128-row tile (the RI transforms have M = nbf5 <= 128 or
  // a multiple of it); everything else uses 64 x 64.
  if (g.M > 64 && g.N >= 64) return dispatch_layout<128, 64>(s, g);
  return dispatch_layout<64, 64>(s, g);
}

int reduce_splits(hipStream_t s, const double* P, long sP, int nsplit, int M, int N, long ld,
                  double* out, long ldo, double beta) {
  const long total = (long)M * N;
  if (total <= 0) return DN_OK;
  const int threads = 256;
  const long blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(reduce_splits_kernel, dim3((unsigned)(blocks < 4096 ? blocks : 4096)),
                     dim3(threads), 0, s, P, sP, nsplit, M, N, ld, out, ldo, beta);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

}  // namespace dn
