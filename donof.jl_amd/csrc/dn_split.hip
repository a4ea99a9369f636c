// Split-precision J/K-MO build: FP16 hi+lo operand pairs on the CDNA4 matrix cores
// (v_mfma_f32_16x16x32_f16), three products hi*hi + hi*lo + lo*hi accumulated in FP32.
//
// Each operand x is stored as hi = f16(x s), lo = f16(x s - hi) with a power-of-two
// scale s (per auxiliary index Q for b, one for C5), which keeps ~22 significant bits
// for every element within 2^-22 of the slab maximum and keeps FP16 out of overflow and
// out of its subnormal range for the lo parts.  The two RI transforms of the reference
// (integrals.jl:159-160) are two launches of one kernel template:
//   bpn[p,(n,Q)] = sum_m C5[m,p] b[m,(n,Q)]       A: C5^T planes, B: b planes
//   bpq[:,:,Q]   = bpn[:,:,Q] C5                   A: bpn (FP32, split on load), B: C5 planes
// and J, K (integrals.jl:163,166) are reduced from the FP32 bpq in FP64.
#include <cmath>

#include "dn_internal.h"

namespace dn {

typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef _Float16 v2h __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ unsigned xcd_remap16(unsigned hw, unsigned total) {
  const unsigned q = total / 8u, r = total % 8u;
  const unsigned x = hw % 8u, i = hw / 8u;
  return x * q + (x < r ? x : r) + i;
}

// power-of-two scale putting max|x| into (2^13, 2^14]
__host__ __device__ __forceinline__ float pow2_scale(float amax) {
  if (!(amax > 0.f)) return 1.0f;
  const int e = (int)ceilf(log2f(amax));
  return ldexpf(1.0f, 14 - e);
}

__device__ __forceinline__ void split_f16(float x, _Float16* hi, _Float16* lo) {
  const _Float16 h = (_Float16)x;
  *hi = h;
  *lo = (_Float16)(x - (float)h);
}

__device__ __forceinline__ void split_f16_d(double x, _Float16* hi, _Float16* lo) {
  const _Float16 h = (_Float16)x;
  *hi = h;
  *lo = (_Float16)(x - (double)h);
}

__device__ __forceinline__ float block_max(float v, float* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  float m = 0.f;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) m = fmaxf(m, red[k]);
  return m;
}

}  // namespace

// ---- operand preparation -------------------------------------------------------------

// per-Q max |b| and the scale of slab Q (one block per Q)
__global__ void b_slab_max_kernel(const double* __restrict__ b, long slab, int naux,
                                  float* __restrict__ scale, float* __restrict__ inv_scale) {
  __shared__ float red[16];
  const int Q = blockIdx.x;
  const double* s = b + (long)Q * slab;
  float m = 0.f;
  for (long e = threadIdx.x; e < slab; e += blockDim.x) m = fmaxf(m, (float)fabs(s[e]));
  m = block_max(m, red);
  if (threadIdx.x == 0) {
    const float sc = pow2_scale(m * 1.0001f);
    scale[Q] = sc;
    inv_scale[Q] = 1.0f / sc;
  }
}

// b planes: hi/lo of b * scale[Q] as [(n,Q)][m] with leading dimension ld8 >= nbf
// (16-byte aligned rows; the padding stays zero)
__global__ void b_split_kernel(const double* __restrict__ b, int nbf, long slab, long total,
                               int ld8, const float* __restrict__ scale,
                               _Float16* __restrict__ hi, _Float16* __restrict__ lo) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const double x = b[e] * (double)scale[e / slab];
    const long col = e / nbf;
    const long o = col * ld8 + (e - col * nbf);
    split_f16_d(x, &hi[o], &lo[o]);
  }
}

// C5^T planes [Mpad][Kpad] (row = orbital p, k = AO index m), zero padded; one block
// computes the scale first.
__global__ void c_split_kernel(const double* __restrict__ C, int nbf, int nbf5, int Mpad,
                               int Kpad, float* __restrict__ cscale, _Float16* __restrict__ hi,
                               _Float16* __restrict__ lo) {
  __shared__ float red[16];
  __shared__ float sc_sh;
  float m = 0.f;
  for (long e = threadIdx.x; e < (long)nbf * nbf5; e += blockDim.x)
    m = fmaxf(m, (float)fabs(C[e]));
  m = block_max(m, red);
  if (threadIdx.x == 0) {
    sc_sh = pow2_scale(m * 1.0001f);
    cscale[0] = sc_sh;
    cscale[1] = 1.0f / sc_sh;
  }
  __syncthreads();
  const double sc = sc_sh;
  for (long e = threadIdx.x; e < (long)Mpad * Kpad; e += blockDim.x) {
    const int p = e / Kpad, k = e % Kpad;
    _Float16 h = 0, l = 0;
    if (p < nbf5 && k < nbf) split_f16_d(C[k + (long)p * nbf] * sc, &h, &l);
    hi[e] = h;
    lo[e] = l;
  }
}

__global__ void fill_f32_kernel(float* p, long n, float v) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x)
    p[e] = v;
}

// ---- the GEMM ----------------------------------------------------------------------------

struct SplitGemm {
  int M, N, K;
  // A: either f16 planes [row][k] (ld = lda) or FP32 column-major A[m + k*lda]
  const _Float16* ahi;
  const _Float16* alo;
  const float* af32;
  long lda, sA;             // batch stride (FP32 source)
  const float* a_scale;     // per batch: scale applied to the FP32 A before splitting
  // B: f16 planes [col][k] (ld = ldb)
  const _Float16* bhi;
  const _Float16* blo;
  long ldb;
  // C: FP32 column-major, batch stride sC
  float* C;
  long ldc, sC;
  // out = acc * cs[1] * (per-Q col factor or per-batch factor); cs = C5 scale pair on
  // the device (scale, 1/scale)
  const float* cs;
  const float* col_inv_scale;  // per Q = col / col_div (K1), nullable
  int col_div;
  const float* batch_inv_scale;  // per batch (K2), nullable
  float* qmax;                   // per Q max |out| (K1), nullable
  int nbatch;
};

constexpr int SB_K = 32;           // k per MFMA / per LDS tile
constexpr int SB_LD = SB_K + 8;    // halves per LDS row (80 bytes: conflict-free b128 reads)

template <int BM, int BN, int WGM, int WGN, bool AF32>
__global__ __launch_bounds__(64 * WGM * WGN) void split_gemm_kernel(SplitGemm g, int tiles_m,
                                                                   int tiles_n) {
  constexpr int NT = 64 * WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN, FM = WM / 16, FN = WN / 16;
  __shared__ _Float16 Ah[2][BM * SB_LD], Al[2][BM * SB_LD];
  __shared__ _Float16 Bh[2][BN * SB_LD], Bl[2][BN * SB_LD];

  const unsigned total = (unsigned)tiles_m * tiles_n * g.nbatch;
  if (blockIdx.x >= total) return;
  const unsigned lid = xcd_remap16(blockIdx.x, total);
  const int mt = lid % tiles_m;
  const int nt = (lid / tiles_m) % tiles_n;
  const int z = lid / (tiles_m * tiles_n);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WGM, wn = wave / WGM;
  const int m0 = mt * BM, n0 = nt * BN;
  const int M = g.M, N = g.N, K = g.K;
  const int nkt = (K + SB_K - 1) / SB_K;

  v4f acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  // staging registers: the loads of tile t+1 are issued before tile t is computed and
  // written to LDS after it (LDS is double-buffered, one barrier per tile)
  constexpr int AV = BM * 4 / NT;         // 16-byte loads per plane (f16 source)
  constexpr int AP = BM * SB_K / 2 / NT;  // k-pairs per thread (FP32 source)
  constexpr int BV = BN * 4 / NT;
  static_assert((AF32 ? AP : AV) >= 1 && BV >= 1, "tile too small for the thread count");
  v8h rah[AF32 ? 1 : AV], ral[AF32 ? 1 : AV];
  float rf0[AF32 ? AP : 1], rf1[AF32 ? AP : 1];
  v8h rbh[BV], rbl[BV];
  const float* Af = AF32 ? g.af32 + (long)z * g.sA : nullptr;
  const float ascale = AF32 ? g.a_scale[z] : 1.f;

  auto load = [&](int t) {
    const int k0 = t * SB_K;
    if constexpr (AF32) {
#pragma unroll
      for (int i = 0; i < AP; ++i) {
        const int idx = tid + i * NT;
        const int m = idx % BM, kp = idx / BM;
        const int gm = min(m0 + m, M - 1);
        const int gk0 = min(k0 + 2 * kp, K - 1), gk1 = min(k0 + 2 * kp + 1, K - 1);
        rf0[i] = Af[gm + (long)gk0 * g.lda];
        rf1[i] = Af[gm + (long)gk1 * g.lda];
      }
    } else {
#pragma unroll
      for (int i = 0; i < AV; ++i) {
        const int idx = tid + i * NT;
        const int r = idx / 4, kq = idx % 4;
        const long off = (long)(m0 + r) * g.lda + k0 + 8 * kq;  // planes are row/K padded
        rah[i] = *reinterpret_cast<const v8h*>(g.ahi + off);
        ral[i] = *reinterpret_cast<const v8h*>(g.alo + off);
      }
    }
#pragma unroll
    for (int i = 0; i < BV; ++i) {
      const int idx = tid + i * NT;
      const int r = idx / 4, kq = idx % 4;
      const int col = min(n0 + r, N - 1);
      const long off = (long)col * g.ldb + k0 + 8 * kq;
      rbh[i] = *reinterpret_cast<const v8h*>(g.bhi + off);
      rbl[i] = *reinterpret_cast<const v8h*>(g.blo + off);
    }
  };
  auto store = [&](int buf) {
    if constexpr (AF32) {
#pragma unroll
      for (int i = 0; i < AP; ++i) {
        const int idx = tid + i * NT;
        const int m = idx % BM, kp = idx / BM;
        _Float16 h0, l0, h1, l1;
        split_f16(rf0[i] * ascale, &h0, &l0);
        split_f16(rf1[i] * ascale, &h1, &l1);
        *reinterpret_cast<v2h*>(&Ah[buf][m * SB_LD + 2 * kp]) = v2h{h0, h1};
        *reinterpret_cast<v2h*>(&Al[buf][m * SB_LD + 2 * kp]) = v2h{l0, l1};
      }
    } else {
#pragma unroll
      for (int i = 0; i < AV; ++i) {
        const int idx = tid + i * NT;
        const int r = idx / 4, kq = idx % 4;
        *reinterpret_cast<v8h*>(&Ah[buf][r * SB_LD + 8 * kq]) = rah[i];
        *reinterpret_cast<v8h*>(&Al[buf][r * SB_LD + 8 * kq]) = ral[i];
      }
    }
#pragma unroll
    for (int i = 0; i < BV; ++i) {
      const int idx = tid + i * NT;
      const int r = idx / 4, kq = idx % 4;
      *reinterpret_cast<v8h*>(&Bh[buf][r * SB_LD + 8 * kq]) = rbh[i];
      *reinterpret_cast<v8h*>(&Bl[buf][r * SB_LD + 8 * kq]) = rbl[i];
    }
  };

  load(0);
  store(0);
  __syncthreads();
  const int fr = lane & 15, fk = 8 * (lane >> 4);
  for (int t = 0; t < nkt; ++t) {
    const int buf = t & 1;
    if (t + 1 < nkt) load(t + 1);
    v8h bh[FN], bl[FN];
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int r = wn * WN + j * 16 + fr;
      bh[j] = *reinterpret_cast<const v8h*>(&Bh[buf][r * SB_LD + fk]);
      bl[j] = *reinterpret_cast<const v8h*>(&Bl[buf][r * SB_LD + fk]);
    }
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int r = wm * WM + i * 16 + fr;
      const v8h ah = *reinterpret_cast<const v8h*>(&Ah[buf][r * SB_LD + fk]);
      const v8h al = *reinterpret_cast<const v8h*>(&Al[buf][r * SB_LD + fk]);
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[j], acc[i][j], 0, 0, 0);
      }
    }
    if (t + 1 < nkt) store(buf ^ 1);
    __syncthreads();
  }

  // D(16x16) of v_mfma_f32_16x16x32_f16: lane l holds column l%16, rows 4*(l/16) + r.
  float* Cz = g.C + (long)z * g.sC;
  const float bscale = g.batch_inv_scale ? g.batch_inv_scale[z] : 1.f;
  const int qa = g.col_div > 0 ? n0 / g.col_div : 0;
  float mx_a = 0.f, mx_b = 0.f;
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    const int col = n0 + wn * WN + j * 16 + fr;
    if (col >= N) continue;
    float f = g.cs[1] * bscale;
    int q = 0;
    if (g.col_inv_scale) {
      q = col / g.col_div;
      f *= g.col_inv_scale[q];
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int row = m0 + wm * WM + i * 16 + 4 * (lane >> 4) + rr;
        if (row < M) {
          const float v = acc[i][j][rr] * f;
          Cz[row + (long)col * g.ldc] = v;
          if (q == qa) mx_a = fmaxf(mx_a, fabsf(v));
          else if (q == qa + 1) mx_b = fmaxf(mx_b, fabsf(v));
          else if (g.qmax) atomicMax(reinterpret_cast<unsigned*>(&g.qmax[q]),
                                     __float_as_uint(fabsf(v)));
        }
      }
  }
  if (g.qmax) {
    for (int off = 32; off > 0; off >>= 1) {
      mx_a = fmaxf(mx_a, __shfl_down(mx_a, off, 64));
      mx_b = fmaxf(mx_b, __shfl_down(mx_b, off, 64));
    }
    if (lane == 0) {
      const int nq = N / g.col_div;
      atomicMax(reinterpret_cast<unsigned*>(&g.qmax[qa]), __float_as_uint(mx_a));
      if (qa + 1 < nq) atomicMax(reinterpret_cast<unsigned*>(&g.qmax[qa + 1]), __float_as_uint(mx_b));
    }
  }
}

// per-batch scale for the FP32 A of the second transform
__global__ void bpn_scale_kernel(const float* __restrict__ qmax, int naux,
                                 float* __restrict__ scale, float* __restrict__ inv) {
  const int Q = blockIdx.x * blockDim.x + threadIdx.x;
  if (Q >= naux) return;
  const float s = pow2_scale(qmax[Q] * 1.0001f);
  scale[Q] = s;
  inv[Q] = 1.0f / s;
}

// ---- host side ---------------------------------------------------------------------------

static unsigned grid_of(long n) {
  long b = (n + 255) / 256;
  return (unsigned)std::max<long>(1, std::min<long>(b, 8192));
}

int split_prepare_b(hipStream_t s, const double* b, int nbf, int naux, int ld8, float* scale,
                    float* inv_scale, _Float16* hi, _Float16* lo) {
  const long slab = (long)nbf * nbf;
  hipLaunchKernelGGL(b_slab_max_kernel, dim3(naux), dim3(256), 0, s, b, slab, naux, scale,
                     inv_scale);
  DN_HIP(hipGetLastError());
  hipLaunchKernelGGL(b_split_kernel, dim3(grid_of(slab * naux)), dim3(256), 0, s, b, nbf, slab,
                     slab * naux, ld8, scale, hi, lo);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int split_prepare_c(hipStream_t s, const double* C, int nbf, int nbf5, int Mpad, int Kpad,
                    float* cscale, _Float16* hi, _Float16* lo) {
  hipLaunchKernelGGL(c_split_kernel, dim3(1), dim3(1024), 0, s, C, nbf, nbf5, Mpad, Kpad,
                     cscale, hi, lo);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

template <int BM, int BN, int WGM, int WGN, bool AF32>
static int launch_split(hipStream_t s, const SplitGemm& g) {
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  const long total = (long)tm * tn * g.nbatch;
  if (total > 0x7fffffffL) return fail(DN_ERR_ARG, "split gemm grid too large");
  hipLaunchKernelGGL((split_gemm_kernel<BM, BN, WGM, WGN, AF32>), dim3((unsigned)total),
                     dim3(64 * WGM * WGN), 0, s, g, tm, tn);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// The split-precision J/K build: C planes must be prepared (split_prepare_c) for the
// current orbitals; bpn (nbf5*nbf*naux floats), bpq (nbf5*nbf5*naux floats), qmax,
// qs/qinv (naux floats each) are workspaces.  b planes have rows of ld8 halves; the
// C planes are [Mpad][Kpad] with Mpad, Kpad multiples of 128 and 32.
int split_k1(hipStream_t s, int nbf, int nbf5, int naux, int Kpad, int ld8,
             const _Float16* bhi, const _Float16* blo, const float* b_inv_scale,
             const _Float16* chi, const _Float16* clo, const float* cscale_dev, float* bpn,
             float* qmax) {
  hipLaunchKernelGGL(fill_f32_kernel, dim3(grid_of(naux)), dim3(256), 0, s, qmax, (long)naux, 0.f);
  DN_HIP(hipGetLastError());
  SplitGemm g1{};
  g1.M = nbf5; g1.N = nbf * naux; g1.K = nbf;
  g1.ahi = chi; g1.alo = clo; g1.lda = Kpad;
  g1.bhi = bhi; g1.blo = blo; g1.ldb = ld8;
  g1.C = bpn; g1.ldc = nbf5; g1.sC = 0;
  g1.cs = cscale_dev; g1.col_inv_scale = b_inv_scale; g1.col_div = nbf;
  g1.qmax = qmax; g1.nbatch = 1;
  return launch_split<128, 128, 2, 4, false>(s, g1);
}

int split_k2(hipStream_t s, int nbf, int nbf5, int naux, int Kpad, const _Float16* chi,
             const _Float16* clo, const float* cscale_dev, const float* bpn, float* bpq,
             const float* qmax, float* qs, float* qinv) {
  hipLaunchKernelGGL(bpn_scale_kernel, dim3((naux + 255) / 256), dim3(256), 0, s, qmax, naux, qs,
                     qinv);
  DN_HIP(hipGetLastError());
  SplitGemm g2{};
  g2.M = nbf5; g2.N = nbf5; g2.K = nbf;
  g2.af32 = bpn; g2.lda = nbf5; g2.sA = (long)nbf5 * nbf; g2.a_scale = qs;
  g2.bhi = chi; g2.blo = clo; g2.ldb = Kpad;
  g2.C = bpq; g2.ldc = nbf5; g2.sC = (long)nbf5 * nbf5;
  g2.cs = cscale_dev; g2.batch_inv_scale = qinv; g2.col_div = 0;
  g2.nbatch = naux;
  return launch_split<128, 128, 2, 4, true>(s, g2);
}

}  // namespace dn
