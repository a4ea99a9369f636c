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
//
// Plane layout ("tile-major"): the contraction index k (the AO index m of b and C5) is cut
// into tiles of 32; element (row r, k) of a plane with R rows sits at
// (k / 32) * R * 32 + r * 32 + k % 32, so one 32-wide k tile of R rows is R * 64
// contiguous bytes and the 16-row x 64-byte block that one load instruction moves is 1 KB
// contiguous (row-major planes spread it over 16 rows of stride 2 nbf bytes; the
// first transform at 800/3200/240 takes 2.9 ms with those and 2.3 ms with these).  The
// per-k-tile stride R * 32 is the operand's lda / ldb; b planes hold one such block per
// Q (batch stride sB = nbf * Kpad), with k padded to Kpad = 32 * ceil(nbf / 32) by zeros.
#include <cmath>
#include <type_traits>

#include "dn_internal.h"
#include "dn_sched.h"

namespace dn {

typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef _Float16 v2h __attribute__((ext_vector_type(2)));
typedef _Float16 v4h __attribute__((ext_vector_type(4)));
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

// b planes: hi/lo of b * scale[Q], tile-major per slab Q: (column n, k = m) at
// Q nbf Kpad + (m / 32) nbf 32 + n 32 + m % 32 (k past nbf stays zero: the planes are
// cleared at creation and b's own pad row ld > nbf holds zeros)
__global__ void b_split_kernel(const double* __restrict__ b, int ld, int nbf, long slab,
                               long total, int Kpad, const float* __restrict__ scale,
                               _Float16* __restrict__ hi, _Float16* __restrict__ lo) {
  const long pslab = (long)nbf * Kpad;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long Q = e / slab;
    const double x = b[e] * (double)scale[Q];
    const long r = e - Q * slab;
    const int n = (int)(r / ld), m = (int)(r - (long)n * ld);
    const long o = Q * pslab + (long)(m >> 5) * nbf * 32 + (long)n * 32 + (m & 31);
    split_f16_d(x, &hi[o], &lo[o]);
  }
}

// C5^T planes (row = orbital p < Mpad, k = AO index m < Kpad), tile-major, zero padded.
// Per column p of C5 (one workgroup each): max |C5[:,p]| and ||C5[:,p]|| into cscale[2] and
// cscale[3] (bits of non-negative floats, which order like the values; cleared first).
__global__ __launch_bounds__(256) void c_stats_kernel(const double* __restrict__ C, int nbf,
                                                     float* __restrict__ cscale) {
  __shared__ double red_n[4];
  __shared__ float red_m[4];
  const int p = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double s2 = 0.0;
  float m = 0.f;
  for (int k = threadIdx.x; k < nbf; k += 256) {
    const double x = C[k + (long)p * nbf];
    s2 = fma(x, x, s2);
    m = fmaxf(m, (float)fabs(x));
  }
  for (int off = 32; off > 0; off >>= 1) {
    s2 += __shfl_xor(s2, off, 64);
    m = fmaxf(m, __shfl_xor(m, off, 64));
  }
  if (lane == 0) {
    red_n[w] = s2;
    red_m[w] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const float n = (float)sqrt((red_n[0] + red_n[1]) + (red_n[2] + red_n[3])) * 1.000001f;
    const float mm = fmaxf(fmaxf(red_m[0], red_m[1]), fmaxf(red_m[2], red_m[3]));
    atomicMax(reinterpret_cast<unsigned*>(cscale + 2), __float_as_uint(mm));
    atomicMax(reinterpret_cast<unsigned*>(cscale + 3), __float_as_uint(n));
  }
}

// The C5 planes under the power-of-two scale of max |C5| (block 0 records the scale pair),
// and, with lhmax, the per-Q scales of the symmetric build's T planes from the
// Cauchy-Schwarz bound ||C5[:,p]|| ||Lh_Q[:,n]|| >= |T[p,n]| (the first blocks).
__global__ void c_planes_kernel(const double* __restrict__ C, int nbf, int nbf5, int Mpad,
                                int Kpad, float* __restrict__ cscale, _Float16* __restrict__ hi,
                                _Float16* __restrict__ lo, const float* __restrict__ lhmax,
                                int naux, float* __restrict__ qs, float* __restrict__ qinv) {
  const float sc = pow2_scale(cscale[2] * 1.0001f);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cscale[0] = sc;
    cscale[1] = 1.0f / sc;
  }
  if (lhmax) {
    const int Q = blockIdx.x * blockDim.x + threadIdx.x;
    if (Q < naux) {
      const float sq = pow2_scale(lhmax[Q] * cscale[3] * 1.0001f);
      qs[Q] = sq;
      qinv[Q] = 1.0f / sq;
    }
  }
  const long total = (long)Mpad * Kpad;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    // e runs over the tile-major plane: k tile, row p, k within the tile
    const int kt = e / ((long)Mpad * 32), rem = e % ((long)Mpad * 32);
    const int p = rem >> 5, k = kt * 32 + (rem & 31);
    _Float16 h = 0, l = 0;
    if (p < nbf5 && k < nbf) split_f16_d(C[k + (long)p * nbf] * (double)sc, &h, &l);
    hi[e] = h;
    lo[e] = l;
  }
}

// Per slab Q: max_n ||Lh_Q[:,n]|| with Lh the strict lower triangle of b_Q plus half its
// diagonal (b_Q = Lh + Lh^T).  With ||C5[:,p]|| it bounds |T[p,n]| = |sum_{m>=n} C5[m,p]
// Lh[m,n]| by Cauchy-Schwarz, which sets the scale of T's FP16 planes before T exists.
// One block per Q, a wavefront per column.
__global__ __launch_bounds__(256) void b_lhnorm_kernel(const double* __restrict__ b, int nbf,
                                                       int ld, float* __restrict__ lhmax) {
  __shared__ double red[4];
  const double* sl = b + (long)blockIdx.x * ld * nbf;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double best = 0.0;
  for (int n = w; n < nbf; n += 4) {
    double s2 = 0.0;
    for (int m = n + lane; m < nbf; m += 64) {
      double x = sl[m + (long)n * ld];
      if (m == n) x *= 0.5;
      s2 = fma(x, x, s2);
    }
    for (int off = 32; off > 0; off >>= 1) s2 += __shfl_down(s2, off, 64);
    best = fmax(best, s2);  // lane 0 holds the column's sum
  }
  if (lane == 0) red[w] = best;
  __syncthreads();
  if (threadIdx.x == 0)
    lhmax[blockIdx.x] =
        (float)sqrt(fmax(fmax(red[0], red[1]), fmax(red[2], red[3]))) * 1.000001f;
}

__global__ void fill_f32_kernel(float* p, long n, float v) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x)
    p[e] = v;
}

// ---- the GEMM ----------------------------------------------------------------------------

constexpr int SB_K = 32;  // k per MFMA / per LDS tile: four 16-byte chunks of 8 halves
// LDS tiles are K-major in 16-byte chunks: chunk c (k = 8c..8c+7) of row r at
// (c * ROWS + r) * 8 halves.  An MFMA fragment read (lane l: row l % 16, chunk l / 16) and
// a staging store (lane l: row l % 16 + 16 (l / 64), chunk (l / 16) % 4) are then runs of
// 16 consecutive 16-byte chunks, free of bank conflicts (a row-major tile with padded
// rows conflicts on one of the two).
__device__ __forceinline__ int chunk_off(int rows, int r, int c) { return (c * rows + r) * 8; }
__device__ __forceinline__ void stage_rc(int idx, int* r, int* c) {
  *r = (idx & 15) + 16 * (idx >> 6);
  *c = (idx >> 4) & 3;
}

// Output of one wave's FM x FN fragments at rows r0.., columns c0..; n0 is the tile's first
// column (for the per-Q maxima).  D(16x16) of v_mfma_f32_16x16x32_f16: lane l holds
// column l % 16, rows 4 (l / 16) + r.
template <int FM, int FN>
__device__ __forceinline__ void split_epilogue(const SplitGemm& g, v4f (&acc)[FM][FN], int r0,
                                               int c0, int n0, int z, int lane) {
  const int fr = lane & 15;
  const int M = g.M, N = g.N;
  float* Cz = g.C + (long)z * g.sC;
  const float bscale = g.batch_inv_scale ? g.batch_inv_scale[z] : 1.f;
  const int qa = g.col_div > 0 ? n0 / g.col_div : 0;
  float mx_a = 0.f, mx_b = 0.f;
  // each lane holds 4 consecutive rows of one column: one 16-byte store when the column
  // stride and the base keep it aligned
  const bool vec = (g.ldc % 4 == 0) && ((uintptr_t)Cz % 16 == 0);
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    const int col = c0 + j * 16 + fr;
    if (col >= N) continue;
    float f = g.cs[1] * bscale;
    int q = 0;
    if (g.col_inv_scale) {
      q = col / g.col_div;
      f *= g.col_inv_scale[q];
    }
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int row0 = r0 + i * 16 + 4 * (lane >> 4);
      const v4f v = acc[i][j] * f;
      float* c = Cz + row0 + (long)col * g.ldc;
      float m = 0.f;
      if (vec && row0 + 3 < M) {
        *reinterpret_cast<v4f*>(c) = v;
        m = fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3])));
      } else {
#pragma unroll
        for (int rr = 0; rr < 4; ++rr)
          if (row0 + rr < M) {
            c[rr] = v[rr];
            m = fmaxf(m, fabsf(v[rr]));
          }
      }
      if (g.col_div == 0 || q == qa) mx_a = fmaxf(mx_a, m);
      else if (q == qa + 1) mx_b = fmaxf(mx_b, m);
      else if (g.qmax) atomicMax(reinterpret_cast<unsigned*>(&g.qmax[q]), __float_as_uint(m));
    }
  }
  if (g.qmax) {
    for (int off = 32; off > 0; off >>= 1) {
      mx_a = fmaxf(mx_a, __shfl_down(mx_a, off, 64));
      mx_b = fmaxf(mx_b, __shfl_down(mx_b, off, 64));
    }
    if (lane == 0) {
      if (g.col_div == 0) {
        atomicMax(reinterpret_cast<unsigned*>(&g.qmax[z]), __float_as_uint(mx_a));
      } else {
        const int nq = N / g.col_div;
        atomicMax(reinterpret_cast<unsigned*>(&g.qmax[qa]), __float_as_uint(mx_a));
        if (qa + 1 < nq)
          atomicMax(reinterpret_cast<unsigned*>(&g.qmax[qa + 1]), __float_as_uint(mx_b));
      }
    }
  }
}

// TRIB: B is the symmetric slab b_Q (col n, k = m) and the tile multiplies by its lower
// half Lh (m > n as is, m == n halved, m < n zero; b_Q = Lh + Lh^T), so the K loop of
// column block n0 starts at m = n0.
// LDS of split_tile: double-buffered hi and lo planes of A (BM rows) and B (BN rows)
template <int BM, int BN>
constexpr int split_tile_bytes() {
  return 2 * 2 * (BM + BN) * SB_K * (int)sizeof(_Float16);
}

template <int BM, int BN, int WGM, int WGN, bool AF32, bool TRIB>
__device__ __forceinline__ void split_tile(const SplitGemm& g, int mt, int nt, int z,
                                           unsigned char* smem) {
  constexpr int NT = 64 * WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN, FM = WM / 16, FN = WN / 16;
  typedef _Float16 PlaneA[2][BM * SB_K];
  typedef _Float16 PlaneB[2][BN * SB_K];
  PlaneA& Ah = *reinterpret_cast<PlaneA*>(smem);
  PlaneA& Al = *reinterpret_cast<PlaneA*>(smem + sizeof(PlaneA));
  PlaneB& Bh = *reinterpret_cast<PlaneB*>(smem + 2 * sizeof(PlaneA));
  PlaneB& Bl = *reinterpret_cast<PlaneB*>(smem + 2 * sizeof(PlaneA) + sizeof(PlaneB));

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WGM, wn = wave / WGM;
  const int m0 = mt * BM, n0 = nt * BN;
  const int M = g.M, N = g.N, K = g.K;
  const int kt0 = TRIB ? n0 / SB_K : 0;
  const int nkt = (K + SB_K - 1) / SB_K - kt0;

  v4f acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  // staging registers: the loads of tile t+1 are issued before tile t is computed and
  // written to LDS after it (LDS is double-buffered, one barrier per tile)
  constexpr int AV = BM * 4 / NT;  // 16-byte chunks per thread and plane (A, B)
  constexpr int BV = BN * 4 / NT;
  static_assert(AV >= 1 && BV >= 1 && BM * 4 % NT == 0 && BN * 4 % NT == 0,
                "tile too small for the thread count");
  static_assert(!AF32 || BM % 64 == 0, "FP32 A staging covers 64-row groups");
  v8h rah[AF32 ? 1 : AV], ral[AF32 ? 1 : AV];
  float rf[AF32 ? AV : 1][8];
  v8h rbh[BV], rbl[BV];
  const float* Af = AF32 ? g.af32 + (long)z * g.sA : nullptr;
  const float ascale = AF32 ? g.a_scale[z] : 1.f;
  const _Float16* Bhi = g.bhi + (long)z * g.sB;
  const _Float16* Blo = g.blo + (long)z * g.sB;

  // FP32 A: task idx = (row m = idx % BM, chunk c = idx / BM), eight column-major loads
  // (consecutive lanes read consecutive rows)
  auto load = [&](int t) {
    const int k0 = (kt0 + t) * SB_K;
    if constexpr (AF32) {
#pragma unroll
      for (int i = 0; i < AV; ++i) {
        const int idx = tid + i * NT;
        const int m = idx % BM, c = idx / BM;
        const int gm = min(m0 + m, M - 1);
#pragma unroll
        for (int e = 0; e < 8; ++e)
          rf[i][e] = Af[gm + (long)min(k0 + 8 * c + e, K - 1) * g.lda];
      }
    } else {
#pragma unroll
      for (int i = 0; i < AV; ++i) {
        int r, c;
        stage_rc(tid + i * NT, &r, &c);
        // tile-major planes (rows and k padded): k tile kt0 + t at (kt0 + t) lda
        const long off = (long)(kt0 + t) * g.lda + (long)(m0 + r) * 32 + 8 * c;
        rah[i] = *reinterpret_cast<const v8h*>(g.ahi + off);
        ral[i] = *reinterpret_cast<const v8h*>(g.alo + off);
      }
    }
#pragma unroll
    for (int i = 0; i < BV; ++i) {
      int r, c;
      stage_rc(tid + i * NT, &r, &c);
      const int col = min(n0 + r, N - 1);
      const long off = (long)(kt0 + t) * g.ldb + (long)col * 32 + 8 * c;  // tile-major
      rbh[i] = *reinterpret_cast<const v8h*>(Bhi + off);
      rbl[i] = *reinterpret_cast<const v8h*>(Blo + off);
    }
  };
  auto store = [&](int buf, int t) {
    if constexpr (AF32) {
      const int k0 = (kt0 + t) * SB_K;
#pragma unroll
      for (int i = 0; i < AV; ++i) {
        const int idx = tid + i * NT;
        const int m = idx % BM, c = idx / BM;
        v8h h, l;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          _Float16 he, le;
          // k beyond K reads a clamped element: zero it (the B planes are zero there too)
          split_f16(k0 + 8 * c + e < K ? rf[i][e] * ascale : 0.f, &he, &le);
          h[e] = he;
          l[e] = le;
        }
        *reinterpret_cast<v8h*>(&Ah[buf][chunk_off(BM, m, c)]) = h;
        *reinterpret_cast<v8h*>(&Al[buf][chunk_off(BM, m, c)]) = l;
      }
    } else {
#pragma unroll
      for (int i = 0; i < AV; ++i) {
        int r, c;
        stage_rc(tid + i * NT, &r, &c);
        *reinterpret_cast<v8h*>(&Ah[buf][chunk_off(BM, r, c)]) = rah[i];
        *reinterpret_cast<v8h*>(&Al[buf][chunk_off(BM, r, c)]) = ral[i];
      }
    }
#pragma unroll
    for (int i = 0; i < BV; ++i) {
      int r, c;
      stage_rc(tid + i * NT, &r, &c);
      if constexpr (TRIB) {
        // m = k0 + 8 c + e against n = n0 + r; only the k tiles that reach into this
        // column block's diagonal need the mask (a uniform branch)
        const int kk = (kt0 + t) * SB_K;
        if (kk < n0 + BN) {
          const int d = kk + 8 * c - (n0 + r);
          v8h f;
#pragma unroll
          for (int e = 0; e < 8; ++e) {  // bit patterns of 1.0, 0.5, 0.0 in FP16
            const unsigned short bits = d + e > 0 ? 0x3C00 : (d + e == 0 ? 0x3800 : 0);
            f[e] = __builtin_bit_cast(_Float16, bits);
          }
          rbh[i] *= f;
          rbl[i] *= f;
        }
      }
      *reinterpret_cast<v8h*>(&Bh[buf][chunk_off(BN, r, c)]) = rbh[i];
      *reinterpret_cast<v8h*>(&Bl[buf][chunk_off(BN, r, c)]) = rbl[i];
    }
  };

  if (nkt > 0) {
    load(0);
    store(0, 0);
  }
  __syncthreads();
  const int fr = lane & 15, fc = lane >> 4;
  bool live[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
      live[i][j] = m0 + wm * WM + 16 * i < M && n0 + wn * WN + 16 * j < N;
  for (int t = 0; t < nkt; ++t) {
    const int buf = t & 1;
    if (t + 1 < nkt) load(t + 1);
    v8h bh[FN], bl[FN], ah[FM], al[FM];
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int r = wn * WN + j * 16 + fr;
      bh[j] = *reinterpret_cast<const v8h*>(&Bh[buf][chunk_off(BN, r, fc)]);
      bl[j] = *reinterpret_cast<const v8h*>(&Bl[buf][chunk_off(BN, r, fc)]);
    }
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int r = wm * WM + i * 16 + fr;
      ah[i] = *reinterpret_cast<const v8h*>(&Ah[buf][chunk_off(BM, r, fc)]);
      al[i] = *reinterpret_cast<const v8h*>(&Al[buf][chunk_off(BM, r, fc)]);
    }
    // the three products, each over all FM x FN accumulators in turn: consecutive MFMAs
    // write different accumulators, so none waits for the previous one's result;
    // fragments that lie wholly past M or N are skipped (wavefront-uniform branches)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        if (live[i][j])
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        if (live[i][j])
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        if (live[i][j])
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
    if (t + 1 < nkt) store(buf ^ 1, t + 1);
    __syncthreads();
  }

  split_epilogue<FM, FN>(g, acc, m0 + wm * WM, n0 + wn * WN, n0, z, lane);
}

template <int BM, int BN, int WGM, int WGN, bool AF32, bool TRIB = false>
__global__ __launch_bounds__(64 * WGM * WGN) void split_gemm_kernel(SplitGemm g, int tiles_m,
                                                                   int tiles_n) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[split_tile_bytes<BM, BN>()];
  const unsigned total = (unsigned)tiles_m * tiles_n * g.nbatch;
  if (blockIdx.x >= total) return;
  const unsigned lid = xcd_remap16(blockIdx.x, total);
  const int mt = lid % tiles_m;
  const int nt = (lid / tiles_m) % tiles_n;
  const int z = lid / (tiles_m * tiles_n);
  split_tile<BM, BN, WGM, WGN, AF32, TRIB>(g, mt, nt, z, smem);
}

// ---- f16-plane GEMM with LDS-direct loads ---------------------------------------------
//
// Both operands are FP16 planes (A [row][k], B [col][k]).  Every k tile of A and B goes
// from HBM straight into LDS (global_load_lds_dwordx4: lane l writes 16 bytes at the
// instruction's LDS base + 16 l), NST - 1 tiles ahead of the one being multiplied, so no
// staging registers are spent on it and the loads of several tiles are in flight.
// LDS tiles come in blocks of 16 rows, K-major inside a block: chunk c (k = 8c..8c+7) of
// row r at (r / 16) * 1024 + c * 256 + (r % 16) * 16 bytes.  One load instruction fills
// one block (lane l: row l % 16, chunk l / 16 -> the 64 contiguous global bytes of each
// of 16 rows), and the 16 lanes of an MFMA fragment read that share a chunk read 256
// contiguous bytes, free of bank conflicts.

// s_waitcnt vmcnt(N), no wait on the other counters (gfx9 encoding: vmcnt in [3:0] and
// [15:14], expcnt [6:4], lgkmcnt [11:8])
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

template <int BM, int BN, int NST>
constexpr int split_lds_bytes() {
  return NST * (BM + BN) * 128;  // per stage: 2 planes x rows x 64 bytes
}

// Epilogue of the symmetric first transform when the second transform takes its output
// as FP16 planes: out = acc * cs[1] * batch_inv_scale[z] * out_scale[z] (columns past N
// written as zero: they are the k padding of the second transform), split into hi/lo and
// stored tile-major.  A lane holds four rows of one column of each fragment; a 4 x 4
// transpose inside each quad of lanes (two xor-shuffle rounds) gives lane q of the quad
// row q of the quad's four columns, four consecutive halves of the plane: one 8-byte store
// per plane, and the 16 lanes of a row write its 32 bytes of the fragment.
template <int BM, int BN, int WGM, int WGN, int FM, int FN>
__device__ __forceinline__ void split_epilogue_planes(const SplitGemm& g, v4f (&acc)[FM][FN],
                                                      int m0, int n0, int wm, int wn, int z,
                                                      int lane) {
  constexpr int WM = BM / WGM, WN = BN / WGN;
  const float f = g.cs[1] * g.batch_inv_scale[z] * g.out_scale[z];
  const int q = lane & 3;
  _Float16* ohi = g.ohi + (long)z * g.out_sb;
  _Float16* olo = g.olo + (long)z * g.out_sb;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      float v0 = acc[i][j][0], v1 = acc[i][j][1], v2 = acc[i][j][2], v3 = acc[i][j][3];
      {  // round 1 (partner lane ^ 1): keep the rows whose bit 0 is this lane's
        const bool odd = q & 1;
        const float s0 = odd ? v0 : v1, s1 = odd ? v2 : v3;
        const float r0 = __shfl_xor(s0, 1, 64), r1 = __shfl_xor(s1, 1, 64);
        const float k0 = odd ? v1 : v0, k1 = odd ? v3 : v2;
        v0 = odd ? r0 : k0;
        v1 = odd ? k0 : r0;
        v2 = odd ? r1 : k1;
        v3 = odd ? k1 : r1;
      }
      {  // round 2 (partner lane ^ 2): the lower row's pair stays in lanes q < 2
        const bool hi2 = q & 2;
        const float s0 = hi2 ? v0 : v2, s1 = hi2 ? v1 : v3;
        const float r0 = __shfl_xor(s0, 2, 64), r1 = __shfl_xor(s1, 2, 64);
        const float k0 = hi2 ? v2 : v0, k1 = hi2 ? v3 : v1;
        v0 = hi2 ? r0 : k0;
        v1 = hi2 ? r1 : k1;
        v2 = hi2 ? k0 : r0;
        v3 = hi2 ? k1 : r1;
      }
      // row m0 + lr, columns n0 + lc0 .. + 3
      const int lc0 = wn * WN + j * 16 + (lane & 12);
      const int lr = wm * WM + i * 16 + 4 * (lane >> 4) + q;
      const int col = n0 + lc0;
      if (col / 32 >= g.out_kt) continue;
      const float vals[4] = {v0, v1, v2, v3};
      v4h h4, l4;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        _Float16 h, l;
        split_f16(col + c < g.N ? vals[c] * f : 0.f, &h, &l);
        h4[c] = h;
        l4[c] = l;
      }
      const long o = (long)(col >> 5) * g.out_ld + (long)(m0 + lr) * 32 + (col & 31);
      *reinterpret_cast<v4h*>(ohi + o) = h4;
      *reinterpret_cast<v4h*>(olo + o) = l4;
    }
}

template <int BM, int BN, int WGM, int WGN, int NST, bool TRIB, bool OUTP = false,
          bool BR = false>
__device__ __forceinline__ void split_tile_lds(const SplitGemm& g, int mt, int nt, int z,
                                               unsigned char* smem) {
  constexpr int NW = WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN, FM = WM / 16, FN = WN / 16;
  constexpr int A_BYTES = BM * 128, B_BYTES = BN * 128;  // 2 planes x rows x 64 B
  constexpr int STAGE = A_BYTES + B_BYTES;
  constexpr int NA = BM / 8, NB = BN / 8;  // load instructions per k tile (16 rows x 64 B)
  constexpr int LW = (NA + NB) / NW;       // per wavefront
  static_assert(BM % 16 == 0 && BN % 16 == 0 && (NA + NB) % NW == 0, "load split");
  static_assert(NST >= 2 && NST <= 4 && (NST - 2) * LW < 64, "pipeline depth");
  static_assert(FM >= 1 && FN >= 1, "wave tile");
  static_assert(split_lds_bytes<BM, BN, NST>() == NST * STAGE, "LDS size");
  static_assert(!BR || (!TRIB && !OUTP), "batch-reduce tiles: plain products, FP32 output");

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WGM, wn = wave / WGM;
  const int m0 = mt * BM, n0 = nt * BN;
  const int kt0 = TRIB ? n0 / SB_K : 0;
  const int nkt = (g.K + SB_K - 1) / SB_K - kt0;
  // BR: output z sums batches q0 .. q1 - 1, the k tiles of each in turn (T steps in all)
  const int q0 = BR ? z * g.bps : 0, q1 = BR ? min(q0 + g.bps, g.nbatch) : 0;
  const int T = BR ? (q1 > q0 ? (q1 - q0) * nkt : 0) : nkt;

  // this wavefront's load instructions: q = wave + NW i; A for q < NA, else B; each moves
  // the 16-row block grp of one plane
  const _Float16* src[LW];
  int dst[LW];
#pragma unroll
  for (int i = 0; i < LW; ++i) {
    const int q = wave + NW * i;
    const bool isA = q < NA;
    const int qq = isA ? q : q - NA;
    const int rows = isA ? BM : BN;
    const int p = qq / (rows / 16), grp = qq % (rows / 16);
    const int r = 16 * grp + (lane & 15), c = lane >> 4;
    // tile-major planes: the block of 16 rows x 64 bytes is 1 KB contiguous
    if (isA) {
      src[i] = (p ? g.alo : g.ahi) + (long)(m0 + r) * 32 + 8 * c;
    } else {
      const int col = min(n0 + r, g.N - 1);
      src[i] = (p ? g.blo : g.bhi) + (BR ? 0L : (long)z * g.sB) + (long)col * 32 + 8 * c;
    }
    dst[i] = (isA ? 0 : A_BYTES) + p * rows * 64 + grp * 1024;
  }
  auto issue = [&](int t) {
    long kofsA = (long)(kt0 + t) * g.lda, kofsB = (long)(kt0 + t) * g.ldb;
    if constexpr (BR) {
      const int Q = q0 + t / nkt, kt = t % nkt;
      kofsA = (long)Q * g.sA + (long)kt * g.lda;
      kofsB = (long)Q * g.sB + (long)kt * g.ldb;
    }
    unsigned char* base = smem + (t % NST) * STAGE;
#if defined(__HIP_DEVICE_COMPILE__)  // a gfx950 builtin: the host pass only needs the stub
#pragma unroll
    for (int i = 0; i < LW; ++i)
      __builtin_amdgcn_global_load_lds(src[i] + (wave + NW * i < NA ? kofsA : kofsB),
                                       base + dst[i], 16, 0, 0);
#else
    (void)kofsA;
    (void)kofsB;
    (void)base;
#endif
  };

  v4f acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

#pragma unroll
  for (int s = 0; s < NST - 1; ++s)
    if (s < T) issue(s);
  const int fr = lane & 15, fc = lane >> 4;
  for (int t = 0; t < T; ++t) {
    // this wavefront's loads of tile t have landed once at most `ahead` tiles' remain
    const int ahead = min(NST - 2, T - 1 - t);
    if (ahead >= 2) wait_vmcnt<2 * LW>();
    else if (ahead == 1) wait_vmcnt<LW>();
    else wait_vmcnt<0>();
    // Everyone's tile t is in LDS and everyone is done reading tile t - 1.  A bare
    // s_barrier: __syncthreads() would also wait for every outstanding load (vmcnt(0)),
    // draining the tiles in flight; the LDS reads of tile t - 1 are complete already
    // (their values went into the MFMAs), and lgkmcnt(0) makes sure of it.
    __builtin_amdgcn_s_waitcnt((15 & 15) | (3 << 14) | (7 << 4) | (0 << 8));  // lgkmcnt(0)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (t + NST - 1 < T) issue(t + NST - 1);
    const unsigned char* A = smem + (t % NST) * STAGE;
    const unsigned char* B = A + A_BYTES;
    v8h bh[FN], bl[FN], ah[FM], al[FM];
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int o = (wn * WN + j * 16) * 64 + fc * 256 + fr * 16;  // 16-row block, chunk fc
      bh[j] = *reinterpret_cast<const v8h*>(B + o);
      bl[j] = *reinterpret_cast<const v8h*>(B + BN * 64 + o);
    }
    if constexpr (TRIB) {
      // Lh: m = kk + 8 fc + e against column n; only k tiles reaching into this column
      // block's diagonal need it (a uniform branch)
      const int kk = (kt0 + t) * SB_K;
      if (kk < n0 + BN) {
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int d = kk + 8 * fc - (n0 + wn * WN + j * 16 + fr);
          v8h f;
#pragma unroll
          for (int e = 0; e < 8; ++e) {  // bit patterns of 1.0, 0.5, 0.0 in FP16
            const unsigned short bits = d + e > 0 ? 0x3C00 : (d + e == 0 ? 0x3800 : 0);
            f[e] = __builtin_bit_cast(_Float16, bits);
          }
          bh[j] *= f;
          bl[j] *= f;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int o = (wm * WM + i * 16) * 64 + fc * 256 + fr * 16;
      ah[i] = *reinterpret_cast<const v8h*>(A + o);
      al[i] = *reinterpret_cast<const v8h*>(A + BM * 64 + o);
    }
    // TRIB: a fragment column whose first column lies past this k tile's last k is all
    // zero after the mask; its MFMAs are skipped (a wavefront-uniform branch)
    // and fragment rows past M (the padding of nbf5 to the tile) likewise
    bool live_col[FN], live_row[FM];
#pragma unroll
    for (int j = 0; j < FN; ++j)
      live_col[j] = !TRIB || (kt0 + t) * SB_K + SB_K - 1 >= n0 + wn * WN + 16 * j;
#pragma unroll
    for (int i = 0; i < FM; ++i) live_row[i] = m0 + wm * WM + 16 * i < g.M;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        if (live_col[j] && live_row[i])
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        if (live_col[j] && live_row[i])
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        if (live_col[j] && live_row[i])
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
  }
  if constexpr (OUTP) {
    split_epilogue_planes<BM, BN, WGM, WGN, FM, FN>(g, acc, m0, n0, wm, wn, z, lane);
  } else {
    split_epilogue<FM, FN>(g, acc, m0 + wm * WM, n0 + wn * WN, n0, z, lane);
  }
}

template <int BM, int BN, int WGM, int WGN, int NST, bool TRIB, bool OUTP = false,
          bool BR = false>
__global__ __launch_bounds__(64 * WGM * WGN) void split_lds_kernel(SplitGemm g, int tiles_m,
                                                                  int tiles_n) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[split_lds_bytes<BM, BN, NST>()];
  const int nz = BR ? (g.nbatch + g.bps - 1) / g.bps : g.nbatch;
  const unsigned total = (unsigned)tiles_m * tiles_n * nz;
  if (blockIdx.x >= total) return;
  const unsigned lid = xcd_remap16(blockIdx.x, total);
  const int mt = lid % tiles_m;
  const int nt = (lid / tiles_m) % tiles_n;
  const int z = lid / (tiles_m * tiles_n);
  split_tile_lds<BM, BN, WGM, WGN, NST, TRIB, OUTP, BR>(g, mt, nt, z, smem);
}

// Persistent grouped split GEMMs (the schedule of dn_sched.h).
template <int BM, int BN, int WGM, int WGN, int NST>
__global__ __launch_bounds__(64 * WGM * WGN) void split_grouped_lds_kernel(
    const GroupedSplit* __restrict__ desc, int nprob, long total_tiles, long total_slots) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[split_lds_bytes<BM, BN, NST>()];
  grouped_schedule(desc, nprob, total_tiles, total_slots,
                   [&](const GroupedSplit& d, int mt, int nt, int z) {
                     split_tile_lds<BM, BN, WGM, WGN, NST, true>(d.g, mt, nt, z, smem);
                   });
}

template <int BM, int BN, int WGM, int WGN>
__global__ __launch_bounds__(64 * WGM * WGN) void split_grouped_f32a_kernel(
    const GroupedSplit* __restrict__ desc, int nprob, long total_tiles, long total_slots) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[split_tile_bytes<BM, BN>()];
  grouped_schedule(desc, nprob, total_tiles, total_slots,
                   [&](const GroupedSplit& d, int mt, int nt, int z) {
                     split_tile<BM, BN, WGM, WGN, true, false>(d.g, mt, nt, z, smem);
                   });
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

// ---- grouped set-up (all systems of a group per launch) ----------------------------------

// index of the descriptor whose block range holds block b (ranges ascending)
template <bool Q>
__device__ __forceinline__ int find_prep(const SplitPrepDesc* __restrict__ d, int n, long b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((Q ? d[mid].q_begin : d[mid].plane_begin) <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Workgroup i: max |C5| of system i (the elements c_max_kernel reads), the scale pair,
// and the cleared per-Q maxima of the first transform.
__global__ __launch_bounds__(256) void group_c_scale_kernel(const SplitPrepDesc* __restrict__ desc) {
  const SplitPrepDesc d = desc[blockIdx.x];
  __shared__ float red[4];
  const long n = (long)d.nbf * d.nbf5;
  float m = 0.f;
  for (long e = threadIdx.x; e < n; e += blockDim.x) m = fmaxf(m, (float)fabs(d.C[e]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  for (int Q = threadIdx.x; Q < d.naux; Q += blockDim.x) d.qmax[Q] = 0.f;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    const float sc = pow2_scale(mx * 1.0001f);
    d.cscale[2] = mx;
    d.cscale[0] = sc;
    d.cscale[1] = 1.0f / sc;
  }
}

// The C5 planes of every system (c_planes_kernel's element map), 256 elements per block.
__global__ __launch_bounds__(256) void group_c_planes_kernel(const SplitPrepDesc* __restrict__ desc,
                                                             int n) {
  const int i = __builtin_amdgcn_readfirstlane(find_prep<false>(desc, n, blockIdx.x));
  const SplitPrepDesc d = desc[i];
  const long e = (blockIdx.x - d.plane_begin) * 256 + threadIdx.x;
  if (e >= (long)d.Mpad * d.Kpad) return;
  const float sc = d.cscale[0];
  const int kt = e / ((long)d.Mpad * 32), rem = e % ((long)d.Mpad * 32);
  const int p = rem >> 5, k = kt * 32 + (rem & 31);  // tile-major, as c_planes_kernel
  _Float16 h = 0, l = 0;
  if (p < d.nbf5 && k < d.nbf) split_f16_d(d.C[k + (long)p * d.nbf] * (double)sc, &h, &l);
  d.chi[e] = h;
  d.clo[e] = l;
}

// The per-Q scales of every system's first-transform output (bpn_scale_kernel's map).
__global__ __launch_bounds__(256) void group_q_scale_kernel(const SplitPrepDesc* __restrict__ desc,
                                                            int n) {
  const int i = __builtin_amdgcn_readfirstlane(find_prep<true>(desc, n, blockIdx.x));
  const SplitPrepDesc d = desc[i];
  const int Q = (int)((blockIdx.x - d.q_begin) * 256 + threadIdx.x);
  if (Q >= d.naux) return;
  const float s = pow2_scale(d.qmax[Q] * 1.0001f);
  d.qs[Q] = s;
  d.qinv[Q] = 1.0f / s;
}

// ---- host side ---------------------------------------------------------------------------

void split_prep_plan(SplitPrepDesc* desc, int n, long* plane_blocks, long* q_blocks) {
  long pb = 0, qb = 0;
  for (int i = 0; i < n; ++i) {
    desc[i].plane_begin = pb;
    desc[i].q_begin = qb;
    pb += ((long)desc[i].Mpad * desc[i].Kpad + 255) / 256;
    qb += (desc[i].naux + 255) / 256;
  }
  *plane_blocks = pb;
  *q_blocks = qb;
}

int split_group_prep(hipStream_t s, const SplitPrepDesc* dev, int n, long plane_blocks) {
  if (n <= 0) return DN_OK;
  hipLaunchKernelGGL(group_c_scale_kernel, dim3(n), dim3(256), 0, s, dev);
  hipLaunchKernelGGL(group_c_planes_kernel, dim3((unsigned)plane_blocks), dim3(256), 0, s, dev, n);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int split_group_scales(hipStream_t s, const SplitPrepDesc* dev, int n, long q_blocks) {
  if (n <= 0) return DN_OK;
  hipLaunchKernelGGL(group_q_scale_kernel, dim3((unsigned)q_blocks), dim3(256), 0, s, dev, n);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

static unsigned grid_of(long n) {
  long b = (n + 255) / 256;
  return (unsigned)std::max<long>(1, std::min<long>(b, 8192));
}

int split_prepare_b(hipStream_t s, const double* b, int nbf, int ld, int naux, int Kpad,
                    float* scale, float* inv_scale, _Float16* hi, _Float16* lo) {
  const long slab = (long)ld * nbf;  // pad rows hold zeros: they change no maximum
  hipLaunchKernelGGL(b_slab_max_kernel, dim3(naux), dim3(256), 0, s, b, slab, naux, scale,
                     inv_scale);
  DN_HIP(hipGetLastError());
  hipLaunchKernelGGL(b_split_kernel, dim3(grid_of(slab * naux)), dim3(256), 0, s, b, ld, nbf,
                     slab, slab * naux, Kpad, scale, hi, lo);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int split_prepare_c(hipStream_t s, const double* C, int nbf, int nbf5, int Mpad, int Kpad,
                    float* cscale, _Float16* hi, _Float16* lo, const float* lhmax, int naux,
                    float* qs, float* qinv) {
  hipLaunchKernelGGL(fill_f32_kernel, dim3(1), dim3(64), 0, s, cscale + 2, 2L, 0.f);
  hipLaunchKernelGGL(c_stats_kernel, dim3(nbf5), dim3(256), 0, s, C, nbf, cscale);
  const unsigned grid = std::max(grid_of((long)Mpad * Kpad), (unsigned)((naux + 255) / 256));
  hipLaunchKernelGGL(c_planes_kernel, dim3(grid), dim3(256), 0, s, C, nbf, nbf5, Mpad, Kpad,
                     cscale, hi, lo, lhmax, naux, qs, qinv);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// ---- split-precision Lagrangian contraction ----------------------------------------
//
// Z[n,b] = sum_{q,Q} bpn[n,q,Q] bmat[q,b,Q] (utils.jl:44-48: the exchange-type term of the
// orbital Lagrangian, the Coulomb term sitting on bmat's diagonal) on FP16 hi/lo MFMA:
// both operands become tile-major planes with q as the contraction index k (bpn from the
// FP32 first transform, bmat evaluated in FP64 from the FP32 second transform and never
// stored), each under one power-of-two scale taken from its
// largest element over every slab, so a slab's rounding stays below 2^-22 of the largest
// slab's terms, as its share of the sum does; bps slabs accumulate in FP32 per output tile
// and the partials are summed in FP64 in batch order (deterministic).

// Operand planes of the contraction (tile-major, q as the contraction index k): element
// (row r, k) of slab Q at Q Rp Kq + (k / 32) 32 Rp + 32 r + k % 32, zero for k in [K, Kq)
// (and for rows in [R, Rp)).

// Planes of a k-contiguous FP32 source, element (row r, k) of slab Q at
// src[Q sS + r rs + k] for r < R, k < K: thread = (row, 8-wide chunk c) with c fastest (four
// threads read a row's 128-byte run of one k tile); work block = (Q, k tile, 64-row block),
// taken grid-stride.  (One element per thread, 32 threads per run, was
// slower: 289 against 206 us on average over a served mix, its 2-byte stores.)
__global__ __launch_bounds__(256) void kplanes_kc_f32_kernel(const float* __restrict__ src,
                                                             long rs, long sS, int R, int Rp,
                                                             int K, int Kq, long nblk,
                                                             const float* __restrict__ scale,
                                                             _Float16* __restrict__ hi,
                                                             _Float16* __restrict__ lo) {
  const int nkt = Kq / 32, nrb = (Rp + 63) / 64;
  const float sc = *scale;
  for (long b = blockIdx.x; b < nblk; b += gridDim.x) {
    const long qk = b / nrb;
    const int rb = (int)(b - qk * nrb);
    const long Q = qk / nkt;
    const int kt = (int)(qk - Q * nkt);
    const int c = (int)(threadIdx.x & 3), r = rb * 64 + (int)(threadIdx.x >> 2);
    if (r >= Rp) continue;
    const float* row = src + Q * sS + (long)r * rs;
    const int k0 = kt * 32 + 8 * c;
    float x[8];
    if (r < R && k0 + 8 <= K && ((uintptr_t)(row + k0) & 15) == 0) {
      // the whole chunk in range and 16-byte aligned: two vector loads
      typedef float v4 __attribute__((ext_vector_type(4)));
      const v4 a = *reinterpret_cast<const v4*>(row + k0);
      const v4 b = *reinterpret_cast<const v4*>(row + k0 + 4);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        x[e] = a[e];
        x[4 + e] = b[e];
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) x[e] = (r < R && k0 + e < K) ? row[k0 + e] : 0.f;
    }
    v8h h8, l8;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      _Float16 h, l;
      split_f16(x[e] * sc, &h, &l);
      h8[e] = h;
      l8[e] = l;
    }
    const long o = Q * Rp * Kq + (long)kt * 32 * Rp + (long)r * 32 + 8 * c;
    *reinterpret_cast<v8h*>(hi + o) = h8;
    *reinterpret_cast<v8h*>(lo + o) = l8;
  }
}

// The bmat operand of the split contraction without an FP64 bmat (utils.jl:44-48):
//   bmat[q,b,Q] = -ck12[b + nbf5 q] b_pq[b,q,Q] + [q == b] w[b + nbf5 Q],
// evaluated in FP64 from the split second transform's FP32 output (b_pq[b,q,Q] at
// q + nbf5 b + nbf5^2 Q: row b's k = q run is contiguous).  PLANES = false leaves each
// workgroup's max |bmat| (FP32) in bmax[blockIdx.x]; PLANES = true writes the hi/lo planes of
// bmat * scale[0] in the operand layout above (rows b, k = q, Rp = nbf5).  Thread = (row b,
// 8-wide chunk c of one k tile), c fastest; work block = (k tile, 64-row block, range of
// qper slabs), nblk of them taken grid-stride, so a thread's eight ck12 values stay in
// registers over a range.  Both passes evaluate
// the same expression, so the scale bounds exactly the values that are split.
template <bool PLANES>
__global__ __launch_bounds__(256) void bmat_kc_kernel(const float* __restrict__ bpq,
                                                      const double* __restrict__ ck12,
                                                      const double* __restrict__ w, int nbf5,
                                                      int naux, int Kq, int qper, long nblk,
                                                      const float* __restrict__ scale,
                                                      float* __restrict__ bmax,
                                                      _Float16* __restrict__ hi,
                                                      _Float16* __restrict__ lo) {
  __shared__ float red[4];
  const int nkt = Kq / 32, nrb = (nbf5 + 63) / 64;
  const long pairs = (long)nkt * nrb;
  const double sc = PLANES ? (double)*scale : 0.0;
  float m = 0.f;
  for (long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int pair = (int)(blk % pairs);
    const int Q0 = (int)(blk / pairs) * qper;
    const int Q1 = min(Q0 + qper, naux);
    const int kt = pair / nrb, rb = pair - kt * nrb;
    const int c = (int)(threadIdx.x & 3), r = rb * 64 + (int)(threadIdx.x >> 2);
    const int k0 = kt * 32 + 8 * c;
    if (r >= nbf5) continue;
    double ck[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) ck[e] = (k0 + e < nbf5) ? -ck12[r + (long)nbf5 * (k0 + e)] : 0.0;
    const int diag = r - k0;  // chunk position of q == b, if in [0, 8)
    for (int Q = Q0; Q < Q1; ++Q) {
      const float* row = bpq + (long)nbf5 * nbf5 * Q + (long)nbf5 * r;
      float v[8];
      if (k0 + 8 <= nbf5 && ((uintptr_t)(row + k0) & 15) == 0) {
        typedef float v4 __attribute__((ext_vector_type(4)));
        const v4 a = *reinterpret_cast<const v4*>(row + k0);
        const v4 b = *reinterpret_cast<const v4*>(row + k0 + 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          v[e] = a[e];
          v[4 + e] = b[e];
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (k0 + e < nbf5) ? row[k0 + e] : 0.f;
      }
      double x[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        double y = ck[e] * (double)v[e];
        if (e == diag) y += w[r + (long)nbf5 * Q];
        x[e] = y;
      }
      if constexpr (PLANES) {
        v8h h8, l8;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          _Float16 h, l;
          split_f16_d(x[e] * sc, &h, &l);
          h8[e] = h;
          l8[e] = l;
        }
        const long o = (long)Q * nbf5 * Kq + (long)kt * 32 * nbf5 + (long)r * 32 + 8 * c;
        *reinterpret_cast<v8h*>(hi + o) = h8;
        *reinterpret_cast<v8h*>(lo + o) = l8;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) m = fmaxf(m, (float)fabs(x[e]));  // the scale takes 1.0001x
      }
    }
  }
  if constexpr (!PLANES) {
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
      bmax[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  }
}

// zs = (scale of bpn, 1 / (scale of bpn * scale of bmat), scale of bmat, -): the bpn scale
// from the first transform's per-Q maxima, the bmat scale from the nb block maxima that
// bmat_kc_kernel<false> left in zs[4 ..]
__global__ __launch_bounds__(256) void contract_scales_kernel(const float* __restrict__ qmax,
                                                              int naux, int nb,
                                                              float* __restrict__ zs) {
  __shared__ float red[2][4];
  float ma = 0.f, mb = 0.f;
  for (int Q = threadIdx.x; Q < naux; Q += blockDim.x) ma = fmaxf(ma, qmax[Q]);
  for (int i = threadIdx.x; i < nb; i += blockDim.x) mb = fmaxf(mb, zs[4 + i]);
  for (int off = 32; off > 0; off >>= 1) {
    ma = fmaxf(ma, __shfl_down(ma, off, 64));
    mb = fmaxf(mb, __shfl_down(mb, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = ma;
    red[1][threadIdx.x >> 6] = mb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const float a = fmaxf(fmaxf(red[0][0], red[0][1]), fmaxf(red[0][2], red[0][3]));
    const float b = fmaxf(fmaxf(red[1][0], red[1][1]), fmaxf(red[1][2], red[1][3]));
    const float sa = pow2_scale(a * 1.0001f), sb = pow2_scale(b * 1.0001f);
    zs[0] = sa;
    zs[1] = 1.0f / (sa * sb);  // powers of two: exact
    zs[2] = sb;
  }
}

// out[e] = sum_z part[z n + e] in FP64, z in order
__global__ __launch_bounds__(256) void reduce_f32_partials_kernel(const float* __restrict__ part,
                                                                  long n, int nz,
                                                                  double* __restrict__ out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    int z = 0;
    for (; z + 8 <= nz; z += 8) {  // eight loads in flight, added in z order
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = part[(long)(z + u) * n + e];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += (double)v[u];
    }
    for (; z < nz; ++z) acc += (double)part[(long)z * n + e];
    out[e] = acc;
  }
}

namespace {
struct ContractPlan {
  int Rp, Kq, tm, tn, bps, nz;
  long planeA, planeB, partial;  // halves per plane of each operand; partial floats
};

ContractPlan contract_plan(int nbf, int nbf5, int naux) {
  ContractPlan c{};
  c.Rp = (nbf + 127) / 128 * 128;
  c.Kq = (nbf5 + 31) / 32 * 32;
  c.tm = c.Rp / 128;
  c.tn = (nbf5 + 127) / 128;
  // about 2048 workgroups, 4 to 16 slabs per FP32 partial: at least four keeps the
  // partials (and the workspace) small against the planes, at most sixteen bounds the FP32
  // accumulation to 16 Kq terms per output element
  const long want = std::max(1L, 2048L / (long)(c.tm * c.tn));
  c.bps = (int)std::min(16L, std::max(4L, (naux + want - 1) / want));
  c.bps = std::min(c.bps, naux);
  c.nz = (naux + c.bps - 1) / c.bps;
  c.planeA = (long)c.Rp * c.Kq * naux;
  c.planeB = (long)nbf5 * c.Kq * naux;
  c.partial = (long)c.nz * nbf * nbf5;
  return c;
}
}  // namespace

size_t split_contract_bytes(int nbf, int nbf5, int naux) {
  const ContractPlan c = contract_plan(nbf, nbf5, naux);
  return (size_t)(2 * c.planeA + 2 * c.planeB) * sizeof(_Float16) +
         (size_t)c.partial * sizeof(float) + 256;
}

int split_contract(hipStream_t s, int nbf, int nbf5, int naux, const float* bpn,
                   const float* bpq, const double* ck12, const double* w, const float* qmax,
                   float* zscale, void* work, size_t work_bytes, double* Z) {
  const ContractPlan c = contract_plan(nbf, nbf5, naux);
  if (split_contract_bytes(nbf, nbf5, naux) > work_bytes)
    return fail(DN_ERR_ARG, "split contraction workspace too small");
  // bmat work blocks: (k tile, 64-row block) pairs times ranges of qper slabs, about 8192
  // of them; at most 8192 workgroups (the max pass leaves one maximum per workgroup in
  // zscale[4 ..])
  const long pairs = (long)(c.Kq / 32) * ((nbf5 + 63) / 64);
  const long ranges = std::max(1L, std::min((long)naux, 8192 / pairs));
  const int qper = (int)((naux + ranges - 1) / ranges);
  const long nblkB = pairs * ((naux + qper - 1) / qper);
  const int nbmax = (int)std::min(nblkB, 8192L);
  _Float16* ahi = static_cast<_Float16*>(work);
  _Float16* alo = ahi + c.planeA;
  _Float16* bhi = alo + c.planeA;
  _Float16* blo = bhi + c.planeB;
  float* part = reinterpret_cast<float*>(blo + c.planeB);  // 64-byte aligned: Kq % 32 == 0
  hipLaunchKernelGGL(bmat_kc_kernel<false>, dim3((unsigned)nbmax), dim3(256), 0, s, bpq, ck12,
                     w, nbf5, naux, c.Kq, qper, nblkB, nullptr, zscale + 4, nullptr, nullptr);
  hipLaunchKernelGGL(contract_scales_kernel, dim3(1), dim3(256), 0, s, qmax, naux, nbmax,
                     zscale);
  // bpn[n,q,Q] at q + nbf5 n + nbf5 nbf Q (rows n, k = q contiguous)
  const long blkA = (long)naux * (c.Kq / 32) * ((c.Rp + 63) / 64);
  hipLaunchKernelGGL(kplanes_kc_f32_kernel, dim3((unsigned)std::min(blkA, 1L << 20)), dim3(256),
                     0, s, bpn, (long)nbf5, (long)nbf5 * nbf, nbf, c.Rp, nbf5, c.Kq, blkA, zscale,
                     ahi, alo);
  // bmat (rows b, k = q) evaluated again, now into planes under its scale
  hipLaunchKernelGGL(bmat_kc_kernel<true>, dim3((unsigned)nbmax), dim3(256), 0, s, bpq, ck12, w,
                     nbf5, naux, c.Kq, qper, nblkB, zscale + 2, nullptr, bhi, blo);
  SplitGemm g{};
  g.M = nbf; g.N = nbf5; g.K = nbf5;
  g.ahi = ahi; g.alo = alo; g.lda = (long)c.Rp * 32; g.sA = (long)c.Rp * c.Kq;
  g.bhi = bhi; g.blo = blo; g.ldb = (long)nbf5 * 32; g.sB = (long)nbf5 * c.Kq;
  g.C = part; g.ldc = nbf; g.sC = (long)nbf * nbf5;
  g.cs = zscale;  // cs[1] = 1 / (scale of bpn * scale of bmat)
  g.col_div = 0;
  g.nbatch = naux; g.bps = c.bps;
  const long total = (long)c.tm * c.tn * c.nz;
  if (total > 0x7fffffffL) return fail(DN_ERR_ARG, "split contraction grid too large");
  hipLaunchKernelGGL((split_lds_kernel<128, 128, 2, 4, 2, false, false, true>),
                     dim3((unsigned)total), dim3(512), 0, s, g, c.tm, c.tn);
  const long n = (long)nbf * nbf5;
  hipLaunchKernelGGL(reduce_f32_partials_kernel, dim3(grid_of(n)), dim3(256), 0, s, part, n,
                     c.nz, Z);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int split_lhnorm(hipStream_t s, const double* b, int nbf, int ld, int naux, float* lhmax) {
  hipLaunchKernelGGL(b_lhnorm_kernel, dim3(naux), dim3(256), 0, s, b, nbf, ld, lhmax);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// The symmetric-b split build with T as FP16 planes (one system):
//   K1: T[p,n,Q] = sum_{m>=n} C5[m,p] Lh[m,n,Q] written as hi/lo planes of T * qs[Q], the
//       scale set before T exists from ||C5[:,p]|| ||Lh_Q[:,n]|| >= |T[p,n]|;
//   K2: W^T[q,p,Q] = sum_n C5[n,q] T[p,n,Q], both operands planes (the LDS-direct tile),
//       out = acc * cs[1] * qinv[Q].
// W^T instead of W: K = sum_Q (W + W^T)^2 and X = 2 diag W read either.  tp holds the two
// planes (Mpad x Kpad halves per Q each).  part 1 launches K1, part 2 K2; split_prepare_c
// sets qs, qinv first.
int split_sym_planes(hipStream_t s, int nbf, int nbf5, int naux, int Mpad, int Kpad,
                     const _Float16* bhi, const _Float16* blo, const float* b_inv_scale,
                     const _Float16* chi, const _Float16* clo, const float* cscale_dev,
                     const float* qs, const float* qinv, _Float16* tp, float* wt, int part) {
  const long pl = (long)Mpad * Kpad * naux;
  SplitGemm g1{};
  g1.M = nbf5; g1.N = nbf; g1.K = nbf;
  g1.ahi = chi; g1.alo = clo; g1.lda = (long)Mpad * 32;
  g1.bhi = bhi; g1.blo = blo; g1.ldb = (long)nbf * 32; g1.sB = (long)nbf * Kpad;
  g1.cs = cscale_dev; g1.batch_inv_scale = b_inv_scale; g1.nbatch = naux;
  g1.ohi = tp; g1.olo = tp + pl; g1.out_ld = (long)Mpad * 32; g1.out_sb = (long)Mpad * Kpad;
  g1.out_kt = Kpad / 32; g1.out_scale = qs;
  if (part == 1) {
    const int tm = (g1.M + 127) / 128, tn = (g1.N + 127) / 128;
    const long total = (long)tm * tn * naux;
    if (total > 0x7fffffffL || tm * 128 > Mpad) return fail(DN_ERR_ARG, "split K1 grid");
    // eight wavefronts (2 x 4, 64 x 32 each): four per SIMD at two workgroups per CU
    // hide more of the load latency than four (2.51-2.59 against 2.61-2.76 ms at
    // 800/3200/240; the same MFMAs in the same order, so bitwise the same T)
    hipLaunchKernelGGL((split_lds_kernel<128, 128, 2, 4, 2, true, true>), dim3((unsigned)total),
                       dim3(512), 0, s, g1, tm, tn);
    DN_HIP(hipGetLastError());
    return DN_OK;
  }
  SplitGemm g2{};
  g2.M = nbf5; g2.N = nbf5; g2.K = nbf;
  g2.ahi = chi; g2.alo = clo; g2.lda = (long)Mpad * 32;
  g2.bhi = tp; g2.blo = tp + pl; g2.ldb = (long)Mpad * 32; g2.sB = (long)Mpad * Kpad;
  g2.C = wt; g2.ldc = nbf5; g2.sC = (long)nbf5 * nbf5;
  g2.cs = cscale_dev; g2.batch_inv_scale = qinv; g2.nbatch = naux;
  if (nbf5 > 128 && Mpad == 256) {
    // one 256 x 256 tile per slab (nbf5 in (128, 256]): T_Q and the C5 planes pass through
    // LDS once per slab instead of twice (1.13-1.21 against 1.23-1.30 ms at 800/3200/240,
    // bitwise the same W); the A rows read stay below Mpad
    if ((long)naux > 0x7fffffffL) return fail(DN_ERR_ARG, "split K2 grid");
    hipLaunchKernelGGL((split_lds_kernel<256, 256, 4, 2, 2, false>), dim3((unsigned)naux),
                       dim3(512), 0, s, g2, 1, 1);
    DN_HIP(hipGetLastError());
    return DN_OK;
  }
  const int tm = (g2.M + 127) / 128, tn = (g2.N + 127) / 128;
  const long total = (long)tm * tn * naux;
  if (total > 0x7fffffffL) return fail(DN_ERR_ARG, "split K2 grid");
  hipLaunchKernelGGL((split_lds_kernel<128, 128, 2, 2, 2, false>), dim3((unsigned)total),
                     dim3(256), 0, s, g2, tm, tn);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

template <int BM, int BN, int WGM, int WGN, bool AF32, bool TRIB = false>
static int launch_split(hipStream_t s, const SplitGemm& g) {
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  const long total = (long)tm * tn * g.nbatch;
  if (total > 0x7fffffffL) return fail(DN_ERR_ARG, "split gemm grid too large");
  hipLaunchKernelGGL((split_gemm_kernel<BM, BN, WGM, WGN, AF32, TRIB>), dim3((unsigned)total),
                     dim3(64 * WGM * WGN), 0, s, g, tm, tn);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// The split-precision J/K build: C planes must be prepared (split_prepare_c) for the
// current orbitals; bpn (nbf5*nbf*naux floats), bpq (nbf5*nbf5*naux floats), qmax,
// qs/qinv (naux floats each) are workspaces.  The planes are tile-major: b per slab with
// nbf columns (k tile stride nbf * 32, slab stride nbf * Kpad), C5^T with Mpad rows (k tile
// stride Mpad * 32); Mpad and Kpad are multiples of 128 and 32.
// General b: bpn[p,n,Q] = sum_m C5[m,p] b[m,n,Q], batched over Q (the LDS-direct tile).
int split_k1(hipStream_t s, int nbf, int nbf5, int naux, int Mpad, int Kpad,
             const _Float16* bhi, const _Float16* blo, const float* b_inv_scale,
             const _Float16* chi, const _Float16* clo, const float* cscale_dev, float* bpn,
             float* qmax) {
  hipLaunchKernelGGL(fill_f32_kernel, dim3(grid_of(naux)), dim3(256), 0, s, qmax, (long)naux, 0.f);
  DN_HIP(hipGetLastError());
  SplitGemm g1{};
  g1.M = nbf5; g1.N = nbf; g1.K = nbf;
  g1.ahi = chi; g1.alo = clo; g1.lda = (long)Mpad * 32;
  g1.bhi = bhi; g1.blo = blo; g1.ldb = (long)nbf * 32; g1.sB = (long)nbf * Kpad;
  g1.C = bpn; g1.ldc = nbf5; g1.sC = (long)nbf5 * nbf;
  g1.cs = cscale_dev; g1.batch_inv_scale = b_inv_scale; g1.col_div = 0;
  g1.qmax = qmax; g1.nbatch = naux;
  const int tm = (g1.M + 127) / 128, tn = (g1.N + 127) / 128;
  const long total = (long)tm * tn * g1.nbatch;
  if (total > 0x7fffffffL) return fail(DN_ERR_ARG, "split gemm grid too large");
  // eight wavefronts per tile (2 x 4, 64 x 32 each): 3.90 against 4.00 ms with four at
  // 800/3200/240, bitwise the same bpn and per-Q maxima
  hipLaunchKernelGGL((split_lds_kernel<128, 128, 2, 4, 2, false>), dim3((unsigned)total),
                     dim3(512), 0, s, g1, tm, tn);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// Symmetric b: bpn[p,n,Q] = sum_{m >= n} C5[m,p] Lh[m,n,Q] (b_Q = Lh + Lh^T), batched
// over Q so that column block n0 reads only rows m >= n0 of the slab; the second
// transform then yields W with b_pq = W + W^T.  split_k1_begin clears the per-Q maxima
// first.
SplitGemm split_k1_sym_gemm(int nbf, int nbf5, int naux, int Mpad, int Kpad,
                            const _Float16* bhi, const _Float16* blo, const float* b_inv_scale,
                            const _Float16* chi, const _Float16* clo, const float* cscale_dev,
                            float* bpn, float* qmax) {
  SplitGemm g1{};
  g1.M = nbf5; g1.N = nbf; g1.K = nbf;
  g1.ahi = chi; g1.alo = clo; g1.lda = (long)Mpad * 32;
  g1.bhi = bhi; g1.blo = blo; g1.ldb = (long)nbf * 32; g1.sB = (long)nbf * Kpad;
  g1.C = bpn; g1.ldc = nbf5; g1.sC = (long)nbf5 * nbf;
  g1.cs = cscale_dev; g1.batch_inv_scale = b_inv_scale; g1.col_div = 0;
  g1.qmax = qmax; g1.nbatch = naux;
  return g1;
}

int split_k1_begin(hipStream_t s, float* qmax, int naux) {
  hipLaunchKernelGGL(fill_f32_kernel, dim3(grid_of(naux)), dim3(256), 0, s, qmax, (long)naux, 0.f);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int split_k1_sym(hipStream_t s, int nbf, int nbf5, int naux, int Mpad, int Kpad,
                 const _Float16* bhi, const _Float16* blo, const float* b_inv_scale,
                 const _Float16* chi, const _Float16* clo, const float* cscale_dev, float* bpn,
                 float* qmax) {
  DN_TRY(split_k1_begin(s, qmax, naux));
  const SplitGemm g1 = split_k1_sym_gemm(nbf, nbf5, naux, Mpad, Kpad, bhi, blo, b_inv_scale, chi,
                                         clo, cscale_dev, bpn, qmax);
  const int tm = (g1.M + 127) / 128, tn = (g1.N + 127) / 128;
  const long total = (long)tm * tn * g1.nbatch;
  if (total > 0x7fffffffL) return fail(DN_ERR_ARG, "split gemm grid too large");
  hipLaunchKernelGGL((split_lds_kernel<128, 128, 2, 4, 2, true>), dim3((unsigned)total), dim3(512),
                     0, s, g1, tm, tn);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// bpq[:,:,Q] = bpn[:,:,Q] C5 with bpn split on load under the per-Q scale that
// split_k2_begin derives from the first transform's maxima.
SplitGemm split_k2_gemm(int nbf, int nbf5, int naux, int Mpad, const _Float16* chi,
                        const _Float16* clo, const float* cscale_dev, const float* bpn,
                        float* bpq, const float* qs, const float* qinv) {
  SplitGemm g2{};
  g2.M = nbf5; g2.N = nbf5; g2.K = nbf;
  g2.af32 = bpn; g2.lda = nbf5; g2.sA = (long)nbf5 * nbf; g2.a_scale = qs;
  g2.bhi = chi; g2.blo = clo; g2.ldb = (long)Mpad * 32;
  g2.C = bpq; g2.ldc = nbf5; g2.sC = (long)nbf5 * nbf5;
  g2.cs = cscale_dev; g2.batch_inv_scale = qinv; g2.col_div = 0;
  g2.nbatch = naux;
  return g2;
}

int split_k2_begin(hipStream_t s, const float* qmax, int naux, float* qs, float* qinv) {
  hipLaunchKernelGGL(bpn_scale_kernel, dim3((naux + 255) / 256), dim3(256), 0, s, qmax, naux, qs,
                     qinv);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int split_k2(hipStream_t s, int nbf, int nbf5, int naux, int Mpad, const _Float16* chi,
             const _Float16* clo, const float* cscale_dev, const float* bpn, float* bpq,
             const float* qmax, float* qs, float* qinv) {
  DN_TRY(split_k2_begin(s, qmax, naux, qs, qinv));
  const SplitGemm g = split_k2_gemm(nbf, nbf5, naux, Mpad, chi, clo, cscale_dev, bpn, bpq, qs, qinv);
  // nbf5 in (128, 256]: one 256-wide N tile reads each FP32 slab of T once (1.37 against
  // 1.47 ms at 800/3200/240 on MI355X, same sums)
  if (nbf5 > 128 && nbf5 <= 256) return launch_split<128, 256, 2, 4, true>(s, g);
  return launch_split<128, 128, 2, 2, true>(s, g);
}

// ---- grouped launches ------------------------------------------------------------------

namespace {

void split_config_tile(int config, int* bm, int* bn) {
  switch (config) {
    case kSplit64x128: *bm = 64; *bn = 128; break;
    case kSplit64x64: *bm = 64; *bn = 64; break;
    case kSplit32x64: *bm = 32; *bn = 64; break;
    default: *bm = 128; *bn = 128; break;
  }
}

template <int BM, int BN>
int launch_split_grouped(hipStream_t s, int kind, const GroupedSplit* d, int n, long total,
                         long slots, int grid) {
  if (kind == 0)
    return launch_persistent(s, split_grouped_lds_kernel<BM, BN, 2, 2, 2>, 256, d, n, total,
                             slots, grid);
  if constexpr (BM % 64 == 0)
    return launch_persistent(s, split_grouped_f32a_kernel<BM, BN, 2, 2>, 256, d, n, total,
                             slots, grid);
  return fail(DN_ERR_ARG, "grouped split GEMM: configuration not available for this stage");
}

}  // namespace

bool split_config_ok(int config, int kind) {
  if (config < 0 || config >= kNumSplitConfigs) return false;
  return kind == 0 || config != kSplit32x64;  // FP32 A staging covers 64-row groups
}

long split_grouped_plan(GroupedSplit* desc, int n, int config) {
  int bm, bn;
  split_config_tile(config, &bm, &bn);
  for (int i = 0; i < n; ++i) desc[i].nsplit = desc[i].g.nbatch;
  return grouped_offsets(desc, n, bm, bn);
}

// kind 0: symmetric-b first transforms (split_k1_sym_gemm), kind 1: second transforms
// (split_k2_gemm).  The per-system set-up kernels (C planes, split_k1_begin,
// split_k2_begin) run before the respective launch.
int split_grouped(hipStream_t s, const GroupedSplit* dh, const GroupedSplit* dd, int n,
                  long total, int config, int kind, int grid) {
  if (n <= 0 || total <= 0) return DN_OK;
  if (!split_config_ok(config, kind)) return fail(DN_ERR_ARG, "grouped split GEMM: bad configuration");
  const long slots = grouped_slots(dh, n);
  switch (config) {
    case kSplit64x128: return launch_split_grouped<64, 128>(s, kind, dd, n, total, slots, grid);
    case kSplit64x64: return launch_split_grouped<64, 64>(s, kind, dd, n, total, slots, grid);
    case kSplit32x64: return launch_split_grouped<32, 64>(s, kind, dd, n, total, slots, grid);
    default: return launch_split_grouped<128, 128>(s, kind, dd, n, total, slots, grid);
  }
}

}  // namespace dn
