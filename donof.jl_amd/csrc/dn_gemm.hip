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

// Logical tile id for the hardware workgroup id `hw` so that consecutive logical ids
// run on the same XCD (the dispatcher deals workgroups round-robin over 8 XCDs).
__device__ __forceinline__ unsigned xcd_remap(unsigned hw, unsigned total) {
  const unsigned q = total / 8u, r = total % 8u;
  const unsigned x = hw % 8u, i = hw / 8u;
  return x * q + (x < r ? x : r) + i;
}

typedef double v2d __attribute__((ext_vector_type(2)));

// LDS layouts (doubles):
//   K-contiguous operand  -> [mn][BK+1]   rows of one m (or n), odd stride
//   MN-contiguous operand -> [k][BMN+16]  rows of one k
// Both make the 16-lane groups of an MFMA fragment read hit distinct banks, and the
// global->LDS copy keeps consecutive threads on consecutive addresses.
template <int BMN, int BK, bool KC, int NT>
struct TileLds {
  static constexpr int LD = KC ? (BK + 1) : (BMN + 16);
  static constexpr int SIZE = KC ? BMN * (BK + 1) : BK * (BMN + 16);
  static constexpr int PAIRS = BMN * BK / 2 / NT;  // 16-byte loads per thread
  __device__ static int off(int mn, int k) { return KC ? mn * LD + k : k * LD + mn; }
  // (mn, k) of the first element of pair idx; pairs run along the contiguous dimension
  __device__ static void coord(int idx, int* mn, int* k) {
    if (KC) { *k = 2 * (idx % (BK / 2)); *mn = idx / (BK / 2); }
    else { *mn = 2 * (idx % (BMN / 2)); *k = idx / (BMN / 2); }
  }
};

// Two consecutive elements along the contiguous dimension from an address clamped into
// the operand (always in bounds); the caller zeroes the out-of-range ones with
// pair_mask() when it writes the tile to LDS, i.e. after the MFMAs of the current tile,
// so the loads stay in flight meanwhile.  With VEC the pair is one 16-byte load; it
// requires an even extent along the contiguous dimension, an even leading dimension and
// a 16-byte aligned base.  (A run-time VEC switch would make the compiler wait for every
// load at the control-flow merge.)
template <bool KC, bool VEC>
__device__ __forceinline__ v2d load_pair(const double* __restrict__ base, long ld, int gmn,
                                         int gk, int MN, int KK) {
  v2d v;
  if constexpr (VEC) {
    const int cmn = min(gmn, MN - (KC ? 1 : 2));
    const int ck = min(gk, KK - (KC ? 2 : 1));
    const double* p = KC ? base + (long)cmn * ld + ck : base + cmn + (long)ck * ld;
    v = *reinterpret_cast<const v2d*>(p);
  } else {
    const int cmn0 = min(gmn, MN - 1), ck0 = min(gk, KK - 1);
    const int gmn1 = KC ? gmn : gmn + 1, gk1 = KC ? gk + 1 : gk;
    const int cmn1 = min(gmn1, MN - 1), ck1 = min(gk1, KK - 1);
    v.x = KC ? base[(long)cmn0 * ld + ck0] : base[cmn0 + (long)ck0 * ld];
    v.y = KC ? base[(long)cmn1 * ld + ck1] : base[cmn1 + (long)ck1 * ld];
  }
  return v;
}

template <bool KC>
__device__ __forceinline__ v2d pair_mask(v2d v, int gmn, int gk, int MN, int KK) {
  const int gmn1 = KC ? gmn : gmn + 1, gk1 = KC ? gk + 1 : gk;
  v.x = (gmn < MN && gk < KK) ? v.x : 0.0;
  v.y = (gmn1 < MN && gk1 < KK) ? v.y : 0.0;
  return v;
}

// One output tile (mt, nt) of split z of a GEMM, by a workgroup of WGM x WGN waves; each
// wave owns (BM/WGM) x (BN/WGN).  As/Bs are the workgroup's LDS double buffers.
// TRI = kTriB (K-contiguous B): B is symmetric, B(k, n) = B(n, k), and the tile
// multiplies by its lower half Lh with B = Lh + Lh^T, i.e. Lh(k, n) = B(k, n) for k > n,
// B(n, n)/2 for k = n and 0 for k < n; the K loop of column block n0 starts at k = n0.
// TRI = kTriA (K-contiguous A): the same for A, A(m, k) masked by k > m / k == m, and
// the K loop of row block m0 starts at k = m0.
template <int BM, int BN, int BK, int WGM, int WGN, bool AK, bool BKC, bool VA, bool VB,
          int TRI = 0>
__device__ __forceinline__ void gemm_tile(const Gemm& g, int mt, int nt, int z, double* As_,
                                          double* Bs_) {
  constexpr int NT = 64 * WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int FM = WM / 16, FN = WN / 16;
  using LA = TileLds<BM, BK, AK, NT>;
  using LB = TileLds<BN, BK, BKC, NT>;
  static_assert(FM >= 1 && FN >= 1 && WM % 16 == 0 && WN % 16 == 0, "wave tile");
  static_assert(TRI != kTriB || BKC, "the symmetric-B mode needs a K-contiguous B");
  static_assert(TRI != kTriA || AK, "the symmetric-A mode needs a K-contiguous A");
  static_assert(BM * BK % (2 * NT) == 0 && BN * BK % (2 * NT) == 0, "tile/threads");
  double (*As)[LA::SIZE] = reinterpret_cast<double (*)[LA::SIZE]>(As_);
  double (*Bs)[LB::SIZE] = reinterpret_cast<double (*)[LB::SIZE]>(Bs_);

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WGM, wn = wave / WGM;
  const int m0 = mt * BM, n0 = nt * BN;
  const int M = g.M, N = g.N, K = g.K;

  const int r0 = z * g.bps;
  const int r1 = min(r0 + g.bps, g.nbatch);
  const int kt0 = TRI == kTriB ? n0 / BK : (TRI == kTriA ? m0 / BK : 0);
  const int nkt = (K + BK - 1) / BK - kt0;
  const int T = nkt > 0 ? (r1 - r0) * nkt : 0;

  v4d acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  v2d ra[LA::PAIRS], rb[LB::PAIRS];

  auto load_tile = [&](int t) {
    const int r = r0 + t / nkt;
    const int k0 = (kt0 + t % nkt) * BK;
    const double* A = g.A + (long)r * g.sA;
    const double* B = g.B + (long)r * g.sB;
#pragma unroll
    for (int i = 0; i < LA::PAIRS; ++i) {
      int mn, k;
      LA::coord(tid + i * NT, &mn, &k);
      ra[i] = load_pair<AK, VA>(A, g.lda, m0 + mn, k0 + k, M, K);
    }
#pragma unroll
    for (int i = 0; i < LB::PAIRS; ++i) {
      int mn, k;
      LB::coord(tid + i * NT, &mn, &k);
      rb[i] = load_pair<BKC, VB>(B, g.ldb, n0 + mn, k0 + k, N, K);
    }
  };
  auto store_tile = [&](int buf, int t) {
    const int k0 = (kt0 + t % nkt) * BK;
#pragma unroll
    for (int i = 0; i < LA::PAIRS; ++i) {
      int mn, k;
      LA::coord(tid + i * NT, &mn, &k);
      ra[i] = pair_mask<AK>(ra[i], m0 + mn, k0 + k, M, K);
      if constexpr (TRI == kTriA) {
        const int gm = m0 + mn, gk = k0 + k;  // pairs run along k
        ra[i].x *= gk > gm ? 1.0 : (gk == gm ? 0.5 : 0.0);
        ra[i].y *= gk + 1 > gm ? 1.0 : (gk + 1 == gm ? 0.5 : 0.0);
      }
      if (AK) {
        As[buf][LA::off(mn, k)] = ra[i].x;
        As[buf][LA::off(mn, k + 1)] = ra[i].y;
      } else {
        *reinterpret_cast<v2d*>(&As[buf][LA::off(mn, k)]) = ra[i];
      }
    }
#pragma unroll
    for (int i = 0; i < LB::PAIRS; ++i) {
      int mn, k;
      LB::coord(tid + i * NT, &mn, &k);
      rb[i] = pair_mask<BKC>(rb[i], n0 + mn, k0 + k, N, K);
      if constexpr (TRI == kTriB) {
        const int gn = n0 + mn, gk = k0 + k;  // pairs run along k
        rb[i].x *= gk > gn ? 1.0 : (gk == gn ? 0.5 : 0.0);
        rb[i].y *= gk + 1 > gn ? 1.0 : (gk + 1 == gn ? 0.5 : 0.0);
      }
      if (BKC) {
        Bs[buf][LB::off(mn, k)] = rb[i].x;
        Bs[buf][LB::off(mn, k + 1)] = rb[i].y;
      } else {
        *reinterpret_cast<v2d*>(&Bs[buf][LB::off(mn, k)]) = rb[i];
      }
    }
  };

  if (T > 0) {
    load_tile(0);
    store_tile(0, 0);
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
      for (int i = 0; i < FM; ++i) a[i] = As[buf][LA::off(wm * WM + i * 16 + fr, kk + fk)];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = Bs[buf][LB::off(wn * WN + j * 16 + fr, kk + fk)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
    if (t + 1 < T) store_tile(buf ^ 1, t + 1);
    __syncthreads();
  }

  // D(16x16) layout of v_mfma_f64_16x16x4f64 (measured on gfx950): lane l holds
  // column l%16 and rows l/16 + 4*r in accumulator register r.  The product is formed
  // transposed (B fragment first), so a lane holds row l%16 of the C tile and columns
  // l/16 + 4*r: one store instruction (fixed r) writes 16 consecutive rows (128 bytes)
  // of each of 4 columns.
  double* Cz = g.C + (long)z * g.sC;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int row = m0 + wm * WM + i * 16 + fr;
    if (row >= M) continue;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int col = n0 + wn * WN + j * 16 + fk + 4 * rr;
        if (col < N) {
          double* c = Cz + row + (long)col * g.ldc;
          const double v = g.alpha * acc[i][j][rr];
          *c = (g.beta == 0.0) ? v : v + g.beta * (*c);
        }
      }
  }
}

template <int BM, int BN, int BK, int WGM, int WGN, bool AK, bool BKC>
struct GemmLds {
  static constexpr int A = TileLds<BM, BK, AK, 64 * WGM * WGN>::SIZE;
  static constexpr int B = TileLds<BN, BK, BKC, 64 * WGM * WGN>::SIZE;
};

template <int BM, int BN, int BK, int WGM, int WGN, bool AK, bool BKC, bool VA, bool VB,
          int TRI = 0>
__global__ __launch_bounds__(64 * WGM * WGN) void dgemm_mfma_kernel(Gemm g, int tiles_m,
                                                                   int tiles_n, int nsplit) {
  using L = GemmLds<BM, BN, BK, WGM, WGN, AK, BKC>;
  __shared__ double As[2 * L::A];
  __shared__ double Bs[2 * L::B];
  const unsigned total = (unsigned)tiles_m * tiles_n * nsplit;
  const unsigned hw = blockIdx.x;
  if (hw >= total) return;
  const unsigned lid = xcd_remap(hw, total);
  const int mt = lid % tiles_m;
  const int nt = (lid / tiles_m) % tiles_n;
  const int z = lid / (tiles_m * tiles_n);
  gemm_tile<BM, BN, BK, WGM, WGN, AK, BKC, VA, VB, TRI>(g, mt, nt, z, As, Bs);
}

// Persistent grouped GEMM: the tiles of all problems form one list (problem i owns tiles
// [desc[i].tile_begin, desc[i].tile_begin + tiles_m*tiles_n*nsplit)); workgroup b takes
// list positions b, b + gridDim.x, ... until the list is exhausted (a static schedule, so
// every workgroup's loop ends).  With a grid that is a multiple of the 8 XCDs, workgroup
// b runs on XCD b % 8, and position t maps to tile (t % 8) * per + t / 8: each XCD works
// through one contiguous eighth of the list, so the M-tiles that share a B block (and
// consecutive B blocks) meet in the same L2.
template <int BM, int BN, int BK, int WGM, int WGN, bool AK, bool BKC, bool VA, bool VB,
          int TRI = 0>
__global__ __launch_bounds__(64 * WGM * WGN) void dgemm_grouped_kernel(const GroupedGemm* __restrict__ desc,
                                                                      int nprob,
                                                                      long total_tiles) {
  using L = GemmLds<BM, BN, BK, WGM, WGN, AK, BKC>;
  __shared__ double As[2 * L::A];
  __shared__ double Bs[2 * L::B];
  const bool chunk = (gridDim.x % 8u) == 0;
  const long per = (total_tiles + 7) / 8;
  const long span = chunk ? 8 * per : total_tiles;
  for (long pos = blockIdx.x; pos < span; pos += gridDim.x) {
    const long t = chunk ? (pos % 8) * per + pos / 8 : pos;
    if (t >= total_tiles) continue;
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (desc[mid].tile_begin <= t) lo = mid;
      else hi = mid - 1;
    }
    // lo is uniform over the workgroup; saying so makes the descriptor loads scalar (SGPRs)
    // instead of per-lane vector loads.  A private copy: the tile's stores to C cannot
    // alias it, so it is not reloaded after every store.
    lo = __builtin_amdgcn_readfirstlane(lo);
    const GroupedGemm d = desc[lo];
    const long local = t - d.tile_begin;
    const int mt = local % d.tiles_m;
    const int nt = (local / d.tiles_m) % d.tiles_n;
    const int z = local / ((long)d.tiles_m * d.tiles_n);
    gemm_tile<BM, BN, BK, WGM, WGN, AK, BKC, VA, VB, TRI>(d.g, mt, nt, z, As, Bs);
    __syncthreads();  // the next tile reuses the LDS buffers
  }
}

__global__ void reduce_splits_kernel(const double* __restrict__ P, long sP, int nsplit, int M,
                                     int N, long ld, double* __restrict__ out, long ldo,
                                     double beta, double scale) {
  const long total = (long)M * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int m = e % M, n = e / M;
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += P[z * sP + m + n * ld];
    double* o = out + m + n * ldo;
    s *= scale;
    *o = (beta == 0.0) ? s : s + beta * (*o);
  }
}

template <int BM, int BN, int BK, int WGM, int WGN, bool AK, bool BKC, bool VA, bool VB,
          int TRI = 0>
int launch(hipStream_t s, const Gemm& g) {
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  const int nsplit = (g.nbatch + g.bps - 1) / g.bps;
  const long total = (long)tm * tn * nsplit;
  if (total <= 0) return DN_OK;
  if (total > 0x7fffffffL) return fail(DN_ERR_ARG, "gemm grid too large");
  hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, BK, WGM, WGN, AK, BKC, VA, VB, TRI>),
                     dim3((unsigned)total), dim3(64 * WGM * WGN), 0, s, g, tm, tn, nsplit);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

template <int BM, int BN, int BK, int WGM, int WGN, bool AK, bool BKC, int TRI = 0>
int dispatch_vec(hipStream_t s, const Gemm& g) {
  if (g.vec_a) {
    return g.vec_b ? launch<BM, BN, BK, WGM, WGN, AK, BKC, true, true, TRI>(s, g)
                   : launch<BM, BN, BK, WGM, WGN, AK, BKC, true, false, TRI>(s, g);
  }
  return g.vec_b ? launch<BM, BN, BK, WGM, WGN, AK, BKC, false, true, TRI>(s, g)
                 : launch<BM, BN, BK, WGM, WGN, AK, BKC, false, false, TRI>(s, g);
}

template <int BM, int BN, int BK, int WGM, int WGN>
int dispatch_layout(hipStream_t s, const Gemm& g) {
  if (g.a_kcontig) {
    return g.b_kcontig ? dispatch_vec<BM, BN, BK, WGM, WGN, true, true>(s, g)
                       : dispatch_vec<BM, BN, BK, WGM, WGN, true, false>(s, g);
  }
  return g.b_kcontig ? dispatch_vec<BM, BN, BK, WGM, WGN, false, true>(s, g)
                     : dispatch_vec<BM, BN, BK, WGM, WGN, false, false>(s, g);
}

}  // namespace

int gemm_f64(hipStream_t s, const Gemm& g_in) {
  if (g_in.M <= 0 || g_in.N <= 0) return DN_OK;
  if (g_in.K <= 0 || g_in.nbatch <= 0 || g_in.bps <= 0)
    return fail(DN_ERR_ARG, "gemm_f64: bad K/nbatch/bps");
  Gemm g = g_in;
  // 16-byte loads need an even leading dimension and batch stride and an aligned base
  auto aligned = [](const double* p, long ld, long sb, int nb, int extent) {
    return ((uintptr_t)p % 16 == 0) && (ld % 2 == 0) && (nb <= 1 || sb % 2 == 0) &&
           (extent % 2 == 0) && extent >= 2;
  };
  g.vec_a = aligned(g.A, g.lda, g.sA, g.nbatch, g.a_kcontig ? g.K : g.M);
  g.vec_b = aligned(g.B, g.ldb, g.sB, g.nbatch, g.b_kcontig ? g.K : g.N);
  if (g.tri == kTriA) {
    if (!g.a_kcontig || !g.b_kcontig || g.M != g.K)
      return fail(DN_ERR_ARG, "gemm_f64: symmetric-A mode needs K-contiguous A and B, M == K");
    // 64-row blocks (the K loop of row block m0 covers k >= m0 only): the short, uneven
    // K loops balance best over small workgroups (measured on MI355X at nbf = 400, 800)
    return dispatch_vec<64, 128, 16, 2, 4, true, true, kTriA>(s, g);
  }
  if (g.tri) {
    if (!g.a_kcontig || !g.b_kcontig || g.N != g.K)
      return fail(DN_ERR_ARG, "gemm_f64: symmetric-B mode needs K-contiguous A and B, N == K");
    // 64 x 64 tiles: the K loop of column block n0 covers k >= n0 only, and the short,
    // uneven K loops balance best over many small workgroups (measured on MI355X at
    // nbf = 400 and 800 against 128 x 64 and 128 x 128 tiles)
    return dispatch_vec<64, 64, 16, 2, 2, true, true, kTriB>(s, g);
  }
  // Operands larger than 64 in both M and N get the 128 x 128 tile (the RI transforms
  // have M = nbf5, N = nbf5 or nbf); everything else uses 64 x 64.
  if (g.M > 64 && g.N > 64) return dispatch_layout<128, 128, 16, 2, 4>(s, g);
  return dispatch_layout<64, 64, 16, 2, 2>(s, g);
}

namespace {

bool vec_ok(const double* p, long ld, long sb, int nb, int extent) {
  return ((uintptr_t)p % 16 == 0) && (ld % 2 == 0) && (nb <= 1 || sb % 2 == 0) &&
         (extent % 2 == 0) && extent >= 2;
}

void config_tile(int config, int* bm, int* bn) {
  switch (config) {
    case kTile64x64: *bm = 64; *bn = 64; break;
    case kTile64x128: *bm = 64; *bn = 128; break;
    case kTile128x64: *bm = 128; *bn = 64; break;
    case kTile64x32: *bm = 64; *bn = 32; break;
    case kTile32x64: *bm = 32; *bn = 64; break;
    default: *bm = 128; *bn = 128; break;
  }
}

template <int BM, int BN, int BK, int WGM, int WGN, bool AK, bool BKC, bool VA, bool VB,
          int TRI = 0>
int launch_grouped(hipStream_t s, const GroupedGemm* d, int n, long total, int grid) {
  auto kern = dgemm_grouped_kernel<BM, BN, BK, WGM, WGN, AK, BKC, VA, VB, TRI>;
  // grid <= 0: -grid times as many workgroups as stay resident (0 = once), a multiple
  // of the 8 XCDs, at most one per tile (rounded up to the XCD multiple)
  if (grid <= 0) {
    int dev = 0, cus = 0, occ = 0;
    DN_HIP(hipGetDevice(&dev));
    DN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    DN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 64 * WGM * WGN, 0));
    const long full = (long)std::max(occ, 1) * cus * std::max(-grid, 1);
    grid = (int)std::min(std::min(full, (total + 7) / 8 * 8), 0x7ffffff8L);
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * WGM * WGN), 0, s, d, n, total);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

template <int BM, int BN, int BK, int WGM, int WGN>
int grouped_layout(hipStream_t s, bool ak, bool bk, bool va, bool vb, int tri,
                   const GroupedGemm* d, int n, long total, int grid) {
#define DN_G(AKv, BKv, VAv, VBv) \
  launch_grouped<BM, BN, BK, WGM, WGN, AKv, BKv, VAv, VBv>(s, d, n, total, grid)
#define DN_GT(VAv, VBv, T) \
  launch_grouped<BM, BN, BK, WGM, WGN, true, true, VAv, VBv, T>(s, d, n, total, grid)
  if (tri) {
    if (!ak || !bk || tri != kTriA)
      return fail(DN_ERR_ARG, "grouped GEMM: symmetric mode is kTriA with K-contiguous A, B");
    return va ? (vb ? DN_GT(true, true, kTriA) : DN_GT(true, false, kTriA))
              : (vb ? DN_GT(false, true, kTriA) : DN_GT(false, false, kTriA));
  }
#undef DN_GT
#define DN_GV(AKv, BKv) \
  (va ? (vb ? DN_G(AKv, BKv, true, true) : DN_G(AKv, BKv, true, false)) \
      : (vb ? DN_G(AKv, BKv, false, true) : DN_G(AKv, BKv, false, false)))
  // the layouts of the J/K build's stages: (A, B) K-contiguous (K1, flattened K2),
  // (MN, K) (batched K2, H C5) and (MN, MN) (J = X X^T)
  if (ak) {
    if (!bk) return fail(DN_ERR_ARG, "grouped GEMM: layout (K-contiguous A, MN B) unsupported");
    return DN_GV(true, true);
  }
  return bk ? DN_GV(false, true) : DN_GV(false, false);
#undef DN_GV
#undef DN_G
}

}  // namespace

long gemm_grouped_plan(GroupedGemm* desc, int n, int config) {
  int bm, bn;
  config_tile(config, &bm, &bn);
  long begin = 0;
  for (int i = 0; i < n; ++i) {
    Gemm& g = desc[i].g;
    g.vec_a = vec_ok(g.A, g.lda, g.sA, g.nbatch, g.a_kcontig ? g.K : g.M);
    g.vec_b = vec_ok(g.B, g.ldb, g.sB, g.nbatch, g.b_kcontig ? g.K : g.N);
    desc[i].tiles_m = (g.M + bm - 1) / bm;
    desc[i].tiles_n = (g.N + bn - 1) / bn;
    desc[i].nsplit = (g.nbatch + g.bps - 1) / g.bps;
    desc[i].tile_begin = begin;
    begin += (long)desc[i].tiles_m * desc[i].tiles_n * desc[i].nsplit;
  }
  return begin;
}

int gemm_f64_grouped(hipStream_t s, const GroupedGemm* dh, const GroupedGemm* dd, int n,
                     long total, int config, int grid) {
  if (n <= 0 || total <= 0) return DN_OK;
  const bool ak = dh[0].g.a_kcontig, bk = dh[0].g.b_kcontig;
  const int tri = dh[0].g.tri;
  bool va = true, vb = true;
  for (int i = 0; i < n; ++i) {
    if (dh[i].g.a_kcontig != ak || dh[i].g.b_kcontig != bk || dh[i].g.tri != tri)
      return fail(DN_ERR_ARG, "grouped GEMM problems must share operand layouts");
    va = va && dh[i].g.vec_a;
    vb = vb && dh[i].g.vec_b;
  }
  switch (config) {
    case kTile64x64: return grouped_layout<64, 64, 16, 2, 2>(s, ak, bk, va, vb, tri, dd, n, total, grid);
    case kTile64x128: return grouped_layout<64, 128, 16, 2, 4>(s, ak, bk, va, vb, tri, dd, n, total, grid);
    case kTile128x64: return grouped_layout<128, 64, 16, 2, 2>(s, ak, bk, va, vb, tri, dd, n, total, grid);
    case kTile64x32: return grouped_layout<64, 32, 16, 2, 1>(s, ak, bk, va, vb, tri, dd, n, total, grid);
    case kTile32x64: return grouped_layout<32, 64, 16, 1, 2>(s, ak, bk, va, vb, tri, dd, n, total, grid);
    default: return grouped_layout<128, 128, 16, 2, 4>(s, ak, bk, va, vb, tri, dd, n, total, grid);
  }
}

int gemm_f64_splitk(hipStream_t s, const Gemm& g, double* part, size_t part_elems) {
  if (g.nbatch != 1) return gemm_f64(s, g);
  const long tiles = (long)((g.M + 63) / 64) * ((g.N + 63) / 64);
  // largest split count (<= 16, dividing K, chunks >= 32) that keeps ~1024 workgroups
  int best = 1;
  for (int d = 2; d <= 16; ++d) {
    if (g.K % d || g.K / d < 32) continue;
    if ((size_t)d * g.M * g.N > part_elems) break;
    if (tiles * d > 2048) break;
    best = d;
  }
  if (best == 1 || g.beta != 0.0) return gemm_f64(s, g);
  const int kc = g.K / best;
  Gemm h = g;
  h.K = kc;
  h.sA = (long)kc * (g.a_kcontig ? 1 : g.lda);
  h.sB = (long)kc * (g.b_kcontig ? 1 : g.ldb);
  h.nbatch = best;
  h.bps = 1;
  h.C = part;
  h.ldc = g.M;
  h.sC = (long)g.M * g.N;
  h.alpha = 1.0;
  h.beta = 0.0;
  DN_TRY(gemm_f64(s, h));
  return reduce_splits(s, part, (long)g.M * g.N, best, g.M, g.N, g.M, g.C, g.ldc, 0.0, g.alpha);
}

int reduce_splits(hipStream_t s, const double* P, long sP, int nsplit, int M, int N, long ld,
                  double* out, long ldo, double beta, double scale) {
  const long total = (long)M * N;
  if (total <= 0) return DN_OK;
  const int threads = 256;
  const long blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(reduce_splits_kernel, dim3((unsigned)(blocks < 4096 ? blocks : 4096)),
                     dim3(threads), 0, s, P, sP, nsplit, M, N, ld, out, ldo, beta, scale);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

}  // namespace dn
