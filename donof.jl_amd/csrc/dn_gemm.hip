// FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64), LDS double-buffered,
// with strided batches, batch-reduce splits and an XCD-aware tile order.
//
// Every dense contraction of the hot path is one of these calls: the first RI
// transform b_pn = C5^T b (one GEMM with N = nbf*naux), the per-Q second transform,
// the Lagrangian's C^K/C^J contraction (batch-reduce over Q), H C5, C^T H C5 and the
// matrix products of the orbital-rotation exponential.
#include "dn_internal.h"
#include "dn_sched.h"

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
  // fragments dealt round-robin (fragment row i of wavefront row wm: tile rows
  // 16 (i WGM + wm) ..; columns likewise) and those wholly past M or N skipped
  // (wavefront-uniform): a tile cut short by the matrix edge keeps its SIMDs evenly busy
  bool live_row[FM], live_col[FN];
#pragma unroll
  for (int i = 0; i < FM; ++i) live_row[i] = m0 + 16 * (i * WGM + wm) < M;
#pragma unroll
  for (int j = 0; j < FN; ++j) live_col[j] = n0 + 16 * (j * WGN + wn) < N;
  for (int t = 0; t < T; ++t) {
    const int buf = t & 1;
    if (t + 1 < T) load_tile(t + 1);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = As[buf][LA::off(16 * (i * WGM + wm) + fr, kk + fk)];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = Bs[buf][LB::off(16 * (j * WGN + wn) + fr, kk + fk)];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        if (!live_row[i]) continue;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          if (!live_col[j]) continue;
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
        }
      }
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
    const int row = m0 + 16 * (i * WGM + wm) + fr;
    if (row >= M) continue;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int col = n0 + 16 * (j * WGN + wn) + fk + 4 * rr;
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

// ---- LDS-direct tile (K-contiguous A and B) -------------------------------------------
//
// Every 16-wide k tile of A and B goes from HBM straight into LDS
// (global_load_lds_dwordx4: lane l writes 16 bytes at the instruction's LDS base + 16 l),
// NST - 1 tiles ahead of the one being multiplied: no staging registers, no VALU copy, and
// the loads of several tiles in flight.  One barrier per k tile.
//
// LDS layout of one operand tile: blocks of 16 rows (2048 bytes); in a block, half h
// (doubles 8h..8h+7 of the k tile) at h * 1024, 16-byte chunk c (doubles 2c, 2c+1) of
// that half at (c % 4) * 256, row r at (r % 16) * 16.  One load instruction fills one half
// of one block (lane l: row l % 16, chunk l / 16, 64 contiguous global bytes per row).
//
// Fragments are dealt to the wavefronts round-robin: fragment row i of wavefront row wm
// covers tile rows 16 (i WGM + wm) .. + 15, fragment column j of wavefront column wn tile
// columns 16 (j WGN + wn) .. + 15.  A tile cut short by M or N (whose fragments past the
// edge are skipped) then leaves every wavefront, and so every SIMD, a similar share.
//
// MFMA k order: in step s (0..3) of a k tile, the lanes of fragment group g = lane / 16
// contract k = 4 g + s (A and B alike, so the sum is the same; only its order differs
// from the register-staged tile).  A lane's four k values are two consecutive chunks, read
// with two 16-byte LDS loads per fragment; the 16 lanes of a group read 256 contiguous
// bytes, free of bank conflicts.
//
// Rows past M (columns past N) load a clamped row and are not stored; chunks past K load a
// clamped chunk and their A values are zeroed (K is even, so a chunk is wholly in or out).

// s_waitcnt vmcnt(N) with no wait on the other counters (gfx9 encoding: vmcnt in bits
// [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8])
template <int N>
__device__ __forceinline__ void wait_vmcnt_only() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

template <int BM, int BN>
constexpr int lds_tile_bytes(int nst) {
  return nst * (BM + BN) * 128;
}

template <int BM, int BN, int WGM, int WGN, int NST, int TRI>
__device__ __forceinline__ void gemm_tile_lds(const Gemm& g, int mt, int nt, int z,
                                              unsigned char* smem) {
  constexpr int NW = WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN, FM = WM / 16, FN = WN / 16;
  constexpr int A_BYTES = BM * 128;
  constexpr int STAGE = (BM + BN) * 128;
  constexpr int NA = BM / 8, NB = BN / 8;  // load instructions per k tile
  constexpr int LW = (NA + NB) / NW;       // per wavefront
  static_assert(BM % 16 == 0 && BN % 16 == 0 && (NA + NB) % NW == 0, "load split");
  static_assert(FM >= 1 && FN >= 1 && WM % 16 == 0 && WN % 16 == 0, "wave tile");
  static_assert(NST >= 2 && NST <= 4 && (NST - 2) * LW < 64, "pipeline depth");
  static_assert(TRI == kTriNone || TRI == kTriA, "LDS-direct tile: symmetric-A mode only");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WGM, wn = wave / WGM;
  const int m0 = mt * BM, n0 = nt * BN;
  const int M = g.M, N = g.N, K = g.K;
  const int r0 = z * g.bps, r1 = min(r0 + g.bps, g.nbatch);
  const int kt0 = TRI == kTriA ? m0 / 16 : 0;
  const int nkt = (K + 15) / 16 - kt0;
  const int T = nkt > 0 ? (r1 - r0) * nkt : 0;

  // this wavefront's load instructions: q = wave + NW i, A for q < NA, else B
  const double* src[LW];
  long sb[LW];
  int ck[LW], dst[LW];
#pragma unroll
  for (int i = 0; i < LW; ++i) {
    const int q = wave + NW * i;
    const bool isA = q < NA;
    const int qq = isA ? q : q - NA;
    const int blk = qq >> 1, h = qq & 1;
    const int r = 16 * blk + (lane & 15);
    src[i] = isA ? g.A + (long)min(m0 + r, M - 1) * g.lda
                 : g.B + (long)min(n0 + r, N - 1) * g.ldb;
    sb[i] = isA ? g.sA : g.sB;
    ck[i] = 2 * (4 * h + (lane >> 4));
    dst[i] = (isA ? 0 : A_BYTES) + blk * 2048 + h * 1024;
  }
  int ir = r0, ikt = 0;  // batch and k tile of the next load
  auto issue = [&](int t) {
    const int k0 = (kt0 + ikt) * 16;
    unsigned char* base = smem + (t % NST) * STAGE;
#if defined(__HIP_DEVICE_COMPILE__)  // a gfx950 builtin: the host pass only needs the stub
#pragma unroll
    for (int i = 0; i < LW; ++i) {
      const int k = min(k0 + ck[i], K - 2);
      __builtin_amdgcn_global_load_lds(src[i] + (long)ir * sb[i] + k, base + dst[i], 16, 0, 0);
    }
#else
    (void)k0;
    (void)base;
#endif
    if (++ikt == nkt) {
      ikt = 0;
      ++ir;
    }
  };

  v4d acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

#pragma unroll
  for (int s = 0; s < NST - 1; ++s)
    if (s < T) issue(s);
  const int fr = lane & 15, fk = lane >> 4;
  int ckt = 0;  // k tile being multiplied
  for (int t = 0; t < T; ++t) {
    // this wavefront's loads of tile t have landed once at most `ahead` tiles' remain
    const int ahead = min(NST - 2, T - 1 - t);
    if (NST >= 4 && ahead >= 2) wait_vmcnt_only<(NST >= 4 ? 2 * LW : 0)>();
    else if (NST >= 3 && ahead == 1) wait_vmcnt_only<(NST >= 3 ? LW : 0)>();
    else wait_vmcnt_only<0>();
    // Everyone's tile t is in LDS and everyone is done reading tile t - 1 (whose stage the
    // next load overwrites).  A bare s_barrier: __syncthreads() would also wait for every
    // outstanding load, draining the tiles in flight; lgkmcnt(0) completes this
    // wavefront's LDS reads of tile t - 1.
    __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4) | (0 << 8));
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (t + NST - 1 < T) issue(t + NST - 1);
    const unsigned char* As = smem + (t % NST) * STAGE;
    const unsigned char* Bs = As + A_BYTES;
    v2d af[FM][2], bf[FN][2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int c = 2 * fk + p;
      const int o = (c >> 2) * 1024 + (c & 3) * 256 + fr * 16;
#pragma unroll
      for (int i = 0; i < FM; ++i)
        af[i][p] = *reinterpret_cast<const v2d*>(As + 16 * (i * WGM + wm) * 128 + o);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        bf[j][p] = *reinterpret_cast<const v2d*>(Bs + 16 * (j * WGN + wn) * 128 + o);
    }
    const int k0 = (kt0 + ckt) * 16;
    if (++ckt == nkt) ckt = 0;
    const int kl = k0 + 4 * fk;  // k of step s for this lane: kl + s
    if constexpr (TRI == kTriA) {
      // Lh(m, k): 1 for k > m, 1/2 for k == m, 0 below; only k tiles reaching into this
      // row block's diagonal need it (a uniform branch)
      if (k0 < m0 + BM) {
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          const int d = kl - (m0 + 16 * (i * WGM + wm) + fr);
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            af[i][p].x *= d + 2 * p > 0 ? 1.0 : (d + 2 * p == 0 ? 0.5 : 0.0);
            af[i][p].y *= d + 2 * p + 1 > 0 ? 1.0 : (d + 2 * p + 1 == 0 ? 0.5 : 0.0);
          }
        }
      }
    }
    if (k0 + 16 > K) {  // chunks past K hold clamped copies
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          af[i][p].x = kl + 2 * p < K ? af[i][p].x : 0.0;
          af[i][p].y = kl + 2 * p + 1 < K ? af[i][p].y : 0.0;
        }
    }
    // MFMAs of fragments that contribute nothing are skipped (wavefront-uniform branches):
    // fragment rows past M and columns past N (clamped copies, never stored), and in
    // symmetric-A mode rows whose first row lies past this k tile's last k (all zero
    // after the mask)
    bool live_row[FM], live_col[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
      live_row[i] = (TRI != kTriA || k0 + 15 >= m0 + 16 * (i * WGM + wm)) &&
                    m0 + 16 * (i * WGM + wm) < M;
#pragma unroll
    for (int j = 0; j < FN; ++j) live_col[j] = n0 + 16 * (j * WGN + wn) < N;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        if (!live_row[i]) continue;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          if (!live_col[j]) continue;
          const double a = (s & 1) ? af[i][s >> 1].y : af[i][s >> 1].x;
          const double b = (s & 1) ? bf[j][s >> 1].y : bf[j][s >> 1].x;
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, acc[i][j], 0, 0, 0);
        }
      }
  }

  // the accumulator layout of gemm_tile, fragments placed round-robin as above
  double* Cz = g.C + (long)z * g.sC;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int row = m0 + 16 * (i * WGM + wm) + fr;
    if (row >= M) continue;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int col = n0 + 16 * (j * WGN + wn) + fk + 4 * rr;
        if (col < N) {
          double* c = Cz + row + (long)col * g.ldc;
          const double v = g.alpha * acc[i][j][rr];
          *c = (g.beta == 0.0) ? v : v + g.beta * (*c);
        }
      }
  }
}

template <int BM, int BN, int WGM, int WGN, int NST, int TRI>
__global__ __launch_bounds__(64 * WGM * WGN) void dgemm_lds_kernel(Gemm g, int tiles_m,
                                                                  int tiles_n, int nsplit) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[lds_tile_bytes<BM, BN>(NST)];
  const unsigned total = (unsigned)tiles_m * tiles_n * nsplit;
  const unsigned hw = blockIdx.x;
  if (hw >= total) return;
  int mt, nt, z;
  const unsigned pairs = (unsigned)tiles_m * nsplit;
  if (TRI == kTriA && pairs % 8u == 0) {
    // row block mt runs k tiles from its own row on: dispatch the longest first, so the
    // run ends on short tiles.  Work items are the (mt, z) pairs in order of mt, dealt
    // round-robin to the XCDs (workgroup hw runs on XCD hw % 8); the tiles_n tiles of a
    // pair, which read the same rows of b, run back to back on one XCD and share its L2.
    // (Blocks of one slab read disjoint rows of it: no reuse is lost between pairs.)
    const unsigned k = hw / 8u, x = hw % 8u;
    const unsigned p = 8u * (k / (unsigned)tiles_n) + x;
    nt = k % (unsigned)tiles_n;
    mt = p / (unsigned)nsplit;
    z = p % (unsigned)nsplit;
  } else {
    const unsigned lid = xcd_remap(hw, total);
    mt = lid % tiles_m;
    nt = (lid / tiles_m) % tiles_n;
    z = lid / (tiles_m * tiles_n);
  }
  gemm_tile_lds<BM, BN, WGM, WGN, NST, TRI>(g, mt, nt, z, smem);
}

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

// Persistent grouped GEMMs (the schedule of dn_sched.h).
template <int BM, int BN, int BK, int WGM, int WGN, bool AK, bool BKC, bool VA, bool VB,
          int TRI = 0>
__global__ __launch_bounds__(64 * WGM * WGN) void dgemm_grouped_kernel(
    const GroupedGemm* __restrict__ desc, int nprob, long total_tiles, long total_slots) {
  using L = GemmLds<BM, BN, BK, WGM, WGN, AK, BKC>;
  __shared__ double As[2 * L::A];
  __shared__ double Bs[2 * L::B];
  grouped_schedule(desc, nprob, total_tiles, total_slots,
                   [&](const GroupedGemm& d, int mt, int nt, int z) {
                     gemm_tile<BM, BN, BK, WGM, WGN, AK, BKC, VA, VB, TRI>(d.g, mt, nt, z, As,
                                                                          Bs);
                   });
}

template <int BM, int BN, int WGM, int WGN, int NST, int TRI>
__global__ __launch_bounds__(64 * WGM * WGN) void dgemm_grouped_lds_kernel(
    const GroupedGemm* __restrict__ desc, int nprob, long total_tiles, long total_slots) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[lds_tile_bytes<BM, BN>(NST)];
  grouped_schedule(desc, nprob, total_tiles, total_slots,
                   [&](const GroupedGemm& d, int mt, int nt, int z) {
                     gemm_tile_lds<BM, BN, WGM, WGN, NST, TRI>(d.g, mt, nt, z, smem);
                   });
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

template <int BM, int BN, int WGM, int WGN, int NST, int TRI>
int launch_lds(hipStream_t s, const Gemm& g) {
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  const int nsplit = (g.nbatch + g.bps - 1) / g.bps;
  const long total = (long)tm * tn * nsplit;
  if (total <= 0) return DN_OK;
  if (total > 0x7fffffffL) return fail(DN_ERR_ARG, "gemm grid too large");
  hipLaunchKernelGGL((dgemm_lds_kernel<BM, BN, WGM, WGN, NST, TRI>), dim3((unsigned)total),
                     dim3(64 * WGM * WGN), 0, s, g, tm, tn, nsplit);
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
  // (K may exceed M by zero pad rows of A's slabs)
  if (g.tri == kTriA && (!g.a_kcontig || !g.b_kcontig || g.M > g.K))
    return fail(DN_ERR_ARG, "gemm_f64: symmetric-A mode needs K-contiguous A and B, M <= K");
  if (g.lds && g.a_kcontig && g.b_kcontig && g.vec_a && g.vec_b && g.tri != kTriB) {
    // 8 wavefronts per workgroup (2 per SIMD) beat 4 with twice the wave tile on every
    // transform shape measured on MI355X; 64-row blocks for the symmetric-A mode (short,
    // uneven K loops), 128 x 256 tiles for long N (the flattened second transform)
    // (also for row counts whose last 128-row block would be less than half used)
    if (g.tri == kTriA || g.M <= 64 || (g.M % 128 != 0 && g.M % 128 <= 64)) {
      if (lds64_config(g.N) == kLds64x192)  // one 192-wide N tile instead of two
        return g.tri == kTriA ? launch_lds<64, 192, 2, 4, 2, kTriA>(s, g)
                              : launch_lds<64, 192, 2, 4, 2, kTriNone>(s, g);
      // symmetric-A mode: a two-stage ring (three workgroups per CU instead of two) hides
      // the short K loops' latency better (0.72-0.76 against 0.75-0.78 ms at 400/1600/120)
      return g.tri == kTriA ? launch_lds<64, 128, 2, 4, 2, kTriA>(s, g)
                            : launch_lds<64, 128, 2, 4, 3, kTriNone>(s, g);
    }
    // (4 x 2 wavefronts of 32 x 128 for the long-N tile: 352-365 against 361-375 us for
    // 2 x 4 on the flattened second transform at 400/1600/120)
    return g.N > 128 ? launch_lds<128, 256, 4, 2, 2, kTriNone>(s, g)
                     : launch_lds<128, 128, 2, 4, 3, kTriNone>(s, g);
  }
  if (g.tri == kTriA) {
    // 64-row blocks (the K loop of row block m0 covers k >= m0 only): the short, uneven
    // K loops balance best over small workgroups (measured on MI355X at nbf = 400, 800)
    if (g.N > 128 && g.N <= 192) return dispatch_vec<64, 192, 16, 2, 4, true, true, kTriA>(s, g);
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
  switch (reg_config(g.M, g.N, (g.nbatch + g.bps - 1) / g.bps)) {
    case kTile64x192: return dispatch_layout<64, 192, 16, 2, 4>(s, g);
    case kTile64x128: return dispatch_layout<64, 128, 16, 2, 4>(s, g);
    case kTile128x128: return dispatch_layout<128, 128, 16, 2, 4>(s, g);
    default: return dispatch_layout<64, 64, 16, 2, 2>(s, g);
  }
}

int reg_config(int M, int N, long nsplit) {
  if (M <= 64 || N <= 64) return kTile64x64;
  // the shapes fitted to the edges only where the launch has tiles to spare: a GEMM with
  // few output tiles (split-K, small squares) needs the parallelism of smaller tiles
  constexpr long kMany = 1024;
  const long rows64 = (M + 63) / 64;
  if (N > 128 && N <= 192 && rows64 * nsplit >= kMany) return kTile64x192;
  if (M % 128 != 0 && M % 128 <= 64 && rows64 * ((N + 127) / 128) * nsplit >= kMany)
    return kTile64x128;
  // fewer 128 x 128 tiles than two per CU: four times as many 64 x 64 ones
  if ((long)((M + 127) / 128) * ((N + 127) / 128) * nsplit < 512) return kTile64x64;
  return kTile128x128;
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
    case kLds64x128: *bm = 64; *bn = 128; break;
    case kLds128x256: *bm = 128; *bn = 256; break;
    case kLds64x192: *bm = 64; *bn = 192; break;
    case kTile64x192: *bm = 64; *bn = 192; break;
    default: *bm = 128; *bn = 128; break;  // kTile128x128, kLds128x128
  }
}

template <int BM, int BN, int BK, int WGM, int WGN, bool AK, bool BKC, bool VA, bool VB,
          int TRI = 0>
int launch_grouped(hipStream_t s, const GroupedGemm* d, int n, long total, long slots,
                   int grid) {
  return launch_persistent(s, dgemm_grouped_kernel<BM, BN, BK, WGM, WGN, AK, BKC, VA, VB, TRI>,
                           64 * WGM * WGN, d, n, total, slots, grid);
}

template <int BM, int BN, int WGM, int WGN, int NST>
int launch_grouped_lds(hipStream_t s, int tri, const GroupedGemm* d, int n, long total,
                       long slots, int grid) {
  return tri == kTriA
             ? launch_persistent(s, dgemm_grouped_lds_kernel<BM, BN, WGM, WGN, NST, kTriA>,
                                 64 * WGM * WGN, d, n, total, slots, grid)
             : launch_persistent(s, dgemm_grouped_lds_kernel<BM, BN, WGM, WGN, NST, kTriNone>,
                                 64 * WGM * WGN, d, n, total, slots, grid);
}

template <int BM, int BN, int BK, int WGM, int WGN>
int grouped_layout(hipStream_t s, bool ak, bool bk, bool va, bool vb, int tri,
                   const GroupedGemm* d, int n, long total, long slots, int grid) {
#define DN_G(AKv, BKv, VAv, VBv) \
  launch_grouped<BM, BN, BK, WGM, WGN, AKv, BKv, VAv, VBv>(s, d, n, total, slots, grid)
#define DN_GT(VAv, VBv, T) \
  launch_grouped<BM, BN, BK, WGM, WGN, true, true, VAv, VBv, T>(s, d, n, total, slots, grid)
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
  for (int i = 0; i < n; ++i) {
    Gemm& g = desc[i].g;
    g.vec_a = vec_ok(g.A, g.lda, g.sA, g.nbatch, g.a_kcontig ? g.K : g.M);
    g.vec_b = vec_ok(g.B, g.ldb, g.sB, g.nbatch, g.b_kcontig ? g.K : g.N);
    desc[i].nsplit = (g.nbatch + g.bps - 1) / g.bps;
  }
  return grouped_offsets(desc, n, bm, bn);
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
  const long slots = grouped_slots(dh, n);
  if (is_lds_config(config)) {
    if (!ak || !bk || !va || !vb || tri == kTriB)
      return fail(DN_ERR_ARG, "grouped GEMM: LDS-direct configurations need K-contiguous, "
                              "16-byte aligned A and B");
    switch (config) {
      case kLds64x128:  // symmetric-A mode on a two-stage ring, as in gemm_f64
        return tri == kTriA ? launch_grouped_lds<64, 128, 2, 4, 2>(s, tri, dd, n, total, slots, grid)
                            : launch_grouped_lds<64, 128, 2, 4, 3>(s, tri, dd, n, total, slots, grid);
      case kLds128x128: return launch_grouped_lds<128, 128, 2, 4, 3>(s, tri, dd, n, total, slots, grid);
      case kLds64x192: return launch_grouped_lds<64, 192, 2, 4, 2>(s, tri, dd, n, total, slots, grid);
      default: return launch_grouped_lds<128, 256, 2, 4, 2>(s, tri, dd, n, total, slots, grid);
    }
  }
  switch (config) {
    case kTile64x64: return grouped_layout<64, 64, 16, 2, 2>(s, ak, bk, va, vb, tri, dd, n, total, slots, grid);
    case kTile64x128: return grouped_layout<64, 128, 16, 2, 4>(s, ak, bk, va, vb, tri, dd, n, total, slots, grid);
    case kTile128x64: return grouped_layout<128, 64, 16, 2, 2>(s, ak, bk, va, vb, tri, dd, n, total, slots, grid);
    case kTile64x32: return grouped_layout<64, 32, 16, 2, 1>(s, ak, bk, va, vb, tri, dd, n, total, slots, grid);
    case kTile32x64: return grouped_layout<32, 64, 16, 1, 2>(s, ak, bk, va, vb, tri, dd, n, total, slots, grid);
    case kTile64x192: return grouped_layout<64, 192, 16, 2, 4>(s, ak, bk, va, vb, tri, dd, n, total, slots, grid);
    default: return grouped_layout<128, 128, 16, 2, 4>(s, ak, bk, va, vb, tri, dd, n, total, slots, grid);
  }
}

int lds64_config(int N) {
  // 192 columns: one N tile for N in (128, 192]; for long N (flattened transforms) the
  // wider tile's operand reuse wins (measured on MI355X)
  return N > 128 && (N <= 192 || N >= 1024) ? kLds64x192 : kLds64x128;
}

bool gemm_config_ok(int config, const Gemm& g) {
  if (config < 0 || config >= kNumGemmConfigs) return false;
  if (!is_lds_config(config)) return g.tri == kTriNone || g.tri == kTriA;
  return g.a_kcontig && g.b_kcontig && g.tri != kTriB &&
         vec_ok(g.A, g.lda, g.sA, g.nbatch, g.K) && vec_ok(g.B, g.ldb, g.sB, g.nbatch, g.K);
}

int splitk_count(const Gemm& g, size_t part_elems) {
  if (g.nbatch != 1 || g.beta != 0.0) return 1;
  const long tiles = (long)((g.M + 63) / 64) * ((g.N + 63) / 64);
  // largest split count (<= 16, dividing K, chunks >= 32) that keeps ~1024 workgroups
  int best = 1;
  for (int d = 2; d <= 16; ++d) {
    if (g.K % d || g.K / d < 32) continue;
    if ((size_t)d * g.M * g.N > part_elems) break;
    if (tiles * d > 2048) break;
    best = d;
  }
  return best;
}

Gemm splitk_partials(const Gemm& g, int nsplit, double* part) {
  const int kc = g.K / nsplit;
  Gemm h = g;
  h.K = kc;
  h.sA = (long)kc * (g.a_kcontig ? 1 : g.lda);
  h.sB = (long)kc * (g.b_kcontig ? 1 : g.ldb);
  h.nbatch = nsplit;
  h.bps = 1;
  h.C = part;
  h.ldc = g.M;
  h.sC = (long)g.M * g.N;
  h.alpha = 1.0;
  h.beta = 0.0;
  return h;
}

int gemm_f64_splitk(hipStream_t s, const Gemm& g, double* part, size_t part_elems) {
  const int best = splitk_count(g, part_elems);
  if (best == 1) return gemm_f64(s, g);
  DN_TRY(gemm_f64(s, splitk_partials(g, best, part)));
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
