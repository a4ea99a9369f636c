// Internal declarations shared by the libdonofhip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "donof_hip.h"

namespace dn {

// ---- errors -------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// (mn|P) of an orbital and a fitting basis on the device (dn_ints.hip): into dI with
// leading dimension ld (>= nbf) and slab stride ld * nbf; pad rows are not written.
int int3c2e_device(hipStream_t st, const dn_basis* basis, const dn_basis* aux, double* dI,
                   long ld);

#define DN_HIP(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::dn::fail(DN_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define DN_TRY(expr)            \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != DN_OK) return rc_; \
  } while (0)

// ---- GEMM descriptor ----------------------------------------------------------------
// C_z = alpha * sum_{r in split z} op(A_r) op(B_r) + beta * C_z   (column-major C)
// A(m,k) = a_kcontig ? A[m*lda + k] : A[m + k*lda]
// B(k,n) = b_kcontig ? B[k + n*ldb] : B[k*ldb + n]
// split z covers batches r in [z*bps, min((z+1)*bps, nbatch)); C_z = C + z*sC.
struct Gemm {
  int M = 0, N = 0, K = 0;
  const double* A = nullptr;
  long lda = 0, sA = 0;
  bool a_kcontig = false;
  const double* B = nullptr;
  long ldb = 0, sB = 0;
  bool b_kcontig = true;
  double* C = nullptr;
  long ldc = 0, sC = 0;
  double alpha = 1.0, beta = 0.0;
  int nbatch = 1;
  int bps = 1;  // batches per split
  bool vec_a = false, vec_b = false;  // set by gemm_f64: 16-byte loads allowed
  // kTriB: symmetric B (B(k, n) = B(n, k), N == K), multiply by its lower half Lh with
  // B = Lh + Lh^T (k > n as is, k == n halved), skipping the K range above each column
  // block.  kTriA: the same for a symmetric A (M == K), masked by k > m.
  int tri = 0;
  // LDS-direct pipeline (gemm_f64 and the kLds* grouped configurations; needs K-contiguous
  // A and B with 16-byte loads, tri none or kTriA).  Its k order inside each 16-wide k
  // tile differs from the register-staged kernel's, so results differ in the last bits:
  // products compared bitwise across code paths must choose the same kernel.
  bool lds = false;
};
enum { kTriNone = 0, kTriB = 1, kTriA = 2 };

// One split-precision GEMM (dn_split.hip): FP16 hi+lo operand planes, three MFMA products
// accumulated in FP32.
struct SplitGemm {
  int M, N, K;
  // A: either tile-major f16 planes (row r, k) at (k/32) lda + 32 r + k%32 (lda = rows x 32)
  // or FP32 column-major A[m + k*lda]
  const _Float16* ahi;
  const _Float16* alo;
  const float* af32;
  long lda, sA;             // batch stride (FP32 source)
  const float* a_scale;     // per batch: scale applied to the FP32 A before splitting
  // B: tile-major f16 planes (col c, k) at (k/32) ldb + 32 c + k%32 (ldb = cols x 32),
  // batch stride sB (halves)
  const _Float16* bhi;
  const _Float16* blo;
  long ldb, sB;
  // C: FP32 column-major, batch stride sC
  float* C;
  long ldc, sC;
  // out = acc * cs[1] * (per-Q col factor or per-batch factor); cs = C5 scale pair on
  // the device (scale, 1/scale)
  const float* cs;
  const float* col_inv_scale;  // per Q = col / col_div (K1), nullable
  int col_div;
  const float* batch_inv_scale;  // per batch (K2, symmetric K1), nullable
  float* qmax;                   // per Q max |out| (K1: by column group, or by batch when
                                 // col_div == 0), nullable
  int nbatch;
  // planes output of the symmetric first transform (kernels instantiated for it): hi/lo
  // FP16 of out * out_scale[z], tile-major with C's rows as plane columns and C's columns
  // as k: element (r, c) of batch z at z out_sb + (c / 32) out_ld + 32 r + c % 32, written
  // for c < 32 out_kt (zero for c >= N); the second transform reads them as its B planes
  _Float16* ohi;
  _Float16* olo;
  long out_ld, out_sb;
  int out_kt;
  const float* out_scale;
  // batch-reduce kernels (the Lagrangian contraction): output z sums batches
  // [z bps, min((z + 1) bps, nbatch)), A with batch stride sA (halves)
  int bps;
};

// One problem of a grouped split-precision launch (offsets as for GroupedGemm; nsplit is
// the batch count).
struct GroupedSplit {
  SplitGemm g;
  int tiles_m = 0, tiles_n = 0, nsplit = 0;
  long tile_begin = 0, slot_begin = 0;
};
// Tile configurations of the grouped split launches; kind 0 is the symmetric-b first
// transform (LDS-direct, all four), kind 1 the second transform (FP32 A split on load:
// the first three).
enum SplitConfig {
  kSplit128x128 = 0,
  kSplit64x128 = 1,
  kSplit64x64 = 2,
  kSplit32x64 = 3,
  kNumSplitConfigs = 4
};
bool split_config_ok(int config, int kind);
// Operands of one system's split transforms (the same for the single-system and the
// grouped builds) and the per-system kernels that precede each launch.
SplitGemm split_k1_sym_gemm(int nbf, int nbf5, int naux, int Mpad, int Kpad,
                            const _Float16* bhi, const _Float16* blo, const float* b_inv_scale,
                            const _Float16* chi, const _Float16* clo, const float* cscale_dev,
                            float* bpn, float* qmax);
SplitGemm split_k2_gemm(int nbf, int nbf5, int naux, int Mpad, const _Float16* chi,
                        const _Float16* clo, const float* cscale_dev, const float* bpn,
                        float* bpq, const float* qs, const float* qinv);
// C5 planes and scale; with lhmax (else nullptr) also the per-Q scales qs, qinv of the
// symmetric build's T planes
int split_prepare_c(hipStream_t s, const double* C, int nbf, int nbf5, int Mpad, int Kpad,
                    float* cscale, _Float16* hi, _Float16* lo, const float* lhmax, int naux,
                    float* qs, float* qinv);
// Per-system set-up of a grouped split build in three launches for all systems (the
// arithmetic of split_prepare_c, split_k1_begin and split_k2_begin).  plane_begin and
// q_begin are block offsets filled by split_prep_plan.
struct SplitPrepDesc {
  const double* C;
  int nbf = 0, nbf5 = 0, Mpad = 0, Kpad = 0, naux = 0;
  float* cscale;
  _Float16* chi;
  _Float16* clo;
  float* qmax;
  float* qs;
  float* qinv;
  long plane_begin = 0, q_begin = 0;
};
void split_prep_plan(SplitPrepDesc* desc, int n, long* plane_blocks, long* q_blocks);
int split_group_prep(hipStream_t s, const SplitPrepDesc* dev, int n, long plane_blocks);
int split_group_scales(hipStream_t s, const SplitPrepDesc* dev, int n, long q_blocks);
int split_k1_begin(hipStream_t s, float* qmax, int naux);
int split_k2_begin(hipStream_t s, const float* qmax, int naux, float* qs, float* qinv);
long split_grouped_plan(GroupedSplit* desc, int n, int config);
int split_grouped(hipStream_t s, const GroupedSplit* desc_host, const GroupedSplit* desc_dev,
                  int n, long total, int config, int kind, int grid);

int gemm_f64(hipStream_t s, const Gemm& g);
// Split-K form of gemm_f64_splitk: the split count it uses for g with part_elems
// doubles of partials (1 = no split), and the batched GEMM that writes the partials
// (reduced in split order by reduce_splits, times g.alpha).
int splitk_count(const Gemm& g, size_t part_elems);
Gemm splitk_partials(const Gemm& g, int nsplit, double* part);

// One problem of a grouped GEMM launch (tile list offsets filled by gemm_grouped_plan).
struct GroupedGemm {
  Gemm g;
  int tiles_m = 0, tiles_n = 0, nsplit = 0;
  long tile_begin = 0;  // prefix sum of the tile counts n_i
  long slot_begin = 0;  // prefix sum of ceil(n_i / 8): one XCD's share of each problem
};
// Tile configurations of the grouped launches (chosen by the autotuner).
enum GemmConfig {
  kTile128x128 = 0,
  kTile64x64 = 1,
  kTile64x128 = 2,
  kTile128x64 = 3,
  kTile64x32 = 4,
  kTile32x64 = 5,
  kLds64x128 = 6,   // LDS-direct pipeline, K-contiguous A and B only
  kLds128x128 = 7,
  kLds128x256 = 8,
  kLds64x192 = 9,   // for N in (128, 192]: one N tile, fragments past N skipped
  kTile64x192 = 10,  // register-staged counterpart (any layout and alignment)
  kNumGemmConfigs = 11
};
inline bool is_lds_config(int c) { return c >= kLds64x128 && c <= kLds64x192; }
// The register-staged configuration for an M x N problem in nsplit batches or splits:
// 64 x 64 when M or N is at most 64; with at least 1024 tiles to run, 64 x 192 for N in
// (128, 192] and 64-row tiles when the last 128-row block would be less than half used;
// 64 x 64 when 128 x 128 would leave fewer than 512 tiles; else 128 x 128 (dn_gemm.hip).
int reg_config(int M, int N, long nsplit);
// The 64-row LDS-direct configuration for N columns: 192-wide tiles for N in (128, 192]
// and for long N (>= 1024), 128-wide otherwise (dn_gemm.hip).
int lds64_config(int N);
// Whether a tile configuration can run a problem of the given layout and alignment.
bool gemm_config_ok(int config, const Gemm& g);
// Fill tile counts/offsets of desc[0..n) for a config; returns the total tile count.
long gemm_grouped_plan(GroupedGemm* desc, int n, int config);
// One persistent launch over all problems (descriptors already on the device).  All
// problems share the operand layouts (a_kcontig, b_kcontig) of desc_host[0].  grid > 0
// is the workgroup count; grid <= 0 asks for -grid (at least 1) times the resident
// count, capped at one workgroup per tile.
int gemm_f64_grouped(hipStream_t s, const GroupedGemm* desc_host, const GroupedGemm* desc_dev,
                     int n, long total_tiles, int config, int grid);
// Sum `nsplit` partial matrices (each M x N, leading dim ld, stride sP) into out.
int reduce_splits(hipStream_t s, const double* P, long sP, int nsplit, int M, int N,
                  long ld, double* out, long ldo, double beta, double scale = 1.0);
// Single GEMM split along K into partials (part: >= 16*M*N doubles) and reduced; for
// small outputs whose tile count alone cannot fill the GPU.
int gemm_f64_splitk(hipStream_t s, const Gemm& g, double* part, size_t part_elems);

// K = sum_Q (W + W^T)^2 and X = 2 diag(W) per Q (b_pq = W + W^T of the symmetric-b
// build) for one or several systems in two launches.
struct SymKDesc {
  const double* W;     // nbf5 x nbf5 x naux (FP64 build)
  const float* Wf;     // the same in FP32 (split-precision build)
  double* K;           // nbf5 x nbf5
  double* X;           // nbf5 x naux
  double* part;        // partial sums
  size_t part_elems = 0;
  int nbf5 = 0, naux = 0;
  int qps = 1, splits = 1;  // set by symk_plan
  long blk_begin = 0, red_begin = 0;
};
// f32: the plan is for the FP32 W of the split-precision build
void symk_plan(SymKDesc* desc, int n, long* blocks, long* red_blocks, bool f32 = false);
int j_from_x(hipStream_t s, const double* X, int nbf5, int naux, double* part2,
             size_t part2_elems, double* J);
// vec4: FP32 W whose nbf5 is a multiple of 4 for every system (16-byte loads)
int symk(hipStream_t s, const SymKDesc* dev, int n, long blocks, long red_blocks,
         bool f32 = false, bool vec4 = false);

// ---- device buffers -----------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace dn

namespace dn {
inline int padded_ld(int nbf) { return nbf + (nbf & 1); }
// padded copies C -> Cp (dn_jk.hip): one system, or every system of a group in one launch
struct PadDesc {
  const double* C;
  double* Cp;
  int nbf, ld;
  long begin;  // first block (pad_plan)
};
int pad_c_launch(hipStream_t s, const double* C, double* Cp, int nbf, int ld);
long pad_plan(PadDesc* d, int n);
int pad_c_group(hipStream_t s, const PadDesc* d, int n, long blocks);
// Cp = C with the zero pad row (a no-op when ldb == nbf), before the transforms read it
int pad_c(dn_sys* s);
// the transforms of b as GEMMs (dn_api.hip): t1 bpn[n,p,Q] = (b_Q^T C5)[n,p] (Lh_Q^T C5
// with sym), t2 flat bt[q,(p,Q)] = (C5^T bpn)[q,(p,Q)], t2 batched
// bt[q,a,Q] = (bpn_Q^T C)[q,a] for a < ncols
Gemm t1_gemm(const dn_sys* s, bool sym);
Gemm t2_flat_gemm(const dn_sys* s);
Gemm t2_batched_gemm(const dn_sys* s, int ncols);
// the Lagrangian's contraction Z[n,b] = sum_{q,Q} bpn[n,q,Q] bmat[q,b,Q] in split partials
// and elag += C^T Z (dn_api.hip)
Gemm lag_contract_gemm(const dn_sys* s);
Gemm lag_final_gemm(const dn_sys* s);
// resident building blocks (dn_api.hip)
int check_sys(dn_sys* s);
int h2d(dn_sys* s, double* dst, const double* src, size_t elems);
int d2h(dn_sys* s, double* dst, const double* src, size_t elems);
int jkh_resident(dn_sys* s);
int lagrange_resident(dn_sys* s, int nalpha, int nbeta);
bool all_finite(const double* x, size_t n);
int prepare_b(dn_sys* s, const double* b);  // symmetry flag (and split planes) of b
// b written by fill into the system's FP64 b, or for a split system into a transient FP64
// buffer that its planes are made from; clears the RI factor
int set_b(dn_sys* s, const std::function<int(double*)>& fill);
int begin_timing(dn_sys* s);  // HIP events around the dominant kernel
int end_timing(dn_sys* s);
}  // namespace dn

namespace dn {
// One system of a lock-step group for the grouped Lagrangian kernels (dn_lag.hip): the
// operands of lag_w, lag_bmat, lag_init, the reduction of the C^J/C^K partials into elag,
// energy_elec and pack_grad, and block offsets filled by lag_group_plan.
struct LagSys {
  const double* bt;    // b_pq[q,p,Q] at p + ldt (q + nbf5 Q)
  long slice;          // its slab stride (ldt nbf5)
  int ldt;
  int nbf, nbf5, naux, nalpha, nbeta;
  const double* cj12;
  const double* ck12;
  const double* n;
  double* X;
  double* w;
  double* bmat;
  const double* Hmat;
  double* elag;
  const double* part;  // nsplit partials of nbf x nbf5
  int nsplit;
  double* Z;           // their sum (nbf x nbf5), contracted with C into elag afterwards
  double* scal;        // scal[0] = E, scal[1] = bits of max|g|
  double* g;           // packed gradient
  long w_begin, bmat_begin, sq_begin, rect_begin;
};
struct LagBlocks {
  long w = 0, bmat = 0, sq = 0, rect = 0;
  size_t lds_w = 0;
};
void lag_group_plan(LagSys* d, int n, LagBlocks* b);
int lag_group_before(hipStream_t s, const LagSys* d, int n, const LagBlocks& b);
// the partials summed into Z; then (after elag += C^T Z) energy, gmax and packed gradient
int lag_group_reduce(hipStream_t s, const LagSys* d, int n, const LagBlocks& b);
int lag_group_after(hipStream_t s, const LagSys* d, int n, const LagBlocks& b);
}  // namespace dn

struct dn_ctx {
  int device = 0;
  uint64_t budget = 0;
  uint64_t used = 0;
  // Device memory of the context's systems and groups: a stream-ordered pool that keeps
  // the blocks it is given back (hipFree waits for every stream of the device; a
  // stream-ordered free waits for nothing), trimmed when an allocation does not fit.
  hipMemPool_t pool = nullptr;
  // pinned host staging blocks of lock-step groups, kept for reuse (hipHostFree waits
  // for the device as hipFree does)
  std::mutex pin_mu;
  std::vector<std::pair<void*, size_t>> pin_free;
};

namespace dn {
// bytes from the context's pool, ordered on stream s
int dev_alloc(dn_ctx* c, hipStream_t s, size_t bytes, void** p);
// give p back to the pool, ordered after the work queued on s
void dev_free(dn_ctx* c, hipStream_t s, void* p);
// group membership (dn_api.hip): a lock-step group claims its systems (none may already
// belong to one, none listed twice) and releases them; grouped builds count their systems
int claim_lockstep(dn_sys* const* systems, int n, void* owner);
// HBM budget of a context: reserve (DN_ERR_BUDGET past the budget) and give back bytes
// held outside the systems' own footprints
int ctx_reserve(dn_ctx* c, uint64_t bytes);
void ctx_unreserve(dn_ctx* c, uint64_t bytes);
// a dn_grad holds (+1) or lets go of (-1) its system
void count_grad(dn_sys* s, int delta);
void release_lockstep(dn_sys* const* systems, int n, void* owner);
void count_build_group(dn_sys* const* systems, int n, int delta);
// a pinned host block of at least `bytes` (a cached one when one is large enough)
int pinned_get(dn_ctx* c, size_t bytes, void** p, size_t* got);
void pinned_put(dn_ctx* c, void* p, size_t bytes);
}  // namespace dn

struct dn_sys {
  dn_ctx* ctx = nullptr;
  dn_dims d{};
  int precision = DN_FP64;
  hipStream_t stream = nullptr;
  // split-precision systems: H C5 and H_core run on a second stream beside the transforms
  // (forked and joined with the two events)
  hipStream_t aux = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  std::vector<dn::DevBuf> owned;
  uint64_t bytes = 0;
  // resident state (column-major)
  // b[m,n,Q] at m + ldb (n + nbf Q); ldb = nbf rounded up to even, the pad row zero, so
  // that every row of b (and of bpn, Cp) starts 16-byte aligned and the transforms take
  // the LDS-direct tiles for odd nbf too (contracting over ldb: the pads add zeros).
  // Split-precision systems keep b only as its FP16 planes (b is null there).
  int ldb = 0;
  double* b = nullptr;      // ldb*nbf*naux (FP64 systems)
  double* C = nullptr;      // nbf*nbf
  double* Cp = nullptr;     // ldb*nbf: C with the zero pad row (C itself when ldb == nbf),
                            // refreshed from C before the transforms (pad_c)
  double* H = nullptr;      // nbf*nbf
  double* n = nullptr;      // nbf5
  double* cj12 = nullptr;   // nbf5*nbf5
  double* ck12 = nullptr;   // nbf5*nbf5
  double* J = nullptr;      // nbf5*nbf5
  double* K = nullptr;      // nbf5*nbf5
  double* Hcore = nullptr;  // nbf5
  double* elag = nullptr;   // nbf*nbf
  double* Hmat = nullptr;   // nbf*nbf5
  // workspaces
  double* bpn = nullptr;    // ldb*nbf5*naux : [n,p,Q] = (b_Q^T C5)[n,p] (Lh^T C5 symmetric);
                            //                 pad row zero
  double* bt = nullptr;     // nbf5*nbf5*naux: b_pq[q,p,Q] at p + nbf5 (q + nbf5 Q) (general
                            //   transform, the Lagrangian) or W^T (symmetric J/K build)
  size_t bpn_elems = 0;     // doubles of bpn (split systems keep the contraction planes there)
  double* bmat = nullptr;   // nbf5*nbf5*naux: [q,b,Q] at b + nbf5 (q + nbf5 Q), the C^K/C^J
                            //   right operand of the Lagrangian GEMM
  double* X = nullptr;      // nbf5*naux     : b_pp diagonal per Q
  double* w = nullptr;      // nbf5*naux     : Coulomb + C^J weights
  double* HC = nullptr;     // nbf*nbf5      : H C5
  double* Z = nullptr;      // nbf*nbf5      : sum_{q,Q} bpn[n,q,Q] bmat[q,b,Q] (Lagrangian)
  double* part = nullptr;   // split partials
  size_t part_elems = 0;
  double* scal = nullptr;   // small scalars (energy, norms)
  // occupation functional
  double* gamma = nullptr;  // nv
  double* dn = nullptr;     // nbf5*nv : dn/dgamma
  double* goc = nullptr;    // nv      : occupation gradient
  double* orb = nullptr;    // 6*nbf5  : fi, nd, nd12, dfic, ndc, ndx
  int* orbi = nullptr;      // 3*nbf5  : strong, pair, role
  // orbital optimiser and rotation
  long nvar = 0;
  double* gv = nullptr;     // nvar: packed gradient
  double* mv = nullptr;     // nvar: first moment
  double* vv = nullptr;     // nvar: second moment
  double* vmax = nullptr;   // nvar: AMSGrad max
  double* yv = nullptr;     // nvar: step
  double* Y = nullptr;      // nbf*nbf generator
  double* U = nullptr;      // nbf*nbf exp(Y)
  double* rwork = nullptr;  // 5*nbf*nbf exp workspace
  double* Ca = nullptr;     // nbf*nbf spare orbitals
  double* bestC = nullptr;  // nbf*nbf
  double* part2 = nullptr;  // 16*nbf*nbf split-K partials
  size_t part2_elems = 0;
  // split-precision J/K build (precision == DN_SPLIT_F16)
  int Kpad = 0, Mpad = 0;
  _Float16* bhi = nullptr;  // tile-major per Q (nbf columns n, k = m < Kpad): hi of b*scale[Q]
  _Float16* blo = nullptr;
  _Float16* chi = nullptr;  // tile-major Mpad rows, k < Kpad: hi of C5^T*cscale
  _Float16* clo = nullptr;
  float* bscale = nullptr;  // naux scales and inverses of b
  float* binv = nullptr;
  float* cscale = nullptr;  // (scale, 1/scale) of C5
  float* bpn32 = nullptr;   // max(nbf5*nbf, Mpad*Kpad)*naux: FP32 T (general b) or the
                            // hi/lo planes of T (symmetric b), tile-major Mpad x Kpad per Q
  float* lhmax = nullptr;   // naux: max_n ||Lh_Q[:,n]|| (the symmetric T's scale bound)
  float* bpq32 = nullptr;   // nbf5*nbf5*naux
  float* qmax = nullptr;    // naux
  float* qs = nullptr;      // naux
  float* qinv = nullptr;    // naux
  // split Lagrangian contraction: (scale of bpn, 1 / (scale of bpn * scale of bmat), scale
  // of bmat, -) and up to 8192 block maxima of |bmat|; its planes and FP32 partials live
  // in bpn's space
  float* zscale = nullptr;
  bool b_ready = false;
  // an ADAM-family run advanced step by step (dn_orbopt_start / dn_ogroup_step)
  struct OptRun {
    bool started = false, done = false, improved = false, success = false;
    int i = 0, nit = 0, method = 0, maxloop = 0, nalpha = 0, nbeta = 0;
    double alpha = 0, beta1 = 0, binit = 0, beta2 = 0, threshgorb = 0;
    double E = 0, gmax = 0, best_E = 0, nrm = 0;
  } opt;
  // group membership: the lock-step group driving the system (it runs on that group's
  // stream meanwhile) and the number of grouped builds that read its buffers; a system
  // that belongs to either is not destroyed
  void* lockstep = nullptr;
  int build_groups = 0;
  int grad_objs = 0;     // live dn_grad objects (they run on the system's stream)
  // the auxiliary-index factor of b for the nuclear gradients, I3c G^{-1} = b Wt:
  // dn_sys_ri_metric keeps riw = L^{-T} (G = L L^T, b = I3c L^{-T}, Wt = riw^T);
  // dn_sys_set_ri_factor stores a caller's Wt as is (column-major)
  double* riw = nullptr;  // naux*naux
  bool riw_set = false;
  bool riw_transposed = false;  // riw holds Wt^T
  // dn_sys_ri_metric took the eigen route (G not positive definite), with an eigenvalue
  // below zero set to zero
  bool eig_route = false, eig_clamped = false;
  bool b_sym = false;    // b[m,n,Q] == b[n,m,Q] (checked whenever b is set)
  dn::SymKDesc* symk_dev = nullptr;  // K/X pass of the symmetric build (in scal + 16;
                                     // the split-precision one at scal + 32)
  long symk_blocks = 0, symk_red = 0;
  bool use_sym = true;   // let the J/K build use that symmetry
  // timing of the dominant transform kernel: one event pair per launch, read back
  // (and recycled) by dn_sys_get_timing so that timing adds no synchronisation.
  bool timing = false;
  std::vector<hipEvent_t> ev_start, ev_stop;
  size_t ev_used = 0;
  double transform_ms = 0.0;
  long transform_launches = 0;
};
