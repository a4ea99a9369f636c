// Grouped J/K/H_core builds: many resident systems of uneven size built together.
//
// Each system's build is the one of integrals.jl:144-172 (JKH_MO).  The group turns the
// per-system stages into grouped launches over all systems:
//   K1  b_pn = C5^T b                 persistent grouped GEMM (the dominant stage); for a
//                                     symmetric b only its lower half: T = C5^T Lh
//   K2  b_pq[:,:,Q] = b_pn[:,:,Q] C5  persistent grouped GEMM, one tile list over (sys, Q)
//   KX  K = sum_Q b_pq^2, X = diag     partial sums over Q ranges, then their reduction;
//                                     for the lower-half systems K and X come straight
//                                     from W with b_pq = W + W^T (never stored)
//   J   J = X X^T                     persistent grouped GEMM
//   HC  H C5                          persistent grouped GEMM
//   HD  H_core_i = C5[:,i] . (H C5)[:,i]
// An autotuner times every tile configuration of each GEMM stage on the actual group,
// keeps the fastest, and the whole sequence is then captured once into a HIP graph that
// every build replays.
#include "dn_internal.h"

namespace dn {
namespace {

// K = sum_Q b_pq^2 and X[p,Q] = b_ppQ for one system of the group, in three launches
// over all systems: partial sums of K over Q ranges (thread per (p,q), coalesced over the
// slice), their reduction, and the diagonal gather.
struct KxDesc {
  const double* bpq;  // nbf5 x nbf5 x naux, slice stride nbf5^2
  double* K;
  double* X;          // nbf5 x naux
  double* part;       // splits x nbf5^2 partial sums
  int nbf5, naux;
  int splits, qps;    // Q ranges of qps slices
  long part_begin;    // first workgroup of this system in each launch
  long red_begin;
  long x_begin;
};
struct DiagDesc {
  const double* C;
  const double* HC;
  double* Hcore;
  int nbf, nbf5;
  long col_begin;
};

template <typename D>
__device__ __forceinline__ int find_desc(const D* __restrict__ d, int n, long col) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].col_begin <= col) lo = mid;
    else hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);  // uniform: scalar descriptor loads
}

template <long KxDesc::*BEGIN>
__device__ __forceinline__ int find_kx(const KxDesc* __restrict__ d, int n, long blk) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].*BEGIN <= blk) lo = mid;
    else hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);  // uniform: scalar descriptor loads
}

__global__ __launch_bounds__(256) void group_kpart_kernel(const KxDesc* __restrict__ desc, int n) {
  const KxDesc d = desc[find_kx<&KxDesc::part_begin>(desc, n, blockIdx.x)];
  const long nn = (long)d.nbf5 * d.nbf5;
  const long chunks = (nn + 255) / 256;
  const long local = blockIdx.x - d.part_begin;
  const long e = (local % chunks) * 256 + threadIdx.x;
  const int z = (int)(local / chunks);
  if (e >= nn) return;
  const int Q0 = z * d.qps, Q1 = min(Q0 + d.qps, d.naux);
  double k = 0.0;
  for (int Q = Q0; Q < Q1; ++Q) {
    const double v = d.bpq[(long)Q * nn + e];
    k = fma(v, v, k);
  }
  d.part[z * nn + e] = k;
}

__global__ __launch_bounds__(256) void group_kred_kernel(const KxDesc* __restrict__ desc, int n) {
  const KxDesc d = desc[find_kx<&KxDesc::red_begin>(desc, n, blockIdx.x)];
  const long nn = (long)d.nbf5 * d.nbf5;
  const long e = (blockIdx.x - d.red_begin) * 256 + threadIdx.x;
  if (e >= nn) return;
  double k = 0.0;
  for (int z = 0; z < d.splits; ++z) k += d.part[z * nn + e];
  d.K[e] = k;
}

__global__ __launch_bounds__(256) void group_diag_kernel(const KxDesc* __restrict__ desc, int n) {
  const KxDesc d = desc[find_kx<&KxDesc::x_begin>(desc, n, blockIdx.x)];
  const long e = (blockIdx.x - d.x_begin) * 256 + threadIdx.x;
  if (e >= (long)d.nbf5 * d.naux) return;
  const long Q = e / d.nbf5;
  const int p = (int)(e - Q * d.nbf5);
  d.X[e] = d.bpq[Q * d.nbf5 * d.nbf5 + (long)p * (d.nbf5 + 1)];
}

// One wavefront per (system, i): H_core_i = sum_m C[m, i] HC[m, i].
__global__ __launch_bounds__(64) void group_hdiag_kernel(const DiagDesc* __restrict__ desc, int n) {
  const long col = blockIdx.x;
  const DiagDesc d = desc[find_desc(desc, n, col)];
  const int i = (int)(col - d.col_begin);
  double acc = 0.0;
  for (int m = threadIdx.x; m < d.nbf; m += 64)
    acc = fma(d.C[m + (long)i * d.nbf], d.HC[m + (long)i * d.nbf], acc);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (threadIdx.x == 0) d.Hcore[i] = acc;
}

enum Stage { kK1 = 0, kK2 = 1, kJ = 2, kHC = 3, kNumStages = 4 };
// Parallel branches of a build: the classes of a stage (and the per-system set-up kernels
// of split groups) run on these streams side by side, joined before the next stage; the
// device serves four hardware queues per process.
constexpr int kAux = 4;
constexpr int kChainMin = 4;
constexpr double kChainWork = 1e11;  // total build work (the weights of dn_group_create)

}  // namespace
}  // namespace dn

using namespace dn;

struct dn_group {
  dn_ctx* ctx = nullptr;
  std::vector<dn_sys*> sys;
  bool counted = false;  // the top-level group: holds its systems' build_groups count
  int precision = DN_FP64;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hipStream_t aux[kAux] = {};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[kAux] = {};
  // Groups of kChainMin or more systems are chains of leaf groups (systems dealt by work,
  // largest first, to the least-loaded chain): each chain runs its own stage sequence on
  // its own branch of the one captured graph, so one chain's first transform overlaps
  // another's second transform and K/X as independent streams would, while every chain
  // still builds its systems with grouped launches.
  std::vector<dn_group*> chains;
  // a leaf that is one chain of a parent: its stages' classes run one after another on
  // its own stream (see launch_stage)
  bool chain_member = false;
  // FP64 systems with odd nbf: C copied into its padded form before the transforms
  PadDesc* pad = nullptr;
  int npad = 0;
  long pad_blocks = 0;
  // Problems of a stage sorted into classes by operand layout and alignment (16-byte
  // loads allowed for A, for B) and by nbf5 (<= 32, <= 64, <= 128, larger): one
  // persistent launch per class, each with the tile shape the autotuner found fastest
  // for it, so neither an odd leading dimension nor a small system sets the tiles of
  // the others.
  struct Launch {
    int first, count;
    long tiles;
    int config;
    double flops;
    int grid = 0;  // gemm_f64_grouped's grid argument
  };
  std::vector<GroupedGemm> host[kNumStages];
  std::vector<Launch> launches[kNumStages];
  GroupedGemm* dev[kNumStages] = {};
  // split-precision groups: the two transforms run as grouped split launches (kind 0 =
  // K1, 1 = K2) over these descriptors instead of host[kK1], host[kK2]
  std::vector<GroupedSplit> shost[2];
  GroupedSplit* sdev[2] = {};
  SplitPrepDesc* sprep = nullptr;  // per-system set-up of the split transforms
  long prep_blocks = 0, q_blocks = 0;
  KxDesc* kx = nullptr;
  long kx_blocks[3] = {0, 0, 0};
  DiagDesc* diag = nullptr;
  SymKDesc* symk = nullptr;  // systems with the symmetric build
  int nsym = 0;
  bool symk_vec4 = false;
  long symk_blocks = 0, symk_red = 0;
  int nkx = 0;               // systems with the full build
  long ncols = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool tuned = false;
  bool timing = false;
  // Placement of the top-level group's systems (dn_group_place / dn_group_tune): `sub`
  // builds the systems with nbf5 <= place_nbf5 by grouped launches in its captured graph,
  // the `indep` systems run the single-system build on their own streams beside it.
  // Without a placement the group itself is the grouped build of every system.
  bool placed = false;
  bool place_fixed = false;  // set by dn_group_place: dn_group_tune does not search
  int place_nbf5 = 0;
  dn_group* sub = nullptr;
  std::vector<dn_sys*> indep;
  std::vector<hipEvent_t> indep_ev;
  std::vector<hipEvent_t> ev_start, ev_stop;
  size_t ev_used = 0;
  double build_ms = 0.0;
  long builds = 0;
};

namespace dn {
namespace {

// The per-system GEMM of a stage (the same operands as the single-system build).
Gemm stage_gemm(dn_sys* s, int stage) {
  const dn_dims& d = s->d;
  Gemm g;
  switch (stage) {
    case kK1:  // bpn = b_Q C5 (or Lh_Q^T C5 for the symmetric build), batched over Q
      g = t1_gemm(s, s->b_sym && s->use_sym);
      break;
    case kK2:  // C5^T bpn with (p, Q) flattened (symmetric), bpn_Q^T C5 per Q (general)
      g = (s->b_sym && s->use_sym) ? t2_flat_gemm(s) : t2_batched_gemm(s, d.nbf5);
      break;
    case kJ:  // J = X X^T, X[p + Q nbf5]
      g.M = d.nbf5; g.N = d.nbf5; g.K = d.nbfaux;
      g.A = s->X; g.lda = d.nbf5; g.a_kcontig = false;
      g.B = s->X; g.ldb = d.nbf5; g.b_kcontig = false;
      g.C = s->J; g.ldc = d.nbf5;
      break;
    default:  // HC = H C5
      g.M = d.nbf; g.N = d.nbf5; g.K = d.nbf;
      g.A = s->H; g.lda = d.nbf; g.a_kcontig = false;
      g.B = s->C; g.ldb = d.nbf; g.b_kcontig = true;
      g.C = s->HC; g.ldc = d.nbf;
      break;
  }
  return g;
}

int size_bucket(const Gemm& g, int stage) {
  const int n5 = (stage == kHC || stage == kK1) ? g.N : g.M;
  return n5 <= 32 ? 0 : n5 <= 64 ? 1 : n5 <= 128 ? 2 : 3;
}

int default_config(int stage, int bucket, const Gemm& g) {
  if (bucket == 0) return kTile64x32;
  // the transforms take the LDS-direct tiles the single-system build uses (K2 flattened
  // over (p, Q) for the symmetric build, batched over Q for a general b)
  const bool lds = (stage == kK1 || stage == kK2) && g.lds;
  if (lds && gemm_config_ok(kLds64x128, g)) {
    if (stage == kK1) return lds64_config(g.N);
    if (bucket == 1) return kLds64x128;
    return g.nbatch > 1 ? kLds128x128 : kLds128x256;
  }
  if (stage == kK2 && bucket >= 2) return kTile128x128;
  if (stage == kK1) return g.N > 128 && g.N <= 192 ? kTile64x192 : kTile64x128;
  return kTile64x64;
}

// Classes of one stage with their default tile configurations; descriptors uploaded.
bool split_stage(const dn_group* G, int stage) {
  return G->precision == DN_SPLIT_F16 && (stage == kK1 || stage == kK2);
}

// The grouped split transforms of one system (the operands of jkh_split's symmetric build).
SplitGemm split_stage_gemm(dn_sys* s, int stage) {
  const dn_dims& d = s->d;
  if (stage == kK1)
    return split_k1_sym_gemm(d.nbf, d.nbf5, d.nbfaux, s->Mpad, s->Kpad, s->bhi, s->blo, s->binv,
                             s->chi, s->clo, s->cscale, s->bpn32, s->qmax);
  return split_k2_gemm(d.nbf, d.nbf5, d.nbfaux, s->Mpad, s->chi, s->clo, s->cscale, s->bpn32,
                       s->bpq32, s->qs, s->qinv);
}

int split_default_config(int stage, int bucket) {
  if (stage == kK1) return bucket == 0 ? kSplit32x64 : bucket == 1 ? kSplit64x64 : kSplit128x128;
  return bucket <= 1 ? kSplit64x64 : kSplit128x128;
}

int plan_split_stage(dn_group* G, int stage) {
  auto& h = G->shost[stage];
  h.resize(G->sys.size());
  for (size_t i = 0; i < G->sys.size(); ++i) h[i].g = split_stage_gemm(G->sys[i], stage);
  auto key = [](const GroupedSplit& d) {  // nbf5 class
    const int n5 = d.g.M;
    return n5 <= 32 ? 0 : n5 <= 64 ? 1 : n5 <= 128 ? 2 : 3;
  };
  std::stable_sort(h.begin(), h.end(),
                   [&](const GroupedSplit& a, const GroupedSplit& b) { return key(a) < key(b); });
  auto& L = G->launches[stage];
  L.clear();
  for (int i = 0; i < (int)h.size();) {
    int j = i;
    double f = 0.0;
    while (j < (int)h.size() && key(h[j]) == key(h[i])) {
      f += 2.0 * h[j].g.M * h[j].g.N * (double)h[j].g.K * h[j].g.nbatch;
      ++j;
    }
    const int cfg = split_default_config(stage, key(h[i]));
    L.push_back({i, j - i, split_grouped_plan(h.data() + i, j - i, cfg), cfg, f, 0});
    i = j;
  }
  DN_HIP(hipMemcpy(G->sdev[stage], h.data(), h.size() * sizeof(GroupedSplit),
                   hipMemcpyHostToDevice));
  return DN_OK;
}

int plan_stage(dn_group* G, int stage) {
  DN_HIP(hipStreamSynchronize(G->stream));  // a replay in flight may read the descriptors
  if (split_stage(G, stage)) return plan_split_stage(G, stage);
  auto& h = G->host[stage];
  h.resize(G->sys.size());
  for (size_t i = 0; i < G->sys.size(); ++i) h[i].g = stage_gemm(G->sys[i], stage);
  gemm_grouped_plan(h.data(), (int)h.size(), kTile64x64);  // sets the alignment flags
  auto key = [stage](const GroupedGemm& d) {
    return 64 * d.g.tri + 32 * d.g.a_kcontig + 16 * d.g.b_kcontig + 8 * !d.g.vec_a +
           4 * !d.g.vec_b + size_bucket(d.g, stage);
  };
  std::stable_sort(h.begin(), h.end(),
                   [&](const GroupedGemm& a, const GroupedGemm& b) { return key(a) < key(b); });
  auto& L = G->launches[stage];
  L.clear();
  for (int i = 0; i < (int)h.size();) {
    int j = i;
    double f = 0.0;
    while (j < (int)h.size() && key(h[j]) == key(h[i])) {
      f += 2.0 * h[j].g.M * h[j].g.N * (double)h[j].g.K * h[j].g.nbatch;
      ++j;
    }
    const int cfg = default_config(stage, size_bucket(h[i].g, stage), h[i].g);
    L.push_back({i, j - i, gemm_grouped_plan(h.data() + i, j - i, cfg), cfg, f, 0});
    i = j;
  }
  DN_HIP(hipMemcpy(G->dev[stage], h.data(), h.size() * sizeof(GroupedGemm),
                   hipMemcpyHostToDevice));
  return DN_OK;
}

// Re-plan one class for another tile configuration.
int set_config(dn_group* G, int stage, size_t li, int config) {
  DN_HIP(hipStreamSynchronize(G->stream));
  auto& l = G->launches[stage][li];
  if (split_stage(G, stage)) {
    GroupedSplit* h = G->shost[stage].data() + l.first;
    l.tiles = split_grouped_plan(h, l.count, config);
    l.config = config;
    DN_HIP(hipMemcpy(G->sdev[stage] + l.first, h, l.count * sizeof(GroupedSplit),
                     hipMemcpyHostToDevice));
    return DN_OK;
  }
  GroupedGemm* h = G->host[stage].data() + l.first;
  l.tiles = gemm_grouped_plan(h, l.count, config);
  l.config = config;
  DN_HIP(hipMemcpy(G->dev[stage] + l.first, h, l.count * sizeof(GroupedGemm),
                   hipMemcpyHostToDevice));
  return DN_OK;
}

int launch_one(dn_group* G, int stage, size_t li, hipStream_t st = nullptr) {
  const auto& l = G->launches[stage][li];
  if (!st) st = G->stream;
  if (split_stage(G, stage))
    return split_grouped(st, G->shost[stage].data() + l.first, G->sdev[stage] + l.first, l.count,
                         l.tiles, l.config, stage == kK1 ? 0 : 1, l.grid);
  return gemm_f64_grouped(st, G->host[stage].data() + l.first, G->dev[stage] + l.first, l.count,
                          l.tiles, l.config, l.grid);
}

// The first k branch streams start after everything issued on the group stream so far;
// join makes the group stream wait for all of them (also inside a graph capture, where
// this records parallel branches).
int fork(dn_group* G, int k) {
  DN_HIP(hipEventRecord(G->fork_ev, G->stream));
  for (int i = 0; i < k; ++i) DN_HIP(hipStreamWaitEvent(G->aux[i], G->fork_ev, 0));
  return DN_OK;
}

int join(dn_group* G, int k) {
  for (int i = 0; i < k; ++i) {
    DN_HIP(hipEventRecord(G->join_ev[i], G->aux[i]));
    DN_HIP(hipStreamWaitEvent(G->stream, G->join_ev[i], 0));
  }
  return DN_OK;
}

int launch_stage(dn_group* G, int stage) {
  const int L = (int)G->launches[stage].size();
  // A chain's classes stay on the chain's stream: when each chain forked its classes over
  // the four branch streams, the graph ran the chains one after another (each chain's
  // classes side by side on the four hardware queues); linear chains run side by side.
  if (L <= 1 || G->chain_member) {
    for (int li = 0; li < L; ++li) DN_TRY(launch_one(G, stage, li));
    return DN_OK;
  }
  const int k = std::min(L, kAux);
  DN_TRY(fork(G, k));
  for (int li = 0; li < L; ++li) DN_TRY(launch_one(G, stage, li, G->aux[li % k]));
  return join(G, k);
}

int launch_build(dn_group* G);

// The chains' builds as parallel branches of the parent stream.
int launch_chains(dn_group* G) {
  DN_HIP(hipEventRecord(G->fork_ev, G->stream));
  for (dn_group* c : G->chains) DN_HIP(hipStreamWaitEvent(c->stream, G->fork_ev, 0));
  for (dn_group* c : G->chains) DN_TRY(launch_build(c));
  for (dn_group* c : G->chains) {
    DN_HIP(hipEventRecord(c->done, c->stream));
    DN_HIP(hipStreamWaitEvent(G->stream, c->done, 0));
  }
  return DN_OK;
}

int launch_build(dn_group* G) {
  if (!G->chains.empty()) return launch_chains(G);
  const int n = (int)G->sys.size();
  const bool split = G->precision == DN_SPLIT_F16;
  // split groups: C5 planes for the current orbitals and cleared per-Q maxima, then the
  // per-Q scales of the first transform's output, each one launch for all systems
  if (split) DN_TRY(split_group_prep(G->stream, G->sprep, n, G->prep_blocks));
  if (!split) DN_TRY(pad_c_group(G->stream, G->pad, G->npad, G->pad_blocks));
  DN_TRY(launch_stage(G, kK1));
  if (split) DN_TRY(split_group_scales(G->stream, G->sprep, n, G->q_blocks));
  DN_TRY(launch_stage(G, kK2));
  DN_TRY(symk(G->stream, G->symk, G->nsym, G->symk_blocks, G->symk_red, split,
              split && G->symk_vec4));
  if (G->nkx) {
    const int m = G->nkx;
    hipLaunchKernelGGL(group_kpart_kernel, dim3(G->kx_blocks[0]), dim3(256), 0, G->stream, G->kx, m);
    hipLaunchKernelGGL(group_kred_kernel, dim3(G->kx_blocks[1]), dim3(256), 0, G->stream, G->kx, m);
    hipLaunchKernelGGL(group_diag_kernel, dim3(G->kx_blocks[2]), dim3(256), 0, G->stream, G->kx, m);
    DN_HIP(hipGetLastError());
  }
  DN_TRY(launch_stage(G, kJ));
  DN_TRY(launch_stage(G, kHC));
  hipLaunchKernelGGL(group_hdiag_kernel, dim3(G->ncols), dim3(64), 0, G->stream, G->diag, n);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// Each system's own stream may still hold uploads: the group stream waits for them, and
// the systems' streams wait for the build before they read J/K/H_core.
int join_systems(dn_group* G) {
  for (dn_sys* s : G->sys) {
    DN_HIP(hipEventRecord(G->done, s->stream));
    DN_HIP(hipStreamWaitEvent(G->stream, G->done, 0));
  }
  return DN_OK;
}

int release_systems(dn_group* G) {
  DN_HIP(hipEventRecord(G->done, G->stream));
  for (dn_sys* s : G->sys) DN_HIP(hipStreamWaitEvent(s->stream, G->done, 0));
  return DN_OK;
}

void drop_graph(dn_group* G) {
  if (G->exec) (void)hipGraphExecDestroy(G->exec);
  if (G->graph) (void)hipGraphDestroy(G->graph);
  G->exec = nullptr;
  G->graph = nullptr;
}

int capture(dn_group* G) {
  DN_HIP(hipStreamSynchronize(G->stream));
  drop_graph(G);
  DN_HIP(hipStreamBeginCapture(G->stream, hipStreamCaptureModeThreadLocal));
  const int rc = launch_build(G);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(G->stream, &graph);
  if (rc != DN_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  DN_HIP(e);
  G->graph = graph;
  DN_HIP(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
  return DN_OK;
}

int time_launch(dn_group* G, int stage, size_t li, float* ms) {
  hipEvent_t a, b;
  DN_HIP(hipEventCreate(&a));
  DN_HIP(hipEventCreate(&b));
  int rc = launch_one(G, stage, li);  // warm-up (code object load, caches)
  if (rc == DN_OK) {
    (void)hipEventRecord(a, G->stream);
    for (int r = 0; r < 2 && rc == DN_OK; ++r) rc = launch_one(G, stage, li);
    (void)hipEventRecord(b, G->stream);
    (void)hipEventSynchronize(b);
    (void)hipEventElapsedTime(ms, a, b);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

}  // namespace
}  // namespace dn

extern "C" {

static int group_create_leaf(dn_ctx* ctx, dn_sys* const* systems, int nsys, dn_group** out);

// A leaf group, or a parent over chains of leaves, for these systems (build-group counts
// untouched).
static int group_create_any(dn_ctx* ctx, dn_sys* const* systems, int nsys, dn_group** out) {
  *out = nullptr;
  // build work of each system (FLOP of its two transforms, up to a factor)
  std::vector<int> order(nsys);
  std::vector<double> w(nsys);
  double total = 0.0;
  for (int i = 0; i < nsys; ++i) {
    order[i] = i;
    const dn_sys* s = systems[i];
    if (!s || s->ctx != ctx) return fail(DN_ERR_ARG, "systems must share the context");
    w[i] = (double)s->d.nbf * s->d.nbf5 * s->d.nbfaux * (s->d.nbf + 2.0 * s->d.nbf5);
    total += w[i];
  }
  // chains pay when their persistent launches are long enough to overlap; for small
  // systems the extra graph nodes cost more than the overlap gains (measured on MI355X:
  // 64 systems of nbf 16-64 build in 1.0 ms as one chain, 1.2 ms as four)
  if (nsys < kChainMin || total < kChainWork) return group_create_leaf(ctx, systems, nsys, out);
  // a parent over chains: the systems dealt by build work, largest first, to the
  // least-loaded chain
  const int nch = std::min(kAux, nsys / 2);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] > w[b]; });
  std::vector<std::vector<dn_sys*>> part(nch);
  std::vector<double> load(nch, 0.0);
  for (int i : order) {
    const int c = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    part[c].push_back(systems[i]);
    load[c] += w[i];
  }
  dn_group* G = nullptr;
  DN_TRY(group_create_leaf(ctx, systems, nsys, &G));  // validation, streams, events
  for (int c = 0; c < nch; ++c) {
    dn_group* leaf = nullptr;
    const int rc = group_create_leaf(ctx, part[c].data(), (int)part[c].size(), &leaf);
    if (rc != DN_OK) {
      dn_group_destroy(G);
      return rc;
    }
    leaf->chain_member = true;
    G->chains.push_back(leaf);
  }
  *out = G;
  return DN_OK;
}

int dn_group_create(dn_ctx* ctx, dn_sys* const* systems, int nsys, dn_group** out) {
  if (!ctx || !systems || nsys <= 0 || !out) return fail(DN_ERR_ARG, "bad dn_group_create arguments");
  DN_TRY(group_create_any(ctx, systems, nsys, out));
  count_build_group(systems, nsys, +1);  // the systems may not be destroyed meanwhile
  (*out)->counted = true;
  return DN_OK;
}

static int group_create_leaf(dn_ctx* ctx, dn_sys* const* systems, int nsys, dn_group** out) {
  *out = nullptr;
  DN_HIP(hipSetDevice(ctx->device));
  auto* G = new dn_group();
  G->ctx = ctx;
  auto cleanup = [&](int rc) {
    dn_group_destroy(G);
    return rc;
  };
  for (int i = 0; i < nsys; ++i) {
    dn_sys* s = systems[i];
    if (!s || s->ctx != ctx) return cleanup(fail(DN_ERR_ARG, "systems must share the context"));
    if (i == 0) G->precision = s->precision;
    if (s->precision != G->precision)
      return cleanup(fail(DN_ERR_ARG, "systems of a group must share the precision"));
    if (!s->b_ready) return cleanup(fail(DN_ERR_STATE, "b has not been uploaded or generated"));
    G->sys.push_back(s);
  }
  if (G->precision != DN_FP64 && G->precision != DN_SPLIT_F16)
    return cleanup(fail(DN_ERR_ARG, "grouped builds support DN_FP64 and DN_SPLIT_F16 systems"));
  if (G->precision == DN_SPLIT_F16)
    for (dn_sys* s : G->sys)
      if (!(s->b_sym && s->use_sym))
        return cleanup(fail(DN_ERR_ARG, "grouped split-precision builds need the symmetric-b "
                                        "build (a symmetric b with symmetry enabled)"));
  hipError_t e = hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&G->done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&G->fork_ev, hipEventDisableTiming);
  for (int i = 0; i < kAux && e == hipSuccess; ++i) {
    e = hipStreamCreateWithFlags(&G->aux[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&G->join_ev[i], hipEventDisableTiming);
  }
  for (int st = 0; st < kNumStages && e == hipSuccess; ++st)
    e = hipMalloc(&G->dev[st], nsys * sizeof(GroupedGemm));
  if (e == hipSuccess) e = hipMalloc(&G->kx, nsys * sizeof(KxDesc));
  if (e == hipSuccess) e = hipMalloc(&G->diag, nsys * sizeof(DiagDesc));
  if (e == hipSuccess) e = hipMalloc(&G->symk, nsys * sizeof(SymKDesc));
  if (e == hipSuccess && G->precision == DN_FP64) {
    std::vector<PadDesc> pd;
    for (dn_sys* s : G->sys)
      if (s->Cp != s->C) pd.push_back(PadDesc{s->C, s->Cp, s->d.nbf, s->ldb, 0});
    G->npad = (int)pd.size();
    if (G->npad) {
      G->pad_blocks = pad_plan(pd.data(), G->npad);
      e = hipMalloc(&G->pad, pd.size() * sizeof(PadDesc));
      if (e == hipSuccess)
        e = hipMemcpy(G->pad, pd.data(), pd.size() * sizeof(PadDesc), hipMemcpyHostToDevice);
    }
  }
  for (int k = 0; k < 2 && e == hipSuccess && G->precision == DN_SPLIT_F16; ++k)
    e = hipMalloc(&G->sdev[k], nsys * sizeof(GroupedSplit));
  if (e == hipSuccess && G->precision == DN_SPLIT_F16) {
    std::vector<SplitPrepDesc> sp(nsys);
    for (int i = 0; i < nsys; ++i) {
      dn_sys* s = G->sys[i];
      SplitPrepDesc& q = sp[i];
      q.C = s->C;
      q.nbf = s->d.nbf;
      q.nbf5 = s->d.nbf5;
      q.Mpad = s->Mpad;
      q.Kpad = s->Kpad;
      q.naux = s->d.nbfaux;
      q.cscale = s->cscale;
      q.chi = s->chi;
      q.clo = s->clo;
      q.qmax = s->qmax;
      q.qs = s->qs;
      q.qinv = s->qinv;
    }
    split_prep_plan(sp.data(), nsys, &G->prep_blocks, &G->q_blocks);
    e = hipMalloc(&G->sprep, nsys * sizeof(SplitPrepDesc));
    if (e == hipSuccess)
      e = hipMemcpy(G->sprep, sp.data(), nsys * sizeof(SplitPrepDesc), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, std::string("dn_group_create: ") +
                                                          hipGetErrorString(e)));
  std::vector<KxDesc> kx;
  std::vector<SymKDesc> sk;
  std::vector<DiagDesc> dg(nsys);
  long col = 0;
  for (int i = 0; i < nsys; ++i) {
    dn_sys* s = G->sys[i];
    dg[i] = DiagDesc{s->C, s->HC, s->Hcore, s->d.nbf, s->d.nbf5, col};
    col += s->d.nbf5;
    if (s->b_sym && s->use_sym) {  // K, X straight from W (b_pq = W + W^T)
      SymKDesc q;
      q.W = G->precision == DN_SPLIT_F16 ? nullptr : s->bt;
      q.Wf = G->precision == DN_SPLIT_F16 ? s->bpq32 : nullptr;
      q.K = s->K;
      q.X = s->X;
      q.part = s->part;
      q.part_elems = s->part_elems;
      q.nbf5 = s->d.nbf5;
      q.naux = s->d.nbfaux;
      sk.push_back(q);
      continue;
    }
    KxDesc k;
    const long nn = (long)s->d.nbf5 * s->d.nbf5;
    k.bpq = s->bt;
    k.K = s->K;
    k.X = s->X;
    k.part = s->part;
    k.nbf5 = s->d.nbf5;
    k.naux = s->d.nbfaux;
    k.qps = 64;  // Q slices per partial sum
    k.splits = (k.naux + k.qps - 1) / k.qps;
    while ((size_t)k.splits * nn > s->part_elems) {
      k.qps *= 2;
      k.splits = (k.naux + k.qps - 1) / k.qps;
    }
    k.part_begin = G->kx_blocks[0];
    k.red_begin = G->kx_blocks[1];
    k.x_begin = G->kx_blocks[2];
    G->kx_blocks[0] += (long)k.splits * ((nn + 255) / 256);
    G->kx_blocks[1] += (nn + 255) / 256;
    G->kx_blocks[2] += ((long)k.nbf5 * k.naux + 255) / 256;
    kx.push_back(k);
  }
  G->ncols = col;
  G->nsym = (int)sk.size();
  G->nkx = (int)kx.size();
  G->symk_vec4 = true;  // 16-byte loads of the FP32 W when every nbf5 is a multiple of 4
  for (const SymKDesc& k : sk) G->symk_vec4 = G->symk_vec4 && k.nbf5 % 4 == 0;
  symk_plan(sk.data(), G->nsym, &G->symk_blocks, &G->symk_red, G->precision == DN_SPLIT_F16);
  if (G->nsym)
    e = hipMemcpy(G->symk, sk.data(), sk.size() * sizeof(SymKDesc), hipMemcpyHostToDevice);
  if (e == hipSuccess && G->nkx)
    e = hipMemcpy(G->kx, kx.data(), kx.size() * sizeof(KxDesc), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(G->diag, dg.data(), nsys * sizeof(DiagDesc), hipMemcpyHostToDevice);
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, hipGetErrorString(e)));
  for (int st = 0; st < kNumStages; ++st) {
    const int rc = plan_stage(G, st);
    if (rc != DN_OK) return cleanup(rc);
  }
  *out = G;
  return DN_OK;
}

int dn_group_destroy(dn_group* G) {
  if (!G) return DN_OK;
  (void)hipSetDevice(G->ctx->device);
  if (G->stream) (void)hipStreamSynchronize(G->stream);
  if (G->counted) count_build_group(G->sys.data(), (int)G->sys.size(), -1);
  drop_graph(G);
  for (dn_group* c : G->chains) dn_group_destroy(c);
  G->chains.clear();
  if (G->sub) dn_group_destroy(G->sub);
  G->sub = nullptr;
  for (dn_sys* s : G->indep) (void)hipStreamSynchronize(s->stream);
  for (auto e : G->indep_ev) (void)hipEventDestroy(e);
  for (int st = 0; st < kNumStages; ++st)
    if (G->dev[st]) (void)hipFree(G->dev[st]);
  if (G->kx) (void)hipFree(G->kx);
  if (G->diag) (void)hipFree(G->diag);
  if (G->symk) (void)hipFree(G->symk);
  if (G->pad) (void)hipFree(G->pad);
  for (int k = 0; k < 2; ++k)
    if (G->sdev[k]) (void)hipFree(G->sdev[k]);
  if (G->sprep) (void)hipFree(G->sprep);
  for (auto e : G->ev_start) (void)hipEventDestroy(e);
  for (auto e : G->ev_stop) (void)hipEventDestroy(e);
  if (G->done) (void)hipEventDestroy(G->done);
  for (int i = 0; i < kAux; ++i) {
    if (G->aux[i]) (void)hipStreamSynchronize(G->aux[i]);
    if (G->aux[i]) (void)hipStreamDestroy(G->aux[i]);
    if (G->join_ev[i]) (void)hipEventDestroy(G->join_ev[i]);
  }
  if (G->fork_ev) (void)hipEventDestroy(G->fork_ev);
  if (G->stream) (void)hipStreamDestroy(G->stream);
  delete G;
  return DN_OK;
}

}  // extern "C"

namespace dn {
namespace {

// Times every tile configuration x workgroup count of every class of a leaf group (its
// inputs must hold real values: one build has run) and keeps the fastest.
int tune_leaf(dn_group* G) {
  for (int st = 0; st < kNumStages; ++st) {
    for (size_t li = 0; li < G->launches[st].size(); ++li) {
      // tile shape x workgroup count (1, 2 or 4 times the resident count, or one per tile)
      static const int grids[] = {0, -2, -4, -(1 << 20)};
      int best = G->launches[st][li].config, best_grid = 0;
      float best_ms = 1e30f;
      const bool sp = split_stage(G, st);
      for (int c = 0; c < (sp ? (int)kNumSplitConfigs : (int)kNumGemmConfigs); ++c) {
        if (sp ? !split_config_ok(c, st == kK1 ? 0 : 1)
               : !gemm_config_ok(c, G->host[st][G->launches[st][li].first].g))
          continue;
        DN_TRY(set_config(G, st, li, c));
        for (int gr : grids) {
          G->launches[st][li].grid = gr;
          float ms = 0.f;
          DN_TRY(time_launch(G, st, li, &ms));
          if (ms < best_ms) {
            best_ms = ms;
            best = c;
            best_grid = gr;
          }
        }
      }
      DN_TRY(set_config(G, st, li, best));
      G->launches[st][li].grid = best_grid;
    }
  }
  G->tuned = true;
  return DN_OK;
}

// The placement with the systems of nbf5 <= max_nbf5 grouped and the others independent
// (every system grouped: no placement).  Replaces an earlier placement.
int place(dn_group* G, int max_nbf5) {
  if (G->sub) DN_TRY(dn_group_destroy(G->sub));
  G->sub = nullptr;
  for (dn_sys* s : G->indep) DN_HIP(hipStreamSynchronize(s->stream));
  G->indep.clear();
  std::vector<dn_sys*> grouped;
  for (dn_sys* s : G->sys) (s->d.nbf5 <= max_nbf5 ? grouped : G->indep).push_back(s);
  G->place_nbf5 = max_nbf5;
  G->placed = !G->indep.empty();
  if (!G->placed) return DN_OK;
  while (G->indep_ev.size() < G->indep.size()) {
    hipEvent_t e = nullptr;
    DN_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    G->indep_ev.push_back(e);
  }
  if (!grouped.empty()) DN_TRY(group_create_any(G->ctx, grouped.data(), (int)grouped.size(), &G->sub));
  return DN_OK;
}

// Tunes a group's launches on its current inputs (one eager build first) and captures it.
int tune_group(dn_group* G) {
  DN_TRY(join_systems(G));
  DN_TRY(launch_build(G));
  DN_HIP(hipStreamSynchronize(G->stream));
  std::vector<dn_group*> leaves = G->chains;
  if (leaves.empty()) leaves.push_back(G);
  for (dn_group* c : leaves) DN_TRY(tune_leaf(c));
  DN_TRY(capture(G));
  DN_HIP(hipStreamSynchronize(G->stream));
  G->tuned = true;
  return DN_OK;
}

// One build of every system under the placement: the grouped part's graph on its stream
// and each independent system's single-system build on its own stream, all after the
// work issued on G->stream so far, which then waits for every part.
int launch_placed(dn_group* G) {
  DN_HIP(hipEventRecord(G->fork_ev, G->stream));
  if (G->sub) {
    DN_HIP(hipStreamWaitEvent(G->sub->stream, G->fork_ev, 0));
    DN_TRY(dn_group_jkh(G->sub));
    DN_HIP(hipEventRecord(G->sub->done, G->sub->stream));
    DN_HIP(hipStreamWaitEvent(G->stream, G->sub->done, 0));
  }
  for (size_t i = 0; i < G->indep.size(); ++i) {
    dn_sys* s = G->indep[i];
    DN_HIP(hipStreamWaitEvent(s->stream, G->fork_ev, 0));
    DN_TRY(jkh_resident(s));
    DN_HIP(hipEventRecord(G->indep_ev[i], s->stream));
    DN_HIP(hipStreamWaitEvent(G->stream, G->indep_ev[i], 0));
  }
  return DN_OK;
}

// Event time of one build of every system under the current placement (after a warm-up).
int time_placed(dn_group* G, float* ms) {
  hipEvent_t a = nullptr, b = nullptr;
  DN_HIP(hipEventCreate(&a));
  DN_HIP(hipEventCreate(&b));
  auto one = [&]() { return G->placed ? launch_placed(G) : dn_group_jkh(G); };
  int rc = one();
  if (rc == DN_OK && hipStreamSynchronize(G->stream) != hipSuccess) rc = DN_ERR_HIP;
  if (rc == DN_OK && hipEventRecord(a, G->stream) != hipSuccess) rc = DN_ERR_HIP;
  for (int r = 0; r < 3 && rc == DN_OK; ++r) rc = one();
  if (rc == DN_OK && hipEventRecord(b, G->stream) != hipSuccess) rc = DN_ERR_HIP;
  if (rc == DN_OK && hipEventSynchronize(b) != hipSuccess) rc = DN_ERR_HIP;
  if (rc == DN_OK && hipEventElapsedTime(ms, a, b) != hipSuccess) rc = DN_ERR_HIP;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc == DN_ERR_HIP) return fail(rc, "timing a placement of the group");
  *ms /= 3.f;
  return rc;
}

// Grouped launches win where a system's build is too small to fill the GPU (launch-bound
// stages); a large system's single-system build on its own stream reaches the same
// throughput, and independent streams overlap its stages with other systems' at a finer
// grain than a stage-ordered grouped chain does.  The placement search times the whole
// build of all systems with the smallest systems grouped up to a few nbf5 limits (their
// share of the build work at most a half), with every system grouped and with none, and
// keeps the fastest.
int search_placement(dn_group* G) {
  float best_ms = 0.f;
  DN_TRY(time_placed(G, &best_ms));  // every system grouped (tuned above)
  int best = INT32_MAX;
  double total = 0.0;
  for (dn_sys* s : G->sys) total += (double)s->d.nbf * s->d.nbf5 * s->d.nbfaux * (s->d.nbf + 2.0 * s->d.nbf5);
  static const int limits[] = {-1, 32, 64, 96, 128, 160, 192};
  int prev_count = -1;
  for (int lim : limits) {
    double w = 0.0;
    int count = 0;
    for (dn_sys* s : G->sys)
      if (s->d.nbf5 <= lim) {
        w += (double)s->d.nbf * s->d.nbf5 * s->d.nbfaux * (s->d.nbf + 2.0 * s->d.nbf5);
        ++count;
      }
    if (count == prev_count || count == (int)G->sys.size() || w > 0.5 * total) continue;
    prev_count = count;
    DN_TRY(place(G, lim));
    if (G->sub) DN_TRY(tune_group(G->sub));
    float ms = 0.f;
    DN_TRY(time_placed(G, &ms));
    if (ms < best_ms) {
      best_ms = ms;
      best = lim;
    }
  }
  if (best != G->place_nbf5 || !G->placed) {
    DN_TRY(place(G, best));
    if (G->sub) DN_TRY(tune_group(G->sub));
  }
  return DN_OK;
}

}  // namespace
}  // namespace dn

extern "C" {

int dn_group_place(dn_group* G, int max_grouped_nbf5) {
  if (!G) return fail(DN_ERR_ARG, "null dn_group");
  if (!G->counted) return fail(DN_ERR_ARG, "dn_group_place: not a group made by dn_group_create");
  DN_HIP(hipSetDevice(G->ctx->device));
  DN_HIP(hipStreamSynchronize(G->stream));
  DN_TRY(place(G, max_grouped_nbf5));
  G->place_fixed = true;
  G->tuned = false;
  return DN_OK;
}

int dn_group_tune(dn_group* G, int* configs) {
  if (!G) return fail(DN_ERR_ARG, "null dn_group");
  DN_HIP(hipSetDevice(G->ctx->device));
  // one eager build first: every stage's inputs (C planes and per-Q scales of split
  // groups included) then hold real values while candidate launches are timed
  if (G->place_fixed && G->placed) {
    if (G->sub) DN_TRY(tune_group(G->sub));
  } else {
    DN_TRY(tune_group(G));
    if (!G->place_fixed) DN_TRY(search_placement(G));
  }
  G->tuned = true;
  dn_group* T = G->placed ? G->sub : G;  // the grouped part
  std::vector<dn_group*> leaves;
  if (T) {
    leaves = T->chains;
    if (leaves.empty()) leaves.push_back(T);
  }
  if (configs)  // the choice for each stage's class with the most work (-1: none grouped)
    for (int st = 0; st < kNumStages; ++st) {
      configs[st] = -1;
      double fmax = -1.0;
      for (dn_group* c : leaves)
        for (const auto& l : c->launches[st])
          if (l.flops > fmax) {
            fmax = l.flops;
            configs[st] = split_stage(c, st) ? 100 + l.config : l.config;
          }
    }
  return DN_OK;
}

int dn_group_plan_text(dn_group* G, char* buf, int len) {
  if (!G || !buf || len <= 0) return fail(DN_ERR_ARG, "bad dn_group_plan_text arguments");
  static const char* names[] = {"128x128", "64x64", "64x128", "128x64", "64x32", "32x64",
                                "lds64x128", "lds128x128", "lds128x256", "lds64x192",
                                "64x192"};
  static const char* snames[] = {"s128x128", "s64x128", "s64x64", "s32x64"};
  static_assert(sizeof(names) / sizeof(names[0]) == kNumGemmConfigs, "a name per config");
  static_assert(sizeof(snames) / sizeof(snames[0]) == kNumSplitConfigs, "a name per config");
  static const char* stages[] = {"K1", "K2", "J", "HC"};
  std::string t;
  if (G->placed) {
    t = "placement: " + std::to_string(G->sys.size() - G->indep.size()) + " grouped (nbf5 <= " +
        std::to_string(G->place_nbf5) + "), " + std::to_string(G->indep.size()) +
        " single-system builds on their own streams";
    if (G->sub) {
      std::vector<char> one(16384);
      DN_TRY(dn_group_plan_text(G->sub, one.data(), (int)one.size()));
      t += " | " + std::string(one.data());
    }
    std::snprintf(buf, (size_t)len, "%s", t.c_str());
    return DN_OK;
  }
  if (!G->chains.empty()) {
    for (size_t c = 0; c < G->chains.size(); ++c) {
      std::vector<char> one(8192);
      DN_TRY(dn_group_plan_text(G->chains[c], one.data(), (int)one.size()));
      t += "chain " + std::to_string(c) + " (" + std::to_string(G->chains[c]->sys.size()) +
           " systems): " + one.data() + (c + 1 < G->chains.size() ? " | " : "");
    }
    std::snprintf(buf, (size_t)len, "%s", t.c_str());
    return DN_OK;
  }
  for (int st = 0; st < kNumStages; ++st) {
    t += stages[st];
    t += ":";
    for (const auto& l : G->launches[st]) {
      char one[96];
      char grid[16];
      if (l.grid <= -(1 << 20)) std::snprintf(grid, sizeof(grid), "all");
      else std::snprintf(grid, sizeof(grid), "r%d", l.grid < 0 ? -l.grid : 1);
      std::snprintf(one, sizeof(one), " %dx%s/%s(%.1fGF)", l.count,
                    split_stage(G, st) ? snames[l.config] : names[l.config], grid,
                    l.flops * 1e-9);
      t += one;
    }
    t += st + 1 < kNumStages ? "; " : "";
  }
  std::snprintf(buf, (size_t)len, "%s", t.c_str());
  return DN_OK;
}

int dn_group_jkh(dn_group* G) {
  if (!G) return fail(DN_ERR_ARG, "null dn_group");
  DN_HIP(hipSetDevice(G->ctx->device));
  if (G->placed) {  // grouped part and independent systems (dn_group_place / dn_group_tune)
    if (G->timing) {
      if (G->ev_used == G->ev_start.size()) {
        hipEvent_t a = nullptr, b = nullptr;
        DN_HIP(hipEventCreate(&a));
        DN_HIP(hipEventCreate(&b));
        G->ev_start.push_back(a);
        G->ev_stop.push_back(b);
      }
      DN_HIP(hipEventRecord(G->ev_start[G->ev_used], G->stream));
    }
    DN_TRY(launch_placed(G));
    if (G->timing) DN_HIP(hipEventRecord(G->ev_stop[G->ev_used++], G->stream));
    return DN_OK;
  }
  if (!G->exec) DN_TRY(capture(G));
  DN_TRY(join_systems(G));
  if (G->timing) {
    if (G->ev_used == G->ev_start.size()) {
      hipEvent_t a = nullptr, b = nullptr;
      DN_HIP(hipEventCreate(&a));
      DN_HIP(hipEventCreate(&b));
      G->ev_start.push_back(a);
      G->ev_stop.push_back(b);
    }
    DN_HIP(hipEventRecord(G->ev_start[G->ev_used], G->stream));
  }
  DN_HIP(hipGraphLaunch(G->exec, G->stream));
  if (G->timing) DN_HIP(hipEventRecord(G->ev_stop[G->ev_used++], G->stream));
  return release_systems(G);
}

int dn_group_sync(dn_group* G) {
  if (!G) return fail(DN_ERR_ARG, "null dn_group");
  DN_HIP(hipSetDevice(G->ctx->device));
  DN_HIP(hipStreamSynchronize(G->stream));
  return DN_OK;
}

int dn_group_set_timing(dn_group* G, int on) {
  if (!G) return fail(DN_ERR_ARG, "null dn_group");
  G->timing = on != 0;
  return DN_OK;
}

int dn_group_get_timing(dn_group* G, double* total_ms, long* builds) {
  if (!G) return fail(DN_ERR_ARG, "null dn_group");
  DN_HIP(hipSetDevice(G->ctx->device));
  DN_HIP(hipStreamSynchronize(G->stream));
  for (size_t i = 0; i < G->ev_used; ++i) {
    float ms = 0.f;
    DN_HIP(hipEventElapsedTime(&ms, G->ev_start[i], G->ev_stop[i]));
    G->build_ms += ms;
    G->builds += 1;
  }
  G->ev_used = 0;
  if (total_ms) *total_ms = G->build_ms;
  if (builds) *builds = G->builds;
  G->build_ms = 0.0;
  G->builds = 0;
  return DN_OK;
}

}  // extern "C"
