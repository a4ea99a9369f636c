// Grouped J/K/H_core builds: many resident systems of uneven size built together.
//
// Each system's build is the one of integrals.jl:144-172 (JKH_MO).  The group turns the
// per-system stages into grouped launches over all systems:
//   K1  b_pn = C5^T b                 persistent grouped GEMM (the dominant stage); for a
//                                     symmetric b only its lower half: T = C5^T Lh
//   K2  b_pq[:,:,Q] = b_pn[:,:,Q] C5  persistent grouped GEMM, one tile list over (sys, Q)
//   SY  b_pq = W + W^T                 systems whose K1 used the lower half
//   KX  K = sum_Q b_pq^2, X = diag     one workgroup per (system, column q)
//   J   J = X X^T                     persistent grouped GEMM
//   HC  H C5                          persistent grouped GEMM
//   HD  H_core_i = C5[:,i] . (H C5)[:,i]
// An autotuner times every tile configuration of each GEMM stage on the actual group,
// keeps the fastest, and the whole sequence is then captured once into a HIP graph that
// every build replays.
#include "dn_internal.h"

namespace dn {
namespace {

struct KxDesc {
  const double* bpq;  // nbf5 x nbf5 x naux, slice stride nbf5^2
  double* K;
  double* X;          // nbf5 x naux
  int nbf5, naux;
  long col_begin;     // first column of this system in the grouped column list
};

struct DiagDesc {
  const double* C;
  const double* HC;
  double* Hcore;
  int nbf, nbf5;
  long col_begin;
};

template <typename D>
__device__ __forceinline__ int find_desc(const D* d, int n, long col) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].col_begin <= col) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One workgroup of 256 threads per (system, q): K[:, q] and X[q, :].
__global__ __launch_bounds__(256) void group_kx_kernel(const KxDesc* desc, int n) {
  __shared__ double red[4][64];
  const long col = blockIdx.x;
  const KxDesc& d = desc[find_desc(desc, n, col)];
  const int q = (int)(col - d.col_begin);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long slice = (long)d.nbf5 * d.nbf5;
  const double* base = d.bpq + (long)q * d.nbf5;
  for (int p0 = 0; p0 < d.nbf5; p0 += 64) {
    const int p = p0 + lane;
    double acc = 0.0;
    if (p < d.nbf5)
      for (int Q = w; Q < d.naux; Q += 4) {
        const double v = base[p + Q * slice];
        acc = fma(v, v, acc);
      }
    red[w][lane] = acc;
    __syncthreads();
    if (w == 0 && p < d.nbf5)
      d.K[p + (long)q * d.nbf5] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    __syncthreads();
  }
  for (int Q = threadIdx.x; Q < d.naux; Q += blockDim.x)
    d.X[q + (long)Q * d.nbf5] = base[q + Q * slice];
}

// One wavefront per (system, i): H_core_i = sum_m C[m, i] HC[m, i].
__global__ __launch_bounds__(64) void group_hdiag_kernel(const DiagDesc* desc, int n) {
  const long col = blockIdx.x;
  const DiagDesc& d = desc[find_desc(desc, n, col)];
  const int i = (int)(col - d.col_begin);
  double acc = 0.0;
  for (int m = threadIdx.x; m < d.nbf; m += 64)
    acc = fma(d.C[m + (long)i * d.nbf], d.HC[m + (long)i * d.nbf], acc);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (threadIdx.x == 0) d.Hcore[i] = acc;
}

enum Stage { kK1 = 0, kK2 = 1, kJ = 2, kHC = 3, kNumStages = 4 };

}  // namespace
}  // namespace dn

using namespace dn;

struct dn_group {
  dn_ctx* ctx = nullptr;
  std::vector<dn_sys*> sys;
  int precision = DN_FP64;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // Problems of a stage sorted into classes by operand alignment (16-byte loads allowed
  // for A, for B): one persistent launch per class, so one odd leading dimension does
  // not take the vector loads away from every other system.
  struct Launch {
    int first, count;
    long tiles;
  };
  std::vector<GroupedGemm> host[kNumStages];
  std::vector<Launch> launches[kNumStages];
  GroupedGemm* dev[kNumStages] = {};
  int config[kNumStages] = {kTile128x128, kTile128x128, kTile64x64, kTile64x64};
  KxDesc* kx = nullptr;
  DiagDesc* diag = nullptr;
  SymDesc* sym = nullptr;
  int nsym = 0;
  long sym_blocks = 0;
  long ncols = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool tuned = false;
  bool timing = false;
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
    case kK1:
      if (s->b_sym && s->use_sym) {  // T[n,p,Q] = sum_{m >= n} Lh[m,n,Q] C5[m,p]
        g.M = d.nbf; g.N = d.nbf5; g.K = d.nbf;
        g.A = s->b; g.lda = d.nbf; g.sA = (long)d.nbf * d.nbf; g.a_kcontig = true;
        g.B = s->C; g.ldb = d.nbf; g.sB = 0; g.b_kcontig = true;
        g.C = s->bpn; g.ldc = d.nbf; g.sC = (long)d.nbf * d.nbf5;
        g.nbatch = d.nbfaux; g.bps = 1;
        g.tri = kTriA;
      } else {
        g.M = d.nbf5; g.N = d.nbf * d.nbfaux; g.K = d.nbf;
        g.A = s->C; g.lda = d.nbf; g.a_kcontig = true;
        g.B = s->b; g.ldb = d.nbf; g.b_kcontig = true;
        g.C = s->bpn; g.ldc = d.nbf5;
      }
      break;
    case kK2:
      if (s->b_sym && s->use_sym) {  // W^T = C5^T T with (p, Q) flattened
        g.M = d.nbf5; g.N = d.nbf5 * d.nbfaux; g.K = d.nbf;
        g.A = s->C; g.lda = d.nbf; g.a_kcontig = true;
        g.B = s->bpn; g.ldb = d.nbf; g.b_kcontig = true;
        g.C = s->bt; g.ldc = d.nbf5;
      } else {
        g.M = d.nbf5; g.N = d.nbf5; g.K = d.nbf;
        g.A = s->bpn; g.lda = d.nbf5; g.sA = (long)d.nbf5 * d.nbf; g.a_kcontig = false;
        g.B = s->C; g.ldb = d.nbf; g.sB = 0; g.b_kcontig = true;
        g.C = s->bt; g.ldc = d.nbf5; g.sC = (long)d.nbf5 * d.nbf5;
        g.nbatch = d.nbfaux; g.bps = 1;
      }
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

int plan_stage(dn_group* G, int stage, int config) {
  DN_HIP(hipStreamSynchronize(G->stream));  // a replay in flight may read the descriptors
  auto& h = G->host[stage];
  h.resize(G->sys.size());
  for (size_t i = 0; i < G->sys.size(); ++i) h[i].g = stage_gemm(G->sys[i], stage);
  gemm_grouped_plan(h.data(), (int)h.size(), config);  // sets the alignment flags
  auto key = [](const GroupedGemm& d) {
    return 16 * d.g.tri + 8 * d.g.a_kcontig + 4 * d.g.b_kcontig + 2 * !d.g.vec_a + !d.g.vec_b;
  };
  std::stable_sort(h.begin(), h.end(),
                   [&](const GroupedGemm& a, const GroupedGemm& b) { return key(a) < key(b); });
  auto& L = G->launches[stage];
  L.clear();
  for (int i = 0; i < (int)h.size();) {
    int j = i;
    while (j < (int)h.size() && key(h[j]) == key(h[i])) ++j;
    L.push_back({i, j - i, gemm_grouped_plan(h.data() + i, j - i, config)});
    i = j;
  }
  G->config[stage] = config;
  DN_HIP(hipMemcpy(G->dev[stage], h.data(), h.size() * sizeof(GroupedGemm),
                   hipMemcpyHostToDevice));
  return DN_OK;
}

int launch_stage(dn_group* G, int stage) {
  for (const auto& l : G->launches[stage])
    DN_TRY(gemm_f64_grouped(G->stream, G->host[stage].data() + l.first,
                            G->dev[stage] + l.first, l.count, l.tiles, G->config[stage], 0));
  return DN_OK;
}

int launch_build(dn_group* G) {
  const int n = (int)G->sys.size();
  DN_TRY(launch_stage(G, kK1));
  DN_TRY(launch_stage(G, kK2));
  DN_TRY(sym_grouped(G->stream, G->sym, G->nsym, G->sym_blocks));
  hipLaunchKernelGGL(group_kx_kernel, dim3(G->ncols), dim3(256), 0, G->stream, G->kx, n);
  DN_HIP(hipGetLastError());
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

int time_stage(dn_group* G, int stage, float* ms) {
  hipEvent_t a, b;
  DN_HIP(hipEventCreate(&a));
  DN_HIP(hipEventCreate(&b));
  int rc = launch_stage(G, stage);  // warm-up (code object load, caches)
  if (rc == DN_OK) {
    (void)hipEventRecord(a, G->stream);
    for (int r = 0; r < 2 && rc == DN_OK; ++r) rc = launch_stage(G, stage);
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

int dn_group_create(dn_ctx* ctx, dn_sys* const* systems, int nsys, dn_group** out) {
  if (!ctx || !systems || nsys <= 0 || !out) return fail(DN_ERR_ARG, "bad dn_group_create arguments");
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
  if (G->precision != DN_FP64)
    return cleanup(fail(DN_ERR_ARG, "grouped builds support DN_FP64 systems"));
  hipError_t e = hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&G->done, hipEventDisableTiming);
  for (int st = 0; st < kNumStages && e == hipSuccess; ++st)
    e = hipMalloc(&G->dev[st], nsys * sizeof(GroupedGemm));
  if (e == hipSuccess) e = hipMalloc(&G->kx, nsys * sizeof(KxDesc));
  if (e == hipSuccess) e = hipMalloc(&G->diag, nsys * sizeof(DiagDesc));
  if (e == hipSuccess) e = hipMalloc(&G->sym, nsys * sizeof(SymDesc));
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, std::string("dn_group_create: ") +
                                                          hipGetErrorString(e)));
  std::vector<KxDesc> kx(nsys);
  std::vector<DiagDesc> dg(nsys);
  long col = 0;
  for (int i = 0; i < nsys; ++i) {
    dn_sys* s = G->sys[i];
    kx[i] = KxDesc{s->bt, s->K, s->X, s->d.nbf5, s->d.nbfaux, col};
    dg[i] = DiagDesc{s->C, s->HC, s->Hcore, s->d.nbf, s->d.nbf5, col};
    col += s->d.nbf5;
  }
  G->ncols = col;
  std::vector<SymDesc> sy;
  for (dn_sys* s : G->sys)
    if (s->b_sym && s->use_sym) {
      SymDesc q;
      q.bpq = s->bt;
      q.nbf5 = s->d.nbf5;
      q.naux = s->d.nbfaux;
      sy.push_back(q);
    }
  G->nsym = (int)sy.size();
  G->sym_blocks = sym_grouped_plan(sy.data(), G->nsym);
  if (G->nsym)
    e = hipMemcpy(G->sym, sy.data(), sy.size() * sizeof(SymDesc), hipMemcpyHostToDevice);
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, hipGetErrorString(e)));
  e = hipMemcpy(G->kx, kx.data(), nsys * sizeof(KxDesc), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(G->diag, dg.data(), nsys * sizeof(DiagDesc), hipMemcpyHostToDevice);
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, hipGetErrorString(e)));
  for (int st = 0; st < kNumStages; ++st) {
    const int rc = plan_stage(G, st, G->config[st]);
    if (rc != DN_OK) return cleanup(rc);
  }
  *out = G;
  return DN_OK;
}

int dn_group_destroy(dn_group* G) {
  if (!G) return DN_OK;
  (void)hipSetDevice(G->ctx->device);
  if (G->stream) (void)hipStreamSynchronize(G->stream);
  drop_graph(G);
  for (int st = 0; st < kNumStages; ++st)
    if (G->dev[st]) (void)hipFree(G->dev[st]);
  if (G->kx) (void)hipFree(G->kx);
  if (G->diag) (void)hipFree(G->diag);
  if (G->sym) (void)hipFree(G->sym);
  for (auto e : G->ev_start) (void)hipEventDestroy(e);
  for (auto e : G->ev_stop) (void)hipEventDestroy(e);
  if (G->done) (void)hipEventDestroy(G->done);
  if (G->stream) (void)hipStreamDestroy(G->stream);
  delete G;
  return DN_OK;
}

int dn_group_tune(dn_group* G, int* configs) {
  if (!G) return fail(DN_ERR_ARG, "null dn_group");
  DN_HIP(hipSetDevice(G->ctx->device));
  DN_TRY(join_systems(G));
  for (int st = 0; st < kNumStages; ++st) {
    int best = G->config[st];
    float best_ms = 1e30f;
    for (int c = 0; c < kNumGemmConfigs; ++c) {
      DN_TRY(plan_stage(G, st, c));
      float ms = 0.f;
      DN_TRY(time_stage(G, st, &ms));
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    DN_TRY(plan_stage(G, st, best));
  }
  DN_TRY(capture(G));
  DN_HIP(hipStreamSynchronize(G->stream));
  G->tuned = true;
  if (configs)
    for (int st = 0; st < kNumStages; ++st) configs[st] = G->config[st];
  return DN_OK;
}

int dn_group_jkh(dn_group* G) {
  if (!G) return fail(DN_ERR_ARG, "null dn_group");
  DN_HIP(hipSetDevice(G->ctx->device));
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
