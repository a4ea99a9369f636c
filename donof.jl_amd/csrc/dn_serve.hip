// Lock-step ADAM-family optimisation of several resident systems: every step advances all
// live solves together, the orbital Lagrangian of all of them (the dominant work,
// utils.jl:1-59) running as grouped persistent GEMMs captured once in a HIP graph.
//
// Per system the recurrence is the one of the single-system loop (minimization.jl:344-438,
// dn_orbopt_resident): the same convergence test, moments, step, rotation and best-energy
// bookkeeping, so a solve gives the same result in any group.  The tile shape of every
// grouped GEMM depends only on that system's own sizes, which keeps its arithmetic
// independent of the group it runs in.
#include <cmath>
#include <cstring>

#include "dn_internal.h"
#include "dn_occ.h"

namespace dn {

// dn_rot.hip / dn_lag.hip / dn_solve.hip
int norm1(hipStream_t s, const double* Y, int nbf, double* out);
int expm_taylor(hipStream_t s, const double* Y, double nrm, int nbf, double* work, double* U,
                double* part, size_t part_elems);
int matmul_sq(hipStream_t s, const double* A, const double* B, double* out, int nbf,
              double* part, size_t part_elems);
int pack_grad(hipStream_t s, const double* elag, int nbf, int nbf5, double* g,
              unsigned long long* gmax_bits);
int adam_step(hipStream_t s, const double* g, double* m, double* v, double* vmax, double* y,
              int method, double beta1, double beta2, double bc1, double bc2, double alpha,
              int nbf, int nbf5, double* Y);
int lag_w(hipStream_t s, const double* bt, long slice, int nbf5, int naux, const double* cj12,
          const double* n, int nalpha, int nbeta, double* X, double* w);
int lag_bmat(hipStream_t s, const double* bt, long slice, int nbf5, int naux,
             const double* ck12, const double* w, double* bmat);
int lag_init(hipStream_t s, const double* Hmat, const double* n, int nbf, int nbf5,
             double* elag);
int energy_elec(hipStream_t s, const double* elag, const double* Hmat, const double* n,
                int nbf, int nbf5, double* out);
int orbopt_begin(dn_sys* s, const dn_functional* f, const dn_orbopt_state* st);
int fill_dev(hipStream_t s, double* p, long n, double v);
int copy_dev(hipStream_t s, double* dst, const double* src, long n);
size_t lag_split_count(const dn_dims& d);

namespace {

enum LStage { kT1 = 0, kT2 = 1, kHC = 2, kHM = 3, kLG = 4, kNumL = 5 };

// out[2 i + k] = scal_i[k] (E and the bits of max|g|) for every system
__global__ void gather2_kernel(const double* const* __restrict__ src, int n, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    out[2 * i] = src[i][0];
    out[2 * i + 1] = src[i][1];
  }
}

Gemm lag_gemm(dn_sys* s, int stage, int* config) {
  const dn_dims& d = s->d;
  Gemm g;
  switch (stage) {
    case kT1:  // b_pn = C5^T b  (integrals.jl:159 / utils.jl:19)
      g.M = d.nbf5; g.N = d.nbf * d.nbfaux; g.K = d.nbf;
      g.A = s->C; g.lda = d.nbf; g.a_kcontig = true;
      g.B = s->b; g.ldb = d.nbf; g.b_kcontig = true;
      g.C = s->bpn; g.ldc = d.nbf5;
      *config = d.nbf5 > 64 ? kTile128x128 : kTile64x128;
      break;
    case kT2:  // b_MO^T[:, :, Q] = b_pn[:, :, Q] C  (utils.jl:19-20)
      g.M = d.nbf5; g.N = d.nbf; g.K = d.nbf;
      g.A = s->bpn; g.lda = d.nbf5; g.sA = (long)d.nbf5 * d.nbf; g.a_kcontig = false;
      g.B = s->C; g.ldb = d.nbf; g.sB = 0; g.b_kcontig = true;
      g.C = s->bt; g.ldc = d.nbf5; g.sC = (long)d.nbf5 * d.nbf;
      g.nbatch = d.nbfaux; g.bps = 1;
      *config = d.nbf5 > 64 ? kTile128x128 : kTile64x64;
      break;
    case kHC:  // H C5
      g.M = d.nbf; g.N = d.nbf5; g.K = d.nbf;
      g.A = s->H; g.lda = d.nbf; g.a_kcontig = false;
      g.B = s->C; g.ldb = d.nbf; g.b_kcontig = true;
      g.C = s->HC; g.ldc = d.nbf;
      *config = kTile64x64;
      break;
    case kHM:  // H_mat = C^T H C5
      g.M = d.nbf; g.N = d.nbf5; g.K = d.nbf;
      g.A = s->C; g.lda = d.nbf; g.a_kcontig = true;
      g.B = s->HC; g.ldb = d.nbf; g.b_kcontig = true;
      g.C = s->Hmat; g.ldc = d.nbf;
      *config = kTile64x64;
      break;
    default: {  // elag partials: sum_Q sum_q bt[q,a,Q] bmat[q,b,Q] over Q ranges
      const int nsplit = (int)lag_split_count(d);
      const int bps = (d.nbfaux + nsplit - 1) / nsplit;
      g.M = d.nbf; g.N = d.nbf5; g.K = d.nbf5;
      g.A = s->bt; g.lda = d.nbf5; g.sA = (long)d.nbf5 * d.nbf; g.a_kcontig = true;
      g.B = s->bmat; g.ldb = d.nbf5; g.sB = (long)d.nbf5 * d.nbf5; g.b_kcontig = true;
      g.C = s->part; g.ldc = d.nbf; g.sC = (long)d.nbf * d.nbf5;
      g.nbatch = d.nbfaux; g.bps = bps;
      *config = kTile64x64;
      break;
    }
  }
  return g;
}

}  // namespace
}  // namespace dn

using namespace dn;

struct dn_ogroup {
  dn_ctx* ctx = nullptr;
  std::vector<dn_sys*> sys;
  std::vector<hipStream_t> saved_streams;
  hipStream_t stream = nullptr;
  struct Launch {
    int first, count, config;
    long tiles;
  };
  std::vector<GroupedGemm> host[kNumL];
  std::vector<Launch> launches[kNumL];
  GroupedGemm* dev[kNumL] = {};
  double** scal_ptrs = nullptr;  // device array of each system's scal
  double* gathered = nullptr;    // device: E, max|g| bits per system
  double* pinned = nullptr;      // host (pinned) copy
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace dn {
namespace {

int plan_lstage(dn_ogroup* G, int stage) {
  auto& h = G->host[stage];
  const int n = (int)G->sys.size();
  h.resize(n);
  std::vector<int> cfg(n);
  for (int i = 0; i < n; ++i) h[i].g = lag_gemm(G->sys[i], stage, &cfg[i]);
  gemm_grouped_plan(h.data(), n, kTile64x64);  // alignment flags
  // one launch per class: tile configuration, operand layout and alignment
  std::vector<int> key(n), order(n);
  for (int i = 0; i < n; ++i) {
    const Gemm& g = h[i].g;
    key[i] = cfg[i] * 16 + 8 * g.a_kcontig + 4 * g.b_kcontig + 2 * !g.vec_a + !g.vec_b;
    order[i] = i;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  std::vector<GroupedGemm> sorted(n);
  for (int i = 0; i < n; ++i) sorted[i] = h[order[i]];
  h = sorted;
  auto& L = G->launches[stage];
  L.clear();
  for (int i = 0; i < n;) {
    int j = i;
    while (j < n && key[order[j]] == key[order[i]]) ++j;
    const int c = cfg[order[i]];
    L.push_back({i, j - i, c, gemm_grouped_plan(h.data() + i, j - i, c)});
    i = j;
  }
  DN_HIP(hipMemcpy(G->dev[stage], h.data(), n * sizeof(GroupedGemm), hipMemcpyHostToDevice));
  return DN_OK;
}

int launch_lstage(dn_ogroup* G, int stage) {
  for (const auto& l : G->launches[stage])
    DN_TRY(gemm_f64_grouped(G->stream, G->host[stage].data() + l.first, G->dev[stage] + l.first,
                            l.count, l.tiles, l.config, 0));
  return DN_OK;
}

// The Lagrangian, energy and packed gradient of every system (the graph body).
int lagrangian_all(dn_ogroup* G) {
  DN_TRY(launch_lstage(G, kT1));
  DN_TRY(launch_lstage(G, kT2));
  DN_TRY(launch_lstage(G, kHC));
  DN_TRY(launch_lstage(G, kHM));
  for (dn_sys* s : G->sys) {
    const dn_dims& d = s->d;
    const long slice = (long)d.nbf5 * d.nbf;
    DN_TRY(lag_w(G->stream, s->bt, slice, d.nbf5, d.nbfaux, s->cj12, s->n, s->opt.nalpha,
                 s->opt.nbeta, s->X, s->w));
    DN_TRY(lag_bmat(G->stream, s->bt, slice, d.nbf5, d.nbfaux, s->ck12, s->w, s->bmat));
    DN_TRY(lag_init(G->stream, s->Hmat, s->n, d.nbf, d.nbf5, s->elag));
  }
  DN_TRY(launch_lstage(G, kLG));
  for (dn_sys* s : G->sys) {
    const dn_dims& d = s->d;
    const int nsplit = (int)lag_split_count(d);
    const int bps = (d.nbfaux + nsplit - 1) / nsplit;
    const int used = (d.nbfaux + bps - 1) / bps;
    DN_TRY(reduce_splits(G->stream, s->part, (long)d.nbf * d.nbf5, used, d.nbf, d.nbf5, d.nbf,
                         s->elag, d.nbf, 1.0));
    DN_TRY(energy_elec(G->stream, s->elag, s->Hmat, s->n, d.nbf, d.nbf5, s->scal));
    DN_TRY(pack_grad(G->stream, s->elag, d.nbf, d.nbf5, s->gv,
                     reinterpret_cast<unsigned long long*>(s->scal + 1)));
  }
  const int n = (int)G->sys.size();
  hipLaunchKernelGGL(gather2_kernel, dim3((n + 63) / 64), dim3(64), 0, G->stream,
                     (const double* const*)G->scal_ptrs, n, G->gathered);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

void drop_lgraph(dn_ogroup* G) {
  if (G->exec) (void)hipGraphExecDestroy(G->exec);
  if (G->graph) (void)hipGraphDestroy(G->graph);
  G->exec = nullptr;
  G->graph = nullptr;
}

int capture_lgraph(dn_ogroup* G) {
  DN_HIP(hipStreamSynchronize(G->stream));
  drop_lgraph(G);
  DN_HIP(hipStreamBeginCapture(G->stream, hipStreamCaptureModeThreadLocal));
  const int rc = lagrangian_all(G);
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

}  // namespace
}  // namespace dn

extern "C" {

int dn_orbopt_start(dn_sys* s, const dn_functional* f, const dn_orbopt_state* st) {
  DN_TRY(check_sys(s));
  return orbopt_begin(s, f, st);
}

int dn_orbopt_done(dn_sys* s, int* done) {
  DN_TRY(check_sys(s));
  if (!done) return fail(DN_ERR_ARG, "null output");
  *done = (s->opt.started && s->opt.done) ? 1 : 0;
  return DN_OK;
}

int dn_orbopt_finish(dn_sys* s, dn_orbopt_state* st) {
  DN_TRY(check_sys(s));
  if (!st || !s->opt.started) return fail(DN_ERR_STATE, "no optimisation started");
  auto& o = s->opt;
  const long nn = (long)s->d.nbf * s->d.nbf;
  st->alpha = o.alpha;
  st->maxloop = o.maxloop;
  if (!o.improved) {  // minimization.jl:426-431
    st->alpha = 0.3 * st->alpha;
    st->maxloop = st->maxloop + 10;
  }
  if (st->maxloop < 100) {
    DN_TRY(copy_dev(s->stream, s->C, s->bestC, nn));
    st->E = o.best_E;
  } else {
    st->E = o.E;
  }
  st->nit = o.nit;
  st->success = o.success ? 1 : 0;
  DN_HIP(hipStreamSynchronize(s->stream));
  o.started = false;
  return DN_OK;
}

int dn_ogroup_create(dn_ctx* ctx, dn_sys* const* systems, int nsys, dn_ogroup** out) {
  if (!ctx || !systems || nsys <= 0 || !out) return fail(DN_ERR_ARG, "bad dn_ogroup_create arguments");
  *out = nullptr;
  DN_HIP(hipSetDevice(ctx->device));
  auto* G = new dn_ogroup();
  G->ctx = ctx;
  auto cleanup = [&](int rc) {
    dn_ogroup_destroy(G);
    return rc;
  };
  hipError_t e = hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, hipGetErrorString(e)));
  for (int i = 0; i < nsys; ++i) {
    dn_sys* s = systems[i];
    if (!s || s->ctx != ctx) return cleanup(fail(DN_ERR_ARG, "systems must share the context"));
    if (!s->opt.started || s->opt.done)
      return cleanup(fail(DN_ERR_STATE, "every system needs a started, unfinished optimisation"));
    if (s->precision != DN_FP64) return cleanup(fail(DN_ERR_ARG, "lock-step groups are FP64"));
    // the group drives the system on its own stream until it is destroyed
    (void)hipStreamSynchronize(s->stream);
    G->sys.push_back(s);
    G->saved_streams.push_back(s->stream);
    s->stream = G->stream;
  }
  for (int st = 0; st < kNumL && e == hipSuccess; ++st)
    e = hipMalloc(&G->dev[st], nsys * sizeof(GroupedGemm));
  if (e == hipSuccess) e = hipMalloc(&G->scal_ptrs, nsys * sizeof(double*));
  if (e == hipSuccess) e = hipMalloc(&G->gathered, 2 * nsys * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(&G->pinned, 2 * nsys * sizeof(double), hipHostMallocDefault);
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, std::string("dn_ogroup_create: ") + hipGetErrorString(e)));
  std::vector<double*> sp(nsys);
  for (int i = 0; i < nsys; ++i) sp[i] = G->sys[i]->scal;
  e = hipMemcpy(G->scal_ptrs, sp.data(), nsys * sizeof(double*), hipMemcpyHostToDevice);
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, hipGetErrorString(e)));
  for (int st = 0; st < kNumL; ++st) {
    const int rc = plan_lstage(G, st);
    if (rc != DN_OK) return cleanup(rc);
  }
  const int rc = capture_lgraph(G);
  if (rc != DN_OK) return cleanup(rc);
  *out = G;
  return DN_OK;
}

int dn_ogroup_destroy(dn_ogroup* G) {
  if (!G) return DN_OK;
  (void)hipSetDevice(G->ctx->device);
  if (G->stream) (void)hipStreamSynchronize(G->stream);
  for (size_t i = 0; i < G->sys.size(); ++i) G->sys[i]->stream = G->saved_streams[i];
  drop_lgraph(G);
  for (int st = 0; st < kNumL; ++st)
    if (G->dev[st]) (void)hipFree(G->dev[st]);
  if (G->scal_ptrs) (void)hipFree(G->scal_ptrs);
  if (G->gathered) (void)hipFree(G->gathered);
  if (G->pinned) (void)hipHostFree(G->pinned);
  if (G->stream) (void)hipStreamDestroy(G->stream);
  delete G;
  return DN_OK;
}

int dn_ogroup_step(dn_ogroup* G, int* active) {
  if (!G) return fail(DN_ERR_ARG, "null dn_ogroup");
  DN_HIP(hipSetDevice(G->ctx->device));
  const int n = (int)G->sys.size();
  std::vector<char> go(n, 0);
  // the top of the loop body of minimization.jl:344-438 for every live system
  for (int k = 0; k < n; ++k) {
    dn_sys* s = G->sys[k];
    auto& o = s->opt;
    if (o.done) continue;
    if (o.i >= o.maxloop) {
      o.done = true;
      continue;
    }
    ++o.i;
    ++o.nit;
    if (2 * o.gmax < o.threshgorb && o.improved) {  // minimization.jl:400-403
      o.success = true;
      o.done = true;
      continue;
    }
    if (o.method == DN_DEMON) {  // minimization.jl:307-308
      const double rate = (o.i - 1.0) / (o.maxloop - 1.0);
      o.beta1 = o.binit * (1.0 - rate) / ((1.0 - o.binit) + o.binit * (1.0 - rate));
    }
    const double bc1 = 1.0 - std::pow(o.beta1, o.i), bc2 = 1.0 - std::pow(o.beta2, o.i);
    const dn_dims& d = s->d;
    DN_TRY(adam_step(G->stream, s->gv, s->mv, s->vv, s->vmax, s->yv, o.method, o.beta1, o.beta2,
                     bc1, bc2, o.alpha, d.nbf, d.nbf5, s->Y));
    DN_TRY(norm1(G->stream, s->Y, d.nbf, s->scal + 4));
    DN_HIP(hipMemcpyAsync(G->pinned + 2 * k, s->scal + 4, sizeof(double), hipMemcpyDeviceToHost,
                          G->stream));
    go[k] = 1;
  }
  int live = 0;
  for (int k = 0; k < n; ++k) live += go[k];
  if (active) *active = live;
  if (!live) return DN_OK;
  DN_HIP(hipStreamSynchronize(G->stream));
  // C <- C exp(Y)  (utils.jl:499-521); C keeps its buffer, which the graph reads
  for (int k = 0; k < n; ++k) {
    if (!go[k]) continue;
    dn_sys* s = G->sys[k];
    const dn_dims& d = s->d;
    const double nrm = G->pinned[2 * k];
    if (!std::isfinite(nrm)) return fail(DN_ERR_NONFINITE, "non-finite rotation generator");
    DN_TRY(expm_taylor(G->stream, s->Y, nrm, d.nbf, s->rwork, s->U, s->part2, s->part2_elems));
    DN_TRY(matmul_sq(G->stream, s->C, s->U, s->Ca, d.nbf, s->part2, s->part2_elems));
    DN_TRY(copy_dev(G->stream, s->C, s->Ca, (long)d.nbf * d.nbf));
  }
  // every system's Lagrangian, energy and gradient in one graph (systems that stopped
  // this step only repeat the evaluation at their unchanged C)
  DN_HIP(hipGraphLaunch(G->exec, G->stream));
  DN_HIP(hipMemcpyAsync(G->pinned, G->gathered, 2 * n * sizeof(double), hipMemcpyDeviceToHost,
                        G->stream));
  DN_HIP(hipStreamSynchronize(G->stream));
  for (int k = 0; k < n; ++k) {
    if (!go[k]) continue;
    dn_sys* s = G->sys[k];
    auto& o = s->opt;
    o.E = G->pinned[2 * k];
    std::memcpy(&o.gmax, &G->pinned[2 * k + 1], sizeof(double));
    if (!std::isfinite(o.E) || !std::isfinite(o.gmax))
      return fail(DN_ERR_NONFINITE, "non-finite energy or orbital gradient");
    if (o.E < o.best_E) {
      o.best_E = o.E;
      DN_TRY(copy_dev(G->stream, s->bestC, s->C, (long)s->d.nbf * s->d.nbf));
      o.improved = true;
    }
  }
  return DN_OK;
}

}  // extern "C"
