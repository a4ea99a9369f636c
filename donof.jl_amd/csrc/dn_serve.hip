// Lock-step ADAM-family optimisation of several resident systems: every step advances all
// live solves together, the orbital Lagrangian of all of them (the dominant work,
// utils.jl:1-59) running as grouped GEMMs (one workgroup per tile) captured once in a HIP
// graph.
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
#include "dn_rot_dev.h"
#include "dn_sched.h"

namespace dn {

// dn_rot.hip / dn_lag.hip / dn_solve.hip
int orbopt_begin(dn_sys* s, const dn_functional* f, const dn_orbopt_state* st);
int copy_dev(hipStream_t s, double* dst, const double* src, long n);
size_t lag_split_count(const dn_dims& d);

int squarings_for(double nrm);

namespace {

enum LStage { kT1 = 0, kT2 = 1, kHC = 2, kHM = 3, kLG = 4, kLZ = 5, kNumL = 6 };

// ---- grouped rotation: the per-system work of a step for all live systems at once ------
//
// The arithmetic of every element is that of the single-system kernels (dn_rot_dev.h), and
// each square GEMM of exp(Y) and of C exp(Y) is split along K exactly as gemm_f64_splitk
// splits it (splitk_count / splitk_partials; the FP64 tile's per-element K order does not
// depend on the tile shape), reduced in the same split order: a system's rotation is
// bitwise the one of the single-system loop.

// One system's operands; the ADAM coefficients and f = 2^-s are set every step.
struct RotSys {
  const double* g;
  double *m, *v, *vmax, *y, *Y;
  double *X, *X2, *X3, *X4, *T, *U;  // exp workspace (rwork) and exp(Y)
  double *C, *Ca;
  double* nrm;                       // ||Y||_1 (bits of a non-negative double)
  int nbf, nbf5, method;
  int copy_u;                        // exp(Y) ended in T (odd squaring count)
  double beta1, beta2, bc1, bc2, alpha, f;
  long pack_begin, col_begin, elem_begin;  // block / column offsets
};

// out = sum_z P[z sP + e], z in order (reduce_splits with beta 0 and scale 1)
struct RedDesc {
  const double* P;
  long sP, elems, begin;
  double* out;
  int nsplit;
};

template <typename D>
__device__ __forceinline__ int find_block(const D* __restrict__ d, int n, long b,
                                          long D::*begin) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].*begin <= b) lo = mid;
    else hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

__global__ __launch_bounds__(256) void rot_adam_kernel(const RotSys* __restrict__ d, int n) {
  const RotSys r = d[find_block(d, n, blockIdx.x, &RotSys::pack_begin)];
  const long e = (blockIdx.x - r.pack_begin) * 256 + threadIdx.x;
  if (e < (long)r.nbf5 * r.nbf)
    adam_update(e, r.g, r.m, r.v, r.vmax, r.y, r.method, r.beta1, r.beta2, r.bc1, r.bc2,
                r.alpha, r.nbf, r.Y);
}

__global__ void rot_zero_norm_kernel(const RotSys* __restrict__ d, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) *reinterpret_cast<unsigned long long*>(d[i].nrm) = 0ull;
}

// one wavefront per column of every system's Y (four columns per workgroup)
__global__ __launch_bounds__(256) void rot_norm_kernel(const RotSys* __restrict__ d, int n,
                                                       long ncols) {
  const long c = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncols) return;
  const int lane = threadIdx.x & 63;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].col_begin <= c) lo = mid;
    else hi = mid - 1;
  }
  const RotSys& r = d[lo];
  const int j = (int)(c - r.col_begin);
  const double sum = column_abs_sum(r.Y + (long)j * r.nbf, r.nbf, lane);
  if (lane == 0)
    atomicMax(reinterpret_cast<unsigned long long*>(r.nrm),
              (unsigned long long)__double_as_longlong(sum));
}

__global__ void rot_gather_kernel(const RotSys* __restrict__ d, int n, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = *d[i].nrm;
}

// element-wise steps of exp(Y) and of the orbital update over every system's nbf^2 block
enum RotElem { kScale, kPolyB2, kPolyB1, kPolyB0, kAddTU, kAddUT, kCopyUT, kPolish, kCopyCCa };

__global__ __launch_bounds__(256) void rot_elem_kernel(const RotSys* __restrict__ d, int n,
                                                       int op, double c0, double c1, double c2,
                                                       double c3, double c4) {
  const RotSys r = d[find_block(d, n, blockIdx.x, &RotSys::elem_begin)];
  const long e = (blockIdx.x - r.elem_begin) * 256 + threadIdx.x;
  if (e >= (long)r.nbf * r.nbf) return;
  switch (op) {
    case kScale: r.X[e] = r.f * r.Y[e]; break;
    case kPolyB2: r.T[e] = poly_element(e, r.nbf, r.X, r.X2, r.X3, r.X4, c0, c1, c2, c3, c4); break;
    case kPolyB1:
    case kPolyB0:
      r.T[e] = poly_element(e, r.nbf, r.X, r.X2, r.X3, nullptr, c0, c1, c2, c3, 0.0);
      break;
    case kAddTU: r.T[e] += r.U[e]; break;
    case kAddUT: r.U[e] += r.T[e]; break;
    case kCopyUT:
      if (r.copy_u) r.U[e] = r.T[e];
      break;
    case kPolish: r.U[e] = ns_polish(r.U[e], r.X2[e]); break;
    default: r.C[e] = r.Ca[e]; break;
  }
}

__global__ __launch_bounds__(256) void rot_reduce_kernel(const RedDesc* __restrict__ d, int n) {
  const RedDesc r = d[find_block(d, n, blockIdx.x, &RedDesc::begin)];
  const long e = (blockIdx.x - r.begin) * 256 + threadIdx.x;
  if (e >= r.elems) return;
  double s = 0.0;
  for (int z = 0; z < r.nsplit; ++z) s += r.P[z * r.sP + e];
  s *= 1.0;
  r.out[e] = s;
}

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
    case kT1:  // bpn = b_Q C5 (integrals.jl:159 / utils.jl:19)
      g = t1_gemm(s, false);
      *config = gemm_config_ok(kLds64x128, g) ? lds64_config(g.N)
                                              : reg_config(g.M, g.N, g.nbatch / g.bps);
      break;
    case kT2:  // b_pq = C5^T bpn, (q, Q) flattened (the first nbf5 columns of b_MO)
      g = t2_flat_gemm(s);
      *config = !gemm_config_ok(kLds64x128, g) ? reg_config(g.M, g.N, 1)
                : (g.M <= 64 || (g.M % 128 != 0 && g.M % 128 <= 64)) ? lds64_config(g.N)
                                                                    : kLds128x256;
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
    case kLG:  // Z partials: sum_Q sum_q bpn[n,q,Q] bmat[q,b,Q] over Q ranges
      g = lag_contract_gemm(s);
      *config = kTile64x64;
      break;
    default:  // elag += C^T Z
      g = lag_final_gemm(s);
      *config = kTile64x64;
      break;
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
  std::vector<dn_sys*> claimed;  // the systems whose membership this group holds
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
  size_t pinned_bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // grouped rotation: descriptors built on the host each step (pinned staging) and
  // copied to the device in one transfer per phase
  unsigned char* rot_h = nullptr;
  unsigned char* rot_d = nullptr;
  size_t rot_bytes = 0, rot_h_bytes = 0;
  // grouped Lagrangian tail (lag_w, lag_bmat, lag_init, the partial sum, energy, gradient)
  dn::LagSys* lag_d = nullptr;
  dn::LagBlocks lag_b;
  // systems with odd nbf: C copied into its padded form before the transforms
  dn::PadDesc* pad_d = nullptr;
  int npad = 0;
  long pad_blocks = 0;
};

namespace dn {
namespace {

// ---- grouped rotation, host side -------------------------------------------------------

// square-GEMM stages of a step: exp(Y)'s Taylor products, the two squaring directions and
// C exp(Y)
enum RotGemm { kXX, kX2X, kX2X2, kX4T, kSqUT, kSqTU, kUtU, kUX, kCU, kNumRotGemm };

struct RotLayout {  // offsets into the staging area for n systems
  size_t sys, gemm, red, bytes;
  static size_t align(size_t x) { return (x + 255) / 256 * 256; }
  explicit RotLayout(int n) {
    sys = 0;
    gemm = align(sys + n * sizeof(RotSys));
    red = align(gemm + (size_t)kNumRotGemm * n * sizeof(GroupedGemm));
    bytes = align(red + (size_t)kNumRotGemm * n * sizeof(RedDesc));
  }
};

struct RotClass {
  int first, count, config;
};

// One stage's problems (sorted into classes of tile shape and alignment; within a class by
// squaring count, descending, so that the systems of squaring level k are a prefix) and
// the reductions of the split ones (same order).
struct RotStage {
  GroupedGemm* h = nullptr;
  GroupedGemm* d = nullptr;
  RedDesc* rh = nullptr;
  RedDesc* rd = nullptr;
  int n = 0, nred = 0;
  std::vector<RotClass> classes;
  std::vector<int> sq;       // squaring count of each problem (sorted order)
  std::vector<int> red_sq;   // of each reduction
};

// The square GEMM out = A B of sq_gemm (dn_rot.hip) for nbf x nbf operands (out = A^T B,
// tsq_gemm's, with transpose_a).
Gemm rot_square_gemm(const double* A, const double* B, double* out, int nbf,
                     bool transpose_a = false) {
  Gemm g;
  g.M = nbf; g.N = nbf; g.K = nbf;
  g.A = A; g.lda = nbf; g.a_kcontig = transpose_a;
  g.B = B; g.ldb = nbf; g.b_kcontig = true;
  g.C = out; g.ldc = nbf;
  return g;
}

void rot_operands(const RotSys& r, int kind, const double** A, const double** B, double** out) {
  switch (kind) {
    case kXX: *A = r.X; *B = r.X; *out = r.X2; break;
    case kX2X: *A = r.X2; *B = r.X; *out = r.X3; break;
    case kX2X2: *A = r.X2; *B = r.X2; *out = r.X4; break;
    case kX4T: *A = r.X4; *B = r.T; *out = r.U; break;
    case kSqUT: *A = r.U; *B = r.U; *out = r.T; break;
    case kSqTU: *A = r.T; *B = r.T; *out = r.U; break;
    case kUtU: *A = r.U; *B = r.U; *out = r.X; break;    // U^T U (A transposed)
    case kUX: *A = r.U; *B = r.X; *out = r.X2; break;
    default: *A = r.C; *B = r.U; *out = r.Ca; break;
  }
}

// Builds stage `kind` for the systems rs[0..n) into the staging area (host and device
// addresses of the same layout); sq[i] is system i's squaring count.
void rot_build_stage(dn_ogroup* G, const RotLayout& L, int kind, const RotSys* rs,
                     const std::vector<dn_sys*>& sys, const std::vector<int>& sq, int n,
                     RotStage* st) {
  st->h = reinterpret_cast<GroupedGemm*>(G->rot_h + L.gemm) + (size_t)kind * n;
  st->d = reinterpret_cast<GroupedGemm*>(G->rot_d + L.gemm) + (size_t)kind * n;
  st->rh = reinterpret_cast<RedDesc*>(G->rot_h + L.red) + (size_t)kind * n;
  st->rd = reinterpret_cast<RedDesc*>(G->rot_d + L.red) + (size_t)kind * n;
  struct Item {
    GroupedGemm gg;
    RedDesc rd;
    int nsplit, key, sq;
  };
  std::vector<Item> items(n);
  for (int i = 0; i < n; ++i) {
    const double *A, *B;
    double* out;
    rot_operands(rs[i], kind, &A, &B, &out);
    const Gemm g = rot_square_gemm(A, B, out, rs[i].nbf, kind == kUtU);
    Item& it = items[i];
    it.nsplit = splitk_count(g, sys[i]->part2_elems);
    it.gg.g = it.nsplit > 1 ? splitk_partials(g, it.nsplit, sys[i]->part2) : g;
    it.rd = RedDesc{sys[i]->part2, (long)rs[i].nbf * rs[i].nbf, (long)rs[i].nbf * rs[i].nbf, 0,
                    out, it.nsplit};
    gemm_grouped_plan(&it.gg, 1, kTile64x64);  // alignment flags
    const int cfg = (g.M > 64 && g.N > 64) ? kTile128x128 : kTile64x64;
    it.key = cfg * 4 + 2 * !it.gg.g.vec_a + !it.gg.g.vec_b;
    it.sq = sq[i];
  }
  std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
    return a.key != b.key ? a.key < b.key : a.sq > b.sq;
  });
  st->n = n;
  st->classes.clear();
  st->sq.resize(n);
  for (int i = 0; i < n; ++i) {
    st->h[i] = items[i].gg;
    st->sq[i] = items[i].sq;
  }
  for (int i = 0; i < n;) {
    int j = i;
    while (j < n && items[j].key == items[i].key) ++j;
    const int cfg = items[i].key / 4;
    gemm_grouped_plan(st->h + i, j - i, cfg);
    st->classes.push_back({i, j - i, cfg});
    i = j;
  }
  // reductions, by squaring count descending
  std::vector<int> ridx;
  for (int i = 0; i < n; ++i)
    if (items[i].nsplit > 1) ridx.push_back(i);
  std::stable_sort(ridx.begin(), ridx.end(),
                   [&](int a, int b) { return items[a].sq > items[b].sq; });
  st->nred = (int)ridx.size();
  st->red_sq.resize(ridx.size());
  long blk = 0;
  for (size_t k = 0; k < ridx.size(); ++k) {
    RedDesc r = items[ridx[k]].rd;
    r.begin = blk;
    blk += (r.elems + 255) / 256;
    st->rh[k] = r;
    st->red_sq[k] = items[ridx[k]].sq;
  }
}

// Grid of a lock-step group's grouped GEMM: one workgroup per tile slot.  Lock-step
// engines run beside each other and beside the set-up of new solves; a persistent grid
// would hold its CUs until its last tile and hold up every other engine's short kernels
// (measured on MI355X, 8 solves of nbf 50-120: 0.91 ms per step with one workgroup per
// tile against 1.38-1.39 ms with 1 or 2 x the resident count; beside a 4-solve engine of
// nbf 130-155, 1.66 + 1.59 ms against 1.44-1.47 + 1.99-2.12 ms).
int tile_grid(const GroupedGemm* h, int n) {
  return (int)std::min(8 * grouped_slots(h, n), 0x7ffffff8L);
}

// Runs a stage for the problems with squaring count > level (all of them for level -1).
int rot_run_stage(dn_ogroup* G, RotStage* st, int level) {
  for (const RotClass& c : st->classes) {
    int p = 0;
    while (p < c.count && st->sq[c.first + p] > level) ++p;
    if (!p) continue;
    // the class was planned whole; its first p problems keep their offsets (the staging
    // area is not written again while its copy to the device may be pending)
    const GroupedGemm& last = st->h[c.first + p - 1];
    const long total = last.tile_begin + (long)last.tiles_m * last.tiles_n * last.nsplit;
    DN_TRY(gemm_f64_grouped(G->stream, st->h + c.first, st->d + c.first, p, total, c.config,
                            tile_grid(st->h + c.first, p)));
  }
  int r = 0;
  while (r < st->nred && st->red_sq[r] > level) ++r;
  if (r) {
    const long blocks = st->rh[r - 1].begin + (st->rh[r - 1].elems + 255) / 256;
    hipLaunchKernelGGL(rot_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, G->stream,
                       (const RedDesc*)st->rd, r);
    DN_HIP(hipGetLastError());
  }
  return DN_OK;
}

int rot_elem(dn_ogroup* G, const RotSys* d, int n, long blocks, int op, double c0 = 0.0,
             double c1 = 0.0, double c2 = 0.0, double c3 = 0.0, double c4 = 0.0) {
  hipLaunchKernelGGL(rot_elem_kernel, dim3((unsigned)blocks), dim3(256), 0, G->stream, d, n, op,
                     c0, c1, c2, c3, c4);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

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
                            l.count, l.tiles, l.config,
                            tile_grid(G->host[stage].data() + l.first, l.count)));
  return DN_OK;
}

// The Lagrangian, energy and packed gradient of every system (the graph body).
int lagrangian_all(dn_ogroup* G) {
  DN_TRY(pad_c_group(G->stream, G->pad_d, G->npad, G->pad_blocks));
  DN_TRY(launch_lstage(G, kT1));
  DN_TRY(launch_lstage(G, kT2));
  DN_TRY(launch_lstage(G, kHC));
  DN_TRY(launch_lstage(G, kHM));
  const int n = (int)G->sys.size();
  DN_TRY(lag_group_before(G->stream, G->lag_d, n, G->lag_b));
  DN_TRY(launch_lstage(G, kLG));
  DN_TRY(lag_group_reduce(G->stream, G->lag_d, n, G->lag_b));
  DN_TRY(launch_lstage(G, kLZ));
  DN_TRY(lag_group_after(G->stream, G->lag_d, n, G->lag_b));
  hipLaunchKernelGGL(gather2_kernel, dim3((n + 63) / 64), dim3(64), 0, G->stream,
                     (const double* const*)G->scal_ptrs, n, G->gathered);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// the Lagrangian-tail descriptors of every system (fixed for the group's lifetime)
int plan_lag_tail(dn_ogroup* G) {
  const int n = (int)G->sys.size();
  std::vector<LagSys> h(n);
  for (int i = 0; i < n; ++i) {
    dn_sys* s = G->sys[i];
    const dn_dims& d = s->d;
    const int nsplit = (int)lag_split_count(d);
    const int bps = (d.nbfaux + nsplit - 1) / nsplit;
    LagSys& r = h[i];
    r.bt = s->bt;
    r.ldt = d.nbf5;
    r.slice = (long)d.nbf5 * d.nbf5;
    r.nbf = d.nbf;
    r.nbf5 = d.nbf5;
    r.naux = d.nbfaux;
    r.nalpha = s->opt.nalpha;
    r.nbeta = s->opt.nbeta;
    r.cj12 = s->cj12;
    r.ck12 = s->ck12;
    r.n = s->n;
    r.X = s->X;
    r.w = s->w;
    r.bmat = s->bmat;
    r.Hmat = s->Hmat;
    r.elag = s->elag;
    r.part = s->part;
    r.nsplit = (d.nbfaux + bps - 1) / bps;
    r.scal = s->scal;
    r.g = s->gv;
    r.Z = s->Z;
  }
  lag_group_plan(h.data(), n, &G->lag_b);
  DN_HIP(hipMemcpy(G->lag_d, h.data(), n * sizeof(LagSys), hipMemcpyHostToDevice));
  std::vector<PadDesc> pd;
  for (dn_sys* s : G->sys)
    if (s->Cp != s->C) pd.push_back(PadDesc{s->C, s->Cp, s->d.nbf, s->ldb, 0});
  G->npad = (int)pd.size();
  if (G->npad) {
    G->pad_blocks = pad_plan(pd.data(), G->npad);
    DN_HIP(hipMemcpy(G->pad_d, pd.data(), pd.size() * sizeof(PadDesc), hipMemcpyHostToDevice));
  }
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
  for (int i = 0; i < nsys; ++i) {
    dn_sys* s = systems[i];
    if (!s || s->ctx != ctx) return cleanup(fail(DN_ERR_ARG, "systems must share the context"));
    if (!s->opt.started || s->opt.done)
      return cleanup(fail(DN_ERR_STATE, "every system needs a started, unfinished optimisation"));
    if (s->precision != DN_FP64) return cleanup(fail(DN_ERR_ARG, "lock-step groups are FP64"));
  }
  // a system runs on one group's stream at a time: none may belong to another group, and
  // none may be listed twice
  {
    const int rc = claim_lockstep(systems, nsys, G);
    if (rc != DN_OK) return cleanup(rc);
    G->claimed.assign(systems, systems + nsys);
  }
  hipError_t e = hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, hipGetErrorString(e)));
  for (int i = 0; i < nsys; ++i) {
    dn_sys* s = systems[i];
    // the group drives the system on its own stream until it is destroyed
    (void)hipStreamSynchronize(s->stream);
    G->sys.push_back(s);
    G->saved_streams.push_back(s->stream);
    s->stream = G->stream;
  }
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, hipGetErrorString(e)));
  // descriptors and staging from the context's pools: a group is rebuilt whenever its
  // membership changes, and neither may wait for other groups' work on the device
  {
    int rc = DN_OK;
    auto D = [&](size_t bytes, void** p) {
      if (rc == DN_OK) rc = dev_alloc(ctx, G->stream, bytes, p);
    };
    for (int st = 0; st < kNumL; ++st)
      D(nsys * sizeof(GroupedGemm), reinterpret_cast<void**>(&G->dev[st]));
    D(nsys * sizeof(double*), reinterpret_cast<void**>(&G->scal_ptrs));
    D(2 * nsys * sizeof(double), reinterpret_cast<void**>(&G->gathered));
    G->rot_bytes = RotLayout(nsys).bytes;
    D(G->rot_bytes, reinterpret_cast<void**>(&G->rot_d));
    D(nsys * sizeof(LagSys), reinterpret_cast<void**>(&G->lag_d));
    D(nsys * sizeof(PadDesc), reinterpret_cast<void**>(&G->pad_d));
    if (rc == DN_OK)
      rc = pinned_get(ctx, 2 * nsys * sizeof(double), reinterpret_cast<void**>(&G->pinned),
                      &G->pinned_bytes);
    if (rc == DN_OK)
      rc = pinned_get(ctx, G->rot_bytes, reinterpret_cast<void**>(&G->rot_h), &G->rot_h_bytes);
    if (rc != DN_OK) return cleanup(rc);
    e = hipStreamSynchronize(G->stream);  // the blocks are usable from any stream now
    if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, hipGetErrorString(e)));
  }
  std::vector<double*> sp(nsys);
  for (int i = 0; i < nsys; ++i) sp[i] = G->sys[i]->scal;
  e = hipMemcpy(G->scal_ptrs, sp.data(), nsys * sizeof(double*), hipMemcpyHostToDevice);
  if (e != hipSuccess) return cleanup(fail(DN_ERR_HIP, hipGetErrorString(e)));
  for (int st = 0; st < kNumL; ++st) {
    const int rc = plan_lstage(G, st);
    if (rc != DN_OK) return cleanup(rc);
  }
  const int lrc = plan_lag_tail(G);
  if (lrc != DN_OK) return cleanup(lrc);
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
  release_lockstep(G->claimed.data(), (int)G->claimed.size(), G);
  drop_lgraph(G);
  if (G->stream) {
    for (int st = 0; st < kNumL; ++st) dev_free(G->ctx, G->stream, G->dev[st]);
    dev_free(G->ctx, G->stream, G->scal_ptrs);
    dev_free(G->ctx, G->stream, G->gathered);
    dev_free(G->ctx, G->stream, G->rot_d);
    dev_free(G->ctx, G->stream, G->lag_d);
    dev_free(G->ctx, G->stream, G->pad_d);
    (void)hipStreamSynchronize(G->stream);
  }
  pinned_put(G->ctx, G->pinned, G->pinned_bytes);
  pinned_put(G->ctx, G->rot_h, G->rot_h_bytes);
  if (G->stream) (void)hipStreamDestroy(G->stream);
  delete G;
  return DN_OK;
}

int dn_ogroup_step(dn_ogroup* G, int* active) {
  if (!G) return fail(DN_ERR_ARG, "null dn_ogroup");
  DN_HIP(hipSetDevice(G->ctx->device));
  const int n = (int)G->sys.size();
  std::vector<char> go(n, 0);
  std::vector<RotSys> rs;        // the rotating systems, in group order
  std::vector<dn_sys*> gsys;
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
    RotSys r{};
    r.g = s->gv;
    r.m = s->mv;
    r.v = s->vv;
    r.vmax = s->vmax;
    r.y = s->yv;
    r.Y = s->Y;
    const long nn = (long)s->d.nbf * s->d.nbf;
    r.X = s->rwork;
    r.X2 = s->rwork + nn;
    r.X3 = s->rwork + 2 * nn;
    r.X4 = s->rwork + 3 * nn;
    r.T = s->rwork + 4 * nn;
    r.U = s->U;
    r.C = s->C;
    r.Ca = s->Ca;
    r.nrm = s->scal + 4;
    r.nbf = s->d.nbf;
    r.nbf5 = s->d.nbf5;
    r.method = o.method;
    r.beta1 = o.beta1;
    r.beta2 = o.beta2;
    r.bc1 = bc1;
    r.bc2 = bc2;
    r.alpha = o.alpha;
    rs.push_back(r);
    gsys.push_back(s);
    go[k] = 1;
  }
  const int live = (int)rs.size();
  if (active) *active = live;
  if (!live) return DN_OK;
  // the ADAM-family step and ||Y||_1 of every rotating system (one synchronisation)
  const RotLayout L(live);
  long pack = 0, cols = 0, elems = 0;
  for (RotSys& r : rs) {
    r.pack_begin = pack;
    r.col_begin = cols;
    r.elem_begin = elems;
    pack += ((long)r.nbf5 * r.nbf + 255) / 256;
    cols += r.nbf;
    elems += ((long)r.nbf * r.nbf + 255) / 256;
  }
  RotSys* rs_h = reinterpret_cast<RotSys*>(G->rot_h + L.sys);
  const RotSys* rs_d = reinterpret_cast<const RotSys*>(G->rot_d + L.sys);
  std::memcpy(rs_h, rs.data(), live * sizeof(RotSys));
  DN_HIP(hipMemcpyAsync(G->rot_d + L.sys, rs_h, live * sizeof(RotSys), hipMemcpyHostToDevice,
                        G->stream));
  hipLaunchKernelGGL(rot_zero_norm_kernel, dim3((live + 63) / 64), dim3(64), 0, G->stream, rs_d,
                     live);
  hipLaunchKernelGGL(rot_adam_kernel, dim3((unsigned)pack), dim3(256), 0, G->stream, rs_d, live);
  hipLaunchKernelGGL(rot_norm_kernel, dim3((unsigned)((cols + 3) / 4)), dim3(256), 0, G->stream,
                     rs_d, live, cols);
  hipLaunchKernelGGL(rot_gather_kernel, dim3((live + 63) / 64), dim3(64), 0, G->stream, rs_d,
                     live, G->gathered);
  DN_HIP(hipGetLastError());
  DN_HIP(hipMemcpyAsync(G->pinned, G->gathered, live * sizeof(double), hipMemcpyDeviceToHost,
                        G->stream));
  DN_HIP(hipStreamSynchronize(G->stream));
  // C <- C exp(Y) for all of them (utils.jl:499-521); C keeps its buffer, which the graph
  // reads.  exp by scaling and squaring: X = Y 2^-s, degree-12 Taylor polynomial by
  // Paterson-Stockmeyer, s squarings (the steps of expm_taylor).
  std::vector<int> sq(live);
  int smax = 0;
  for (int t = 0; t < live; ++t) {
    const double nrm = G->pinned[t];
    if (!std::isfinite(nrm)) return fail(DN_ERR_NONFINITE, "non-finite rotation generator");
    sq[t] = squarings_for(nrm);
    smax = std::max(smax, sq[t]);
    rs_h[t].f = std::ldexp(1.0, -sq[t]);
    rs_h[t].copy_u = sq[t] & 1;
  }
  RotStage st[kNumRotGemm];
  for (int k = 0; k < kNumRotGemm; ++k) rot_build_stage(G, L, k, rs_h, gsys, sq, live, &st[k]);
  DN_HIP(hipMemcpyAsync(G->rot_d, G->rot_h, L.bytes, hipMemcpyHostToDevice, G->stream));
  double c[13];
  c[0] = 1.0;
  for (int k = 1; k <= 12; ++k) c[k] = c[k - 1] / k;
  DN_TRY(rot_elem(G, rs_d, live, elems, kScale));
  DN_TRY(rot_run_stage(G, &st[kXX], -1));
  DN_TRY(rot_run_stage(G, &st[kX2X], -1));
  DN_TRY(rot_run_stage(G, &st[kX2X2], -1));
  DN_TRY(rot_elem(G, rs_d, live, elems, kPolyB2, c[8], c[9], c[10], c[11], c[12]));
  DN_TRY(rot_run_stage(G, &st[kX4T], -1));
  DN_TRY(rot_elem(G, rs_d, live, elems, kPolyB1, c[4], c[5], c[6], c[7]));
  DN_TRY(rot_elem(G, rs_d, live, elems, kAddTU));
  DN_TRY(rot_run_stage(G, &st[kX4T], -1));
  DN_TRY(rot_elem(G, rs_d, live, elems, kPolyB0, c[0], c[1], c[2], c[3]));
  DN_TRY(rot_elem(G, rs_d, live, elems, kAddUT));
  for (int k = 0; k < smax; ++k)  // squarings: U <- U^2 ping-ponging with T
    DN_TRY(rot_run_stage(G, &st[(k & 1) ? kSqTU : kSqUT], k));
  bool odd = false;
  for (int t = 0; t < live; ++t) odd = odd || (sq[t] & 1);
  if (odd) DN_TRY(rot_elem(G, rs_d, live, elems, kCopyUT));  // results that ended in T
  // re-orthonormalisation (expm_taylor's): U <- U (3 - U^T U) / 2
  DN_TRY(rot_run_stage(G, &st[kUtU], -1));
  DN_TRY(rot_run_stage(G, &st[kUX], -1));
  DN_TRY(rot_elem(G, rs_d, live, elems, kPolish));
  DN_TRY(rot_run_stage(G, &st[kCU], -1));
  DN_TRY(rot_elem(G, rs_d, live, elems, kCopyCCa));
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

int dn_ogroup_run(dn_ogroup* G, int max_steps, const volatile int* stop, int* steps,
                  int* active) {
  if (!G || max_steps < 0) return fail(DN_ERR_ARG, "bad dn_ogroup_run arguments");
  auto finished = [&] {
    int k = 0;
    for (dn_sys* s : G->sys) k += s->opt.done ? 1 : 0;
    return k;
  };
  int k = 0, act = 0;
  while (k < max_steps && !(stop && *stop)) {
    const int before = finished();
    DN_TRY(dn_ogroup_step(G, &act));
    ++k;
    if (act == 0 || finished() != before) break;
  }
  if (steps) *steps = k;
  if (active) *active = act;
  return DN_OK;
}

}  // extern "C"
