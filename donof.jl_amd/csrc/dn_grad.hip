// RI nuclear gradients on the device: the densities and contractions of
// compute_geom_gradients' RI branch (grads.jl:1-100).
//
// The reference forms IG = I3c inv(G) (nbf^2 x naux^2 work), I2 = C5^T IG C5, the
// densities D (nbf5 x nbf5 per Q) and d = C5 D C5^T, and m[Q,R] = sum_ls d[l,s,Q]
// IG[s,l,R] (another nbf^2 naux^2).  Here:
//   IG = b Wt with the RI factor of b (b = I3c L^{-T} from dn_sys_ri_metric: Wt = L^{-1};
//        any other b = I3c W: Wt = W^{-1} G^{-1}, given by the caller), so that
//   I2_Q = sum_P (C5^T b_P C5) Wt[P,Q]: the J/K build's own transform of b, then one
//        GEMM of nbf5^2 x naux x naux;
//   m[Q,R] = sum_pq D[p,q,Q] I2[q,p,R] (C5 moved from d onto IG): nbf5^2 naux^2, a
//        factor (nbf/nbf5)^2 below the reference's contraction over (l, s);
//   d_Q = C5 D_Q C5^T once (two GEMMs, the second flattened over (s, Q)), kept for the
//        per-atom contraction with the derivative integrals geri3c.
// Per atom (grads.jl:13-21, 91-96) the electronic gradient is five flat dot products:
//   -<lag, dS_k> + <RDM1, dT_k> + <RDM1, dV_k> + 2 <d, geri3c_k> - <m, geri2c_k>,
// HBM-bound (geri3c is nbf^2 naux x 3); host arrays are streamed through a staging
// buffer, device arrays are read in place.  All sums have a fixed order.
#include <cmath>

#include "dn_internal.h"

struct dn_grad {
  dn_sys* s = nullptr;
  int nbf = 0, naux = 0, nbf5 = 0;
  uint64_t bytes = 0;       // reserved in the context's budget
  void* mem = nullptr;      // one device block holding every array below
  double* d = nullptr;      // nbf*nbf*naux : d[l,s,Q] at l + nbf (s + nbf Q)
  double* m = nullptr;      // naux*naux    : m[Q,R] at Q + naux R
  double* rdm1 = nullptr;   // nbf*nbf
  double* lag = nullptr;    // nbf*nbf
  double* stage = nullptr;  // 3 * stage_len : host derivative integrals, one chunk
  long stage_len = 0;
  double* part = nullptr;   // dot partials, part_len
  long part_len = 0;
  double* res = nullptr;    // 16 dot results
};

namespace dn {
namespace {

constexpr int kDotBlocks = 1024;  // 16 wavefronts per CU for the HBM-bound dots
constexpr int kDotThreads = 256;
constexpr long kStageLen = 1L << 24;  // doubles per component and chunk (128 MiB)

// X[p,Q] = I2[p,p,Q]; dg[q,Q] = sum_{p != q} cj12[p,q] X[p,Q] + n_q X[q,Q] for q < nbeta
// and nalpha <= q < nbf5 (grads.jl:72-84; the diagonal of cj12 zeroed as 24-25 do).
// I2 at q + nbf5 p + nbf5^2 Q.  One block per Q.
__global__ __launch_bounds__(256) void grad_diag_kernel(const double* __restrict__ I2,
                                                        const double* __restrict__ cj12,
                                                        const double* __restrict__ n, int nbf5,
                                                        int nalpha, int nbeta,
                                                        double* __restrict__ dg) {
  extern __shared__ double x[];
  const int Q = blockIdx.x;
  const long slab = (long)nbf5 * nbf5;
  const double* I = I2 + Q * slab;
  for (int p = threadIdx.x; p < nbf5; p += blockDim.x) x[p] = I[(long)p * (nbf5 + 1)];
  __syncthreads();
  for (int q = threadIdx.x; q < nbf5; q += blockDim.x) {
    double acc = 0.0;
    const double* cq = cj12 + (long)q * nbf5;
    for (int p = 0; p < nbf5; ++p)
      if (p != q) acc += cq[p] * x[p];
    if (q < nbeta || q >= nalpha) acc += n[q] * x[q];
    dg[q + (long)nbf5 * Q] = acc;
  }
}

// D[p,q,Q] at p + nbf5 q + nbf5^2 Q: -ck12[p,q] I2[p,q,Q] off the diagonal (grads.jl:68),
// dg[p,Q] on it.
__global__ __launch_bounds__(256) void grad_dens_kernel(const double* __restrict__ I2,
                                                        const double* __restrict__ ck12,
                                                        const double* __restrict__ dg, int nbf5,
                                                        int naux, double* __restrict__ D) {
  const long slab = (long)nbf5 * nbf5;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= slab * naux) return;
  const long Q = i / slab;
  const int j = (int)(i - Q * slab);
  const int p = j % nbf5, q = j / nbf5;
  D[i] = p == q ? dg[p + (long)nbf5 * Q] : -ck12[j] * I2[Q * slab + q + (long)nbf5 * p];
}

// B[j] = W[j] + W[j^T] within each nbf5 x nbf5 slab: b_pq = W + W^T of the half
// contraction
__global__ __launch_bounds__(256) void grad_sym_kernel(const double* __restrict__ W, int nbf5,
                                                       int naux, double* __restrict__ B) {
  const long slab = (long)nbf5 * nbf5;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= slab * naux) return;
  const long Q = i / slab;
  const int j = (int)(i - Q * slab);
  const int a = j % nbf5, c = j / nbf5;
  B[i] = W[i] + W[Q * slab + c + (long)nbf5 * a];
}

// CT[p + nbf5 l] = C[l,p], CnT[p + nbf5 l] = n_p C[l,p] for p < nbf5
__global__ void grad_ct_kernel(const double* __restrict__ C, const double* __restrict__ n,
                               int nbf, int nbf5, double* __restrict__ CT,
                               double* __restrict__ CnT) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)nbf * nbf5) return;
  const int p = (int)(i % nbf5);
  const long l = i / nbf5;
  const double c = C[l + (long)nbf * p];
  CT[i] = c;
  CnT[i] = n[p] * c;
}

__device__ __forceinline__ double block_sum256(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kDotThreads / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

// part[blockIdx.x + kDotBlocks k] = this block's share of sum_i x[i] y[i + k ystride],
// k < 3, over i < len.  Grid-stride with a fixed grid: the order is fixed by len.  VEC:
// 16-byte loads of element pairs (x, y 16-byte aligned, ystride and len even).
template <bool VEC>
__global__ __launch_bounds__(kDotThreads) void dot3_kernel(const double* __restrict__ x,
                                                           const double* __restrict__ y,
                                                           long ystride, long len,
                                                           double* __restrict__ part) {
  __shared__ double red[kDotThreads / 64];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  const long step = (long)gridDim.x * kDotThreads;
  if constexpr (VEC) {
    const double2* x2 = reinterpret_cast<const double2*>(x);
    const double2* y0 = reinterpret_cast<const double2*>(y);
    const double2* y1 = reinterpret_cast<const double2*>(y + ystride);
    const double2* y2 = reinterpret_cast<const double2*>(y + 2 * ystride);
    for (long i = blockIdx.x * (long)kDotThreads + threadIdx.x; i < len / 2; i += step) {
      const double2 v = x2[i], u0 = y0[i], u1 = y1[i], u2 = y2[i];
      a0 += v.x * u0.x + v.y * u0.y;
      a1 += v.x * u1.x + v.y * u1.y;
      a2 += v.x * u2.x + v.y * u2.y;
    }
  } else {
    for (long i = blockIdx.x * (long)kDotThreads + threadIdx.x; i < len; i += step) {
      const double v = x[i];
      a0 += v * y[i];
      a1 += v * y[i + ystride];
      a2 += v * y[i + 2 * ystride];
    }
  }
  a0 = block_sum256(a0, red);
  a1 = block_sum256(a1, red);
  a2 = block_sum256(a2, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a0;
    part[blockIdx.x + kDotBlocks] = a1;
    part[blockIdx.x + 2 * kDotBlocks] = a2;
  }
}

// res[k] = sum over chunks c < nchunk and blocks b of part[(3 c + k) kDotBlocks + b]
__global__ __launch_bounds__(kDotThreads) void dot3_reduce_kernel(const double* __restrict__ part,
                                                                  int nchunk,
                                                                  double* __restrict__ res) {
  __shared__ double red[kDotThreads / 64];
  for (int k = 0; k < 3; ++k) {
    double a = 0.0;
    for (int c = 0; c < nchunk; ++c)
      for (int b = threadIdx.x; b < kDotBlocks; b += kDotThreads)
        a += part[(3L * c + k) * kDotBlocks + b];
    a = block_sum256(a, red);
    if (threadIdx.x == 0) res[k] = a;
  }
}

// one chunk's partials: the vector kernel where the operands allow 16-byte loads
int launch_dot3(hipStream_t st, const double* x, const double* y, long ystride, long len,
                double* part) {
  const bool vec = (uintptr_t)x % 16 == 0 && (uintptr_t)y % 16 == 0 && ystride % 2 == 0 &&
                   len % 2 == 0;
  if (vec)
    hipLaunchKernelGGL(dot3_kernel<true>, dim3(kDotBlocks), dim3(kDotThreads), 0, st, x, y,
                       ystride, len, part);
  else
    hipLaunchKernelGGL(dot3_kernel<false>, dim3(kDotBlocks), dim3(kDotThreads), 0, st, x, y,
                       ystride, len, part);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

bool on_device(const void* p, int device) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable host memory
    return false;
  }
  return a.type == hipMemoryTypeDevice && a.device == device;
}

// res[0..2] = sum_i x[i] y[i + k len] (y: len x 3, host or device)
int dot3(dn_grad* g, const double* x, const double* y, long len, double* res) {
  hipStream_t st = g->s->stream;
  if (on_device(y, g->s->ctx->device)) {
    DN_TRY(launch_dot3(st, x, y, len, len, g->part));
    hipLaunchKernelGGL(dot3_reduce_kernel, dim3(1), dim3(kDotThreads), 0, st, g->part, 1, res);
    DN_HIP(hipGetLastError());
    return DN_OK;
  }
  const long nchunk = (len + g->stage_len - 1) / g->stage_len;
  if (nchunk * 3 * kDotBlocks > g->part_len) return fail(DN_ERR_ARG, "gradient partials overflow");
  for (long c = 0; c < nchunk; ++c) {
    const long off = c * g->stage_len, cl = std::min(g->stage_len, len - off);
    for (int k = 0; k < 3; ++k)
      DN_HIP(hipMemcpyAsync(g->stage + k * cl, y + k * len + off, cl * sizeof(double),
                            hipMemcpyHostToDevice, st));
    DN_TRY(launch_dot3(st, x + off, g->stage, cl, cl, g->part + c * 3 * kDotBlocks));
  }
  hipLaunchKernelGGL(dot3_reduce_kernel, dim3(1), dim3(kDotThreads), 0, st, g->part,
                     (int)nchunk, res);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

size_t r256(size_t doubles) { return (doubles + 31) / 32 * 32; }

}  // namespace
}  // namespace dn

using namespace dn;

extern "C" {

int dn_grad_create(dn_sys* s, const double* C, const double* n, const double* cj12,
                   const double* ck12, const double* elag, dn_grad** out) {
  DN_TRY(check_sys(s));
  if (!C || !n || !cj12 || !ck12 || !elag || !out) return fail(DN_ERR_ARG, "null argument");
  *out = nullptr;
  if (!s->b_ready) return fail(DN_ERR_STATE, "b has not been set");
  if (!s->b)
    return fail(DN_ERR_STATE, "nuclear gradients need an FP64 system (a split-precision "
                              "system keeps b only as FP16 planes)");
  if (!s->riw_set && s->eig_clamped)
    return fail(DN_ERR_STATE, "the RI metric had eigenvalues below zero (set to zero in "
                              "b): I3c G^{-1} cannot be formed from b");
  if (!s->riw_set)
    return fail(DN_ERR_STATE, "no RI factor for b: set b with dn_sys_ri_metric, or give "
                              "the factor of an uploaded b with dn_sys_set_ri_factor");
  if (s->lockstep) return fail(DN_ERR_STATE, "the system is driven by a lock-step group");
  const dn_dims& dd = s->d;
  const size_t nbf = dd.nbf, naux = dd.nbfaux, nbf5 = dd.nbf5, nn = nbf5 * nbf5;
  const long len3 = (long)(nbf * nbf * naux);
  // the longest dot product (d with geri3c, m with geri2c or the nbf^2 densities)
  const long maxlen = std::max<long>(len3, (long)std::max(naux * naux, nbf * nbf));
  const long stage_len = std::min<long>(kStageLen, maxlen);
  const long part_len = (maxlen + stage_len - 1) / stage_len * 3 * kDotBlocks;
  // split-K of the m GEMM (K = nbf5^2, an naux x naux output: too few tiles alone)
  const int msplit = nn % 4 == 0 && nn / 4 >= 64 ? 4 : (nn % 2 == 0 && nn >= 128 ? 2 : 1);
  // persistent: d, m, rdm1, lag, stage, part, res; transient: I2, D, dg, CT, CnT, T and
  // the m partials
  const size_t keep = r256(len3) + r256(naux * naux) + 2 * r256(nbf * nbf) +
                      r256(3 * stage_len) + r256(part_len) + r256(16);
  const size_t tmp = 2 * r256(nn * naux) + r256(nbf5 * naux) + 2 * r256(nbf * nbf5) +
                     r256(nbf * nbf) + (msplit > 1 ? r256(msplit * naux * naux) : 0);
  auto* g = new dn_grad;
  g->s = s;
  g->nbf = dd.nbf; g->naux = dd.nbfaux; g->nbf5 = dd.nbf5;
  g->bytes = keep * sizeof(double);
  int rc = ctx_reserve(s->ctx, g->bytes + tmp * sizeof(double));
  if (rc != DN_OK) {
    delete g;
    return rc;
  }
  double* w = nullptr;
  rc = dev_alloc(s->ctx, s->stream, keep * sizeof(double), &g->mem);
  if (rc == DN_OK) rc = dev_alloc(s->ctx, s->stream, tmp * sizeof(double),
                                  reinterpret_cast<void**>(&w));
  auto run = [&]() -> int {
    double* p = static_cast<double*>(g->mem);
    g->d = p;       p += r256(len3);
    g->m = p;       p += r256(naux * naux);
    g->rdm1 = p;    p += r256(nbf * nbf);
    g->lag = p;     p += r256(nbf * nbf);
    g->stage = p;   p += r256(3 * stage_len);
    g->stage_len = stage_len;
    g->part = p;    p += r256(part_len);
    g->part_len = part_len;
    g->res = p;
    double* I2 = w;
    double* D = I2 + r256(nn * naux);
    double* dg = D + r256(nn * naux);
    double* CT = dg + r256(nbf5 * naux);
    double* CnT = CT + r256(nbf * nbf5);
    double* T = CnT + r256(nbf * nbf5);
    double* mpart = T + r256(nbf * nbf);
    hipStream_t st = s->stream;
    DN_TRY(h2d(s, s->C, C, nbf * nbf));
    DN_TRY(h2d(s, s->n, n, nbf5));
    DN_TRY(h2d(s, s->cj12, cj12, nn));
    DN_TRY(h2d(s, s->ck12, ck12, nn));
    DN_TRY(h2d(s, s->elag, elag, nbf * nbf));
    // b_pq[p,q,Q] = (C5^T b_Q C5)[p,q] at q + nbf5 p + nbf5^2 Q: the J/K build's transform
    // (for a symmetric b its half contraction W_Q, then b_pq = W + W^T into D's buffer)
    DN_TRY(pad_c(s));
    const double* bpq = s->bt;
    if (s->b_sym && s->use_sym) {
      DN_TRY(gemm_f64(st, t1_gemm(s, true)));
      DN_TRY(gemm_f64(st, t2_flat_gemm(s)));
      const long nB = (long)(nn * naux);
      hipLaunchKernelGGL(grad_sym_kernel, dim3((nB + 255) / 256), dim3(256), 0, st, s->bt,
                         (int)nbf5, (int)naux, D);
      DN_HIP(hipGetLastError());
      bpq = D;
    } else {
      DN_TRY(gemm_f64(st, t1_gemm(s, false)));
      DN_TRY(gemm_f64(st, t2_flat_gemm(s)));
    }
    // I2[p,q,R] = sum_P b_pq[p,q,P] Wt[P,R], same layout (grads.jl:61,66)
    Gemm i2;
    i2.M = (int)nn; i2.N = (int)naux; i2.K = (int)naux;
    i2.A = bpq; i2.lda = (long)nn; i2.a_kcontig = false;
    i2.B = s->riw; i2.ldb = (long)naux;
    i2.b_kcontig = !s->riw_transposed;  // Wt[P + naux R], or riw = Wt^T: Wt[P,R] at R + naux P
    i2.C = I2; i2.ldc = (long)nn;
    DN_TRY(gemm_f64(st, i2));
    // D (grads.jl:68-84)
    hipLaunchKernelGGL(grad_diag_kernel, dim3(naux), dim3(256), nbf5 * sizeof(double), st, I2,
                       s->cj12, s->n, (int)nbf5, dd.nalpha, dd.nbeta, dg);
    DN_HIP(hipGetLastError());
    const long nD = (long)(nn * naux);
    hipLaunchKernelGGL(grad_dens_kernel, dim3((nD + 255) / 256), dim3(256), 0, st, I2, s->ck12,
                       dg, (int)nbf5, (int)naux, D);
    DN_HIP(hipGetLastError());
    // m[Q,R] = sum_{(p,q)} D[p,q,Q] I2[q,p,R]: D at p + nbf5 q, I2[q,p] at p + nbf5 q (89)
    Gemm gm;
    gm.M = (int)naux; gm.N = (int)naux; gm.K = (int)nn;
    gm.A = D; gm.lda = (long)nn; gm.a_kcontig = true;
    gm.B = I2; gm.ldb = (long)nn; gm.b_kcontig = true;
    gm.C = g->m; gm.ldc = (long)naux;
    gm.lds = true;
    if (msplit > 1) {
      DN_TRY(gemm_f64(st, splitk_partials(gm, msplit, mpart)));
      DN_TRY(reduce_splits(st, mpart, (long)(naux * naux), msplit, (int)naux, (int)naux,
                           (long)naux, g->m, (long)naux, 0.0, 1.0));
    } else {
      DN_TRY(gemm_f64(st, gm));
    }
    // d_Q = C5 D_Q C5^T (grads.jl:69,74,79,84): Zt[p,s,Q] = sum_q D[p,q,Q] C5[s,q] into
    // bpn (free after the transform), then d[l,(s,Q)] = sum_p C5[l,p] Zt[p,(s,Q)]
    const long nC = (long)(nbf * nbf5);
    hipLaunchKernelGGL(grad_ct_kernel, dim3((nC + 255) / 256), dim3(256), 0, st, s->C, s->n,
                       (int)nbf, (int)nbf5, CT, CnT);
    DN_HIP(hipGetLastError());
    double* Zt = s->bpn;  // nbf5*nbf*naux <= ldb*nbf5*naux
    Gemm za;
    za.M = (int)nbf5; za.N = (int)nbf; za.K = (int)nbf5;
    za.A = D; za.lda = (long)nbf5; za.sA = (long)nn; za.a_kcontig = false;
    za.B = CT; za.ldb = (long)nbf5; za.sB = 0; za.b_kcontig = true;  // C5[s,q] = CT[q + nbf5 s]
    za.C = Zt; za.ldc = (long)nbf5; za.sC = (long)(nbf5 * nbf);
    za.nbatch = (int)naux; za.bps = 1;
    DN_TRY(gemm_f64(st, za));
    Gemm zb;
    zb.M = (int)nbf; zb.N = (int)(nbf * naux); zb.K = (int)nbf5;
    zb.A = CT; zb.lda = (long)nbf5; zb.a_kcontig = true;
    zb.B = Zt; zb.ldb = (long)nbf5; zb.b_kcontig = true;
    zb.C = g->d; zb.ldc = (long)nbf;
    zb.lds = true;
    DN_TRY(gemm_f64(st, zb));
    // RDM1 = 2 C5 diag(n) C5^T (grads.jl:5)
    Gemm r1;
    r1.M = (int)nbf; r1.N = (int)nbf; r1.K = (int)nbf5;
    r1.A = CnT; r1.lda = (long)nbf5; r1.a_kcontig = true;
    r1.B = CT; r1.ldb = (long)nbf5; r1.b_kcontig = true;
    r1.C = g->rdm1; r1.ldc = (long)nbf;
    r1.alpha = 2.0;
    DN_TRY(gemm_f64(st, r1));
    // lag = 2 C elag C^T (grads.jl:6)
    Gemm l1;
    l1.M = (int)nbf; l1.N = (int)nbf; l1.K = (int)nbf;
    l1.A = s->C; l1.lda = (long)nbf; l1.a_kcontig = false;
    l1.B = s->elag; l1.ldb = (long)nbf; l1.b_kcontig = true;
    l1.C = T; l1.ldc = (long)nbf;
    DN_TRY(gemm_f64(st, l1));
    Gemm l2;
    l2.M = (int)nbf; l2.N = (int)nbf; l2.K = (int)nbf;
    l2.A = T; l2.lda = (long)nbf; l2.a_kcontig = false;
    l2.B = s->C; l2.ldb = (long)nbf; l2.b_kcontig = false;  // C[l,p] at l + nbf p
    l2.C = g->lag; l2.ldc = (long)nbf;
    l2.alpha = 2.0;
    DN_TRY(gemm_f64(st, l2));
    DN_HIP(hipStreamSynchronize(st));
    return DN_OK;
  };
  if (rc == DN_OK) rc = run();
  if (w) dev_free(s->ctx, s->stream, w);
  (void)hipStreamSynchronize(s->stream);
  ctx_unreserve(s->ctx, tmp * sizeof(double));
  if (rc != DN_OK) {
    const std::string msg = dn_last_error();
    if (g->mem) dev_free(s->ctx, s->stream, g->mem);
    ctx_unreserve(s->ctx, g->bytes);
    delete g;
    return fail(rc, msg);
  }
  count_grad(s, +1);
  *out = g;
  return DN_OK;
}

int dn_grad_atom(dn_grad* g, const double* dS, const double* dT, const double* dV,
                 const double* geri3c, const double* geri2c, double* grad3) {
  if (!g) return fail(DN_ERR_ARG, "null dn_grad");
  if (!dS || !dT || !dV || !geri3c || !geri2c || !grad3) return fail(DN_ERR_ARG, "null argument");
  DN_TRY(check_sys(g->s));
  const long n2 = (long)g->nbf * g->nbf;
  DN_TRY(dot3(g, g->lag, dS, n2, g->res));
  DN_TRY(dot3(g, g->rdm1, dT, n2, g->res + 3));
  DN_TRY(dot3(g, g->rdm1, dV, n2, g->res + 6));
  DN_TRY(dot3(g, g->d, geri3c, n2 * g->naux, g->res + 9));
  DN_TRY(dot3(g, g->m, geri2c, (long)g->naux * g->naux, g->res + 12));
  double r[15];
  DN_HIP(hipMemcpyAsync(r, g->res, sizeof(r), hipMemcpyDeviceToHost, g->s->stream));
  DN_HIP(hipStreamSynchronize(g->s->stream));
  for (int k = 0; k < 3; ++k) {
    grad3[k] = -r[k] + r[3 + k] + r[6 + k] + 2.0 * r[9 + k] - r[12 + k];
    if (!std::isfinite(grad3[k])) return fail(DN_ERR_NONFINITE, "non-finite gradient");
  }
  return DN_OK;
}

int dn_grad_get_densities(dn_grad* g, double* RDM1, double* lag, double* m, double* d) {
  if (!g) return fail(DN_ERR_ARG, "null dn_grad");
  dn_sys* s = g->s;
  DN_TRY(check_sys(s));
  const size_t n2 = (size_t)g->nbf * g->nbf;
  if (RDM1) DN_TRY(d2h(s, RDM1, g->rdm1, n2));
  if (lag) DN_TRY(d2h(s, lag, g->lag, n2));
  if (m) DN_TRY(d2h(s, m, g->m, (size_t)g->naux * g->naux));
  if (d) DN_TRY(d2h(s, d, g->d, n2 * g->naux));
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

int dn_grad_destroy(dn_grad* g) {
  if (!g) return DN_OK;
  dn_sys* s = g->s;
  (void)hipSetDevice(s->ctx->device);
  dev_free(s->ctx, s->stream, g->mem);
  (void)hipStreamSynchronize(s->stream);
  ctx_unreserve(s->ctx, g->bytes);
  count_grad(s, -1);
  delete g;
  return DN_OK;
}

}  // extern "C"
