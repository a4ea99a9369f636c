// RI metric on the device: b[m,n,:] = I3c[m,n,:] * G^{-1/2} up to an orthogonal
// rotation of the auxiliary index (integrals.jl:68-91, Iaux; ext/DoNOFGPUExt.jl:9-33).
//
// The reference diagonalises G and forms G^{-1/2}.  Here G = L L^T (blocked Cholesky,
// FP64) and b = I3c L^{-T}.  The two results differ by b_ref = b Q with the orthogonal
// Q = L^T G^{-1/2}; every consumer of b (J, K, the Lagrangian, the energy) contracts the
// auxiliary index in pairs, sum_Q b[.,Q] b[.,Q], so all of them are identical.  G must be
// positive definite (a Coulomb metric is); the eigen route's clamp of negative
// eigenvalues has no Cholesky counterpart, so such a G is rejected.
//
// A G that is not positive definite takes the reference's own route instead: G = V L V^T
// (rocSOLVER syevd on the device, upper triangle as Symmetric(G) reads it), 1/sqrt(l)
// for l >= 0 and 0 for l < 0, G^{-1/2} = V diag V^T and b = I3c G^{-1/2}: the
// reference's b itself.
//
// Layout: G column-major naux x naux (lower triangle read by the Cholesky route),
// I3c column-major (nbf, nbf, naux), i.e. an (nbf^2 x naux) matrix with the auxiliary
// index as columns.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "dn_internal.h"

namespace dn {
namespace {

constexpr int NB = 64;  // Cholesky block size

// Factor the kb x kb diagonal block at (k, k) in LDS, write L_kk back into G and
// L_kk^{-1} (full kb x kb, zero above the diagonal) into rows k..k+kb of Dinv.
__global__ __launch_bounds__(256) void potrf_block_kernel(double* G, long ld, int k, int kb,
                                                          double* Dinv, int* flag) {
  __shared__ double a[NB][NB + 1];
  __shared__ double inv[NB][NB + 1];
  __shared__ double piv;
  const int tid = threadIdx.x;
  for (int i = tid; i < kb * kb; i += blockDim.x) {
    const int r = i % kb, c = i / kb;
    a[r][c] = G[(k + r) + (long)(k + c) * ld];
  }
  for (int j = 0; j < kb; ++j) {
    __syncthreads();
    if (tid == 0) {
      double d = a[j][j];
      if (!(d > 0.0) || !isfinite(d)) {
        atomicOr(flag, 1);
        d = 1.0;
      }
      piv = sqrt(d);
      a[j][j] = piv;
    }
    __syncthreads();
    for (int r = j + 1 + tid; r < kb; r += blockDim.x) a[r][j] /= piv;
    __syncthreads();
    const int m = kb - j - 1;
    for (int i = tid; i < m * m; i += blockDim.x) {
      const int r = j + 1 + i % m, c = j + 1 + i / m;
      if (c <= r) a[r][c] -= a[r][j] * a[c][j];
    }
  }
  __syncthreads();
  for (int i = tid; i < kb * kb; i += blockDim.x) {
    const int r = i % kb, c = i / kb;
    if (r >= c) G[(k + r) + (long)(k + c) * ld] = a[r][c];
  }
  // column c of L^{-1} by forward substitution: L x = e_c
  if (tid < kb) {
    const int c = tid;
    for (int r = 0; r < c; ++r) inv[r][c] = 0.0;
    inv[c][c] = 1.0 / a[c][c];
    for (int r = c + 1; r < kb; ++r) {
      double acc = 0.0;
      for (int i = c; i < r; ++i) acc += a[r][i] * inv[i][c];
      inv[r][c] = -acc / a[r][r];
    }
  }
  __syncthreads();
  for (int i = tid; i < kb * kb; i += blockDim.x) {
    const int r = i % kb, c = i / kb;
    Dinv[(k + r) + (long)c * ld] = inv[r][c];
  }
}

// dst[r + c*ldd] = src[r + c*lds] for r < rows, c < cols
__global__ void copy_block_kernel(double* dst, long ldd, const double* src, long lds, int rows,
                                  int cols) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)rows * cols) return;
  const int r = i % rows, c = i / rows;
  dst[r + c * ldd] = src[r + c * lds];
}

int copy_block(hipStream_t s, double* dst, long ldd, const double* src, long lds, int rows,
               int cols) {
  const long n = (long)rows * cols;
  if (n <= 0) return DN_OK;
  hipLaunchKernelGGL(copy_block_kernel, dim3((n + 255) / 256), dim3(256), 0, s, dst, ldd, src,
                     lds, rows, cols);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// VS[:, i] = V[:, i] * (l_i < 0 ? 0 : 1 / sqrt(l_i)) (integrals.jl:76-82)
__global__ void scale_evec_kernel(const double* __restrict__ V, const double* __restrict__ lam,
                                  int n, double* __restrict__ VS) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)n * n) return;
  const double l = lam[i / n];
  VS[i] = l < 0.0 ? 0.0 : V[i] / sqrt(l);
}

// the n x n identity, column-major
__global__ void identity_kernel(double* X, int n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)n * n) return;
  X[i] = (i % n == i / n) ? 1.0 : 0.0;
}

}  // namespace

// In place: the lower triangle of G (naux x naux, ld = naux) becomes L with G = L L^T;
// Dinv (naux x NB) receives the inverses of the diagonal blocks; T (naux x NB) is the
// panel workspace.  *flag (device) is set when a pivot is not positive.
int cholesky_f64(hipStream_t s, double* G, int naux, double* Dinv, double* T, int* flag) {
  const long ld = naux;
  for (int k = 0; k < naux; k += NB) {
    const int kb = std::min(NB, naux - k);
    hipLaunchKernelGGL(potrf_block_kernel, dim3(1), dim3(256), 0, s, G, ld, k, kb, Dinv, flag);
    DN_HIP(hipGetLastError());
    const int rem = naux - k - kb;
    if (rem <= 0) break;
    // panel: T = G[k+kb:, k:k+kb] L_kk^{-T}
    Gemm p;
    p.M = rem; p.N = kb; p.K = kb;
    p.A = G + (k + kb) + (long)k * ld; p.lda = ld; p.a_kcontig = false;
    p.B = Dinv + k; p.ldb = ld; p.b_kcontig = false;  // B(q, c) = Dinv[(k + c) + q ld]
    p.C = T; p.ldc = ld;
    DN_TRY(gemm_f64(s, p));
    // trailing update: G[k+kb:, k+kb:] -= T T^T
    Gemm u;
    u.M = rem; u.N = rem; u.K = kb;
    u.A = T; u.lda = ld; u.a_kcontig = false;
    u.B = T; u.ldb = ld; u.b_kcontig = false;
    u.C = G + (k + kb) * (ld + 1); u.ldc = ld;
    u.alpha = -1.0; u.beta = 1.0;
    DN_TRY(gemm_f64(s, u));
    DN_TRY(copy_block(s, G + (k + kb) + (long)k * ld, ld, T, ld, rem, kb));
  }
  return DN_OK;
}

// X (rows x naux, ld = rows) is overwritten with X - (solved columns) and the solution
// Y = X L^{-T} is written to Y (rows x naux, ld = rows): column block k is
// Y_k = (X_k - Y_{<k} L_{k,<k}^T) L_kk^{-T}.
int trsm_right_lt(hipStream_t s, double* X, double* Y, long rows, int naux, const double* L,
                  const double* Dinv) {
  if (rows > INT32_MAX) return fail(DN_ERR_ARG, "RI metric: nbf^2 exceeds 2^31");
  const long ld = naux;
  for (int k = 0; k < naux; k += NB) {
    const int kb = std::min(NB, naux - k);
    if (k > 0) {
      Gemm g;
      g.M = (int)rows; g.N = kb; g.K = k;
      g.A = Y; g.lda = rows; g.a_kcontig = false;
      g.B = L + k; g.ldb = ld; g.b_kcontig = false;  // B(q, c) = L[(k + c) + q ld]
      g.C = X + k * rows; g.ldc = rows;
      g.alpha = -1.0; g.beta = 1.0;
      DN_TRY(gemm_f64(s, g));
    }
    Gemm d;
    d.M = (int)rows; d.N = kb; d.K = kb;
    d.A = X + k * rows; d.lda = rows; d.a_kcontig = false;
    d.B = Dinv + k; d.ldb = ld; d.b_kcontig = false;
    d.C = Y + k * rows; d.ldc = rows;
    DN_TRY(gemm_f64(s, d));
  }
  return DN_OK;
}

}  // namespace dn

using namespace dn;

extern "C" int dn_sys_set_ri_factor(dn_sys* s, const double* Wt) {
  DN_TRY(check_sys(s));
  if (!Wt) return fail(DN_ERR_ARG, "null host pointer");
  if (!s->b_ready) return fail(DN_ERR_STATE, "set b before its RI factor");
  const size_t naux = s->d.nbfaux;
  DN_TRY(h2d(s, s->riw, Wt, naux * naux));
  DN_HIP(hipStreamSynchronize(s->stream));
  s->riw_set = true;
  s->riw_transposed = false;
  return DN_OK;
}

namespace dn {
namespace {

// G (in dG, overwritten by V) = V diag(lam) V^T on the device; *bad = syevd's info
int syevd(hipStream_t st, double* dG, int naux, double* lam, double* E, int* info, int* bad) {
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success)
    return fail(DN_ERR_HIP, "rocblas_create_handle failed");
  rocblas_status r = rocblas_set_stream(h, st);
  if (r == rocblas_status_success)
    r = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, naux, dG, naux, lam, E,
                         info);
  if (r == rocblas_status_success) {
    const hipError_t e = hipMemcpyAsync(bad, info, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) (void)hipStreamSynchronize(st);
    else r = rocblas_status_internal_error;
  }
  rocblas_destroy_handle(h);
  if (r != rocblas_status_success)
    return fail(DN_ERR_HIP, std::string("rocsolver_dsyevd: ") + rocblas_status_to_string(r));
  if (*bad) return fail(DN_ERR_ARG, "RI metric: the eigensolver did not converge");
  return DN_OK;
}

}  // namespace
}  // namespace dn

// The metric solve of dn_sys_ri_metric with the three-centre integrals written into the
// device workspace by fill_I (rows ldb per column, slab stride ldb nbf; pad rows zero).
static int ri_metric_impl(dn_sys* s, const double* G,
                          const std::function<int(double*)>& fill_I) {
  const int naux = s->d.nbfaux;
  // rows of b in its resident layout: ldb per column of each slab, the pad rows zero (the
  // solve maps zero rows of I3c to zero rows of b)
  const long rows = (long)s->ldb * s->d.nbf;
  const size_t nG = (size_t)naux * naux, nP = (size_t)naux * NB;
  const size_t nI = std::max((size_t)rows * naux, nG);  // I3c, then the identity
  // + the eigen route: V diag(s) and G^{-1/2} (naux^2 each), eigenvalues, syevd's E, info
  const size_t bytes = (nI + 3 * nG + 2 * nP + 2 * (size_t)naux) * sizeof(double) + 512;
  DN_TRY(ctx_reserve(s->ctx, bytes));
  char* w = nullptr;
  {
    const int rc0 = dev_alloc(s->ctx, s->stream, bytes, reinterpret_cast<void**>(&w));
    if (rc0 != DN_OK) {
      ctx_unreserve(s->ctx, bytes);
      return rc0;
    }
  }
  double* dI = reinterpret_cast<double*>(w);
  double* dG = dI + nI;
  double* Dinv = dG + nG;
  double* T = Dinv + nP;
  double* VS = T + nP;
  double* Gmh = VS + nG;
  double* lam = Gmh + nG;
  double* E = lam + naux;
  int* flag = reinterpret_cast<int*>(E + naux);
  int* info = flag + 2;
  int rc = DN_OK, host_flag = 0, host_info = 0;
  auto run = [&]() -> int {
    if (s->ldb != s->d.nbf) DN_HIP(hipMemsetAsync(dI, 0, nI * sizeof(double), s->stream));
    DN_TRY(fill_I(dI));
    DN_TRY(h2d(s, dG, G, nG));
    DN_HIP(hipMemsetAsync(flag, 0, sizeof(int), s->stream));
    DN_TRY(cholesky_f64(s->stream, dG, naux, Dinv, T, flag));
    DN_HIP(hipMemcpyAsync(&host_flag, flag, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    DN_HIP(hipStreamSynchronize(s->stream));
    if (host_flag) {
      // not positive definite: the reference's eigen route with its clamp
      DN_TRY(h2d(s, dG, G, nG));
      DN_TRY(syevd(s->stream, dG, naux, lam, E, info, &host_info));
      std::vector<double> l(naux);
      DN_TRY(d2h(s, l.data(), lam, naux));
      DN_HIP(hipStreamSynchronize(s->stream));
      bool clamped = false;
      for (double x : l) clamped = clamped || x < 0.0;
      hipLaunchKernelGGL(scale_evec_kernel, dim3((nG + 255) / 256), dim3(256), 0, s->stream, dG,
                         lam, naux, VS);
      DN_HIP(hipGetLastError());
      Gemm gm;  // G^{-1/2} = (V diag(s)) V^T
      gm.M = naux; gm.N = naux; gm.K = naux;
      gm.A = VS; gm.lda = naux; gm.a_kcontig = false;
      gm.B = dG; gm.ldb = naux; gm.b_kcontig = false;  // V^T(k, c) = V[c + naux k]
      gm.C = Gmh; gm.ldc = naux;
      DN_TRY(gemm_f64(s->stream, gm));
      DN_TRY(set_b(s, [&](double* dst) {
        Gemm gb;  // b = I3c G^{-1/2} (integrals.jl:86-88)
        gb.M = (int)rows; gb.N = naux; gb.K = naux;
        gb.A = dI; gb.lda = rows; gb.a_kcontig = false;
        gb.B = Gmh; gb.ldb = naux; gb.b_kcontig = true;
        gb.C = dst; gb.ldc = rows;
        return gemm_f64(s->stream, gb);
      }));
      if (!clamped) {
        // I3c G^{-1} = b G^{-1/2}: the factor of the nuclear gradients; with a clamped
        // eigenvalue I3c is not recoverable from b and the gradients are refused
        DN_HIP(hipMemcpyAsync(s->riw, Gmh, nG * sizeof(double), hipMemcpyDeviceToDevice,
                              s->stream));
      }
      DN_HIP(hipStreamSynchronize(s->stream));
      s->eig_route = true;
      s->eig_clamped = clamped;
      return DN_OK;
    }
    s->eig_route = false;
    s->eig_clamped = false;
    DN_TRY(set_b(s, [&](double* dst) {
      return trsm_right_lt(s->stream, dI, dst, rows, naux, dG, Dinv);
    }));
    // riw = 1 L^{-T}: I3c G^{-1} = b L^{-1} = b riw^T for the nuclear gradients
    hipLaunchKernelGGL(identity_kernel, dim3((nG + 255) / 256), dim3(256), 0, s->stream, dI,
                       naux);
    DN_HIP(hipGetLastError());
    DN_TRY(trsm_right_lt(s->stream, dI, s->riw, naux, naux, dG, Dinv));
    DN_HIP(hipStreamSynchronize(s->stream));
    return DN_OK;
  };
  rc = run();
  dev_free(s->ctx, s->stream, w);
  (void)hipStreamSynchronize(s->stream);
  ctx_unreserve(s->ctx, bytes);
  if (rc == DN_OK) {
    s->riw_set = !s->eig_clamped;
    s->riw_transposed = !s->eig_route;  // L^{-T} (Cholesky) or G^{-1/2} itself
  }
  return rc;
}

extern "C" int dn_sys_ri_metric(dn_sys* s, const double* G, const double* I3c) {
  DN_TRY(check_sys(s));
  if (!G || !I3c) return fail(DN_ERR_ARG, "null host pointer");
  return ri_metric_impl(s, G, [&](double* dI) -> int {
    const size_t nbf = s->d.nbf;
    DN_HIP(hipMemcpy2DAsync(dI, s->ldb * sizeof(double), I3c, nbf * sizeof(double),
                            nbf * sizeof(double), nbf * s->d.nbfaux, hipMemcpyHostToDevice,
                            s->stream));
    return DN_OK;
  });
}

extern "C" int dn_sys_ri_from_basis(dn_sys* s, const dn_basis* basis, const dn_basis* aux) {
  DN_TRY(check_sys(s));
  if (!basis || !aux) return fail(DN_ERR_ARG, "null basis");
  int nbf = 0, naux = 0;
  DN_TRY(dn_basis_nbf(basis, &nbf));
  DN_TRY(dn_basis_nbf(aux, &naux));
  if (nbf != s->d.nbf || naux != s->d.nbfaux)
    return fail(DN_ERR_ARG, "dn_sys_ri_from_basis: basis sizes differ from the system's");
  std::vector<double> G((size_t)naux * naux);
  DN_TRY(dn_int2c2e(aux, G.data()));
  for (int i = 0; i < naux; ++i)  // symmetrised as integrals.jl:51 does
    for (int j = 0; j < i; ++j) {
      const double v = 0.5 * (G[i + (size_t)naux * j] + G[j + (size_t)naux * i]);
      G[i + (size_t)naux * j] = G[j + (size_t)naux * i] = v;
    }
  return ri_metric_impl(s, G.data(), [&](double* dI) {
    return int3c2e_device(s->stream, basis, aux, dI, s->ldb);
  });
}

extern "C" int dn_int3c2e_device(const dn_basis* basis, const dn_basis* aux, double* I3c) {
  if (!basis || !aux || !I3c) return fail(DN_ERR_ARG, "dn_int3c2e_device: null argument");
  int nbf = 0, naux = 0;
  DN_TRY(dn_basis_nbf(basis, &nbf));
  DN_TRY(dn_basis_nbf(aux, &naux));
  const size_t n = (size_t)nbf * nbf * naux;
  hipStream_t st = nullptr;
  DN_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  double* d = nullptr;
  int rc = DN_OK;
  hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&d), n * sizeof(double), st);
  if (e == hipSuccess) {
    rc = int3c2e_device(st, basis, aux, d, nbf);
    if (rc == DN_OK) e = hipMemcpyAsync(I3c, d, n * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFreeAsync(d, st);
    (void)hipStreamSynchronize(st);
  }
  (void)hipStreamDestroy(st);
  if (rc != DN_OK) return rc;
  if (e != hipSuccess) return fail(DN_ERR_HIP, std::string("dn_int3c2e_device: ") + hipGetErrorString(e));
  return DN_OK;
}
