// Orbital rotation Cnew = C exp(Y) (utils.jl:499-521) and the ADAM-family moment step
// of the orbital optimisers (minimization.jl:242-630), in FP64 on the device.
//
// exp(Y) uses scaling and squaring with a degree-12 Taylor polynomial evaluated by the
// Paterson-Stockmeyer scheme (Paterson & Stockmeyer 1973; Higham, "Functions of
// Matrices", SIAM 2008, sec. 4.2 and 10.3): X = Y / 2^s with ||X||_1 <= 1/4, so the
// truncation error is below ||X||^13/13! < 3e-18, then s squarings.  Y is antisymmetric,
// so U is orthogonal to rounding.
#include <cmath>

#include "dn_internal.h"
#include "dn_rot_dev.h"

namespace dn {

// Y[a,b] from y (antisymmetric, zero outside rows/columns < nbf5)
__global__ void build_y_kernel(const double* __restrict__ y, int nbf, int nbf5,
                               double* __restrict__ Y) {
  const long total = (long)nbf * nbf;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int a = e % nbf, b = e / nbf;
    double v = 0.0;
    if (a < b && a < nbf5) v = y[pair_index(a, b, nbf)];
    else if (b < a && b < nbf5) v = -y[pair_index(b, a, nbf)];
    Y[e] = v;
  }
}

// max_j sum_i |Y[i,j]|: one wavefront per column (coalesced reads); the maximum is taken
// on the bit patterns (non-negative doubles order as their bits, a NaN above all), so the
// result does not depend on the order the columns finish in.  *out must start at zero.
__global__ void norm1_kernel(const double* __restrict__ Y, int nbf,
                             unsigned long long* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (j >= nbf) return;
  const double s = column_abs_sum(Y + (long)j * nbf, nbf, lane);
  if (lane == 0) atomicMax(out, (unsigned long long)__double_as_longlong(s));
}

// out = c0 I + c1 X + c2 X2 + c3 X3 + c4 X4 (X4 may be null when c4 == 0), all scaled
// matrices already formed; `scale` multiplies X (X = Y / 2^s is applied here so that Y
// itself is never rescaled in place).
__global__ void poly_block_kernel(const double* __restrict__ X, const double* __restrict__ X2,
                                  const double* __restrict__ X3, const double* __restrict__ X4,
                                  double c0, double c1, double c2, double c3, double c4,
                                  int nbf, double* __restrict__ out) {
  const long total = (long)nbf * nbf;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x)
    out[e] = poly_element(e, nbf, X, X2, X3, X4, c0, c1, c2, c3, c4);
}

__global__ void scale_kernel(const double* __restrict__ in, double f, long n,
                             double* __restrict__ out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x)
    out[e] = f * in[e];
}

__global__ void add_kernel(double* __restrict__ a, const double* __restrict__ b, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x)
    a[e] += b[e];
}

// ---- ADAM family ----------------------------------------------------------------------

// g = pack(4 elag - 4 elag^T), and the running max |g| (as ordered bits of a
// non-negative double) into *gmax_bits.
__global__ void pack_grad_kernel(const double* __restrict__ elag, int nbf, int nbf5,
                                 double* __restrict__ g, unsigned long long* gmax_bits) {
  const long total = (long)nbf5 * nbf;
  double mx = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int j = e % nbf, i = e / nbf;  // consecutive threads -> consecutive j
    if (j <= i) continue;
    const double v = 4.0 * elag[i + (long)j * nbf] - 4.0 * elag[j + (long)i * nbf];
    g[pair_index(i, j, nbf)] = v;
    mx = fmax(mx, fabs(v));
  }
  // wave reduction then one atomic per wave (max is order independent)
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(gmax_bits, (unsigned long long)__double_as_longlong(mx));
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

// One AMSGrad-family step (minimization.jl:406-411 ADAM, 503-508 ADABelief, 598-603
// YOGI, 307-316 Demon) on the packed variables; also writes Y[i,j] = y, Y[j,i] = -y.
__global__ void adam_step_kernel(const double* __restrict__ g, double* __restrict__ m,
                                 double* __restrict__ v, double* __restrict__ vmax,
                                 double* __restrict__ y, int method, double beta1, double beta2,
                                 double bc1, double bc2, double alpha, int nbf, int nbf5,
                                 double* __restrict__ Y) {
  const long total = (long)nbf5 * nbf;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x)
    adam_update(e, g, m, v, vmax, y, method, beta1, beta2, bc1, bc2, alpha, nbf, Y);
}

namespace {
inline unsigned grid_for(long n) {
  long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}
}  // namespace

int build_y(hipStream_t s, const double* y, int nbf, int nbf5, double* Y) {
  hipLaunchKernelGGL(build_y_kernel, dim3(grid_for((long)nbf * nbf)), dim3(256), 0, s, y, nbf,
                     nbf5, Y);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int norm1(hipStream_t s, const double* Y, int nbf, double* out) {
  auto* bits = reinterpret_cast<unsigned long long*>(out);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, bits, 1);
  DN_HIP(hipGetLastError());
  hipLaunchKernelGGL(norm1_kernel, dim3((nbf + 3) / 4), dim3(256), 0, s, Y, nbf, bits);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

// Square GEMM out = A B (nbf x nbf, column-major), split along K
static int sq_gemm(hipStream_t s, const double* A, const double* B, double* out, int nbf,
                   double* part, size_t part_elems) {
  Gemm g;
  g.M = nbf; g.N = nbf; g.K = nbf;
  g.A = A; g.lda = nbf; g.a_kcontig = false;
  g.B = B; g.ldb = nbf; g.b_kcontig = true;
  g.C = out; g.ldc = nbf;
  return gemm_f64_splitk(s, g, part, part_elems);
}

// out = A^T B (A read along its columns), split along K as sq_gemm splits
static int tsq_gemm(hipStream_t s, const double* A, const double* B, double* out, int nbf,
                    double* part, size_t part_elems) {
  Gemm g;
  g.M = nbf; g.N = nbf; g.K = nbf;
  g.A = A; g.lda = nbf; g.a_kcontig = true;
  g.B = B; g.ldb = nbf; g.b_kcontig = true;
  g.C = out; g.ldc = nbf;
  return gemm_f64_splitk(s, g, part, part_elems);
}

__global__ void ns_polish_kernel(double* __restrict__ U, const double* __restrict__ X, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) U[i] = ns_polish(U[i], X[i]);
}

int matmul_sq(hipStream_t s, const double* A, const double* B, double* out, int nbf,
              double* part, size_t part_elems) {
  return sq_gemm(s, A, B, out, nbf, part, part_elems);
}

// Number of squarings for ||Y||_1 = nrm so that ||Y / 2^s||_1 <= 1/4.
int squarings_for(double nrm) {
  if (!(nrm > 0.25)) return 0;
  int s = (int)std::ceil(std::log2(nrm / 0.25));
  return s < 0 ? 0 : s;
}

// U = exp(Y) for ||Y||_1 = nrm.  work holds 5 nbf x nbf matrices.
int expm_taylor(hipStream_t s, const double* Y, double nrm, int nbf, double* work, double* U,
                double* part, size_t part_elems) {
  const long nn = (long)nbf * nbf;
  double* X = work;
  double* X2 = work + nn;
  double* X3 = work + 2 * nn;
  double* X4 = work + 3 * nn;
  double* T = work + 4 * nn;
  const int sq = squarings_for(nrm);
  const double f = std::ldexp(1.0, -sq);
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(nn)), dim3(256), 0, s, Y, f, nn, X);
  DN_HIP(hipGetLastError());
  DN_TRY(sq_gemm(s, X, X, X2, nbf, part, part_elems));
  DN_TRY(sq_gemm(s, X2, X, X3, nbf, part, part_elems));
  DN_TRY(sq_gemm(s, X2, X2, X4, nbf, part, part_elems));
  // Taylor coefficients c_k = 1/k!
  double c[13];
  c[0] = 1.0;
  for (int k = 1; k <= 12; ++k) c[k] = c[k - 1] / k;
  // p = B0 + X4 (B1 + X4 B2)
  hipLaunchKernelGGL(poly_block_kernel, dim3(grid_for(nn)), dim3(256), 0, s, X, X2, X3, X4, c[8],
                     c[9], c[10], c[11], c[12], nbf, T);  // T = B2
  DN_HIP(hipGetLastError());
  DN_TRY(sq_gemm(s, X4, T, U, nbf, part, part_elems));                         // U = X4 B2
  hipLaunchKernelGGL(poly_block_kernel, dim3(grid_for(nn)), dim3(256), 0, s, X, X2, X3,
                     (const double*)nullptr, c[4], c[5], c[6], c[7], 0.0, nbf, T);  // B1
  DN_HIP(hipGetLastError());
  hipLaunchKernelGGL(add_kernel, dim3(grid_for(nn)), dim3(256), 0, s, T, U, nn);  // B1 + X4 B2
  DN_HIP(hipGetLastError());
  DN_TRY(sq_gemm(s, X4, T, U, nbf, part, part_elems));                         // X4 (B1 + X4 B2)
  hipLaunchKernelGGL(poly_block_kernel, dim3(grid_for(nn)), dim3(256), 0, s, X, X2, X3,
                     (const double*)nullptr, c[0], c[1], c[2], c[3], 0.0, nbf, T);  // B0
  DN_HIP(hipGetLastError());
  hipLaunchKernelGGL(add_kernel, dim3(grid_for(nn)), dim3(256), 0, s, U, T, nn);
  DN_HIP(hipGetLastError());
  // squarings: U <- U^2, ping-ponging with T
  double* cur = U;
  double* oth = T;
  for (int k = 0; k < sq; ++k) {
    DN_TRY(sq_gemm(s, cur, cur, oth, nbf, part, part_elems));
    std::swap(cur, oth);
  }
  if (cur != U) DN_HIP(hipMemcpyAsync(U, cur, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
  // re-orthonormalisation: one Newton-Schulz step, U <- U (3 - U^T U) / 2, so that the
  // rounding of the products (and of the squarings) does not accumulate in C over the
  // iterations of the optimiser
  DN_TRY(tsq_gemm(s, U, U, X, nbf, part, part_elems));   // X = U^T U
  DN_TRY(sq_gemm(s, U, X, X2, nbf, part, part_elems));   // X2 = U X
  hipLaunchKernelGGL(ns_polish_kernel, dim3(grid_for(nn)), dim3(256), 0, s, U, X2, nn);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int pack_grad(hipStream_t s, const double* elag, int nbf, int nbf5, double* g,
              unsigned long long* gmax_bits) {
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, gmax_bits, 1);
  DN_HIP(hipGetLastError());
  hipLaunchKernelGGL(pack_grad_kernel, dim3(grid_for((long)nbf5 * nbf)), dim3(256), 0, s, elag,
                     nbf, nbf5, g, gmax_bits);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int adam_step(hipStream_t s, const double* g, double* m, double* v, double* vmax, double* y,
              int method, double beta1, double beta2, double bc1, double bc2, double alpha,
              int nbf, int nbf5, double* Y) {
  hipLaunchKernelGGL(adam_step_kernel, dim3(grid_for((long)nbf5 * nbf)), dim3(256), 0, s, g, m,
                     v, vmax, y, method, beta1, beta2, bc1, bc2, alpha, nbf, nbf5, Y);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

}  // namespace dn
