// C ABI of the occupation functional, the orbital rotation and the ADAM-family orbital
// optimiser.  Declarations and reference citations are in include/donof_hip.h.
#include <cmath>

#include "dn_internal.h"
#include "dn_occ.h"

namespace dn {

// dn_rot.hip
int build_y(hipStream_t s, const double* y, int nbf, int nbf5, double* Y);
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

namespace {

__global__ void fill_kernel(double* __restrict__ p, long n, double v) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x)
    p[e] = v;
}

__global__ void copy_kernel(double* __restrict__ dst, const double* __restrict__ src, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

int fill(hipStream_t s, double* p, long n, double v) {
  long b = (n + 255) / 256;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)std::min<long>(std::max<long>(b, 1), 4096)),
                     dim3(256), 0, s, p, n, v);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int copy(hipStream_t s, double* dst, const double* src, long n) {
  long b = (n + 255) / 256;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)std::min<long>(std::max<long>(b, 1), 4096)),
                     dim3(256), 0, s, dst, src, n);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int functional_of(dn_sys* s, const dn_functional* f, OccDims* P) {
  if (!f) return fail(DN_ERR_ARG, "null dn_functional");
  if (f->ipnof != 4 && f->ipnof != 5 && f->ipnof != 7 && f->ipnof != 8 && f->ipnof != 9)
    return fail(DN_ERR_ARG, "ipnof must be 4, 5, 7, 8 or 9");
  if (f->ista < 0 || f->ista > 5) return fail(DN_ERR_ARG, "ista out of range");
  if (f->occ_method != DN_OCC_SOFTMAX && f->occ_method != DN_OCC_TRIGONOMETRIC)
    return fail(DN_ERR_ARG, "occ_method must be Softmax or Trigonometric");
  const dn_dims& d = s->d;
  P->nbf5 = d.nbf5; P->no1 = d.no1; P->ndoc = d.ndoc; P->nsoc = d.nsoc;
  P->ndns = d.ndoc + d.nsoc; P->nbeta = d.nbeta; P->nalpha = d.nalpha;
  P->ncwo = d.ncwo; P->nv = d.nv;
  P->ipnof = f->ipnof; P->ista = f->ista; P->occ_method = f->occ_method;
  P->highspin = f->highspin; P->mspin = f->mspin; P->h_cut = f->h_cut;
  if (f->occ_method == DN_OCC_TRIGONOMETRIC && d.ncwo < 1)
    return fail(DN_ERR_ARG, "Trigonometric occupations need ncwo >= 1");
  return DN_OK;
}

OccWork work_of(dn_sys* s) {
  const int n5 = s->d.nbf5;
  OccWork w;
  w.n = s->n;
  w.fi = s->orb;
  w.nd = s->orb + n5;
  w.nd12 = s->orb + 2 * n5;
  w.dfic = s->orb + 3 * n5;
  w.ndc = s->orb + 4 * n5;
  w.ndx = s->orb + 5 * n5;
  w.strong = s->orbi;
  w.pair = s->orbi + n5;
  w.role = s->orbi + 2 * n5;
  w.scal = s->scal + 8;
  return w;
}

// n, dn/dgamma, per-orbital quantities and cj12/ck12 from the resident gamma
int occupations_resident(dn_sys* s, const OccDims& P) {
  DN_TRY(occ_ocupacion(s->stream, P, s->gamma, s->n, s->dn));
  const OccWork w = work_of(s);
  DN_TRY(occ_prepare(s->stream, P, w));
  DN_TRY(occ_coefficients(s->stream, P, w, s->cj12, s->ck12));
  return DN_OK;
}

int upload_jkh(dn_sys* s, const double* J, const double* K, const double* Hcore) {
  const size_t nn = (size_t)s->d.nbf5 * s->d.nbf5;
  if ((J == nullptr) != (K == nullptr) || (J == nullptr) != (Hcore == nullptr))
    return fail(DN_ERR_ARG, "J, K and H_core must all be given or all be NULL (resident)");
  if (J) {
    DN_TRY(h2d(s, s->J, J, nn));
    DN_TRY(h2d(s, s->K, K, nn));
    DN_TRY(h2d(s, s->Hcore, Hcore, s->d.nbf5));
  }
  return DN_OK;
}

// Cnew = C exp(Y(y)) on the device; y in s->yv, C and Cnew device buffers.
int rotate_resident(dn_sys* s, const double* C, double* Cnew, bool build) {
  const dn_dims& d = s->d;
  if (build) DN_TRY(build_y(s->stream, s->yv, d.nbf, d.nbf5, s->Y));
  DN_TRY(norm1(s->stream, s->Y, d.nbf, s->scal + 4));
  double nrm = 0.0;
  DN_TRY(d2h(s, &nrm, s->scal + 4, 1));
  DN_HIP(hipStreamSynchronize(s->stream));
  if (!std::isfinite(nrm)) return fail(DN_ERR_NONFINITE, "non-finite rotation generator");
  DN_TRY(expm_taylor(s->stream, s->Y, nrm, d.nbf, s->rwork, s->U, s->part2, s->part2_elems));
  DN_TRY(matmul_sq(s->stream, C, s->U, Cnew, d.nbf, s->part2, s->part2_elems));
  return DN_OK;
}

// Lagrangian at the resident C, then the packed gradient; returns E and max|g|.
int lagrangian_and_grad(dn_sys* s, int nalpha, int nbeta, double* E, double* gmax) {
  const dn_dims& d = s->d;
  DN_TRY(lagrange_resident(s, nalpha, nbeta));
  DN_TRY(pack_grad(s->stream, s->elag, d.nbf, d.nbf5, s->gv,
                   reinterpret_cast<unsigned long long*>(s->scal + 1)));
  double h[2];
  DN_TRY(d2h(s, h, s->scal, 2));
  DN_HIP(hipStreamSynchronize(s->stream));
  *E = h[0];
  *gmax = h[1];  // non-negative double stored as its bit pattern
  if (!std::isfinite(*E) || !std::isfinite(*gmax))
    return fail(DN_ERR_NONFINITE, "non-finite energy or orbital gradient");
  return DN_OK;
}

}  // namespace

// Start of an ADAM-family run (minimization.jl:344-360): occupations and coefficients from
// the resident gamma, the Lagrangian and gradient at the resident C, zero moments.
int orbopt_begin(dn_sys* s, const dn_functional* f, const dn_orbopt_state* st) {
  const dn_dims& d = s->d;
  OccDims P;
  DN_TRY(functional_of(s, f, &P));
  if (!st) return fail(DN_ERR_ARG, "null dn_orbopt_state");
  if (st->method < DN_ADAM || st->method > DN_DEMON) return fail(DN_ERR_ARG, "unknown method");
  if (!s->b_ready) return fail(DN_ERR_STATE, "b has not been uploaded or generated");
  const long nn = (long)d.nbf * d.nbf;
  DN_TRY(occupations_resident(s, P));
  double E = 0.0, gmax = 0.0;
  DN_TRY(lagrangian_and_grad(s, d.nalpha, d.nbeta, &E, &gmax));
  DN_TRY(fill(s->stream, s->mv, s->nvar, 0.0));
  DN_TRY(fill(s->stream, s->vv, s->nvar, 0.0));
  DN_TRY(fill(s->stream, s->vmax, s->nvar, 0.0));
  DN_TRY(fill(s->stream, s->Y, nn, 0.0));
  DN_TRY(copy(s->stream, s->bestC, s->C, nn));
  auto& o = s->opt;
  o = dn_sys::OptRun{};
  o.started = true;
  o.method = st->method;
  o.maxloop = st->maxloop;
  o.alpha = st->alpha;
  o.threshgorb = st->threshgorb;
  o.beta1 = o.binit = (st->method == DN_DEMON) ? 0.6 : 0.5;
  o.beta2 = (st->method == DN_DEMON) ? 0.999 : 0.9;
  o.E = o.best_E = E;
  o.gmax = gmax;
  o.nalpha = d.nalpha;
  o.nbeta = d.nbeta;
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

int fill_dev(hipStream_t s, double* p, long n, double v) { return fill(s, p, n, v); }
int copy_dev(hipStream_t s, double* dst, const double* src, long n) { return copy(s, dst, src, n); }

// The ADAM-family loop of minimization.jl:242-630 with C resident on the device.
int orbopt_resident(dn_sys* s, const dn_functional* f, dn_orbopt_state* st) {
  const dn_dims& d = s->d;
  DN_TRY(orbopt_begin(s, f, st));
  const long nn = (long)d.nbf * d.nbf;
  double E = s->opt.E, gmax = s->opt.gmax;
  const double alpha = st->alpha;
  double beta1 = s->opt.beta1;
  const double beta2 = s->opt.beta2;
  const double binit = beta1;
  bool improved = false, success = false;
  double best_E = E;
  int nit = 0;
  const int maxloop = st->maxloop;
  for (int i = 1; i <= maxloop; ++i) {
    ++nit;
    if (2 * gmax < st->threshgorb && improved) {  // minimization.jl:400-403
      success = true;
      break;
    }
    if (st->method == DN_DEMON) {  // minimization.jl:307-308
      const double rate = (i - 1.0) / (maxloop - 1.0);
      beta1 = binit * (1.0 - rate) / ((1.0 - binit) + binit * (1.0 - rate));
    }
    const double bc1 = 1.0 - std::pow(beta1, i), bc2 = 1.0 - std::pow(beta2, i);
    DN_TRY(adam_step(s->stream, s->gv, s->mv, s->vv, s->vmax, s->yv, st->method, beta1, beta2,
                     bc1, bc2, alpha, d.nbf, d.nbf5, s->Y));
    DN_TRY(rotate_resident(s, s->C, s->Ca, false));  // C <- C exp(Y)   (utils.jl:499-521)
    DN_TRY(copy(s->stream, s->C, s->Ca, nn));
    DN_TRY(lagrangian_and_grad(s, d.nalpha, d.nbeta, &E, &gmax));
    if (E < best_E) {
      best_E = E;
      DN_TRY(copy(s->stream, s->bestC, s->C, nn));
      improved = true;
    }
  }
  s->opt.started = false;
  if (!improved) {  // minimization.jl:426-431
    st->alpha = 0.3 * st->alpha;
    st->maxloop = st->maxloop + 10;
  }
  if (st->maxloop < 100) {
    DN_TRY(copy(s->stream, s->C, s->bestC, nn));
    st->E = best_E;
  } else {
    st->E = E;
  }
  st->nit = nit;
  st->success = success ? 1 : 0;
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

}  // namespace dn

using namespace dn;

extern "C" {

int dn_ocupacion(dn_sys* s, const dn_functional* f, const double* gamma, double* n,
                 double* dn_dgamma) {
  DN_TRY(check_sys(s));
  OccDims P;
  DN_TRY(functional_of(s, f, &P));
  DN_TRY(h2d(s, s->gamma, gamma, s->d.nv));
  DN_TRY(occ_ocupacion(s->stream, P, s->gamma, s->n, s->dn));
  DN_TRY(d2h(s, n, s->n, s->d.nbf5));
  DN_TRY(d2h(s, dn_dgamma, s->dn, (size_t)s->d.nbf5 * s->d.nv));
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

int dn_pnof_coefficients(dn_sys* s, const dn_functional* f, const double* n, double* cj12,
                         double* ck12) {
  DN_TRY(check_sys(s));
  OccDims P;
  DN_TRY(functional_of(s, f, &P));
  const size_t nn = (size_t)s->d.nbf5 * s->d.nbf5;
  DN_TRY(h2d(s, s->n, n, s->d.nbf5));
  const OccWork w = work_of(s);
  DN_TRY(occ_prepare(s->stream, P, w));
  DN_TRY(occ_coefficients(s->stream, P, w, s->cj12, s->ck12));
  DN_TRY(d2h(s, cj12, s->cj12, nn));
  DN_TRY(d2h(s, ck12, s->ck12, nn));
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

int dn_der_pnof_coefficients(dn_sys* s, const dn_functional* f, const double* n,
                             const double* dn_dgamma, double* Dcj12r, double* Dck12r) {
  DN_TRY(check_sys(s));
  OccDims P;
  DN_TRY(functional_of(s, f, &P));
  const size_t total = (size_t)s->d.nbf5 * s->d.nbf5 * s->d.nv;
  if (!Dcj12r || !Dck12r) return fail(DN_ERR_ARG, "null output");
  DN_TRY(h2d(s, s->n, n, s->d.nbf5));
  DN_TRY(h2d(s, s->dn, dn_dgamma, (size_t)s->d.nbf5 * s->d.nv));
  double* tmp = nullptr;
  DN_TRY(dev_alloc(s->ctx, s->stream, 2 * std::max<size_t>(total, 1) * sizeof(double),
                   reinterpret_cast<void**>(&tmp)));
  const OccWork w = work_of(s);
  int rc = occ_prepare(s->stream, P, w);
  if (rc == DN_OK) rc = occ_der_coefficients(s->stream, P, w, s->dn, tmp, tmp + total);
  if (rc == DN_OK) rc = d2h(s, Dcj12r, tmp, total);
  if (rc == DN_OK) rc = d2h(s, Dck12r, tmp + total, total);
  dev_free(s->ctx, s->stream, tmp);
  hipError_t e = hipStreamSynchronize(s->stream);
  if (rc != DN_OK) return rc;
  if (e != hipSuccess) return fail(DN_ERR_HIP, hipGetErrorString(e));
  return DN_OK;
}

int dn_calce(dn_sys* s, const dn_functional* f, const double* n, const double* cj12,
             const double* ck12, const double* J, const double* K, const double* Hcore,
             double* E) {
  DN_TRY(check_sys(s));
  OccDims P;
  DN_TRY(functional_of(s, f, &P));
  if (!E) return fail(DN_ERR_ARG, "null E");
  const size_t nn = (size_t)s->d.nbf5 * s->d.nbf5;
  DN_TRY(h2d(s, s->n, n, s->d.nbf5));
  DN_TRY(h2d(s, s->cj12, cj12, nn));
  DN_TRY(h2d(s, s->ck12, ck12, nn));
  DN_TRY(upload_jkh(s, J, K, Hcore));
  DN_TRY(occ_calce(s->stream, P, s->n, s->cj12, s->ck12, s->J, s->K, s->Hcore, s->scal + 2));
  DN_TRY(d2h(s, E, s->scal + 2, 1));
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

int dn_calcocce(dn_sys* s, const dn_functional* f, const double* gamma, const double* J,
                const double* K, const double* Hcore, double* E) {
  DN_TRY(check_sys(s));
  OccDims P;
  DN_TRY(functional_of(s, f, &P));
  if (!E) return fail(DN_ERR_ARG, "null E");
  DN_TRY(h2d(s, s->gamma, gamma, s->d.nv));
  DN_TRY(upload_jkh(s, J, K, Hcore));
  DN_TRY(occupations_resident(s, P));
  DN_TRY(occ_calce(s->stream, P, s->n, s->cj12, s->ck12, s->J, s->K, s->Hcore, s->scal + 2));
  DN_TRY(d2h(s, E, s->scal + 2, 1));
  DN_HIP(hipStreamSynchronize(s->stream));
  if (!std::isfinite(*E)) return fail(DN_ERR_NONFINITE, "non-finite occupation energy");
  return DN_OK;
}

int dn_occoptr(dn_sys* s, const dn_functional* f, const double* gamma, const double* J,
               const double* K, const double* Hcore, double threshen, int maxiter,
               double* gamma_out, double* E, int* nit, int* converged) {
  DN_TRY(check_sys(s));
  OccDims P;
  DN_TRY(functional_of(s, f, &P));
  if (!gamma || !gamma_out || !E || !nit || !converged) return fail(DN_ERR_ARG, "null argument");
  if (!(threshen > 0.0) || maxiter < 1) return fail(DN_ERR_ARG, "need threshen > 0, maxiter >= 1");
  const int nv = s->d.nv;
  if (nv <= 0) return fail(DN_ERR_ARG, "no occupation variables (ndoc = 0)");
  DN_TRY(upload_jkh(s, J, K, Hcore));
  const size_t bytes = sizeof(OccCG) + (kCgVecs * (size_t)nv + occ_cg_energy_blocks(P)) *
                                          sizeof(double) + 256;
  char* mem = nullptr;
  DN_TRY(dev_alloc(s->ctx, s->stream, bytes, reinterpret_cast<void**>(&mem)));
  OccCG* c = reinterpret_cast<OccCG*>(mem);
  double* vec = reinterpret_cast<double*>(mem + (sizeof(OccCG) + 255) / 256 * 256);
  OccCG h{};
  h.state = 0;  // first slot: evaluate at gamma
  h.tol = threshen;
  h.maxiter = maxiter;
  auto run = [&]() -> int {
    DN_HIP(hipMemcpyAsync(c, &h, sizeof(h), hipMemcpyHostToDevice, s->stream));
    DN_TRY(h2d(s, vec + kCgX * (size_t)nv, gamma, nv));
    const OccWork w = work_of(s);
    // evaluation slots in batches; the done flag is read once per batch (a solve of
    // ordinary length fits in the first)
    const int batch = 64;
    for (long evals = 0;; evals += batch) {
      DN_TRY(occ_cg_slots(s->stream, P, w, c, vec, s->dn, s->J, s->K, s->Hcore, batch));
      DN_HIP(hipMemcpyAsync(&h, c, sizeof(h), hipMemcpyDeviceToHost, s->stream));
      DN_HIP(hipStreamSynchronize(s->stream));
      if (h.done) break;
      if (evals > 64L * maxiter + 4096)
        return fail(DN_ERR_STATE, "occupation optimiser: evaluation limit reached");
    }
    DN_TRY(d2h(s, gamma_out, vec + kCgX * (size_t)nv, nv));
    DN_HIP(hipStreamSynchronize(s->stream));
    return DN_OK;
  };
  const int rc = run();
  dev_free(s->ctx, s->stream, mem);
  (void)hipStreamSynchronize(s->stream);
  if (rc != DN_OK) return rc;
  *E = h.f;
  *nit = h.it;
  *converged = h.converged;
  if (!std::isfinite(h.f) || !all_finite(gamma_out, nv))
    return fail(DN_ERR_NONFINITE, "non-finite occupation optimiser result");
  return DN_OK;
}

namespace {

// order_occupations_softmax (utils.jl:586-616), one block: when some gamma > 0, each
// pair's strong orbital and its ncwo weak orbitals are sorted by gamma (0 for the strong
// one) in the reverse of a stable ascending order, the gammas shifted by their maximum;
// perm[c] = the old column that becomes column c (identity elsewhere).  gt / idx: ndoc
// rows of ncwo + 1 (scratch).
__global__ void order_softmax_kernel(int nbf5, int no1, int ndoc, int ndns, int ncwo, int nv,
                                     const double* __restrict__ gamma,
                                     double* __restrict__ gamma_out, int* __restrict__ perm,
                                     double* __restrict__ gt, int* __restrict__ idx) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < nv; k += blockDim.x) {
    gamma_out[k] = gamma[k];
    if (gamma[k] > 0.0) atomicOr(&any, 1);
  }
  for (int c = threadIdx.x; c < nbf5; c += blockDim.x) perm[c] = c;
  __syncthreads();
  if (!any) return;
  const int m = ncwo + 1;
  for (int i = 1 + threadIdx.x; i <= ndoc; i += blockDim.x) {
    double* g = gt + (long)(i - 1) * m;
    int* id = idx + (long)(i - 1) * m;
    const int ll = no1 + ndns + (ndoc - i);  // first weak column (0-based)
    const int gx = ndoc - i;                 // first gamma of the pair
    g[0] = 0.0;
    for (int j = 0; j < ncwo; ++j) g[1 + j] = gamma[gx + ndoc * j];
    // stable ascending argsort (insertion: equal keys keep their order)
    for (int t = 0; t < m; ++t) id[t] = t;
    for (int t = 1; t < m; ++t) {
      const int v = id[t];
      int u = t;
      while (u > 0 && g[id[u - 1]] > g[v]) {
        id[u] = id[u - 1];
        --u;
      }
      id[u] = v;
    }
    double gmax = g[0];
    for (int t = 1; t < m; ++t) gmax = fmax(gmax, g[t]);
    // sort_idx = reverse; column t of the pair takes old column sort_idx[t]
    for (int t = 0; t < m; ++t) {
      const int src = id[m - 1 - t];
      const int dst_col = t == 0 ? no1 + i - 1 : ll + ndoc * (t - 1);
      const int src_col = src == 0 ? no1 + i - 1 : ll + ndoc * (src - 1);
      perm[dst_col] = src_col;
      if (t > 0) gamma_out[gx + ndoc * (t - 1)] = g[src] - gmax;
    }
  }
}

// Cout[:, c] = C[:, perm[c]] for c < nbf5, Cout[:, c] = C[:, c] beyond
__global__ void permute_cols_kernel(const double* __restrict__ C, const int* __restrict__ perm,
                                    int nbf, int nbf5, double* __restrict__ Cout) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)nbf * nbf) return;
  const int r = (int)(e % nbf), c = (int)(e / nbf);
  const int src = c < nbf5 ? perm[c] : c;
  Cout[e] = C[r + (long)nbf * src];
}

}  // namespace

int dn_order_occupations_softmax(dn_sys* s, const double* C, const double* gamma,
                                 double* C_out, double* gamma_out) {
  DN_TRY(check_sys(s));
  if (!C || !gamma || !C_out || !gamma_out) return fail(DN_ERR_ARG, "null argument");
  const dn_dims& d = s->d;
  if (d.ndoc <= 0 || d.nv <= 0) return fail(DN_ERR_ARG, "no pairs to order (ndoc = 0)");
  const size_t nn = (size_t)d.nbf * d.nbf, m = (size_t)d.ncwo + 1;
  const size_t bytes = (d.ndoc * m + d.nv) * sizeof(double) + (d.ndoc * m + d.nbf5) * sizeof(int) + 64;
  char* w = nullptr;
  DN_TRY(dev_alloc(s->ctx, s->stream, bytes, reinterpret_cast<void**>(&w)));
  double* gt = reinterpret_cast<double*>(w);
  double* gout = gt + d.ndoc * m;
  int* idx = reinterpret_cast<int*>(gout + d.nv);
  int* perm = idx + d.ndoc * m;
  auto run = [&]() -> int {
    DN_TRY(h2d(s, s->Ca, C, nn));
    DN_TRY(h2d(s, s->gamma, gamma, d.nv));
    hipLaunchKernelGGL(order_softmax_kernel, dim3(1), dim3(256), 0, s->stream, d.nbf5, d.no1,
                       d.ndoc, d.ndoc + d.nsoc, d.ncwo, d.nv, s->gamma, gout, perm, gt, idx);
    DN_HIP(hipGetLastError());
    hipLaunchKernelGGL(permute_cols_kernel, dim3((nn + 255) / 256), dim3(256), 0, s->stream,
                       s->Ca, perm, d.nbf, d.nbf5, s->bestC);
    DN_HIP(hipGetLastError());
    DN_TRY(d2h(s, C_out, s->bestC, nn));
    DN_TRY(d2h(s, gamma_out, gout, d.nv));
    DN_HIP(hipStreamSynchronize(s->stream));
    return DN_OK;
  };
  const int rc = run();
  dev_free(s->ctx, s->stream, w);
  (void)hipStreamSynchronize(s->stream);
  return rc;
}

int dn_calcoccg(dn_sys* s, const dn_functional* f, const double* gamma, const double* J,
                const double* K, const double* Hcore, double* grad) {
  DN_TRY(check_sys(s));
  OccDims P;
  DN_TRY(functional_of(s, f, &P));
  if (!grad) return fail(DN_ERR_ARG, "null grad");
  DN_TRY(h2d(s, s->gamma, gamma, s->d.nv));
  DN_TRY(upload_jkh(s, J, K, Hcore));
  DN_TRY(occupations_resident(s, P));
  DN_TRY(occ_calcoccg(s->stream, P, work_of(s), s->dn, s->J, s->K, s->Hcore, s->goc));
  DN_TRY(d2h(s, grad, s->goc, s->d.nv));
  DN_HIP(hipStreamSynchronize(s->stream));
  if (!all_finite(grad, s->d.nv)) return fail(DN_ERR_NONFINITE, "non-finite occupation gradient");
  return DN_OK;
}

int dn_rotate_orbital(dn_sys* s, const double* y, const double* C, double* Cnew) {
  DN_TRY(check_sys(s));
  if (!Cnew) return fail(DN_ERR_ARG, "null Cnew");
  const size_t nn = (size_t)s->d.nbf * s->d.nbf;
  DN_TRY(h2d(s, s->yv, y, s->nvar));
  DN_TRY(h2d(s, s->Ca, C, nn));
  DN_TRY(rotate_resident(s, s->Ca, s->bestC, true));
  DN_TRY(d2h(s, Cnew, s->bestC, nn));
  DN_HIP(hipStreamSynchronize(s->stream));
  if (!all_finite(Cnew, nn)) return fail(DN_ERR_NONFINITE, "non-finite rotated orbitals");
  return DN_OK;
}

int dn_orbopt(dn_sys* s, const dn_functional* f, dn_orbopt_state* st, const double* gamma,
              double* C, const double* H) {
  DN_TRY(check_sys(s));
  if (!C) return fail(DN_ERR_ARG, "null C");
  const size_t nn = (size_t)s->d.nbf * s->d.nbf;
  DN_TRY(h2d(s, s->gamma, gamma, s->d.nv));
  DN_TRY(h2d(s, s->C, C, nn));
  DN_TRY(h2d(s, s->H, H, nn));
  DN_TRY(orbopt_resident(s, f, st));
  DN_TRY(d2h(s, C, s->C, nn));
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

int dn_sys_set_gamma(dn_sys* s, const double* gamma) {
  DN_TRY(check_sys(s));
  DN_TRY(h2d(s, s->gamma, gamma, s->d.nv));
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

int dn_sys_get_orbitals(dn_sys* s, double* C) {
  DN_TRY(check_sys(s));
  if (!C) return fail(DN_ERR_ARG, "null C");
  DN_TRY(d2h(s, C, s->C, (size_t)s->d.nbf * s->d.nbf));
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

int dn_orbopt_resident(dn_sys* s, const dn_functional* f, dn_orbopt_state* st) {
  DN_TRY(check_sys(s));
  return orbopt_resident(s, f, st);
}

}  // extern "C"
