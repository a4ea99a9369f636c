// The J/K-MO build (integrals.jl:144-172) for callers that cannot allocate device memory:
// every input and output is a device pointer the caller owns, and the intermediates live
// in one file-scope device array of DN_NOALLOC_SCRATCH_BYTES per GPU.
//
// The kernels are the general-b path of the resident build (dn_api.hip: the batched first
// transform bpn_Q = b_Q^T C5, the batched second transform b_pq,Q = bpn_Q^T C5, K and X
// from b_pq, J = X X^T, H C5 and the H_core diagonal); b is read in the ABI layout
// (leading dimension nbf, no pad row), so for even nbf the GEMM launches are those of
// dn_jkh_mo with the symmetric path switched off, and for odd nbf the GEMMs take the
// register-staged tiles (rows of odd length are not 16-byte aligned).  The K reduction
// may split Q into other ranges than the resident build's (a smaller partials buffer), so
// results agree with dn_jkh_mo to rounding, not bit for bit.
//
// Nothing in the scratch array has to start at zero: every kernel writes the elements it
// later reads (GEMMs with beta = 0, the K partials, X, the split-K partials).  Calls that
// overlap on one GPU share that GPU's array and race; the caller orders them (one stream,
// or its own events).
#include <mutex>

#include "dn_internal.h"

namespace dn {
int jk_from_bpq(hipStream_t s, const double* bpq, long slice, int nbf5, int naux, double* X,
                double* part, size_t part_elems, double* part2, size_t part2_elems, double* J,
                double* K);
int hcore_diag(hipStream_t s, const double* C, const double* HC, int nbf, int nbf5,
               double* Hcore);
}  // namespace dn

namespace {

constexpr size_t kScratchDoubles = (size_t)DN_NOALLOC_SCRATCH_BYTES / sizeof(double);
__device__ __align__(256) double scratch[kScratchDoubles];

// The device address of `scratch` on each GPU, taken once per device with that device
// current.  The host-side symbol `scratch` is not a device address and is never passed to
// a kernel or a copy.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_of(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return dn::fail(DN_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(e));
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return DN_OK;
}

size_t round256(size_t doubles) { return (doubles + 31) / 32 * 32; }  // 256-byte blocks

// Offsets (doubles) of the intermediates inside the scratch array.
struct NoAllocLayout {
  size_t bpn, bt, X, HC, part, part2, total;
  size_t part_elems, part2_elems;
};

NoAllocLayout noalloc_layout(int nbf, int naux, int nbf5) {
  NoAllocLayout L{};
  const size_t n = nbf, a = naux, p = nbf5, nn = p * p;
  L.part_elems = 2 * 64 * nn;  // the resident build's K partials and split-K partials
  L.part2_elems = 16 * n * n;
  size_t off = 0;
  L.bpn = off;   off += round256(n * p * a);  // bpn[n,p,Q] = (b_Q^T C5)[n,p]
  L.bt = off;    off += round256(nn * a);     // b_pq[p,q,Q]
  L.X = off;     off += round256(p * a);      // b_ppQ
  L.HC = off;    off += round256(n * p);      // H C5
  L.part = off;  off += round256(L.part_elems);
  L.part2 = off; off += round256(L.part2_elems);
  L.total = off;
  return L;
}

}  // namespace

using namespace dn;

extern "C" {

int dn_jkh_mo_noalloc(int device, int nbf, int nbfaux, int nbf5, const double* b,
                      const double* C, const double* H, double* J, double* K, double* Hcore,
                      void* stream) {
  if (!b || !C || !H || !J || !K || !Hcore) return fail(DN_ERR_ARG, "null device pointer");
  if (nbf <= 0 || nbfaux <= 0 || nbf5 <= 0 || nbf5 > nbf || nbf > 46340)
    return fail(DN_ERR_ARG, "need 0 < nbf5 <= nbf <= 46340 and nbfaux > 0");
  const uint64_t need = (uint64_t)noalloc_layout(nbf, nbfaux, nbf5).total * sizeof(double);
  if (need > (uint64_t)DN_NOALLOC_SCRATCH_BYTES)
    return fail(DN_ERR_ARG, "nbf=" + std::to_string(nbf) + " nbfaux=" + std::to_string(nbfaux) +
                                " nbf5=" + std::to_string(nbf5) + " needs " +
                                std::to_string(need) + " bytes of scratch; the no-alloc build "
                                "holds at most " +
                                std::to_string((uint64_t)DN_NOALLOC_SCRATCH_BYTES));
  int count = 0;
  DN_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count || device >= kMaxDevices)
    return fail(DN_ERR_ARG, "device index out of range");
  DN_HIP(hipSetDevice(device));
  double* base = nullptr;
  DN_TRY(scratch_of(device, &base));
  const NoAllocLayout L = noalloc_layout(nbf, nbfaux, nbf5);
  double* bpn = base + L.bpn;
  double* bt = base + L.bt;
  double* X = base + L.X;
  double* HC = base + L.HC;
  double* part = base + L.part;
  double* part2 = base + L.part2;
  hipStream_t s = static_cast<hipStream_t>(stream);

  // bpn[n,p,Q] = sum_m b[m,n,Q] C5[m,p]  (integrals.jl:159)
  Gemm t1;
  t1.M = nbf; t1.N = nbf5; t1.K = nbf;
  t1.A = b; t1.lda = nbf; t1.sA = (long)nbf * nbf; t1.a_kcontig = true;
  t1.B = C; t1.ldb = nbf; t1.sB = 0; t1.b_kcontig = true;
  t1.C = bpn; t1.ldc = nbf; t1.sC = (long)nbf * nbf5;
  t1.nbatch = nbfaux; t1.bps = 1;
  t1.lds = true;  // taken only where rows are 16-byte aligned (even nbf)
  DN_TRY(gemm_f64(s, t1));
  // b_pq[p,q,Q] = sum_n bpn[n,p,Q] C5[n,q]  (integrals.jl:160)
  Gemm t2;
  t2.M = nbf5; t2.N = nbf5; t2.K = nbf;
  t2.A = bpn; t2.lda = nbf; t2.sA = (long)nbf * nbf5; t2.a_kcontig = true;
  t2.B = C; t2.ldb = nbf; t2.sB = 0; t2.b_kcontig = true;
  t2.C = bt; t2.ldc = nbf5; t2.sC = (long)nbf5 * nbf5;
  t2.nbatch = nbfaux; t2.bps = 1;
  t2.lds = true;
  DN_TRY(gemm_f64(s, t2));
  // K, X = diag(b_pq), J = X X^T  (integrals.jl:163,166)
  DN_TRY(jk_from_bpq(s, bt, (long)nbf5 * nbf5, nbf5, nbfaux, X, part, L.part_elems, part2,
                     L.part2_elems, J, K));
  // H_core_p = (C5^T H C5)_pp  (integrals.jl:156,169)
  Gemm hc;
  hc.M = nbf; hc.N = nbf5; hc.K = nbf;
  hc.A = H; hc.lda = nbf; hc.a_kcontig = false;
  hc.B = C; hc.ldb = nbf; hc.b_kcontig = true;
  hc.C = HC; hc.ldc = nbf;
  DN_TRY(gemm_f64_splitk(s, hc, part2, L.part2_elems));
  return hcore_diag(s, C, HC, nbf, nbf5, Hcore);
}

}  // extern "C"
