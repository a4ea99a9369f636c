// Internal declarations shared by the libdonofhip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "donof_hip.h"

namespace dn {

// ---- errors -------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

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
};

int gemm_f64(hipStream_t s, const Gemm& g);
// Sum `nsplit` partial matrices (each M x N, leading dim ld, stride sP) into out.
int reduce_splits(hipStream_t s, const double* P, long sP, int nsplit, int M, int N,
                  long ld, double* out, long ldo, double beta);

// ---- device buffers -----------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace dn

struct dn_ctx {
  int device = 0;
  uint64_t budget = 0;
  uint64_t used = 0;
};

struct dn_sys {
  dn_ctx* ctx = nullptr;
  dn_dims d{};
  int precision = DN_FP64;
  hipStream_t stream = nullptr;
  std::vector<dn::DevBuf> owned;
  uint64_t bytes = 0;
  // resident state (column-major)
  double* b = nullptr;      // nbf*nbf*naux
  double* C = nullptr;      // nbf*nbf
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
  double* bpn = nullptr;    // nbf5*nbf*naux : b_pn[p,n,Q] = sum_m C5[m,p] b[m,n,Q]
  double* bt = nullptr;     // nbf5*nbf*naux : b_MO^T[q,a,Q]; also b_pq for the J/K build
  double* bmat = nullptr;   // nbf5*nbf5*naux: C^K/C^J right operand of the Lagrangian GEMM
  double* X = nullptr;      // nbf5*naux     : b_pp diagonal per Q
  double* w = nullptr;      // nbf5*naux     : Coulomb + C^J weights
  double* HC = nullptr;     // nbf*nbf5      : H C5
  double* part = nullptr;   // split partials
  size_t part_elems = 0;
  double* scal = nullptr;   // small scalars (energy, norms)
  bool b_ready = false;
  // timing of the dominant transform kernel: one event pair per launch, read back
  // (and recycled) by dn_sys_get_timing so that timing adds no synchronisation.
  bool timing = false;
  std::vector<hipEvent_t> ev_start, ev_stop;
  size_t ev_used = 0;
  double transform_ms = 0.0;
  long transform_launches = 0;
};
