// C ABI of libdonofhip: contexts, resident systems, the J/K-MO build and the
// Lagrangian.  Declarations and reference citations are in include/donof_hip.h.
#include <cmath>
#include <cstring>
#include <mutex>

#include <functional>

#include "dn_internal.h"

namespace dn {

// kernels of dn_jk.hip / dn_lag.hip
int jk_from_bpq(hipStream_t s, const double* bpq, long slice, int nbf5, int naux, double* X,
                double* part, size_t part_elems, double* part2, size_t part2_elems, double* J,
                double* K);
int hcore_diag(hipStream_t s, const double* C, const double* HC, int nbf, int nbf5,
               double* Hcore);
int hcore_from_hmat(hipStream_t s, const double* Hmat, int nbf, int nbf5, double* Hcore);
int synth_b(hipStream_t s, double* b, int nbf, int ld, int naux, uint64_t seed);
int b_symmetry(hipStream_t s, const double* b, int nbf, int ld, int naux, unsigned long long* dev2,
               double* dmax, double* bmax);
int lag_w(hipStream_t s, const double* bt, long slice, int nbf, int nbf5, int naux,
          const double* cj12, const double* n, int nalpha, int nbeta, double* X, double* w);
int lag_bmat(hipStream_t s, const double* bt, long slice, int nbf, int nbf5, int naux,
             const double* ck12, const double* w, double* bmat);
int lag_w_split(hipStream_t s, const float* bpq, int nbf5, int naux, const double* cj12,
                const double* n, int nalpha, int nbeta, double* X, double* w);
int lag_init(hipStream_t s, const double* Hmat, const double* n, int nbf, int nbf5,
             double* elag);
int energy_elec(hipStream_t s, const double* elag, const double* Hmat, const double* n,
                int nbf, int nbf5, double* out);
// dn_split.hip
int split_prepare_b(hipStream_t s, const double* b, int nbf, int ld, int naux, int Kpad, float* scale,
                    float* inv_scale, _Float16* hi, _Float16* lo);
int split_k1(hipStream_t s, int nbf, int nbf5, int naux, int Mpad, int Kpad,
             const _Float16* bhi, const _Float16* blo, const float* b_inv_scale,
             const _Float16* chi, const _Float16* clo, const float* cscale_dev, float* bpn,
             float* qmax);
int split_k1_sym(hipStream_t s, int nbf, int nbf5, int naux, int Mpad, int Kpad,
                 const _Float16* bhi, const _Float16* blo, const float* b_inv_scale,
                 const _Float16* chi, const _Float16* clo, const float* cscale_dev, float* bpn,
                 float* qmax);
int split_lhnorm(hipStream_t s, const double* b, int nbf, int ld, int naux, float* lhmax);
int split_sym_planes(hipStream_t s, int nbf, int nbf5, int naux, int Mpad, int Kpad,
                     const _Float16* bhi, const _Float16* blo, const float* b_inv_scale,
                     const _Float16* chi, const _Float16* clo, const float* cscale_dev,
                     const float* qs, const float* qinv, _Float16* tp, float* wt, int part);
int split_k2(hipStream_t s, int nbf, int nbf5, int naux, int Mpad, const _Float16* chi,
             const _Float16* clo, const float* cscale_dev, const float* bpn, float* bpq,
             const float* qmax, float* qs, float* qinv);
size_t split_contract_bytes(int nbf, int nbf5, int naux);
int split_contract(hipStream_t s, int nbf, int nbf5, int naux, const float* bpn,
                   const float* bpq, const double* ck12, const double* w, const float* qmax,
                   float* zscale, void* work, size_t work_bytes, double* Z);
int jk_from_bpq_f32(hipStream_t s, const float* bpq, long slice, int nbf5, int naux, double* X,
                    double* part, size_t part_elems, double* part2, size_t part2_elems,
                    double* J, double* K);

namespace {
thread_local std::string g_last_error;
std::mutex g_ctx_mutex;

bool dims_valid(const dn_dims& d, std::string* why) {
  auto bad = [&](const char* m) { *why = m; return false; };
  if (d.nbf <= 0 || d.nbfaux <= 0 || d.nbf5 <= 0) return bad("nbf, nbfaux, nbf5 must be > 0");
  if (d.nbf5 > d.nbf) return bad("nbf5 must be <= nbf");
  if (d.nbf > 46340) return bad("nbf too large");
  if (d.no1 < 0 || d.ndoc < 0 || d.nsoc < 0 || d.ncwo < 0 || d.nv < 0)
    return bad("negative occupation sizes");
  if (d.nbeta < 0 || d.nalpha < d.nbeta || d.nalpha > d.nbf5)
    return bad("need 0 <= nbeta <= nalpha <= nbf5");
  if (d.nv != d.ncwo * d.ndoc) return bad("nv must equal ncwo*ndoc");
  if (d.no1 + d.ndoc * (1 + d.ncwo) + d.nsoc != d.nbf5)
    return bad("nbf5 must equal no1 + ndoc*(1+ncwo) + nsoc");
  return true;
}

// Element counts of every device buffer of a system.
struct Layout {
  size_t b, C, H, n, nn, nbf2, nbf_nbf5, big, bpn, cp, bmat, xw, part, scal;
  size_t nv, dnm, nvar, part2, riw;
  // split precision (0 when FP64): half counts of the b and C planes, float counts
  size_t bplane, cplane, f_big, f_bpq, f_naux;
  static size_t d8(size_t halves) { return (halves + 3) / 4; }   // halves -> doubles
  static size_t f8(size_t floats) { return (floats + 1) / 2; }   // floats -> doubles
  size_t split_doubles() const {
    if (!bplane) return 0;  // FP64 system
    // + the C5 scale pair (cscale) and the contraction's scales and block maxima (zscale)
    return 2 * d8(bplane) + 2 * d8(cplane) + f8(f_big) + f8(f_bpq) + 6 * f8(f_naux) + 2 +
           f8(4 + 8192);
  }
  size_t total_bytes() const {
    return 8 * (b + C + H + n + 4 * nn + nbf2 + nbf_nbf5 * 3 + n + big + bpn + cp + bmat + 2 * xw +
                part + scal + 2 * nv + dnm + 9 * n + 5 * nvar + 9 * nbf2 + part2 + riw +
                split_doubles());
  }
};

int round_up(int x, int m) { return (x + m - 1) / m * m; }

size_t lag_splits(const dn_dims& d) {
  // batch-reduce splits over Q for the Lagrangian GEMM: aim at ~4096 workgroups of 64 x 64
  // (its K loops are short, nbf5 per batch: more, shorter splits keep the CUs fuller;
  // 0.67 against 0.76 ms at 400/1600/134 on MI355X for 178 against 48 splits)
  const int tm = (d.nbf + 63) / 64, tn = (d.nbf5 + 63) / 64;
  long want = 4096 / std::max(1, tm * tn);
  if (want < 1) want = 1;
  if (want > d.nbfaux) want = d.nbfaux;
  const long bps = (d.nbfaux + want - 1) / want;
  return (size_t)((d.nbfaux + bps - 1) / bps);
}

Layout layout_of(const dn_dims& d, int precision) {
  Layout L{};
  const size_t nbf = d.nbf, nbf5 = d.nbf5, naux = d.nbfaux, ld = padded_ld(d.nbf);
  L.b = precision == DN_SPLIT_F16 ? 0 : ld * nbf * naux;  // split: FP16 planes only
  L.C = nbf * nbf;
  L.H = nbf * nbf;
  L.n = nbf5;
  L.nn = nbf5 * nbf5;
  L.nbf2 = nbf * nbf;
  L.nbf_nbf5 = nbf * nbf5;
  L.big = precision == DN_SPLIT_F16 ? 0 : nbf5 * nbf5 * naux;  // split: b_pq in FP32 only
  L.bpn = ld * nbf5 * naux;
  L.cp = ld != nbf ? ld * nbf : 0;
  L.bmat = precision == DN_SPLIT_F16 ? 0 : nbf5 * nbf5 * naux;  // split: never stored
  L.xw = nbf5 * naux;
  const size_t jk_part = 2 * 64 * L.nn;
  const size_t lag_part = lag_splits(d) * L.nbf_nbf5;
  L.part = std::max(jk_part, lag_part);
  L.scal = 64;
  L.nv = std::max<size_t>(d.nv, 1);
  L.dnm = std::max<size_t>(nbf5 * d.nv, 1);
  L.nvar = nbf * nbf5 - nbf5 * (nbf5 + 1) / 2;
  L.part2 = 16 * nbf * nbf;
  L.riw = naux * naux;
  if (precision == DN_SPLIT_F16) {
    L.bplane = (size_t)round_up(d.nbf, 32) * nbf * naux + 64;
    L.cplane = (size_t)round_up(d.nbf5, 128) * round_up(d.nbf, 32);
    // FP32 T (general b) or the two FP16 planes of T (symmetric b), whichever is larger
    L.f_big = std::max(nbf5 * nbf, (size_t)round_up(d.nbf5, 128) * round_up(d.nbf, 32)) * naux;
    L.f_bpq = nbf5 * nbf5 * naux;
    L.f_naux = naux;
    // the contraction's planes and FP32 partials live where the FP64 path keeps bpn
    L.bpn = std::max(L.bpn, (split_contract_bytes(d.nbf, d.nbf5, d.nbfaux) + 7) / 8);
  }
  return L;
}

int alloc(dn_sys* s, size_t elems, double** out) {
  const size_t bytes = std::max<size_t>(elems, 1) * sizeof(double);
  void* p = nullptr;
  DN_TRY(dev_alloc(s->ctx, s->stream, bytes, &p));
  s->owned.push_back(DevBuf{p, bytes});
  *out = static_cast<double*>(p);
  return DN_OK;
}

// the second stream and its events, once its queued work is done
void drop_aux(dn_sys* s) {
  if (s->aux) {
    (void)hipStreamSynchronize(s->aux);
    (void)hipStreamDestroy(s->aux);
  }
  if (s->fork_ev) (void)hipEventDestroy(s->fork_ev);
  if (s->join_ev) (void)hipEventDestroy(s->join_ev);
  s->aux = nullptr;
  s->fork_ev = s->join_ev = nullptr;
}

// every buffer back to the pool once the system's queued work is done
void release(dn_sys* s) {
  for (auto& b : s->owned) dev_free(s->ctx, s->stream, b.p);
  s->owned.clear();
  if (s->stream) (void)hipStreamSynchronize(s->stream);
}


// The transforms of b (integrals.jl:159-160, utils.jl:19-20), as GEMM descriptors shared
// by the single-system paths and the grouped builds and lock-step groups (the same
// operands and tile family give the same arithmetic):
//   t1: bpn[n,p,Q] = sum_m b[m,n,Q] C5[m,p], batched over Q (= (b_Q C5)[n,p], b_Q^T C5
//       for a general b).  With `sym` (b_Q = Lh_Q + Lh_Q^T, Lh the strict lower triangle
//       plus half the diagonal) the sum runs over m >= n with Lh: T_Q = Lh_Q^T C5, and
//       row block n0 reads only rows m >= n0 of b, half of b;
//   t2 flat: bt[q,(p,Q)] = sum_n C5[n,q] bpn[n,(p,Q)], one GEMM with (p, Q) flattened:
//       W_Q^T = (C5^T Lh_Q C5)^T after the symmetric t1 (C5^T b_Q C5 = W_Q + W_Q^T);
//   t2 batched: bt[q,a,Q] = sum_n bpn[n,q,Q] C[n,a] for a < ncols, (C5^T b_Q C)[q,a] after
//       the general t1 (slice stride nbf5 * ncols): C5^T b_Q C5 for ncols = nbf5;
// All read A and B along K, the layout of the LDS-direct tile.
}  // namespace

Gemm t1_gemm(const dn_sys* s, bool sym) {
  const dn_dims& d = s->d;
  const int P = s->ldb;  // contraction over the padded rows (the pads are zero)
  Gemm g;
  g.M = d.nbf; g.N = d.nbf5; g.K = P;
  g.A = s->b; g.lda = P; g.sA = (long)P * d.nbf; g.a_kcontig = true;
  g.B = s->Cp; g.ldb = P; g.sB = 0; g.b_kcontig = true;
  g.C = s->bpn; g.ldc = P; g.sC = (long)P * d.nbf5;
  g.nbatch = d.nbfaux; g.bps = 1;
  g.tri = sym ? kTriA : kTriNone;
  g.lds = true;
  return g;
}

Gemm t2_flat_gemm(const dn_sys* s) {
  const dn_dims& d = s->d;
  const int P = s->ldb;
  Gemm g;
  g.M = d.nbf5; g.N = d.nbf5 * d.nbfaux; g.K = P;
  g.A = s->Cp; g.lda = P; g.a_kcontig = true;
  g.B = s->bpn; g.ldb = P; g.b_kcontig = true;
  g.C = s->bt; g.ldc = d.nbf5;
  g.lds = true;
  return g;
}

Gemm t2_batched_gemm(const dn_sys* s, int ncols) {
  const dn_dims& d = s->d;
  const int P = s->ldb;
  Gemm g;
  g.M = d.nbf5; g.N = ncols; g.K = P;
  g.A = s->bpn; g.lda = P; g.sA = (long)P * d.nbf5; g.a_kcontig = true;
  g.B = s->Cp; g.ldb = P; g.sB = 0; g.b_kcontig = true;
  g.C = s->bt; g.ldc = d.nbf5; g.sC = (long)d.nbf5 * ncols;
  g.nbatch = d.nbfaux; g.bps = 1;
  g.lds = true;
  return g;
}

// The Lagrangian's two-electron terms (utils.jl:19-55) without the full transform b_MO:
// elag[a,b] += sum_{q,Q} b_MO^T[q,a,Q] bmat[q,b,Q] with b_MO^T[:,:,Q] = bpn_Q^T C, so
//   Z[n,b]     = sum_{q,Q} bpn[n,q,Q] bmat[q,b,Q]   (split partials over Q ranges),
//   elag[a,b] += sum_n C[n,a] Z[n,b];
// bmat needs b_MO only on its first nbf5 columns, the flattened second transform C5^T bpn
// (t2_flat_gemm: b_pq[q,p,Q] at p + nbf5 (q + nbf5 Q)).  The full transform (nbf columns,
// 2 nbf^2 nbf5 naux FLOP) becomes two of 2 nbf nbf5^2 naux.
Gemm lag_contract_gemm(const dn_sys* s) {
  const dn_dims& d = s->d;
  const int nsplit = (int)lag_splits(d);
  const int P = s->ldb;
  Gemm g;
  g.M = d.nbf; g.N = d.nbf5; g.K = d.nbf5;
  g.A = s->bpn; g.lda = P; g.sA = (long)P * d.nbf5; g.a_kcontig = false;
  g.B = s->bmat; g.ldb = d.nbf5; g.sB = (long)d.nbf5 * d.nbf5; g.b_kcontig = false;
  g.C = s->part; g.ldc = d.nbf; g.sC = (long)d.nbf * d.nbf5;
  g.nbatch = d.nbfaux;
  g.bps = (d.nbfaux + nsplit - 1) / nsplit;
  return g;
}

Gemm lag_final_gemm(const dn_sys* s) {
  const dn_dims& d = s->d;
  Gemm g;
  g.M = d.nbf; g.N = d.nbf5; g.K = d.nbf;
  g.A = s->C; g.lda = d.nbf; g.a_kcontig = true;
  g.B = s->Z; g.ldb = d.nbf; g.b_kcontig = true;
  g.C = s->elag; g.ldc = d.nbf;
  g.beta = 1.0;  // onto the one-electron term n[b] Hmat[a,b]
  return g;
}

int pad_c(dn_sys* s) {
  if (s->Cp == s->C) return DN_OK;
  return pad_c_launch(s->stream, s->C, s->Cp, s->d.nbf, s->ldb);
}

namespace {

int transform1(dn_sys* s, bool sym) {
  DN_TRY(begin_timing(s));
  DN_TRY(gemm_f64(s->stream, t1_gemm(s, sym)));
  return end_timing(s);
}

// HC = H C5 ; optionally Hmat = C^T HC.  The J/K build splits K for parallelism; the
// Lagrangian uses plain GEMMs, the arithmetic the grouped lock-step optimiser repeats.
int hc_products(dn_sys* s, bool want_hmat, bool splitk = true) {
  const dn_dims& d = s->d;
  Gemm g;
  g.M = d.nbf; g.N = d.nbf5; g.K = d.nbf;
  g.A = s->H; g.lda = d.nbf; g.a_kcontig = false;
  g.B = s->C; g.ldb = d.nbf; g.b_kcontig = true;
  g.C = s->HC; g.ldc = d.nbf;
  DN_TRY(splitk ? gemm_f64_splitk(s->stream, g, s->part2, s->part2_elems) : gemm_f64(s->stream, g));
  if (want_hmat) {
    Gemm h;
    h.M = d.nbf; h.N = d.nbf5; h.K = d.nbf;
    h.A = s->C; h.lda = d.nbf; h.a_kcontig = true;
    h.B = s->HC; h.ldb = d.nbf; h.b_kcontig = true;
    h.C = s->Hmat; h.ldc = d.nbf;
    DN_TRY(splitk ? gemm_f64_splitk(s->stream, h, s->part2, s->part2_elems) : gemm_f64(s->stream, h));
  }
  return DN_OK;
}

}  // namespace

size_t lag_split_count(const dn_dims& d) { return lag_splits(d); }

int check_sys(dn_sys* s) {
  if (!s) return fail(DN_ERR_ARG, "null dn_sys");
  // the calling thread may not have selected the system's device yet
  DN_HIP(hipSetDevice(s->ctx->device));
  return DN_OK;
}

int prepare_b(dn_sys* s, const double* b) {
  // b[m,n,Q] = b[n,m,Q] (to 2^-50 of max|b|) lets the J/K build read half of b
  double dmax = 0.0, bmax = 0.0;
  DN_TRY(b_symmetry(s->stream, b, s->d.nbf, s->ldb, s->d.nbfaux,
                    reinterpret_cast<unsigned long long*>(s->scal + 62), &dmax, &bmax));
  s->b_sym = dmax <= std::ldexp(bmax, -50);
  if (s->precision != DN_SPLIT_F16) return DN_OK;
  DN_TRY(split_lhnorm(s->stream, b, s->d.nbf, s->ldb, s->d.nbfaux, s->lhmax));
  return split_prepare_b(s->stream, b, s->d.nbf, s->ldb, s->d.nbfaux, s->Kpad, s->bscale, s->binv,
                         s->bhi, s->blo);
}

int set_b(dn_sys* s, const std::function<int(double*)>& fill) {
  s->b_ready = false;
  s->riw_set = false;  // a new b: its RI factor is set with it, if at all
  s->eig_clamped = false;
  if (s->b) {
    DN_TRY(fill(s->b));
    DN_TRY(prepare_b(s, s->b));
    DN_HIP(hipStreamSynchronize(s->stream));
    s->b_ready = true;
    return DN_OK;
  }
  // split precision: b exists in FP64 only while its planes are made
  const size_t elems = (size_t)s->ldb * s->d.nbf * s->d.nbfaux;
  const uint64_t bytes = elems * sizeof(double);
  DN_TRY(ctx_reserve(s->ctx, bytes));
  double* w = nullptr;
  int rc = dev_alloc(s->ctx, s->stream, bytes, reinterpret_cast<void**>(&w));
  if (rc == DN_OK) {
    auto run = [&]() -> int {
      if (s->ldb != s->d.nbf) DN_HIP(hipMemsetAsync(w, 0, bytes, s->stream));  // pad rows
      DN_TRY(fill(w));
      DN_TRY(prepare_b(s, w));
      DN_HIP(hipStreamSynchronize(s->stream));
      return DN_OK;
    };
    rc = run();
    dev_free(s->ctx, s->stream, w);
    (void)hipStreamSynchronize(s->stream);
  }
  ctx_unreserve(s->ctx, bytes);
  if (rc == DN_OK) s->b_ready = true;
  return rc;
}

int begin_timing(dn_sys* s) {
  if (!s->timing) return DN_OK;
  if (s->ev_used == s->ev_start.size()) {
    hipEvent_t a = nullptr, b = nullptr;
    DN_HIP(hipEventCreate(&a));
    DN_HIP(hipEventCreate(&b));
    s->ev_start.push_back(a);
    s->ev_stop.push_back(b);
  }
  DN_HIP(hipEventRecord(s->ev_start[s->ev_used], s->stream));
  return DN_OK;
}

int end_timing(dn_sys* s) {
  if (!s->timing) return DN_OK;
  DN_HIP(hipEventRecord(s->ev_stop[s->ev_used], s->stream));
  s->ev_used += 1;
  return DN_OK;
}

int jkh_split(dn_sys* s) {
  const dn_dims& d = s->d;
  const bool sym = s->b_sym && s->use_sym;
  DN_TRY(split_prepare_c(s->stream, s->C, d.nbf, d.nbf5, s->Mpad, s->Kpad, s->cscale, s->chi,
                         s->clo, sym ? s->lhmax : nullptr, d.nbfaux, s->qs, s->qinv));
  if (sym) {
    // H C5 and H_core need only C and H: on the second stream, beside the transforms (a
    // plain GEMM: its split-K workspace is the J GEMM's)
    DN_HIP(hipEventRecord(s->fork_ev, s->stream));
    DN_HIP(hipStreamWaitEvent(s->aux, s->fork_ev, 0));
    {
      Gemm g;
      g.M = d.nbf; g.N = d.nbf5; g.K = d.nbf;
      g.A = s->H; g.lda = d.nbf; g.a_kcontig = false;
      g.B = s->C; g.ldb = d.nbf; g.b_kcontig = true;
      g.C = s->HC; g.ldc = d.nbf;
      DN_TRY(gemm_f64(s->aux, g));
      DN_TRY(hcore_diag(s->aux, s->C, s->HC, d.nbf, d.nbf5, s->Hcore));
    }
    DN_HIP(hipEventRecord(s->join_ev, s->aux));
    // both transforms on FP16 planes; bpq32 then holds W^T with b_pq = W + W^T
    _Float16* tp = reinterpret_cast<_Float16*>(s->bpn32);
    DN_TRY(begin_timing(s));
    DN_TRY(split_sym_planes(s->stream, d.nbf, d.nbf5, d.nbfaux, s->Mpad, s->Kpad, s->bhi,
                            s->blo, s->binv, s->chi, s->clo, s->cscale, s->qs, s->qinv, tp,
                            s->bpq32, 1));
    DN_TRY(end_timing(s));
    DN_TRY(split_sym_planes(s->stream, d.nbf, d.nbf5, d.nbfaux, s->Mpad, s->Kpad, s->bhi,
                            s->blo, s->binv, s->chi, s->clo, s->cscale, s->qs, s->qinv, tp,
                            s->bpq32, 2));
    DN_TRY(symk(s->stream, reinterpret_cast<const SymKDesc*>(s->scal + 32), 1, s->symk_blocks,
                s->symk_red, true, d.nbf5 % 4 == 0));
    DN_TRY(j_from_x(s->stream, s->X, d.nbf5, d.nbfaux, s->part2, s->part2_elems, s->J));
    DN_HIP(hipStreamWaitEvent(s->stream, s->join_ev, 0));
    return DN_OK;
  }
  DN_TRY(begin_timing(s));
  DN_TRY(split_k1(s->stream, d.nbf, d.nbf5, d.nbfaux, s->Mpad, s->Kpad, s->bhi, s->blo,
                  s->binv, s->chi, s->clo, s->cscale, s->bpn32, s->qmax));
  DN_TRY(end_timing(s));
  DN_TRY(split_k2(s->stream, d.nbf, d.nbf5, d.nbfaux, s->Mpad, s->chi, s->clo, s->cscale,
                  s->bpn32, s->bpq32, s->qmax, s->qs, s->qinv));
  DN_TRY(jk_from_bpq_f32(s->stream, s->bpq32, (long)d.nbf5 * d.nbf5, d.nbf5, d.nbfaux, s->X,
                         s->part, s->part_elems, s->part2, s->part2_elems, s->J, s->K));
  DN_TRY(hc_products(s, false));
  DN_TRY(hcore_diag(s->stream, s->C, s->HC, d.nbf, d.nbf5, s->Hcore));
  return DN_OK;
}

int jkh_resident(dn_sys* s) {
  const dn_dims& d = s->d;
  if (!s->b_ready) return fail(DN_ERR_STATE, "b has not been uploaded or generated");
  if (s->precision == DN_SPLIT_F16) return jkh_split(s);
  DN_TRY(pad_c(s));
  if (s->b_sym && s->use_sym) {
    DN_TRY(transform1(s, true));
    DN_TRY(gemm_f64(s->stream, t2_flat_gemm(s)));
    DN_TRY(symk(s->stream, s->symk_dev, 1, s->symk_blocks, s->symk_red));
    DN_TRY(j_from_x(s->stream, s->X, d.nbf5, d.nbfaux, s->part2, s->part2_elems, s->J));
    DN_TRY(hc_products(s, false));
    DN_TRY(hcore_diag(s->stream, s->C, s->HC, d.nbf, d.nbf5, s->Hcore));
    return DN_OK;
  }
  DN_TRY(transform1(s, false));
  DN_TRY(gemm_f64(s->stream, t2_batched_gemm(s, d.nbf5)));
  DN_TRY(jk_from_bpq(s->stream, s->bt, (long)d.nbf5 * d.nbf5, d.nbf5, d.nbfaux, s->X, s->part,
                     s->part_elems, s->part2, s->part2_elems, s->J, s->K));
  DN_TRY(hc_products(s, false));
  DN_TRY(hcore_diag(s->stream, s->C, s->HC, d.nbf, d.nbf5, s->Hcore));
  return DN_OK;
}

// The Lagrangian's two transforms in split precision (FP16 hi+lo planes of b and C5, FP32
// accumulation): T[p,n,Q] = sum_m C5[m,p] b[m,n,Q] (the general first transform) and
// b_pq[p,q,Q] = sum_n T[p,n,Q] C5[n,q], both left in FP32 (bpn32, bpq32) for the C^J/C^K
// terms and the contraction, which read them as they are.
int lagrange_split_transforms(dn_sys* s) {
  const dn_dims& d = s->d;
  DN_TRY(split_prepare_c(s->stream, s->C, d.nbf, d.nbf5, s->Mpad, s->Kpad, s->cscale, s->chi,
                         s->clo, nullptr, d.nbfaux, s->qs, s->qinv));
  DN_TRY(split_k1(s->stream, d.nbf, d.nbf5, d.nbfaux, s->Mpad, s->Kpad, s->bhi, s->blo,
                  s->binv, s->chi, s->clo, s->cscale, s->bpn32, s->qmax));
  return split_k2(s->stream, d.nbf, d.nbf5, d.nbfaux, s->Mpad, s->chi, s->clo, s->cscale,
                  s->bpn32, s->bpq32, s->qmax, s->qs, s->qinv);
}

int lagrange_resident(dn_sys* s, int nalpha, int nbeta) {
  const dn_dims& d = s->d;
  if (!s->b_ready) return fail(DN_ERR_STATE, "b has not been uploaded or generated");
  if (nbeta < 0 || nalpha < 0 || nbeta > d.nbf5 || nalpha > d.nbf5)
    return fail(DN_ERR_ARG, "nalpha/nbeta out of range");
  const long slice = (long)d.nbf5 * d.nbf5;
  if (s->precision == DN_SPLIT_F16) {
    DN_TRY(lagrange_split_transforms(s));
  } else {
    DN_TRY(pad_c(s));
    DN_TRY(transform1(s, false));
    DN_TRY(gemm_f64(s->stream, t2_flat_gemm(s)));  // b_pq[q,p,Q] at p + nbf5 (q + nbf5 Q)
  }
  DN_TRY(hc_products(s, true, false));
  if (s->precision == DN_SPLIT_F16) {  // b_pq in FP32 (bpq32, stored transposed)
    DN_TRY(lag_w_split(s->stream, s->bpq32, d.nbf5, d.nbfaux, s->cj12, s->n, nalpha, nbeta,
                       s->X, s->w));
  } else {
    DN_TRY(lag_w(s->stream, s->bt, slice, d.nbf5, d.nbf5, d.nbfaux, s->cj12, s->n, nalpha,
                 nbeta, s->X, s->w));
    DN_TRY(lag_bmat(s->stream, s->bt, slice, d.nbf5, d.nbf5, d.nbfaux, s->ck12, s->w, s->bmat));
  }
  DN_TRY(lag_init(s->stream, s->Hmat, s->n, d.nbf, d.nbf5, s->elag));
  // Z[n,b] = sum_Q sum_q bpn[n,q,Q] bmat[q,b,Q], then elag += C^T Z
  if (s->precision == DN_SPLIT_F16) {
    DN_TRY(split_contract(s->stream, d.nbf, d.nbf5, d.nbfaux, s->bpn32, s->bpq32, s->ck12,
                          s->w, s->qmax, s->zscale, s->bpn, s->bpn_elems * sizeof(double),
                          s->Z));
  } else {
    const Gemm g = lag_contract_gemm(s);
    const int used = (g.nbatch + g.bps - 1) / g.bps;
    if ((size_t)used * d.nbf * d.nbf5 > s->part_elems)
      return fail(DN_ERR_ARG, "partials overflow");
    DN_TRY(gemm_f64(s->stream, g));
    DN_TRY(reduce_splits(s->stream, s->part, (long)d.nbf * d.nbf5, used, d.nbf, d.nbf5, d.nbf,
                         s->Z, d.nbf, 0.0));
  }
  DN_TRY(gemm_f64(s->stream, lag_final_gemm(s)));
  DN_TRY(energy_elec(s->stream, s->elag, s->Hmat, s->n, d.nbf, d.nbf5, s->scal));
  return DN_OK;
}

int h2d(dn_sys* s, double* dst, const double* src, size_t elems) {
  if (!src) return fail(DN_ERR_ARG, "null host pointer");
  DN_HIP(hipMemcpyAsync(dst, src, elems * sizeof(double), hipMemcpyHostToDevice, s->stream));
  return DN_OK;
}

int d2h(dn_sys* s, double* dst, const double* src, size_t elems) {
  if (!dst) return DN_OK;
  DN_HIP(hipMemcpyAsync(dst, src, elems * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  return DN_OK;
}

bool all_finite(const double* x, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(x[i])) return false;
  return true;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int claim_lockstep(dn_sys* const* systems, int n, void* owner) {
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  for (int i = 0; i < n; ++i) {
    if (!systems[i]) return fail(DN_ERR_ARG, "null system");
    if (systems[i]->lockstep)
      return fail(DN_ERR_STATE, "a system already belongs to a lock-step group");
    for (int j = 0; j < i; ++j)
      if (systems[j] == systems[i]) return fail(DN_ERR_ARG, "a system is listed twice");
  }
  for (int i = 0; i < n; ++i) systems[i]->lockstep = owner;
  return DN_OK;
}

void release_lockstep(dn_sys* const* systems, int n, void* owner) {
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  for (int i = 0; i < n; ++i)
    if (systems[i] && systems[i]->lockstep == owner) systems[i]->lockstep = nullptr;
}

void count_build_group(dn_sys* const* systems, int n, int delta) {
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  for (int i = 0; i < n; ++i)
    if (systems[i]) systems[i]->build_groups += delta;
}

int ctx_reserve(dn_ctx* c, uint64_t bytes) {
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  if (c->used + bytes > c->budget)
    return fail(DN_ERR_BUDGET, "needs " + std::to_string(bytes) + " bytes; budget left " +
                                   std::to_string(c->budget - c->used));
  c->used += bytes;
  return DN_OK;
}

void ctx_unreserve(dn_ctx* c, uint64_t bytes) {
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  c->used -= bytes;
}

void count_grad(dn_sys* s, int delta) {
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  s->grad_objs += delta;
}

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int dev_alloc(dn_ctx* c, hipStream_t s, size_t bytes, void** p) {
  *p = nullptr;
  hipError_t e = hipMallocFromPoolAsync(p, bytes, c->pool, s);
  if (e == hipErrorOutOfMemory) {  // hand the pool's idle blocks back and try once more
    (void)hipGetLastError();
    (void)hipStreamSynchronize(s);
    (void)hipMemPoolTrimTo(c->pool, 0);
    e = hipMallocFromPoolAsync(p, bytes, c->pool, s);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(DN_ERR_NOMEM, std::string("device allocation of ") + std::to_string(bytes) +
                                  " bytes: " + hipGetErrorString(e));
  }
  return DN_OK;
}

void dev_free(dn_ctx* c, hipStream_t s, void* p) {
  (void)c;
  if (p) (void)hipFreeAsync(p, s);
}

int pinned_get(dn_ctx* c, size_t bytes, void** p, size_t* got) {
  {
    std::lock_guard<std::mutex> lock(c->pin_mu);
    size_t best = c->pin_free.size();
    for (size_t i = 0; i < c->pin_free.size(); ++i)
      if (c->pin_free[i].second >= bytes &&
          (best == c->pin_free.size() || c->pin_free[i].second < c->pin_free[best].second))
        best = i;
    if (best < c->pin_free.size()) {
      *p = c->pin_free[best].first;
      *got = c->pin_free[best].second;
      c->pin_free.erase(c->pin_free.begin() + best);
      return DN_OK;
    }
  }
  DN_HIP(hipHostMalloc(p, bytes, hipHostMallocDefault));
  *got = bytes;
  return DN_OK;
}

void pinned_put(dn_ctx* c, void* p, size_t bytes) {
  if (!p) return;
  std::lock_guard<std::mutex> lock(c->pin_mu);
  c->pin_free.emplace_back(p, bytes);
}

}  // namespace dn

using namespace dn;

extern "C" {

const char* dn_last_error(void) { return g_last_error.c_str(); }

int dn_device_count(int* count) {
  if (!count) return fail(DN_ERR_ARG, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return DN_OK;
}

int dn_ctx_create(int device, uint64_t hbm_budget_bytes, dn_ctx** out) {
  if (!out) return fail(DN_ERR_ARG, "null out");
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(DN_ERR_HIP, "no HIP device visible");
  if (device < 0 || device >= count) return fail(DN_ERR_ARG, "device index out of range");
  DN_HIP(hipSetDevice(device));
  size_t free_b = 0, total_b = 0;
  DN_HIP(hipMemGetInfo(&free_b, &total_b));
  hipMemPoolProps props{};
  props.allocType = hipMemAllocationTypePinned;
  props.handleTypes = hipMemHandleTypeNone;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = device;
  hipMemPool_t pool = nullptr;
  DN_HIP(hipMemPoolCreate(&pool, &props));
  uint64_t keep = UINT64_MAX;  // freed blocks stay in the pool until trimmed
  const hipError_t e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  if (e != hipSuccess) {
    (void)hipMemPoolDestroy(pool);
    return fail(DN_ERR_HIP, std::string("hipMemPoolSetAttribute: ") + hipGetErrorString(e));
  }
  auto* c = new dn_ctx;
  c->device = device;
  c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)free_b;
  c->used = 0;
  c->pool = pool;
  *out = c;
  return DN_OK;
}

int dn_ctx_destroy(dn_ctx* ctx) {
  if (!ctx) return DN_OK;
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  (void)hipSetDevice(ctx->device);
  {
    std::lock_guard<std::mutex> pl(ctx->pin_mu);
    for (auto& b : ctx->pin_free) (void)hipHostFree(b.first);
    ctx->pin_free.clear();
  }
  if (ctx->pool) {
    (void)hipMemPoolTrimTo(ctx->pool, 0);
    (void)hipMemPoolDestroy(ctx->pool);
  }
  delete ctx;
  return DN_OK;
}

int dn_ctx_hbm_used(const dn_ctx* ctx, uint64_t* used) {
  if (!ctx || !used) return fail(DN_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  *used = ctx->used;
  return DN_OK;
}

int dn_ctx_hbm_budget(const dn_ctx* ctx, uint64_t* budget) {
  if (!ctx || !budget) return fail(DN_ERR_ARG, "null argument");
  *budget = ctx->budget;
  return DN_OK;
}

int dn_sys_footprint(const dn_dims* dims, int precision, uint64_t* bytes) {
  if (!dims || !bytes) return fail(DN_ERR_ARG, "null argument");
  std::string why;
  if (!dims_valid(*dims, &why)) return fail(DN_ERR_ARG, why);
  if (precision != DN_FP64 && precision != DN_SPLIT_F16)
    return fail(DN_ERR_ARG, "unknown precision");
  *bytes = layout_of(*dims, precision).total_bytes();
  return DN_OK;
}

int dn_sys_create(dn_ctx* ctx, const dn_dims* dims, int precision, dn_sys** out) {
  if (!ctx || !dims || !out) return fail(DN_ERR_ARG, "null argument");
  std::string why;
  if (!dims_valid(*dims, &why)) return fail(DN_ERR_ARG, why);
  if (precision != DN_FP64 && precision != DN_SPLIT_F16)
    return fail(DN_ERR_ARG, "unknown precision");
  const Layout L = layout_of(*dims, precision);
  const uint64_t need = L.total_bytes();
  {
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    if (ctx->used + need > ctx->budget)
      return fail(DN_ERR_BUDGET, "system needs " + std::to_string(need) +
                                     " bytes; budget left " +
                                     std::to_string(ctx->budget - ctx->used));
    ctx->used += need;
  }
  DN_HIP(hipSetDevice(ctx->device));
  auto* s = new dn_sys;
  s->ctx = ctx;
  s->d = *dims;
  s->precision = precision;
  s->bytes = need;
  int rc = DN_OK;
  {
    const hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      s->stream = nullptr;
      rc = fail(DN_ERR_HIP, hipGetErrorString(e));
    }
  }
  auto A = [&](size_t n, double** p) {
    if (rc == DN_OK) rc = alloc(s, n, p);
  };
  s->ldb = padded_ld(dims->nbf);
  if (L.b) A(L.b, &s->b);
  A(L.C, &s->C);
  if (L.cp) A(L.cp, &s->Cp);
  else s->Cp = s->C;
  A(L.H, &s->H);
  A(L.n, &s->n);
  A(L.nn, &s->cj12);
  A(L.nn, &s->ck12);
  A(L.nn, &s->J);
  A(L.nn, &s->K);
  A(L.n, &s->Hcore);
  A(L.nbf2, &s->elag);
  A(L.nbf_nbf5, &s->Hmat);
  A(L.nbf_nbf5, &s->HC);
  A(L.nbf_nbf5, &s->Z);
  A(L.bpn, &s->bpn);
  s->bpn_elems = L.bpn;
  A(L.big, &s->bt);
  A(L.bmat, &s->bmat);
  A(L.xw, &s->X);
  A(L.xw, &s->w);
  A(L.part, &s->part);
  A(L.scal, &s->scal);
  A(L.nv, &s->gamma);
  A(L.dnm, &s->dn);
  A(L.nv, &s->goc);
  A(6 * L.n, &s->orb);
  double* orbi = nullptr;
  A(3 * L.n, &orbi);  // 3*nbf5 ints fit in 3*nbf5 doubles
  s->orbi = reinterpret_cast<int*>(orbi);
  s->nvar = (long)L.nvar;
  A(L.nvar, &s->gv);
  A(L.nvar, &s->mv);
  A(L.nvar, &s->vv);
  A(L.nvar, &s->vmax);
  A(L.nvar, &s->yv);
  A(L.nbf2, &s->Y);
  A(L.nbf2, &s->U);
  A(5 * L.nbf2, &s->rwork);
  A(L.nbf2, &s->Ca);
  A(L.nbf2, &s->bestC);
  A(L.part2, &s->part2);
  s->part2_elems = L.part2;
  A(L.riw, &s->riw);
  if (precision == DN_SPLIT_F16 && rc == DN_OK) {
    if (hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->join_ev, hipEventDisableTiming) != hipSuccess)
      rc = fail(DN_ERR_HIP, "second stream of a split-precision system");
  }
  if (precision == DN_SPLIT_F16) {
    s->Kpad = round_up(dims->nbf, 32);
    s->Mpad = round_up(dims->nbf5, 128);
    double* q = nullptr;
    A(Layout::d8(L.bplane), &q); s->bhi = reinterpret_cast<_Float16*>(q);
    A(Layout::d8(L.bplane), &q); s->blo = reinterpret_cast<_Float16*>(q);
    A(Layout::d8(L.cplane), &q); s->chi = reinterpret_cast<_Float16*>(q);
    A(Layout::d8(L.cplane), &q); s->clo = reinterpret_cast<_Float16*>(q);
    A(Layout::f8(L.f_big), &q); s->bpn32 = reinterpret_cast<float*>(q);
    A(Layout::f8(L.f_bpq), &q); s->bpq32 = reinterpret_cast<float*>(q);
    A(Layout::f8(L.f_naux), &q); s->bscale = reinterpret_cast<float*>(q);
    A(Layout::f8(L.f_naux), &q); s->binv = reinterpret_cast<float*>(q);
    A(Layout::f8(L.f_naux), &q); s->qmax = reinterpret_cast<float*>(q);
    A(Layout::f8(L.f_naux), &q); s->qs = reinterpret_cast<float*>(q);
    A(Layout::f8(L.f_naux), &q); s->qinv = reinterpret_cast<float*>(q);
    A(Layout::f8(L.f_naux), &q); s->lhmax = reinterpret_cast<float*>(q);
    A(2, &q); s->cscale = reinterpret_cast<float*>(q);  // scale, 1/scale, max bits
    A(Layout::f8(4 + 8192), &q); s->zscale = reinterpret_cast<float*>(q);
    if (rc == DN_OK) {
      // plane padding must read as zero
      hipError_t e1 = hipMemsetAsync(s->bhi, 0, Layout::d8(L.bplane) * 8, s->stream);
      hipError_t e2 = hipMemsetAsync(s->blo, 0, Layout::d8(L.bplane) * 8, s->stream);
      if (e1 != hipSuccess || e2 != hipSuccess) rc = fail(DN_ERR_HIP, "hipMemset of b planes");
    }
  }
  s->part_elems = L.part;
  if (rc == DN_OK && s->ldb != dims->nbf) {  // the pad rows of b, bpn and Cp read as zero
    hipError_t e1 = s->b ? hipMemsetAsync(s->b, 0, L.b * sizeof(double), s->stream)
                         : hipSuccess;
    hipError_t e2 = hipMemsetAsync(s->bpn, 0, L.bpn * sizeof(double), s->stream);
    hipError_t e3 = hipMemsetAsync(s->Cp, 0, L.cp * sizeof(double), s->stream);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
      rc = fail(DN_ERR_HIP, "hipMemset of the padded buffers");
  }
  if (rc == DN_OK) {  // descriptor of the symmetric build's K/X pass, kept in scal[16..31]
    static_assert(sizeof(SymKDesc) <= 16 * sizeof(double), "SymKDesc fits scal[16..31]");
    SymKDesc k;
    k.Wf = nullptr;
    k.W = s->bt;
    k.K = s->K;
    k.X = s->X;
    k.part = s->part;
    k.part_elems = s->part_elems;
    k.nbf5 = dims->nbf5;
    k.naux = dims->nbfaux;
    symk_plan(&k, 1, &s->symk_blocks, &s->symk_red, s->precision == DN_SPLIT_F16);
    s->symk_dev = reinterpret_cast<SymKDesc*>(s->scal + 16);
    SymKDesc k32 = k;  // the split-precision build's W (FP32), at scal[32..47]
    k32.W = nullptr;
    k32.Wf = s->bpq32;
    static_assert(sizeof(SymKDesc) % sizeof(double) == 0 && sizeof(SymKDesc) <= 128, "layout");
    if (hipMemcpyAsync(s->symk_dev, &k, sizeof(k), hipMemcpyHostToDevice, s->stream) !=
            hipSuccess ||
        hipMemcpyAsync(s->scal + 32, &k32, sizeof(k32), hipMemcpyHostToDevice, s->stream) !=
            hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)  // k, k32 live on this frame
      rc = fail(DN_ERR_HIP, "descriptor upload");
  }
  if (rc != DN_OK) {
    std::string msg = g_last_error;
    drop_aux(s);
    release(s);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    {
      std::lock_guard<std::mutex> lock(g_ctx_mutex);
      ctx->used -= need;
    }
    delete s;
    return fail(rc, msg);
  }
  *out = s;
  return DN_OK;
}

int dn_sys_destroy(dn_sys* s) {
  if (!s) return DN_OK;
  {
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    if (s->lockstep || s->build_groups > 0)
      return fail(DN_ERR_STATE, "the system belongs to a group; destroy the group first");
    if (s->grad_objs > 0)
      return fail(DN_ERR_STATE, "a gradient object holds the system; destroy it first");
  }
  (void)hipSetDevice(s->ctx->device);
  drop_aux(s);
  release(s);  // ordered after the system's queued work
  if (s->stream) (void)hipStreamDestroy(s->stream);
  for (auto ev : s->ev_start) (void)hipEventDestroy(ev);
  for (auto ev : s->ev_stop) (void)hipEventDestroy(ev);
  {
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    s->ctx->used -= s->bytes;
  }
  delete s;
  return DN_OK;
}

int dn_sys_upload_b(dn_sys* s, const double* b) {
  DN_TRY(check_sys(s));
  if (!b) return fail(DN_ERR_ARG, "null host pointer");
  const size_t nbf = s->d.nbf;
  // into rows of ldb doubles (the pad rows keep their zeros)
  return set_b(s, [&](double* dst) -> int {
    DN_HIP(hipMemcpy2DAsync(dst, s->ldb * sizeof(double), b, nbf * sizeof(double),
                            nbf * sizeof(double), nbf * s->d.nbfaux, hipMemcpyHostToDevice,
                            s->stream));
    return DN_OK;
  });
}

int dn_sys_download_b(dn_sys* s, double* b) {
  DN_TRY(check_sys(s));
  if (!b) return fail(DN_ERR_ARG, "null host pointer");
  if (!s->b)
    return fail(DN_ERR_STATE, "a split-precision system keeps b only as FP16 hi/lo planes");
  const size_t nbf = s->d.nbf;
  DN_HIP(hipMemcpy2DAsync(b, nbf * sizeof(double), s->b, s->ldb * sizeof(double),
                          nbf * sizeof(double), nbf * s->d.nbfaux, hipMemcpyDeviceToHost,
                          s->stream));
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

int dn_sys_generate_b(dn_sys* s, uint64_t seed) {
  DN_TRY(check_sys(s));
  return set_b(s, [&](double* dst) {
    return synth_b(s->stream, dst, s->d.nbf, s->ldb, s->d.nbfaux, seed);
  });
}

int dn_sys_set_orbitals(dn_sys* s, const double* C) {
  DN_TRY(check_sys(s));
  return h2d(s, s->C, C, (size_t)s->d.nbf * s->d.nbf);
}

int dn_sys_set_hcore(dn_sys* s, const double* H) {
  DN_TRY(check_sys(s));
  return h2d(s, s->H, H, (size_t)s->d.nbf * s->d.nbf);
}

int dn_jkh_mo_resident(dn_sys* s) {
  DN_TRY(check_sys(s));
  return jkh_resident(s);
}

int dn_sys_get_jkh(dn_sys* s, double* J, double* K, double* Hcore) {
  DN_TRY(check_sys(s));
  const size_t nn = (size_t)s->d.nbf5 * s->d.nbf5;
  DN_TRY(d2h(s, J, s->J, nn));
  DN_TRY(d2h(s, K, s->K, nn));
  DN_TRY(d2h(s, Hcore, s->Hcore, s->d.nbf5));
  DN_HIP(hipStreamSynchronize(s->stream));
  if ((J && !all_finite(J, nn)) || (K && !all_finite(K, nn)) ||
      (Hcore && !all_finite(Hcore, s->d.nbf5)))
    return fail(DN_ERR_NONFINITE, "non-finite value in J/K/H_core");
  return DN_OK;
}

int dn_sys_set_symmetry(dn_sys* s, int use) {
  DN_TRY(check_sys(s));
  s->use_sym = use != 0;
  return DN_OK;
}

int dn_sys_b_symmetric(dn_sys* s, int* is_symmetric) {
  DN_TRY(check_sys(s));
  if (!is_symmetric) return fail(DN_ERR_ARG, "null output pointer");
  *is_symmetric = s->b_sym ? 1 : 0;
  return DN_OK;
}

int dn_sys_sync(dn_sys* s) {
  DN_TRY(check_sys(s));
  DN_HIP(hipStreamSynchronize(s->stream));
  return DN_OK;
}

int dn_jkh_mo(dn_sys* s, const double* C, const double* H, double* J, double* K,
              double* Hcore) {
  DN_TRY(check_sys(s));
  DN_TRY(dn_sys_set_orbitals(s, C));
  DN_TRY(dn_sys_set_hcore(s, H));
  DN_TRY(jkh_resident(s));
  return dn_sys_get_jkh(s, J, K, Hcore);
}

int dn_lagrange2(dn_sys* s, const double* C, const double* n, const double* H,
                 const double* cj12, const double* ck12, int nalpha, int nbeta, double* elag,
                 double* Hmat, double* E) {
  DN_TRY(check_sys(s));
  const dn_dims& d = s->d;
  const size_t nn = (size_t)d.nbf5 * d.nbf5;
  DN_TRY(h2d(s, s->C, C, (size_t)d.nbf * d.nbf));
  DN_TRY(h2d(s, s->H, H, (size_t)d.nbf * d.nbf));
  DN_TRY(h2d(s, s->n, n, d.nbf5));
  DN_TRY(h2d(s, s->cj12, cj12, nn));
  DN_TRY(h2d(s, s->ck12, ck12, nn));
  DN_TRY(lagrange_resident(s, nalpha, nbeta));
  DN_TRY(d2h(s, elag, s->elag, (size_t)d.nbf * d.nbf));
  DN_TRY(d2h(s, Hmat, s->Hmat, (size_t)d.nbf * d.nbf5));
  double e = 0.0;
  DN_TRY(d2h(s, &e, s->scal, 1));
  DN_HIP(hipStreamSynchronize(s->stream));
  if (E) *E = e;
  if (!std::isfinite(e) || (elag && !all_finite(elag, (size_t)d.nbf * d.nbf)))
    return fail(DN_ERR_NONFINITE, "non-finite value in the Lagrangian");
  return DN_OK;
}

int dn_lagrange2_jkh(dn_sys* s, const double* C, const double* n, const double* H,
                     const double* cj12, const double* ck12, int nalpha, int nbeta,
                     double* elag, double* Hmat, double* E, double* J, double* K,
                     double* Hcore) {
  DN_TRY(check_sys(s));
  if (!J || !K || !Hcore) return fail(DN_ERR_ARG, "null J/K/Hcore output");
  const dn_dims& d = s->d;
  const size_t nn = (size_t)d.nbf5 * d.nbf5;
  DN_TRY(h2d(s, s->C, C, (size_t)d.nbf * d.nbf));
  DN_TRY(h2d(s, s->H, H, (size_t)d.nbf * d.nbf));
  DN_TRY(h2d(s, s->n, n, d.nbf5));
  DN_TRY(h2d(s, s->cj12, cj12, nn));
  DN_TRY(h2d(s, s->ck12, ck12, nn));
  DN_TRY(lagrange_resident(s, nalpha, nbeta));
  // b_pq[q,p,Q] of the Lagrangian's second transform gives K and X = diag(b_pq), and
  // J = X X^T; H_core is the diagonal of Hmat's leading nbf5 block (C5^T H C5)
  if (s->precision == DN_SPLIT_F16)  // FP32 b_pq (K comes out as the transpose: b_pq = b_qp)
    DN_TRY(jk_from_bpq_f32(s->stream, s->bpq32, (long)nn, d.nbf5, d.nbfaux, s->X, s->part,
                           s->part_elems, s->part2, s->part2_elems, s->J, s->K));
  else
    DN_TRY(jk_from_bpq(s->stream, s->bt, (long)nn, d.nbf5, d.nbfaux, s->X, s->part,
                       s->part_elems, s->part2, s->part2_elems, s->J, s->K));
  DN_TRY(hcore_from_hmat(s->stream, s->Hmat, d.nbf, d.nbf5, s->Hcore));
  std::vector<double> hm;
  double* hm_out = Hmat;
  if (!hm_out) {
    hm.resize((size_t)d.nbf * d.nbf5);
    hm_out = hm.data();
  }
  DN_TRY(d2h(s, elag, s->elag, (size_t)d.nbf * d.nbf));
  DN_TRY(d2h(s, hm_out, s->Hmat, (size_t)d.nbf * d.nbf5));
  DN_TRY(d2h(s, J, s->J, nn));
  DN_TRY(d2h(s, K, s->K, nn));
  DN_TRY(d2h(s, Hcore, s->Hcore, d.nbf5));
  double e = 0.0;
  DN_TRY(d2h(s, &e, s->scal, 1));
  DN_HIP(hipStreamSynchronize(s->stream));
  if (E) *E = e;
  if (!std::isfinite(e) || (elag && !all_finite(elag, (size_t)d.nbf * d.nbf)) ||
      !all_finite(J, nn) || !all_finite(K, nn) || !all_finite(Hcore, d.nbf5))
    return fail(DN_ERR_NONFINITE, "non-finite value in the Lagrangian or J/K/H_core");
  return DN_OK;
}

int dn_sys_set_timing(dn_sys* s, int enable) {
  DN_TRY(check_sys(s));
  DN_HIP(hipStreamSynchronize(s->stream));
  s->timing = enable != 0;
  s->ev_used = 0;
  s->transform_ms = 0.0;
  s->transform_launches = 0;
  return DN_OK;
}

int dn_sys_get_timing(dn_sys* s, double* ms_total, long* launches) {
  DN_TRY(check_sys(s));
  DN_HIP(hipStreamSynchronize(s->stream));
  for (size_t i = 0; i < s->ev_used; ++i) {
    float ms = 0.f;
    DN_HIP(hipEventElapsedTime(&ms, s->ev_start[i], s->ev_stop[i]));
    s->transform_ms += ms;
    s->transform_launches += 1;
  }
  s->ev_used = 0;
  if (ms_total) *ms_total = s->transform_ms;
  if (launches) *launches = s->transform_launches;
  return DN_OK;
}

}  // extern "C"
