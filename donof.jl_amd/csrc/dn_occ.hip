// Occupation functional on the device: occupations and their gamma-Jacobian
// (pnof.jl:1024-1163), the PNOF4/5/7/8(GNOF)/9 coefficient matrices cj12/ck12
// (pnof.jl:1-53, 119-1022), the occupation energy calce (pnof.jl:1273-1300) and the
// occupation gradient calcoccg (pnof.jl:1324-1377).
//
// The reference materialises the derivative arrays Dcj12r/Dck12r (nbf5 x nbf5 x nv) and
// contracts them with J and K.  Here every element Dc[i,j,k] is rebuilt from per-orbital
// quantities while the contraction runs: one block per gamma index k, coalesced reads of
// J and K, a block reduction per k.  Element formulas keep the reference's block
// structure (core / beta / singly occupied / weak "cwo" orbitals, intra-pair overrides,
// zero diagonal) so each branch cites the reference lines it restates.
#include <cmath>

#include "dn_internal.h"
#include "dn_occ.h"

namespace dn {

namespace {

enum Role { kCore = 0, kBeta = 1, kSoc = 2, kCwo = 3 };

// Per-orbital quantities shared by the coefficient and derivative kernels.
struct Orb {
  const double* n;     // occupations
  const double* fi;    // static factor fi (variant dependent)
  const double* nd;    // dynamic occupation n_d (GNOF, PNOF9)
  const double* nd12;  // sqrt(n_d)
  const double* dfic;  // dfi/dn  (dfi[i,k] = dfic[i] dn[i,k])
  const double* ndc;   // dn_d[i,k] = ndc[i] dn[i,k] + ndx[i] dn[strong[i],k]
  const double* ndx;
  const int* strong;   // strong orbital of i's pair (or -1)
  const int* pair;     // pair id (1-based as the reference's l) or 0
  const int* role;
};

__device__ __forceinline__ bool in_pair_block(const OccDims& P, const Orb& o, int i, int j,
                                              int* kind) {
  // kind 1: strong(i)-weak(j), 2: weak(i)-strong(j), 3: weak-weak, same pair
  const int li = o.pair[i], lj = o.pair[j];
  if (li == 0 || li != lj || i == j) return false;
  const bool si = i < P.no1 + P.ndoc, sj = j < P.no1 + P.ndoc;
  *kind = si ? 1 : (sj ? 2 : 3);
  return true;
}

// cj12[i,j], ck12[i,j] (before the diagonal is zeroed)
__device__ void coef_elem(const OccDims& P, const Orb& o, double S_F, int i, int j, double* cj,
                          double* ck) {
  const double ni = o.n[i], nj = o.n[j];
  const int ri = o.role[i], rj = o.role[j];
  if (P.ipnof == 4) {  // pnof.jl:119-173 (occupied block = first ndoc orbitals)
    const bool oi = i < P.ndoc, oj = j < P.ndoc;
    const double hi = 1.0 - ni, hj = 1.0 - nj;
    const double c = (1 - S_F) / S_F;
    double D, Pi;
    if (oi && oj) {
      D = hi * hj;
      Pi = -sqrt(hi * hj);
    } else if (oi) {
      const double R = (hi * nj) / S_F;
      D = c * (hi * nj);
      Pi = -sqrt((hi * nj) / S_F) * sqrt((ni - nj) + R);
    } else if (oj) {
      const double Rt = (hj * ni) / S_F;
      D = c * (ni * hj);
      Pi = -sqrt((ni * hj) / S_F) * sqrt(-(ni - nj) + Rt);
    } else {
      D = ni * nj;
      Pi = sqrt(ni * nj);
    }
    const double nn = ni * nj;
    *cj = 2.0 * (nn - D);
    *ck = nn - D - Pi;
    return;
  }
  double cjv = 2.0 * ni * nj;
  double ckv = ni * nj;
  const bool beta_cwo = (ri == kBeta && rj == kCwo), cwo_beta = (ri == kCwo && rj == kBeta);
  const bool cwo_cwo = (ri == kCwo && rj == kCwo);
  if (P.ipnof == 5) {  // pnof.jl:274-310
    if (P.mspin == 0 && P.nsoc > 1 && ri == kSoc && rj == kSoc) ckv = 2.0 * ni * nj;
  } else if (P.ipnof == 7) {  // pnof.jl:356-411
    ckv = ni * nj + o.fi[i] * o.fi[j];
    if (P.nsoc > 1 && ri == kSoc && rj == kSoc) ckv = 2.0 * ni * nj;
  } else if (P.ipnof == 8 || P.ipnof == 9) {
    if (P.ipnof == 8 && P.ista == 0) {  // pnof.jl:543-562
      if (beta_cwo || cwo_beta || cwo_cwo) ckv += o.fi[i] * o.fi[j];
      if (P.nsoc > 0) {
        if (ri == kBeta && rj == kSoc) ckv += 0.25 * o.fi[i];
        if (ri == kSoc && rj == kBeta) ckv += 0.25 * o.fi[j];
        if (ri == kSoc && rj == kCwo) ckv += 0.5 * o.fi[j];
        if (ri == kCwo && rj == kSoc) ckv += 0.5 * o.fi[i];
      }
      if (P.nsoc > 1 && ri == kSoc && rj == kSoc) ckv = 0.5;
    } else {  // pnof.jl:563-568 (ista > 0), 827 and 846-848 (PNOF9)
      ckv = ni * nj + o.fi[i] * o.fi[j];
      if (P.ipnof == 9 && P.nsoc > 1 && ri == kSoc && rj == kSoc) ckv = 0.5;
    }
    // dynamic terms (pnof.jl:572-578)
    if (beta_cwo || cwo_beta) ckv += o.nd12[i] * o.nd12[j] - o.nd[i] * o.nd[j];
    if (cwo_cwo) ckv += -o.nd12[i] * o.nd12[j] - o.nd[i] * o.nd[j];
  }
  int kind;
  if (in_pair_block(P, o, i, j, &kind)) {  // intra-pair overrides (e.g. pnof.jl:580-597)
    cjv = 0.0;
    ckv = (kind == 3) ? -sqrt(ni * nj) : sqrt(ni * nj);
  }
  *cj = cjv;
  *ck = ckv;
}

// Row-derivative coefficients Dcj12r[i,j,k], Dck12r[i,j,k] (before the diagonal is
// zeroed).  dni = dn[i,k], dsi = dn[strong(i),k], dnj = dn[j,k] (PNOF4 only),
// DSF = dS_F/dgamma_k (PNOF4 only).
__device__ void der_elem(const OccDims& P, const Orb& o, double S_F, double DSF, int i, int j,
                         double dni, double dsi, double dnj, double* dcj, double* dck) {
  const double ni = o.n[i], nj = o.n[j];
  const int ri = o.role[i], rj = o.role[j];
  if (P.ipnof == 4) {  // pnof.jl:175-272 (S_F clamped at 1e-7 as there)
    const bool oi = i < P.ndoc, oj = j < P.ndoc;
    const double hi = 1.0 - ni, hj = 1.0 - nj;
    const double dhi = -dni, dhj = -dnj;
    const double SF = S_F;
    const double T = (1 - SF) / SF;
    const double DT = -DSF / (SF * SF);
    double DD, DPi = 0.0;
    if (oi && oj) {
      DD = dhi * hj;
      DPi = -(1.0 / fmax(2 * sqrt(hi), 1e-7) * dhi) * sqrt(hj);
    } else if (oi) {
      const double R = (hi * nj) / SF;
      const double DR = ((dhi * nj + hi * dnj) * SF - (hi * nj) * DSF) / (SF * SF);
      DD = DT * (hi * nj) + T * (dhi * nj);
      DPi = -(1.0 / fmax(2 * sqrt(R), 1e-7) * DR * sqrt((ni - nj) + R) +
              sqrt(R) * ((dni - dnj) + DR) / fmax(2 * sqrt((ni * nj) + R), 1e-7));
    } else if (oj) {
      DD = T * (dni * hj);
    } else {
      DD = dni * nj;
      DPi = (1.0 / fmax(2 * sqrt(ni), 1e-7) * dni) * sqrt(nj);
    }
    *dcj = 2.0 * ((dni * nj) - DD);
    *dck = (dni * nj) - DD - DPi;
    return;
  }
  double dcjv = 2.0 * dni * nj;
  double dckv = dni * nj;
  const bool beta_cwo = (ri == kBeta && rj == kCwo), cwo_beta = (ri == kCwo && rj == kBeta);
  const bool cwo_cwo = (ri == kCwo && rj == kCwo);
  const double dfi = o.dfic[i] * dni;
  if (P.ipnof == 7) {  // pnof.jl:448-449
    dckv = dni * nj + dfi * o.fi[j];
  } else if (P.ipnof == 8 || P.ipnof == 9) {
    if (P.ipnof == 8 && P.ista == 0) {  // pnof.jl:709-731
      if (beta_cwo) dckv += dfi * o.fi[j];
      if (cwo_beta) dckv += o.fi[j] * dfi;
      if (cwo_cwo) dckv += o.fi[j] * dfi;
      if (P.nsoc > 0) {
        if (ri == kBeta && rj == kSoc) dckv += 0.5 * dfi * 0.5;
        if (ri == kCwo && rj == kSoc) dckv += dfi * 0.5;
      }
      if (P.nsoc > 1 && ri == kSoc && rj == kSoc) dckv = 0.0;
    } else {  // pnof.jl:732-740 and 947, 967-969
      dckv = dni * nj + dfi * o.fi[j];
      if (P.nsoc > 1 && ri == kSoc && rj == kSoc) dckv = 0.0;
    }
    if (beta_cwo || cwo_beta || cwo_cwo) {  // pnof.jl:750-755
      const double dnd = o.ndc[i] * dni + o.ndx[i] * dsi;
      const double dnd12 = 0.5 * dnd / fmax(o.nd12[i], 1e-15);
      if (beta_cwo) dckv += dnd12 * o.nd12[j] - dnd * o.nd[j];
      else if (cwo_beta) dckv += o.nd12[j] * dnd12 - o.nd[j] * dnd;
      else dckv += -o.nd12[j] * dnd12 - o.nd[j] * dnd;
    }
  }
  int kind;
  if (in_pair_block(P, o, i, j, &kind)) {  // pnof.jl:759-789
    dcjv = 0.0;
    const double a = fmax(ni, 1e-15);
    const double v = 1.0 / 2.0 * 1.0 / sqrt(a) * dni * sqrt(nj);
    dckv = (kind == 3) ? -v : v;
  }
  *dcj = dcjv;
  *dck = dckv;
}

__device__ __forceinline__ Orb orb_view(const OccWork& w) {
  Orb o;
  o.n = w.n; o.fi = w.fi; o.nd = w.nd; o.nd12 = w.nd12; o.dfic = w.dfic;
  o.ndc = w.ndc; o.ndx = w.ndx; o.strong = w.strong; o.pair = w.pair; o.role = w.role;
  return o;
}

__device__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) s += red[k];
  return s;  // valid in thread 0
}

}  // namespace

// ---- occupations ------------------------------------------------------------------------

// pnof.jl:1122-1163 (Softmax) and 1047-1120 (Trigonometric); one block.
__device__ void ocupacion_block(const OccDims& P, const double* __restrict__ gamma,
                                double* __restrict__ n, double* __restrict__ dn) {
  const int nbf5 = P.nbf5, nv = P.nv, ndoc = P.ndoc, ncwo = P.ncwo, no1 = P.no1;
  for (long e = threadIdx.x; e < (long)nbf5 * nv; e += blockDim.x) dn[e] = 0.0;
  for (int i = threadIdx.x; i < nbf5; i += blockDim.x) {
    if (i < no1) n[i] = 1.0;
    else if (i >= no1 + ndoc && i < no1 + P.ndns) n[i] = P.highspin ? 1.0 : 0.5;
  }
  __syncthreads();
  if (P.occ_method == DN_OCC_SOFTMAX) {
    for (int l = 1 + threadIdx.x; l <= ndoc; l += blockDim.x) {
      const int w0 = no1 + P.ndns + (ndoc - l), g0 = ndoc - l;
      double den = 1.0;
      for (int j = 0; j < ncwo; ++j) den += exp(gamma[g0 + ndoc * j]);
      n[no1 + l - 1] = 1.0 / den;
      for (int j = 0; j < ncwo; ++j) n[w0 + ndoc * j] = exp(gamma[g0 + ndoc * j]) / den;
    }
    __syncthreads();
    // dn[no1+l-1, g_j] = -n[l-1] n_wj (the reference reads n[i], pnof.jl:1152);
    // dn[w_j, g_k] = -n_wj n_wk + [j==k] n_wj
    const long per = (long)ncwo * (ncwo + 1);
    for (long e = threadIdx.x; e < per * ndoc; e += blockDim.x) {
      const int l = 1 + (int)(e / per);
      const int r = (int)(e % per);
      const int w0 = no1 + P.ndns + (ndoc - l), g0 = ndoc - l;
      const int col = r % ncwo, row = r / ncwo;  // row 0: strong orbital, 1..: weak
      const int gk = g0 + ndoc * col;
      const double nwk = n[w0 + ndoc * col];
      if (row == 0) {
        dn[(no1 + l - 1) + (long)gk * nbf5] = -n[l - 1] * nwk;
      } else {
        const int wj = w0 + ndoc * (row - 1);
        double v = -n[wj] * nwk;
        if (row - 1 == col) v += n[wj];
        dn[wj + (long)gk * nbf5] = v;
      }
    }
  } else {
    // Trigonometric: one thread per pair, the reference's loop order
    for (int l = 1 + threadIdx.x; l <= ndoc; l += blockDim.x) {
      const int ig = no1 + l - 1;
      const double gl = gamma[l - 1];
      const double cg = cos(gl);
      const double ng = 0.5 * (1.0 + cg * cg);
      const double dni = -0.5 * sin(2.0 * gl);
      n[ig] = ng;
      const double h = 1.0 - ng;
      const int w0 = no1 + P.ndns + (ndoc - l);
      const int gi0 = ndoc + (ndoc - l);
      // n_pi
      for (int r = 0; r < ncwo; ++r) {
        double v = h;
        for (int kw = 0; kw < ncwo - 1 && kw <= r; ++kw) {
          const double g = gamma[gi0 + ndoc * kw];
          if (kw == r) v *= sin(g) * sin(g);
          else v *= cos(g) * cos(g);
        }
        n[w0 + ndoc * r] = v;
      }
      dn[ig + (long)(l - 1) * nbf5] = dni;
      for (int r = 0; r < ncwo; ++r) {  // dn_pi/dgamma_g
        double v = -dni;
        for (int kw = 0; kw < ncwo - 1 && kw <= r; ++kw) {
          const double g = gamma[gi0 + ndoc * kw];
          if (kw == r) v *= sin(g) * sin(g);
          else v *= cos(g) * cos(g);
        }
        dn[(w0 + ndoc * r) + (long)(l - 1) * nbf5] = v;
      }
      for (int jw = 0; jw < ncwo - 1; ++jw) {  // dn_pi/dgamma_pj
        const long col = (long)(gi0 + ndoc * jw) * nbf5;
        const double gj = gamma[gi0 + ndoc * jw];
        for (int r = jw + 1; r < ncwo; ++r) {
          double v = ng - 1.0;
          for (int kw = 0; kw < jw; ++kw) {
            const double g = gamma[gi0 + ndoc * kw];
            v *= cos(g) * cos(g);
          }
          v *= sin(2.0 * gj);
          for (int kw = jw + 1; kw < ncwo - 1 && kw <= r; ++kw) {
            const double g = gamma[gi0 + ndoc * kw];
            if (kw == r) v *= sin(g) * sin(g);
            else v *= cos(g) * cos(g);
          }
          dn[(w0 + ndoc * r) + col] = v;
        }
        double d = 1.0 - ng;  // dn_pi/dgamma_pi
        for (int kw = 0; kw < jw; ++kw) {
          const double g = gamma[gi0 + ndoc * kw];
          d *= cos(g) * cos(g);
        }
        d *= sin(2.0 * gj);
        dn[(w0 + ndoc * jw) + col] = d;
      }
    }
  }
}

__global__ void ocupacion_kernel(OccDims P, const double* __restrict__ gamma,
                                 double* __restrict__ n, double* __restrict__ dn) {
  ocupacion_block(P, gamma, n, dn);
}

// ---- per-orbital quantities ---------------------------------------------------------------

// one block: role, pair, fi, n_d, sqrt(n_d), dfi/dn, dn_d coefficients, S_F
__device__ void orb_prep_block(const OccDims& P, const OccWork& w) {
  const int nbf5 = P.nbf5, no1 = P.no1, ndoc = P.ndoc;
  const double* n = w.n;
  const double c = (P.ipnof == 8 && P.ista == 5) ? 0.8 : 1.0;
  for (int i = threadIdx.x; i < nbf5; i += blockDim.x) {
    int role = kCwo;
    if (i < no1) role = kCore;
    else if (i < P.nbeta) role = kBeta;
    else if (i < P.nalpha) role = kSoc;
    w.role[i] = role;
    int l = 0, st = -1;
    if (i >= no1 && i < no1 + ndoc) {
      l = i - no1 + 1;
      st = i;
    } else if (i >= no1 + P.ndns && ndoc > 0) {
      l = ndoc - ((i - no1 - P.ndns) % ndoc);
      st = no1 + l - 1;
    }
    w.pair[i] = l;
    w.strong[i] = st;
    const double ni = n[i];
    // fi and dfi/dn by variant (pnof.jl:370-375/434-444 PNOF7; 522-538/656-702 GNOF;
    // 818-820/929-939 PNOF9)
    double x = ni * (1.0 - ni);
    if (x <= 0) x = 0.0;
    double fi = 0.0, dfic = 0.0;
    if (P.ipnof == 7) {
      if (P.ista == 0) {
        fi = sqrt(x);
        dfic = 1.0 / (2 * fmax(fi, 1e-15)) * (1 - 2 * ni);
      } else {
        fi = 2 * ni * (1.0 - ni);
        dfic = 2 * (1 - 2 * ni);
      }
    } else if (P.ipnof == 9 || (P.ipnof == 8 && (P.ista == 0 || P.ista == 1 || P.ista == 5))) {
      fi = sqrt(x);
      dfic = 1.0 / (2 * fmax(fi, 1e-15)) * (1 - 2 * ni);
    } else if (P.ipnof == 8 && P.ista == 2) {
      fi = 0.9 * sqrt(x);
      dfic = 0.81 / (2 * fmax(fi, 1e-15)) * (1 - 2 * ni);
    } else if (P.ipnof == 8 && P.ista == 3) {
      const double alpha = 0.55;
      const double aa = pow(pow(2.0, 2.0 * alpha - 1), 1 / alpha);
      fi = pow(aa * x, alpha);
      const double a = fmax(aa * ni * (1 - ni), 1e-15);
      dfic = alpha * pow(a, alpha - 1) * aa * (1 - 2 * ni);
    } else if (P.ipnof == 8 && P.ista == 4) {
      const double f0 = sqrt(x);
      const double e = exp(-(ni - 0.5) * (ni - 0.5));
      dfic = 1.0 / (2 * fmax(f0, 1e-15)) * (1 - 2 * ni) * e - f0 * 2 * (ni - 0.5) * e;
      fi = f0 * e;
    }
    if (i < no1) dfic = 0.0;
    w.fi[i] = fi;
    w.dfic[i] = dfic;
    // dynamic occupations (pnof.jl:506-521, 631-654)
    double nd = 0.0, ndc = 0.0, ndx = 0.0;
    if (l > 0 && (P.ipnof == 8 || P.ipnof == 9)) {
      const double ns = n[no1 + l - 1];
      const double coc = (1.0 - ns) / P.h_cut;
      const double F = c * exp(-coc * coc);
      nd = ni * F;
      if (st == i) {
        ndc = F * (1 - ni * (-2 * coc / P.h_cut));
      } else {
        ndc = F;
        ndx = F * (2 * coc / P.h_cut) * ni;
      }
    }
    w.nd[i] = nd;
    w.nd12[i] = sqrt(nd);
    w.ndc[i] = ndc;
    w.ndx[i] = ndx;
  }
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < ndoc; ++i) s += 1.0 - n[i];
    w.scal[0] = s;
  }
}

__global__ void orb_prep_kernel(OccDims P, OccWork w) { orb_prep_block(P, w); }

// cj12, ck12 with the diagonal zeroed (pnof.jl:49-50)
__global__ void coef_kernel(OccDims P, OccWork w, double* __restrict__ cj12,
                            double* __restrict__ ck12) {
  const Orb o = orb_view(w);
  const double S_F = w.scal[0];
  const long nn = (long)P.nbf5 * P.nbf5;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nn;
       e += (long)gridDim.x * blockDim.x) {
    const int i = e % P.nbf5, j = e / P.nbf5;
    double cj = 0.0, ck = 0.0;
    if (i != j) coef_elem(P, o, S_F, i, j, &cj, &ck);
    cj12[e] = cj;
    ck12[e] = ck;
  }
}

// Dcj12r, Dck12r materialised (nbf5 x nbf5 x nv), diagonal zeroed (pnof.jl:109-114).
__global__ void der_coef_kernel(OccDims P, OccWork w, const double* __restrict__ dn,
                                double* __restrict__ Dcj, double* __restrict__ Dck) {
  const Orb o = orb_view(w);
  const double S_F = fmax(w.scal[0], 1e-7);
  const int nbf5 = P.nbf5;
  const long nn = (long)nbf5 * nbf5;
  const long total = nn * P.nv;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int k = (int)(e / nn);
    const long r = e - (long)k * nn;
    const int i = r % nbf5, j = r / nbf5;
    const double* dk = dn + (long)k * nbf5;
    double DSF = 0.0;
    if (P.ipnof == 4)
      for (int p = 0; p < P.ndoc; ++p) DSF += -dk[p];
    double a = 0.0, b = 0.0;
    if (i != j) {
      const int st = o.strong[i];
      der_elem(P, o, S_F, DSF, i, j, dk[i], st >= 0 ? dk[st] : 0.0, dk[j], &a, &b);
    }
    Dcj[e] = a;
    Dck[e] = b;
  }
}

// grad[k] of calcoccg (pnof.jl:1324-1377): one block per k.
__device__ void calcoccg_block(const OccDims& P, const OccWork& w, const double* __restrict__ dn,
                               const double* __restrict__ J, const double* __restrict__ K,
                               const double* __restrict__ Hcore, double* __restrict__ grad,
                               int k, double* red) {
  const Orb o = orb_view(w);
  const double S_F = fmax(w.scal[0], 1e-7);
  const int nbf5 = P.nbf5;
  const double* dk = dn + (long)k * nbf5;
  double DSF = 0.0;
  if (P.ipnof == 4)
    for (int p = 0; p < P.ndoc; ++p) DSF += -dk[p];
  const int nb = P.nbeta - P.no1;             // rows no1 .. nbeta-1
  const int nr = nb + (nbf5 - P.nalpha);      // plus rows nalpha .. nbf5-1
  double acc = 0.0;
  for (int i = threadIdx.x; i < nr; i += blockDim.x) {  // dn (2 H + J_ii)
    const int row = (i < nb) ? P.no1 + i : P.nalpha + (i - nb);
    acc += dk[row] * (2 * Hcore[row] + J[row + (long)row * nbf5]);
  }
  const long cells = (long)nr * nbf5;
  for (long e = threadIdx.x; e < cells; e += blockDim.x) {
    const int jj = e % nbf5;
    const int i = (int)(e / nbf5);
    const int row = (i < nb) ? P.no1 + i : P.nalpha + (i - nb);
    if (row == jj) continue;
    const int st = o.strong[row];
    double dcj, dck;
    der_elem(P, o, S_F, DSF, row, jj, dk[row], st >= 0 ? dk[st] : 0.0, dk[jj], &dcj, &dck);
    const long ji = jj + (long)row * nbf5;  // J[j,i]
    acc += 2 * dcj * J[ji] - 2 * dck * K[ji];
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) grad[k] = s;
}

__global__ void calcoccg_kernel(OccDims P, OccWork w, const double* __restrict__ dn,
                                const double* __restrict__ J, const double* __restrict__ K,
                                const double* __restrict__ Hcore, double* __restrict__ grad) {
  __shared__ double red[16];
  calcoccg_block(P, w, dn, J, K, Hcore, grad, blockIdx.x, red);
}

// calce (pnof.jl:1273-1300): one block.
__global__ void calce_kernel(OccDims P, const double* __restrict__ n,
                             const double* __restrict__ cj12, const double* __restrict__ ck12,
                             const double* __restrict__ J, const double* __restrict__ K,
                             const double* __restrict__ Hcore, double* __restrict__ out) {
  __shared__ double red[16];
  const int nbf5 = P.nbf5;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nbf5; i += blockDim.x) {
    const double jii = J[i + (long)i * nbf5];
    if (i < P.nbeta) acc += n[i] * (2 * Hcore[i] + jii);
    else if (i < P.nalpha) {
      if (P.nsoc > 0) acc += n[i] * 2 * Hcore[i];
    } else acc += n[i] * (2 * Hcore[i] + jii);
  }
  const long nn = (long)nbf5 * nbf5;
  for (long e = threadIdx.x; e < nn; e += blockDim.x) {
    const int i = e % nbf5, j = e / nbf5;
    const long ji = j + (long)i * nbf5;
    acc += cj12[e] * J[ji] - ck12[e] * K[ji];
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) out[0] = s;
}

// ---- occupation optimiser: nonlinear CG on the device ------------------------------------
//
// The iteration of the host conjugate_gradient (Hager & Zhang, SIAM J. Optim. 16 (2005)
// 170-192: eta-truncated beta; Wolfe line search with bracketing and a safeguarded cubic
// zoom, Nocedal & Wright Alg. 3.5/3.6, accepting the approximate Wolfe conditions) as a
// state machine advanced once per evaluation: every thread of the block holds the same
// scalars (block reductions broadcast), so all take the same branches.

namespace {

constexpr int kCgThreads = 256;
enum { kCgInit = 0, kCgAfterInit = 1, kCgLs = 2, kCgZoom = 3 };
constexpr double kC1 = 0.1, kC2 = 0.9, kEpsWolfe = 1e-6, kAmax = 1e6, kEta = 0.4, kPsi0 = 0.01;
constexpr int kLsMax = 40, kZoomMax = 40;

__device__ double bsum_all(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < kCgThreads / 64; ++k) s += red[k];
  return s;
}

__device__ double bmax_all(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = red[0];
  for (int k = 1; k < kCgThreads / 64; ++k) s = fmax(s, red[k]);
  return s;
}

struct CgVec {
  double *x, *g, *d, *xt, *gt, *gb;
  int nv;
};

__device__ double dot(const double* a, const double* b, int nv, double* red) {
  double v = 0.0;
  for (int i = threadIdx.x; i < nv; i += kCgThreads) v += a[i] * b[i];
  return bsum_all(v, red);
}

__device__ double amax(const double* a, int nv, double* red) {
  double v = 0.0;
  for (int i = threadIdx.x; i < nv; i += kCgThreads) v = fmax(v, fabs(a[i]));
  return bmax_all(v, red);
}

__device__ void set_trial(const CgVec& v, double a) {  // xt = x + a d
  for (int i = threadIdx.x; i < v.nv; i += kCgThreads) v.xt[i] = v.x[i] + a * v.d[i];
  __syncthreads();
}

__device__ bool approx_wolfe(const OccCG& c, double f, double d) {
  return (2 * kC1 - 1) * c.d0 >= d && d >= kC2 * c.d0 && f <= c.f0 + kEpsWolfe * fabs(c.f0);
}

// a trial step inside the bracket: cubic interpolation safeguarded to its interior
__device__ double zoom_trial(const OccCG& c) {
  double a = 0.0;
  bool have = false;
  if (isfinite(c.fhi) && isfinite(c.dhi) && c.ahi != c.alo) {
    const double d1 = c.dlo + c.dhi - 3 * (c.flo - c.fhi) / (c.alo - c.ahi);
    const double rad = d1 * d1 - c.dlo * c.dhi;
    if (rad >= 0) {
      const double d2 = copysign(sqrt(rad), c.ahi - c.alo);
      const double den = c.dhi - c.dlo + 2 * d2;
      if (den != 0) {
        a = c.ahi - (c.ahi - c.alo) * (c.dhi + d2 - d1) / den;
        have = true;
      }
    }
  }
  const double lo = fmin(c.alo, c.ahi), hi = fmax(c.alo, c.ahi);
  if (!have || !(lo + 0.1 * (hi - lo) <= a && a <= hi - 0.1 * (hi - lo))) a = 0.5 * (c.alo + c.ahi);
  return a;
}

// start the line search of an iteration from step a (xt set)
__device__ void begin_iteration(OccCG& c, const CgVec& v, double* red) {
  c.d0 = dot(v.g, v.d, v.nv, red);
  if (c.d0 >= 0) {
    for (int i = threadIdx.x; i < v.nv; i += kCgThreads) v.d[i] = -v.g[i];
    __syncthreads();
    c.d0 = dot(v.g, v.d, v.nv, red);
  }
  c.f0 = c.f;
  c.a_prev = 0.0;
  c.f_prev = c.f0;
  c.d_prev = c.d0;
  c.ls_i = 0;
  c.state = kCgLs;
  set_trial(v, c.a);
}

__device__ void start_zoom(OccCG& c, const CgVec& v, double alo, double flo, double dlo,
                           double ahi, double fhi, double dhi) {
  c.alo = alo; c.flo = flo; c.dlo = dlo;
  c.ahi = ahi; c.fhi = fhi; c.dhi = dhi;
  c.zoom_i = 0;
  c.best_valid = 0;
  c.a = zoom_trial(c);
  c.state = kCgZoom;
  set_trial(v, c.a);
}

// the line search returned (a, fn, gn): the CG step (x += a d), convergence, the
// Hager-Zhang direction, the next iteration
__device__ void accept(OccCG& c, const CgVec& v, double a, double fn, const double* gn,
                       double* red) {
  if (!(isfinite(fn) && fn <= c.f + 1e-6 * fabs(c.f)) || a == 0.0) {
    c.done = 1;  // no acceptable step: stagnation at the round-off floor
    return;
  }
  for (int i = threadIdx.x; i < v.nv; i += kCgThreads) v.x[i] += a * v.d[i];
  __syncthreads();
  if (amax(gn, v.nv, red) <= c.tol) {
    c.f = fn;
    for (int i = threadIdx.x; i < v.nv; i += kCgThreads) v.g[i] = gn[i];
    __syncthreads();
    c.converged = 1;
    c.done = 1;
    return;
  }
  // y = gn - g (in xt's buffer: the trial point is not needed again)
  double* y = v.xt;
  for (int i = threadIdx.x; i < v.nv; i += kCgThreads) y[i] = gn[i] - v.g[i];
  __syncthreads();
  const double dy = dot(v.d, y, v.nv, red);
  double beta = 0.0;
  if (dy != 0.0 && isfinite(dy)) {
    const double yy = dot(y, y, v.nv, red);
    double t = 0.0;
    for (int i = threadIdx.x; i < v.nv; i += kCgThreads)
      t += (y[i] - 2.0 * v.d[i] * yy / dy) * gn[i];
    beta = bsum_all(t, red) / dy;
    const double dn = sqrt(dot(v.d, v.d, v.nv, red));
    const double gnorm = sqrt(dot(v.g, v.g, v.nv, red));
    beta = fmax(beta, -1.0 / (dn * fmin(kEta, gnorm)));
  }
  for (int i = threadIdx.x; i < v.nv; i += kCgThreads) {
    v.d[i] = -gn[i] + beta * v.d[i];
    v.g[i] = gn[i];
  }
  __syncthreads();
  c.f = fn;
  if (c.it >= c.maxiter) {
    c.done = 1;
    return;
  }
  c.it += 1;
  c.a = a;  // the next search starts from the accepted step
  begin_iteration(c, v, red);
}

// one controller step: consume the evaluation at xt (c.ft, gt), set the next xt
__device__ void cg_control(OccCG& c, const CgVec& v, double* red) {
  switch (c.state) {
    case kCgInit:
      for (int i = threadIdx.x; i < v.nv; i += kCgThreads) v.xt[i] = v.x[i];
      __syncthreads();
      c.state = kCgAfterInit;
      return;
    case kCgAfterInit: {
      c.f = c.ft;
      for (int i = threadIdx.x; i < v.nv; i += kCgThreads) v.g[i] = v.gt[i];
      __syncthreads();
      const double gmax = amax(v.g, v.nv, red);
      c.it = 0;
      if (gmax <= c.tol) {
        c.converged = 1;
        c.done = 1;
        return;
      }
      for (int i = threadIdx.x; i < v.nv; i += kCgThreads) v.d[i] = -v.g[i];
      __syncthreads();
      const double xn = amax(v.x, v.nv, red);
      c.a = xn > 0 ? kPsi0 * xn / gmax
                   : kPsi0 * fmax(fabs(c.f), 1.0) / dot(v.g, v.g, v.nv, red);
      c.it = 1;
      begin_iteration(c, v, red);
      return;
    }
    case kCgLs: {
      const double f = c.ft, d = dot(v.gt, v.d, v.nv, red);
      if (isfinite(f) && approx_wolfe(c, f, d)) return accept(c, v, c.a, f, v.gt, red);
      if (!isfinite(f) || f > c.f0 + kC1 * c.a * c.d0 || (c.ls_i > 0 && f >= c.f_prev))
        return start_zoom(c, v, c.a_prev, c.f_prev, c.d_prev, c.a, f, d);
      if (fabs(d) <= -kC2 * c.d0) return accept(c, v, c.a, f, v.gt, red);
      if (d >= 0) return start_zoom(c, v, c.a, f, d, c.a_prev, c.f_prev, c.d_prev);
      c.a_prev = c.a;
      c.f_prev = f;
      c.d_prev = d;
      c.ls_i += 1;
      if (c.ls_i >= kLsMax) return accept(c, v, c.a_prev, f, v.gt, red);
      c.a = fmin(2.0 * c.a, kAmax);
      set_trial(v, c.a);
      return;
    }
    case kCgZoom: {
      const double f = c.ft, d = dot(v.gt, v.d, v.nv, red);
      if (!c.best_valid || f < c.best_f) {
        c.best_valid = 1;
        c.best_a = c.a;
        c.best_f = f;
        for (int i = threadIdx.x; i < v.nv; i += kCgThreads) v.gb[i] = v.gt[i];
        __syncthreads();
      }
      if (isfinite(f) && approx_wolfe(c, f, d)) return accept(c, v, c.a, f, v.gt, red);
      if (!isfinite(f) || f > c.f0 + kC1 * c.a * c.d0 || f >= c.flo) {
        c.ahi = c.a; c.fhi = f; c.dhi = d;
      } else {
        if (fabs(d) <= -kC2 * c.d0) return accept(c, v, c.a, f, v.gt, red);
        if (d * (c.ahi - c.alo) >= 0) {
          c.ahi = c.alo; c.fhi = c.flo; c.dhi = c.dlo;
        }
        c.alo = c.a; c.flo = f; c.dlo = d;
      }
      c.zoom_i += 1;
      if (fabs(c.ahi - c.alo) < 1e-16 * fmax(1.0, fabs(c.alo)) || c.zoom_i >= kZoomMax)
        return accept(c, v, c.best_a, c.best_f, v.gb, red);
      c.a = zoom_trial(c);
      set_trial(v, c.a);
      return;
    }
    default:
      c.done = 1;
  }
}

}  // namespace

// slot part 1 (one block): the occupation energy at the previous trial point from part
// 2's partials (fixed order), the controller, then n, dn and the per-orbital quantities
// at the new trial point xt
__global__ __launch_bounds__(kCgThreads) void occ_cg_step_kernel(OccDims P, OccWork w,
                                                                 OccCG* __restrict__ gc,
                                                                 double* __restrict__ vec,
                                                                 double* __restrict__ dn,
                                                                 int nE) {
  __shared__ double red[kCgThreads / 64];
  OccCG c = *gc;  // every thread: the same copy
  if (c.done) return;
  CgVec v{vec + kCgX * P.nv, vec + kCgG * P.nv, vec + kCgD * P.nv,
          vec + kCgXT * P.nv, vec + kCgGT * P.nv, vec + kCgGB * P.nv, P.nv};
  if (c.state != kCgInit) {
    const double* part = vec + kCgVecs * P.nv;
    double f = 0.0;
    for (int b = 0; b < nE; ++b) f += part[b];
    c.ft = f;
    c.nfev += 1;
  }
  cg_control(c, v, red);
  if (!c.done) {
    ocupacion_block(P, v.xt, w.n, dn);
    __syncthreads();
    orb_prep_block(P, w);
  }
  __syncthreads();
  if (threadIdx.x == 0) *gc = c;
}

// slot part 2: blocks k < nv the gradient calcoccg[k] at xt (pnof.jl:1324-1377), blocks
// nv .. nv + nE - 1 slices of the occupation energy calce at xt (pnof.jl:1273-1300,
// cj12/ck12 built in place), J and K read along their columns
__global__ __launch_bounds__(512) void occ_cg_eval_kernel(OccDims P, OccWork w,
                                                          const OccCG* __restrict__ gc,
                                                          double* __restrict__ vec,
                                                          const double* __restrict__ dn,
                                                          const double* __restrict__ J,
                                                          const double* __restrict__ K,
                                                          const double* __restrict__ Hcore,
                                                          int nE) {
  __shared__ double red[16];
  if (gc->done) return;
  if ((int)blockIdx.x < P.nv) {
    calcoccg_block(P, w, dn, J, K, Hcore, vec + kCgGT * P.nv, blockIdx.x, red);
    return;
  }
  const int b = blockIdx.x - P.nv;
  const Orb o = orb_view(w);
  const double S_F = w.scal[0];
  const int nbf5 = P.nbf5;
  double acc = 0.0;
  if (b == 0) {
    for (int i = threadIdx.x; i < nbf5; i += blockDim.x) {
      const double jii = J[i + (long)i * nbf5];
      if (i < P.nbeta) acc += w.n[i] * (2 * Hcore[i] + jii);
      else if (i < P.nalpha) {
        if (P.nsoc > 0) acc += w.n[i] * 2 * Hcore[i];
      } else acc += w.n[i] * (2 * Hcore[i] + jii);
    }
  }
  const long nn = (long)nbf5 * nbf5;
  for (long e = b * (long)blockDim.x + threadIdx.x; e < nn; e += (long)nE * blockDim.x) {
    const int j = e % nbf5, i = e / nbf5;  // J[j,i] at e: coalesced
    if (i == j) continue;
    double cj, ck;
    coef_elem(P, o, S_F, i, j, &cj, &ck);
    acc += cj * J[e] - ck * K[e];
  }
  const double sum = block_sum(acc, red);
  if (threadIdx.x == 0) vec[kCgVecs * P.nv + b] = sum;
}

// ---- launchers ------------------------------------------------------------------------

int occ_cg_energy_blocks(const OccDims& P) {
  const long nn = (long)P.nbf5 * P.nbf5;
  return (int)std::min<long>(64, std::max<long>(1, (nn + 4095) / 4096));
}

int occ_cg_slots(hipStream_t s, const OccDims& P, const OccWork& w, OccCG* c, double* vec,
                 double* dn, const double* J, const double* K, const double* Hcore,
                 int nslots) {
  if (P.nv <= 0) return fail(DN_ERR_ARG, "occupation optimiser: no gamma variables");
  const int nE = occ_cg_energy_blocks(P);
  for (int i = 0; i < nslots; ++i) {
    hipLaunchKernelGGL(occ_cg_step_kernel, dim3(1), dim3(kCgThreads), 0, s, P, w, c, vec, dn,
                       nE);
    hipLaunchKernelGGL(occ_cg_eval_kernel, dim3(P.nv + nE), dim3(512), 0, s, P, w, c, vec, dn,
                       J, K, Hcore, nE);
  }
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int occ_ocupacion(hipStream_t s, const OccDims& P, const double* gamma, double* n, double* dn) {
  hipLaunchKernelGGL(ocupacion_kernel, dim3(1), dim3(256), 0, s, P, gamma, n, dn);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int occ_prepare(hipStream_t s, const OccDims& P, const OccWork& w) {
  hipLaunchKernelGGL(orb_prep_kernel, dim3(1), dim3(256), 0, s, P, w);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int occ_coefficients(hipStream_t s, const OccDims& P, const OccWork& w, double* cj12,
                     double* ck12) {
  const long nn = (long)P.nbf5 * P.nbf5;
  long blocks = (nn + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(coef_kernel, dim3((unsigned)blocks), dim3(256), 0, s, P, w, cj12, ck12);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int occ_der_coefficients(hipStream_t s, const OccDims& P, const OccWork& w, const double* dn,
                         double* Dcj, double* Dck) {
  const long total = (long)P.nbf5 * P.nbf5 * P.nv;
  long blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(der_coef_kernel, dim3((unsigned)blocks), dim3(256), 0, s, P, w, dn, Dcj,
                     Dck);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int occ_calcoccg(hipStream_t s, const OccDims& P, const OccWork& w, const double* dn,
                 const double* J, const double* K, const double* Hcore, double* grad) {
  if (P.nv <= 0) return DN_OK;
  hipLaunchKernelGGL(calcoccg_kernel, dim3(P.nv), dim3(256), 0, s, P, w, dn, J, K, Hcore, grad);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

int occ_calce(hipStream_t s, const OccDims& P, const double* n, const double* cj12,
              const double* ck12, const double* J, const double* K, const double* Hcore,
              double* out) {
  hipLaunchKernelGGL(calce_kernel, dim3(1), dim3(256), 0, s, P, n, cj12, ck12, J, K, Hcore, out);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

}  // namespace dn
