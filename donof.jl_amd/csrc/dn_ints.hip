// Gaussian integrals over contracted shells: the integral source of an RI system
// (compute_integrals, integrals.jl:1-65, which calls GaussianBasis for overlap, kinetic,
// nuclear, ERI_2e2c, ERI_2e3c and ERI_2e4c).
//
// Method: McMurchie-Davidson.  A product of two Cartesian Gaussians is expanded in
// Hermite Gaussians with the coefficients E^{ij}_t of the Hermite recursion; overlap and
// kinetic integrals follow from E^{ij}_0, Coulomb-type integrals from the Hermite
// Coulomb integrals R_{tuv} of the Boys function F_n(T) (series plus downward recursion
// for T <= 40, asymptote plus upward recursion above).  Reference: T. Helgaker,
// P. Jorgensen, J. Olsen, Molecular Electronic-Structure Theory (Wiley 2000), sections
// 9.5-9.9; real solid harmonics from its eq. 6.4.47-6.4.50.
//
// Functions: every shell is evaluated as Cartesian components x^i y^j z^k exp(-a r^2)
// sharing one radial contraction (primitives normalised as x^l exp(-a r^2), the
// contraction normalised to one), then mapped by the shell's transform matrix to its
// functions: real solid harmonics m = -l..l for pure shells (p as x, y, z), each
// normalised to one, or the Cartesian components x^l, x^(l-1) y, x^(l-1) z, ... each
// normalised to one.  All of it is host code: the sizes an RI system needs the
// integrals at (nbf^2 naux three-index integrals) are small next to the b the device
// builds from them, and the metric solve runs on the device (dn_sys_ri_metric).
#include <atomic>
#include <cmath>
#include <thread>

#include "dn_internal.h"

struct dn_basis {
  struct Shell {
    int l, pure, nprim, nfunc, ncart, offset;
    double ctr[3];
    std::vector<double> exps, coef;  // coef: normalised primitive times contraction factor
    std::vector<double> tr;          // nfunc x ncart transform (row-major)
  };
  std::vector<Shell> shells;
  int nbf = 0, lmax = 0;
};

namespace dn {
namespace {

constexpr int kMaxL = 6;  // i functions per shell at most

double dfact(int n) {  // n!! for n >= -1
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

constexpr int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian components of shell l in the order x^l, x^(l-1)y, x^(l-1)z, x^(l-2)y^2, ...
void cart_list(int l, int (*c)[3]) {
  int k = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      c[k][0] = lx;
      c[k][1] = ly;
      c[k][2] = l - lx - ly;
      ++k;
    }
}

int cart_index(int l, int lx, int ly) {  // position of (lx, ly, l-lx-ly) in cart_list
  const int a = l - lx;                  // components before: sum_{d<a} (d+1)
  return a * (a + 1) / 2 + (a - ly);
}

// Boys function F_0..F_nmax at T
void boys(int nmax, double T, double* F) {
  if (T < 1e-15) {
    for (int n = 0; n <= nmax; ++n) F[n] = 1.0 / (2 * n + 1);
    return;
  }
  const double eT = std::exp(-T);
  if (T > 40.0) {
    F[0] = 0.5 * std::sqrt(M_PI / T);
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - eT) / (2.0 * T);
    return;
  }
  double term = 1.0 / (2 * nmax + 1), sum = term;
  for (int k = 1; k < 400; ++k) {
    term *= 2.0 * T / (2 * nmax + 2 * k + 1);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  F[nmax] = eT * sum;
  for (int n = nmax - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + eT) / (2 * n + 1);
}

// Hermite coefficients of one Cartesian direction: E[(i (lb+1) + j) (la+lb+1) + t] for the
// product x_A^i x_B^j of exponents a, b with XAB = A - B.
void hermite_e(int la, int lb, double a, double b, double XAB, double* E) {
  const double p = a + b, mu = a * b / p;
  const double XPA = -b * XAB / p, XPB = a * XAB / p, h = 0.5 / p;
  const int LT = la + lb + 1;
  auto at = [&](int i, int j, int t) -> double& { return E[(i * (lb + 1) + j) * LT + t]; };
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j)
      for (int t = 0; t < LT; ++t) at(i, j, t) = 0.0;
  at(0, 0, 0) = std::exp(-mu * XAB * XAB);
  for (int i = 0; i < la; ++i)
    for (int t = 0; t <= i + 1; ++t) {
      double v = XPA * at(i, 0, t);
      if (t > 0) v += h * at(i, 0, t - 1);
      if (t + 1 <= i) v += (t + 1) * at(i, 0, t + 1);
      at(i + 1, 0, t) = v;
    }
  for (int j = 0; j < lb; ++j)
    for (int i = 0; i <= la; ++i)
      for (int t = 0; t <= i + j + 1; ++t) {
        double v = XPB * at(i, j, t);
        if (t > 0) v += h * at(i, j, t - 1);
        if (t + 1 <= i + j) v += (t + 1) * at(i, j, t + 1);
        at(i, j + 1, t) = v;
      }
}

// Hermite Coulomb integrals R^0_{tuv}(alpha, PC) for t+u+v <= L, at R[(t (L+1) + u) (L+1) + v].
// Work array w holds the levels n (size (L+1)^4).
void hermite_r(int L, double alpha, const double* PC, double* R, std::vector<double>& w) {
  const int D = L + 1;
  w.assign((size_t)D * D * D * D, 0.0);
  auto at = [&](int n, int t, int u, int v) -> double& {
    return w[(((size_t)n * D + t) * D + u) * D + v];
  };
  double F[4 * kMaxL + 8];
  const double T = alpha * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
  boys(L, T, F);
  double f = 1.0;
  for (int n = 0; n <= L; ++n) {
    at(n, 0, 0, 0) = f * F[n];
    f *= -2.0 * alpha;
  }
  for (int n = L - 1; n >= 0; --n)
    for (int t = 0; t <= L - n; ++t)
      for (int u = 0; u <= L - n - t; ++u)
        for (int v = 0; v <= L - n - t - u; ++v) {
          if (t + u + v == 0) continue;
          double r;
          if (t > 0) {
            r = PC[0] * at(n + 1, t - 1, u, v);
            if (t > 1) r += (t - 1) * at(n + 1, t - 2, u, v);
          } else if (u > 0) {
            r = PC[1] * at(n + 1, t, u - 1, v);
            if (u > 1) r += (u - 1) * at(n + 1, t, u - 2, v);
          } else {
            r = PC[2] * at(n + 1, t, u, v - 1);
            if (v > 1) r += (v - 1) * at(n + 1, t, u, v - 2);
          }
          at(n, t, u, v) = r;
        }
  for (int t = 0; t <= L; ++t)
    for (int u = 0; u <= L - t; ++u)
      for (int v = 0; v <= L - t - u; ++v) R[((size_t)t * D + u) * D + v] = at(0, t, u, v);
}

// Hermite expansion of a shell pair: for every Cartesian pair (ca, cb) the coefficients
// E_{tuv} = E^x_t E^y_u E^z_v over t+u+v <= la+lb, stored densely in [ncart_a][ncart_b][H]
// with H = (L+1)^3 indexed (t (L+1) + u) (L+1) + v.
struct PairExp {
  double p, P[3], pref;  // pref: the primitive pair's coefficient product
  std::vector<double> e;
};

void pair_expansion(const dn_basis::Shell& A, int ia, const dn_basis::Shell& B, int ib,
                    PairExp& out) {
  const double a = A.exps[ia], b = B.exps[ib];
  const int la = A.l, lb = B.l, L = la + lb, D = L + 1;
  out.p = a + b;
  for (int d = 0; d < 3; ++d) out.P[d] = (a * A.ctr[d] + b * B.ctr[d]) / out.p;
  out.pref = A.coef[ia] * B.coef[ib];
  double Ex[(kMaxL + 1) * (kMaxL + 1) * (2 * kMaxL + 1)], Ey[sizeof(Ex) / sizeof(double)],
      Ez[sizeof(Ex) / sizeof(double)];
  hermite_e(la, lb, a, b, A.ctr[0] - B.ctr[0], Ex);
  hermite_e(la, lb, a, b, A.ctr[1] - B.ctr[1], Ey);
  hermite_e(la, lb, a, b, A.ctr[2] - B.ctr[2], Ez);
  int ca[ncart_of(kMaxL)][3], cb[ncart_of(kMaxL)][3];
  cart_list(la, ca);
  cart_list(lb, cb);
  const int na = A.ncart, nb = B.ncart, H = D * D * D;
  out.e.assign((size_t)na * nb * H, 0.0);
  auto ex = [&](const double* E, int i, int j, int t) { return E[(i * (lb + 1) + j) * D + t]; };
  for (int x = 0; x < na; ++x)
    for (int y = 0; y < nb; ++y) {
      double* e = &out.e[((size_t)x * nb + y) * H];
      const int ix = ca[x][0], iy = ca[x][1], iz = ca[x][2];
      const int jx = cb[y][0], jy = cb[y][1], jz = cb[y][2];
      for (int t = 0; t <= ix + jx; ++t) {
        const double et = ex(Ex, ix, jx, t);
        for (int u = 0; u <= iy + jy; ++u) {
          const double eu = et * ex(Ey, iy, jy, u);
          for (int v = 0; v <= iz + jz; ++v) e[(t * D + u) * D + v] = eu * ex(Ez, iz, jz, v);
        }
      }
    }
}

// The unit shell (l = 0, one primitive of exponent 0, coefficient 1) that turns a pair
// routine into a single-shell one: x^i x_B^0 with b = 0 is the single Gaussian at A.
dn_basis::Shell unit_shell(const double* ctr) {
  dn_basis::Shell s;
  s.l = 0;
  s.pure = 0;
  s.nprim = 1;
  s.nfunc = s.ncart = 1;
  s.offset = 0;
  for (int d = 0; d < 3; ++d) s.ctr[d] = ctr[d];
  s.exps = {0.0};
  s.coef = {1.0};
  s.tr = {1.0};
  return s;
}

// ---- Cartesian shell-level integrals (before the transform to functions) --------------

// overlap (cart_a x cart_b, row-major)
void cart_overlap(const dn_basis::Shell& A, const dn_basis::Shell& B, double* out) {
  const int na = A.ncart, nb = B.ncart;
  std::fill(out, out + na * nb, 0.0);
  const int la = A.l, lb = B.l, D = la + lb + 1;
  int ca[ncart_of(kMaxL)][3], cb[ncart_of(kMaxL)][3];
  cart_list(la, ca);
  cart_list(lb, cb);
  double Ex[(kMaxL + 1) * (kMaxL + 1) * (2 * kMaxL + 1)], Ey[sizeof(Ex) / 8], Ez[sizeof(Ex) / 8];
  for (int i = 0; i < A.nprim; ++i)
    for (int j = 0; j < B.nprim; ++j) {
      const double a = A.exps[i], b = B.exps[j], p = a + b;
      hermite_e(la, lb, a, b, A.ctr[0] - B.ctr[0], Ex);
      hermite_e(la, lb, a, b, A.ctr[1] - B.ctr[1], Ey);
      hermite_e(la, lb, a, b, A.ctr[2] - B.ctr[2], Ez);
      const double f = A.coef[i] * B.coef[j] * std::pow(M_PI / p, 1.5);
      for (int x = 0; x < na; ++x)
        for (int y = 0; y < nb; ++y)
          out[x * nb + y] += f * Ex[(ca[x][0] * (lb + 1) + cb[y][0]) * D] *
                             Ey[(ca[x][1] * (lb + 1) + cb[y][1]) * D] *
                             Ez[(ca[x][2] * (lb + 1) + cb[y][2]) * D];
    }
}

// kinetic energy -1/2 <a|nabla^2|b>
void cart_kinetic(const dn_basis::Shell& A, const dn_basis::Shell& B, double* out) {
  const int na = A.ncart, nb = B.ncart;
  std::fill(out, out + na * nb, 0.0);
  const int la = A.l, lb = B.l, lb2 = lb + 2, D = la + lb2 + 1;
  int ca[ncart_of(kMaxL)][3], cb[ncart_of(kMaxL)][3];
  cart_list(la, ca);
  cart_list(lb, cb);
  double Ex[(kMaxL + 1) * (kMaxL + 3) * (2 * kMaxL + 3)], Ey[sizeof(Ex) / 8], Ez[sizeof(Ex) / 8];
  for (int i = 0; i < A.nprim; ++i)
    for (int j = 0; j < B.nprim; ++j) {
      const double a = A.exps[i], b = B.exps[j], p = a + b;
      hermite_e(la, lb2, a, b, A.ctr[0] - B.ctr[0], Ex);
      hermite_e(la, lb2, a, b, A.ctr[1] - B.ctr[1], Ey);
      hermite_e(la, lb2, a, b, A.ctr[2] - B.ctr[2], Ez);
      const double s1 = std::sqrt(M_PI / p);
      auto S = [&](const double* E, int ii, int jj) {
        return jj < 0 ? 0.0 : E[(ii * (lb2 + 1) + jj) * D] * s1;
      };
      auto K = [&](const double* E, int ii, int jj) {  // -1/2 <i| d^2/dx^2 |j>
        return -0.5 * (jj * (jj - 1) * S(E, ii, jj - 2) - 2.0 * b * (2 * jj + 1) * S(E, ii, jj) +
                       4.0 * b * b * S(E, ii, jj + 2));
      };
      const double f = A.coef[i] * B.coef[j];
      for (int x = 0; x < na; ++x)
        for (int y = 0; y < nb; ++y) {
          const int ix = ca[x][0], iy = ca[x][1], iz = ca[x][2];
          const int jx = cb[y][0], jy = cb[y][1], jz = cb[y][2];
          const double sx = S(Ex, ix, jx), sy = S(Ey, iy, jy), sz = S(Ez, iz, jz);
          out[x * nb + y] +=
              f * (K(Ex, ix, jx) * sy * sz + sx * K(Ey, iy, jy) * sz + sx * sy * K(Ez, iz, jz));
        }
    }
}

// nuclear attraction sum_C -Z_C <a| 1/|r - C| |b>
void cart_nuclear(const dn_basis::Shell& A, const dn_basis::Shell& B, int natom,
                  const double* Z, const double* xyz, double* out, std::vector<double>& w) {
  const int na = A.ncart, nb = B.ncart, L = A.l + B.l, D = L + 1, H = D * D * D;
  std::fill(out, out + na * nb, 0.0);
  std::vector<double> R(H);
  PairExp pe;
  for (int i = 0; i < A.nprim; ++i)
    for (int j = 0; j < B.nprim; ++j) {
      pair_expansion(A, i, B, j, pe);
      for (int c = 0; c < natom; ++c) {
        const double PC[3] = {pe.P[0] - xyz[3 * c], pe.P[1] - xyz[3 * c + 1],
                              pe.P[2] - xyz[3 * c + 2]};
        hermite_r(L, pe.p, PC, R.data(), w);
        const double f = -Z[c] * 2.0 * M_PI / pe.p * pe.pref;
        for (int xy = 0; xy < na * nb; ++xy) {
          const double* e = &pe.e[(size_t)xy * H];
          double s = 0.0;
          for (int h = 0; h < H; ++h) s += e[h] * R[h];
          out[xy] += f * s;
        }
      }
    }
}

// Coulomb integral of two shell pairs (ab|cd) over Cartesian components,
// out[((a nb + b) nc + c) nd + d]
void cart_eri(const dn_basis::Shell& A, const dn_basis::Shell& B, const dn_basis::Shell& C,
              const dn_basis::Shell& Dd, double* out, std::vector<double>& w) {
  const int na = A.ncart, nb = B.ncart, nc = C.ncart, nd = Dd.ncart;
  const int Lab = A.l + B.l, Lcd = C.l + Dd.l, L = Lab + Lcd, DL = L + 1;
  const int Dab = Lab + 1, Dcd = Lcd + 1, Hab = Dab * Dab * Dab, Hcd = Dcd * Dcd * Dcd;
  const size_t nab = (size_t)na * nb, ncd = (size_t)nc * nd;
  std::fill(out, out + nab * ncd, 0.0);
  std::vector<double> R((size_t)DL * DL * DL);
  std::vector<PairExp> ket(C.nprim * Dd.nprim);
  for (int k = 0; k < C.nprim; ++k)
    for (int l = 0; l < Dd.nprim; ++l) pair_expansion(C, k, Dd, l, ket[k * Dd.nprim + l]);
  // ket Hermite indices (tau, nu, phi) with the sign (-1)^(tau+nu+phi)
  std::vector<int> kt, ku, kv;
  for (int t = 0; t <= Lcd; ++t)
    for (int u = 0; u <= Lcd - t; ++u)
      for (int v = 0; v <= Lcd - t - u; ++v) {
        kt.push_back(t);
        ku.push_back(u);
        kv.push_back(v);
      }
  std::vector<int> bt, bu, bv;
  for (int t = 0; t <= Lab; ++t)
    for (int u = 0; u <= Lab - t; ++u)
      for (int v = 0; v <= Lab - t - u; ++v) {
        bt.push_back(t);
        bu.push_back(u);
        bv.push_back(v);
      }
  std::vector<double> G(bt.size() * ncd);  // sum over ket Hermite terms for each bra term
  PairExp bra;
  for (int i = 0; i < A.nprim; ++i)
    for (int j = 0; j < B.nprim; ++j) {
      pair_expansion(A, i, B, j, bra);
      for (const PairExp& q : ket) {
        const double PQ[3] = {bra.P[0] - q.P[0], bra.P[1] - q.P[1], bra.P[2] - q.P[2]};
        const double alpha = bra.p * q.p / (bra.p + q.p);
        hermite_r(L, alpha, PQ, R.data(), w);
        const double f = 2.0 * std::pow(M_PI, 2.5) / (bra.p * q.p * std::sqrt(bra.p + q.p)) *
                         bra.pref * q.pref;
        for (size_t h = 0; h < bt.size(); ++h)
          for (size_t cd = 0; cd < ncd; ++cd) {
            const double* e = &q.e[cd * Hcd];
            double s = 0.0;
            for (size_t g = 0; g < kt.size(); ++g) {
              const int t = kt[g], u = ku[g], v = kv[g];
              const double c = e[(t * Dcd + u) * Dcd + v];
              if (c == 0.0) continue;
              const double r = R[((bt[h] + t) * DL + bu[h] + u) * DL + bv[h] + v];
              s += ((t + u + v) & 1) ? -c * r : c * r;
            }
            G[h * ncd + cd] = s;
          }
        for (size_t ab = 0; ab < nab; ++ab) {
          const double* e = &bra.e[ab * Hab];
          double* o = out + ab * ncd;
          for (size_t h = 0; h < bt.size(); ++h) {
            const double c = e[(bt[h] * Dab + bu[h]) * Dab + bv[h]];
            if (c == 0.0) continue;
            const double fc = f * c;
            const double* g = &G[h * ncd];
            for (size_t cd = 0; cd < ncd; ++cd) o[cd] += fc * g[cd];
          }
        }
      }
    }
}

// ---- shells ---------------------------------------------------------------------------

// transform rows for shell l: real solid harmonics (HJO eq. 6.4.48, up to normalisation),
// or unit rows for Cartesian shells; normalised afterwards from the self-overlap
void shell_transform(dn_basis::Shell& s) {
  const int l = s.l, nc = s.ncart;
  std::vector<double> tr;
  if (s.pure) {
    std::vector<int> ms;
    if (l == 1) ms = {1, -1, 0};  // p as x, y, z
    else
      for (int m = -l; m <= l; ++m) ms.push_back(m);
    for (int m : ms) {
      std::vector<double> row(nc, 0.0);
      const int am = std::abs(m);
      for (int t = 0; t <= (l - am) / 2; ++t)
        for (int u = 0; u <= t; ++u)
          for (int w = (m < 0 ? 1 : 0); w <= am; w += 2) {  // w = 2v
            const int sgn_exp = t + (w - (m < 0 ? 1 : 0)) / 2;
            const double c = ((sgn_exp & 1) ? -1.0 : 1.0) * std::pow(0.25, t) * binom(l, t) *
                             binom(l - t, am + t) * binom(t, u) * binom(am, w);
            const int px = 2 * t + am - 2 * u - w, py = 2 * u + w;
            row[cart_index(l, px, py)] += c;
          }
      tr.insert(tr.end(), row.begin(), row.end());
    }
  } else {
    tr.assign((size_t)nc * nc, 0.0);
    for (int k = 0; k < nc; ++k) tr[(size_t)k * nc + k] = 1.0;
  }
  s.nfunc = (int)tr.size() / nc;
  std::vector<double> S((size_t)nc * nc);
  cart_overlap(s, s, S.data());
  for (int f = 0; f < s.nfunc; ++f) {
    double n2 = 0.0;
    for (int x = 0; x < nc; ++x)
      for (int y = 0; y < nc; ++y) n2 += tr[(size_t)f * nc + x] * S[(size_t)x * nc + y] * tr[(size_t)f * nc + y];
    const double inv = 1.0 / std::sqrt(n2);
    for (int x = 0; x < nc; ++x) tr[(size_t)f * nc + x] *= inv;
  }
  s.tr = tr;
}

// functions of two shells from Cartesian integrals: out(fa, fb) = Ta cart Tb^T, stored
// column-major in the nbf x nbf matrix M (leading dimension ld)
void put_pair(const dn_basis::Shell& A, const dn_basis::Shell& B, const double* cart,
              double* M, long ld, bool mirror) {
  for (int fa = 0; fa < A.nfunc; ++fa)
    for (int fb = 0; fb < B.nfunc; ++fb) {
      double s = 0.0;
      for (int x = 0; x < A.ncart; ++x) {
        const double ta = A.tr[(size_t)fa * A.ncart + x];
        if (ta == 0.0) continue;
        for (int y = 0; y < B.ncart; ++y) s += ta * cart[x * B.ncart + y] * B.tr[(size_t)fb * B.ncart + y];
      }
      const long i = A.offset + fa, j = B.offset + fb;
      M[i + j * ld] = s;
      if (mirror) M[j + i * ld] = s;
    }
}

// run f(k) for k in [0, n) on the host's threads, dynamically scheduled
void parallel_for(long n, const std::function<void(long)>& f) {
  unsigned nt = std::max(1u, std::thread::hardware_concurrency());
  nt = (unsigned)std::min<long>(nt, std::max<long>(1, n));
  std::atomic<long> next(0);
  auto body = [&]() {
    for (long k; (k = next.fetch_add(1)) < n;) f(k);
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(body);
  body();
  for (auto& t : th) t.join();
}

// ---- derivatives with respect to a shell's centre --------------------------------------
//
// d/dA_x of (x - A_x)^i exp(-a |r - A|^2) is 2a (x - A_x)^(i+1) exp(..) - i (x - A_x)^(i-1)
// exp(..): the derivative of a Cartesian component of shell A is a combination of the
// components of the same contraction at l + 1 (primitive coefficients times 2a) and l - 1.

dn_basis::Shell shifted(const dn_basis::Shell& A, int dl) {
  dn_basis::Shell s = A;
  s.l = A.l + dl;
  s.ncart = ncart_of(s.l);
  s.nfunc = s.ncart;
  if (dl > 0)
    for (int i = 0; i < s.nprim; ++i) s.coef[i] = 2.0 * A.exps[i] * A.coef[i];
  s.tr.clear();
  return s;
}

// out[k][c][r] = d/dA_k of f's Cartesian integrals of shell A (rows c, nrest columns r),
// where f(shell, buf) writes [ncart(shell)][nrest] for a shell on A's centre.
template <class F>
void deriv_center(const dn_basis::Shell& A, size_t nrest, F&& f, double* out) {
  const int na = A.ncart;
  const dn_basis::Shell Ap = shifted(A, +1);
  std::vector<double> P((size_t)Ap.ncart * nrest), M;
  f(Ap, P.data());
  if (A.l > 0) {
    const dn_basis::Shell Am = shifted(A, -1);
    M.resize((size_t)Am.ncart * nrest);
    f(Am, M.data());
  }
  int ca[ncart_of(kMaxL)][3];
  cart_list(A.l, ca);
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < na; ++c) {
      int up[3] = {ca[c][0], ca[c][1], ca[c][2]};
      up[k] += 1;
      const double* pp = &P[(size_t)cart_index(A.l + 1, up[0], up[1]) * nrest];
      double* o = out + ((size_t)k * na + c) * nrest;
      for (size_t r = 0; r < nrest; ++r) o[r] = pp[r];
      if (ca[c][k] > 0) {
        int dn_[3] = {ca[c][0], ca[c][1], ca[c][2]};
        dn_[k] -= 1;
        const double* mm = &M[(size_t)cart_index(A.l - 1, dn_[0], dn_[1]) * nrest];
        for (size_t r = 0; r < nrest; ++r) o[r] -= ca[c][k] * mm[r];
      }
    }
}

// ---- three-centre integrals on the device ---------------------------------------------
//
// One workgroup per (orbital shell pair A >= B, auxiliary shell P); a launch per angular
// momentum class (la, lb, lc), which fixes every size at compile time.  For each primitive
// triple the workgroup builds the Hermite Coulomb integrals R_tuv (t + u + v <= L) in LDS,
// level by level of the Boys-function recursion; thread (a, b) then contracts them with its
// Cartesian pair's Hermite coefficients and every component c of P, accumulating the
// contracted Cartesian block in registers.  The block goes through LDS to the shells'
// transform matrices and is written to I3c[m + ld n + ld nbf Q] and its mirror.

struct DevShell {
  double x, y, z;
  int l, nprim, prim, func, tr;  // prim: offset of exps/coefs, func: first function, tr: offset
  int nfunc;
};

constexpr int kDevMaxLab = 2;   // orbital shells up to d
constexpr int kDevMaxLc = 4;    // fitting shells up to g

__host__ __device__ constexpr int nc_of(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian exponents of component k of shell l (the order of cart_list)
__device__ __forceinline__ void comp_of(int l, int k, int* ex) {
  int c = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      if (c == k) {
        ex[0] = lx;
        ex[1] = ly;
        ex[2] = l - lx - ly;
        return;
      }
      ++c;
    }
}

__device__ void boys_dev(int nmax, double T, double* F) {
  if (T < 1e-15) {
    for (int n = 0; n <= nmax; ++n) F[n] = 1.0 / (2 * n + 1);
    return;
  }
  const double eT = exp(-T);
  if (T > 40.0) {
    F[0] = 0.5 * sqrt(M_PI / T);
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - eT) / (2.0 * T);
    return;
  }
  double term = 1.0 / (2 * nmax + 1), sum = term;
  for (int k = 1; k < 400; ++k) {
    term *= 2.0 * T / (2 * nmax + 2 * k + 1);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  F[nmax] = eT * sum;
  for (int n = nmax - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + eT) / (2 * n + 1);
}

// 1D Hermite coefficients of x_A^i x_B^j (i <= LI, j <= LJ) for one Cartesian exponent pair
template <int LI, int LJ>
__device__ __forceinline__ void herm1(int i, int j, double a, double b, double XAB,
                                      double (&E)[LI + LJ + 1]) {
  const double p = a + b, h = 0.5 / p;
  const double XPA = -b * XAB / p, XPB = a * XAB / p;
  double cur[LI + LJ + 1];
  for (int t = 0; t <= LI + LJ; ++t) cur[t] = 0.0;
  cur[0] = exp(-a * b / p * XAB * XAB);
  for (int ii = 0; ii < i; ++ii) {  // raise i
    double nx[LI + LJ + 1];
    for (int t = 0; t <= LI + LJ; ++t) {
      double v = XPA * cur[t];
      if (t > 0) v += h * cur[t - 1];
      if (t + 1 <= ii) v += (t + 1) * cur[t + 1];
      nx[t] = t <= ii + 1 ? v : 0.0;
    }
    for (int t = 0; t <= LI + LJ; ++t) cur[t] = nx[t];
  }
  for (int jj = 0; jj < j; ++jj) {  // raise j
    double nx[LI + LJ + 1];
    for (int t = 0; t <= LI + LJ; ++t) {
      double v = XPB * cur[t];
      if (t > 0) v += h * cur[t - 1];
      if (t + 1 <= i + jj) v += (t + 1) * cur[t + 1];
      nx[t] = t <= i + jj + 1 ? v : 0.0;
    }
    for (int t = 0; t <= LI + LJ; ++t) cur[t] = nx[t];
  }
  for (int t = 0; t <= LI + LJ; ++t) E[t] = cur[t];
}

template <int LA, int LB, int LC>
__global__ __launch_bounds__(64) void int3c_kernel(const DevShell* __restrict__ shells,
                                                   const DevShell* __restrict__ ashells,
                                                   const double* __restrict__ prim,
                                                   const double* __restrict__ trm,
                                                   const int2* __restrict__ pairs, int npairs,
                                                   const int* __restrict__ pshell, int np,
                                                   double* __restrict__ out, long ld,
                                                   long slab) {
  constexpr int NA = nc_of(LA), NB = nc_of(LB), NC = nc_of(LC);
  constexpr int L = LA + LB + LC, D = L + 1;
  __shared__ double Rb[2][D * D * D];
  __shared__ double Fn[L + 1];
  __shared__ double cart[NA * NB * NC];
  const long wg = blockIdx.x;
  if (wg >= (long)npairs * np) return;
  const int2 pr = pairs[wg / np];
  const DevShell A = shells[pr.x], B = shells[pr.y];
  const DevShell P = ashells[pshell[wg % np]];
  const int tid = threadIdx.x;
  const bool act = tid < NA * NB;
  int ea[3] = {0, 0, 0}, eb[3] = {0, 0, 0};
  if (act) {
    comp_of(LA, tid / NB, ea);
    comp_of(LB, tid % NB, eb);
  }
  double acc[NC];
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
  for (int i = 0; i < A.nprim; ++i)
    for (int j = 0; j < B.nprim; ++j) {
      const double a = prim[2 * (A.prim + i)], ca = prim[2 * (A.prim + i) + 1];
      const double b = prim[2 * (B.prim + j)], cb = prim[2 * (B.prim + j) + 1];
      const double p = a + b;
      const double Px = (a * A.x + b * B.x) / p, Py = (a * A.y + b * B.y) / p,
                   Pz = (a * A.z + b * B.z) / p;
      double Ex[LA + LB + 1], Ey[LA + LB + 1], Ez[LA + LB + 1];
      if (act) {
        herm1<LA, LB>(ea[0], eb[0], a, b, AB[0], Ex);
        herm1<LA, LB>(ea[1], eb[1], a, b, AB[1], Ey);
        herm1<LA, LB>(ea[2], eb[2], a, b, AB[2], Ez);
      }
      for (int k = 0; k < P.nprim; ++k) {
        const double c = prim[2 * (P.prim + k)], cc = prim[2 * (P.prim + k) + 1];
        const double alpha = p * c / (p + c);
        const double X = Px - P.x, Y = Py - P.y, Z = Pz - P.z;
        __syncthreads();  // the previous triple's R is no longer read
        if (tid == 0) boys_dev(L, alpha * (X * X + Y * Y + Z * Z), Fn);
        __syncthreads();
        // levels n = L .. 0 of R^(n)_tuv, t + u + v <= L - n, ping-pong buffers
        for (int n = L; n >= 0; --n) {
          double* cur = Rb[n & 1];
          const double* prv = Rb[(n + 1) & 1];
          const int m = L - n;
          const int cnt = (m + 1) * (m + 2) * (m + 3) / 6;
          for (int e = tid; e < cnt; e += 64) {
            // e-th (t, u, v) with t + u + v <= m
            int t = 0, u = 0, v = 0, r = e;
            for (t = 0; t <= m; ++t) {
              const int k2 = (m - t + 1) * (m - t + 2) / 2;
              if (r < k2) break;
              r -= k2;
            }
            for (u = 0; u <= m - t; ++u) {
              const int k1 = m - t - u + 1;
              if (r < k1) break;
              r -= k1;
            }
            v = r;
            double val;
            if (t + u + v == 0) {
              val = Fn[n];
              for (int q = 0; q < n; ++q) val *= -2.0 * alpha;
            } else if (t > 0) {
              val = X * prv[((t - 1) * D + u) * D + v];
              if (t > 1) val += (t - 1) * prv[((t - 2) * D + u) * D + v];
            } else if (u > 0) {
              val = Y * prv[(t * D + u - 1) * D + v];
              if (u > 1) val += (u - 1) * prv[(t * D + u - 2) * D + v];
            } else {
              val = Z * prv[(t * D + u) * D + v - 1];
              if (v > 1) val += (v - 1) * prv[(t * D + u) * D + v - 2];
            }
            cur[(t * D + u) * D + v] = val;
          }
          __syncthreads();
        }
        if (act) {
          const double* R = Rb[0];
          const double f = 2.0 * pow(M_PI, 2.5) / (p * c * sqrt(p + c)) * ca * cb * cc;
          for (int kc = 0; kc < NC; ++kc) {
            int ec[3];
            comp_of(LC, kc, ec);
            double Fx[LC + 1], Fy[LC + 1], Fz[LC + 1];
            herm1<LC, 0>(ec[0], 0, c, 0.0, 0.0, Fx);
            herm1<LC, 0>(ec[1], 0, c, 0.0, 0.0, Fy);
            herm1<LC, 0>(ec[2], 0, c, 0.0, 0.0, Fz);
            double s = 0.0;
            for (int t = 0; t <= ea[0] + eb[0]; ++t)
              for (int u = 0; u <= ea[1] + eb[1]; ++u)
                for (int v = 0; v <= ea[2] + eb[2]; ++v) {
                  const double e = Ex[t] * Ey[u] * Ez[v];
                  double g = 0.0;
                  for (int tt = 0; tt <= ec[0]; ++tt)
                    for (int uu = 0; uu <= ec[1]; ++uu)
                      for (int vv = 0; vv <= ec[2]; ++vv) {
                        const double w = Fx[tt] * Fy[uu] * Fz[vv];
                        const double r = R[((t + tt) * D + u + uu) * D + v + vv];
                        g += ((tt + uu + vv) & 1) ? -w * r : w * r;
                      }
                  s += e * g;
                }
            acc[kc] += f * s;
          }
        }
      }
    }
  if (act)
    for (int kc = 0; kc < NC; ++kc) cart[tid * NC + kc] = acc[kc];
  __syncthreads();
  const int nfa = A.nfunc, nfb = B.nfunc, nfc = P.nfunc;
  const bool diag = pr.x == pr.y;
  for (int o = tid; o < nfa * nfb * nfc; o += 64) {
    const int fa = o / (nfb * nfc), fb = (o / nfc) % nfb, fc = o % nfc;
    double s = 0.0;
    for (int x = 0; x < NA; ++x) {
      const double ta = trm[A.tr + fa * NA + x];
      if (ta == 0.0) continue;
      for (int y = 0; y < NB; ++y) {
        const double tb = trm[B.tr + fb * NB + y];
        if (tb == 0.0) continue;
        double q = 0.0;
        for (int z = 0; z < NC; ++z) q += trm[P.tr + fc * NC + z] * cart[(x * NB + y) * NC + z];
        s += ta * tb * q;
      }
    }
    const long m = A.func + fa, n = B.func + fb, Q = P.func + fc;
    out[m + ld * n + slab * Q] = s;
    if (!diag) out[n + ld * m + slab * Q] = s;
  }
}

template <int LA, int LB, int LC>
int launch_int3c(hipStream_t st, const DevShell* sh, const DevShell* ash, const double* prim,
                 const double* trm, const int2* pairs, int npairs, const int* ps, int np,
                 double* out, long ld, long slab) {
  const long total = (long)npairs * np;
  if (total <= 0) return DN_OK;
  if (total > 0x7fffffffL) return fail(DN_ERR_ARG, "three-centre integral grid too large");
  hipLaunchKernelGGL((int3c_kernel<LA, LB, LC>), dim3((unsigned)total), dim3(64), 0, st, sh, ash,
                     prim, trm, pairs, npairs, ps, np, out, ld, slab);
  DN_HIP(hipGetLastError());
  return DN_OK;
}

template <int LA, int LB>
int dispatch_lc(int lc, hipStream_t st, const DevShell* sh, const DevShell* ash, const double* prim,
                const double* trm, const int2* pairs, int npairs, const int* ps, int np,
                double* out, long ld, long slab) {
  switch (lc) {
    case 0: return launch_int3c<LA, LB, 0>(st, sh, ash, prim, trm, pairs, npairs, ps, np, out, ld, slab);
    case 1: return launch_int3c<LA, LB, 1>(st, sh, ash, prim, trm, pairs, npairs, ps, np, out, ld, slab);
    case 2: return launch_int3c<LA, LB, 2>(st, sh, ash, prim, trm, pairs, npairs, ps, np, out, ld, slab);
    case 3: return launch_int3c<LA, LB, 3>(st, sh, ash, prim, trm, pairs, npairs, ps, np, out, ld, slab);
    case 4: return launch_int3c<LA, LB, 4>(st, sh, ash, prim, trm, pairs, npairs, ps, np, out, ld, slab);
  }
  return fail(DN_ERR_ARG, "fitting shell angular momentum above 4");
}

int dispatch_int3c(int la, int lb, int lc, hipStream_t st, const DevShell* sh,
                   const DevShell* ash, const double* prim, const double* trm,
                   const int2* pairs, int npairs, const int* ps, int np, double* out, long ld,
                   long slab) {
#define DN_CASE(A_, B_)                                                                   \
  if (la == A_ && lb == B_)                                                               \
    return dispatch_lc<A_, B_>(lc, st, sh, ash, prim, trm, pairs, npairs, ps, np, out, ld, slab);
  DN_CASE(0, 0) DN_CASE(1, 0) DN_CASE(0, 1) DN_CASE(1, 1) DN_CASE(2, 0) DN_CASE(0, 2)
  DN_CASE(2, 1) DN_CASE(1, 2) DN_CASE(2, 2)
#undef DN_CASE
  return fail(DN_ERR_ARG, "orbital shell angular momentum above 2 on the device path");
}

}  // namespace

// (mn|P) on the device into dI (device, column-major with leading dimension ld per slab
// column and slab stride ld * nbf); the caller clears dI's pad rows.  Orbital shells up to
// d, fitting shells up to g.
int int3c2e_device(hipStream_t st, const dn_basis* b, const dn_basis* aux, double* dI, long ld) {
  if (!b || !aux || !dI) return fail(DN_ERR_ARG, "int3c2e_device: null argument");
  for (const auto& s : b->shells)
    if (s.l > kDevMaxLab) return fail(DN_ERR_ARG, "device three-centre integrals: orbital l > 2");
  for (const auto& s : aux->shells)
    if (s.l > kDevMaxLc) return fail(DN_ERR_ARG, "device three-centre integrals: fitting l > 4");
  std::vector<DevShell> hs, ha;
  std::vector<double> prim, trm;
  auto add = [&](const dn_basis* bs, std::vector<DevShell>& out) {
    for (const auto& s : bs->shells) {
      DevShell d;
      d.x = s.ctr[0];
      d.y = s.ctr[1];
      d.z = s.ctr[2];
      d.l = s.l;
      d.nprim = s.nprim;
      d.prim = (int)(prim.size() / 2);
      for (int i = 0; i < s.nprim; ++i) {
        prim.push_back(s.exps[i]);
        prim.push_back(s.coef[i]);
      }
      d.func = s.offset;
      d.nfunc = s.nfunc;
      d.tr = (int)trm.size();
      trm.insert(trm.end(), s.tr.begin(), s.tr.end());
      out.push_back(d);
    }
  };
  add(b, hs);
  add(aux, ha);
  // pairs (i >= j) by (la, lb), fitting shells by lc
  std::vector<int2> pairs[kDevMaxLab + 1][kDevMaxLab + 1];
  for (int i = 0; i < (int)hs.size(); ++i)
    for (int j = 0; j <= i; ++j) pairs[hs[i].l][hs[j].l].push_back(make_int2(i, j));
  std::vector<int> ps[kDevMaxLc + 1];
  for (int q = 0; q < (int)ha.size(); ++q) ps[ha[q].l].push_back(q);
  std::vector<int2> allp;
  std::vector<int> allq;
  size_t poff[kDevMaxLab + 1][kDevMaxLab + 1], qoff[kDevMaxLc + 1];
  for (int la = 0; la <= kDevMaxLab; ++la)
    for (int lb = 0; lb <= kDevMaxLab; ++lb) {
      poff[la][lb] = allp.size();
      allp.insert(allp.end(), pairs[la][lb].begin(), pairs[la][lb].end());
    }
  for (int lc = 0; lc <= kDevMaxLc; ++lc) {
    qoff[lc] = allq.size();
    allq.insert(allq.end(), ps[lc].begin(), ps[lc].end());
  }
  const size_t bytes = hs.size() * sizeof(DevShell) + ha.size() * sizeof(DevShell) +
                       prim.size() * 8 + trm.size() * 8 + allp.size() * sizeof(int2) +
                       allq.size() * sizeof(int) + 64 * 6;
  char* w = nullptr;
  DN_HIP(hipMallocAsync(reinterpret_cast<void**>(&w), bytes, st));
  char* q = w;
  auto put = [&](const void* src, size_t n) -> void* {
    void* dst = q;
    if (n) (void)hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st);
    q += (n + 63) / 64 * 64;
    return dst;
  };
  auto* dsh = static_cast<const DevShell*>(put(hs.data(), hs.size() * sizeof(DevShell)));
  auto* dash = static_cast<const DevShell*>(put(ha.data(), ha.size() * sizeof(DevShell)));
  auto* dprim = static_cast<const double*>(put(prim.data(), prim.size() * 8));
  auto* dtrm = static_cast<const double*>(put(trm.data(), trm.size() * 8));
  auto* dpairs = static_cast<const int2*>(put(allp.data(), allp.size() * sizeof(int2)));
  auto* dq = static_cast<const int*>(put(allq.data(), allq.size() * sizeof(int)));
  int rc = DN_OK;
  const long slab = ld * b->nbf;
  for (int la = 0; la <= kDevMaxLab && rc == DN_OK; ++la)
    for (int lb = 0; lb <= kDevMaxLab && rc == DN_OK; ++lb)
      for (int lc = 0; lc <= kDevMaxLc && rc == DN_OK; ++lc) {
        const int npr = (int)pairs[la][lb].size(), nq = (int)ps[lc].size();
        if (!npr || !nq) continue;
        rc = dispatch_int3c(la, lb, lc, st, dsh, dash, dprim, dtrm, dpairs + poff[la][lb], npr,
                            dq + qoff[lc], nq, dI, ld, slab);
      }
  // the host vectors are read by the copies: wait for them before they are freed
  const hipError_t e1 = hipStreamSynchronize(st);
  (void)hipFreeAsync(w, st);
  const hipError_t e2 = hipStreamSynchronize(st);
  if (rc != DN_OK) return rc;
  if (e1 != hipSuccess || e2 != hipSuccess)
    return fail(DN_ERR_HIP, std::string("int3c2e_device: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
  return DN_OK;
}

}  // namespace dn

using namespace dn;

extern "C" {

int dn_basis_create(int nshell, const int* l, const int* nprim, const int* pure,
                    const double* centers, const double* exps, const double* coefs,
                    dn_basis** out) {
  if (!out || nshell < 0 || (nshell > 0 && (!l || !nprim || !pure || !centers || !exps || !coefs)))
    return fail(DN_ERR_ARG, "dn_basis_create: null argument");
  *out = nullptr;
  dn_basis* b = new dn_basis();
  long pofs = 0;
  for (int s = 0; s < nshell; ++s) {
    if (l[s] < 0 || l[s] > kMaxL || nprim[s] < 1) {
      delete b;
      return fail(DN_ERR_ARG, "dn_basis_create: shell angular momentum outside 0..6 or no primitives");
    }
    dn_basis::Shell sh;
    sh.l = l[s];
    sh.pure = pure[s] ? 1 : 0;
    sh.nprim = nprim[s];
    sh.ncart = ncart_of(sh.l);
    for (int d = 0; d < 3; ++d) sh.ctr[d] = centers[3 * s + d];
    // primitives normalised as x^l exp(-a r^2), then the contraction normalised to one
    const int L = sh.l;
    double norm2 = 0.0;
    for (int i = 0; i < sh.nprim; ++i) {
      const double a = exps[pofs + i];
      if (!(a > 0.0)) {
        delete b;
        return fail(DN_ERR_ARG, "dn_basis_create: exponents must be positive");
      }
      sh.exps.push_back(a);
      sh.coef.push_back(coefs[pofs + i] * std::pow(2.0 * a / M_PI, 0.75) *
                        std::pow(4.0 * a, 0.5 * L) / std::sqrt(dfact(2 * L - 1)));
    }
    for (int i = 0; i < sh.nprim; ++i)
      for (int j = 0; j < sh.nprim; ++j) {
        const double p = sh.exps[i] + sh.exps[j];
        norm2 += sh.coef[i] * sh.coef[j] * std::pow(M_PI / p, 1.5) * dfact(2 * L - 1) /
                 std::pow(2.0 * p, L);
      }
    for (double& c : sh.coef) c /= std::sqrt(norm2);
    pofs += sh.nprim;
    shell_transform(sh);
    sh.offset = b->nbf;
    b->nbf += sh.nfunc;
    b->lmax = std::max(b->lmax, sh.l);
    b->shells.push_back(std::move(sh));
  }
  *out = b;
  return DN_OK;
}

int dn_basis_destroy(dn_basis* b) {
  delete b;
  return DN_OK;
}

int dn_basis_nbf(const dn_basis* b, int* nbf) {
  if (!b || !nbf) return fail(DN_ERR_ARG, "dn_basis_nbf: null argument");
  *nbf = b->nbf;
  return DN_OK;
}

int dn_int1e(const dn_basis* b, int natom, const double* Z, const double* xyz, double* S,
             double* T, double* V) {
  if (!b || natom < 0 || (natom > 0 && (!Z || !xyz)))
    return fail(DN_ERR_ARG, "dn_int1e: null argument");
  const int ns = (int)b->shells.size();
  const long nbf = b->nbf, npair = (long)ns * (ns + 1) / 2;
  parallel_for(npair, [&](long k) {
    int i = (int)((std::sqrt(8.0 * k + 1.0) - 1.0) / 2.0);
    while ((long)i * (i + 1) / 2 > k) --i;
    while ((long)(i + 1) * (i + 2) / 2 <= k) ++i;
    const int j = (int)(k - (long)i * (i + 1) / 2);
    const auto& A = b->shells[i];
    const auto& B = b->shells[j];
    std::vector<double> cart((size_t)A.ncart * B.ncart), w;
    if (S) {
      cart_overlap(A, B, cart.data());
      put_pair(A, B, cart.data(), S, nbf, true);
    }
    if (T) {
      cart_kinetic(A, B, cart.data());
      put_pair(A, B, cart.data(), T, nbf, true);
    }
    if (V) {
      cart_nuclear(A, B, natom, Z, xyz, cart.data(), w);
      put_pair(A, B, cart.data(), V, nbf, true);
    }
  });
  return DN_OK;
}

// (P|Q) over the auxiliary basis, naux x naux column-major (ERI_2e2c)
int dn_int2c2e(const dn_basis* aux, double* G) {
  if (!aux || !G) return fail(DN_ERR_ARG, "dn_int2c2e: null argument");
  const int ns = (int)aux->shells.size();
  const long n = aux->nbf, npair = (long)ns * (ns + 1) / 2;
  parallel_for(npair, [&](long k) {
    int i = (int)((std::sqrt(8.0 * k + 1.0) - 1.0) / 2.0);
    while ((long)i * (i + 1) / 2 > k) --i;
    while ((long)(i + 1) * (i + 2) / 2 <= k) ++i;
    const int j = (int)(k - (long)i * (i + 1) / 2);
    const auto& P = aux->shells[i];
    const auto& Q = aux->shells[j];
    const dn_basis::Shell up = unit_shell(P.ctr), uq = unit_shell(Q.ctr);
    std::vector<double> cart((size_t)P.ncart * Q.ncart), w;
    cart_eri(P, up, Q, uq, cart.data(), w);
    put_pair(P, Q, cart.data(), G, n, true);
  });
  return DN_OK;
}

// (mn|P), nbf x nbf x naux column-major (ERI_2e3c)
int dn_int3c2e(const dn_basis* b, const dn_basis* aux, double* I3c) {
  if (!b || !aux || !I3c) return fail(DN_ERR_ARG, "dn_int3c2e: null argument");
  const int ns = (int)b->shells.size(), na = (int)aux->shells.size();
  const long nbf = b->nbf, npair = (long)ns * (ns + 1) / 2;
  parallel_for(npair * na, [&](long k) {
    const long pr = k / na;
    const int q = (int)(k % na);
    int i = (int)((std::sqrt(8.0 * pr + 1.0) - 1.0) / 2.0);
    while ((long)i * (i + 1) / 2 > pr) --i;
    while ((long)(i + 1) * (i + 2) / 2 <= pr) ++i;
    const int j = (int)(pr - (long)i * (i + 1) / 2);
    const auto& A = b->shells[i];
    const auto& B = b->shells[j];
    const auto& P = aux->shells[q];
    const dn_basis::Shell up = unit_shell(P.ctr);
    std::vector<double> cart((size_t)A.ncart * B.ncart * P.ncart), w;
    cart_eri(A, B, P, up, cart.data(), w);
    // functions: (fa fb | fp) = sum Ta Tb Tp cart
    std::vector<double> tmp((size_t)A.ncart * B.ncart);
    for (int fp = 0; fp < P.nfunc; ++fp) {
      for (size_t ab = 0; ab < tmp.size(); ++ab) {
        double s = 0.0;
        for (int z = 0; z < P.ncart; ++z) s += cart[ab * P.ncart + z] * P.tr[(size_t)fp * P.ncart + z];
        tmp[ab] = s;
      }
      double* slab = I3c + (long)(P.offset + fp) * nbf * nbf;
      put_pair(A, B, tmp.data(), slab, nbf, true);
    }
  });
  return DN_OK;
}

// (mn|ls), nbf^4 column-major (ERI_2e4c)
int dn_int4c2e(const dn_basis* b, double* eri) {
  if (!b || !eri) return fail(DN_ERR_ARG, "dn_int4c2e: null argument");
  const int ns = (int)b->shells.size();
  const long nbf = b->nbf, npair = (long)ns * (ns + 1) / 2;
  parallel_for(npair * npair, [&](long k) {
    const long p1 = k / npair, p2 = k % npair;
    if (p2 > p1) return;
    auto unpack = [](long pr, int* i, int* j) {
      int a = (int)((std::sqrt(8.0 * pr + 1.0) - 1.0) / 2.0);
      while ((long)a * (a + 1) / 2 > pr) --a;
      while ((long)(a + 1) * (a + 2) / 2 <= pr) ++a;
      *i = a;
      *j = (int)(pr - (long)a * (a + 1) / 2);
    };
    int i, j, kk, l;
    unpack(p1, &i, &j);
    unpack(p2, &kk, &l);
    const auto& A = b->shells[i];
    const auto& B = b->shells[j];
    const auto& C = b->shells[kk];
    const auto& D = b->shells[l];
    std::vector<double> cart((size_t)A.ncart * B.ncart * C.ncart * D.ncart), w;
    cart_eri(A, B, C, D, cart.data(), w);
    const int nc = C.ncart, nd = D.ncart;
    std::vector<double> half((size_t)A.ncart * B.ncart * C.nfunc * D.nfunc);
    for (size_t ab = 0; ab < (size_t)A.ncart * B.ncart; ++ab)
      for (int fc = 0; fc < C.nfunc; ++fc)
        for (int fd = 0; fd < D.nfunc; ++fd) {
          double s = 0.0;
          for (int z = 0; z < nc; ++z) {
            const double tc = C.tr[(size_t)fc * nc + z];
            if (tc == 0.0) continue;
            for (int y = 0; y < nd; ++y)
              s += tc * cart[(ab * nc + z) * nd + y] * D.tr[(size_t)fd * nd + y];
          }
          half[(ab * C.nfunc + fc) * D.nfunc + fd] = s;
        }
    for (int fa = 0; fa < A.nfunc; ++fa)
      for (int fb = 0; fb < B.nfunc; ++fb)
        for (int fc = 0; fc < C.nfunc; ++fc)
          for (int fd = 0; fd < D.nfunc; ++fd) {
            double s = 0.0;
            for (int x = 0; x < A.ncart; ++x) {
              const double ta = A.tr[(size_t)fa * A.ncart + x];
              if (ta == 0.0) continue;
              for (int y = 0; y < B.ncart; ++y)
                s += ta * B.tr[(size_t)fb * B.ncart + y] *
                     half[(((size_t)x * B.ncart + y) * C.nfunc + fc) * D.nfunc + fd];
            }
            const long m = A.offset + fa, n = B.offset + fb, r = C.offset + fc, q = D.offset + fd;
            const long idx[8][4] = {{m, n, r, q}, {n, m, r, q}, {m, n, q, r}, {n, m, q, r},
                                    {r, q, m, n}, {q, r, m, n}, {r, q, n, m}, {q, r, n, m}};
            for (auto& e : idx) eri[e[0] + nbf * (e[1] + nbf * (e[2] + nbf * e[3]))] = s;
          }
  });
  return DN_OK;
}

// d/dR of the one-electron integrals for the nuclear position R of atom `atom` (bohr):
// the basis functions on that atom move, and for dV its nucleus moves too
// (translational invariance: d/dC = -(d/dA + d/dB) for each nucleus C); nbf x nbf x 3
// each (the Cartesian component last); a NULL output is skipped.  shell_atom[s] is the
// atom of shell s (0-based).
int dn_int1e_deriv(const dn_basis* b, const int* shell_atom, int natom, const double* Z,
                   const double* xyz, int atom, double* dS, double* dT, double* dV) {
  if (!b || !shell_atom || natom <= 0 || !Z || !xyz || atom < 0 || atom >= natom)
    return fail(DN_ERR_ARG, "dn_int1e_deriv: bad argument");
  const int ns = (int)b->shells.size();
  const long nbf = b->nbf, nn = nbf * nbf;
  for (double* o : {dS, dT, dV})
    if (o) std::fill(o, o + 3 * nn, 0.0);
  parallel_for((long)ns * ns, [&](long k) {
    const int i = (int)(k / ns), j = (int)(k % ns);
    const auto& A = b->shells[i];
    const auto& B = b->shells[j];
    const bool ai = shell_atom[i] == atom, aj = shell_atom[j] == atom;
    const size_t nb = B.ncart, na = A.ncart;
    std::vector<double> d(3 * na * nb), cart(na * nb), w;
    auto put = [&](double* M, double f) {  // M[.., .., k] += f * (Ta d_k Tb^T)
      for (int kk = 0; kk < 3; ++kk) {
        for (size_t c = 0; c < na * nb; ++c) cart[c] = f * d[kk * na * nb + c];
        for (int fa = 0; fa < A.nfunc; ++fa)
          for (int fb = 0; fb < B.nfunc; ++fb) {
            double s = 0.0;
            for (size_t x = 0; x < na; ++x) {
              const double ta = A.tr[(size_t)fa * na + x];
              if (ta == 0.0) continue;
              for (size_t y = 0; y < nb; ++y) s += ta * cart[x * nb + y] * B.tr[(size_t)fb * nb + y];
            }
            M[(A.offset + fa) + nbf * (B.offset + fb) + nn * kk] += s;
          }
      }
    };
    // two-centre integrals: d/dB = -d/dA, so atom-pairs move only when exactly one is on atom
    if ((dS || dT) && ai != aj) {
      const double f = ai ? 1.0 : -1.0;
      if (dS) {
        deriv_center(A, nb, [&](const dn_basis::Shell& X, double* o) { cart_overlap(X, B, o); }, d.data());
        put(dS, f);
      }
      if (dT) {
        deriv_center(A, nb, [&](const dn_basis::Shell& X, double* o) { cart_kinetic(X, B, o); }, d.data());
        put(dT, f);
      }
    }
    if (dV) {
      // per nucleus c: (a_i - [c == atom]) dA V^c + (a_j - [c == atom]) dB V^c, with
      // dB V^c(A, B) the transpose of dA V^c(B, A)
      std::vector<double> dA(3 * na * nb), dBt(3 * nb * na);
      for (int c = 0; c < natom; ++c) {
        const double fc = c == atom ? 1.0 : 0.0;
        const double wa = (ai ? 1.0 : 0.0) - fc, wb = (aj ? 1.0 : 0.0) - fc;
        if (wa == 0.0 && wb == 0.0) continue;
        if (wa != 0.0)
          deriv_center(A, nb, [&](const dn_basis::Shell& X, double* o) {
            cart_nuclear(X, B, 1, Z + c, xyz + 3 * c, o, w);
          }, dA.data());
        if (wb != 0.0)
          deriv_center(B, na, [&](const dn_basis::Shell& X, double* o) {
            cart_nuclear(X, A, 1, Z + c, xyz + 3 * c, o, w);
          }, dBt.data());
        for (int kk = 0; kk < 3; ++kk)
          for (size_t x = 0; x < na; ++x)
            for (size_t y = 0; y < nb; ++y)
              d[(kk * na + x) * nb + y] = (wa != 0.0 ? wa * dA[(kk * na + x) * nb + y] : 0.0) +
                                          (wb != 0.0 ? wb * dBt[(kk * nb + y) * na + x] : 0.0);
        put(dV, 1.0);
      }
    }
  });
  return DN_OK;
}

// d/dR of (P|Q) for atom `atom`: naux x naux x 3 (d/dQ = -d/dP)
int dn_int2c2e_deriv(const dn_basis* aux, const int* aux_atom, int atom, double* dG) {
  if (!aux || !aux_atom || !dG || atom < 0) return fail(DN_ERR_ARG, "dn_int2c2e_deriv: bad argument");
  const int ns = (int)aux->shells.size();
  const long n = aux->nbf, nn = n * n;
  std::fill(dG, dG + 3 * nn, 0.0);
  parallel_for((long)ns * ns, [&](long k) {
    const int i = (int)(k / ns), j = (int)(k % ns);
    const bool ai = aux_atom[i] == atom, aj = aux_atom[j] == atom;
    if (ai == aj) return;
    const auto& P = aux->shells[i];
    const auto& Q = aux->shells[j];
    const dn_basis::Shell up = unit_shell(P.ctr), uq = unit_shell(Q.ctr);
    const size_t np = P.ncart, nq = Q.ncart;
    std::vector<double> d(3 * np * nq), w;
    deriv_center(P, nq, [&](const dn_basis::Shell& X, double* o) { cart_eri(X, up, Q, uq, o, w); },
                 d.data());
    const double f = ai ? 1.0 : -1.0;
    for (int kk = 0; kk < 3; ++kk)
      for (int fp = 0; fp < P.nfunc; ++fp)
        for (int fq = 0; fq < Q.nfunc; ++fq) {
          double s = 0.0;
          for (size_t x = 0; x < np; ++x) {
            const double tp = P.tr[(size_t)fp * np + x];
            if (tp == 0.0) continue;
            for (size_t y = 0; y < nq; ++y) s += tp * d[(kk * np + x) * nq + y] * Q.tr[(size_t)fq * nq + y];
          }
          dG[(P.offset + fp) + n * (Q.offset + fq) + nn * kk] = f * s;
        }
  });
  return DN_OK;
}

// d/dR of (mn|P) for atom `atom`: nbf x nbf x naux x 3 (d/dP = -(d/dA + d/dB))
int dn_int3c2e_deriv(const dn_basis* b, const int* shell_atom, const dn_basis* aux,
                     const int* aux_atom, int atom, double* dI) {
  if (!b || !aux || !shell_atom || !aux_atom || !dI || atom < 0)
    return fail(DN_ERR_ARG, "dn_int3c2e_deriv: bad argument");
  const int ns = (int)b->shells.size(), nas = (int)aux->shells.size();
  const long nbf = b->nbf, naux = aux->nbf, slab = nbf * nbf, comp = slab * naux;
  std::fill(dI, dI + 3 * comp, 0.0);
  parallel_for((long)ns * ns * nas, [&](long k) {
    const int q = (int)(k % nas);
    const int i = (int)((k / nas) / ns), j = (int)((k / nas) % ns);
    const auto& A = b->shells[i];
    const auto& B = b->shells[j];
    const auto& P = aux->shells[q];
    const double fp_ = aux_atom[q] == atom ? 1.0 : 0.0;
    const double wa = (shell_atom[i] == atom ? 1.0 : 0.0) - fp_;
    const double wb = (shell_atom[j] == atom ? 1.0 : 0.0) - fp_;
    if (wa == 0.0 && wb == 0.0) return;
    const dn_basis::Shell u = unit_shell(P.ctr);
    const size_t na = A.ncart, nb = B.ncart, np = P.ncart;
    std::vector<double> dA, dB, w;
    if (wa != 0.0) {
      dA.resize(3 * na * nb * np);
      deriv_center(A, nb * np, [&](const dn_basis::Shell& X, double* o) { cart_eri(X, B, P, u, o, w); },
                   dA.data());
    }
    if (wb != 0.0) {
      dB.resize(3 * nb * na * np);
      deriv_center(B, na * np, [&](const dn_basis::Shell& X, double* o) { cart_eri(X, A, P, u, o, w); },
                   dB.data());
    }
    for (int kk = 0; kk < 3; ++kk)
      for (int fa = 0; fa < A.nfunc; ++fa)
        for (int fb = 0; fb < B.nfunc; ++fb)
          for (int fq = 0; fq < P.nfunc; ++fq) {
            double s = 0.0;
            for (size_t x = 0; x < na; ++x) {
              const double ta = A.tr[(size_t)fa * na + x];
              if (ta == 0.0) continue;
              for (size_t y = 0; y < nb; ++y) {
                const double tb = B.tr[(size_t)fb * nb + y];
                if (tb == 0.0) continue;
                for (size_t z = 0; z < np; ++z) {
                  double v = 0.0;
                  if (wa != 0.0) v += wa * dA[((kk * na + x) * nb + y) * np + z];
                  if (wb != 0.0) v += wb * dB[((kk * nb + y) * na + x) * np + z];
                  s += ta * tb * P.tr[(size_t)fq * np + z] * v;
                }
              }
            }
            dI[(A.offset + fa) + nbf * (B.offset + fb) + slab * (P.offset + fq) + comp * kk] = s;
          }
  });
  return DN_OK;
}

}  // extern "C"
