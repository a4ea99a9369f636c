// Occupation-functional kernels (dn_occ.hip): sizes, functional options and the
// per-orbital work arrays they share.
#pragma once

#include "dn_internal.h"

namespace dn {

struct OccDims {
  int nbf5, no1, ndoc, nsoc, ndns, nbeta, nalpha, ncwo, nv;
  int ipnof, ista, occ_method, highspin, mspin;
  double h_cut;
};

// Device arrays of length nbf5 (n, fi, nd, nd12, dfic, ndc, ndx, strong, pair, role) and
// scal[0] = S_F (sum of holes of the first ndoc orbitals, PNOF4).
struct OccWork {
  double* n = nullptr;
  double* fi = nullptr;
  double* nd = nullptr;
  double* nd12 = nullptr;
  double* dfic = nullptr;
  double* ndc = nullptr;
  double* ndx = nullptr;
  int* strong = nullptr;
  int* pair = nullptr;
  int* role = nullptr;
  double* scal = nullptr;
};

int occ_ocupacion(hipStream_t s, const OccDims& P, const double* gamma, double* n, double* dn);
int occ_prepare(hipStream_t s, const OccDims& P, const OccWork& w);
int occ_coefficients(hipStream_t s, const OccDims& P, const OccWork& w, double* cj12,
                     double* ck12);
int occ_der_coefficients(hipStream_t s, const OccDims& P, const OccWork& w, const double* dn,
                         double* Dcj, double* Dck);
int occ_calcoccg(hipStream_t s, const OccDims& P, const OccWork& w, const double* dn,
                 const double* J, const double* K, const double* Hcore, double* grad);
int occ_calce(hipStream_t s, const OccDims& P, const double* n, const double* cj12,
              const double* ck12, const double* J, const double* K, const double* Hcore,
              double* out);

}  // namespace dn
