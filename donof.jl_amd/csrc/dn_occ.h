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

// State of the occupation optimiser's nonlinear conjugate gradient on the device (the
// iteration of minimization.jl:100-139's optimize(..., ConjugateGradient()) with the
// Hager-Zhang direction and a Wolfe line search); every field is written by the
// controller at the start of an evaluation slot.  Vectors (nv each, `vec`): x, g (at x),
// d, the trial point xt, the gradient at xt, the gradient at the best zoom point.
struct OccCG {
  int state, done, converged, it, ls_i, zoom_i, maxiter, best_valid, nfev, pad;
  double tol, f, f0, d0, a, a_prev, f_prev, d_prev, alo, flo, dlo, ahi, fhi, dhi;
  double best_a, best_f, ft;
};
enum { kCgX = 0, kCgG = 1, kCgD = 2, kCgXT = 3, kCgGT = 4, kCgGB = 5, kCgVecs = 6 };
// nslots evaluation slots (each: the controller's next trial point and its occupations
// in one block, then calcoccg's gradient and slices of the energy at it); slots after
// convergence exit at once.  vec holds kCgVecs * nv doubles and then the energy slices
// (occ_cg_energy_blocks of them).
int occ_cg_energy_blocks(const OccDims& P);
int occ_cg_slots(hipStream_t s, const OccDims& P, const OccWork& w, OccCG* c, double* vec,
                 double* dn, const double* J, const double* K, const double* Hcore,
                 int nslots);

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
