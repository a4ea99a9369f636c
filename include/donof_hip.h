/* libdonofhip: MI355X (gfx950) implementation of DoNOF.jl's orbital-optimization hot
 * path behind a C ABI.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Host arrays are column-major exactly as Julia
 *    stores them: b[m,n,Q] at m + nbf*n + nbf*nbf*Q, C[m,p] at m + nbf*p.
 *  - Host pointers are borrowed for the duration of the call; outputs go to
 *    caller-owned buffers.  The library owns every device buffer of a dn_sys.
 *  - Every entry point returns 0 (DN_OK) on success or a DN_ERR_* code; the message of
 *    the last failure on the calling thread is returned by dn_last_error().
 *  - A dn_sys is used by one host thread at a time and runs on its own HIP stream.
 *
 * Each entry point names the reference interface it replaces (file:line in DoNOF.jl
 * v1.1.2).  The reference binds its GPU path by Julia dispatch on CuArray
 * (ext/DoNOFGPUExt.jl); the ccall bindings that replace it are in INTEGRATION.md.
 */
#ifndef DONOF_HIP_H
#define DONOF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DN_OK = 0,
  DN_ERR_ARG = 1,       /* invalid argument or size */
  DN_ERR_HIP = 2,       /* HIP runtime error */
  DN_ERR_NOMEM = 3,     /* device allocation failed */
  DN_ERR_BUDGET = 4,    /* allocation would exceed the context's HBM budget */
  DN_ERR_NONFINITE = 5, /* NaN/Inf found in an output */
  DN_ERR_STATE = 6      /* call not valid in the current state (e.g. b not uploaded) */
};

/* Arithmetic of the RI transforms and contractions of a system.
 * DN_SPLIT_F16: b and C5 as FP16 hi+lo pairs with power-of-two scales, three MFMA
 * products accumulated in FP32: the J/K build's two transforms (J/K reduced in FP64) and
 * the Lagrangian's two transforms and contraction sum_Q bpn bmat (FP32 partials summed in
 * FP64); H_core, the Lagrangian's one-electron and C^J/C^K terms (evaluated in FP64 from
 * the FP32 transforms), the rotation and the occupation functional stay FP64.  b, b_pq
 * and the contraction's right operand are not kept in FP64. */
enum { DN_FP64 = 0, DN_SPLIT_F16 = 1 };

/* Orbital optimiser of the ADAM family: p.orb_method "ADAM", "ADABelief", "YOGI",
 * "Demon" (minimization.jl:242-630). */
enum { DN_ADAM = 0, DN_ADABELIEF = 1, DN_YOGI = 2, DN_DEMON = 3 };

/* Occupation parametrisation: p.occ_method "Softmax", "Trigonometric" (pnof.jl:1024-1045). */
enum { DN_OCC_SOFTMAX = 0, DN_OCC_TRIGONOMETRIC = 1 };

typedef struct dn_ctx dn_ctx;
typedef struct dn_sys dn_sys;

/* Sizes of one system; field meaning as DoNOF's Param (param.jl:3-57). */
typedef struct dn_dims {
  int32_t nbf;    /* AO basis size */
  int32_t nbfaux; /* RI auxiliary basis size */
  int32_t nbf5;   /* active natural orbitals */
  int32_t no1;    /* frozen core */
  int32_t ndoc;   /* strongly doubly occupied */
  int32_t nsoc;   /* singly occupied */
  int32_t nbeta;
  int32_t nalpha;
  int32_t ncwo;   /* weak orbitals per pair */
  int32_t nv;     /* occupation variables = ncwo*ndoc */
} dn_dims;

const char* dn_last_error(void);
int dn_device_count(int* count);

/* Device context with an HBM budget (bytes; 0 = the device's free memory). */
int dn_ctx_create(int device, uint64_t hbm_budget_bytes, dn_ctx** out);
int dn_ctx_destroy(dn_ctx* ctx);
int dn_ctx_hbm_used(const dn_ctx* ctx, uint64_t* used_bytes);
int dn_ctx_hbm_budget(const dn_ctx* ctx, uint64_t* budget_bytes);

/* One resident system: b, C, H, occupations, J/K/Lagrangian outputs and workspaces. */
int dn_sys_create(dn_ctx* ctx, const dn_dims* dims, int precision, dn_sys** out);
int dn_sys_destroy(dn_sys* sys);
/* Bytes dn_sys_create would reserve for these sizes (admission control). */
int dn_sys_footprint(const dn_dims* dims, int precision, uint64_t* bytes);

/* b becomes device-resident once: replaces eris_to_gpu (ext/DoNOFGPUExt.jl:36-50,
 * called from integrals.jl:59-62). */
int dn_sys_upload_b(dn_sys* sys, const double* b);
int dn_sys_download_b(dn_sys* sys, double* b);
/* Synthetic b generated in HBM (Philox4x32-10 normals, see donof.jl_amd/synthetic.py). */
int dn_sys_generate_b(dn_sys* sys, uint64_t seed);
/* RI setup on the device: b = I3c G^{-1/2} from the 2-centre metric G (naux x naux,
 * column-major, lower triangle read) and the 3-centre integrals I3c (column-major
 * nbf x nbf x naux); replaces Iaux (integrals.jl:68-91) and Iaux_gpu
 * (ext/DoNOFGPUExt.jl:9-33).  FP64 blocked Cholesky G = L L^T and b = I3c L^{-T}: b
 * equals the reference's eigen-route b up to an orthogonal rotation of the auxiliary
 * index, which leaves J, K, the Lagrangian and every energy unchanged.  A G that is
 * not positive definite takes the reference's eigen route (rocSOLVER syevd of the upper
 * triangle, integrals.jl:74-83: 1/sqrt(l) for eigenvalues l >= 0, 0 below): b is then
 * the reference's b itself; with an eigenvalue below zero no RI factor is kept for the
 * nuclear gradients.  Host arrays are not modified; a transient device workspace of the
 * size of b is taken for the call. */
int dn_sys_ri_metric(dn_sys* sys, const double* G, const double* I3c);
/* RI nuclear gradients: replaces compute_geom_gradients' RI branch (grads.jl:1-100).
 * dn_grad_create forms on the device, from the resident b at orbitals C (nbf x nbf),
 * occupations n (nbf5), cj12, ck12 (nbf5 x nbf5; their diagonals are ignored, as
 * grads.jl:24-25 zero them; the caller's arrays are not modified) and the Lagrangian
 * elag (nbf x nbf): RDM1 = 2 C5 diag(n) C5^T (grads.jl:5), lag = 2 C elag C^T (6), the
 * three-index density d[l,s,Q] (66-84) and the metric density m[Q,R] (89).  IG =
 * I3c G^{-1} is taken as b Wt with the RI factor Wt of the resident b: dn_sys_ri_metric
 * keeps it (b = I3c L^{-T}, Wt = L^{-1}); for an uploaded b = I3c W give it with
 * dn_sys_set_ri_factor (Wt = W^{-1} G^{-1}, naux x naux, column-major: G^{-1/2} for the
 * reference's eigen-route b of integrals.jl:68-91; setting b again drops it).  The
 * object holds nbf^2 naux + naux^2 + 2 nbf^2 doubles of HBM (charged to the context's
 * budget) and keeps the system from being destroyed.
 * dn_grad_atom returns one atom's electronic gradient (grads.jl:13-21, 91-96)
 *   grad3[k] = -<lag, dS_k> + <RDM1, dT_k + dV_k> + 2 <d, geri3c_k> - <m, geri2c_k>
 * from the derivative integrals of that atom (dS, dT, dV: nbf x nbf x 3; geri3c:
 * nbf x nbf x naux x 3; geri2c: naux x naux x 3; column-major, Cartesian index last),
 * each a host or a device pointer (device arrays are read in place, host arrays streamed
 * through a staging buffer).  The nuclear-repulsion term compute_grad_nuc
 * (utils.jl:315-336) is the caller's.  dn_grad_get_densities copies out any of RDM1,
 * lag, m, d (NULL skips). */
typedef struct dn_grad dn_grad;
int dn_sys_set_ri_factor(dn_sys* sys, const double* Wt);
int dn_grad_create(dn_sys* sys, const double* C, const double* n, const double* cj12,
                   const double* ck12, const double* elag, dn_grad** out);
int dn_grad_atom(dn_grad* grad, const double* dS, const double* dT, const double* dV,
                 const double* geri3c, const double* geri2c, double* grad3);
int dn_grad_get_densities(dn_grad* grad, double* RDM1, double* lag, double* m, double* d);
int dn_grad_destroy(dn_grad* grad);

/* Whenever b is set, the library checks b[m,n,Q] == b[n,m,Q] (to 2^-50 of max|b|; the
 * RI tensor is symmetric in m, n, integrals.jl:86-88).  A symmetric b lets the FP64
 * J/K build contract only its lower half: C5^T b_Q C5 = W + W^T with W = C5^T Lh_Q C5,
 * b_Q = Lh_Q + Lh_Q^T.  use = 0 forces the full contraction of integrals.jl:159-160. */
int dn_sys_set_symmetry(dn_sys* sys, int use);
int dn_sys_b_symmetric(dn_sys* sys, int* is_symmetric);

/* J_MO, K_MO (nbf5 x nbf5) and H_core (nbf5) from C (nbf x nbf) and H (nbf x nbf):
 * replaces computeJKH_MO / JKH_MO (integrals.jl:131-172; ext/DoNOFGPUExt.jl:52-94). */
int dn_jkh_mo(dn_sys* sys, const double* C, const double* H, double* J, double* K,
              double* Hcore);

/* elag (nbf x nbf, columns >= nbf5 zero) and H_mat = C^T H C5 (nbf x nbf5):
 * replaces compute_Lagrange2 (utils.jl:1-59; ext/DoNOFGPUExt.jl:150-245).
 * E_elec of computeE_elec(Hmat, n, elag) (utils.jl:216-227) is returned in *E if E is
 * not NULL. */
int dn_lagrange2(dn_sys* sys, const double* C, const double* n, const double* H,
                 const double* cj12, const double* ck12, int nalpha, int nbeta,
                 double* elag, double* Hmat, double* E);

/* The diagnostics of one outer iteration of energy (energy.jl:172-190: ENERGY1r's
 * Lagrangian, and computeJKH_MO at the same orbitals) from one pass over b: dn_lagrange2
 * plus J, K (nbf5 x nbf5) and Hcore (nbf5) taken from the Lagrangian's own transform
 * b_pq = C5^T b C5 and from H_mat, instead of a second J/K-MO build
 * (integrals.jl:131-172).  Hmat may be NULL. */
int dn_lagrange2_jkh(dn_sys* sys, const double* C, const double* n, const double* H,
                     const double* cj12, const double* ck12, int nalpha, int nbeta,
                     double* elag, double* Hmat, double* E, double* J, double* K,
                     double* Hcore);

/* Resident variants: inputs already in HBM (set once), outputs stay in HBM.  They
 * enqueue on the system's stream and return without synchronising. */
int dn_sys_set_orbitals(dn_sys* sys, const double* C);
int dn_sys_set_hcore(dn_sys* sys, const double* H);
int dn_jkh_mo_resident(dn_sys* sys);
int dn_sys_get_jkh(dn_sys* sys, double* J, double* K, double* Hcore);
int dn_sys_sync(dn_sys* sys);

/* Per-launch timing of the dominant kernel (the first RI transform) with HIP events on
 * the system's stream. */
int dn_sys_set_timing(dn_sys* sys, int enable);
int dn_sys_get_timing(dn_sys* sys, double* transform_ms_total, long* launches);

/* ---- occupation functional --------------------------------------------------------- */

/* The functional options of Param read by the occupation code (param.jl:26-27,49-51,56). */
typedef struct dn_functional {
  int32_t ipnof;      /* p.ipnof: 4, 5, 7, 8 (GNOF) or 9 */
  int32_t ista;       /* p.ista */
  int32_t occ_method; /* DN_OCC_SOFTMAX or DN_OCC_TRIGONOMETRIC */
  int32_t highspin;   /* p.HighSpin */
  int32_t mspin;      /* p.MSpin */
  double h_cut;       /* p.h_cut */
} dn_functional;

/* n (nbf5) and dn/dgamma (nbf5 x nv) from gamma (nv): ocupacion (pnof.jl:1024-1163). */
int dn_ocupacion(dn_sys* sys, const dn_functional* f, const double* gamma, double* n,
                 double* dn_dgamma);
/* cj12, ck12 (nbf5 x nbf5, zero diagonal): PNOFi_selector (pnof.jl:1-53). */
int dn_pnof_coefficients(dn_sys* sys, const dn_functional* f, const double* n, double* cj12,
                         double* ck12);
/* Dcj12r, Dck12r (nbf5 x nbf5 x nv): der_PNOFi_selector (pnof.jl:55-117). */
int dn_der_pnof_coefficients(dn_sys* sys, const dn_functional* f, const double* n,
                             const double* dn_dgamma, double* Dcj12r, double* Dck12r);
/* Occupation energy: calce (pnof.jl:1273-1300).  J, K, Hcore NULL = the resident ones
 * of the last J/K build (all three given or all NULL). */
int dn_calce(dn_sys* sys, const dn_functional* f, const double* n, const double* cj12,
             const double* ck12, const double* J, const double* K, const double* Hcore,
             double* E);
/* calcocce (pnof.jl:1302-1322) and calcoccg (pnof.jl:1324-1377). */
int dn_calcocce(dn_sys* sys, const dn_functional* f, const double* gamma, const double* J,
                const double* K, const double* Hcore, double* E);
int dn_calcoccg(dn_sys* sys, const dn_functional* f, const double* gamma, const double* J,
                const double* K, const double* Hcore, double* grad);
/* The occupation optimiser of occoptr (minimization.jl:100-139): optimize(calcocce,
 * calcoccg, gamma, ConjugateGradient(), Optim.Options(g_abstol = threshen)) run on the
 * device with J, K, H_core (NULL = resident, e.g. from dn_jkh_mo_resident): nonlinear CG
 * with the Hager-Zhang direction and a Wolfe line search whose state stays in HBM; each
 * evaluation is calcocce + calcoccg at a trial gamma, the host reads one flag per 64
 * evaluations.  gamma_out (nv), E (the minimum), nit (CG iterations, res.iterations) and
 * converged (res.stopped_by.g_converged) are returned. */
int dn_occoptr(dn_sys* sys, const dn_functional* f, const double* gamma, const double* J,
               const double* K, const double* Hcore, double threshen, int maxiter,
               double* gamma_out, double* E, int* nit, int* converged);
/* order_occupations_softmax (utils.jl:586-616): when some gamma > 0, each pair's
 * orbitals (its strong one and ncwo weak ones) sorted by gamma, strongest first, and the
 * gammas shifted by their maximum; C (nbf x nbf) and gamma (nv) in, the re-ordered C
 * and gamma out (host arrays; the permutation runs on the device). */
int dn_order_occupations_softmax(dn_sys* sys, const double* C, const double* gamma,
                                 double* C_out, double* gamma_out);

/* ---- orbital rotation and optimiser ------------------------------------------------ */

/* Cnew = C exp(Y(y)), y packed row-wise over i < nbf5, j > i (nvar entries):
 * rotate_orbital (utils.jl:499-521; ext/DoNOFGPUExt.jl:355-382, there in FP32). */
int dn_rotate_orbital(dn_sys* sys, const double* y, const double* C, double* Cnew);

/* State of one orbital-optimiser call; maxloop and alpha are updated as the reference
 * mutates p.maxloop and p.alpha (minimization.jl:426-431). */
typedef struct dn_orbopt_state {
  int32_t method;     /* DN_ADAM, DN_ADABELIEF, DN_YOGI or DN_DEMON */
  int32_t maxloop;    /* in/out: p.maxloop */
  double alpha;       /* in/out: p.alpha */
  double threshgorb;  /* p.threshgorb */
  int32_t nit;        /* out: iterations */
  int32_t success;    /* out: gradient criterion met */
  double E;           /* out: returned electronic energy */
} dn_orbopt_state;

/* orbopt_adam / orbopt_adabelief / orbopt_yogi / orbopt_demon (minimization.jl:242-630):
 * C (nbf x nbf) is replaced by the returned orbitals. */
int dn_orbopt(dn_sys* sys, const dn_functional* f, dn_orbopt_state* st, const double* gamma,
              double* C, const double* H);
/* Resident form: gamma, C and H already set on the system (dn_sys_set_gamma,
 * dn_sys_set_orbitals, dn_sys_set_hcore); C stays in HBM. */
int dn_sys_set_gamma(dn_sys* sys, const double* gamma);
int dn_sys_get_orbitals(dn_sys* sys, double* C);
int dn_orbopt_resident(dn_sys* sys, const dn_functional* f, dn_orbopt_state* st);

/* ---- step-wise optimisation and lock-step groups ---------------------------------
 * The loop of dn_orbopt_resident (orbopt_adam & co., minimization.jl:242-630) split into
 * start / steps / finish, so that several systems can advance together: a dn_ogroup
 * steps every started system by one iteration per dn_ogroup_step, the orbital Lagrangian
 * of all of them (utils.jl:1-59) running as grouped GEMMs in one HIP graph.
 * The reference has no grouped form (it optimises one system per call); per system the
 * recurrence and its decisions are the single-system loop's, with the same arithmetic.
 * dn_orbopt_start takes gamma, C and H already set on the system (dn_sys_set_gamma,
 * dn_sys_set_orbitals, dn_sys_set_hcore); dn_orbopt_finish fills st as dn_orbopt_resident
 * does (E, nit, success, updated alpha and maxloop) and leaves the orbitals in the system
 * (dn_sys_get_orbitals).  While a system belongs to a dn_ogroup it runs on the group's
 * stream; destroy the group before using the system on its own again. */
int dn_orbopt_start(dn_sys* sys, const dn_functional* f, const dn_orbopt_state* st);
int dn_orbopt_done(dn_sys* sys, int* done);
int dn_orbopt_finish(dn_sys* sys, dn_orbopt_state* st);
typedef struct dn_ogroup dn_ogroup;
int dn_ogroup_create(dn_ctx* ctx, dn_sys* const* systems, int nsys, dn_ogroup** out);
int dn_ogroup_destroy(dn_ogroup* group);
/* One iteration of every unfinished system; *active = systems that stepped. */
int dn_ogroup_step(dn_ogroup* group, int* active);
/* Repeated dn_ogroup_step without returning to the caller between iterations: stops after
 * the iteration in which a system finished, when no system stepped, when *stop is non-zero
 * at an iteration boundary (another thread may set it, e.g. when solves wait to join), or
 * after max_steps.  *steps = iterations run, *active = systems that stepped in the last.
 * (No reference counterpart: the serving loop's native inner loop.) */
int dn_ogroup_run(dn_ogroup* group, int max_steps, const volatile int* stop, int* steps,
                  int* active);

/* ---- J/K-MO build without allocation ---------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * build of dn_jkh_mo (computeJKH_MO, integrals.jl:131-172) on caller-owned DEVICE arrays,
 * b (nbf x nbf x nbfaux), C and H (nbf x nbf) in, J, K (nbf5 x nbf5) and Hcore (nbf5) out,
 * all column-major with leading dimension nbf, enqueued on `stream` (a hipStream_t; NULL
 * is the null stream) without synchronising.  It calls no hipMalloc/hipFree or any
 * allocating routine: its intermediates live in one static device array of
 * DN_NOALLOC_SCRATCH_BYTES per GPU, held from the library's first use of that GPU until
 * the process ends.  The intermediates take 8 * (nbf*nbf5*nbfaux + nbf5^2*nbfaux +
 * nbf5*nbfaux + nbf*nbf5 + 128*nbf5^2 + 16*nbf^2) bytes (each part rounded up to 256
 * bytes); larger sizes return DN_ERR_ARG.  Every call on a GPU uses that GPU's one array:
 * calls that overlap on one GPU (other streams, other threads) race, so the caller
 * orders them.  b need not be symmetric (the general contraction runs). */
#define DN_NOALLOC_SCRATCH_BYTES (128ull << 20)
int dn_jkh_mo_noalloc(int device, int nbf, int nbfaux, int nbf5, const double* b,
                      const double* C, const double* H, double* J, double* K, double* Hcore,
                      void* stream);

/* ---- grouped builds --------------------------------------------------------------
 * Many resident systems of uneven size built in one sequence of grouped, persistent
 * launches, captured once into a HIP graph.  Each system's build is computeJKH_MO /
 * JKH_MO (integrals.jl:131-172) on its resident b, C (dn_sys_set_orbitals) and H
 * (dn_sys_set_hcore); the reference has no grouped form, it builds one system per call.
 * Results stay in each system (read them with dn_sys_get_jkh).  The systems must outlive
 * the group and share its context and precision.  DN_SPLIT_F16 groups run the split
 * transforms of the symmetric-b build (systems with a symmetric b and symmetry enabled)
 * as grouped split launches; J and H_core stay FP64 as in the single-system build. */
typedef struct dn_group dn_group;
int dn_group_create(dn_ctx* ctx, dn_sys* const* systems, int nsys, dn_group** out);
int dn_group_destroy(dn_group* group);
/* The systems of each grouped GEMM stage (K1, K2, J, H C5) are split into classes by
 * operand layout, alignment and size; this times every tile configuration of every
 * class on this group, keeps the fastest, and captures the build graph; then, unless
 * dn_group_place fixed it, it chooses the placement (see dn_group_place).  configs
 * (4 ints, may be NULL) receives each stage's choice for its grouped class with the most
 * work (FP64 tiles 0-10; split-precision transforms 100 + their split configuration; -1
 * when no system is grouped). */
int dn_group_tune(dn_group* group, int* configs);
/* Placement of the group's systems (no reference counterpart): the systems with nbf5 <=
 * max_grouped_nbf5 are built by the grouped launches, the others by the single-system
 * build (dn_jkh_mo_resident's kernels) on their own streams beside them; a limit at or
 * above every system's nbf5 groups all of them, one below every nbf5 none.  Without this
 * call dn_group_tune chooses the placement by timing whole builds (every system grouped,
 * none, and the smallest systems grouped up to a few nbf5 limits); after it, dn_group_tune
 * tunes the grouped part only.  Results do not depend on the placement beyond the last
 * bits (each GEMM family sums in its own fixed order). */
int dn_group_place(dn_group* group, int max_grouped_nbf5);
/* The launch plan as text: per stage, each class's system count, tile shape and GFLOP. */
int dn_group_plan_text(dn_group* group, char* buf, int len);
/* One J/K/H_core build of every system (a graph replay; asynchronous). */
int dn_group_jkh(dn_group* group);
int dn_group_sync(dn_group* group);
int dn_group_set_timing(dn_group* group, int on);
/* Summed HIP-event time of the builds since the last call, and their count. */
int dn_group_get_timing(dn_group* group, double* total_ms, long* builds);

/* ---- Integral source of an RI system (compute_integrals, integrals.jl:1-65) ----------
 * The reference evaluates its integrals with GaussianBasis (overlap, kinetic, nuclear,
 * ERI_2e2c, ERI_2e3c, ERI_2e4c; integrals.jl:4-10,50-53) and hands G and I3c to Iaux
 * (dn_sys_ri_metric here).  A dn_basis holds contracted Gaussian shells: shell s has
 * angular momentum l[s] (0..6), nprim[s] primitives, pure[s] != 0 for real solid
 * harmonics (m = -l..l; p as x, y, z) or 0 for Cartesian components (x^l, x^(l-1)y,
 * x^(l-1)z, ...), centre centers[3s..3s+2] in bohr, and its exponents and contraction
 * coefficients (for normalised primitives, as basis-set files list them) consecutively
 * in exps/coefs.  Every function is normalised to one.  Matrices are column-major over
 * the functions in shell order; the integrals are computed on the host's threads
 * (McMurchie-Davidson). */
typedef struct dn_basis dn_basis;
int dn_basis_create(int nshell, const int* l, const int* nprim, const int* pure,
                    const double* centers, const double* exps, const double* coefs,
                    dn_basis** out);
int dn_basis_destroy(dn_basis* basis);
int dn_basis_nbf(const dn_basis* basis, int* nbf);
/* overlap, kinetic and nuclear attraction (charges Z[natom] at xyz[3 natom], bohr),
 * nbf x nbf each; a NULL output is skipped (integrals.jl:4-9) */
int dn_int1e(const dn_basis* basis, int natom, const double* Z, const double* xyz, double* S,
             double* T, double* V);
/* (P|Q), naux x naux (ERI_2e2c, integrals.jl:50) */
int dn_int2c2e(const dn_basis* aux, double* G);
/* (mn|P), nbf x nbf x naux (ERI_2e3c, integrals.jl:52) */
int dn_int3c2e(const dn_basis* basis, const dn_basis* aux, double* I3c);
/* (mn|ls), nbf^4 (ERI_2e4c, integrals.jl:16; the non-RI integrals) */
int dn_int4c2e(const dn_basis* basis, double* eri);
/* (mn|P) computed on the device (orbital shells up to d, fitting shells up to g), copied
 * to the host array I3c (nbf x nbf x naux); the same integrals as dn_int3c2e. */
int dn_int3c2e_device(const dn_basis* basis, const dn_basis* aux, double* I3c);
/* The RI setup on the device from a basis and a fitting basis: G = (P|Q) on the host
 * (symmetrised), (mn|P) on the device straight into the metric solve's workspace, then
 * b = I3c G^{-1/2} as dn_sys_ri_metric (integrals.jl:47-53, 68-91).  The system's nbf and
 * nbfaux must be the two bases' sizes. */
int dn_sys_ri_from_basis(dn_sys* sys, const dn_basis* basis, const dn_basis* aux);
/* Derivatives for the nuclear gradients (GaussianBasis' ∇overlap, ∇kinetic, ∇nuclear,
 * ∇ERI_2e2c, ∇ERI_2e3c as grads.jl:13-19,88-89 calls them): d/dR of each integral for
 * the position R (bohr) of atom `atom`, the Cartesian component last (nbf x nbf x 3,
 * naux x naux x 3, nbf x nbf x naux x 3).  shell_atom / aux_atom give each shell's atom
 * (0-based); for dV the nuclei Z, xyz are those of dn_int1e and the moving atom's nucleus
 * moves with its functions.  A NULL output of dn_int1e_deriv is skipped. */
int dn_int1e_deriv(const dn_basis* basis, const int* shell_atom, int natom, const double* Z,
                   const double* xyz, int atom, double* dS, double* dT, double* dV);
int dn_int2c2e_deriv(const dn_basis* aux, const int* aux_atom, int atom, double* dG);
int dn_int3c2e_deriv(const dn_basis* basis, const int* shell_atom, const dn_basis* aux,
                     const int* aux_atom, int atom, double* dI);

#ifdef __cplusplus
}
#endif

#endif /* DONOF_HIP_H */
