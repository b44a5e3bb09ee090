/*
 * ORACLE — test infrastructure only.
 *
 * Plain-C CPU restatement of Dendro-GR's per-block BSSN right-hand side, used
 * as the checker for the HIP path. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; nothing in the product links it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the Dendro-GR checkout root).
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_deriv644_* / orc_ko_deriv42_*   closed forms: exactness on polynomials
 *                                       and brute force vs textbook stencil
 *                                       weights derived independently
 *                                       (tests/test_oracle_stencils.py)
 *   orc_bssn_point                      golden vectors from the reference's own
 *                                       generated equations compiled in place
 *                                       (oracle/_ref, tests/golden/)
 *   orc_bssnrhs_block                   golden block outputs (reference algebra
 *                                       + these stencils), Minkowski => 0,
 *                                       axis-permutation covariance; its glue
 *                                       by closed forms (tests/
 *                                       test_oracle_block_glue.py): the
 *                                       radiative BC on f_inf + c r^-p for p =
 *                                       falloff and falloff +- 1 on every face
 *                                       plane and nowhere else, KO added after
 *                                       it; eta(r) from B_rhs = -eta B on flat
 *                                       data; the CAKO sqrt(chi) eps_gauge /
 *                                       eps_other split from the KO of a
 *                                       product of sines; each check fails on
 *                                       a copy with a planted mistake
 *   orc_ham_point                       golden vectors from the reference's
 *                                       generated physconeqs.cpp compiled in
 *                                       place (oracle/_ref, tests/golden/);
 *                                       Minkowski => 0; Brill-Lindquist data
 *                                       (an exact vacuum solution) => 0 up to
 *                                       truncation, converging at 4th order
 *   orc_constraint_point_xyz (mom)      golden vectors (same build); Minkowski
 *                                       => 0; closed forms on linear data
 *   orc_constraint_point_xyz (psi4)     golden vectors (same build, per-point
 *                                       coordinates); Minkowski => 0; a pure
 *                                       At_xx / R_xx probe against the tetrad
 *                                       closed form on the flat metric
 *   orc_constraint_l2_norm              brute-force mask/mean in the tests
 *   orc_enforce_constraints             det(gt)=1, tr(At)=0, idempotence
 *   orc_rk4_step                        exact on polynomial-in-t RHS (closed
 *                                       form), fixture of one step
 */
#ifndef DGR_BSSN_ORACLE_H
#define DGR_BSSN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Variable order: BSSN_GR/include/grDef.h:35-60 (VAR enum). */
enum {
    ORC_ALPHA = 0, ORC_CHI = 1, ORC_K = 2,
    ORC_GT0 = 3, ORC_GT1 = 4, ORC_GT2 = 5,          /* Gt^i (conformal connection) */
    ORC_BETA0 = 6, ORC_BETA1 = 7, ORC_BETA2 = 8,
    ORC_B0 = 9, ORC_B1 = 10, ORC_B2 = 11,
    ORC_SYMGT0 = 12,                                /* gt_ij: 00 01 02 11 12 22 */
    ORC_SYMAT0 = 18,                                /* At_ij: 00 01 02 11 12 22 */
    ORC_NVARS = 24
};

/* Face bits of bflag: OCT_DIR_{LEFT,RIGHT,DOWN,UP,BACK,FRONT} = 0..5
 * (usage BSSN_GR/src/derivs.cpp:2654,2676; values are Dendro's TreeNode.h). */
enum { ORC_DIR_LEFT = 0, ORC_DIR_RIGHT, ORC_DIR_DOWN, ORC_DIR_UP, ORC_DIR_BACK, ORC_DIR_FRONT };

/* Parameters the path reads (BSSN_GR/src/parameters.cpp:36-42,123-136). */
typedef struct orc_params {
    double lambda[4];      /* BSSN_LAMBDA (unsigned ints in the reference) */
    double lambda_f[2];    /* BSSN_LAMBDA_F */
    double eta_central;    /* RIT_ETA_CENTRAL */
    double eta_outer;      /* RIT_ETA_OUTER */
    double eta_width;      /* RIT_ETA_WIDTH */
    double ko_sigma;       /* KO_DISS_SIGMA */
    double chi_floor;      /* CHI_FLOOR */
    int32_t cako_enabled;  /* BSSN_CAKO_ENABLED */
    int32_t eta_func;      /* 1: the eta of the reference's BSSN_ETA_FUNCTION build (CodeGen/bssn.py:71-75) */
    double eps_cako_gauge; /* BSSN_EPSILON_CAKO_GAUGE */
    double eps_cako_other; /* BSSN_EPSILON_CAKO_OTHER */
    double eta_r0;         /* BSSN_ETA_R0 */
    double eta_power[2];   /* BSSN_ETA_POWER */
    int32_t rochester;     /* 1: the Rochester puncture gauge (USE_ROCHESTER_GAUGE, CodeGen/bssn.py:415-455) */
    int32_t ssl;           /* 1: the slow-start lapse (BSSN_ENABLE_SSL_IN_RHS, CodeGen/bssn.py:90-99) */
    double xi[3];          /* BSSN_XI (parameters.cpp:157) */
    double time;           /* t of bssnrhs (rhs.cpp:82), read by the slow-start lapse only */
} orc_params;

/* One padded block, the arguments bssnrhs() receives (BSSN_GR/src/rhs.cpp:49-70). */
typedef struct orc_block {
    uint64_t offset;       /* start of the block inside each variable array */
    uint32_t sz[3];        /* padded allocation size nx, ny, nz */
    uint32_t bflag;        /* face bits */
    double pmin[3];        /* coordinate of padded point (0,0,0) */
    double pmax[3];        /* coordinate of padded point (nx-1,ny-1,nz-1) */
} orc_block;

/* Per-point algebra input, flat layout of ORC_POINT_NIN doubles:
 *   [0,24)    fields in VAR order
 *   [24,96)   first derivatives, index 24 + 3*v + d   (grad_d_v)
 *   [96,162)  second derivatives of the 11 vars alpha, chi, beta0..2, gt0..5
 *             (in that order), components 00 01 02 11 12 22: 96 + 6*m + c
 *   [162]     eta                                                          */
enum { ORC_POINT_NIN = 163 };
/* The eta of the reference's BSSN_ETA_FUNCTION build at one point's algebra input
 * (CodeGen/bssn.py:71-75): eta_r0 sqrt(gt^ij d_i chi d_j chi) / (1 - chi^p0)^p1. */
double orc_eta_func(const double *in, const orc_params *p);

/* ---- stencils: BSSN_GR/src/derivs.cpp ---------------------------------- */
void orc_deriv644_x(double *Du, const double *u, double dx, const uint32_t *sz, uint32_t bflag);
void orc_deriv644_y(double *Du, const double *u, double dy, const uint32_t *sz, uint32_t bflag);
void orc_deriv644_z(double *Du, const double *u, double dz, const uint32_t *sz, uint32_t bflag);
void orc_deriv644_xx(double *Du, const double *u, double dx, const uint32_t *sz, uint32_t bflag);
void orc_deriv644_yy(double *Du, const double *u, double dy, const uint32_t *sz, uint32_t bflag);
void orc_deriv644_zz(double *Du, const double *u, double dz, const uint32_t *sz, uint32_t bflag);
void orc_ko_deriv42_x(double *Du, const double *u, double dx, const uint32_t *sz, uint32_t bflag);
void orc_ko_deriv42_y(double *Du, const double *u, double dy, const uint32_t *sz, uint32_t bflag);
void orc_ko_deriv42_z(double *Du, const double *u, double dz, const uint32_t *sz, uint32_t bflag);

/* ---- algebra ------------------------------------------------------------ */
void orc_default_params(orc_params *p);
void orc_bssn_point(const double *in, const orc_params *p, double *rhs24);
void orc_bssn_points(const double *in, int64_t npts, const orc_params *p, double *rhs);
/* Hamiltonian constraint at a point, same input layout (eta unused). */
void orc_ham_point(const double *in, double *ham);
void orc_ham_points(const double *in, int64_t npts, double *ham);
/* C_HAM and the momentum constraints mom0..2 (CodeGen/constraints.py:256-282)
 * of one point / npts points (mom: three per point). */
void orc_constraint_point(const double *in, double *ham, double *mom);
void orc_constraint_points(const double *in, int64_t npts, double *ham, double *mom);
/* The same plus psi4 (CodeGen/constraints.py:42-244; psi4: real, imaginary)
 * at the coordinates xyz (three per point); mom and psi4 may be NULL. */
void orc_constraint_point_xyz(const double *in, const double *xyz, double *ham, double *mom, double *psi4);
void orc_constraint_points_xyz(const double *in, const double *xyz, int64_t npts, double *ham, double *mom,
                               double *psi4);
/* C_HAM of physical_constraints (physcon.cpp:12-89) at the interior points of
 * every block; the padding of `ham` is not written. */
int orc_ham_block(double *ham, const double *const *u, const orc_block *blk);
int orc_ham_batch(double *ham, const double *const *u, const orc_block *blks, int64_t nblocks, int nthreads);
/* orc_ham_batch plus the momentum constraints into mom[0..2] and (psi4 !=
 * NULL) psi4 real/imaginary into psi4[0..1] (laid out like a variable; padding
 * not written; psi4 = 0 within 1e-7 of the z axis as physcon.cpp:82-85). */
int orc_constraints_batch(double *ham, double *const *mom, double *const *psi4, const double *const *u,
                          const orc_block *blks, int64_t nblocks, int nthreads);
/* Padding refresh from same-level neighbours (the copy part of unzip): nbr
 * has 27 entries per block, index (di+1)+3(dj+1)+9(dk+1), -1 for none. */
void orc_halo_fill(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr);
/* orc_halo_fill, then (general != 0) the padding points without a same-size
 * neighbour in their direction from the block whose interior holds them:
 * same spacing (copy), half the spacing (injection) or twice the spacing
 * (order-6 Lagrange interpolation in that block's element). */
void orc_padding_fill(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr, int general);
/* The 7 Lagrange weights of an order-6 element at t = code/2, code 0..12. */
void orc_lagrange6_half(int code, double w[7]);
/* orc_rk4_step_halo with orc_padding_fill(general) as the refresh. */
int orc_rk4_step_fill(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr, int general,
                      double dt, const orc_params *p, int enforce_per_stage, int nthreads);
/* One step of the reference's three-stage SSP Runge-Kutta (rkBSSN.cpp:691-762),
 * constraints enforced after every stage; with nbr (may be NULL) the padding is
 * refreshed by orc_padding_fill(general) before every stage and after the step. */
int orc_rk3_step(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr, int general,
                 double dt, const orc_params *p, int nthreads);
/* One step of the reference's embedded Runge-Kutta 4(5) with step-size control
 * (rkBSSN.cpp:861-1085): repeated from the step's start with a smaller h while
 * the largest |y4 - y5| exceeds tol (at most max_attempts tries); u = y5 of the
 * accepted try. Returns 0 when accepted, 1 when max_attempts tries all
 * exceeded tol (u then holds the last try), -1 on a constraint violation;
 * h_used, h_next, err_norm and attempts report the step. */
int orc_rk45_step(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr, int general,
                  double h, double tol, const orc_params *p, int max_attempts, int nthreads, double *h_used,
                  double *h_next, double *err_norm, int *attempts);
/* One RK4 step with every stage state's padding refreshed by orc_halo_fill
 * before its right-hand side, and u's padding after the step. */
int orc_rk4_step_halo(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr, double dt,
                      const orc_params *p, int enforce_per_stage, int nthreads);
/* computeConstraintL2Norm (grUtils.tcc:37-134) over interior points outside
 * r1 of BH 1 and r2 of BH 2 (bh = x1,y1,z1,x2,y2,z2); -1 if none qualifies. */
double orc_constraint_l2_norm(const double *c, const orc_block *blks, int64_t nblocks, const double *bh,
                              double r1, double r2);
/* The same with the neighbour table of orc_halo_fill (27 per block, -1 for
 * none): a node shared by neighbouring blocks is counted once. */
double orc_constraint_l2_norm_nbr(const double *c, const orc_block *blks, int64_t nblocks, const int32_t *nbr,
                                  const double *bh, double r1, double r2);
/* Replace the algebra orc_bssnrhs_block evaluates per point (NULL restores
 * orc_bssn_point). Process-global: fixture generation only, single-threaded. */
typedef void (*orc_point_fn)(const double *in, const orc_params *p, double *rhs24);
void orc_set_point_algebra(orc_point_fn fn);

/* ---- per-block RHS (rhs.cpp:80-504) and batch (rhs.cpp:10-73) ----------- */
int orc_bssnrhs_block(double *const *rhs, const double *const *u, const orc_block *blk,
                      const orc_params *p);
/* Whole-block replacement of the point loop (NULL restores it): the fields,
 * the 72 first-derivative arrays (24 vars x 3, d1[3v+d]) and the 66 second-
 * derivative arrays (11 vars x 6, d2[6m+c]) of one block, its sizes, padded
 * corner and spacing; writes the 24 right-hand sides at the interior points.
 * Process-global: set before the threads start. */
typedef void (*orc_block_algebra_fn)(const double *const *u, const double *const *d1, const double *const *d2,
                                     const uint32_t *sz, const double *pmin, const double *h,
                                     const orc_params *p, double *const *rhs);
void orc_set_block_algebra(orc_block_algebra_fn fn);
/* Seconds per phase (deriv, rhs, bc, ko) of orc_bssnrhs_block summed over
 * blocks and threads since the last reset; reset != 0 clears them. */
void orc_phase_times(double *out4, int reset);
int orc_bssnrhs_batch(double *const *rhs, const double *const *u, const orc_block *blks,
                      int64_t nblocks, const orc_params *p, int nthreads);

/* ---- constraints / RK4 -------------------------------------------------- */
int orc_enforce_constraints_point(double *const *u, uint64_t node, double chi_floor);
int orc_rk4_step(double *const *u, const orc_block *blks, int64_t nblocks, double dt,
                 const orc_params *p, int enforce_per_stage, int nthreads);

/* ---- synthetic inputs --------------------------------------------------- */
uint64_t orc_splitmix64(uint64_t *state);
void orc_init_perturbed_minkowski(double *const *u, const orc_block *blks, int64_t nblocks,
                                  uint64_t seed, double amp);
void orc_init_brill_lindquist(double *const *u, const orc_block *blks, int64_t nblocks,
                              uint64_t seed, double noise, double chi_floor);

#ifdef __cplusplus
}
#endif
#endif
