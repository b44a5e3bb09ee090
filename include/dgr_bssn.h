/*
 * dgr_bssn.h — C-ABI of the MI355X BSSN right-hand side (Dendro-GR hot path).
 *
 * Everything here is plain C: device pointers are caller-owned `double*`
 * (hipMalloc'd or any device allocation), streams are passed as `void*`
 * (a hipStream_t; NULL = the null stream). All functions return 0 on success
 * or a negative DGR_E* code; dgr_last_error() describes the last failure of
 * the calling thread. Nothing here calls exit().
 *
 * Layout: 24 BSSN variables in the order of BSSN_GR/include/grDef.h:35-60
 * (alpha, chi, K, Gt0-2, beta0-2, B0-2, gt00..gt22, At00..At22), each a
 * separate device array of `szpdof` doubles holding the padded ("unzipped")
 * blocks; block b occupies [offset_b, offset_b + nx*ny*nz) with point
 * (i,j,k) at offset_b + i + nx*(j + ny*k) — the reference's unzip layout
 * (BSSN_GR/src/rhs.cpp:84-132, 208).
 */
#ifndef DGR_BSSN_H
#define DGR_BSSN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGR_NUM_VARS 24
#define DGR_PADDING_WIDTH 3

enum {
    DGR_OK = 0,
    DGR_E_INVALID = -1,  /* bad argument (sizes, offsets, null pointers) */
    DGR_E_HIP = -2,      /* HIP runtime error */
    DGR_E_NOMEM = -3,    /* device allocation failed */
    DGR_E_CONSTRAINT = -4, /* det(gt) <= 0 met while enforcing constraints */
    DGR_E_STEP = -5        /* dgr_rk45_step: tolerance not met within max_attempts tries */
};

/* Face bits of bflag: OCT_DIR_LEFT/RIGHT (-x/+x) = bits 0/1, DOWN/UP (-y/+y) =
 * bits 2/3, BACK/FRONT (-z/+z) = bits 4/5 (Dendro's OCT_DIR_* numbering, used as
 * 1u << OCT_DIR_* at BSSN_GR/src/derivs.cpp:2654,2676 and rhs.cpp:542-629). */
#define DGR_FACE_LEFT (1u << 0)
#define DGR_FACE_RIGHT (1u << 1)
#define DGR_FACE_DOWN (1u << 2)
#define DGR_FACE_UP (1u << 3)
#define DGR_FACE_BACK (1u << 4)
#define DGR_FACE_FRONT (1u << 5)

/* One padded block: exactly what bssnRHS() derives from ot::Block before it
 * calls bssnrhs() (BSSN_GR/src/rhs.cpp:49-70): offset, allocation size,
 * face flags and the coordinates of the first and last padded points.
 * Grid spacing is (pmax - pmin)/(n - 1) per axis (rhs.cpp:144-146). */
typedef struct dgr_block {
    uint64_t offset;
    uint32_t sz[3];
    uint32_t bflag;
    double pmin[3];
    double pmax[3];
} dgr_block;

/* Parameters bssnrhs() reads from the bssn:: globals (BSSN_GR/src/rhs.cpp:
 * 144-168, 209-215, 338-355; defaults BSSN_GR/src/parameters.cpp:36-42,
 * 123-136). dgr_default_params() fills the reference defaults. */
typedef struct dgr_bssn_params {
    double lambda[4];       /* BSSN_LAMBDA */
    double lambda_f[2];     /* BSSN_LAMBDA_F */
    double eta_central;     /* RIT_ETA_CENTRAL */
    double eta_outer;       /* RIT_ETA_OUTER */
    double eta_width;       /* RIT_ETA_WIDTH */
    double ko_sigma;        /* KO_DISS_SIGMA */
    double chi_floor;       /* CHI_FLOOR */
    int32_t cako_enabled;   /* BSSN_CAKO_ENABLED */
    int32_t eta_func;       /* 0: the RIT eta(r) above (rhs.cpp:209-215); 1: the eta of the
                               reference's BSSN_ETA_FUNCTION build (USE_ETA_FUNC, equations
                               bssneqs_eta_func_standard_gauge.cpp): eta_r0 |d chi| / (1 - chi^p0)^p1,
                               |d chi|^2 = gt^ij d_i chi d_j chi (CodeGen/bssn.py:71-75) */
    double eps_cako_gauge;  /* BSSN_EPSILON_CAKO_GAUGE */
    double eps_cako_other;  /* BSSN_EPSILON_CAKO_OTHER */
    double eta_r0;          /* BSSN_ETA_R0 */
    double eta_power[2];    /* BSSN_ETA_POWER */
    int32_t rochester;      /* 1: the Rochester puncture gauge of the reference's USE_ROCHESTER_GAUGE
                               build (bssneqs_*_rochester_gauge*.cpp, CodeGen/bssn.py:415-455):
                               beta_rhs^i = xi[1] beta^j d_j beta^i + 3/4 xi[2] Gt^i - eta beta^i,
                               B frozen (B's right-hand side is 0 before Kreiss-Oliger; the
                               reference's generated code leaves it unwritten) */
    int32_t ssl;            /* 1: the slow-start lapse of the BSSN_ENABLE_SSL_IN_RHS build
                               (bssneqs_*_SSL.cpp, CodeGen/bssn.py:90-99): alpha's right-hand side
                               gains -0.6 sqrt(chi) (alpha - sqrt(chi)) exp(-t^2/800) */
    double xi[3];           /* BSSN_XI (unsigned ints in the reference, parameters.cpp:157) */
    double time;            /* the t the slow-start lapse reads in dgr_rk3_step, dgr_rk4_step(s)
                               and dgr_rk45_step: the step's start time at every stage, as
                               rkBSSN.cpp:694-836 passes current_t; dgr_rk4_steps adds dt per
                               step (on the device, inside its captured graph). dgr_bssn_rhs
                               reads its t argument instead. */
} dgr_bssn_params;

typedef struct dgr_plan dgr_plan;

/* Reference defaults (BSSN_GR/src/parameters.cpp:36-42,123-136). */
void dgr_default_params(dgr_bssn_params *p);

/* Last error message of the calling thread ("" if none). */
const char *dgr_last_error(void);

/* Validates the block table (sizes >= 13 per axis, blocks inside [0, szpdof)),
 * splits every block into tiles and orders the work for the device. The
 * default is the single-kernel path (tiles of at most 7 x 7 x 5 points, or
 * 7 x 7 x 7 in mixed precision, see DGR_TILING_*; no derivative workspace);
 * the environment variable DGR_KERNEL=split selects the
 * two-kernel path (tiles of at most 7^3 points; a derivative workspace of 162
 * doubles per point in chunks of DGR_CHUNK_POINTS points).
 * Replaces the reference's global derivative workspace and its per-block
 * loop setup (BSSN_GR/src/rhs.cpp:10-73, grUtils.cpp:2374-2396). The table is
 * copied; the plan does not keep `blocks`. */
int dgr_plan_create(int device, const dgr_block *blocks, uint32_t nblocks, uint64_t szpdof,
                    dgr_plan **out);
int dgr_plan_destroy(dgr_plan *plan);

/* Arithmetic of dgr_bssn_rhs / dgr_rk4_step(s) (no counterpart in the
 * reference, which is FP64 only). DGR_PRECISION_FP64 (default): everything in
 * FP64, rounding like the reference. DGR_PRECISION_MIXED: derivatives,
 * equations and KO in FP32 on fields staged with a per-tile offset subtracted
 * (the value at the tile region's first point), while the state, the RK stage
 * accumulation and the constraint enforcement stay FP64; the right-hand side
 * is widened to FP64 before it is stored or used. Fused path only. */
#define DGR_PRECISION_FP64 0
#define DGR_PRECISION_MIXED 1
int dgr_plan_set_precision(dgr_plan *plan, int precision);

/* Tile shapes of the fused kernel (no counterpart in the reference, whose
 * kernels take fixed shapes). Every block is cut into columns of at most 7 x 7
 * points in x and y; along z a block without flagged faces is cut by one of:
 *   DGR_TILING_WT    parts of 5, a 1-2 point remainder in 128-thread tiles
 *   DGR_TILING_W     near-equal parts of at most 5
 *   DGR_TILING_FULL  near-equal parts of at most 7 (384-thread tiles; mixed
 *                    precision only)
 * Blocks with flagged faces always use near-equal parts of at most 5. Each
 * precision keeps one tiling per block size; before tuning: WT in FP64; in
 * mixed precision FULL for blocks with at most 7 interior points along z and
 * WT for the others. */
#define DGR_TILING_WT 0
#define DGR_TILING_W 1
#define DGR_TILING_FULL 2

/* Tile autotuner: for the plan's current precision and every size of the
 * blocks without flagged faces, times one right-hand-side evaluation of a
 * sample of those blocks (at most 2048) under each applicable tiling (best of
 * three after a warm-up, HIP events on `stream`) and keeps the fastest. The
 * evaluations read d_u and overwrite d_rhs (both caller-owned, laid out as
 * for dgr_bssn_rhs). Allocates temporary device memory, waits for `stream`,
 * and drops a captured dgr_rk4_steps graph. Results do not depend on the
 * tiling in FP64 (bitwise); in mixed precision they move within its error
 * (the staging offset is per tile). */
int dgr_plan_autotune(dgr_plan *plan, double *const *d_rhs, const double *const *d_u,
                      const dgr_bssn_params *params, void *stream);

/* Sets the tiling of blocks of size nx x ny x nz (without flagged faces) for
 * `precision` and rebuilds that precision's tile list (DGR_TILING_FULL is for
 * mixed precision only). Drops a captured dgr_rk4_steps graph. */
int dgr_plan_set_tiling(dgr_plan *plan, int precision, uint32_t nx, uint32_t ny, uint32_t nz, int scheme);

/* The tiling per block size for `precision`: returns the number of distinct
 * sizes of blocks without flagged faces and writes the first max_sizes of
 * them (sizes[3i..3i+2] = nx, ny, nz; scheme[i] = DGR_TILING_*), in
 * increasing (nx, ny, nz) order. */
int dgr_plan_tiling(const dgr_plan *plan, int precision, int max_sizes, uint32_t *sizes, int *scheme);

/* Interior points (sum over blocks of (nx-6)(ny-6)(nz-6)) the plan covers. */
int64_t dgr_plan_interior_points(const dgr_plan *plan);

/* Moves block b of the plan to another place in the unzipped arrays and in
 * space (offset, pmin, pmax as in dgr_block), keeping its sizes and face
 * flags, so one plan serves every block of one shape: the per-block form of
 * the reference's loop (BSSN_GR/src/rhs.cpp:49-71 calls bssnrhs once per
 * block with that block's offset, corner coordinates and spacing). Waits for
 * the device (earlier launches may still read the block table); rebuilds the
 * neighbour table when one was built. */
int dgr_plan_move_block(dgr_plan *plan, uint32_t b, uint64_t offset, const double *pmin, const double *pmax);

/* Kernel timing (no counterpart in the reference, whose profiler_t timers wrap
 * the phases of bssnrhs, include/profile_params.h:15-53): when enabled, every
 * launch is bracketed by HIP events on the caller's stream.
 * dgr_plan_kernel_times waits for them and returns, per kernel kind, the
 * summed milliseconds and the number of launches since the previous call;
 * `ms` and `launches` have DGR_TIMING_KINDS entries: 0 derivatives,
 * 1 equations, 2 boundary faces (the split path); 3 fused kernel on regular
 * tiles of blocks without flagged faces, 4 on thin tiles, 5 on tiles of
 * blocks with flagged faces, 6 the mixed-precision kernel on 7 x 7 x 7 tiles
 * of blocks without flagged faces (it replaces 3 and 4 in that precision). */
#define DGR_TIMING_KINDS 7
int dgr_plan_set_timing(dgr_plan *plan, int enable);
int dgr_plan_kernel_times(dgr_plan *plan, double *ms, int64_t *launches);

/* Right-hand side of all blocks: for each block the work of bssnrhs()
 * (BSSN_GR/src/rhs.cpp:80-504): 6th-order derivatives, the BSSN equations,
 * the radiative boundary fix-up on flagged faces, Kreiss-Oliger dissipation.
 * d_rhs[v] receives interior points only; its padding is not written.
 * `t` is the time of the slow-start lapse term (params->ssl; the other
 * equations do not read it). Asynchronous on `stream`. Replaces bssnRHS() (rhs.h:28-30) and
 * device::bssnrhs3 (cuda/include/bssn_kernels.cuh:1526-1538). */
int dgr_bssn_rhs(dgr_plan *plan, double *const *d_rhs, const double *const *d_u, double t,
                 const dgr_bssn_params *params, void *stream);

/* One classic RK4 step (BSSN_GR/include/rkBSSN.h:50-53, src/rkBSSN.cpp:764-860)
 * of every block, in place on d_u. Padding points keep their values (blocks are
 * advanced without re-filling padding between stages) unless the padding
 * refresh is on (dgr_plan_set_halo_refresh). With
 * enforce_per_stage != 0 the algebraic constraints are enforced on every stage
 * state as rkBSSN.cpp:810-812 does; they are always enforced on the result
 * (rkBSSN.cpp:858, bssnCtx.cpp:1174-1182). Uses plan-owned buffers,
 * allocated on first use: a padded work state (24 x szpdof doubles, its
 * padding refreshed from d_u's at the start of every call) and two
 * interior-only states (the step's initial state and the accumulator,
 * 2 x 24 x interior-points doubles). The stages alternate between d_u and the
 * work state, so d_u's interior holds an intermediate stage state while the
 * call runs (its padding is written only by the padding refresh, when that is
 * on). */
int dgr_rk4_step(dgr_plan *plan, double *const *d_u, double dt, const dgr_bssn_params *params,
                 int enforce_per_stage, void *stream);

/* `nsteps` RK4 steps as dgr_rk4_step does them, with the launches of one step
 * captured once into a HIP graph and the graph replayed on `stream` (the
 * reference's solver loop drives the stages from the host each time,
 * rkBSSN.cpp:764-860). The graph is kept in the plan and captured again when
 * d_u, dt, params or enforce_per_stage differ from the previous call. With
 * kernel timing enabled, or on the split path, the steps run without a graph. */
int dgr_rk4_steps(dgr_plan *plan, double *const *d_u, double dt, const dgr_bssn_params *params,
                  int enforce_per_stage, int nsteps, void *stream);

/* One step of the three-stage SSP Runge-Kutta of the reference's legacy solver
 * (BSSN_RK_TYPE = 0, BSSN_GR/src/rkBSSN.cpp:691-762): u1 = u0 + dt k(u0),
 * u2 = 3/4 u0 + 1/4 u1 + dt/4 k(u1), u3 = 1/3 u0 + 2/3 u2 + 2/3 dt k(u2), with
 * enforce_bssn_constraints after every stage as there. Each stage is the
 * right-hand side (as dgr_bssn_rhs, into the plan's padded work state) and a
 * pointwise combination in place on d_u's interior; the padding refresh, when
 * on (dgr_plan_set_halo_refresh), runs on d_u before every stage and after the
 * step, otherwise the padding keeps its values. Uses the plan buffers of
 * dgr_rk4_step. */
int dgr_rk3_step(dgr_plan *plan, double *const *d_u, double dt, const dgr_bssn_params *params, void *stream);

/* The smallest and largest finite value of one variable (a szpdof array on
 * the plan's device) over every block's interior points, and the number of
 * non-finite values among them: the failure check the reference makes after
 * each step (BSSNCtx::terminal_output, BSSN_GR/src/bssnCtx.cpp:1240-1262, the
 * range of alpha, stopping the run on a NaN), here returned to the caller
 * instead of aborting. NaN for both when no value is finite. Synchronizes
 * `stream`. */
int dgr_variable_range(dgr_plan *plan, const double *d_var, double *vmin, double *vmax, int64_t *nonfinite,
                       void *stream);

/* One step of the embedded Runge-Kutta 4(5) with step-size control of the
 * reference's legacy solver (BSSN_RK_TYPE = 2, BSSN_GR/src/rkBSSN.cpp:861-1085,
 * tableau rkBSSN.h:57-71 as the reference has it, including its fourth-order
 * weight 2197/4101 of stage 3). Six stages, each the right-hand side (as
 * dgr_bssn_rhs) and a pointwise combination u = u0 + sum U[s+1][m] h k_m with
 * enforce_bssn_constraints; then the fourth- and fifth-order solutions, both
 * enforced, and err = max |y4 - y5| over every variable and interior point.
 * While err > tol the step is repeated from the step's start with
 * h = 0.8 h (tol/err)^(1/4) (at most max_attempts tries; the reference has no
 * limit). On return d_u holds y5 of the last try, *dt_used the h it took,
 * *err_norm its err, *attempts the tries (each may be NULL), and *dt the next
 * step size, 0.8 h (tol/err)^(1/5) (or the retry size when the limit was hit,
 * which returns DGR_E_STEP). The padding refresh, when on, runs before every
 * stage and after the step. Synchronizes `stream` once per try (the accept
 * test is on the host, as the reference's is after its all-reduce). Uses the
 * plan buffers of dgr_rk4_step and 6 x 24 x interior doubles more. */
int dgr_rk45_step(dgr_plan *plan, double *const *d_u, double *dt, double tol, const dgr_bssn_params *params,
                  int max_attempts, double *dt_used, double *err_norm, int *attempts, void *stream);

/* Padding refresh, the reference's unzip (BSSN_GR/src/bssnCtx.cpp:75; the
 * octree code is in dendrolib) restated on the block table.
 *
 * Same-level neighbours: block nb is block b's neighbour in direction
 * (di,dj,dk) in {-1,0,1}^3 when both have the same sizes and spacing and
 * pmin_nb = pmin_b + d (n - 7) h (the two share their boundary nodes); the
 * table is found from the block geometry. Padding points of b inside such a
 * neighbour take its value at the same coordinates (a copy).
 *
 * The rest (DGR_REFRESH_FULL, dgr_padding_fill): a padding point whose
 * direction has no such neighbour takes its value from the block whose
 * interior holds it: a block of the same spacing and another size (copy), of
 * half the spacing (a finer block; the point is one of its nodes: injection)
 * or of twice the spacing (a coarser block: order-6 Lagrange interpolation in
 * its element of 6 intervals holding the point); the same spacing wins over
 * half, half over twice, ties go to the lower block index. Padding with no
 * such block (the domain boundary) is left as it is.
 *
 * dgr_plan_link_neighbors builds (again) the table; nbr (27 int32 per block,
 * index (di+1)+3(dj+1)+9(dk+1), -1 for none; may be NULL) and nlinks (may be
 * NULL) receive it. dgr_halo_fill refreshes the padding of d_u from the
 * same-level neighbours in place, dgr_padding_fill all of it (asynchronous on
 * `stream`; both build what they need on first use). dgr_plan_fill_sources
 * builds the table of the rest (kept in the plan) and reports how many
 * padding points it copies and how many it interpolates.
 * dgr_plan_set_halo_refresh(plan, mode) makes every stage of
 * dgr_rk4_step(s) see its input state's padding refreshed, as the
 * reference's unzip between stages does (rkBSSN.cpp:764-860):
 * DGR_REFRESH_LINKED from same-level neighbours (the stage kernels of both
 * precisions read that padding from the neighbours when szpdof is below 2^29;
 * larger plans copy it in before each stage), DGR_REFRESH_FULL also the rest
 * (written into the stage input before each stage). When the call returns, the caller's padding
 * is refreshed from the new interiors the same way. DGR_REFRESH_OFF (the
 * default) keeps every block's padding at its value at the start of the call.
 * Needs the fused path. */
#define DGR_REFRESH_OFF 0
#define DGR_REFRESH_LINKED 1
#define DGR_REFRESH_FULL 2
int dgr_plan_link_neighbors(dgr_plan *plan, int32_t *nbr, int64_t *nlinks);
int dgr_halo_fill(dgr_plan *plan, double *const *d_u, void *stream);
int dgr_padding_fill(dgr_plan *plan, double *const *d_u, void *stream);
int dgr_plan_fill_sources(dgr_plan *plan, int64_t *ncopy, int64_t *ninterp);
int dgr_plan_set_halo_refresh(dgr_plan *plan, int mode);

/* enforce_bssn_constraints (BSSN_GR/include/bssn_constraints.h:15-172) on the
 * interior points of every block: det(gt) = 1, tr(At) = 0, chi and alpha
 * floored at chi_floor. */
int dgr_enforce_constraints(dgr_plan *plan, double *const *d_u, double chi_floor, void *stream);

/* Hamiltonian constraint chi igt^jk R_jk - At_ij At^ij + 2/3 K^2 at the interior
 * points of every block, into d_ham (one array of szpdof doubles laid out like a
 * variable; padding not written). Replaces the C_HAM output of
 * physical_constraints (BSSN_GR/src/physcon.cpp:12-89, generated equations
 * physconeqs.cpp:1373, derivatives constraint_derivs.h). Asynchronous on
 * `stream`; dgr_constraints with only d_ham. */
int dgr_ham_constraint(dgr_plan *plan, double *d_ham, const double *const *d_u, void *stream);

/* physical_constraints (BSSN_GR/src/physcon.cpp:12-89) at the interior points
 * of every block, each output laid out like a variable (padding not written):
 *   d_ham      the Hamiltonian constraint, as dgr_ham_constraint;
 *   d_mom[i]   (i = 0,1,2) the momentum constraint igt^jk (D_k At_ij
 *              - C2^m_ki At_jm) - Gt^j At_ij - 3/2 igt^jk At_ki d_j chi / chi
 *              - 2/3 d_i K (C_MOM);
 *   d_psi4[0], d_psi4[1]  real and imaginary part of the Weyl scalar psi4 in
 *              the Gram-Schmidt tetrad of (x,y,z) at x = pmin + i*h
 *              (C_PSI4_REAL, C_PSI4_IMG), zero within 1e-7 of the z axis as
 *              physcon.cpp:82-85.
 * Equations from CodeGen/constraints.py:42-282 (generated into physconeqs.cpp).
 * d_mom and d_psi4 are host arrays of 3 and 2 device pointers, or NULL; d_ham
 * and any entry may be NULL (that output is skipped), but not all six.
 * Asynchronous on `stream`. The fused path evaluates derivatives and
 * constraints in one kernel per tile kind with no workspace (all six are
 * computed whatever is asked for); DGR_KERNEL=split uses the plan's
 * derivative workspace (allocated on first use). */
int dgr_constraints(dgr_plan *plan, double *d_ham, double *const *d_mom, double *const *d_psi4,
                    const double *const *d_u, void *stream);

/* computeConstraintL2Norm (BSSN_GR/include/grUtils.tcc:37-134): sqrt of the mean
 * of c^2 over the interior points whose distance to (bh[0],bh[1],bh[2]) is >= r1
 * and to (bh[3],bh[4],bh[5]) is >= r2 (the reference's BSSN_BH_LOC and
 * BSSN_BH{1,2}_CONSTRAINT_R, default 5.0, parameters.cpp:71-72). `d_c` is laid
 * out like a variable (e.g. the output of dgr_ham_constraint). A node shared
 * by same-level neighbours (see dgr_plan_link_neighbors; the table is built on
 * first use) is counted once, by the neighbour with the lowest block index, as
 * the reference counts each zipped node once (grUtils.tcc:60-110); nodes
 * shared across a level jump or between blocks of different sizes count once
 * per block. Waits for `stream`; *norm = -1 when no point qualifies. */
int dgr_constraint_l2_norm(dgr_plan *plan, const double *d_c, const double *bh, double r1, double r2,
                           double *norm, void *stream);

/* Largest plan dgr_constraint_l2_norm_noalloc serves: workgroups of 256
 * interior points, counted per chunk of blocks (about 268 M interior points). */
#define DGR_NOALLOC_NORM_MAX_WORKGROUPS 1048576

/* dgr_constraint_l2_norm that allocates nothing and does not wait. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same kernels and same result as
 * dgr_constraint_l2_norm, but asynchronous on `stream`: the norm (or -1) is
 * written to d_norm[0], a caller-owned device double. The partial sums live in
 * a fixed device array of the library (16 MiB per GPU, held until the process
 * ends) instead of plan memory; every call on one GPU shares it, so calls
 * that overlap on one GPU (other streams or host threads) race and must be
 * ordered by the caller. Needs the plan's neighbour table to exist
 * (dgr_plan_link_neighbors, or one call of dgr_constraint_l2_norm, at set-up);
 * DGR_E_INVALID for plans of more than DGR_NOALLOC_NORM_MAX_WORKGROUPS norm
 * workgroups or without the table. */
int dgr_constraint_l2_norm_noalloc(dgr_plan *plan, const double *d_c, const double *bh, double r1, double r2,
                                   double *d_norm, void *stream);


/* Waits for `stream`, then reports and clears constraint violations met by
 * dgr_rk4_step / dgr_enforce_constraints since the last call: DGR_E_CONSTRAINT
 * for det(gt) < 0 or a projection off by more than 1e-6 (where the reference
 * calls exit(0), bssn_constraints.h:42-47, 76-88, 121-132), else DGR_OK. */
int dgr_plan_status(dgr_plan *plan, void *stream);

#ifdef __cplusplus
}
#endif
#endif
