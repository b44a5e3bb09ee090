/*
 * ORACLE — test infrastructure only (see bssn_oracle.h for the rules and for
 * what pins each function). Compiled with -ffp-contract=off so that every
 * expression rounds in the order written, which is the order of the reference
 * expressions cited beside it.
 */
#define _POSIX_C_SOURCE 199309L /* clock_gettime */
#include "bssn_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX3(i, j, k) ((int64_t)(i) + (int64_t)nx * ((int64_t)(j) + (int64_t)ny * (int64_t)(k)))

/* ======================================================================== */
/* 1-D stencil lines. A line is the set of points base + a*s, a = 0..n-1.   */
/* ======================================================================== */

/* deriv644_{x,y,z} interior + face closures: BSSN_GR/src/derivs.cpp:2646-2697,
 * 2746-2808, 2860-2918. */
static void d1_line(double *Du, const double *u, int64_t base, int64_t s, int n, double h,
                    int lo_face, int hi_face) {
    const double ih = 1.0 / h;
    const double ih_by_12 = ih / 12.0;
    const double ih_by_60 = ih / 60.0;
    const int ie = n - 3;
#define U(a) u[base + (int64_t)(a) * s]
#define D(a) Du[base + (int64_t)(a) * s]
    for (int a = 3; a < ie; a++) {
        D(a) = (-U(a - 3) + 9.0 * U(a - 2) - 45.0 * U(a - 1) + 45.0 * U(a + 1) - 9.0 * U(a + 2) +
                U(a + 3)) *
               ih_by_60;
    }
    if (lo_face) {
        D(3) = (-25.0 * U(3) + 48.0 * U(4) - 36.0 * U(5) + 16.0 * U(6) - 3.0 * U(7)) * ih_by_12;
        D(4) = (-3.0 * U(3) - 10.0 * U(4) + 18.0 * U(5) - 6.0 * U(6) + U(7)) * ih_by_12;
        D(5) = (+U(3) - 8.0 * U(4) + 8.0 * U(6) - U(7)) * ih_by_12;
    }
    if (hi_face) {
        D(ie - 3) = (+U(ie - 5) - 8.0 * U(ie - 4) + 8.0 * U(ie - 2) - U(ie - 1)) * ih_by_12;
        D(ie - 2) = (-U(ie - 5) + 6.0 * U(ie - 4) - 18.0 * U(ie - 3) + 10.0 * U(ie - 2) +
                     3.0 * U(ie - 1)) *
                    ih_by_12;
        D(ie - 1) = (3.0 * U(ie - 5) - 16.0 * U(ie - 4) + 36.0 * U(ie - 3) - 48.0 * U(ie - 2) +
                     25.0 * U(ie - 1)) *
                    ih_by_12;
    }
#undef U
#undef D
}

/* deriv644_{xx,yy,zz}: BSSN_GR/src/derivs.cpp:2975-3047, 3089-3160, 3210-3278. */
static void d2_line(double *Du, const double *u, int64_t base, int64_t s, int n, double h,
                    int lo_face, int hi_face) {
    const double ih_sqrd = 1.0 / (h * h);
    const double by_180 = ih_sqrd / 180.0;
    const double by_12 = ih_sqrd / 12.0;
    const int ie = n - 3;
#define U(a) u[base + (int64_t)(a) * s]
#define D(a) Du[base + (int64_t)(a) * s]
    for (int a = 3; a < ie; a++) {
        D(a) = (2.0 * U(a - 3) - 27.0 * U(a - 2) + 270.0 * U(a - 1) - 490.0 * U(a) +
                270.0 * U(a + 1) - 27.0 * U(a + 2) + 2.0 * U(a + 3)) *
               by_180;
    }
    if (lo_face) {
        D(3) = (45.0 * U(3) - 154.0 * U(4) + 214.0 * U(5) - 156.0 * U(6) + 61.0 * U(7) -
                10.0 * U(8)) *
               by_12;
        D(4) = (10.0 * U(3) - 15.0 * U(4) - 4.0 * U(5) + 14.0 * U(6) - 6.0 * U(7) + U(8)) * by_12;
        D(5) = (-U(3) + 16.0 * U(4) - 30.0 * U(5) + 16.0 * U(6) - U(7)) * by_12;
    }
    if (hi_face) {
        D(ie - 3) = (-U(ie - 5) + 16.0 * U(ie - 4) - 30.0 * U(ie - 3) + 16.0 * U(ie - 2) -
                     U(ie - 1)) *
                    by_12;
        D(ie - 2) = (U(ie - 6) - 6.0 * U(ie - 5) + 14.0 * U(ie - 4) - 4.0 * U(ie - 3) -
                     15.0 * U(ie - 2) + 10.0 * U(ie - 1)) *
                    by_12;
        D(ie - 1) = (-10.0 * U(ie - 6) + 61.0 * U(ie - 5) - 156.0 * U(ie - 4) +
                     214.0 * U(ie - 3) - 154.0 * U(ie - 2) + 45.0 * U(ie - 1)) *
                    by_12;
    }
#undef U
#undef D
}

/* ko_deriv42_{x,y,z}: BSSN_GR/src/derivs.cpp:964-1056, 1079-1171, 1194-1288. */
static void ko_line(double *Du, const double *u, int64_t base, int64_t s, int n, double h,
                    int lo_face, int hi_face) {
    const double pre_factor_6 = -1.0 / 64.0 / h;
    const double smr3 = 59.0 / 48.0 * 64 * h;
    const double smr2 = 43.0 / 48.0 * 64 * h;
    const double smr1 = 49.0 / 48.0 * 64 * h;
    const int ie = n - 3;
#define U(a) u[base + (int64_t)(a) * s]
#define D(a) Du[base + (int64_t)(a) * s]
    for (int a = 3; a < ie; a++) {
        D(a) = pre_factor_6 * (-U(a - 3) + 6.0 * U(a - 2) - 15.0 * U(a - 1) + 20.0 * U(a) -
                               15.0 * U(a + 1) + 6.0 * U(a + 2) - U(a + 3));
    }
    if (lo_face) {
        D(3) = (U(6) - 3.0 * U(5) + 3.0 * U(4) - U(3)) / smr3;
        D(4) = (U(7) - 6.0 * U(6) + 12.0 * U(5) - 10.0 * U(4) + 3.0 * U(3)) / smr2;
        D(5) = (U(8) - 6.0 * U(7) + 15.0 * U(6) - 19.0 * U(5) + 12.0 * U(4) - 3.0 * U(3)) / smr1;
    }
    if (hi_face) {
        D(ie - 3) = (U(ie - 6) - 6.0 * U(ie - 5) + 15.0 * U(ie - 4) - 19.0 * U(ie - 3) +
                     12.0 * U(ie - 2) - 3.0 * U(ie - 1)) /
                    smr1;
        D(ie - 2) = (U(ie - 5) - 6.0 * U(ie - 4) + 12.0 * U(ie - 3) - 10.0 * U(ie - 2) +
                     3.0 * U(ie - 1)) /
                    smr2;
        D(ie - 1) = (U(ie - 4) - 3.0 * U(ie - 3) + 3.0 * U(ie - 2) - U(ie - 1)) / smr3;
    }
#undef U
#undef D
}

typedef void (*line_fn)(double *, const double *, int64_t, int64_t, int, double, int, int);

/* Sweep a line operator along one axis. The other two indices run over
 * [lo_other, n_other - lo_other) — 0 for an axis the reference loops fully,
 * 3 where it restricts to the interior (loop bounds at derivs.cpp:2628-2633,
 * 2729-2734, 2843-2848, 2957-2962, 985-990). */
static void sweep(line_fn fn, double *Du, const double *u, double h, const uint32_t *sz,
                  uint32_t bflag, int axis, int lo_a, int lo_b) {
    const int nx = (int)sz[0], ny = (int)sz[1], nz = (int)sz[2];
    const int lo_face = (bflag >> (2 * axis)) & 1u;
    const int hi_face = (bflag >> (2 * axis + 1)) & 1u;
    if (axis == 0) {
        for (int k = lo_b; k < nz - lo_b; k++)
            for (int j = lo_a; j < ny - lo_a; j++)
                fn(Du, u, IDX3(0, j, k), 1, nx, h, lo_face, hi_face);
    } else if (axis == 1) {
        for (int k = lo_b; k < nz - lo_b; k++)
            for (int i = lo_a; i < nx - lo_a; i++)
                fn(Du, u, IDX3(i, 0, k), nx, ny, h, lo_face, hi_face);
    } else {
        for (int j = lo_b; j < ny - lo_b; j++)
            for (int i = lo_a; i < nx - lo_a; i++)
                fn(Du, u, IDX3(i, j, 0), (int64_t)nx * ny, nz, h, lo_face, hi_face);
    }
}

/* deriv644_x: i in [3,nx-3), all j,k (derivs.cpp:2619-2713). */
void orc_deriv644_x(double *Du, const double *u, double dx, const uint32_t *sz, uint32_t bflag) {
    sweep(d1_line, Du, u, dx, sz, bflag, 0, 0, 0);
}
/* deriv644_y: i,j in [3,n-3), all k (derivs.cpp:2720-2833). */
void orc_deriv644_y(double *Du, const double *u, double dy, const uint32_t *sz, uint32_t bflag) {
    sweep(d1_line, Du, u, dy, sz, bflag, 1, 3, 0);
}
/* deriv644_z: i,j,k in [3,n-3) (derivs.cpp:2834-2941). */
void orc_deriv644_z(double *Du, const double *u, double dz, const uint32_t *sz, uint32_t bflag) {
    sweep(d1_line, Du, u, dz, sz, bflag, 2, 3, 3);
}
/* deriv644_xx / _yy / _zz: [3,n-3)^3 (derivs.cpp:2948-3306). */
void orc_deriv644_xx(double *Du, const double *u, double dx, const uint32_t *sz, uint32_t bflag) {
    sweep(d2_line, Du, u, dx, sz, bflag, 0, 3, 3);
}
void orc_deriv644_yy(double *Du, const double *u, double dy, const uint32_t *sz, uint32_t bflag) {
    sweep(d2_line, Du, u, dy, sz, bflag, 1, 3, 3);
}
void orc_deriv644_zz(double *Du, const double *u, double dz, const uint32_t *sz, uint32_t bflag) {
    sweep(d2_line, Du, u, dz, sz, bflag, 2, 3, 3);
}
/* ko_deriv42_x / _y / _z: [3,n-3)^3 (derivs.cpp:964-1304). */
void orc_ko_deriv42_x(double *Du, const double *u, double dx, const uint32_t *sz, uint32_t bflag) {
    sweep(ko_line, Du, u, dx, sz, bflag, 0, 3, 3);
}
void orc_ko_deriv42_y(double *Du, const double *u, double dy, const uint32_t *sz, uint32_t bflag) {
    sweep(ko_line, Du, u, dy, sz, bflag, 1, 3, 3);
}
void orc_ko_deriv42_z(double *Du, const double *u, double dz, const uint32_t *sz, uint32_t bflag) {
    sweep(ko_line, Du, u, dz, sz, bflag, 2, 3, 3);
}

/* ======================================================================== */
/* Pointwise BSSN algebra: the equations of CodeGen/bssn.py:78-218 with the   */
/* tensor helpers of CodeGen/dendro.py:173-550, evaluated as explicit tensor  */
/* loops (no CSE), one point at a time.                                       */
/* ======================================================================== */

void orc_default_params(orc_params *p) {
    /* BSSN_GR/src/parameters.cpp:36-42,123-136 */
    memset(p, 0, sizeof(*p));
    for (int i = 0; i < 4; i++) p->lambda[i] = 1.0;
    p->lambda_f[0] = 1.0;
    p->lambda_f[1] = 0.0;
    p->eta_central = 2.0;
    p->eta_outer = 0.25;
    p->eta_width = 40.0;
    p->ko_sigma = 0.01;
    p->chi_floor = 0.1;
    p->cako_enabled = 0;
    p->eps_cako_gauge = 0.99;
    p->eps_cako_other = 0.3;
    p->eta_func = 0;          /* the reference's default build (BSSN_ETA_FUNCTION off) */
    p->eta_r0 = 1.31;         /* parameters.cpp:59-60 */
    p->eta_power[0] = 2.0;
    p->eta_power[1] = 2.0;
    p->rochester = 0;         /* standard gauge, no slow-start lapse: the default build */
    p->ssl = 0;
    p->xi[0] = p->xi[1] = p->xi[2] = 0.0; /* parameters.cpp:157 */
    p->time = 0.0;
}

static const int SYM6[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
/* index of each second-derivative variable in the d2 block of the input */
static int d2slot(int v) {
    switch (v) {
    case ORC_ALPHA: return 0;
    case ORC_CHI: return 1;
    case ORC_BETA0: return 2;
    case ORC_BETA0 + 1: return 3;
    case ORC_BETA0 + 2: return 4;
    default: return 5 + (v - ORC_SYMGT0); /* gt0..gt5 */
    }
}

/* The pointwise algebra shared by the right-hand side (rhs != NULL, needs p),
 * the Hamiltonian constraint (ham != NULL), the momentum constraints
 * (mom != NULL, three values) and psi4 (psi4 != NULL: real and imaginary
 * part, at the coordinates xyz). */
static void point_algebra(const double *in, const orc_params *p, double *rhs, double *ham, double *mom,
                          const double *xyz, double *psi4) {
    const double *f = in;
#define D1(v, d) in[24 + 3 * (v) + (d)]
#define D2(v, a, b) in[96 + 6 * d2slot(v) + SYM6[a][b]]
    const double eta = in[162];
    const double l1 = p ? p->lambda[0] : 0.0, l2 = p ? p->lambda[1] : 0.0;
    const double l3 = p ? p->lambda[2] : 0.0, l4 = p ? p->lambda[3] : 0.0;
    const double lf0 = p ? p->lambda_f[0] : 0.0, lf1 = p ? p->lambda_f[1] : 0.0;

    const double alpha = f[ORC_ALPHA], chi = f[ORC_CHI], K = f[ORC_K];
    double g[3][3], A[3][3], b[3], B[3];
    double dg[3][3][3], dA[3][3][3], db[3][3], dGt[3][3], dB[3][3], da[3], dchi[3], dK[3];
    double d2g[3][3][3][3], d2b[3][3][3], d2a[3][3], d2chi[3][3];
    for (int i = 0; i < 3; i++) {
        /* the evolved Gt^i enters only through its derivatives (dendro.py:518-522) */
        b[i] = f[ORC_BETA0 + i];
        B[i] = f[ORC_B0 + i];
        for (int j = 0; j < 3; j++) {
            g[i][j] = f[ORC_SYMGT0 + SYM6[i][j]];
            A[i][j] = f[ORC_SYMAT0 + SYM6[i][j]];
        }
    }
    for (int k = 0; k < 3; k++) { /* k: derivative direction */
        da[k] = D1(ORC_ALPHA, k);
        dchi[k] = D1(ORC_CHI, k);
        dK[k] = D1(ORC_K, k);
        for (int i = 0; i < 3; i++) {
            db[k][i] = D1(ORC_BETA0 + i, k);
            dGt[k][i] = D1(ORC_GT0 + i, k);
            dB[k][i] = D1(ORC_B0 + i, k);
            for (int j = 0; j < 3; j++) {
                dg[k][i][j] = D1(ORC_SYMGT0 + SYM6[i][j], k);
                dA[k][i][j] = D1(ORC_SYMAT0 + SYM6[i][j], k);
            }
        }
    }
    for (int l = 0; l < 3; l++)
        for (int m = 0; m < 3; m++) {
            d2a[l][m] = D2(ORC_ALPHA, l, m);
            d2chi[l][m] = D2(ORC_CHI, l, m);
            for (int i = 0; i < 3; i++) {
                d2b[i][l][m] = D2(ORC_BETA0 + i, l, m);
                for (int j = 0; j < 3; j++) d2g[l][m][i][j] = D2(ORC_SYMGT0 + SYM6[i][j], l, m);
            }
        }
#undef D1
#undef D2

    /* inverse conformal metric from the adjugate (dendro.py:375-393) */
    double igt[3][3];
    {
        const double c00 = g[1][1] * g[2][2] - g[1][2] * g[2][1];
        const double c01 = g[1][2] * g[2][0] - g[1][0] * g[2][2];
        const double c02 = g[1][0] * g[2][1] - g[1][1] * g[2][0];
        const double det = g[0][0] * c00 + g[0][1] * c01 + g[0][2] * c02;
        const double idet = 1.0 / det;
        igt[0][0] = c00 * idet;
        igt[0][1] = igt[1][0] = c01 * idet;
        igt[0][2] = igt[2][0] = c02 * idet;
        igt[1][1] = (g[0][0] * g[2][2] - g[0][2] * g[2][0]) * idet;
        igt[1][2] = igt[2][1] = (g[0][2] * g[1][0] - g[0][0] * g[1][2]) * idet;
        igt[2][2] = (g[0][0] * g[1][1] - g[0][1] * g[1][0]) * idet;
    }

    /* first Christoffel C1[k][i][j] (dendro.py:418-421) */
    double C1[3][3][3], C2[3][3][3], C3[3][3][3], CalGt[3];
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                C1[k][i][j] = 0.5 * (dg[j][k][i] + dg[i][k][j] - dg[k][i][j]);
    /* second Christoffel C2[i][j][k] = igt^il C1_ljk (dendro.py:445-446) */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                double s = 0.0;
                for (int l = 0; l < 3; l++) s += igt[i][l] * C1[l][j][k];
                C2[i][j][k] = s;
            }
    /* Christoffel of the physical metric (dendro.py:471-480) */
    double igt_dchi[3];
    for (int i = 0; i < 3; i++) {
        double s = 0.0;
        for (int m = 0; m < 3; m++) s += igt[i][m] * dchi[m];
        igt_dchi[i] = s;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++)
                C3[i][j][k] = C2[i][j][k] - 0.5 / chi *
                                                ((i == j ? dchi[k] : 0.0) + (i == k ? dchi[j] : 0.0) -
                                                 g[j][k] * igt_dchi[i]);
    /* CalGt^i = igt^kl C2^i_kl (dendro.py:512) */
    for (int i = 0; i < 3; i++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++) s += igt[k][l] * C2[i][k][l];
        CalGt[i] = s;
    }

    /* Ricci tensor R = Rt + Rphi (dendro.py:514-541) */
    double R[3][3];
    double lap_chi_term = 0.0, calgt_dchi = 0.0;
    for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++)
            lap_chi_term += igt[k][l] * (d2chi[k][l] - 3.0 / (2.0 * chi) * dchi[k] * dchi[l]);
    for (int m = 0; m < 3; m++) calgt_dchi += CalGt[m] * dchi[m];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
            for (int l = 0; l < 3; l++)
                for (int m = 0; m < 3; m++) t1 += igt[l][m] * d2g[l][m][i][j];
            for (int k = 0; k < 3; k++) t2 += g[k][i] * dGt[j][k] + g[k][j] * dGt[i][k];
            for (int k = 0; k < 3; k++) t3 += CalGt[k] * (C1[i][j][k] + C1[j][i][k]);
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++)
                    for (int m = 0; m < 3; m++)
                        t4 += igt[l][m] * (C2[k][l][i] * C1[j][k][m] + C2[k][l][j] * C1[i][k][m] +
                                           C2[k][i][m] * C1[k][l][j]);
            const double Rt = -0.5 * t1 + 0.5 * t2 + 0.5 * t3 + t4;
            double c2dchi = 0.0;
            for (int k = 0; k < 3; k++) c2dchi += C2[k][j][i] * dchi[k];
            const double xRphi = 1.0 / (2.0 * chi) * (d2chi[i][j] - c2dchi) -
                                 1.0 / (4.0 * chi * chi) * dchi[i] * dchi[j];
            const double Rphi = xRphi + 1.0 / (2.0 * chi) * g[i][j] * (lap_chi_term - calgt_dchi);
            R[i][j] = Rt + Rphi;
        }

    /* D_i D_j alpha with the physical Christoffels (dendro.py:181-183) */
    double DDa[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int l = 0; l < 3; l++) s += C3[l][i][j] * da[l];
            DDa[i][j] = d2a[i][j] - s;
        }

    /* A^ij, A_ik A^k_j, tr(A^2) (dendro.py:203-208, 305-315; bssn.py:122-133) */
    double AUU[3][3], AikAkj[3][3], sqrA = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) s += igt[i][k] * igt[j][l] * A[k][l];
            AUU[i][j] = s;
            double q = 0.0;
            for (int k = 0; k < 3; k++) {
                double r = 0.0;
                for (int l = 0; l < 3; l++) r += igt[k][l] * A[l][j];
                q += A[i][k] * r;
            }
            AikAkj[i][j] = q;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) sqrA += A[i][j] * AUU[i][j];

    /* Hamiltonian constraint chi igt^jk R_jk - At_ij At^ij + 2/3 K^2
     * (CodeGen/constraints.py:250-254; dendro.sqr, dendro.py:309-318) */
    if (ham) {
        double s = 0.0;
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) s += chi * igt[j][k] * R[j][k];
        *ham = s - sqrA + 2.0 / 3.0 * K * K;
    }
    /* momentum constraints (CodeGen/constraints.py:256-282):
     * mom_i = igt^jk (d_k At_ij - C2^m_ki At_jm) - Gt^j At_ij
     *         - 3/2 igt^jk At_ki d_j chi / chi - 2/3 d_i K, with the evolved Gt^j */
    if (mom) {
        for (int i = 0; i < 3; i++) {
            double s1 = 0.0, s2 = 0.0, s3 = 0.0;
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    double cA = 0.0;
                    for (int m = 0; m < 3; m++) cA += C2[m][k][i] * A[j][m];
                    s1 += igt[j][k] * (dA[k][i][j] - cA);
                    s3 += igt[j][k] * A[k][i] * dchi[j] / chi;
                }
            for (int j = 0; j < 3; j++) s2 += f[ORC_GT0 + j] * A[i][j];
            mom[i] = s1 - s2 - 1.5 * s3 - 2.0 / 3.0 * dK[i];
        }
    }
    /* psi4 (CodeGen/constraints.py:42-244): the tetrad from Gram-Schmidt on
     * r = (x,y,z), theta = (xz, yz, -(x^2+y^2)), phi = (-y, x, 0) in the
     * physical metric gt/chi; m = (theta + i phi)/sqrt(2); the real part
     * carries the LazEv sign convention of constraints.py:238-244. */
    if (psi4) {
        const double x = xyz[0], y = xyz[1], z = xyz[2];
        double gd[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) gd[i][j] = g[i][j] / chi;
#define GDOT(u, v) (gd[0][0] * (u)[0] * (v)[0] + gd[0][1] * (u)[0] * (v)[1] + gd[0][2] * (u)[0] * (v)[2] + \
                    gd[1][0] * (u)[1] * (v)[0] + gd[1][1] * (u)[1] * (v)[1] + gd[1][2] * (u)[1] * (v)[2] + \
                    gd[2][0] * (u)[2] * (v)[0] + gd[2][1] * (u)[2] * (v)[1] + gd[2][2] * (u)[2] * (v)[2])
        double rv[3] = {x, y, z};
        double th[3] = {x * z, y * z, -(x * x + y * y)};
        double ph[3] = {-y, x, 0.0};
        const double nr = sqrt(GDOT(rv, rv));
        for (int i = 0; i < 3; i++) rv[i] /= nr;
        {
            const double i1 = GDOT(th, th), i2 = GDOT(rv, th);
            for (int i = 0; i < 3; i++) th[i] -= i2 * rv[i];
            const double n = sqrt(i1 - i2 * i2);
            for (int i = 0; i < 3; i++) th[i] /= n;
        }
        {
            const double i1 = GDOT(ph, ph), i2 = GDOT(rv, ph), i3 = GDOT(th, ph);
            for (int i = 0; i < 3; i++) ph[i] -= i2 * rv[i] + i3 * th[i];
            const double n = sqrt(i1 - i2 * i2 - i3 * i3);
            for (int i = 0; i < 3; i++) ph[i] /= n;
        }
#undef GDOT
        const double is2 = 1.0 / sqrt(2.0);
        double mr[3], mi[3];
        for (int i = 0; i < 3; i++) {
            mr[i] = th[i] * is2;
            mi[i] = ph[i] * is2;
        }
        double MM[3][3], NN[3][3], MR[3][3], NR[3][3], Uu[3][3], Vv[3][3], Av[3];
        for (int i = 0; i < 3; i++) {
            Av[i] = 0.0;
            for (int j = 0; j < 3; j++) Av[i] += A[i][j] * rv[j];
            for (int j = 0; j < 3; j++) {
                MM[i][j] = mr[i] * mr[j] - mi[i] * mi[j];
                NN[i][j] = mr[i] * mi[j] + mr[j] * mi[i];
                MR[i][j] = mr[i] * rv[j] - mr[j] * rv[i];
                NR[i][j] = mi[i] * rv[j] - mi[j] * rv[i];
                /* U_ij = m_k (d_j At_ki + C2^m_ki At_mj) */
                double u = 0.0, v = 0.0;
                for (int k = 0; k < 3; k++) {
                    double t = dA[j][k][i];
                    for (int m = 0; m < 3; m++) t += C2[m][k][i] * A[m][j];
                    u += mr[k] * t;
                    v += mi[k] * t;
                }
                Uu[i][j] = u;
                Vv[i][j] = v;
            }
        }
        double rA = 0.0, rdchi = 0.0, mrd = 0.0, mid = 0.0, mrA = 0.0, miA = 0.0;
        for (int i = 0; i < 3; i++) {
            rA += Av[i] * rv[i];
            rdchi += rv[i] * dchi[i];
            mrd += mr[i] * dchi[i];
            mid += mi[i] * dchi[i];
            mrA += mr[i] * Av[i];
            miA += mi[i] * Av[i];
        }
        const double Atmp = (rA + K * chi / 3.0 + 0.5 * rdchi) / (chi * chi);
        double p1r = 0.0, p1i = 0.0, p2r = 0.0, p2i = 0.0, p3r = 0.0, p3i = 0.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                p1r += R[i][j] * MM[i][j];
                p1i += R[i][j] * NN[i][j];
                p2r += A[i][j] * MM[i][j];
                p2i += A[i][j] * NN[i][j];
                p3r += MR[i][j] * Uu[i][j] - NR[i][j] * Vv[i][j];
                p3i += NR[i][j] * Uu[i][j] + MR[i][j] * Vv[i][j];
            }
        p2r *= Atmp;
        p2i *= Atmp;
        p3r /= chi;
        p3i /= chi;
        const double p4r = (mrA * (mrA + 0.5 * mrd) - miA * (miA + 0.5 * mid)) / (chi * chi);
        const double p4i = (mrA * (miA + 0.5 * mid) + miA * (mrA + 0.5 * mrd)) / (chi * chi);
        psi4[0] = -(p1r + p2r - p3r - p4r);
        psi4[1] = p1i + p2i - p3i - p4i;
    }
    if (!rhs) return;

    double divb = db[0][0] + db[1][1] + db[2][2];
    const double w = -2.0 / 3.0; /* Lie-derivative weight (bssn.py:39) */

    /* a_rhs (bssn.py:100) */
    double adv_a = 0.0, adv_chi = 0.0, adv_K = 0.0;
    for (int i = 0; i < 3; i++) {
        adv_a += b[i] * da[i];
        adv_chi += b[i] * dchi[i];
        adv_K += b[i] * dK[i];
    }
    rhs[ORC_ALPHA] = l1 * adv_a - 2.0 * alpha * K;
    if (p->ssl) { /* slow-start lapse (bssn.py:90-99): - W h exp(-t^2/(2 sig^2)) (alpha - W), W = chi^(1/2) */
        const double W = sqrt(chi), hs = 0.6, sig = 20.0;
        rhs[ORC_ALPHA] -= W * (hs * exp(-(p->time * p->time) / (2.0 * sig * sig))) * (alpha - W);
    }

    /* b_rhs (bssn.py:102-105); the Rochester gauge's (bssn.py:443-450) */
    for (int i = 0; i < 3; i++) {
        double adv = 0.0;
        for (int j = 0; j < 3; j++) adv += b[j] * db[j][i];
        if (p->rochester)
            rhs[ORC_BETA0 + i] = p->xi[1] * adv + 0.75 * p->xi[2] * f[ORC_GT0 + i] - eta * b[i];
        else
            rhs[ORC_BETA0 + i] = 0.75 * (lf0 + lf1 * alpha) * B[i] + l2 * adv;
    }

    /* chi_rhs (bssn.py:109) */
    rhs[ORC_CHI] = adv_chi + w * chi * divb + 2.0 / 3.0 * chi * alpha * K;

    /* trace-free part of alpha R - D_iD_j alpha (dendro.py:352-365) */
    double X[3][3], trX = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) X[i][j] = alpha * R[i][j] - DDa[i][j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) trX += igt[i][j] * X[i][j];

    /* gt_rhs, At_rhs (bssn.py:107, 135-139) with lie() of dendro.py:238-239 */
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) {
            double lg = 0.0, lA = 0.0;
            for (int k = 0; k < 3; k++) {
                lg += b[k] * dg[k][i][j] + g[i][k] * db[j][k] + g[k][j] * db[i][k] + w * g[i][j] * db[k][k];
                lA += b[k] * dA[k][i][j] + A[i][k] * db[j][k] + A[k][j] * db[i][k] + w * A[i][j] * db[k][k];
            }
            rhs[ORC_SYMGT0 + SYM6[i][j]] = lg - 2.0 * alpha * A[i][j];
            const double tf = X[i][j] - g[i][j] * trX / 3.0;
            rhs[ORC_SYMAT0 + SYM6[i][j]] = lA + chi * tf + alpha * (K * A[i][j] - 2.0 * AikAkj[i][j]);
        }

    /* K_rhs (bssn.py:141-145): laplacian with the physical inverse metric chi*igt (dendro.py:288-290) */
    double lap = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) lap += chi * igt[i][j] * DDa[i][j];
    rhs[ORC_K] = adv_K - lap + alpha * (K * K / 3.0 + sqrA);

    /* Gt_rhs (bssn.py:149-200), B_rhs (bssn.py:204-212) */
    for (int i = 0; i < 3; i++) {
        double s = 0.0;
        for (int j = 0; j < 3; j++) s += b[j] * dGt[j][i];
        for (int j = 0; j < 3; j++) s -= CalGt[j] * db[j][i];
        s += 2.0 / 3.0 * CalGt[i] * divb;
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) s += igt[j][k] * d2b[i][j][k] + igt[i][j] * d2b[k][j][k] / 3.0;
        for (int j = 0; j < 3; j++) s -= 2.0 * AUU[i][j] * da[j];
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) s += 2.0 * alpha * C2[i][j][k] * AUU[j][k];
        for (int j = 0; j < 3; j++) s -= alpha * (3.0 / chi * AUU[i][j] * dchi[j] + 4.0 / 3.0 * igt[i][j] * dK[j]);
        rhs[ORC_GT0 + i] = s;
        double advB = 0.0, advGt = 0.0;
        for (int j = 0; j < 3; j++) {
            advB += b[j] * dB[j][i];
            advGt += b[j] * dGt[j][i];
        }
        /* the Rochester gauge freezes B (bssn.py:466, B_rhs = 0) */
        rhs[ORC_B0 + i] = p->rochester ? 0.0 : s - eta * B[i] + l3 * advB - l4 * advGt;
    }
}

void orc_bssn_point(const double *in, const orc_params *p, double *rhs) { point_algebra(in, p, rhs, NULL, NULL, NULL, NULL); }

void orc_ham_point(const double *in, double *ham) { point_algebra(in, NULL, NULL, ham, NULL, NULL, NULL); }

void orc_constraint_point(const double *in, double *ham, double *mom) {
    point_algebra(in, NULL, NULL, ham, mom, NULL, NULL);
}

void orc_constraint_point_xyz(const double *in, const double *xyz, double *ham, double *mom, double *psi4) {
    point_algebra(in, NULL, NULL, ham, mom, xyz, psi4);
}

void orc_constraint_points_xyz(const double *in, const double *xyz, int64_t npts, double *ham, double *mom,
                               double *psi4) {
    for (int64_t q = 0; q < npts; q++)
        orc_constraint_point_xyz(in + q * ORC_POINT_NIN, xyz + 3 * q, ham + q, mom + 3 * q, psi4 + 2 * q);
}

void orc_constraint_points(const double *in, int64_t npts, double *ham, double *mom) {
    for (int64_t q = 0; q < npts; q++) orc_constraint_point(in + q * ORC_POINT_NIN, ham + q, mom + 3 * q);
}

void orc_ham_points(const double *in, int64_t npts, double *ham) {
    for (int64_t q = 0; q < npts; q++) orc_ham_point(in + q * ORC_POINT_NIN, ham + q);
}

/* The per-point algebra used by orc_bssnrhs_block. Fixture generation swaps in
 * the reference's generated equations (oracle/_ref) through this hook. */
static orc_point_fn g_point_fn = orc_bssn_point;
void orc_set_point_algebra(orc_point_fn fn) { g_point_fn = fn ? fn : orc_bssn_point; }

/* The whole-block algebra hook: when set, orc_bssnrhs_block hands the block's
 * field and derivative arrays to it instead of running the point loop (the
 * CPU baseline runs the reference's generated equations over the block this
 * way, as rhs.cpp:193-269 does). */
static orc_block_algebra_fn g_block_fn = NULL;
void orc_set_block_algebra(orc_block_algebra_fn fn) { g_block_fn = fn; }

/* Wall-clock seconds per phase of orc_bssnrhs_block, summed over blocks and
 * threads, at the reference's timer boundaries: t_deriv around the derivative
 * calls (rhs.cpp:171-182), t_rhs around the equations (:193-270), t_bdyc
 * around the boundary pass (:273-329), then the KO derivatives and sum
 * (:332-492; the reference books those under t_deriv and t_rhs, kept apart
 * here as "ko"). */
static double g_phase[4];
static double wall(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
static void phase_add(int k, double dt) {
#ifdef _OPENMP
#pragma omp atomic
#endif
    g_phase[k] += dt;
}
void orc_phase_times(double *out4, int reset) {
    for (int k = 0; k < 4; k++) {
        out4[k] = g_phase[k];
        if (reset) g_phase[k] = 0.0;
    }
}

void orc_bssn_points(const double *in, int64_t npts, const orc_params *p, double *rhs) {
    for (int64_t q = 0; q < npts; q++) orc_bssn_point(in + q * ORC_POINT_NIN, p, rhs + q * 24);
}

/* ======================================================================== */
/* Per-block RHS                                                              */
/* ======================================================================== */

/* The 11 variables with second derivatives (BSSN_GR/scripts/bssnrhs_derivs.h). */
/* The damping of the reference's BSSN_ETA_FUNCTION build (CodeGen/bssn.py:
 * 71-75, generated as bssneqs_eta_func_standard_gauge.cpp, which replaces the
 * RIT eta(r) of rhs.cpp:209-215): R0 sqrt(gt^ij d_i chi d_j chi) / (1 - chi^p0)^p1
 * at the point whose algebra input is `in` (fields, first derivatives). */
double orc_eta_func(const double *in, const orc_params *p) {
    const double *g = in + ORC_SYMGT0;  /* gt 00 01 02 11 12 22 */
    const double c00 = g[3] * g[5] - g[4] * g[4], c01 = g[2] * g[4] - g[1] * g[5], c02 = g[1] * g[4] - g[2] * g[3];
    const double c11 = g[0] * g[5] - g[2] * g[2], c12 = g[1] * g[2] - g[0] * g[4], c22 = g[0] * g[3] - g[1] * g[1];
    const double det = g[0] * c00 + g[1] * c01 + g[2] * c02;
    const double *dc = in + 24 + 3 * ORC_CHI;
    const double q = (c00 * dc[0] * dc[0] + c11 * dc[1] * dc[1] + c22 * dc[2] * dc[2] +
                      2.0 * (c01 * dc[0] * dc[1] + c02 * dc[0] * dc[2] + c12 * dc[1] * dc[2])) / det;
    return p->eta_r0 * sqrt(q) * pow(1.0 - pow(in[ORC_CHI], p->eta_power[0]), -p->eta_power[1]);
}

static const int D2VARS[11] = {ORC_ALPHA, ORC_CHI, ORC_BETA0, ORC_BETA0 + 1, ORC_BETA0 + 2,
                               ORC_SYMGT0, ORC_SYMGT0 + 1, ORC_SYMGT0 + 2, ORC_SYMGT0 + 3,
                               ORC_SYMGT0 + 4, ORC_SYMGT0 + 5};

/* Radiative boundary condition on the flagged faces (rhs.cpp:511-631). */
static void bssn_bcs(double *f_rhs, const double *f, const double *dxf, const double *dyf,
                     const double *dzf, const double *pmin, const double *pmax, double f_falloff,
                     double f_asymptotic, const uint32_t *sz, uint32_t bflag) {
    const int nx = (int)sz[0], ny = (int)sz[1], nz = (int)sz[2];
    const double hx = (pmax[0] - pmin[0]) / (nx - 1);
    const double hy = (pmax[1] - pmin[1]) / (ny - 1);
    const double hz = (pmax[2] - pmin[2]) / (nz - 1);
    const int PW = 3;
    const int ie = nx - PW, je = ny - PW, ke = nz - PW;
    for (int face = 0; face < 6; face++) {
        if (!(bflag & (1u << face))) continue;
        const int axis = face >> 1;
        const int fixed = (face & 1) ? ((axis == 0 ? ie : axis == 1 ? je : ke) - 1) : PW;
        for (int b2 = PW; b2 < (axis == 2 ? ny : nz) - PW; b2++)
            for (int b1 = PW; b1 < (axis == 0 ? ny : nx) - PW; b1++) {
                int i, j, k;
                if (axis == 0) { i = fixed; j = b1; k = b2; }
                else if (axis == 1) { i = b1; j = fixed; k = b2; }
                else { i = b1; j = b2; k = fixed; }
                const double x = pmin[0] + i * hx, y = pmin[1] + j * hy, z = pmin[2] + k * hz;
                const int64_t pp = IDX3(i, j, k);
                const double inv_r = 1.0 / sqrt(x * x + y * y + z * z);
                f_rhs[pp] = -inv_r * (x * dxf[pp] + y * dyf[pp] + z * dzf[pp] +
                                      f_falloff * (f[pp] - f_asymptotic));
            }
    }
}

/* (falloff, asymptotic value) per variable: rhs.cpp:275-327 */
static void bc_consts(int v, double *falloff, double *asym) {
    switch (v) {
    case ORC_ALPHA: case ORC_CHI: *falloff = 1.0; *asym = 1.0; return;
    case ORC_K: *falloff = 1.0; *asym = 0.0; return;
    case ORC_BETA0: case ORC_BETA0 + 1: case ORC_BETA0 + 2: *falloff = 1.0; *asym = 0.0; return;
    case ORC_GT0: case ORC_GT0 + 1: case ORC_GT0 + 2: *falloff = 2.0; *asym = 0.0; return;
    case ORC_B0: case ORC_B0 + 1: case ORC_B0 + 2: *falloff = 1.0; *asym = 0.0; return;
    default: break;
    }
    if (v >= ORC_SYMAT0) { *falloff = 2.0; *asym = 0.0; return; }
    /* gt: diagonal -> 1, off-diagonal -> 0 */
    const int c = v - ORC_SYMGT0;
    *falloff = 1.0;
    *asym = (c == 0 || c == 3 || c == 5) ? 1.0 : 0.0;
}

static int is_gauge_var(int v) {
    return v == ORC_ALPHA || (v >= ORC_BETA0 && v <= ORC_B0 + 2);
}

/* Per-thread scratch reused across blocks: slot 0 the derivative workspace,
 * slot 1 the RK4 block buffers. The reference allocates its workspace once
 * (grUtils.cpp:2374-2396); a fresh allocation per block would make the large
 * blocks page-fault bound. Held until the thread ends. */
static _Thread_local double *g_scratch[2];
static _Thread_local size_t g_scratch_n[2];
static double *scratch(int slot, size_t n) {
    if (g_scratch_n[slot] < n) {
        free(g_scratch[slot]);
        g_scratch[slot] = (double *)malloc(sizeof(double) * n);
        g_scratch_n[slot] = g_scratch[slot] ? n : 0;
    }
    return g_scratch[slot];
}

/* All first derivatives and the second/mixed derivatives of the 11 D2VARS
 * into `ws` (bssnrhs_derivs.h; the constraints use the same calls,
 * constraint_derivs.h): 72 + 66 arrays of the block's size. */
static void block_derivs(const double *const *u, const uint32_t *sz, uint32_t bflag, double hx, double hy,
                         double hz, double *ws, double *d1[24][3], double *d2[11][6]) {
    const int64_t n = (int64_t)sz[0] * sz[1] * sz[2];
    for (int v = 0; v < 24; v++)
        for (int d = 0; d < 3; d++) d1[v][d] = ws + (size_t)n * (3 * v + d);
    for (int m = 0; m < 11; m++)
        for (int c = 0; c < 6; c++) d2[m][c] = ws + (size_t)n * (72 + 6 * m + c);
    for (int v = 0; v < 24; v++) {
        orc_deriv644_x(d1[v][0], u[v], hx, sz, bflag);
        orc_deriv644_y(d1[v][1], u[v], hy, sz, bflag);
        orc_deriv644_z(d1[v][2], u[v], hz, sz, bflag);
    }
    for (int m = 0; m < 11; m++) {
        const int v = D2VARS[m];
        orc_deriv644_xx(d2[m][0], u[v], hx, sz, bflag);
        orc_deriv644_y(d2[m][1], d1[v][0], hy, sz, bflag);  /* grad2_0_1 = D_y(D_x u) */
        orc_deriv644_z(d2[m][2], d1[v][0], hz, sz, bflag);  /* grad2_0_2 = D_z(D_x u) */
        orc_deriv644_yy(d2[m][3], u[v], hy, sz, bflag);
        orc_deriv644_z(d2[m][4], d1[v][1], hz, sz, bflag);  /* grad2_1_2 = D_z(D_y u) */
        orc_deriv644_zz(d2[m][5], u[v], hz, sz, bflag);
    }
}

int orc_bssnrhs_block(double *const *rhs_all, const double *const *u_all, const orc_block *blk,
                      const orc_params *p) {
    const uint32_t *sz = blk->sz;
    const int nx = (int)sz[0], ny = (int)sz[1], nz = (int)sz[2];
    if (nx < 13 || ny < 13 || nz < 13) return -1; /* 6th-order closures need >= 7 interior */
    const int64_t n = (int64_t)nx * ny * nz;
    const double *pmin = blk->pmin, *pmax = blk->pmax;
    const double hx = (pmax[0] - pmin[0]) / (nx - 1);
    const double hy = (pmax[1] - pmin[1]) / (ny - 1);
    const double hz = (pmax[2] - pmin[2]) / (nz - 1);
    const uint32_t bflag = blk->bflag;
    const int PW = 3;

    const double *u[24];
    double *rhs[24];
    for (int v = 0; v < 24; v++) {
        u[v] = u_all[v] + blk->offset;
        rhs[v] = rhs_all[v] + blk->offset;
    }

    /* derivative workspace: 72 first + 66 second (rhs.cpp:171-180) */
    double t0 = wall();
    double *ws = scratch(0, (size_t)n * (72 + 66 + 3));
    if (!ws) return -2;
    double *d1[24][3], *d2[11][6];
    block_derivs(u, sz, bflag, hx, hy, hz, ws, d1, d2);
    double *kox = ws + (size_t)n * 138, *koy = kox + n, *koz = koy + n;
    double t1 = wall();
    phase_add(0, t1 - t0);

    /* pointwise equations over the interior (rhs.cpp:194-269) */
    double in[ORC_POINT_NIN], out[24];
    if (g_block_fn) {
        const double hs[3] = {hx, hy, hz};
        g_block_fn(u, (const double *const *)&d1[0][0], (const double *const *)&d2[0][0], sz, pmin, hs, p, rhs);
    } else
    for (int k = PW; k < nz - PW; k++)
        for (int j = PW; j < ny - PW; j++)
            for (int i = PW; i < nx - PW; i++) {
                const double x = pmin[0] + i * hx;
                const double y = pmin[1] + j * hy;
                const double z = pmin[2] + k * hz;
                const int64_t pp = IDX3(i, j, k);
                const double r_coord = sqrt(x * x + y * y + z * z);
                const double wr = r_coord / p->eta_width;
                const double arg = -wr * wr * wr * wr;
                const double eta = (p->eta_central - p->eta_outer) * exp(arg) + p->eta_outer;
                for (int v = 0; v < 24; v++) {
                    in[v] = u[v][pp];
                    for (int d = 0; d < 3; d++) in[24 + 3 * v + d] = d1[v][d][pp];
                }
                for (int m = 0; m < 11; m++)
                    for (int c = 0; c < 6; c++) in[96 + 6 * m + c] = d2[m][c][pp];
                in[162] = p->eta_func ? orc_eta_func(in, p) : eta;
                g_point_fn(in, p, out);
                for (int v = 0; v < 24; v++) rhs[v][pp] = out[v];
            }

    t0 = wall();
    phase_add(1, t0 - t1);

    /* boundary fix-up (rhs.cpp:272-330) */
    if (bflag != 0) {
        for (int v = 0; v < 24; v++) {
            double fo, fa;
            bc_consts(v, &fo, &fa);
            bssn_bcs(rhs[v], u[v], d1[v][0], d1[v][1], d1[v][2], pmin, pmax, fo, fa, sz, bflag);
        }
    }

    t1 = wall();
    phase_add(2, t1 - t0);

    /* Kreiss-Oliger dissipation (rhs.cpp:333, 338-492) */
    for (int v = 0; v < 24; v++) {
        orc_ko_deriv42_x(kox, u[v], hx, sz, bflag);
        orc_ko_deriv42_y(koy, u[v], hy, sz, bflag);
        orc_ko_deriv42_z(koz, u[v], hz, sz, bflag);
        for (int k = PW; k < nz - PW; k++)
            for (int j = PW; j < ny - PW; j++)
                for (int i = PW; i < nx - PW; i++) {
                    const int64_t pp = IDX3(i, j, k);
                    double sigma = p->ko_sigma;
                    if (p->cako_enabled) {
                        sigma = sqrt(u[ORC_CHI][pp]);
                        sigma *= is_gauge_var(v) ? p->eps_cako_gauge : p->eps_cako_other;
                    }
                    rhs[v][pp] += sigma * (kox[pp] + koy[pp] + koz[pp]);
                }
    }
    phase_add(3, wall() - t1);
    return 0;
}

int orc_bssnrhs_batch(double *const *rhs, const double *const *u, const orc_block *blks,
                      int64_t nblocks, const orc_params *p, int nthreads) {
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (int64_t b = 0; b < nblocks; b++) err |= (orc_bssnrhs_block(rhs, u, &blks[b], p) != 0);
    (void)nthreads;
    return err ? -1 : 0;
}

/* ======================================================================== */
/* Hamiltonian constraint per block: physical_constraints, physcon.cpp:12-89  */
/* (C_HAM only; interior points, padding of `ham` untouched)                  */
/* ======================================================================== */
static int constraint_block(double *ham_all, double *const *mom_all, double *const *psi4_all,
                            const double *const *u_all, const orc_block *blk);

int orc_ham_block(double *ham_all, const double *const *u_all, const orc_block *blk) {
    return constraint_block(ham_all, NULL, NULL, u_all, blk);
}

/* physical_constraints (physcon.cpp:12-89): C_HAM, with mom_all != NULL
 * (three arrays laid out like a variable) C_MOM, with psi4_all != NULL (two
 * arrays) psi4 real and imaginary part at x = pmin + i*h (physcon.cpp:75-77),
 * zero within 1e-7 of the z axis (physcon.cpp:82-85). */
static int constraint_block(double *ham_all, double *const *mom_all, double *const *psi4_all,
                            const double *const *u_all, const orc_block *blk) {
    const uint32_t *sz = blk->sz;
    const int nx = (int)sz[0], ny = (int)sz[1], nz = (int)sz[2];
    if (nx < 13 || ny < 13 || nz < 13) return -1;
    const int64_t n = (int64_t)nx * ny * nz;
    const double hx = (blk->pmax[0] - blk->pmin[0]) / (nx - 1);
    const double hy = (blk->pmax[1] - blk->pmin[1]) / (ny - 1);
    const double hz = (blk->pmax[2] - blk->pmin[2]) / (nz - 1);
    const int PW = 3;
    const double *u[24];
    for (int v = 0; v < 24; v++) u[v] = u_all[v] + blk->offset;
    double *ham = ham_all + blk->offset;
    double *ws = scratch(0, (size_t)n * (72 + 66 + 3));
    if (!ws) return -2;
    double *d1[24][3], *d2[11][6];
    block_derivs(u, sz, blk->bflag, hx, hy, hz, ws, d1, d2);
    double in[ORC_POINT_NIN];
    for (int k = PW; k < nz - PW; k++)
        for (int j = PW; j < ny - PW; j++)
            for (int i = PW; i < nx - PW; i++) {
                const int64_t pp = IDX3(i, j, k);
                for (int v = 0; v < 24; v++) {
                    in[v] = u[v][pp];
                    for (int d = 0; d < 3; d++) in[24 + 3 * v + d] = d1[v][d][pp];
                }
                for (int m = 0; m < 11; m++)
                    for (int c = 0; c < 6; c++) in[96 + 6 * m + c] = d2[m][c][pp];
                in[162] = 0.0;
                double m[3], ps[2];
                const double xyz[3] = {blk->pmin[0] + i * hx, blk->pmin[1] + j * hy, blk->pmin[2] + k * hz};
                orc_constraint_point_xyz(in, xyz, &ham[pp], m, psi4_all ? ps : NULL);
                if (mom_all)
                    for (int c = 0; c < 3; c++) mom_all[c][blk->offset + pp] = m[c];
                if (psi4_all) {
                    if (fabs(xyz[0]) <= 1e-7 && fabs(xyz[1]) <= 1e-7) ps[0] = ps[1] = 0.0;
                    for (int c = 0; c < 2; c++) psi4_all[c][blk->offset + pp] = ps[c];
                }
            }
    return 0;
}

int orc_constraints_batch(double *ham, double *const *mom, double *const *psi4, const double *const *u,
                          const orc_block *blks, int64_t nblocks, int nthreads) {
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (int64_t b = 0; b < nblocks; b++) err |= (constraint_block(ham, mom, psi4, u, &blks[b]) != 0);
    (void)nthreads;
    return err ? -1 : 0;
}

int orc_ham_batch(double *ham, const double *const *u, const orc_block *blks, int64_t nblocks, int nthreads) {
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (int64_t b = 0; b < nblocks; b++) err |= (orc_ham_block(ham, u, &blks[b]) != 0);
    (void)nthreads;
    return err ? -1 : 0;
}

/* computeConstraintL2Norm (BSSN_GR/include/grUtils.tcc:37-134): sqrt of the
 * mean of c^2 over the nodes (here: interior points of every block) whose
 * distance to BH 1 is >= r1 and to BH 2 is >= r2. bh = {x1,y1,z1,x2,y2,z2}.
 * Returns -1 when no point qualifies. */
/* A node on a face, edge or corner of block b's interior (index PW or n-PW-1
 * along each such axis, side s_a = -1 or +1) is also a node of every
 * neighbour lying in a direction d with d_a in {0, s_a} along those axes and
 * d_a = 0 along the others; same-size neighbours holding one node are all
 * linked to each other, so each of them sees the same set, and the node is
 * counted by the one with the lowest block index. A node shared by
 * neighbouring blocks is thus counted once, as computeConstraintL2Norm counts
 * each zipped node once (grUtils.tcc:60-110, its `accumulated` flags), also
 * where a lattice has holes. nbr: 27 per block, -1 for none (NULL: blocks
 * are taken as disjoint). */
static int owned_elsewhere(const int32_t *nbr, int64_t b, int sx, int sy, int sz) {
    if (!nbr || !(sx | sy | sz)) return 0;
    for (int dz = sz < 0 ? sz : 0; dz <= (sz > 0 ? sz : 0); dz++)
        for (int dy = sy < 0 ? sy : 0; dy <= (sy > 0 ? sy : 0); dy++)
            for (int dx = sx < 0 ? sx : 0; dx <= (sx > 0 ? sx : 0); dx++) {
                if (!(dx | dy | dz)) continue;
                const int32_t o = nbr[27 * b + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
                if (o >= 0 && o < b) return 1;
            }
    return 0;
}

double orc_constraint_l2_norm_nbr(const double *c_all, const orc_block *blks, int64_t nblocks, const int32_t *nbr,
                                  const double *bh, double r1, double r2) {
    double l2 = 0.0;
    int64_t cnt = 0;
    for (int64_t b = 0; b < nblocks; b++) {
        const orc_block *blk = &blks[b];
        const uint32_t *sz = blk->sz;
        const int nx = (int)sz[0], ny = (int)sz[1], nz = (int)sz[2];
        const double hx = (blk->pmax[0] - blk->pmin[0]) / (nx - 1);
        const double hy = (blk->pmax[1] - blk->pmin[1]) / (ny - 1);
        const double hz = (blk->pmax[2] - blk->pmin[2]) / (nz - 1);
        const double *c = c_all + blk->offset;
        for (int k = 3; k < nz - 3; k++)
            for (int j = 3; j < ny - 3; j++)
                for (int i = 3; i < nx - 3; i++) {
                    if (owned_elsewhere(nbr, b, i == 3 ? -1 : (i == nx - 4 ? 1 : 0), j == 3 ? -1 : (j == ny - 4 ? 1 : 0),
                                        k == 3 ? -1 : (k == nz - 4 ? 1 : 0)))
                        continue;
                    const double x = blk->pmin[0] + i * hx, y = blk->pmin[1] + j * hy, z = blk->pmin[2] + k * hz;
                    const double d1 = sqrt((x - bh[0]) * (x - bh[0]) + (y - bh[1]) * (y - bh[1]) + (z - bh[2]) * (z - bh[2]));
                    const double d2 = sqrt((x - bh[3]) * (x - bh[3]) + (y - bh[4]) * (y - bh[4]) + (z - bh[5]) * (z - bh[5]));
                    if (d1 < r1 || d2 < r2) continue;
                    const double v = c[IDX3(i, j, k)];
                    l2 += v * v;
                    cnt++;
                }
    }
    return cnt ? sqrt(l2 / (double)cnt) : -1.0;
}

double orc_constraint_l2_norm(const double *c_all, const orc_block *blks, int64_t nblocks, const double *bh,
                              double r1, double r2) {
    return orc_constraint_l2_norm_nbr(c_all, blks, nblocks, NULL, bh, r1, r2);
}

/* ======================================================================== */
/* Algebraic constraints: BSSN_GR/include/bssn_constraints.h:15-172           */
/* Returns 0, or -1 for det(gt) < 0, -2 for |det-1| > 1e-6 after projection,  */
/* -3 for |tr A| > 1e-6 (where the reference calls exit(0)).                  */
/* ======================================================================== */
int orc_enforce_constraints_point(double *const *u, uint64_t node, double chi_floor) {
    const double one_third = 1.0 / 3.0;
    double gtd[3][3], Atd[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            gtd[i][j] = u[ORC_SYMGT0 + SYM6[i][j]][node];
            Atd[i][j] = u[ORC_SYMAT0 + SYM6[i][j]][node];
        }
    double det_gtd = gtd[0][0] * (gtd[1][1] * gtd[2][2] - gtd[1][2] * gtd[1][2]) -
                     gtd[0][1] * gtd[0][1] * gtd[2][2] + 2.0 * gtd[0][1] * gtd[0][2] * gtd[1][2] -
                     gtd[0][2] * gtd[0][2] * gtd[1][1];
    if (det_gtd < 0.0) return -1;
    const double det_gtd_to_neg_third = 1.0 / pow(det_gtd, one_third);
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) gtd[i][j] *= det_gtd_to_neg_third;
    det_gtd = gtd[0][0] * (gtd[1][1] * gtd[2][2] - gtd[1][2] * gtd[1][2]) -
              gtd[0][1] * gtd[0][1] * gtd[2][2] + 2.0 * gtd[0][1] * gtd[0][2] * gtd[1][2] -
              gtd[0][2] * gtd[0][2] * gtd[1][1];
    const double detgt_m1 = det_gtd - 1.0;
    if (fabs(detgt_m1) > 1.0e-6) return -2;
    double gtu[3][3];
    const double idet_gtd = 1.0 / det_gtd;
    gtu[0][0] = idet_gtd * (gtd[1][1] * gtd[2][2] - gtd[1][2] * gtd[1][2]);
    gtu[0][1] = idet_gtd * (-gtd[0][1] * gtd[2][2] + gtd[0][2] * gtd[1][2]);
    gtu[0][2] = idet_gtd * (gtd[0][1] * gtd[1][2] - gtd[0][2] * gtd[1][1]);
    gtu[1][1] = idet_gtd * (gtd[0][0] * gtd[2][2] - gtd[0][2] * gtd[0][2]);
    gtu[1][2] = idet_gtd * (-gtd[0][0] * gtd[1][2] + gtd[0][1] * gtd[0][2]);
    gtu[2][2] = idet_gtd * (gtd[0][0] * gtd[1][1] - gtd[0][1] * gtd[0][1]);
    const double one_third_trace_Atd =
        one_third * (Atd[0][0] * gtu[0][0] + Atd[1][1] * gtu[1][1] + Atd[2][2] * gtu[2][2] +
                     2.0 * (Atd[0][1] * gtu[0][1] + Atd[0][2] * gtu[0][2] + Atd[1][2] * gtu[1][2]));
    Atd[0][0] -= one_third_trace_Atd * gtd[0][0];
    Atd[0][1] -= one_third_trace_Atd * gtd[0][1];
    Atd[0][2] -= one_third_trace_Atd * gtd[0][2];
    Atd[1][1] -= one_third_trace_Atd * gtd[1][1];
    Atd[1][2] -= one_third_trace_Atd * gtd[1][2];
    Atd[2][2] -= one_third_trace_Atd * gtd[2][2];
    const double tr_A = Atd[0][0] * gtu[0][0] + Atd[1][1] * gtu[1][1] + Atd[2][2] * gtu[2][2] +
                        2.0 * (Atd[0][1] * gtu[0][1] + Atd[0][2] * gtu[0][2] + Atd[1][2] * gtu[1][2]);
    if (fabs(tr_A) > 1.0e-6) return -3;
    u[ORC_SYMAT0 + 0][node] = Atd[0][0];
    u[ORC_SYMAT0 + 1][node] = Atd[0][1];
    u[ORC_SYMAT0 + 2][node] = Atd[0][2];
    u[ORC_SYMAT0 + 3][node] = Atd[1][1];
    u[ORC_SYMAT0 + 4][node] = Atd[1][2];
    u[ORC_SYMAT0 + 5][node] = Atd[2][2];
    u[ORC_SYMGT0 + 0][node] = gtd[0][0];
    u[ORC_SYMGT0 + 1][node] = gtd[0][1];
    u[ORC_SYMGT0 + 2][node] = gtd[0][2];
    u[ORC_SYMGT0 + 3][node] = gtd[1][1];
    u[ORC_SYMGT0 + 4][node] = gtd[1][2];
    u[ORC_SYMGT0 + 5][node] = gtd[2][2];
    if (u[ORC_CHI][node] < chi_floor) u[ORC_CHI][node] = chi_floor;
    /* std::max(alpha, CHI_FLOOR), bssn_constraints.h:171: (alpha < floor) ? floor : alpha, so a NaN stays */
    if (u[ORC_ALPHA][node] < chi_floor) u[ORC_ALPHA][node] = chi_floor;
    return 0;
}

/* ======================================================================== */
/* Classic RK4 with the reference's stage table (include/rkBSSN.h:50-53) and  */
/* stage loop (src/rkBSSN.cpp:764-860), per block with the padding frozen at  */
/* its value at the start of the step (blocks have no neighbours to unzip).   */
/* ======================================================================== */
static const double RK4_C[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
static const double RK4_U[4] = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};

static int rk4_block(double *const *u_all, const orc_block *blk, double dt, const orc_params *p,
                     int enforce_per_stage) {
    const int nx = (int)blk->sz[0], ny = (int)blk->sz[1], nz = (int)blk->sz[2];
    const int64_t n = (int64_t)nx * ny * nz;
    const int PW = 3;
    double *buf = scratch(1, (size_t)n * 24 * 3);
    if (!buf) return -2;
    double *u0[24], *us[24], *k[24], *ucur[24];
    for (int v = 0; v < 24; v++) {
        u0[v] = buf + (size_t)n * v;
        us[v] = buf + (size_t)n * (24 + v);
        k[v] = buf + (size_t)n * (48 + v);
        memcpy(u0[v], u_all[v] + blk->offset, sizeof(double) * (size_t)n);
        memcpy(us[v], u0[v], sizeof(double) * (size_t)n);
        memset(k[v], 0, sizeof(double) * (size_t)n);
        ucur[v] = u_all[v] + blk->offset; /* accumulates u0 + sum C_s dt k_s */
    }
    orc_block local = *blk;
    local.offset = 0;
    int err = 0;
    for (int s = 0; s < 4 && !err; s++) {
        err = orc_bssnrhs_block(k, (const double *const *)us, &local, p);
        if (err) break;
        for (int kk = PW; kk < nz - PW; kk++)
            for (int j = PW; j < ny - PW; j++)
                for (int i = PW; i < nx - PW; i++) {
                    const int64_t pp = IDX3(i, j, kk);
                    for (int v = 0; v < 24; v++) ucur[v][pp] += (RK4_C[s] * dt * k[v][pp]);
                    if (s < 3) {
                        for (int v = 0; v < 24; v++) us[v][pp] = u0[v][pp] + (RK4_U[s + 1] * dt * k[v][pp]);
                        if (enforce_per_stage) {
                            const int e = orc_enforce_constraints_point(us, (uint64_t)pp, p->chi_floor);
                            if (e) err = e;
                        }
                    }
                }
    }
    if (!err) {
        for (int kk = PW; kk < nz - PW; kk++)
            for (int j = PW; j < ny - PW; j++)
                for (int i = PW; i < nx - PW; i++) {
                    const int e = orc_enforce_constraints_point(ucur, (uint64_t)IDX3(i, j, kk), p->chi_floor);
                    if (e) err = e;
                }
    }
    return err;
}

int orc_rk4_step(double *const *u, const orc_block *blks, int64_t nblocks, double dt,
                 const orc_params *p, int enforce_per_stage, int nthreads) {
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (int64_t b = 0; b < nblocks; b++) err |= (rk4_block(u, &blks[b], dt, p, enforce_per_stage) != 0);
    (void)nthreads;
    return err ? -1 : 0;
}

/* ======================================================================== */
/* Padding refresh from same-level neighbours and RK4 with it                 */
/* ======================================================================== */
/* The part of unzip (bssnCtx.cpp:75, dendrolib) that needs no interpolation:
 * a padding point of block b that lies in the interior of the neighbouring
 * block nb in direction (di,dj,dk) (same size and spacing, the two blocks
 * sharing their boundary nodes, so nb starts n-2*PW-1 nodes further along)
 * takes nb's value at the same coordinates. nbr holds 27 entries per block,
 * index (di+1) + 3(dj+1) + 9(dk+1), -1 where there is no such neighbour;
 * those padding points keep their values. */
void orc_halo_fill(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr) {
    const int PW = 3;
    for (int64_t b = 0; b < nblocks; b++) {
        const orc_block *blk = &blks[b];
        const int nx = (int)blk->sz[0], ny = (int)blk->sz[1], nz = (int)blk->sz[2];
        const int st[3] = {nx - 2 * PW - 1, ny - 2 * PW - 1, nz - 2 * PW - 1};
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    const int di = i < PW ? -1 : (i >= nx - PW ? 1 : 0);
                    const int dj = j < PW ? -1 : (j >= ny - PW ? 1 : 0);
                    const int dk = k < PW ? -1 : (k >= nz - PW ? 1 : 0);
                    if (!di && !dj && !dk) continue;
                    const int32_t nb = nbr[27 * b + (di + 1) + 3 * (dj + 1) + 9 * (dk + 1)];
                    if (nb < 0) continue;
                    const int si = i - di * st[0], sj = j - dj * st[1], sk = k - dk * st[2];
                    const int64_t src = (int64_t)blks[nb].offset + si + (int64_t)nx * (sj + (int64_t)ny * sk);
                    const int64_t dst = (int64_t)blk->offset + i + (int64_t)nx * (j + (int64_t)ny * k);
                    for (int v = 0; v < 24; v++) u[v][dst] = u[v][src];
                }
    }
}

/* Lagrange weights of the 7 nodes of an order-6 element at t = code/2 (code in
 * 0..12), the element-local interpolation an unzip uses where a block's
 * padding lies inside a coarser block (the octree code is in dendrolib,
 * absent; this is the interpolation of the element's nodal polynomial). */
void orc_lagrange6_half(int code, double w[7]) {
    const double t = 0.5 * (double)code;
    for (int m = 0; m < 7; m++) {
        double num = 1.0, den = 1.0;
        for (int l = 0; l < 7; l++) {
            if (l == m) continue;
            num *= (t - (double)l);
            den *= (double)(m - l);
        }
        w[m] = num / den;
    }
}

/* Where padding point X of block b takes its value when no same-size
 * neighbour in its direction holds it: the block whose interior contains X,
 * preferring the same spacing (a neighbour of another size: copy), then half
 * the spacing (a finer neighbour: X is one of its nodes, injection), then
 * twice the spacing (a coarser neighbour: order-6 interpolation inside the
 * element of that block containing X); ties go to the lower block index.
 * Returns the kind (0 none, 1 copy, 2 interpolation) and the source. */
static int padding_source(const orc_block *blks, int64_t nblocks, int64_t b, const double X[3], int64_t *src_blk,
                          int q_out[3], int code[3]) {
    double hb[3];
    for (int a = 0; a < 3; a++) hb[a] = (blks[b].pmax[a] - blks[b].pmin[a]) / (blks[b].sz[a] - 1);
    int best = 3;
    for (int64_t c = 0; c < nblocks; c++) {
        if (c == b) continue;
        const orc_block *nb = &blks[c];
        double hc[3];
        for (int a = 0; a < 3; a++) hc[a] = (nb->pmax[a] - nb->pmin[a]) / (nb->sz[a] - 1);
        int rel = -1; /* 0 same, 1 finer, 2 coarser */
        if (fabs(hc[0] - hb[0]) <= 1e-9 * hb[0]) rel = 0;
        else if (fabs(2.0 * hc[0] - hb[0]) <= 1e-9 * hb[0]) rel = 1;
        else if (fabs(hc[0] - 2.0 * hb[0]) <= 1e-9 * hb[0]) rel = 2;
        if (rel < 0 || rel >= best) continue;
        int ok = 1, q[3], cd[3];
        for (int a = 0; a < 3 && ok; a++) {
            const double qa = (X[a] - nb->pmin[a]) / hc[a];
            const double q2 = 2.0 * qa;
            const double r2 = floor(q2 + 0.5);
            if (fabs(q2 - r2) > 2e-6) { ok = 0; break; }
            const long i2 = (long)r2;
            if (i2 < 2 * 3 || i2 > 2 * ((long)nb->sz[a] - 4)) { ok = 0; break; }
            if (rel < 2 && (i2 & 1)) { ok = 0; break; }
            if (rel < 2) {
                q[a] = (int)(i2 / 2);
            } else {
                /* element e of the block's interior (6 intervals each) holding X */
                const long t2 = i2 - 2 * 3, ne = ((long)nb->sz[a] - 7) / 6;
                long e = t2 / 12;
                if (e > ne - 1) e = ne - 1;
                q[a] = (int)(3 + 6 * e);     /* the element's first node */
                cd[a] = (int)(t2 - 12 * e);  /* 2 t, t the element-local coordinate */
            }
        }
        if (!ok) continue;
        best = rel;
        *src_blk = c;
        for (int a = 0; a < 3; a++) {
            q_out[a] = q[a];
            code[a] = rel == 2 ? cd[a] : 0;
        }
    }
    return best == 3 ? 0 : (best < 2 ? 1 : 2);
}

/* orc_halo_fill, then (general != 0) every padding point without a same-size
 * neighbour in its direction from padding_source: the copy part of the
 * reference's unzip for neighbours of any size and across one level jump
 * (BSSN_GR/src/bssnCtx.cpp:75; the octree code is in dendrolib, absent).
 * Reads interior points only and writes padding points only. */
void orc_padding_fill(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr, int general) {
    const int PW = 3;
    orc_halo_fill(u, blks, nblocks, nbr);
    if (!general) return;
    for (int64_t b = 0; b < nblocks; b++) {
        const orc_block *blk = &blks[b];
        const int nx = (int)blk->sz[0], ny = (int)blk->sz[1], nz = (int)blk->sz[2];
        double h[3];
        for (int a = 0; a < 3; a++) h[a] = (blk->pmax[a] - blk->pmin[a]) / (blk->sz[a] - 1);
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    const int di = i < PW ? -1 : (i >= nx - PW ? 1 : 0);
                    const int dj = j < PW ? -1 : (j >= ny - PW ? 1 : 0);
                    const int dk = k < PW ? -1 : (k >= nz - PW ? 1 : 0);
                    if (!di && !dj && !dk) continue;
                    if (nbr[27 * b + (di + 1) + 3 * (dj + 1) + 9 * (dk + 1)] >= 0) continue;
                    const double X[3] = {blk->pmin[0] + i * h[0], blk->pmin[1] + j * h[1], blk->pmin[2] + k * h[2]};
                    int64_t c = -1;
                    int q[3], cd[3];
                    const int kind = padding_source(blks, nblocks, b, X, &c, q, cd);
                    if (!kind) continue;
                    const orc_block *nb = &blks[c];
                    const int64_t mx = nb->sz[0], my = nb->sz[1];
                    const int64_t dst = (int64_t)blk->offset + i + (int64_t)nx * (j + (int64_t)ny * k);
                    const int64_t s0 = (int64_t)nb->offset + q[0] + mx * (q[1] + my * q[2]);
                    if (kind == 1) {
                        for (int v = 0; v < 24; v++) u[v][dst] = u[v][s0];
                        continue;
                    }
                    double wx[7], wy[7], wz[7];
                    orc_lagrange6_half(cd[0], wx);
                    orc_lagrange6_half(cd[1], wy);
                    orc_lagrange6_half(cd[2], wz);
                    for (int v = 0; v < 24; v++) {
                        double sz_ = 0.0;
                        for (int c3 = 0; c3 < 7; c3++) {
                            double sy = 0.0;
                            for (int b2 = 0; b2 < 7; b2++) {
                                double sx = 0.0;
                                const double *row = u[v] + s0 + mx * (b2 + my * c3);
                                for (int a1 = 0; a1 < 7; a1++) sx += wx[a1] * row[a1];
                                sy += wy[b2] * sx;
                            }
                            sz_ += wz[c3] * sy;
                        }
                        u[v][dst] = sz_;
                    }
                }
    }
}

/* The three-stage SSP Runge-Kutta of the reference's legacy solver
 * (BSSN_RK_TYPE = 0, rkBSSN.cpp:691-762): u1 = u0 + dt k(u0);
 * u2 = 3/4 u0 + 1/4 u1 + dt/4 k(u1); u3 = 1/3 u0 + 2/3 u2 + (2/3) dt k(u2), each
 * stage state followed by enforce_bssn_constraints (rkBSSN.cpp:705, 736, 760),
 * the products and sums in the order written there. With nbr, every stage
 * state's padding is refreshed by orc_padding_fill(general) before its right-hand
 * side and u's after the step (the reference unzips each stage state); without,
 * padding stays as it is. Returns 0, or -1 on a constraint violation. */
int orc_rk3_step(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr, int general,
                 double dt, const orc_params *p, int nthreads) {
    const int PW = 3;
    int64_t n = 0;
    for (int64_t b = 0; b < nblocks; b++) {
        const int64_t e = (int64_t)blks[b].offset + (int64_t)blks[b].sz[0] * blks[b].sz[1] * blks[b].sz[2];
        if (e > n) n = e;
    }
    double *buf = (double *)malloc(sizeof(double) * (size_t)n * 24 * 2);
    if (!buf) return -2;
    double *u0[24], *k[24];
    for (int v = 0; v < 24; v++) {
        u0[v] = buf + (size_t)n * v;
        k[v] = buf + (size_t)n * (24 + v);
        memset(k[v], 0, sizeof(double) * (size_t)n);
    }
    if (nbr) orc_padding_fill(u, blks, nblocks, nbr, general);
    for (int v = 0; v < 24; v++) memcpy(u0[v], u[v], sizeof(double) * (size_t)n);
    int err = 0;
    for (int s = 0; s < 3 && !err; s++) {
        if (s > 0 && nbr) orc_padding_fill(u, blks, nblocks, nbr, general);
#ifdef _OPENMP
        if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
        for (int64_t b = 0; b < nblocks; b++) {
            const int e = orc_bssnrhs_block(k, (const double *const *)u, &blks[b], p);
            err |= (e != 0);
        }
        /* the combination reads u at the same point only, so all blocks'
         * right-hand sides are taken first, then u is overwritten */
        for (int64_t b = 0; b < nblocks && !err; b++) {
            const orc_block *blk = &blks[b];
            const int nx = (int)blk->sz[0], ny = (int)blk->sz[1], nz = (int)blk->sz[2];
            for (int kk = PW; kk < nz - PW; kk++)
                for (int j = PW; j < ny - PW; j++)
                    for (int i = PW; i < nx - PW; i++) {
                        const int64_t pp = (int64_t)blk->offset + i + (int64_t)nx * (j + (int64_t)ny * kk);
                        for (int v = 0; v < 24; v++) {
                            if (s == 0)
                                u[v][pp] = u0[v][pp] + dt * k[v][pp];
                            else if (s == 1)
                                u[v][pp] = (0.75) * u0[v][pp] + 0.25 * u[v][pp] + dt * 0.25 * k[v][pp];
                            else
                                u[v][pp] = (1.0 / 3.0) * u0[v][pp] + (2.0 / 3.0) * u[v][pp] +
                                           dt * (2.0 / 3.0) * k[v][pp];
                        }
                        if (orc_enforce_constraints_point(u, (uint64_t)pp, p->chi_floor)) err = 1;
                    }
        }
    }
    (void)nthreads;
    if (!err && nbr) orc_padding_fill(u, blks, nblocks, nbr, general);
    free(buf);
    return err ? -1 : 0;
}

/* The embedded Runge-Kutta 4(5) of the reference (rkBSSN.cpp:861-1085), with
 * its coefficients (rkBSSN.h:57-71) as the reference writes them; note that
 * its fourth-order weight of stage 3 is 2197/4101 where Fehlberg's is
 * 2197/4104, so its fourth-order solution is not consistent and the error
 * estimate carries a term of order dt. */
static const double RK45_U[6][6] = {
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {1.0 / 4.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {3.0 / 32.0, 9.0 / 32.0, 0.0, 0.0, 0.0, 0.0},
    {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0.0, 0.0, 0.0},
    {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0, 0.0, 0.0},
    {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0, 0.0}};
static const double RK45_C5[6] = {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0};
static const double RK45_C4[5] = {25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4101.0, -1.0 / 5.0};

/* One accepted step: stages 0..4 set u = u0 + sum_j (U[s+1][j] h) k_j and
 * enforce the constraints; stage 5 forms the fourth- and fifth-order
 * solutions (stage 1 skipped, its weights being 0), enforces both, and takes
 * the largest |y4 - y5| over every variable and interior point. Above tol the
 * step is repeated from u0 with h = 0.8 h (tol/err)^(1/4); else u = y5 and the
 * next step size is 0.8 h (tol/err)^(1/5) (the reference's BSSN_SAFETY_FAC
 * and exponents). With nbr (may be NULL) the padding is refreshed by
 * orc_padding_fill(general) before every stage and after the step. */
int orc_rk45_step(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr, int general,
                  double h, double tol, const orc_params *p, int max_attempts, int nthreads, double *h_used,
                  double *h_next, double *err_norm, int *attempts) {
    const int PW = 3;
    int64_t n = 0;
    for (int64_t b = 0; b < nblocks; b++) {
        const int64_t e = (int64_t)blks[b].offset + (int64_t)blks[b].sz[0] * blks[b].sz[1] * blks[b].sz[2];
        if (e > n) n = e;
    }
    double *buf = (double *)malloc(sizeof(double) * (size_t)n * 24 * 7);
    if (!buf) return -2;
    double *u0[24], *k[6][24];
    for (int v = 0; v < 24; v++) {
        u0[v] = buf + (size_t)n * v;
        for (int s = 0; s < 6; s++) {
            k[s][v] = buf + (size_t)n * (24 * (s + 1) + v);
            memset(k[s][v], 0, sizeof(double) * (size_t)n);
        }
    }
    if (nbr) orc_padding_fill(u, blks, nblocks, nbr, general);
    for (int v = 0; v < 24; v++) memcpy(u0[v], u[v], sizeof(double) * (size_t)n);
    int err = 0, tries = 0;
    double en = 0.0;
    for (;;) {
        tries++;
        en = 0.0;
        for (int s = 0; s < 6 && !err; s++) {
            if (s > 0 && nbr) orc_padding_fill(u, blks, nblocks, nbr, general);
#ifdef _OPENMP
            if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
            for (int64_t b = 0; b < nblocks; b++) {
                const int e = orc_bssnrhs_block(k[s], (const double *const *)u, &blks[b], p);
                err |= (e != 0);
            }
            for (int64_t b = 0; b < nblocks && !err; b++) {
                const orc_block *blk = &blks[b];
                const int nx = (int)blk->sz[0], ny = (int)blk->sz[1], nz = (int)blk->sz[2];
                for (int kk = PW; kk < nz - PW; kk++)
                    for (int j = PW; j < ny - PW; j++)
                        for (int i = PW; i < nx - PW; i++) {
                            const int64_t pp = (int64_t)blk->offset + i + (int64_t)nx * (j + (int64_t)ny * kk);
                            if (s < 5) {
                                for (int v = 0; v < 24; v++) {
                                    double a = u0[v][pp];
                                    for (int m = 0; m <= s; m++) a += (RK45_U[s + 1][m] * h) * k[m][v][pp];
                                    u[v][pp] = a;
                                }
                                if (orc_enforce_constraints_point(u, (uint64_t)pp, p->chi_floor)) err = 1;
                                continue;
                            }
                            double y4[24];
                            for (int v = 0; v < 24; v++) {
                                double a = u0[v][pp];
                                for (int m = 0; m < 5; m++)
                                    if (m != 1) a += (RK45_C4[m] * h) * k[m][v][pp];
                                u[v][pp] = a;
                            }
                            if (orc_enforce_constraints_point(u, (uint64_t)pp, p->chi_floor)) err = 1;
                            for (int v = 0; v < 24; v++) y4[v] = u[v][pp];
                            for (int v = 0; v < 24; v++) {
                                double a = u0[v][pp];
                                for (int m = 0; m < 6; m++)
                                    if (m != 1) a += (RK45_C5[m] * h) * k[m][v][pp];
                                u[v][pp] = a;
                            }
                            if (orc_enforce_constraints_point(u, (uint64_t)pp, p->chi_floor)) err = 1;
                            for (int v = 0; v < 24; v++) {
                                const double d = fabs(y4[v] - u[v][pp]);
                                if (d > en) en = d;
                            }
                        }
            }
        }
        if (err) break;
        if (en > tol && tries < max_attempts) {
            h = 0.8 * h * pow(fabs(tol / en), 0.25);
            for (int v = 0; v < 24; v++) memcpy(u[v], u0[v], sizeof(double) * (size_t)n);
            continue;
        }
        break;
    }
    (void)nthreads;
    if (!err && nbr) orc_padding_fill(u, blks, nblocks, nbr, general);
    free(buf);
    if (err) return -1;
    if (h_used) *h_used = h;
    if (h_next) *h_next = en > tol ? 0.8 * h * pow(fabs(tol / en), 0.25) : 0.8 * h * pow(fabs(tol / en), 0.20);
    if (err_norm) *err_norm = en;
    if (attempts) *attempts = tries;
    return en > tol ? 1 : 0;
}

/* RK4 (rkBSSN.cpp:764-860) with the padding of every stage state refreshed
 * from the neighbours before its right-hand side, as the reference's unzip
 * does between stages: u's padding before stage 0, each later stage's state
 * before its stage, and u's padding again after the step. Padding without a
 * neighbour is frozen at its value at the start of the step. */
int orc_rk4_step_halo(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr, double dt,
                      const orc_params *p, int enforce_per_stage, int nthreads) {
    return orc_rk4_step_fill(u, blks, nblocks, nbr, 0, dt, p, enforce_per_stage, nthreads);
}

int orc_rk4_step_fill(double *const *u, const orc_block *blks, int64_t nblocks, const int32_t *nbr, int general,
                      double dt, const orc_params *p, int enforce_per_stage, int nthreads) {
    const int PW = 3;
    int64_t n = 0;
    for (int64_t b = 0; b < nblocks; b++) {
        const int64_t e = (int64_t)blks[b].offset + (int64_t)blks[b].sz[0] * blks[b].sz[1] * blks[b].sz[2];
        if (e > n) n = e;
    }
    double *buf = (double *)malloc(sizeof(double) * (size_t)n * 24 * 3);
    if (!buf) return -2;
    double *u0[24], *us[24], *k[24];
    orc_padding_fill(u, blks, nblocks, nbr, general);
    for (int v = 0; v < 24; v++) {
        u0[v] = buf + (size_t)n * v;
        us[v] = buf + (size_t)n * (24 + v);
        k[v] = buf + (size_t)n * (48 + v);
        memcpy(u0[v], u[v], sizeof(double) * (size_t)n);
        memcpy(us[v], u[v], sizeof(double) * (size_t)n);
        memset(k[v], 0, sizeof(double) * (size_t)n);
    }
    int err = 0;
    for (int s = 0; s < 4 && !err; s++) {
        if (s > 0) orc_padding_fill(us, blks, nblocks, nbr, general);
#ifdef _OPENMP
        if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
        for (int64_t b = 0; b < nblocks; b++) {
            const orc_block *blk = &blks[b];
            int e = orc_bssnrhs_block(k, (const double *const *)us, blk, p);
            const int nx = (int)blk->sz[0], ny = (int)blk->sz[1], nz = (int)blk->sz[2];
            for (int kk = PW; kk < nz - PW && !e; kk++)
                for (int j = PW; j < ny - PW; j++)
                    for (int i = PW; i < nx - PW; i++) {
                        const int64_t pp = (int64_t)blk->offset + i + (int64_t)nx * (j + (int64_t)ny * kk);
                        for (int v = 0; v < 24; v++) u[v][pp] += (RK4_C[s] * dt * k[v][pp]);
                        if (s < 3) {
                            for (int v = 0; v < 24; v++) us[v][pp] = u0[v][pp] + (RK4_U[s + 1] * dt * k[v][pp]);
                            if (enforce_per_stage) {
                                const int e2 = orc_enforce_constraints_point(us, (uint64_t)pp, p->chi_floor);
                                if (e2) e = e2;
                            }
                        }
                    }
            err |= (e != 0);
        }
    }
    (void)nthreads;
    for (int64_t b = 0; b < nblocks && !err; b++) {
        const orc_block *blk = &blks[b];
        const int nx = (int)blk->sz[0], ny = (int)blk->sz[1], nz = (int)blk->sz[2];
        for (int kk = PW; kk < nz - PW; kk++)
            for (int j = PW; j < ny - PW; j++)
                for (int i = PW; i < nx - PW; i++) {
                    const int64_t pp = (int64_t)blk->offset + i + (int64_t)nx * (j + (int64_t)ny * kk);
                    if (orc_enforce_constraints_point(u, (uint64_t)pp, p->chi_floor)) err = 1;
                }
    }
    if (!err) orc_padding_fill(u, blks, nblocks, nbr, general);
    free(buf);
    return err ? -1 : 0;
}

/* ======================================================================== */
/* Synthetic inputs                                                           */
/* ======================================================================== */

/* splitmix64 (Steele, Lea, Flood 2014). */
uint64_t orc_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double u01(uint64_t *s) { return (double)(orc_splitmix64(s) >> 11) * 0x1.0p-53; }
static uint64_t stream_key(uint64_t seed, int64_t block, int var) {
    return seed + 0x632BE59BD9B4E019ull * (uint64_t)(block + 1) + 0x8CB92BA72F3D8DD7ull * (uint64_t)(var + 1);
}

/* Minkowski (grUtils.cpp:1621-1648) plus amp * sum of 4 seeded sine modes per
 * variable and block: A in U[-1,1], wave numbers k in {1..4}^3 per box length,
 * phase in U[0,2pi). gt is stored as 6 components so it stays symmetric. */
void orc_init_perturbed_minkowski(double *const *u, const orc_block *blks, int64_t nblocks,
                                  uint64_t seed, double amp) {
    static const double bg[24] = {1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                  1, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0};
    const double twopi = 6.283185307179586476925286766559;
    for (int64_t b = 0; b < nblocks; b++) {
        const orc_block *blk = &blks[b];
        const int nx = (int)blk->sz[0], ny = (int)blk->sz[1], nz = (int)blk->sz[2];
        const double hx = (blk->pmax[0] - blk->pmin[0]) / (nx - 1);
        const double hy = (blk->pmax[1] - blk->pmin[1]) / (ny - 1);
        const double hz = (blk->pmax[2] - blk->pmin[2]) / (nz - 1);
        const double Lx = blk->pmax[0] - blk->pmin[0];
        const double Ly = blk->pmax[1] - blk->pmin[1];
        const double Lz = blk->pmax[2] - blk->pmin[2];
        for (int v = 0; v < 24; v++) {
            uint64_t st = stream_key(seed, b, v);
            double A[4], kx[4], ky[4], kz[4], ph[4];
            for (int m = 0; m < 4; m++) {
                A[m] = 2.0 * u01(&st) - 1.0;
                kx[m] = (double)(1 + (orc_splitmix64(&st) & 3u)) * twopi / Lx;
                ky[m] = (double)(1 + (orc_splitmix64(&st) & 3u)) * twopi / Ly;
                kz[m] = (double)(1 + (orc_splitmix64(&st) & 3u)) * twopi / Lz;
                ph[m] = twopi * u01(&st);
            }
            double *dst = u[v] + blk->offset;
            for (int k = 0; k < nz; k++)
                for (int j = 0; j < ny; j++)
                    for (int i = 0; i < nx; i++) {
                        const double x = blk->pmin[0] + i * hx;
                        const double y = blk->pmin[1] + j * hy;
                        const double z = blk->pmin[2] + k * hz;
                        double s = 0.0;
                        for (int m = 0; m < 4; m++) s += A[m] * sin(kx[m] * x + ky[m] * y + kz[m] * z + ph[m]);
                        dst[IDX3(i, j, k)] = bg[v] + amp * s;
                    }
        }
    }
}

/* Brill-Lindquist two-puncture data, m1 = m2 = 0.5 at (+-3,0,0): alpha = psi^-2
 * and chi = psi^-4, each floored at chi_floor (grUtils.cpp:1081-1096), gt = delta;
 * K, Gt, beta, B, At get noise * U[-1,1] seeded per block and variable. */
void orc_init_brill_lindquist(double *const *u, const orc_block *blks, int64_t nblocks,
                              uint64_t seed, double noise, double chi_floor) {
    const double m1 = 0.5, m2 = 0.5, x1 = 3.0, x2 = -3.0;
    for (int64_t b = 0; b < nblocks; b++) {
        const orc_block *blk = &blks[b];
        const int nx = (int)blk->sz[0], ny = (int)blk->sz[1], nz = (int)blk->sz[2];
        const double hx = (blk->pmax[0] - blk->pmin[0]) / (nx - 1);
        const double hy = (blk->pmax[1] - blk->pmin[1]) / (ny - 1);
        const double hz = (blk->pmax[2] - blk->pmin[2]) / (nz - 1);
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    const double x = blk->pmin[0] + i * hx;
                    const double y = blk->pmin[1] + j * hy;
                    const double z = blk->pmin[2] + k * hz;
                    const double r1 = sqrt((x - x1) * (x - x1) + y * y + z * z);
                    const double r2 = sqrt((x - x2) * (x - x2) + y * y + z * z);
                    const double psi = 1.0 + m1 / (2.0 * r1) + m2 / (2.0 * r2);
                    double al = 1.0 / (psi * psi);
                    al = al > chi_floor ? al : chi_floor;
                    double ch = 1.0 / pow(psi, 4);
                    if (ch < chi_floor) ch = chi_floor;
                    const int64_t pp = blk->offset + IDX3(i, j, k);
                    u[ORC_ALPHA][pp] = al;
                    u[ORC_CHI][pp] = ch;
                    for (int c = 0; c < 6; c++) u[ORC_SYMGT0 + c][pp] = (c == 0 || c == 3 || c == 5) ? 1.0 : 0.0;
                }
        static const int noisy[16] = {ORC_K, ORC_GT0, ORC_GT0 + 1, ORC_GT0 + 2, ORC_BETA0, ORC_BETA0 + 1,
                                      ORC_BETA0 + 2, ORC_B0, ORC_B0 + 1, ORC_B0 + 2, ORC_SYMAT0,
                                      ORC_SYMAT0 + 1, ORC_SYMAT0 + 2, ORC_SYMAT0 + 3, ORC_SYMAT0 + 4,
                                      ORC_SYMAT0 + 5};
        const int64_t n = (int64_t)nx * ny * nz;
        for (int q = 0; q < 16; q++) {
            const int v = noisy[q];
            uint64_t st = stream_key(seed, b, v);
            for (int64_t pp = 0; pp < n; pp++) u[v][blk->offset + pp] = noise * (2.0 * u01(&st) - 1.0);
        }
    }
}
