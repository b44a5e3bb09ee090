// ORACLE — test infrastructure only.
//
// Harness that compiles the reference's own generated BSSN equations
// (BSSN_GR/src/bssneqs_eta_const_standard_gauge.cpp, included in place from the
// reference checkout, never copied) inside a point loop, so the fixtures pin the
// oracle's algebra to the reference's arithmetic. It supplies exactly the names
// the generated fragment reads, as BSSN_GR/src/rhs.cpp:84-156 does:
// per-variable arrays, grad_/grad2_ arrays (scripts/bssnrhs_evar_derivs.h),
// lambda[4] (unsigned, rhs.cpp:154), lambda_f[2] and a per-point eta.
//
// The reference's derivative stencils (BSSN_GR/src/derivs.cpp) are not built
// here: derivs.h includes dendrolib's TreeNode.h, absent from the checkout.
//
// Input layout per point: the oracle's ORC_POINT_NIN = 163 doubles
//   [0,24) fields (VAR order), [24,96) grad_d_v at 24+3v+d,
//   [96,162) grad2 of alpha, chi, beta0..2, gt0..5 (00 01 02 11 12 22), [162] eta.
#include <cmath>
#include <cstdint>
#include <vector>

using std::pow;
using std::sqrt;
using std::exp;

namespace {
enum { NIN = 163 };
}

extern "C" int ref_bssn_points(const double *in, int64_t npts, const double *lambda_d,
                               const double *lambda_f_in, double *out) {
    std::vector<double> soa((size_t)NIN * (size_t)npts);
    std::vector<double> res((size_t)24 * (size_t)npts);
    for (int64_t q = 0; q < npts; q++)
        for (int c = 0; c < NIN; c++) soa[(size_t)c * npts + q] = in[(size_t)q * NIN + c];

    const unsigned int lambda[4] = {(unsigned)lambda_d[0], (unsigned)lambda_d[1],
                                    (unsigned)lambda_d[2], (unsigned)lambda_d[3]};
    const double lambda_f[2] = {lambda_f_in[0], lambda_f_in[1]};

#define FIELD(name, v) const double *const name = &soa[(size_t)(v)*npts];
    FIELD(alpha, 0) FIELD(chi, 1) FIELD(K, 2) FIELD(Gt0, 3) FIELD(Gt1, 4) FIELD(Gt2, 5)
    FIELD(beta0, 6) FIELD(beta1, 7) FIELD(beta2, 8) FIELD(B0, 9) FIELD(B1, 10) FIELD(B2, 11)
    FIELD(gt0, 12) FIELD(gt1, 13) FIELD(gt2, 14) FIELD(gt3, 15) FIELD(gt4, 16) FIELD(gt5, 17)
    FIELD(At0, 18) FIELD(At1, 19) FIELD(At2, 20) FIELD(At3, 21) FIELD(At4, 22) FIELD(At5, 23)
#define GRAD(name, v)                                                     \
    FIELD(grad_0_##name, 24 + 3 * (v) + 0)                                \
    FIELD(grad_1_##name, 24 + 3 * (v) + 1) FIELD(grad_2_##name, 24 + 3 * (v) + 2)
    GRAD(alpha, 0) GRAD(chi, 1) GRAD(K, 2) GRAD(Gt0, 3) GRAD(Gt1, 4) GRAD(Gt2, 5)
    GRAD(beta0, 6) GRAD(beta1, 7) GRAD(beta2, 8) GRAD(B0, 9) GRAD(B1, 10) GRAD(B2, 11)
    GRAD(gt0, 12) GRAD(gt1, 13) GRAD(gt2, 14) GRAD(gt3, 15) GRAD(gt4, 16) GRAD(gt5, 17)
    GRAD(At0, 18) GRAD(At1, 19) GRAD(At2, 20) GRAD(At3, 21) GRAD(At4, 22) GRAD(At5, 23)
#define GRAD2(name, m)                                                    \
    FIELD(grad2_0_0_##name, 96 + 6 * (m) + 0) FIELD(grad2_0_1_##name, 96 + 6 * (m) + 1) \
    FIELD(grad2_0_2_##name, 96 + 6 * (m) + 2) FIELD(grad2_1_1_##name, 96 + 6 * (m) + 3) \
    FIELD(grad2_1_2_##name, 96 + 6 * (m) + 4) FIELD(grad2_2_2_##name, 96 + 6 * (m) + 5)
    GRAD2(alpha, 0) GRAD2(chi, 1) GRAD2(beta0, 2) GRAD2(beta1, 3) GRAD2(beta2, 4)
    GRAD2(gt0, 5) GRAD2(gt1, 6) GRAD2(gt2, 7) GRAD2(gt3, 8) GRAD2(gt4, 9) GRAD2(gt5, 10)
    const double *const eta_arr = &soa[(size_t)162 * npts];
#define OUT(name, v) double *const name = &res[(size_t)(v)*npts];
    OUT(a_rhs, 0) OUT(chi_rhs, 1) OUT(K_rhs, 2) OUT(Gt_rhs0, 3) OUT(Gt_rhs1, 4) OUT(Gt_rhs2, 5)
    OUT(b_rhs0, 6) OUT(b_rhs1, 7) OUT(b_rhs2, 8) OUT(B_rhs0, 9) OUT(B_rhs1, 10) OUT(B_rhs2, 11)
    OUT(gt_rhs00, 12) OUT(gt_rhs01, 13) OUT(gt_rhs02, 14) OUT(gt_rhs11, 15) OUT(gt_rhs12, 16)
    OUT(gt_rhs22, 17) OUT(At_rhs00, 18) OUT(At_rhs01, 19) OUT(At_rhs02, 20) OUT(At_rhs11, 21)
    OUT(At_rhs12, 22) OUT(At_rhs22, 23)

    for (int64_t pp = 0; pp < npts; pp++) {
        const double eta = eta_arr[pp];
#include "bssneqs_eta_const_standard_gauge.cpp"
    }

    for (int64_t q = 0; q < npts; q++)
        for (int v = 0; v < 24; v++) out[(size_t)q * 24 + v] = res[(size_t)v * npts + q];
    return 0;
}

// The same point loop over the reference's equations of its BSSN_ETA_FUNCTION
// build (bssneqs_eta_func_standard_gauge.cpp, CMakeLists.txt:95-97), whose B
// damping is eta_r0 |d chi| / (1 - chi^p0)^p1 (the per-point eta is unused).
extern "C" int ref_bssn_points_etafunc(const double *in, int64_t npts, const double *lambda_d,
                                       const double *lambda_f_in, double eta_r0, double eta_p0, double eta_p1,
                                       double *out) {
    // the names the eta-function variant reads (parameters.h:288-289)
    const double BSSN_ETA_R0 = eta_r0;
    const double BSSN_ETA_POWER[2] = {eta_p0, eta_p1};
    std::vector<double> soa((size_t)NIN * (size_t)npts);
    std::vector<double> res((size_t)24 * (size_t)npts);
    for (int64_t q = 0; q < npts; q++)
        for (int c = 0; c < NIN; c++) soa[(size_t)c * npts + q] = in[(size_t)q * NIN + c];

    const unsigned int lambda[4] = {(unsigned)lambda_d[0], (unsigned)lambda_d[1],
                                    (unsigned)lambda_d[2], (unsigned)lambda_d[3]};
    const double lambda_f[2] = {lambda_f_in[0], lambda_f_in[1]};

    FIELD(alpha, 0) FIELD(chi, 1) FIELD(K, 2) FIELD(Gt0, 3) FIELD(Gt1, 4) FIELD(Gt2, 5)
    FIELD(beta0, 6) FIELD(beta1, 7) FIELD(beta2, 8) FIELD(B0, 9) FIELD(B1, 10) FIELD(B2, 11)
    FIELD(gt0, 12) FIELD(gt1, 13) FIELD(gt2, 14) FIELD(gt3, 15) FIELD(gt4, 16) FIELD(gt5, 17)
    FIELD(At0, 18) FIELD(At1, 19) FIELD(At2, 20) FIELD(At3, 21) FIELD(At4, 22) FIELD(At5, 23)
    GRAD(alpha, 0) GRAD(chi, 1) GRAD(K, 2) GRAD(Gt0, 3) GRAD(Gt1, 4) GRAD(Gt2, 5)
    GRAD(beta0, 6) GRAD(beta1, 7) GRAD(beta2, 8) GRAD(B0, 9) GRAD(B1, 10) GRAD(B2, 11)
    GRAD(gt0, 12) GRAD(gt1, 13) GRAD(gt2, 14) GRAD(gt3, 15) GRAD(gt4, 16) GRAD(gt5, 17)
    GRAD(At0, 18) GRAD(At1, 19) GRAD(At2, 20) GRAD(At3, 21) GRAD(At4, 22) GRAD(At5, 23)
    GRAD2(alpha, 0) GRAD2(chi, 1) GRAD2(beta0, 2) GRAD2(beta1, 3) GRAD2(beta2, 4)
    GRAD2(gt0, 5) GRAD2(gt1, 6) GRAD2(gt2, 7) GRAD2(gt3, 8) GRAD2(gt4, 9) GRAD2(gt5, 10)
    const double *const eta_arr = &soa[(size_t)162 * npts];
    OUT(a_rhs, 0) OUT(chi_rhs, 1) OUT(K_rhs, 2) OUT(Gt_rhs0, 3) OUT(Gt_rhs1, 4) OUT(Gt_rhs2, 5)
    OUT(b_rhs0, 6) OUT(b_rhs1, 7) OUT(b_rhs2, 8) OUT(B_rhs0, 9) OUT(B_rhs1, 10) OUT(B_rhs2, 11)
    OUT(gt_rhs00, 12) OUT(gt_rhs01, 13) OUT(gt_rhs02, 14) OUT(gt_rhs11, 15) OUT(gt_rhs12, 16)
    OUT(gt_rhs22, 17) OUT(At_rhs00, 18) OUT(At_rhs01, 19) OUT(At_rhs02, 20) OUT(At_rhs11, 21)
    OUT(At_rhs12, 22) OUT(At_rhs22, 23)

    for (int64_t pp = 0; pp < npts; pp++) {
        const double eta = eta_arr[pp];
        (void)eta;
#include "bssneqs_eta_func_standard_gauge.cpp"
    }

    for (int64_t q = 0; q < npts; q++)
        for (int v = 0; v < 24; v++) out[(size_t)q * 24 + v] = res[(size_t)v * npts + q];
    return 0;
}

// All eight builds of the reference's equations (rhs.cpp:219-265), chosen by
// variant = eta_func (1) | Rochester gauge (2) | slow-start lapse (4), at the
// time t; the Rochester builds write no B_rhs (left at 0 here).
extern "C" int ref_bssn_points_variant(const double *in, int64_t npts, const double *lambda_d,
                                       const double *lambda_f_in, double eta_r0, double eta_p0, double eta_p1,
                                       const double *xi, double t, int variant, double *out) {
    // the names the variants read (parameters.h:288-289, 340)
    const double BSSN_ETA_R0 = eta_r0;
    const double BSSN_ETA_POWER[2] = {eta_p0, eta_p1};
    const unsigned int BSSN_XI[3] = {(unsigned)xi[0], (unsigned)xi[1], (unsigned)xi[2]};
    if (variant < 0 || variant > 7) return -1;
    std::vector<double> soa((size_t)NIN * (size_t)npts);
    std::vector<double> res((size_t)24 * (size_t)npts, 0.0);
    for (int64_t q = 0; q < npts; q++)
        for (int c = 0; c < NIN; c++) soa[(size_t)c * npts + q] = in[(size_t)q * NIN + c];

    const unsigned int lambda[4] = {(unsigned)lambda_d[0], (unsigned)lambda_d[1],
                                    (unsigned)lambda_d[2], (unsigned)lambda_d[3]};
    const double lambda_f[2] = {lambda_f_in[0], lambda_f_in[1]};

    FIELD(alpha, 0) FIELD(chi, 1) FIELD(K, 2) FIELD(Gt0, 3) FIELD(Gt1, 4) FIELD(Gt2, 5)
    FIELD(beta0, 6) FIELD(beta1, 7) FIELD(beta2, 8) FIELD(B0, 9) FIELD(B1, 10) FIELD(B2, 11)
    FIELD(gt0, 12) FIELD(gt1, 13) FIELD(gt2, 14) FIELD(gt3, 15) FIELD(gt4, 16) FIELD(gt5, 17)
    FIELD(At0, 18) FIELD(At1, 19) FIELD(At2, 20) FIELD(At3, 21) FIELD(At4, 22) FIELD(At5, 23)
    GRAD(alpha, 0) GRAD(chi, 1) GRAD(K, 2) GRAD(Gt0, 3) GRAD(Gt1, 4) GRAD(Gt2, 5)
    GRAD(beta0, 6) GRAD(beta1, 7) GRAD(beta2, 8) GRAD(B0, 9) GRAD(B1, 10) GRAD(B2, 11)
    GRAD(gt0, 12) GRAD(gt1, 13) GRAD(gt2, 14) GRAD(gt3, 15) GRAD(gt4, 16) GRAD(gt5, 17)
    GRAD(At0, 18) GRAD(At1, 19) GRAD(At2, 20) GRAD(At3, 21) GRAD(At4, 22) GRAD(At5, 23)
    GRAD2(alpha, 0) GRAD2(chi, 1) GRAD2(beta0, 2) GRAD2(beta1, 3) GRAD2(beta2, 4)
    GRAD2(gt0, 5) GRAD2(gt1, 6) GRAD2(gt2, 7) GRAD2(gt3, 8) GRAD2(gt4, 9) GRAD2(gt5, 10)
    const double *const eta_arr = &soa[(size_t)162 * npts];
    OUT(a_rhs, 0) OUT(chi_rhs, 1) OUT(K_rhs, 2) OUT(Gt_rhs0, 3) OUT(Gt_rhs1, 4) OUT(Gt_rhs2, 5)
    OUT(b_rhs0, 6) OUT(b_rhs1, 7) OUT(b_rhs2, 8) OUT(B_rhs0, 9) OUT(B_rhs1, 10) OUT(B_rhs2, 11)
    OUT(gt_rhs00, 12) OUT(gt_rhs01, 13) OUT(gt_rhs02, 14) OUT(gt_rhs11, 15) OUT(gt_rhs12, 16)
    OUT(gt_rhs22, 17) OUT(At_rhs00, 18) OUT(At_rhs01, 19) OUT(At_rhs02, 20) OUT(At_rhs11, 21)
    OUT(At_rhs12, 22) OUT(At_rhs22, 23)

    switch (variant) {
    case 0:
        for (int64_t pp = 0; pp < npts; pp++) {
            const double eta = eta_arr[pp];
            (void)eta;
#include "bssneqs_eta_const_standard_gauge.cpp"
        }
        break;
    case 1:
        for (int64_t pp = 0; pp < npts; pp++) {
            const double eta = eta_arr[pp];
            (void)eta;
#include "bssneqs_eta_func_standard_gauge.cpp"
        }
        break;
    case 2:
        for (int64_t pp = 0; pp < npts; pp++) {
            const double eta = eta_arr[pp];
            (void)eta;
#include "bssneqs_eta_const_rochester_gauge.cpp"
        }
        break;
    case 3:
        for (int64_t pp = 0; pp < npts; pp++) {
            const double eta = eta_arr[pp];
            (void)eta;
#include "bssneqs_eta_func_rochester_gauge.cpp"
        }
        break;
    case 4:
        for (int64_t pp = 0; pp < npts; pp++) {
            const double eta = eta_arr[pp];
            (void)eta;
#include "bssneqs_eta_const_standard_gauge_SSL.cpp"
        }
        break;
    case 5:
        for (int64_t pp = 0; pp < npts; pp++) {
            const double eta = eta_arr[pp];
            (void)eta;
#include "bssneqs_eta_func_standard_gauge_SSL.cpp"
        }
        break;
    case 6:
        for (int64_t pp = 0; pp < npts; pp++) {
            const double eta = eta_arr[pp];
            (void)eta;
#include "bssneqs_eta_const_rochester_gauge_SSL.cpp"
        }
        break;
    case 7:
        for (int64_t pp = 0; pp < npts; pp++) {
            const double eta = eta_arr[pp];
            (void)eta;
#include "bssneqs_eta_func_rochester_gauge_SSL.cpp"
        }
        break;
    }

    for (int64_t q = 0; q < npts; q++)
        for (int v = 0; v < 24; v++) out[(size_t)q * 24 + v] = res[(size_t)v * npts + q];
    return 0;
}

// Single point with the oracle's parameter struct, whose first six doubles are
// lambda[4] and lambda_f[2]; usable as the oracle's orc_point_fn hook.
extern "C" void ref_bssn_point_orc(const double *in, const void *params, double *out) {
    const double *p = static_cast<const double *>(params);
    ref_bssn_points(in, 1, p, p + 4, out);
}

// The reference's generated constraint fragment (BSSN_GR/src/physconeqs.cpp,
// included in place as physcon.cpp:78-86 does) over the same point layout:
// ham[pp], and with mom_out mom0..2[pp] (three per point), with psi4_out
// psi4_real, psi4_img[pp] (two per point). xyz (three per point) gives the
// coordinates the psi4 part reads; NULL sets them to (1, 2, 3).
extern "C" int ref_constraints_xyz(const double *in, const double *xyz, int64_t npts, double *ham_out,
                                   double *mom_out, double *psi4_out) {
    std::vector<double> soa((size_t)NIN * (size_t)npts);
    std::vector<double> res((size_t)6 * (size_t)npts);
    for (int64_t q = 0; q < npts; q++)
        for (int c = 0; c < NIN; c++) soa[(size_t)c * npts + q] = in[(size_t)q * NIN + c];
    FIELD(alpha, 0) FIELD(chi, 1) FIELD(K, 2) FIELD(Gt0, 3) FIELD(Gt1, 4) FIELD(Gt2, 5)
    FIELD(beta0, 6) FIELD(beta1, 7) FIELD(beta2, 8) FIELD(B0, 9) FIELD(B1, 10) FIELD(B2, 11)
    FIELD(gt0, 12) FIELD(gt1, 13) FIELD(gt2, 14) FIELD(gt3, 15) FIELD(gt4, 16) FIELD(gt5, 17)
    FIELD(At0, 18) FIELD(At1, 19) FIELD(At2, 20) FIELD(At3, 21) FIELD(At4, 22) FIELD(At5, 23)
    GRAD(alpha, 0) GRAD(chi, 1) GRAD(K, 2) GRAD(Gt0, 3) GRAD(Gt1, 4) GRAD(Gt2, 5)
    GRAD(beta0, 6) GRAD(beta1, 7) GRAD(beta2, 8) GRAD(B0, 9) GRAD(B1, 10) GRAD(B2, 11)
    GRAD(gt0, 12) GRAD(gt1, 13) GRAD(gt2, 14) GRAD(gt3, 15) GRAD(gt4, 16) GRAD(gt5, 17)
    GRAD(At0, 18) GRAD(At1, 19) GRAD(At2, 20) GRAD(At3, 21) GRAD(At4, 22) GRAD(At5, 23)
    GRAD2(alpha, 0) GRAD2(chi, 1) GRAD2(beta0, 2) GRAD2(beta1, 3) GRAD2(beta2, 4)
    GRAD2(gt0, 5) GRAD2(gt1, 6) GRAD2(gt2, 7) GRAD2(gt3, 8) GRAD2(gt4, 9) GRAD2(gt5, 10)
    OUT(ham, 0) OUT(mom0, 1) OUT(mom1, 2) OUT(mom2, 3) OUT(psi4_real, 4) OUT(psi4_img, 5)
    for (int64_t pp = 0; pp < npts; pp++) {
        const double x = xyz ? xyz[3 * pp] : 1.0;
        const double y = xyz ? xyz[3 * pp + 1] : 2.0;
        const double z = xyz ? xyz[3 * pp + 2] : 3.0;
#include "physconeqs.cpp"
    }
    for (int64_t q = 0; q < npts; q++) ham_out[q] = ham[q];
    if (mom_out)
        for (int64_t q = 0; q < npts; q++) {
            mom_out[3 * q] = mom0[q];
            mom_out[3 * q + 1] = mom1[q];
            mom_out[3 * q + 2] = mom2[q];
        }
    if (psi4_out)
        for (int64_t q = 0; q < npts; q++) {
            psi4_out[2 * q] = psi4_real[q];
            psi4_out[2 * q + 1] = psi4_img[q];
        }
    return 0;
}

extern "C" int ref_constraint_points(const double *in, int64_t npts, double *ham_out, double *mom_out) {
    return ref_constraints_xyz(in, nullptr, npts, ham_out, mom_out, nullptr);
}

extern "C" int ref_ham_points(const double *in, int64_t npts, double *ham_out) {
    return ref_constraint_points(in, npts, ham_out, nullptr);
}

// The reference's generated equations over one block, with the loop and the
// per-point coordinates and RIT eta of BSSN_GR/src/rhs.cpp:193-269, on the
// field and derivative arrays the oracle computed for the block (its
// restatement of derivs.cpp). Installed with orc_set_block_algebra, it makes
// orc_bssnrhs_block the reference's bssnrhs with only the stencils, the
// boundary pass and the KO sum restated: the CPU baseline times it.
#include "bssn_oracle.h"

extern "C" void ref_bssn_block_algebra(const double *const *u, const double *const *d1, const double *const *d2,
                                       const uint32_t *sz, const double *pmin, const double *hs,
                                       const orc_params *p, double *const *rhs) {
    const unsigned int nx = sz[0], ny = sz[1], nz = sz[2];
    const unsigned int PW = 3;
    const double hx = hs[0], hy = hs[1], hz = hs[2];
    const unsigned int lambda[4] = {(unsigned)p->lambda[0], (unsigned)p->lambda[1], (unsigned)p->lambda[2],
                                    (unsigned)p->lambda[3]};
    const double lambda_f[2] = {p->lambda_f[0], p->lambda_f[1]};
#define BF(name, v) const double *const name = u[v];
#define BG(name, v) \
    const double *const grad_0_##name = d1[3 * (v)]; \
    const double *const grad_1_##name = d1[3 * (v) + 1]; \
    const double *const grad_2_##name = d1[3 * (v) + 2];
#define BG2(name, m) \
    const double *const grad2_0_0_##name = d2[6 * (m)]; const double *const grad2_0_1_##name = d2[6 * (m) + 1]; \
    const double *const grad2_0_2_##name = d2[6 * (m) + 2]; const double *const grad2_1_1_##name = d2[6 * (m) + 3]; \
    const double *const grad2_1_2_##name = d2[6 * (m) + 4]; const double *const grad2_2_2_##name = d2[6 * (m) + 5];
#define BO(name, v) double *const name = rhs[v];
    BF(alpha, 0) BF(chi, 1) BF(K, 2) BF(Gt0, 3) BF(Gt1, 4) BF(Gt2, 5)
    BF(beta0, 6) BF(beta1, 7) BF(beta2, 8) BF(B0, 9) BF(B1, 10) BF(B2, 11)
    BF(gt0, 12) BF(gt1, 13) BF(gt2, 14) BF(gt3, 15) BF(gt4, 16) BF(gt5, 17)
    BF(At0, 18) BF(At1, 19) BF(At2, 20) BF(At3, 21) BF(At4, 22) BF(At5, 23)
    BG(alpha, 0) BG(chi, 1) BG(K, 2) BG(Gt0, 3) BG(Gt1, 4) BG(Gt2, 5)
    BG(beta0, 6) BG(beta1, 7) BG(beta2, 8) BG(B0, 9) BG(B1, 10) BG(B2, 11)
    BG(gt0, 12) BG(gt1, 13) BG(gt2, 14) BG(gt3, 15) BG(gt4, 16) BG(gt5, 17)
    BG(At0, 18) BG(At1, 19) BG(At2, 20) BG(At3, 21) BG(At4, 22) BG(At5, 23)
    BG2(alpha, 0) BG2(chi, 1) BG2(beta0, 2) BG2(beta1, 3) BG2(beta2, 4)
    BG2(gt0, 5) BG2(gt1, 6) BG2(gt2, 7) BG2(gt3, 8) BG2(gt4, 9) BG2(gt5, 10)
    BO(a_rhs, 0) BO(chi_rhs, 1) BO(K_rhs, 2) BO(Gt_rhs0, 3) BO(Gt_rhs1, 4) BO(Gt_rhs2, 5)
    BO(b_rhs0, 6) BO(b_rhs1, 7) BO(b_rhs2, 8) BO(B_rhs0, 9) BO(B_rhs1, 10) BO(B_rhs2, 11)
    BO(gt_rhs00, 12) BO(gt_rhs01, 13) BO(gt_rhs02, 14) BO(gt_rhs11, 15) BO(gt_rhs12, 16)
    BO(gt_rhs22, 17) BO(At_rhs00, 18) BO(At_rhs01, 19) BO(At_rhs02, 20) BO(At_rhs11, 21)
    BO(At_rhs12, 22) BO(At_rhs22, 23)
#undef BF
#undef BG
#undef BG2
#undef BO
    for (unsigned int k = PW; k < nz - PW; k++)
        for (unsigned int j = PW; j < ny - PW; j++)
            for (unsigned int i = PW; i < nx - PW; i++) {
                const double x = pmin[0] + i * hx;
                const double y = pmin[1] + j * hy;
                const double z = pmin[2] + k * hz;
                const unsigned int pp = i + nx * (j + ny * k);
                const double r_coord = sqrt(x * x + y * y + z * z);
                const double w = r_coord / p->eta_width;
                const double arg = -w * w * w * w;
                const double eta = (p->eta_central - p->eta_outer) * exp(arg) + p->eta_outer;
#include "bssneqs_eta_const_standard_gauge.cpp"
            }
}
