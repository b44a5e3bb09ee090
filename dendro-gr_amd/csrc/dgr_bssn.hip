// MI355X (gfx950) BSSN right-hand side over a batch of padded octree blocks.
//
// Work decomposition
//   * every block's interior is cut into tiles of at most 7x7x7 points
//     (a 13^3 block is one tile, a 55^3 block 343 tiles);
//   * tiles are grouped into chunks of whole blocks whose derivative
//     workspace (162 doubles per interior point) stays cache-sized;
//   * per chunk, k_derivs stages one variable of one tile (with its 3-point
//     halo) in LDS and writes the 72 first, 66 second/mixed derivatives and
//     the Kreiss-Oliger sum of that variable for the tile's points;
//   * k_rhs runs one thread per interior point: the generated BSSN equations
//     and the KO term, then either a plain store of the right-hand side or the
//     RK4 stage update; k_face does the same for the points on flagged faces
//     with the radiative boundary condition in place of the equations.
//
// Reference behaviour restated: BSSN_GR/src/rhs.cpp:80-504 (bssnrhs),
// :511-631 (bssn_bcs), BSSN_GR/src/derivs.cpp (deriv644_*, ko_deriv42_*),
// BSSN_GR/include/bssn_constraints.h:15-172, BSSN_GR/src/rkBSSN.cpp:764-860.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "dgr_bssn.h"

#define DGR_HD __host__ __device__
#define DGR_INLINE __forceinline__
// keeps a value's computation where it is written (see bssn_rhs_phased.h)
#define DGR_PIN(x) asm volatile("" : "+v"(x))
#define DGR_GEN_SQRT
template <typename T>
__host__ __device__ inline T dgr_sqrt(T x) {
    return sqrt(x);
}
#include "bssn_rhs_generated.h"
#include "stencils.h"
#include "bssn_rhs_phased.h"
#include "bssn_rhs_gauge_generated.h"
#include "bssn_rhs_phased_gauge.h"
#include "bssn_cons_phased.h"
#include "bssn_ham_generated.h"

namespace {

constexpr int NV = 24;
constexpr int PW = 3;
constexpr int TMAX = 7;              // max tile edge (interior points)
constexpr int RMAX = TMAX + 2 * PW;  // max staged region edge
constexpr int SLOT_D2 = 72;          // 24 vars x 3 first derivatives
constexpr int SLOT_KO = 72 + 66;     // 11 vars x 6 second derivatives
constexpr int NSLOT = 72 + 66 + 24;
constexpr int NTHREADS = 256;

struct VarPtrs {
    double *p[NV];
};

struct DBlock {  // device copy of a block with its grid spacing
    uint64_t offset;
    int64_t q0;  // first interior point of the block in interior-indexed buffers
    int n[3];
    uint32_t bflag;
    double pmin[3];
    double h[3];
};

struct Tile {
    int block;
    int q0;       // first workspace point of the tile within its chunk (split path)
    int o[3];     // absolute index of the tile's first point along each axis
    int e[3];     // tile extents
    int lanetab;  // fused path: offset of the tile shape's lane -> point table
};

struct Chunk {
    int tile_begin, tile_end;  // tiles [begin, end)
    int64_t npoints;           // interior points in the chunk
    int wg_begin;              // first entry in the k_rhs workgroup->tile table
    int nwg;
    int face_begin;            // first entry in the flagged-block tile list
    int nface;
};

// The 11 variables with second derivatives: alpha, chi, beta0..2, gt0..5.
__host__ __device__ constexpr int d2slot(int v) {
    return v == 0 ? 0 : v == 1 ? 1 : (v >= 6 && v <= 8) ? v - 4 : (v >= 12 && v <= 17) ? v - 7 : -1;
}

// Radiative BC constants (falloff, asymptotic value) per variable, rhs.cpp:275-327.
__device__ constexpr double BC_FALLOFF[NV] = {1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1,
                                              1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2};
__device__ constexpr double BC_ASYM[NV] = {1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                           1, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0};
// gauge variables take the gauge CAKO coefficient (rhs.cpp:356-372, 425-432)
__host__ __device__ constexpr bool is_gauge(int v) { return v == 0 || (v >= 6 && v <= 11); }

struct RhsParams {
    double lam[4];
    double lamf[2];
    double eta_central, eta_outer, eta_width;
    double ko_sigma;
    double chi_floor;
    int cako;
    double eps_gauge, eps_other;
    int eta_func;                    // eta from chi's gradient (BSSN_ETA_FUNCTION), else RIT eta(r)
    double eta_r0, eta_p0, eta_p1;
    int rochester;                   // the Rochester puncture gauge (USE_ROCHESTER_GAUGE)
    int ssl;                         // the slow-start lapse term (BSSN_ENABLE_SSL_IN_RHS)
    double xi[3];                    // BSSN_XI (the Rochester gauge's shift coefficients)
    double ssl_e;                    // exp(-t^2/800) at the time the call evaluates at
    const double *ssl_dev;           // if set, {t, exp(-t^2/800)} in device memory: read [1] instead
};

// The right-hand side of one of the reference's other builds (eta function,
// Rochester gauge, slow-start lapse): its own kernels run it.
inline bool gauge_variant(const RhsParams &prm) { return prm.eta_func || prm.rochester || prm.ssl; }

// The slow-start lapse term 0.6 sqrt(chi) (alpha - sqrt(chi)) exp(-t^2/800)
// (CodeGen/bssn.py:90-99), in the factor order of the generated
// bssneqs_*_SSL.cpp; e = exp(-t^2/800).
template <typename T>
__device__ __forceinline__ T ssl_term(T alpha, T chi, double e) {
    const T w = sqrt(chi);
    return T(0.6) * w * (alpha - w) * T(e);
}

// -------------------------------------------------------------------------
// k_derivs: one workgroup per (tile, variable)
// -------------------------------------------------------------------------
__global__ __launch_bounds__(NTHREADS) void k_derivs(VarPtrs u, const DBlock *__restrict__ blocks,
                                                    const Tile *__restrict__ tiles, int tile_base,
                                                    double *__restrict__ W, int64_t P) {
    __shared__ double s_u[RMAX * RMAX * RMAX];
    __shared__ double s_dx[TMAX * RMAX * RMAX];  // D_x u on tile-x x region-y x region-z
    __shared__ double s_dy[TMAX * TMAX * RMAX];  // D_y u on tile-x x tile-y x region-z

    const Tile T = tiles[tile_base + blockIdx.x];
    const int v = blockIdx.y;
    const DBlock B = blocks[T.block];
    const int nx = B.n[0], ny = B.n[1];
    const int ex = T.e[0], ey = T.e[1], ez = T.e[2];
    const int rx = ex + 2 * PW, ry = ey + 2 * PW, rz = ez + 2 * PW;
    const int tid = threadIdx.x;

    // stage the region [o-3, o+e+3) of variable v
    const double *__restrict__ src = u.p[v] + B.offset;
    const int gx0 = T.o[0] - PW, gy0 = T.o[1] - PW, gz0 = T.o[2] - PW;
    const int nreg = rx * ry * rz;
    for (int idx = tid; idx < nreg; idx += NTHREADS) {
        const int ii = idx % rx;
        const int jj = (idx / rx) % ry;
        const int kk = idx / (rx * ry);
        s_u[idx] = src[(int64_t)(gx0 + ii) + (int64_t)nx * ((gy0 + jj) + (int64_t)ny * (gz0 + kk))];
    }
    __syncthreads();

    const bool lox = B.bflag & 1u, hix = B.bflag & 2u;
    const bool loy = B.bflag & 4u, hiy = B.bflag & 8u;
    const bool loz = B.bflag & 16u, hiz = B.bflag & 32u;
    const dgr::Axis1<double> c1x(B.h[0]), c1y(B.h[1]), c1z(B.h[2]);
    const int m2 = d2slot(v);

    if (m2 >= 0) {
        // D_x over tile-x x all region rows (the mixed derivatives need D_x on the
        // y and z halo, like deriv644_x over all j,k: derivs.cpp:2628-2633)
        const int ndx = ex * ry * rz;
        for (int idx = tid; idx < ndx; idx += NTHREADS) {
            const int li = idx % ex;
            const int jj = (idx / ex) % ry;
            const int kk = idx / (ex * ry);
            const int a = T.o[0] + li;
            s_dx[idx] = dgr::d1(&s_u[(PW + li) + rx * (jj + ry * kk)], 1, a, nx, lox, hix, c1x);
        }
        // D_y over tile-x x tile-y x region-z (deriv644_y range: derivs.cpp:2729-2734)
        const int ndy = ex * ey * rz;
        for (int idx = tid; idx < ndy; idx += NTHREADS) {
            const int li = idx % ex;
            const int lj = (idx / ex) % ey;
            const int kk = idx / (ex * ey);
            const int a = T.o[1] + lj;
            s_dy[idx] = dgr::d1(&s_u[(PW + li) + rx * ((PW + lj) + ry * kk)], rx, a, ny, loy, hiy, c1y);
        }
        __syncthreads();
    }

    const int npts = ex * ey * ez;
    const int nz = B.n[2];
    for (int p = tid; p < npts; p += NTHREADS) {
        const int li = p % ex;
        const int lj = (p / ex) % ey;
        const int lk = p / (ex * ey);
        const int ax = T.o[0] + li, ay = T.o[1] + lj, az = T.o[2] + lk;
        const double *c = &s_u[(PW + li) + rx * ((PW + lj) + ry * (PW + lk))];
        const int64_t q = (int64_t)T.q0 + p;
        double gx, gy;
        if (m2 >= 0) {
            gx = s_dx[li + ex * ((PW + lj) + ry * (PW + lk))];
            gy = s_dy[li + ex * (lj + ey * (PW + lk))];
        } else {
            gx = dgr::d1(c, 1, ax, nx, lox, hix, c1x);
            gy = dgr::d1(c, rx, ay, ny, loy, hiy, c1y);
        }
        const double gz = dgr::d1(c, rx * ry, az, nz, loz, hiz, c1z);
        W[(int64_t)(3 * v + 0) * P + q] = gx;
        W[(int64_t)(3 * v + 1) * P + q] = gy;
        W[(int64_t)(3 * v + 2) * P + q] = gz;
        if (m2 >= 0) {
            const dgr::Axis2<double> c2x(B.h[0]), c2y(B.h[1]), c2z(B.h[2]);
            const double *cx = &s_dx[li + ex * ((PW + lj) + ry * (PW + lk))];
            const double *cy = &s_dy[li + ex * (lj + ey * (PW + lk))];
            double *w2 = W + (int64_t)(SLOT_D2 + 6 * m2) * P + q;
            w2[0 * P] = dgr::d2(c, 1, ax, nx, lox, hix, c2x);
            w2[1 * P] = dgr::d1(cx, ex, ay, ny, loy, hiy, c1y);        // D_y(D_x u)
            w2[2 * P] = dgr::d1(cx, ex * ry, az, nz, loz, hiz, c1z);   // D_z(D_x u)
            w2[3 * P] = dgr::d2(c, rx, ay, ny, loy, hiy, c2y);
            w2[4 * P] = dgr::d1(cy, ex * ey, az, nz, loz, hiz, c1z);   // D_z(D_y u)
            w2[5 * P] = dgr::d2(c, rx * ry, az, nz, loz, hiz, c2z);
        }
        const dgr::AxisKO<double> kx(B.h[0]), ky(B.h[1]), kz(B.h[2]);
        const double kox = dgr::ko(c, 1, ax, nx, lox, hix, kx);
        const double koy = dgr::ko(c, rx, ay, ny, loy, hiy, ky);
        const double koz = dgr::ko(c, rx * ry, az, nz, loz, hiz, kz);
        W[(int64_t)(SLOT_KO + v) * P + q] = kox + koy + koz;
    }
}

// -------------------------------------------------------------------------
// k_rhs: one thread per interior point
// -------------------------------------------------------------------------
struct PointIn {
    const VarPtrs *u;
    uint64_t idx;
    const double *W;
    int64_t P, q;
    template <int V>
    __device__ __forceinline__ double fld() const { return u->p[V][idx]; }
    template <int V, int D>
    __device__ __forceinline__ double d1() const { return W[(int64_t)(3 * V + D) * P + q]; }
    template <int M, int C>
    __device__ __forceinline__ double d2() const { return W[(int64_t)(SLOT_D2 + 6 * M + C) * P + q]; }
    double x[3];  // coordinates (read by the psi4 expressions only)
    template <int D>
    __device__ __forceinline__ double crd() const { return x[D]; }
};

enum { MODE_RHS = 0, MODE_RK = 1 };

struct RKArgs {
    VarPtrs u0;     // state at the start of the step (read; written at stage 3)
    VarPtrs acc;    // running u0 + sum c_s dt k_s
    VarPtrs snext;  // next stage state (written at stages 0..2)
    double cacc;    // RK4_C[s] * dt
    double cnext;   // RK4_U[s+1] * dt
    int stage;
    int enforce_stage;
    int *err;
};

// enforce_bssn_constraints on one point's values: bssn_constraints.h:15-172.
// Order in v14: alpha, chi, gt00..gt22, At00..At22. Returns error bits.
__device__ int enforce_point(double v14[14], double chi_floor) {
#pragma clang fp contract(off)
    const double one_third = 1.0 / 3.0;
    double g00 = v14[2], g01 = v14[3], g02 = v14[4], g11 = v14[5], g12 = v14[6], g22 = v14[7];
    double A00 = v14[8], A01 = v14[9], A02 = v14[10], A11 = v14[11], A12 = v14[12], A22 = v14[13];
    int err = 0;
    double det = g00 * (g11 * g22 - g12 * g12) - g01 * g01 * g22 + 2.0 * g01 * g02 * g12 - g02 * g02 * g11;
    if (det < 0.0) return 1;
    const double s = 1.0 / pow(det, one_third);
    g00 *= s; g01 *= s; g02 *= s; g11 *= s; g12 *= s; g22 *= s;
    det = g00 * (g11 * g22 - g12 * g12) - g01 * g01 * g22 + 2.0 * g01 * g02 * g12 - g02 * g02 * g11;
    if (fabs(det - 1.0) > 1.0e-6) err |= 2;
    const double idet = 1.0 / det;
    const double u00 = idet * (g11 * g22 - g12 * g12);
    const double u01 = idet * (-g01 * g22 + g02 * g12);
    const double u02 = idet * (g01 * g12 - g02 * g11);
    const double u11 = idet * (g00 * g22 - g02 * g02);
    const double u12 = idet * (-g00 * g12 + g01 * g02);
    const double u22 = idet * (g00 * g11 - g01 * g01);
    const double tr3 = one_third * (A00 * u00 + A11 * u11 + A22 * u22 + 2.0 * (A01 * u01 + A02 * u02 + A12 * u12));
    A00 -= tr3 * g00; A01 -= tr3 * g01; A02 -= tr3 * g02;
    A11 -= tr3 * g11; A12 -= tr3 * g12; A22 -= tr3 * g22;
    const double trA = A00 * u00 + A11 * u11 + A22 * u22 + 2.0 * (A01 * u01 + A02 * u02 + A12 * u12);
    if (fabs(trA) > 1.0e-6) err |= 4;
    v14[2] = g00; v14[3] = g01; v14[4] = g02; v14[5] = g11; v14[6] = g12; v14[7] = g22;
    v14[8] = A00; v14[9] = A01; v14[10] = A02; v14[11] = A11; v14[12] = A12; v14[13] = A22;
    if (v14[1] < chi_floor) v14[1] = chi_floor;
    // std::max(alpha, CHI_FLOOR) (bssn_constraints.h:171): a NaN stays a NaN
    // (fmax would replace it with the floor and hide it from the NaN check)
    if (v14[0] < chi_floor) v14[0] = chi_floor;
    return err;
}

__device__ __forceinline__ int slot14(int v) { return v <= 1 ? v : (v >= 12 ? v - 10 : -1); }

// Where a finished right-hand side value goes: a plain store (MODE_RHS) or the
// RK4 stage update of rkBSSN.cpp:803-812 / 847-858 (MODE_RK). The 14 variables
// enforce_bssn_constraints touches are held until finish() when the stage
// enforces constraints.
template <int MODE>
struct Epilogue {
    const VarPtrs *rhs;
    const RKArgs *rk;
    uint64_t idx;
    double v14[14];

    template <int V>
    __device__ __forceinline__ void store(double val) {
#pragma clang fp contract(off)
        if constexpr (MODE == MODE_RHS) {
            rhs->p[V][idx] = val;
        } else {
            const int s = rk->stage;
            if (s < 3) {
                const double u0 = rk->u0.p[V][idx];
                const double nxt = u0 + rk->cnext * val;
                rk->acc.p[V][idx] = (s == 0 ? u0 : rk->acc.p[V][idx]) + rk->cacc * val;
                if (slot14(V) >= 0 && rk->enforce_stage) v14[slot14(V)] = nxt;
                else rk->snext.p[V][idx] = nxt;
            } else {
                const double fin = rk->acc.p[V][idx] + rk->cacc * val;
                if (slot14(V) >= 0) v14[slot14(V)] = fin;
                else rk->u0.p[V][idx] = fin;
            }
        }
    }

    __device__ __forceinline__ void finish(double chi_floor) {
        if constexpr (MODE == MODE_RK) {
            const bool final_stage = rk->stage == 3;
            if (final_stage || rk->enforce_stage) {
                const int e = enforce_point(v14, chi_floor);
                if (e) atomicOr(rk->err, e);
                const VarPtrs &dst = final_stage ? rk->u0 : rk->snext;
                dst.p[0][idx] = v14[0];
                dst.p[1][idx] = v14[1];
#pragma unroll
                for (int c = 0; c < 12; c++) dst.p[12 + c][idx] = v14[2 + c];
            }
        }
    }
};

// Receives the generated equations' outputs and adds the Kreiss-Oliger term
// (rhs.cpp:434-489; CAKO variant :356-433).
template <int MODE>
struct PointOut : Epilogue<MODE> {
    const double *W;
    int64_t P, q;
    double sig_gauge, sig_other;

    template <int V>
    __device__ __forceinline__ void put(double val) {
#pragma clang fp contract(off)
        const double ko = W[(int64_t)(SLOT_KO + V) * P + q];
        this->template store<V>(val + (is_gauge(V) ? sig_gauge : sig_other) * ko);
    }
};

// x = pmin + i*h and the RIT eta of rhs.cpp:202-215, rounded as the reference does.
__device__ __forceinline__ double coord(double pmin, int i, double h) {
#pragma clang fp contract(off)
    return pmin + i * h;
}

// The eta of the reference's BSSN_ETA_FUNCTION build (CodeGen/bssn.py:71-75,
// generated as bssneqs_eta_func_standard_gauge.cpp, which takes the place of
// the RIT eta(r) below): r0 sqrt(gt^ij d_i chi d_j chi) (1 - chi^p0)^-p1, from
// the point's gt (00 01 02 11 12 22), chi and chi's first derivatives.
template <typename T>
__device__ __forceinline__ T eta_function(const T *g, const T *dc, T chi, double r0, double p0, double p1) {
    const T c00 = g[3] * g[5] - g[4] * g[4], c01 = g[2] * g[4] - g[1] * g[5], c02 = g[1] * g[4] - g[2] * g[3];
    const T c11 = g[0] * g[5] - g[2] * g[2], c12 = g[1] * g[2] - g[0] * g[4], c22 = g[0] * g[3] - g[1] * g[1];
    const T det = g[0] * c00 + g[1] * c01 + g[2] * c02;
    const T q = (c00 * dc[0] * dc[0] + c11 * dc[1] * dc[1] + c22 * dc[2] * dc[2] +
                 T(2) * (c01 * dc[0] * dc[1] + c02 * dc[0] * dc[2] + c12 * dc[1] * dc[2])) /
                det;
    return T(r0) * sqrt(q) * pow(T(1) - pow(chi, T(p0)), -T(p1));
}

__device__ __forceinline__ double rit_eta(double x, double y, double z, const RhsParams &prm) {
#pragma clang fp contract(off)
    const double r_coord = sqrt(x * x + y * y + z * z);
    const double w = r_coord / prm.eta_width;
    const double arg = -w * w * w * w;
    return (prm.eta_central - prm.eta_outer) * exp(arg) + prm.eta_outer;
}

__device__ __forceinline__ double inv_radius(double x, double y, double z) {
#pragma clang fp contract(off)
    return 1.0 / sqrt(x * x + y * y + z * z);
}

struct PointGeom {
    int tile, i, j, k;
    uint64_t idx;
    bool face;
};

__device__ __forceinline__ PointGeom locate(const Tile &T, const DBlock &B, int p) {
    PointGeom g;
    const int li = p % T.e[0];
    const int lj = (p / T.e[0]) % T.e[1];
    const int lk = p / (T.e[0] * T.e[1]);
    g.i = T.o[0] + li;
    g.j = T.o[1] + lj;
    g.k = T.o[2] + lk;
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    g.idx = B.offset + (uint64_t)g.i + (uint64_t)nx * ((uint64_t)g.j + (uint64_t)ny * g.k);
    const uint32_t bf = B.bflag;
    g.face = ((bf & 1u) && g.i == PW) || ((bf & 2u) && g.i == nx - PW - 1) || ((bf & 4u) && g.j == PW) ||
             ((bf & 8u) && g.j == ny - PW - 1) || ((bf & 16u) && g.k == PW) || ((bf & 32u) && g.k == nz - PW - 1);
    return g;
}

__device__ __forceinline__ void ko_sigmas(const RhsParams &prm, const VarPtrs &u, uint64_t idx, double &sg,
                                          double &so) {
    if (prm.cako) {  // sigma = sqrt(chi) * eps (rhs.cpp:356-360)
        const double s = sqrt(u.p[1][idx]);
        sg = s * prm.eps_gauge;
        so = s * prm.eps_other;
    } else {
        sg = so = prm.ko_sigma;
    }
}

// One thread per interior point of a chunk. Points on a flagged face are left to
// k_face: bssn_bcs overwrites all 24 right-hand sides there (rhs.cpp:272-330).
template <int MODE>
__global__ __launch_bounds__(NTHREADS) void k_rhs(VarPtrs u, VarPtrs rhs, RKArgs rk, RhsParams prm,
                                                 const DBlock *__restrict__ blocks,
                                                 const Tile *__restrict__ tiles, const int *__restrict__ wg_tile,
                                                 int tile_end, const double *__restrict__ W, int64_t P) {
    const int64_t q = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
    if (q >= P) return;
    int t = wg_tile[blockIdx.x];
    while (t + 1 < tile_end && q >= tiles[t + 1].q0) t++;
    const Tile T = tiles[t];
    const DBlock B = blocks[T.block];
    const PointGeom g = locate(T, B, (int)(q - T.q0));
    if (g.face) return;

    // coordinates and eta (rhs.cpp:202-215)
    const double x = coord(B.pmin[0], g.i, B.h[0]);
    const double y = coord(B.pmin[1], g.j, B.h[1]);
    const double z = coord(B.pmin[2], g.k, B.h[2]);
    PointIn in{&u, g.idx, W, P, q};
    double eta;
    if (prm.eta_func) {
        const double gt[6] = {in.fld<12>(), in.fld<13>(), in.fld<14>(), in.fld<15>(), in.fld<16>(), in.fld<17>()};
        const double dc[3] = {in.d1<1, 0>(), in.d1<1, 1>(), in.d1<1, 2>()};
        eta = eta_function<double>(gt, dc, in.fld<1>(), prm.eta_r0, prm.eta_p0, prm.eta_p1);
    } else {
        eta = rit_eta(x, y, z, prm);
    }
    PointOut<MODE> out;
    out.rhs = &rhs;
    out.rk = &rk;
    out.idx = g.idx;
    out.W = W;
    out.P = P;
    out.q = q;
    ko_sigmas(prm, u, g.idx, out.sig_gauge, out.sig_other);
    if (prm.rochester || prm.ssl) {
        const double sslt = prm.ssl ? ssl_term<double>(in.fld<0>(), in.fld<1>(), prm.ssl_e) : 0.0;
        dgr::bssn_rhs_point_gauge<double>(in, eta, prm.lam, prm.lamf, prm.xi, prm.rochester, sslt, out);
    } else {
        dgr::bssn_rhs_point<double>(in, eta, prm.lam, prm.lamf, out);
    }
    out.finish(prm.chi_floor);
}

// Radiative boundary condition at the face points of flagged blocks
// (bssn_bcs, rhs.cpp:511-631), then the KO term (rhs.cpp:434-489). One
// workgroup per tile of a flagged block.
template <int MODE, int V>
__device__ __forceinline__ void face_var(Epilogue<MODE> &ep, const VarPtrs &u, const double *W, int64_t P,
                                         int64_t q, double x, double y, double z, double inv_r, double sg,
                                         double so) {
#pragma clang fp contract(off)
    const double f = u.p[V][ep.idx];
    const double dxf = W[(int64_t)(3 * V + 0) * P + q];
    const double dyf = W[(int64_t)(3 * V + 1) * P + q];
    const double dzf = W[(int64_t)(3 * V + 2) * P + q];
    double val = -inv_r * (x * dxf + y * dyf + z * dzf + BC_FALLOFF[V] * (f - BC_ASYM[V]));
    val += (is_gauge(V) ? sg : so) * W[(int64_t)(SLOT_KO + V) * P + q];
    ep.template store<V>(val);
    if constexpr (V + 1 < NV) face_var<MODE, V + 1>(ep, u, W, P, q, x, y, z, inv_r, sg, so);
}

template <int MODE>
__global__ __launch_bounds__(NTHREADS) void k_face(VarPtrs u, VarPtrs rhs, RKArgs rk, RhsParams prm,
                                                  const DBlock *__restrict__ blocks,
                                                  const Tile *__restrict__ tiles,
                                                  const int *__restrict__ face_tiles,
                                                  const double *__restrict__ W, int64_t P) {
    const Tile T = tiles[face_tiles[blockIdx.x]];
    const DBlock B = blocks[T.block];
    const int npts = T.e[0] * T.e[1] * T.e[2];
    for (int p = threadIdx.x; p < npts; p += NTHREADS) {
        const PointGeom g = locate(T, B, p);
        if (!g.face) continue;
        const int64_t q = (int64_t)T.q0 + p;
        const double x = coord(B.pmin[0], g.i, B.h[0]);
        const double y = coord(B.pmin[1], g.j, B.h[1]);
        const double z = coord(B.pmin[2], g.k, B.h[2]);
        const double inv_r = inv_radius(x, y, z);
        Epilogue<MODE> ep;
        ep.rhs = &rhs;
        ep.rk = &rk;
        ep.idx = g.idx;
        double sg, so;
        ko_sigmas(prm, u, g.idx, sg, so);
        face_var<MODE, 0>(ep, u, W, P, q, x, y, z, inv_r, sg, so);
        ep.finish(prm.chi_floor);
    }
}

// -------------------------------------------------------------------------
// k_fused: derivatives, equations, KO, BC and the stage update in one kernel.
// One workgroup per tile of at most 7 x 7 x 5 points, one point per thread
// (threads beyond the tile's points shadow its last point and store nothing).
// The 24 variables stream through LDS in the phase order of
// bssn_rhs_phased.h: while phase K works on variable V (staged in LDS buffer
// K&1), the next variable's region is already in flight into registers.
// Every thread's own field values sit in LDS for the equations, its KO term
// (on flagged faces: boundary value plus KO term) is left in LDS, and the
// right-hand side is combined with both in one batched epilogue.
//
// Instruction economy matters more than usual here: the equations need about
// 430 registers per lane, so a SIMD runs a single wave and every instruction
// costs a full issue slot. LDS strides are compile-time (region 13 x 13 x
// (ez+6)), stencil constants come precomputed per block, and face closures
// sit behind a branch that is uniform over the tile.
// -------------------------------------------------------------------------
constexpr int FRX = RMAX, FRXY = RMAX * RMAX;  // LDS strides of the staged region (13, 169)
// D_x u (tile x, region y, region z) and D_y u (tile x, tile y, region z) use
// the same strides, so one lane -> point assignment keeps every access pattern
// of the own-point stencils free of LDS bank conflicts (see lane_table)
constexpr int DXS_Y = FRX, DXS_Z = FRXY;
constexpr int DYS_Y = FRX, DYS_Z = FRXY;
constexpr int LANE_SHADOW = 0x8000;  // lane table: the lane repeats a point and stores nothing

// Tile shape of a k_fused launch: FT threads (one point each) and tiles of at
// most 7 x 7 x FZ points. The regular shape is 256 threads / z <= 5 (one
// workgroup per CU: it takes 143 KB of LDS); the thin shape, 128 threads /
// z <= 2, takes the 2-point z remainder of a block (a 13^3 block's interior
// is 7x7x5 + 7x7x2) in half the threads and ~80 KB, so two share a CU.
// PD variables are fetched ahead of the one being differentiated: with a
// single wave per SIMD, the loads in flight are what keeps HBM busy.
template <int FT_, int FZ_, int PD_>
struct FCfg {
    static constexpr int FT = FT_;
    static constexpr int FZ = FZ_;
    static constexpr int PD = PD_;                          // variables in flight ahead of the current one
    static constexpr int NREG = FRXY * (FZ + 2 * PW);       // staged region (13 x 13 x (FZ+6))
    static constexpr int PREF = (NREG + FT - 1) / FT;       // region elements per thread
    // D_x points: D_y(D_x u) at the tile's points reads D_x on its z planes
    // (all 13 region rows), D_z(D_x u) on its y rows (all FZ+6 region planes)
    static constexpr int NDX = TMAX * (RMAX * FZ + TMAX * 2 * PW);
    static constexpr int NDY = TMAX * TMAX * (FZ + 2 * PW);  // D_y points
    static constexpr int DXIT = (NDX + FT - 1) / FT;         // D_x evaluations per thread
    static constexpr int DYIT = (NDY + FT - 1) / FT;
    static constexpr int NSDX = TMAX + FRX * (RMAX - 1) + FRXY * (FZ + 2 * PW - 1);  // D_x array
    static constexpr int NSDY = TMAX + FRX * (TMAX - 1) + FRXY * (FZ + 2 * PW - 1);  // D_y array
    // LDS (FP64 kernel): two region buffers, D_x / D_y, KO terms
    static constexpr size_t LDS = sizeof(double) * (2 * NREG + NSDX + NSDY + NV * FT);
};
using CfgWide = FCfg<256, 5, 3>;
using CfgThin = FCfg<128, 2, 2>;
// FP32 kernel on interior blocks: a whole 7 x 7 x 7 column per workgroup
// (384 threads, ~98 KB of LDS, one workgroup per CU). Its registers allow two
// waves per SIMD, so six waves fit where the FP64 kernel fits four: a 13^3
// block's interior is one tile, and the staging and D_x / D_y work of the
// z halo is shared by 7 planes instead of 5 + 2.
using CfgFull = FCfg<384, 7, 1>;
// The FP32 kernels fit two waves per SIMD in 256 registers only with one
// variable in flight.
template <typename T>
struct CfgOf {
    using Wide = CfgWide;
    using Thin = CfgThin;
};
template <>
struct CfgOf<float> {
    using Wide = FCfg<256, 5, 1>;
    using Thin = FCfg<128, 2, 1>;
};
static_assert(7 * 7 * CfgWide::FZ <= CfgWide::FT && 7 * 7 * CfgThin::FZ <= CfgThin::FT, "one point per thread");
static_assert(2 * CfgThin::LDS <= 160 * 1024, "two thin FP64 workgroups per CU");
// FP32 regular tiles: one FP64 and one FP32 region buffer, FP64-sized D_x /
// D_y arrays (alpha and chi are differentiated in FP64), FP32 KO terms
static_assert(2 * (sizeof(double) * (CfgWide::NREG + (CfgWide::NREG + 1) / 2 + CfgWide::NSDX + CfgWide::NSDY) +
                   sizeof(float) * NV * CfgWide::FT) <= 160 * 1024,
              "two FP32 workgroups per CU");
static_assert(7 * 7 * CfgFull::FZ <= CfgFull::FT, "one point per thread");
// three FP64-sized region buffers and two D_x / D_y buffers (one barrier per phase)
static_assert(sizeof(double) * (3 * CfgFull::NREG + 2 * (CfgFull::NSDX + CfgFull::NSDY)) +
                      sizeof(float) * NV * CfgFull::FT <= 160 * 1024,
              "one FP32 7x7x7 workgroup per CU");
static_assert(sizeof(double) * (3 * CfgWide::NREG + 2 * (CfgWide::NSDX + CfgWide::NSDY) + NV * CfgWide::FT) <=
                  160 * 1024,
              "one FP64 regular workgroup per CU");
// tile groups of the fused path, one launch each (FULL: FP32 only)
enum { FG_WIDE = 0, FG_THIN = 1, FG_FACE = 2, FG_FULL = 3, NFGROUP = 4 };
// Tilings of a block without flagged faces, chosen per block size and
// precision (dgr_plan_autotune measures them). x and y are always cut into
// near-equal parts of at most 7 points; along z:
//   SCH_WT   parts of 5, a remainder of 1-2 points in thin tiles (128 threads)
//   SCH_W    near-equal parts of at most 5, all regular tiles
//   SCH_FULL near-equal parts of at most 7 (384 threads; FP32 only)
enum { SCH_WT = 0, SCH_W = 1, SCH_FULL = 2, NSCHEME = 3 };

// Where a stage reads its state and writes its results. Stage buffers hold
// interior points only, indexed by DBlock::q0 + (i-3) + mx*((j-3) + my*(k-3));
// padding always comes from the caller's state `u`.
// Buffers of one fused launch. An RK4 step alternates two padded states: the
// caller's u and the plan's work state W, whose padding holds a copy of u's
// (k_pad_copy, once per dgr_rk4_step / dgr_rk4_steps call). Stage 0 reads u
// and writes the next stage state into W's interior; stage 1 reads W and
// writes u's interior; stage 2 reads u and writes W; stage 3 reads W and
// writes the result into u. A stage therefore reads its whole region from
// one array, and u's padding is never written. u0 (needed by stages 1 and 2
// once u's interior is reused) and the accumulator are interior-only.
struct StageIO {
    VarPtrs src;   // padded stage input (u or W; u in RHS mode)
    VarPtrs dst;   // padded next-stage state, interior points (W or u)
    VarPtrs u0;    // interior-only copy of the step's initial state
    VarPtrs acc;   // interior-only accumulator
    VarPtrs rhs;   // RHS mode: padded output
    double cacc, cnext;
    int stage;     // -1 in RHS mode
    int enforce_stage;
    int *err;
    // 27 per block: the offset of the same-level neighbour in each direction
    // (-1 for none), read only by the HALO instantiation of k_fused, which
    // stages region points inside a linked neighbour from that neighbour's
    // interior instead of the block's own padding
    const int64_t *nbo;
};

// Phase tables of the fused kernels: the variables in the order they stream
// through LDS (one phase each), and whether a phase also evaluates the KO term
// and, on flagged faces, the boundary value (the right-hand side does; the
// constraint diagnostic reads derivatives only).
struct RhsPhases {
    static constexpr int N = NV;
    static constexpr bool KO = true;
    static constexpr bool GAUGE = false;  // the standard gauge with the RIT eta(r)
    static constexpr int var(int k) { return dgr::kBssnPhaseVar[k]; }
};
// The right-hand side of the reference's other builds, chosen at run time:
// BSSN_ETA_FUNCTION (eta from chi's gradient, formed in chi's phase), the
// Rochester gauge, the slow-start lapse (its term formed in alpha's phase).
// Its own kernels, so the default ones carry none of its registers.
struct RhsPhasesGauge : RhsPhases {
    static constexpr bool GAUGE = true;
};
struct ConsPhases {
    static constexpr int N = dgr::kConsPhases;
    static constexpr bool KO = false;
    static constexpr bool GAUGE = false;
    static constexpr int var(int k) { return dgr::kConsPhaseVar[k]; }
};

// FACE: the tile's block has flagged faces (closures and boundary values are
// needed); interior tiles run a body without any of that code.
// T: the arithmetic type of derivatives, equations and KO (double, or float
// for the mixed-precision path, which stages each variable in LDS with a
// per-tile offset subtracted: the value at the region's first point).
// P: the phase table. HALO: region offsets index the whole variable (the
// in-kernel padding refresh reads neighbours); otherwise they are byte
// offsets within the block, so a staged load is a uniform base plus a 32-bit
// offset (one address register, no 64-bit address arithmetic per load).
template <bool FACE, class C, typename T, class P = RhsPhases, bool HALO = false>
struct FusedCtx {
    static constexpr int FT = C::FT, NREG = C::NREG, PREF = C::PREF, DXIT = C::DXIT, DYIT = C::DYIT, PD = C::PD;
    const StageIO &io;
    const dgr::AxisConsts *K;  // x, y, z constants of this block
    double *s_tile, *s_dx, *s_dy;  // sized for FP64 staging; viewed as the staging type
    T *s_ko;
    // tile (uniform)
    int ox, oy, oz, iex, iey, iez;  // tile origin, ie = n - 3 per axis
    bool lox, hix, loy, hiy, loz, hiz;
    uint64_t boff;
    uint32_t corner;                // padded offset (within the block) of the region's first point
    uint64_t corner_src;            // HALO: where that point's value is read (a neighbour's interior)
    double pmin[3], h[3];
    // this thread
    int tid, ax, ay, az, cofs, dxo, dyo;
    bool face;
    T sig_gauge, sig_other;
    // region elements this thread fetches: offset in the padded block
    uint32_t goff[PREF];
    double pre[PD][PREF];  // fetched region of the variables of phases K+1 .. K+PD
    double poff[PD];       // the staging offset of each fetched variable (mixed precision)
    T f[NV];               // this point's field values
    // the eta the equations read (by reference, after the last phase): the
    // RIT eta(r), or with P::GAUGE and eta_fn the eta of chi's gradient
    // (BSSN_ETA_FUNCTION) set in chi's phase
    T eta_v;
    double eta_r0, eta_p0, eta_p1;
    bool eta_fn, ssl_on;
    T ssl_v;       // P::GAUGE: the slow-start lapse term (0 without it), set in alpha's phase
    double ssl_e;  // exp(-t^2/800)
    // LDS source offsets of this thread's D_x / D_y evaluations
    int dxw[DXIT], dyw[DYIT];  // where this thread's D_x / D_y evaluations go (see dtable)
    T d1v[3], d2v[6];
    // FP64: the interior stencil constants of the block (1/(60h), 1/(180h^2),
    // KO prefactor per axis) held in vector registers. Left to the compiler
    // they are scalar loads re-issued near every use, and since scalar and LDS
    // loads share one wait counter, each re-issue drains the LDS queue (+1.2%
    // measured). The FP32 kernel, short of registers at two waves per SIMD,
    // is faster loading them (-1.4% measured with them in registers).
    static constexpr bool VK = sizeof(T) == sizeof(double);
    double kc[VK ? 9 : 1];
    template <int A, typename S>
    __device__ __forceinline__ S c_d1() const {
        if constexpr (VK) return kc[A];
        else return S(K[A].d1_60);
    }
    template <int A, typename S>
    __device__ __forceinline__ S c_d2() const {
        if constexpr (VK) return kc[3 + A];
        else return S(K[A].d2_180);
    }
    template <int A, typename S>
    __device__ __forceinline__ S c_ko() const {
        if constexpr (VK) return kc[6 + A];
        else return S(K[A].ko_pre);
    }
    __device__ __forceinline__ void load_consts() {
        if constexpr (VK) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                kc[a] = K[a].d1_60;
                kc[3 + a] = K[a].d2_180;
                kc[6 + a] = K[a].ko_pre;
            }
#pragma unroll
            for (int i = 0; i < 9; i++) DGR_PIN(kc[i]);
        }
    }

    __device__ FusedCtx(const StageIO &io_) : io(io_) {}

    // region of the variable of phase K into prefetch slot K % PD
    template <int K>
    __device__ __forceinline__ void fetch() {
        constexpr int v = P::var(K);
        double *dst = pre[K % PD];
        const double *__restrict__ sb = io.src.p[v] + boff;
        if constexpr (sizeof(T) < sizeof(double)) poff[K % PD] = HALO ? sb[corner_src] : sb[corner];
        // raw buffer loads: the base in a resource (scalar registers), the
        // 32-bit byte offset as the only per-lane address operand (the base is
        // the block's, or with HALO the variable's: the host runs the HALO
        // kernels only for variables below 2^29 points, whose byte offsets fit)
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(sb), 0, HALO ? -1 : 0x7fffffff, 0x00020000);
#pragma unroll
        for (int c = 0; c < PREF; c++)
            dst[c] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, goff[c], 0, 0));
    }

    // The type variable V is staged and differentiated in. The mixed path
    // keeps alpha and chi in FP64: near the punctures they vary by O(1)
    // within a tile, and FP32 staging errors, amplified by the second
    // derivatives and the KO stencil, would swamp their small right-hand sides.
    template <int V>
    using Stage = std::conditional_t<(sizeof(T) < sizeof(double) && (V == 0 || V == 1)), double, T>;

    // One barrier per phase where LDS allows it (DEEP: the FP64 256-thread and
    // FP32 384-thread kernels, which hold one workgroup per CU anyway): three
    // region buffers and two D_x / D_y buffers, so phase K stages variable K+1
    // before its single barrier, and phase K+1's D_x / D_y and region stencils
    // follow phase K's equations with no barrier between them (the scheduler
    // overlaps the equations' arithmetic with those LDS round trips). Buffer
    // reuse: region buffer (K+1)%3 was last read by phase K-2 and D_x / D_y
    // buffer K&1 by phase K-2, both before barrier K-1, which every wave has
    // passed when it writes them in phase K. Other kernels keep two region
    // buffers and two barriers in phases with mixed derivatives.
    static constexpr bool DEEP = sizeof(T) == sizeof(double) ? C::FT == CfgWide::FT : C::FT == CfgFull::FT;
    template <int K>
    static constexpr int rbuf() { return DEEP ? K % 3 : K & 1; }

    // region of the variable of phase K (prefetch slot K % PD) into its LDS
    // buffer; buffers sit at multiples of NREG doubles whatever the staging type
    template <int K>
    __device__ __forceinline__ void stash() {
        constexpr int V = P::var(K);
        using S = Stage<V>;
        static_assert(sizeof(S) == sizeof(T) || DEEP || K % 2 == 0,
                      "FP64-staged variables need the FP64-sized (even) buffer");
        S *t = reinterpret_cast<S *>(s_tile + rbuf<K>() * NREG);
        if constexpr (sizeof(S) < sizeof(double)) {
            const double o = poff[K % PD];
#pragma unroll
            for (int c = 0; c < PREF; c++)
                if (c + 1 < PREF || tid + c * FT < NREG) t[tid + c * FT] = S(pre[K % PD][c] - o);
        } else {
#pragma unroll
            for (int c = 0; c < PREF; c++)
                if (c + 1 < PREF || tid + c * FT < NREG) t[tid + c * FT] = pre[K % PD][c];
        }
    }

    // Each phase stages the next phase's variable at its end (after its own
    // stencils), so the equations that follow a phase share a barrier-free
    // stretch with the next phase's D_x / D_y evaluations and the scheduler
    // can overlap them with those LDS round trips. (Staging at the start of
    // each phase instead measured ~1% slower in both precisions.)

    // Stages phase K and synchronizes (the kernel has fetched phases 0 .. PD-1).
    template <int K>
    __device__ __forceinline__ void stage_next() {
        stash<K>();
        if constexpr (K + PD < P::N) fetch<K + PD>();
        __syncthreads();
    }

    __device__ __forceinline__ void begin() {
        stage_next<0>();
    }

    // Phase KPH: variable V in LDS buffer KPH & 1; its derivatives, KO term
    // and boundary value.
    template <int V, int NEXT, int KPH>
    __device__ __forceinline__ void phase() {
        static_assert(P::var(KPH) == V, "phase order");
        using S = Stage<V>;
        S *t = reinterpret_cast<S *>(s_tile + rbuf<KPH>() * NREG);
        const dgr::AxisConsts &kx = K[0], &ky = K[1], &kz = K[2];
        constexpr int m2 = d2slot(V);
        S *sdx = reinterpret_cast<S *>(s_dx + (DEEP ? (KPH & 1) * C::NSDX : 0));
        S *sdy = reinterpret_cast<S *>(s_dy + (DEEP ? (KPH & 1) * C::NSDY : 0));
        if constexpr (m2 >= 0) {
            // D_x u on tile-x x region-y x region-z and D_y u on tile-x x tile-y x region-z
            // (the reference's mixed derivatives are D_y(D_x u), D_z(D_x u), D_z(D_y u))
            // every thread evaluates DXIT / DYIT points; those beyond the
            // tile's region read valid LDS and land in padding nobody reads
#pragma unroll
            for (int it = 0; it < DXIT; it++) {
                const int w = dxw[it];         // i0 + 13 jj + 169 kk
                const S *s = t + (w + PW);     // region point (3+i0, jj, kk)
                S d = dgr::d1_in<1>(s, c_d1<0, S>());
                if (FACE && (lox || hix)) d = dgr::d1_edge<1>(s, d, ox + w % FRX, iex, lox, hix, S(kx.d1_12));
                sdx[w] = d;
            }
#pragma unroll
            for (int it = 0; it < DYIT; it++) {
                const int w = dyw[it];                 // i0 + 13 j0 + 169 kk
                const S *s = t + (w + PW + FRX * PW);  // region point (3+i0, 3+j0, kk)
                S d = dgr::d1_in<FRX>(s, c_d1<1, S>());
                if (FACE && (loy || hiy))
                    d = dgr::d1_edge<FRX>(s, d, oy + (w / FRX) % FRX, iey, loy, hiy, S(ky.d1_12));
                sdy[w] = d;
            }
        }
        const S *c = t + cofs;
        if constexpr (m2 >= 0) {
            // the pure second derivatives, D_z and (interior tiles) the KO
            // term read only the staged region: they go before the barrier
            // that publishes D_x / D_y, in one stretch with those LDS loads
            S r;
            r = dgr::d2_in<1>(c, c_d2<0, S>());
            if (FACE && (lox || hix)) r = dgr::d2_edge<1>(c, r, ax, iex, lox, hix, S(kx.d2_12));
            d2v[0] = T(r);
            r = dgr::d2_in<FRX>(c, c_d2<1, S>());
            if (FACE && (loy || hiy)) r = dgr::d2_edge<FRX>(c, r, ay, iey, loy, hiy, S(ky.d2_12));
            d2v[3] = T(r);
            r = dgr::d2_in<FRXY>(c, c_d2<2, S>());
            if (FACE && (loz || hiz)) r = dgr::d2_edge<FRXY>(c, r, az, iez, loz, hiz, S(kz.d2_12));
            d2v[5] = T(r);
            r = dgr::d1_in<FRXY>(c, c_d1<2, S>());
            if (FACE && (loz || hiz)) r = dgr::d1_edge<FRXY>(c, r, az, iez, loz, hiz, S(kz.d1_12));
            d1v[2] = T(r);
            if constexpr (!FACE && P::KO) ko_term<V, S>(c, kx, ky, kz);
            if constexpr (DEEP && NEXT >= 0) {
                stash<KPH + 1>();
                if constexpr (KPH + 1 + PD < P::N) fetch<KPH + 1 + PD>();
            }
            __syncthreads();
            const S *cx = sdx + dxo;
            const S *cy = sdy + dyo;
            d1v[0] = T(cx[0]);
            d1v[1] = T(cy[0]);
            r = dgr::d1_in<DXS_Y>(cx, c_d1<1, S>());  // D_y(D_x u)
            if (FACE && (loy || hiy)) r = dgr::d1_edge<DXS_Y>(cx, r, ay, iey, loy, hiy, S(ky.d1_12));
            d2v[1] = T(r);
            r = dgr::d1_in<DXS_Z>(cx, c_d1<2, S>());  // D_z(D_x u)
            if (FACE && (loz || hiz)) r = dgr::d1_edge<DXS_Z>(cx, r, az, iez, loz, hiz, S(kz.d1_12));
            d2v[2] = T(r);
            r = dgr::d1_in<DYS_Z>(cy, c_d1<2, S>());  // D_z(D_y u)
            if (FACE && (loz || hiz)) r = dgr::d1_edge<DYS_Z>(cy, r, az, iez, loz, hiz, S(kz.d1_12));
            d2v[4] = T(r);
        } else {
            S r = dgr::d1_in<1>(c, c_d1<0, S>());
            if (FACE && (lox || hix)) r = dgr::d1_edge<1>(c, r, ax, iex, lox, hix, S(kx.d1_12));
            d1v[0] = T(r);
            r = dgr::d1_in<FRX>(c, c_d1<1, S>());
            if (FACE && (loy || hiy)) r = dgr::d1_edge<FRX>(c, r, ay, iey, loy, hiy, S(ky.d1_12));
            d1v[1] = T(r);
            r = dgr::d1_in<FRXY>(c, c_d1<2, S>());
            if (FACE && (loz || hiz)) r = dgr::d1_edge<FRXY>(c, r, az, iez, loz, hiz, S(kz.d1_12));
            d1v[2] = T(r);
        }
        // computed here, not at their last use (that would keep the LDS taps live)
        DGR_PIN(d1v[0]);
        DGR_PIN(d1v[1]);
        DGR_PIN(d1v[2]);
        if constexpr (m2 >= 0) {
#pragma unroll
            for (int k = 0; k < 6; k++) DGR_PIN(d2v[k]);
        }
        if constexpr (V == 1 && P::GAUGE) {  // chi: its first derivatives are at hand
            if (eta_fn) eta_v = eta_function<T>(f + 12, d1v, f[1], eta_r0, eta_p0, eta_p1);
        }
        if constexpr (V == 0 && P::GAUGE) {
            if (ssl_on) ssl_v = ssl_term<T>(f[0], f[1], ssl_e);
        }
        // on a flagged face the boundary value needs the first derivatives
        if constexpr (P::KO && (m2 < 0 || FACE)) ko_term<V, S>(c, kx, ky, kz);
        if constexpr (!DEEP || m2 < 0) {
            if constexpr (NEXT >= 0) stage_next<KPH + 1>();
        }
    }

    // KO term sigma*(ko_x+ko_y+ko_z) (rhs.cpp:434-489); on a flagged face the
    // boundary value replaces the equations (bssn_bcs, rhs.cpp:511-631) and the
    // KO term is added to it, as the reference adds KO after the boundary pass.
    template <int V, typename S>
    __device__ __forceinline__ void ko_term(const S *c, const dgr::AxisConsts &kx, const dgr::AxisConsts &ky,
                                            const dgr::AxisConsts &kz) {
#pragma clang fp contract(off)
        S kox = dgr::ko_in<1>(c, c_ko<0, S>());
        if (FACE && (lox || hix)) kox = dgr::ko_edge<1>(c, kox, ax, iex, lox, hix, kx);
        S koy = dgr::ko_in<FRX>(c, c_ko<1, S>());
        if (FACE && (loy || hiy)) koy = dgr::ko_edge<FRX>(c, koy, ay, iey, loy, hiy, ky);
        S koz = dgr::ko_in<FRXY>(c, c_ko<2, S>());
        if (FACE && (loz || hiz)) koz = dgr::ko_edge<FRXY>(c, koz, az, iez, loz, hiz, kz);
        const S kt = S(is_gauge(V) ? sig_gauge : sig_other) * (kox + koy + koz);
        T v = T(kt);
        if (FACE && face) {
            // the field value itself (the staged one may carry an offset)
            const double x = coord(pmin[0], ax, h[0]);
            const double y = coord(pmin[1], ay, h[1]);
            const double z = coord(pmin[2], az, h[2]);
            const double inv_r = inv_radius(x, y, z);
            const T bc = T(-inv_r) * (T(x) * d1v[0] + T(y) * d1v[1] + T(z) * d1v[2] +
                                      T(BC_FALLOFF[V]) * (f[V] - T(BC_ASYM[V])));
            v = bc + T(kt);
        }
        s_ko[V * FT + tid] = v;
    }

    template <int V, int D>
    __device__ __forceinline__ T d1() const { return d1v[D]; }
    template <int M, int K2>
    __device__ __forceinline__ T d2() const { return d2v[K2]; }
    template <int V>
    __device__ __forceinline__ T fld() const { return f[V]; }
    // the point's coordinates (the constraint kernel sets them)
    double xc[3];
    template <int D>
    __device__ __forceinline__ T crd() const { return T(xc[D]); }
    __device__ __forceinline__ void finish() {}
};

// The equations' outputs: on a flagged face the boundary value (already
// holding its KO term), elsewhere equations + KO term, kept in LDS until the
// epilogue.
template <bool FACE, class C, typename T, bool HALO, class P = RhsPhases>
struct FusedOut {
    static constexpr int FT = C::FT;
    const FusedCtx<FACE, C, T, P, HALO> &ctx;
    __device__ explicit FusedOut(const FusedCtx<FACE, C, T, P, HALO> &c) : ctx(c) {}
    template <int V>
    __device__ __forceinline__ void put(T val) {
#pragma clang fp contract(off)
        T *s = ctx.s_ko + V * FT + ctx.tid;
        const T kb = *s;
        *s = (FACE && ctx.face) ? kb : val + kb;
    }
};

// RHS store or RK4 stage update of one point (rkBSSN.cpp:764-860), with the
// loads of all 24 variables issued before any store. Stage s (U = {0, 1/2,
// 1/2, 1}, C = {1/6, 1/3, 1/3, 1/6}): acc = (s ? acc : u0) + C[s] dt k;
// next = u0 + U[s+1] dt k (s < 3) or the final state acc (s = 3).
// The RK arithmetic is FP64 in both precisions; the mixed path's FP32
// right-hand side is widened first and its u0 read back in FP64.
template <class C, typename T>
__device__ __forceinline__ void fused_epilogue(const StageIO &io, const T *s_ko, const T *f, int tid,
                                               uint64_t idx, int64_t q, double chi_floor) {
#pragma clang fp contract(off)
    constexpr int FT = C::FT;
    const int s = io.stage;
    if (s < 0) {
#pragma unroll
        for (int v = 0; v < NV; v++) io.rhs.p[v][idx] = double(s_ko[v * FT + tid]);
        return;
    }
    double u0[NV], a0[NV], v14[14];
    if (s == 0 && sizeof(T) == sizeof(double)) {
#pragma unroll
        for (int v = 0; v < NV; v++) u0[v] = a0[v] = double(f[v]);
    } else if (s == 0) {
#pragma unroll
        for (int v = 0; v < NV; v++) u0[v] = a0[v] = io.src.p[v][idx];
    } else {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            if (s < 3) u0[v] = io.u0.p[v][q];
            a0[v] = io.acc.p[v][q];
        }
    }
    const bool enf = s == 3 || io.enforce_stage;
#pragma unroll
    for (int v = 0; v < NV; v++) {
        if (s == 0) io.u0.p[v][q] = u0[v];
        const double k = double(s_ko[v * FT + tid]);
        const double acc = a0[v] + io.cacc * k;
        double out;
        if (s < 3) {
            io.acc.p[v][q] = acc;
            out = u0[v] + io.cnext * k;
        } else {
            out = acc;
        }
        if (slot14(v) >= 0 && enf) v14[slot14(v)] = out;
        else io.dst.p[v][idx] = out;
    }
    if (!enf) return;
    const int e = enforce_point(v14, chi_floor);
    if (e) atomicOr(io.err, e);
#pragma unroll
    for (int c = 0; c < 14; c++) {
        const int v = c < 2 ? c : c + 10;
        io.dst.p[v][idx] = v14[c];
    }
}

// Tile geometry, this thread's point, the region elements it fetches and its
// D_x / D_y slots (common to the fused kernels): idx is the point's padded
// index, q its interior index, active false for shadow lanes.
template <bool FACE, class C, typename T, bool HALO, class P>
__device__ __forceinline__ void tile_setup(FusedCtx<FACE, C, T, P, HALO> &ctx, const StageIO &io, const Tile &tl,
                                           const DBlock &B, const uint16_t *__restrict__ tabs, int dtab,
                                           uint64_t &idx_out, int64_t &q_out, bool &active_out) {
    constexpr int FT = C::FT, PREF = C::PREF, DXIT = C::DXIT, DYIT = C::DYIT;
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const int ex = tl.e[0], ey = tl.e[1], ez = tl.e[2];
    ctx.ox = tl.o[0];
    ctx.oy = tl.o[1];
    ctx.oz = tl.o[2];
    ctx.iex = nx - PW;
    ctx.iey = ny - PW;
    ctx.iez = nz - PW;
    const uint32_t bf = B.bflag;
    ctx.lox = bf & 1u; ctx.hix = bf & 2u; ctx.loy = bf & 4u; ctx.hiy = bf & 8u; ctx.loz = bf & 16u; ctx.hiz = bf & 32u;
    ctx.boff = B.offset;
    ctx.corner = (uint32_t)((ctx.ox - PW) + nx * ((ctx.oy - PW) + ny * (ctx.oz - PW)));
    for (int a = 0; a < 3; a++) {
        ctx.pmin[a] = B.pmin[a];
        ctx.h[a] = B.h[a];
    }
    const int mx = nx - 2 * PW, my = ny - 2 * PW;
    const int tid = threadIdx.x;
    ctx.tid = tid;
    // this thread's point: from the tile shape's lane table (conflict-free LDS
    // access for the own-point stencils); shadow lanes repeat a point
    const int lt = tabs[tl.lanetab + tid];
    const bool active = !(lt & LANE_SHADOW);
    const int p = lt & (LANE_SHADOW - 1);
    const int li = p % ex, lj = (p / ex) % ey, lk = p / (ex * ey);
    ctx.ax = ctx.ox + li;
    ctx.ay = ctx.oy + lj;
    ctx.az = ctx.oz + lk;
    ctx.cofs = (PW + li) + FRX * (PW + lj) + FRXY * (PW + lk);
    ctx.dxo = li + DXS_Y * (PW + lj) + DXS_Z * (PW + lk);
    ctx.dyo = li + DYS_Y * lj + DYS_Z * (PW + lk);
    const uint64_t idx = B.offset + (uint64_t)ctx.ax + (uint64_t)nx * ((uint64_t)ctx.ay + (uint64_t)ny * ctx.az);
    const int64_t q = B.q0 + (ctx.ax - PW) + (int64_t)mx * ((ctx.ay - PW) + (int64_t)my * (ctx.az - PW));
    ctx.face = (ctx.lox && ctx.ax == PW) || (ctx.hix && ctx.ax == nx - PW - 1) || (ctx.loy && ctx.ay == PW) ||
               (ctx.hiy && ctx.ay == ny - PW - 1) || (ctx.loz && ctx.az == PW) || (ctx.hiz && ctx.az == nz - PW - 1);
    // region elements: e = tid + c*FT -> (ii, jj, kk) in the 13 x 13 x (ez+6) box;
    // elements outside the region (x or y beyond a short tile, z beyond ez+6)
    // load the block's first point and are never read
    if constexpr (HALO) {
        // offsets are byte offsets into the whole variable (below 2^32: the
        // host runs this kernel only for variables below 2^29 points); a
        // padding point inside a linked
        // same-level neighbour reads that neighbour's interior point at the
        // same coordinates (node index minus d (n - 2 PW - 1) per axis)
        ctx.boff = 0;
        const int64_t *__restrict__ nbo = io.nbo + 27 * (size_t)tl.block;
#pragma unroll
        for (int c = 0; c < PREF; c++) {
            const int e = tid + c * FT;
            const int kk = e / FRXY, r = e - kk * FRXY;
            const int jj = r / FRX, ii = r - jj * FRX;
            const bool in_region = ii < ex + 2 * PW && jj < ey + 2 * PW && kk < ez + 2 * PW;
            const int gi = ctx.ox - PW + ii, gj = ctx.oy - PW + jj, gk = ctx.oz - PW + kk;
            const int di = gi < PW ? -1 : (gi >= nx - PW ? 1 : 0);
            const int dj = gj < PW ? -1 : (gj >= ny - PW ? 1 : 0);
            const int dk = gk < PW ? -1 : (gk >= nz - PW ? 1 : 0);
            const int64_t o = (in_region && (di | dj | dk)) ? nbo[(di + 1) + 3 * (dj + 1) + 9 * (dk + 1)] : -1;
            const uint64_t a =
                o >= 0 ? (uint64_t)o + (uint64_t)((gi - di * (nx - 2 * PW - 1)) +
                                                   nx * ((gj - dj * (ny - 2 * PW - 1)) + ny * (gk - dk * (nz - 2 * PW - 1))))
                       : B.offset + (uint64_t)(in_region ? gi + nx * (gj + ny * gk) : 0);
            ctx.goff[c] = (uint32_t)(a * sizeof(double));
        }
        if constexpr (sizeof(T) < sizeof(double)) {
            // the mixed-precision staging offset: the value at the region's
            // first point, read where the copy path would have copied it from
            const int gi = ctx.ox - PW, gj = ctx.oy - PW, gk = ctx.oz - PW;
            const int di = gi < PW ? -1 : 0, dj = gj < PW ? -1 : 0, dk = gk < PW ? -1 : 0;
            const int64_t o = (di | dj | dk) ? nbo[(di + 1) + 3 * (dj + 1) + 9 * (dk + 1)] : -1;
            ctx.corner_src = o >= 0 ? (uint64_t)o + (uint64_t)((gi - di * (nx - 2 * PW - 1)) +
                                                               nx * ((gj - dj * (ny - 2 * PW - 1)) +
                                                                     ny * (gk - dk * (nz - 2 * PW - 1))))
                                    : B.offset + (uint64_t)(gi + nx * (gj + ny * gk));
        }
    } else {
#pragma unroll
        for (int c = 0; c < PREF; c++) {
            const int e = tid + c * FT;
            const int kk = e / FRXY, r = e - kk * FRXY;
            const int jj = r / FRX, ii = r - jj * FRX;
            const bool in_region = ii < ex + 2 * PW && jj < ey + 2 * PW && kk < ez + 2 * PW;
            const int gi = ctx.ox - PW + ii, gj = ctx.oy - PW + jj, gk = ctx.oz - PW + kk;
            ctx.goff[c] = in_region ? (uint32_t)(gi + nx * (gj + ny * gk)) * (uint32_t)sizeof(double) : 0u;
        }
    }
    // where this thread's D_x / D_y evaluations go (the config's slot table)
#pragma unroll
    for (int it = 0; it < DXIT; it++) ctx.dxw[it] = tabs[dtab + it * FT + tid];
#pragma unroll
    for (int it = 0; it < DYIT; it++) ctx.dyw[it] = tabs[dtab + (DXIT + it) * FT + tid];
    idx_out = idx;
    q_out = q;
    active_out = active;
}

template <bool FACE, class C, typename T, bool HALO = false, class P = RhsPhases>
__device__ __forceinline__ void fused_tile(const StageIO &io, const RhsParams &prm, const Tile &tl, const DBlock &B,
                                           const dgr::AxisConsts *K, const uint16_t *__restrict__ tabs,
                                           int dtab, double *s_tile, double *s_dx, double *s_dy, T *s_ko) {
    FusedCtx<FACE, C, T, P, HALO> ctx(io);
    ctx.K = K;
    ctx.load_consts();
    ctx.s_tile = s_tile;
    ctx.s_dx = s_dx;
    ctx.s_dy = s_dy;
    ctx.s_ko = s_ko;
    uint64_t idx;
    int64_t q;
    bool active;
    tile_setup<FACE, C, T, HALO>(ctx, io, tl, B, tabs, dtab, idx, q, active);
    // this point's field values (registers; the compiler parks them in AGPRs)
    {
        T *f = ctx.f;
#pragma unroll
        for (int v = 0; v < NV; v++) f[v] = T(io.src.p[v][idx]);
        if (prm.cako) {
            const T sg = sqrt(f[1]);
            ctx.sig_gauge = sg * T(prm.eps_gauge);
            ctx.sig_other = sg * T(prm.eps_other);
        } else {
            ctx.sig_gauge = ctx.sig_other = T(prm.ko_sigma);
        }
    }
    // eta: the RIT eta(r), or (P::GAUGE with eta_func) set in chi's phase from
    // chi's gradient
    if (P::GAUGE && prm.eta_func)
        ctx.eta_v = T(0);
    else
        ctx.eta_v = T(rit_eta(coord(B.pmin[0], ctx.ax, B.h[0]), coord(B.pmin[1], ctx.ay, B.h[1]),
                              coord(B.pmin[2], ctx.az, B.h[2]), prm));
    ctx.eta_r0 = prm.eta_r0;
    ctx.eta_p0 = prm.eta_p0;
    ctx.eta_p1 = prm.eta_p1;
    if constexpr (P::GAUGE) {
        ctx.eta_fn = prm.eta_func != 0;
        ctx.ssl_on = prm.ssl != 0;
        ctx.ssl_e = prm.ssl_dev ? prm.ssl_dev[1] : prm.ssl_e;
        ctx.ssl_v = T(0);
    }
    const T lam[4] = {T(prm.lam[0]), T(prm.lam[1]), T(prm.lam[2]), T(prm.lam[3])};
    const T lamf[2] = {T(prm.lamf[0]), T(prm.lamf[1])};
    ctx.template fetch<0>();
    if constexpr (C::PD > 1) ctx.template fetch<1>();
    if constexpr (C::PD > 2) ctx.template fetch<2>();
    static_assert(C::PD <= 3, "prefetch distance");
    ctx.begin();
    FusedOut<FACE, C, T, HALO, P> out(ctx);
    if constexpr (P::GAUGE) {
        const T xi[3] = {T(prm.xi[0]), T(prm.xi[1]), T(prm.xi[2])};
        dgr::bssn_rhs_phased_gauge<T>(ctx, ctx.eta_v, lam, lamf, xi, prm.rochester, ctx.ssl_v, out);
    } else {
        dgr::bssn_rhs_phased<T>(ctx, ctx.eta_v, lam, lamf, out);
    }
    if (active) fused_epilogue<C, T>(io, s_ko, ctx.f, ctx.tid, idx, q, prm.chi_floor);
}

// One launch per tile kind: FACE = false over the tiles of blocks without
// flagged faces, FACE = true over the others; C the tile shape; T the
// arithmetic type (FP32 asks for two waves per SIMD: its registers allow it).
#ifndef DGR_FP32_WAVES
#define DGR_FP32_WAVES 2  // waves per SIMD the FP32 kernels are built for
#endif

template <bool FACE, class C, typename T, bool HALO = false, bool GAUGE = false>
__global__ __launch_bounds__(C::FT, sizeof(T) < sizeof(double) ? DGR_FP32_WAVES : 1) void k_fused(
    const StageIO io, const RhsParams prm, const DBlock *__restrict__ blocks,
    const dgr::AxisConsts *__restrict__ bconst, const Tile *__restrict__ tiles,
    const uint16_t *__restrict__ tabs, int dtab) {
    // the region buffers sit at multiples of NREG doubles; with two buffers the
    // FP32 kernel keeps FP64 data in the even-phase one only, the odd one FP32
    constexpr bool DEEP = FusedCtx<FACE, C, T>::DEEP;
    __shared__ double s_tile[DEEP ? 3 * C::NREG : C::NREG + (sizeof(T) < sizeof(double) ? (C::NREG + 1) / 2 : C::NREG)];
    __shared__ double s_dx[(DEEP ? 2 : 1) * C::NSDX];
    __shared__ double s_dy[(DEEP ? 2 : 1) * C::NSDY];
    __shared__ T s_ko[NV * C::FT];

    const Tile tl = tiles[blockIdx.x];
    const DBlock B = blocks[tl.block];
    fused_tile<FACE, C, T, HALO, std::conditional_t<GAUGE, RhsPhasesGauge, RhsPhases>>(
        io, prm, tl, B, bconst + 3 * (size_t)tl.block, tabs, dtab, s_tile, s_dx, s_dy, s_ko);
}

struct ConsPtrs {
    double *c[6];  // ham, mom0, mom1, mom2, psi4 real, psi4 imaginary (any may be null)
};

// The constraint diagnostic as one fused kernel per tile kind (the k_fused
// machinery with the constraints' phase table, bssn_cons_phased.h): the 17
// variables whose derivatives physical_constraints reads (physcon.cpp:12-89)
// stream through LDS one at a time and the generated constraints run as far
// as those derivatives allow; no derivative workspace. Outputs go straight to
// the caller's arrays (null ones are skipped); psi4 is zero within 1e-7 of
// the z axis (physcon.cpp:82-85).
struct ConsFusedOut {
    double v[6];
    template <int I>
    __device__ __forceinline__ void put(double x) { v[I] = x; }
};

template <bool FACE, class C>
__global__ __launch_bounds__(C::FT, 1) void k_cons_fused(const StageIO io, ConsPtrs o,
                                                        const DBlock *__restrict__ blocks,
                                                        const dgr::AxisConsts *__restrict__ bconst,
                                                        const Tile *__restrict__ tiles,
                                                        const uint16_t *__restrict__ tabs, int dtab) {
    using Ctx = FusedCtx<FACE, C, double, ConsPhases>;
    constexpr bool DEEP = Ctx::DEEP;
    __shared__ double s_tile[(DEEP ? 3 : 2) * C::NREG];
    __shared__ double s_dx[(DEEP ? 2 : 1) * C::NSDX];
    __shared__ double s_dy[(DEEP ? 2 : 1) * C::NSDY];
    const Tile tl = tiles[blockIdx.x];
    const DBlock B = blocks[tl.block];
    Ctx ctx(io);
    ctx.K = bconst + 3 * (size_t)tl.block;
    ctx.load_consts();
    ctx.s_tile = s_tile;
    ctx.s_dx = s_dx;
    ctx.s_dy = s_dy;
    ctx.s_ko = nullptr;
    uint64_t idx;
    int64_t q;
    bool active;
    tile_setup<FACE, C, double, false>(ctx, io, tl, B, tabs, dtab, idx, q, active);
#pragma unroll
    for (int v = 0; v < NV; v++) ctx.f[v] = io.src.p[v][idx];
    // x = pmin + i h as physcon.cpp:75-77 computes it
    ctx.xc[0] = __dadd_rn(B.pmin[0], __dmul_rn((double)ctx.ax, B.h[0]));
    ctx.xc[1] = __dadd_rn(B.pmin[1], __dmul_rn((double)ctx.ay, B.h[1]));
    ctx.xc[2] = __dadd_rn(B.pmin[2], __dmul_rn((double)ctx.az, B.h[2]));
    ctx.template fetch<0>();
    if constexpr (C::PD > 1) ctx.template fetch<1>();
    if constexpr (C::PD > 2) ctx.template fetch<2>();
    ctx.begin();
    ConsFusedOut out;
    dgr::bssn_cons_phased<double>(ctx, out);
    if (!active) return;
    const bool axis = fabs(ctx.xc[0]) <= 1e-7 && fabs(ctx.xc[1]) <= 1e-7;
#pragma unroll
    for (int i = 0; i < 6; i++)
        if (o.c[i]) o.c[i][idx] = (i >= 4 && axis) ? 0.0 : out.v[i];
}

// Copies the padding points of every block of src into dst (one workgroup per
// block and variable): the RK4 work state's padding, refreshed from u once per
// call of the RK4 entry points (u's padding is frozen while they run).
__global__ __launch_bounds__(256) void k_pad_copy(VarPtrs dst, VarPtrs src, const DBlock *__restrict__ blocks) {
    const DBlock &B = blocks[blockIdx.x];
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const int n = nx * ny * nz;
    const double *__restrict__ s = src.p[blockIdx.y] + B.offset;
    double *__restrict__ d = dst.p[blockIdx.y] + B.offset;
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        const int i = p % nx, j = (p / nx) % ny, k = p / (nx * ny);
        const bool inner = i >= PW && i < nx - PW && j >= PW && j < ny - PW && k >= PW && k < nz - PW;
        if (!inner) d[p] = s[p];
    }
}

// Padding refresh from same-level neighbours, the copy part of the
// reference's unzip (bssnCtx.cpp:75): a padding point of block b lying in the
// interior of the neighbour nb in direction (di,dj,dk) takes nb's value at the
// same coordinates. Neighbours share their boundary nodes, so nb's node index
// is b's minus d*(n-2*PW-1) along each axis; every source is an interior point
// and every destination a padding point, so blocks can be filled in place in
// any order. nbr: 27 entries per block, -1 for none (that padding is kept).
// One thread per point, one workgroup per 256-point chunk of a block
// (blockIdx.x = chunk * nblocks + block: consecutive workgroups take the same
// chunk of consecutive blocks, which measured faster than a block's chunks
// in a row): the
// neighbours' offsets are staged in LDS, and each padding point's index
// arithmetic is done once for all 24 variables, whose loads are issued
// together before the stores.
__global__ __launch_bounds__(256) void k_halo_fill(VarPtrs u, const DBlock *__restrict__ blocks,
                                                   const int *__restrict__ nbr, const int2 *__restrict__ wgs) {
    __shared__ int64_t s_off[27];
    const int2 w = wgs[blockIdx.x];  // (block, chunk of 256 points)
    const int b = w.x, chunk = w.y;
    const DBlock &B = blocks[b];
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const int n = nx * ny * nz;
    if (threadIdx.x < 27) {
        const int nb = nbr[27 * (size_t)b + threadIdx.x];
        s_off[threadIdx.x] = (nb >= 0 && threadIdx.x != 13) ? (int64_t)blocks[nb].offset : -1;
    }
    __syncthreads();
    const int p = chunk * blockDim.x + threadIdx.x;
    if (p < n) {
        const int i = p % nx, j = (p / nx) % ny, k = p / (nx * ny);
        const int di = i < PW ? -1 : (i >= nx - PW ? 1 : 0);
        const int dj = j < PW ? -1 : (j >= ny - PW ? 1 : 0);
        const int dk = k < PW ? -1 : (k >= nz - PW ? 1 : 0);
        const int64_t off = s_off[(di + 1) + 3 * (dj + 1) + 9 * (dk + 1)];  // -1 for interior points too
        if (off < 0) return;
        const int si = i - di * (nx - 2 * PW - 1), sj = j - dj * (ny - 2 * PW - 1), sk = k - dk * (nz - 2 * PW - 1);
        const size_t src = (size_t)off + si + (size_t)nx * (sj + (size_t)ny * sk);
        const size_t dst = B.offset + p;
        double t[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) t[v] = u.p[v][src];
#pragma unroll
        for (int v = 0; v < NV; v++) u.p[v][dst] = t[v];
    }
}

// The rest of the padding refresh (the interpolating part of unzip,
// BSSN_GR/src/bssnCtx.cpp:75; the octree code is in dendrolib): one record
// per padding point that has no same-size neighbour in its direction but lies
// in the interior of another block, which gives the value: a block of the same
// spacing (any size) or of half the spacing (a finer block: the point is one of
// its nodes, injection) by a copy, a block of twice the spacing (a coarser
// block) by order-6 Lagrange interpolation in its element holding the point.
// Records read interior points only and write padding points only.
struct FillRec {
    uint64_t dst;   // the padding point
    uint64_t src;   // copy: the source node; interpolation: the element's first node
    uint32_t mx;    // interpolation: x and xy strides of the source block
    uint32_t mxy;
    uint32_t code;  // bits 0-3, 4-7, 8-11: 2t per axis, t in [0, 6] the element-local
                    // coordinate; bit 12: interpolation (else copy)
    uint32_t pad;
};
constexpr uint32_t FILL_INTERP = 1u << 12;

// The weights of the 7 nodes of an order-6 element at t = code/2, code 0..12
// (the Lagrange basis), computed on the host in the oracle's operation order
// (orc_lagrange6_half) and passed by value.
struct Lag6Table {
    double w[13][7];
};

// The copy records: one thread per (record, variable), blockIdx.y the
// variable. They come first in the record list.
__global__ __launch_bounds__(256) void k_fill_copy(VarPtrs u, const FillRec *__restrict__ recs, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double *__restrict__ uv = u.p[blockIdx.y];
    const FillRec r = recs[q];
    uv[r.dst] = uv[r.src];
}

// Interpolation records that read one coarse element: records [first,
// first + count), count <= FILL_GROUP_MAX, the element's first node and strides.
struct FillGroup {
    uint64_t src;
    uint32_t mx, mxy;
    uint32_t first, count;
};

// The value at element coordinates (cx, cy, cz)/2 (codes 0..12) is, per axis,
// the Lagrange sum over the axis' 7 nodes (odd code) or the one node it sits
// on (even code, weight 1), x innermost, then y, then z, each sum started at 0:
//   sx(b, c) = sum_a wx[a] e(a, b, c),  sy(c) = sum_b wy[b] sx(b, c),  v = sum_c wz[c] sy(c).
// k_fill_interp evaluates sx for all 13 x codes and sy for all 13 x 13 code
// pairs once per element and variable, then each record's last sum: the
// same products and additions, in the same order, as evaluating a record
// alone. The x and y passes loop over codes uniformly across a wavefront, so
// their weights are scalar operands. One group holds the records of one
// element (up to FILL_GROUP_MAX), so the element is read once.
constexpr int FILL_WG = 256;           // threads per workgroup
constexpr int FILL_RPT = 4;            // records per thread
constexpr int FILL_GROUP_MAX = FILL_WG * FILL_RPT;  // records per group
constexpr int FILL_EPL = (343 + FILL_WG - 1) / FILL_WG;  // element nodes each thread stages
constexpr int FILL_NVB = 2;  // variables per round of passes
static_assert(NV % FILL_NVB == 0, "rounds of whole variables");

__global__ __launch_bounds__(FILL_WG) void k_fill_interp(VarPtrs u, const FillRec *__restrict__ recs,
                                                         const FillGroup *__restrict__ groups, const Lag6Table tab) {
#pragma clang fp contract(off)
    __shared__ double s_e[FILL_NVB][343];       // the element, node (a, b, c) at a + 7 b + 49 c
    __shared__ double s_x[FILL_NVB][13 * 49];   // sx(b, c) of x code cx at 49 cx + b + 7 c
    __shared__ double s_y[FILL_NVB][13 * 91];   // sy(c) of codes (cx, cy) at 7 (13 cx + cy) + c
    __shared__ double s_w[13 * 7];              // the weights, code-major
    const FillGroup g = groups[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int k = tid; k < 13 * 7; k += FILL_WG) s_w[k] = (&tab.w[0][0])[k];
    // this thread's records (a load inside the z pass would wait for the
    // stores before it: loads and stores share one counter)
    uint64_t rdst[FILL_RPT];
    uint32_t rcode[FILL_RPT];
#pragma unroll
    for (int i = 0; i < FILL_RPT; i++) {
        const int q = tid + FILL_WG * i;
        rdst[i] = 0;
        rcode[i] = 0;
        if (q < (int)g.count) {
            rdst[i] = recs[g.first + q].dst;
            rcode[i] = recs[g.first + q].code;
        }
    }
    size_t off[FILL_EPL];
#pragma unroll
    for (int k = 0; k < FILL_EPL; k++) {
        const int e = min(tid + FILL_WG * k, 342);
        off[k] = g.src + (size_t)(e % 7) + (size_t)g.mx * ((e / 7) % 7) + (size_t)g.mxy * (e / 49);
    }
    // every variable's nodes in flight at once; the empty asm makes each value
    // opaque, so the compiler can neither sink a load to its variable's pass
    // nor issue it again there (it did, and each pass then waited for its
    // loads' full latency)
    double pre[NV][FILL_EPL];
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int k = 0; k < FILL_EPL; k++) pre[v][k] = tid + FILL_WG * k < 343 ? u.p[v][off[k]] : 0.0;
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int k = 0; k < FILL_EPL; k++) asm volatile("" : "+v"(pre[v][k]));
    // FILL_NVB variables per round of passes
#pragma unroll
    for (int v0 = 0; v0 < NV; v0 += FILL_NVB) {
#pragma unroll
        for (int m = 0; m < FILL_NVB; m++)
#pragma unroll
            for (int k = 0; k < FILL_EPL; k++)
                if (tid + FILL_WG * k < 343) s_e[m][tid + FILL_WG * k] = pre[v0 + m][k];
        __syncthreads();
        // x pass: lane (b, c) < 49 of wave w, x codes w, w + 4, ...
        if (lane < 49) {
            double row[FILL_NVB][7];
#pragma unroll
            for (int m = 0; m < FILL_NVB; m++)
#pragma unroll
                for (int a1 = 0; a1 < 7; a1++) row[m][a1] = s_e[m][a1 + 7 * lane];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int x = wave + 4 * j;
                if (x > 12) break;
#pragma unroll
                for (int m = 0; m < FILL_NVB; m++) {
                    double sx = 0.0;
                    if (x & 1) {
#pragma unroll
                        for (int a1 = 0; a1 < 7; a1++) sx += tab.w[x][a1] * row[m][a1];
                    } else {
                        sx += 1.0 * s_e[m][(x >> 1) + 7 * lane];  // (a register index would go to scratch)
                    }
                    s_x[m][49 * x + lane] = sx;
                }
            }
        }
        __syncthreads();
        // y pass: lane + 64 (w & 1) = (x code, c) < 91, y codes (w >> 1), + 2, ...
        {
            const int pc = lane + 64 * (wave & 1);
            if (pc < 91) {
                const int x = pc / 7, c3 = pc - 7 * x;
                double col[FILL_NVB][7];
#pragma unroll
                for (int m = 0; m < FILL_NVB; m++)
#pragma unroll
                    for (int b2 = 0; b2 < 7; b2++) col[m][b2] = s_x[m][49 * x + b2 + 7 * c3];
#pragma unroll
                for (int j = 0; j < 7; j++) {
                    const int y = (wave >> 1) + 2 * j;
                    if (y > 12) break;
#pragma unroll
                    for (int m = 0; m < FILL_NVB; m++) {
                        double sy = 0.0;
                        if (y & 1) {
#pragma unroll
                            for (int b2 = 0; b2 < 7; b2++) sy += tab.w[y][b2] * col[m][b2];
                        } else {
                            sy += 1.0 * s_x[m][49 * x + (y >> 1) + 7 * c3];
                        }
                        s_y[m][7 * (13 * x + y) + c3] = sy;
                    }
                }
            }
        }
        __syncthreads();
        // z pass: the group's records, FILL_WG at a time
#pragma unroll
        for (int i = 0; i < FILL_RPT; i++) {
            if (tid + FILL_WG * i >= (int)g.count) break;
            const uint32_t cd = rcode[i];
            const int cx = (int)(cd & 15u), cy = (int)((cd >> 4) & 15u), cz = (int)((cd >> 8) & 15u);
            const double *w = s_w + 7 * cz;
#pragma unroll
            for (int m = 0; m < FILL_NVB; m++) {
                const double *col = s_y[m] + 7 * (13 * cx + cy);
                double sz = 0.0;
                if (cz & 1) {
#pragma unroll
                    for (int c3 = 0; c3 < 7; c3++) sz += w[c3] * col[c3];
                } else {
                    sz += 1.0 * col[cz >> 1];
                }
                u.p[v0 + m][rdst[i]] = sz;
            }
        }
    }
}

// -------------------------------------------------------------------------
// Hamiltonian constraint (C_HAM of physical_constraints, physcon.cpp:12-89):
// the generated constraint at every interior point of a chunk, from the
// derivative workspace k_derivs fills with the same calls the reference's
// constraint_derivs.h makes. A diagnostic: it reuses the split path rather
// than a fused kernel of its own.
// -------------------------------------------------------------------------
// physical_constraints (physcon.cpp:12-89): the generated constraint header
// produces ham (output 0), mom0..2 (outputs 1-3) and psi4 real/imaginary
// (outputs 4-5). MASK selects the groups kept (CONS_HAM, CONS_MOM, CONS_PSI4);
// the compiler drops the expressions of the others.
enum { CONS_HAM = 1, CONS_MOM = 2, CONS_PSI4 = 4 };

template <int MASK>
struct ConsOut {
    double v[6];
    template <int V>
    __device__ __forceinline__ void put(double x) {
        if ((V == 0 && (MASK & CONS_HAM)) || (V >= 1 && V <= 3 && (MASK & CONS_MOM)) ||
            (V >= 4 && (MASK & CONS_PSI4)))
            v[V] = x;
    }
};


template <int MASK>
__global__ __launch_bounds__(NTHREADS) void k_constraints(VarPtrs u, ConsPtrs o,
                                                          const DBlock *__restrict__ blocks,
                                                          const Tile *__restrict__ tiles,
                                                          const int *__restrict__ wg_tile, int tile_end,
                                                          const double *__restrict__ W, int64_t P) {
    const int64_t q = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
    if (q >= P) return;
    int t = wg_tile[blockIdx.x];
    while (t + 1 < tile_end && q >= tiles[t + 1].q0) t++;
    const Tile T = tiles[t];
    const DBlock B = blocks[T.block];
    const PointGeom g = locate(T, B, (int)(q - T.q0));
    // x = pmin + i h as physcon.cpp:75-77 computes it
    PointIn in{&u, g.idx, W, P, q,
               {__dadd_rn(B.pmin[0], __dmul_rn((double)g.i, B.h[0])),
                __dadd_rn(B.pmin[1], __dmul_rn((double)g.j, B.h[1])),
                __dadd_rn(B.pmin[2], __dmul_rn((double)g.k, B.h[2]))}};
    ConsOut<MASK> out;
    dgr::bssn_ham_point<double>(in, out);
    if (MASK & CONS_HAM)
        if (o.c[0]) o.c[0][g.idx] = out.v[0];
    if (MASK & CONS_MOM) {
#pragma unroll
        for (int i = 1; i < 4; i++)
            if (o.c[i]) o.c[i][g.idx] = out.v[i];
    }
    if (MASK & CONS_PSI4) {
        // psi4 is set to zero within 1e-7 of the z axis (physcon.cpp:82-85)
        const bool axis = fabs(in.x[0]) <= 1e-7 && fabs(in.x[1]) <= 1e-7;
#pragma unroll
        for (int i = 4; i < 6; i++)
            if (o.c[i]) o.c[i][g.idx] = axis ? 0.0 : out.v[i];
    }
}

// computeConstraintL2Norm (grUtils.tcc:37-134): per workgroup, the sum of c^2
// and the count over interior points at distance >= r1 from BH 1 and >= r2
// from BH 2; k_norm_final adds the partials in a fixed order.
struct NormMask {
    double bh[6];
    double r1, r2;
};

// With a neighbour table (27 per block, -1 for none), a node on a face, edge
// or corner of block b's interior (side s_a = -1 at index PW, +1 at n-PW-1) is
// shared with the neighbours in the directions d with d_a in {0, s_a}; these
// are linked to each other, so all of them see the same set, and the one with
// the lowest block index counts the node: each zipped node is counted once, as
// computeConstraintL2Norm does (grUtils.tcc:60-110).
__device__ __forceinline__ bool owned_elsewhere(const int *__restrict__ nbr, int b, const PointGeom &g,
                                                const DBlock &B) {
    if (!nbr) return false;
    const int sx = g.i == PW ? -1 : (g.i == B.n[0] - PW - 1 ? 1 : 0);
    const int sy = g.j == PW ? -1 : (g.j == B.n[1] - PW - 1 ? 1 : 0);
    const int sz = g.k == PW ? -1 : (g.k == B.n[2] - PW - 1 ? 1 : 0);
    if (!(sx | sy | sz)) return false;
    for (int dz = min(sz, 0); dz <= max(sz, 0); dz++)
        for (int dy = min(sy, 0); dy <= max(sy, 0); dy++)
            for (int dx = min(sx, 0); dx <= max(sx, 0); dx++) {
                if (!(dx | dy | dz)) continue;
                const int o = nbr[27 * b + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
                if (o >= 0 && o < b) return true;
            }
    return false;
}

__global__ __launch_bounds__(NTHREADS) void k_norm_partial(const double *__restrict__ c, NormMask m,
                                                          const DBlock *__restrict__ blocks,
                                                          const Tile *__restrict__ tiles,
                                                          const int *__restrict__ wg_tile, int tile_end, int64_t P,
                                                          const int *__restrict__ nbr, double *__restrict__ psum,
                                                          unsigned long long *__restrict__ pcnt) {
#pragma clang fp contract(off)
    __shared__ double s_sum[NTHREADS];
    __shared__ unsigned long long s_cnt[NTHREADS];
    const int64_t q = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
    double v2 = 0.0;
    unsigned long long cnt = 0;
    if (q < P) {
        int t = wg_tile[blockIdx.x];
        while (t + 1 < tile_end && q >= tiles[t + 1].q0) t++;
        const Tile T = tiles[t];
        const DBlock B = blocks[T.block];
        const PointGeom g = locate(T, B, (int)(q - T.q0));
        const double x = coord(B.pmin[0], g.i, B.h[0]);
        const double y = coord(B.pmin[1], g.j, B.h[1]);
        const double z = coord(B.pmin[2], g.k, B.h[2]);
        const double d1 = sqrt((x - m.bh[0]) * (x - m.bh[0]) + (y - m.bh[1]) * (y - m.bh[1]) +
                               (z - m.bh[2]) * (z - m.bh[2]));
        const double d2 = sqrt((x - m.bh[3]) * (x - m.bh[3]) + (y - m.bh[4]) * (y - m.bh[4]) +
                               (z - m.bh[5]) * (z - m.bh[5]));
        if (!(d1 < m.r1 || d2 < m.r2) && !owned_elsewhere(nbr, T.block, g, B)) {
            const double v = c[g.idx];
            v2 = v * v;
            cnt = 1;
        }
    }
    s_sum[threadIdx.x] = v2;
    s_cnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = NTHREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        psum[blockIdx.x] = s_sum[0];
        pcnt[blockIdx.x] = s_cnt[0];
    }
}

// out[0], out[1] = sum and count; with norm_out, also *norm_out = sqrt(sum /
// count), or -1 when no point qualifies.
__global__ __launch_bounds__(NTHREADS) void k_norm_final(const double *__restrict__ psum,
                                                        const unsigned long long *__restrict__ pcnt, int n,
                                                        double *__restrict__ out, double *__restrict__ norm_out) {
#pragma clang fp contract(off)
    __shared__ double s_sum[NTHREADS];
    __shared__ unsigned long long s_cnt[NTHREADS];
    double a = 0.0;
    unsigned long long k = 0;
    for (int i = threadIdx.x; i < n; i += NTHREADS) {
        a += psum[i];
        k += pcnt[i];
    }
    s_sum[threadIdx.x] = a;
    s_cnt[threadIdx.x] = k;
    __syncthreads();
    for (int w = NTHREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s_sum[0];
        out[1] = (double)s_cnt[0];
        if (norm_out) *norm_out = s_cnt[0] ? sqrt(s_sum[0] / (double)s_cnt[0]) : -1.0;
    }
}

// enforce_bssn_constraints on interior points (standalone entry point)
__global__ __launch_bounds__(NTHREADS) void k_enforce(VarPtrs u, const DBlock *__restrict__ blocks,
                                                     const Tile *__restrict__ tiles, const int *__restrict__ wg_tile,
                                                     int tile_end, int64_t P, double chi_floor, int *err) {
    const int64_t q = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
    if (q >= P) return;
    int t = wg_tile[blockIdx.x];
    while (t + 1 < tile_end && q >= tiles[t + 1].q0) t++;
    const Tile T = tiles[t];
    const DBlock B = blocks[T.block];
    const int p = (int)(q - T.q0);
    const int i = T.o[0] + p % T.e[0];
    const int j = T.o[1] + (p / T.e[0]) % T.e[1];
    const int k = T.o[2] + p / (T.e[0] * T.e[1]);
    const uint64_t idx = B.offset + (uint64_t)i + (uint64_t)B.n[0] * ((uint64_t)j + (uint64_t)B.n[1] * k);
    double v14[14];
    v14[0] = u.p[0][idx];
    v14[1] = u.p[1][idx];
    for (int c = 0; c < 12; c++) v14[2 + c] = u.p[12 + c][idx];
    const int e = enforce_point(v14, chi_floor);
    if (e) atomicOr(err, e);
    u.p[0][idx] = v14[0];
    u.p[1][idx] = v14[1];
    for (int c = 0; c < 12; c++) u.p[12 + c][idx] = v14[2 + c];
}

// One stage combination of the three-stage SSP Runge-Kutta (rkBSSN.cpp:
// 691-762) at every interior point, then enforce_bssn_constraints, as the
// reference does after each stage: stage 0 saves u0 (interior-only buffer)
// and sets u = u0 + c k; stages 1 and 2 set u = (a u0 + b u) + c k, the
// products and sums in the reference's order. Pointwise, so u is updated in
// place.
__global__ __launch_bounds__(NTHREADS) void k_rk3_combine(VarPtrs u, VarPtrs k, VarPtrs u0,
                                                         const DBlock *__restrict__ blocks,
                                                         const Tile *__restrict__ tiles,
                                                         const int *__restrict__ wg_tile, int tile_end, int64_t P,
                                                         int stage, double a, double b, double c, double chi_floor,
                                                         int *err) {
#pragma clang fp contract(off)
    const int64_t qq = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
    if (qq >= P) return;
    int t = wg_tile[blockIdx.x];
    while (t + 1 < tile_end && qq >= tiles[t + 1].q0) t++;
    const Tile T = tiles[t];
    const DBlock B = blocks[T.block];
    const int p = (int)(qq - T.q0);
    const int i = T.o[0] + p % T.e[0];
    const int j = T.o[1] + (p / T.e[0]) % T.e[1];
    const int kk = T.o[2] + p / (T.e[0] * T.e[1]);
    const uint64_t idx = B.offset + (uint64_t)i + (uint64_t)B.n[0] * ((uint64_t)j + (uint64_t)B.n[1] * kk);
    const int mx = B.n[0] - 2 * PW, my = B.n[1] - 2 * PW;
    const int64_t q = B.q0 + (i - PW) + (int64_t)mx * ((j - PW) + (int64_t)my * (kk - PW));
    double out[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const double uv = u.p[v][idx], kv = k.p[v][idx];
        if (stage == 0) {
            u0.p[v][q] = uv;
            out[v] = uv + c * kv;
        } else {
            out[v] = (a * u0.p[v][q] + b * uv) + c * kv;
        }
    }
    double v14[14];
    v14[0] = out[0];
    v14[1] = out[1];
    for (int m = 0; m < 12; m++) v14[2 + m] = out[12 + m];
    const int e = enforce_point(v14, chi_floor);
    if (e) atomicOr(err, e);
    out[0] = v14[0];
    out[1] = v14[1];
    for (int m = 0; m < 12; m++) out[12 + m] = v14[2 + m];
#pragma unroll
    for (int v = 0; v < NV; v++) u.p[v][idx] = out[v];
}

// The tableau of the embedded Runge-Kutta 4(5) times the step size, per
// stage: a[m] = U[s+1][m] h for stages 0..4; c4 and c5 the weights of the
// fourth- and fifth-order solutions times h for stage 5 (rkBSSN.h:57-71, with
// the reference's values; each product formed in FP64 as the reference forms
// U h before multiplying by k).
struct Rk45Coef {
    double a[6];
    double c4[5];
    double c5[6];
};

// One stage of the embedded Runge-Kutta 4(5) (rkBSSN.cpp:906-1060) at every
// interior point: stage 0 saves u0 (interior-only); the stage's right-hand
// side r (padded) is kept in k_s (interior-only, stage-major); stages 0..4 set
// u = u0 + sum_m a[m] k_m and enforce the constraints; stage 5 forms y4 and
// y5 (stage 1 has weight 0 in both and is skipped, as there), enforces both,
// leaves y5 in u and takes max |y4 - y5| over the variables into *emax (the
// bits of a non-negative double, so unsigned max orders them). Sums in the
// reference's order with contraction off.
__global__ __launch_bounds__(NTHREADS) void k_rk45_stage(VarPtrs u, VarPtrs r, VarPtrs u0, double *__restrict__ kb,
                                                        int64_t ni, const DBlock *__restrict__ blocks,
                                                        const Tile *__restrict__ tiles,
                                                        const int *__restrict__ wg_tile, int tile_end, int64_t P,
                                                        int stage, Rk45Coef cf, double chi_floor, int *err,
                                                        unsigned long long *emax) {
#pragma clang fp contract(off)
    __shared__ double s_max[NTHREADS / 64];
    const int64_t qq = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
    double dmax = 0.0;
    if (qq < P) {
        int t = wg_tile[blockIdx.x];
        while (t + 1 < tile_end && qq >= tiles[t + 1].q0) t++;
        const Tile T = tiles[t];
        const DBlock B = blocks[T.block];
        const int p = (int)(qq - T.q0);
        const int i = T.o[0] + p % T.e[0];
        const int j = T.o[1] + (p / T.e[0]) % T.e[1];
        const int kk = T.o[2] + p / (T.e[0] * T.e[1]);
        const uint64_t idx = B.offset + (uint64_t)i + (uint64_t)B.n[0] * ((uint64_t)j + (uint64_t)B.n[1] * kk);
        const int mx = B.n[0] - 2 * PW, my = B.n[1] - 2 * PW;
        const int64_t q = B.q0 + (i - PW) + (int64_t)mx * ((j - PW) + (int64_t)my * (kk - PW));
        auto kv = [&](int m, int v) -> double & { return kb[((size_t)m * NV + v) * (size_t)ni + (size_t)q]; };
        double out[NV];
        for (int v = 0; v < NV; v++) {
            if (stage == 0) u0.p[v][q] = u.p[v][idx];
            kv(stage, v) = r.p[v][idx];
        }
        auto enforce = [&](double *o) {
            double v14[14];
            v14[0] = o[0];
            v14[1] = o[1];
            for (int m = 0; m < 12; m++) v14[2 + m] = o[12 + m];
            const int e = enforce_point(v14, chi_floor);
            if (e) atomicOr(err, e);
            o[0] = v14[0];
            o[1] = v14[1];
            for (int m = 0; m < 12; m++) o[12 + m] = v14[2 + m];
        };
        if (stage < 5) {
            for (int v = 0; v < NV; v++) {
                double a = u0.p[v][q];
                for (int m = 0; m <= stage; m++) a += cf.a[m] * kv(m, v);
                out[v] = a;
            }
            enforce(out);
        } else {
            double y4[NV];
            for (int v = 0; v < NV; v++) {
                double a = u0.p[v][q];
                for (int m = 0; m < 5; m++)
                    if (m != 1) a += cf.c4[m] * kv(m, v);
                y4[v] = a;
            }
            enforce(y4);
            for (int v = 0; v < NV; v++) {
                double a = u0.p[v][q];
                for (int m = 0; m < 6; m++)
                    if (m != 1) a += cf.c5[m] * kv(m, v);
                out[v] = a;
            }
            enforce(out);
            for (int v = 0; v < NV; v++) dmax = fmax(dmax, fabs(y4[v] - out[v]));
        }
        for (int v = 0; v < NV; v++) u.p[v][idx] = out[v];
    }
    if (stage == 5) {
        // workgroup maximum, then one atomic per workgroup
        for (int o = 32; o > 0; o >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, o));
        if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = dmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = s_max[0];
            for (int w = 1; w < NTHREADS / 64; w++) m = fmax(m, s_max[w]);
            atomicMax(emax, (unsigned long long)__double_as_longlong(m));
        }
    }
}

// A rejected try of the embedded Runge-Kutta 4(5): u's interior back to u0.
__global__ __launch_bounds__(NTHREADS) void k_rk45_restore(VarPtrs u, VarPtrs u0, const DBlock *__restrict__ blocks,
                                                          const Tile *__restrict__ tiles,
                                                          const int *__restrict__ wg_tile, int tile_end, int64_t P) {
    const int64_t qq = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
    if (qq >= P) return;
    int t = wg_tile[blockIdx.x];
    while (t + 1 < tile_end && qq >= tiles[t + 1].q0) t++;
    const Tile T = tiles[t];
    const DBlock B = blocks[T.block];
    const int p = (int)(qq - T.q0);
    const int i = T.o[0] + p % T.e[0];
    const int j = T.o[1] + (p / T.e[0]) % T.e[1];
    const int kk = T.o[2] + p / (T.e[0] * T.e[1]);
    const uint64_t idx = B.offset + (uint64_t)i + (uint64_t)B.n[0] * ((uint64_t)j + (uint64_t)B.n[1] * kk);
    const int mx = B.n[0] - 2 * PW, my = B.n[1] - 2 * PW;
    const int64_t q = B.q0 + (i - PW) + (int64_t)mx * ((j - PW) + (int64_t)my * (kk - PW));
    for (int v = 0; v < NV; v++) u.p[v][idx] = u0.p[v][q];
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

// The slow-start lapse's time in a captured RK4 step: at the step's start the
// factor exp(-t^2/800) of the time d[0] goes to d[1]; at its end the time moves
// on by dt (rkBSSN.cpp:1087).
__global__ void k_ssl_begin(double *d) { d[1] = exp(-1.0 / 800.0 * (d[0] * d[0])); }
__global__ void k_ssl_advance(double *d, double dt) { d[0] += dt; }

// The range of one variable over every interior point and the number of
// non-finite values there (the reference's check after each step,
// BSSNCtx::terminal_output, bssnCtx.cpp:1240-1262, takes the min and max of
// alpha over the nodes and stops on a NaN). Finite values only enter the
// range; they are reduced as order-preserving 64-bit keys (sign bit set for
// non-negative numbers, all bits flipped for negative ones) with one atomic
// per workgroup. r[0] = min key, r[1] = max key, r[2] = non-finite count.
__device__ __forceinline__ unsigned long long order_key(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__global__ __launch_bounds__(NTHREADS) void k_var_range(const double *__restrict__ v,
                                                       const DBlock *__restrict__ blocks,
                                                       const Tile *__restrict__ tiles,
                                                       const int *__restrict__ wg_tile, int tile_end, int64_t P,
                                                       unsigned long long *r) {
    __shared__ unsigned long long s_lo[NTHREADS / 64], s_hi[NTHREADS / 64], s_bad[NTHREADS / 64];
    const int64_t qq = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
    unsigned long long lo = ~0ull, hi = 0ull, bad = 0ull;
    if (qq < P) {
        int t = wg_tile[blockIdx.x];
        while (t + 1 < tile_end && qq >= tiles[t + 1].q0) t++;
        const Tile T = tiles[t];
        const DBlock B = blocks[T.block];
        const int p = (int)(qq - T.q0);
        const int i = T.o[0] + p % T.e[0];
        const int j = T.o[1] + (p / T.e[0]) % T.e[1];
        const int kk = T.o[2] + p / (T.e[0] * T.e[1]);
        const double x = v[B.offset + (uint64_t)i + (uint64_t)B.n[0] * ((uint64_t)j + (uint64_t)B.n[1] * kk)];
        if (isfinite(x)) {
            lo = hi = order_key(x);
        } else {
            bad = 1;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, (unsigned long long)__shfl_xor(lo, o));
        hi = max(hi, (unsigned long long)__shfl_xor(hi, o));
        bad += (unsigned long long)__shfl_xor(bad, o);
    }
    if ((threadIdx.x & 63) == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
        s_bad[threadIdx.x >> 6] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NTHREADS / 64; w++) {
            lo = min(lo, s_lo[w]);
            hi = max(hi, s_hi[w]);
            bad += s_bad[w];
        }
        atomicMin(&r[0], lo);
        atomicMax(&r[1], hi);
        if (bad) atomicAdd(&r[2], bad);
    }
}

__global__ void k_var_range_init(unsigned long long *r) {
    r[0] = ~0ull;
    r[1] = 0ull;
    r[2] = 0ull;
}

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_OK(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail(DGR_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

#ifdef DGR_F32_KERNELS
// The library compiles this source twice. With DGR_F32_KERNELS the unit holds
// only the FP32 fused kernels and this table of their addresses: they are
// built with the compiler's default instruction scheduler, while the main
// unit (everything else, FP64 kernels included) is built with the iterative
// ILP scheduler, which is 3% faster on the FP64 kernels and spills the FP32
// ones (measured).
extern "C" __attribute__((visibility("hidden"))) void dgr_internal_fused_f32_kernels(const void **k) {
    k[0] = reinterpret_cast<const void *>(&k_fused<false, CfgOf<float>::Wide, float>);
    k[1] = reinterpret_cast<const void *>(&k_fused<false, CfgOf<float>::Thin, float>);
    k[2] = reinterpret_cast<const void *>(&k_fused<false, CfgFull, float>);
    k[3] = reinterpret_cast<const void *>(&k_fused<true, CfgOf<float>::Wide, float>);
    k[4] = reinterpret_cast<const void *>(&k_fused<false, CfgOf<float>::Wide, float, true>);
    k[5] = reinterpret_cast<const void *>(&k_fused<false, CfgOf<float>::Thin, float, true>);
    k[6] = reinterpret_cast<const void *>(&k_fused<false, CfgFull, float, true>);
    k[7] = reinterpret_cast<const void *>(&k_fused<true, CfgOf<float>::Wide, float, true>);
}
#else
extern "C" void dgr_internal_fused_f32_kernels(const void **k);

namespace {
// Device scratch of dgr_constraint_l2_norm_noalloc: the norm's per-workgroup
// partial sums (NORM_SCRATCH_WG doubles), counts (as many 64-bit integers)
// and the final sum and count. Each GPU holds one copy until the process
// ends (16 MiB), so that entry point allocates nothing; every call on a GPU
// shares it. In host code `scratch` is only a stand-in for the symbol: kernels
// get the device address that scratch_address() looked up for the device.
constexpr int NORM_SCRATCH_WG = DGR_NOALLOC_NORM_MAX_WORKGROUPS;
static __device__ __align__(256) double scratch[2 * (size_t)NORM_SCRATCH_WG + 2];

std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_ptr;  // device -> address of its `scratch`

// The device address of `scratch` on `device`, looked up once per device
// with that device current (dgr_plan_create calls this after hipSetDevice).
hipError_t scratch_address(int device, double **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    double *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    *out = ptr;
    return hipSuccess;
}
}  // namespace


struct dgr_plan {
    int device = 0;
    uint64_t szpdof = 0;
    int64_t interior = 0;
    std::vector<Chunk> chunks;
    DBlock *d_blocks = nullptr;
    Tile *d_tiles = nullptr;
    int *d_wg_tile = nullptr;
    int *d_face_tiles = nullptr;
    double *d_W = nullptr;
    int64_t max_chunk_points = 0;
    double *d_stage = nullptr;  // split path RK4: 2 * 24 * szpdof (stage state, accumulator)
    bool fused = true;          // k_fused (default) or the split k_derivs/k_rhs path (DGR_KERNEL=split)
    int precision = DGR_PRECISION_FP64;  // arithmetic of the fused kernel
    // k_fused tiles per precision (0 FP64, 1 mixed), XCD-grouped, group k in
    // [fgroup[p][k], fgroup[p][k+1]), each interior block tiled by the scheme
    // tiling[p] holds for its size (else the default of tiling_of)
    Tile *d_ftiles[2] = {};
    int fgroup[2][NFGROUP + 1] = {};
    std::map<std::array<int, 3>, int> tiling[2];
    std::vector<DBlock> hblocks;                  // host copy of the block table
    std::map<std::array<int, 4>, int> lane_off;   // lane table offset per (ex, ey, ez, threads)
    dgr::AxisConsts *d_bconst = nullptr;  // k_fused: x, y, z stencil constants per block
    uint16_t *d_tabs = nullptr;           // k_fused: lane tables per tile shape, slot tables per config
    int dtab_wide = 0, dtab_thin = 0, dtab_full = 0;  // offsets of the configurations' slot tables
    std::vector<int> hnbr;      // same-level neighbour table, 27 per block (built on first use)
    int *d_nbr = nullptr;       // its device copy
    int64_t *d_nbo = nullptr;   // the same as neighbour block offsets (-1 for none), for k_fused
    int64_t nlinks = -1;        // links in the table (-1: not built yet)
    int2 *d_halo_wg = nullptr;  // k_halo_fill's workgroups: (block, chunk) of the linked blocks
    int64_t nhalo_wg = 0;
    bool halo = false;          // dgr_rk4_step(s): refresh padding from neighbours before each stage
    bool fill_general = false;  // ... and from blocks of other sizes or levels (k_fill_copy, k_fill_interp)
    FillRec *d_fill = nullptr;  // general fill records (built on first use)
    int64_t nfill = -1;         // records (-1: not built yet)
    int64_t nfill_interp = 0;   // of which interpolations (they follow the copies)
    FillGroup *d_fill_groups = nullptr;  // the interpolations grouped by coarse element
    int64_t nfill_groups = 0;
    double *d_istage = nullptr; // k_fused RK4: 2 * 24 * interior (u0 copy, accumulator)
    double *d_rk45 = nullptr;   // dgr_rk45_step: the six stages' right-hand sides, 6 * 24 * interior
    unsigned long long *d_rk45_emax = nullptr;  // dgr_rk45_step: the error estimate's bits
    double *d_ssl = nullptr;    // dgr_rk4_steps with the slow-start lapse: {time, exp(-time^2/800)}
    unsigned long long *d_range = nullptr;      // dgr_variable_range: min key, max key, non-finite count
    double *d_wstate = nullptr; // k_fused RK4: 24 * szpdof, the padded work state W
    int *d_err = nullptr;
    // dgr_rk4_steps: one RK4 step captured as a graph, valid for the arguments in rk4_key
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t rk4_exec = nullptr;
    struct {
        double *u[NV];
        double dt;
        dgr_bssn_params prm;
        int enforce;
        bool halo;
        bool fill_general;
    } rk4_key{};
    int nwg_total = 0;                    // k_rhs-style workgroups over all chunks
    double *d_norm = nullptr;             // norm partials: nwg_total sums, nwg_total counts, 2 results
    double *d_scratch = nullptr;          // this device's `scratch` (dgr_constraint_l2_norm_noalloc)
    // optional per-kernel timing: event pairs recorded around each launch
    bool timing = false;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[DGR_TIMING_KINDS];

    hipEvent_t take_event() {
        if (event_pool.empty()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            return e;
        }
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
};

namespace {

// Lane -> point table for a tile of ex x ey x ez points on ft lanes. Lane
// 32g + r takes a point whose LDS address i + 13 j + 169 k is congruent to r
// mod 32, so the 32 lanes of a group touch 32 different bank pairs at every
// stencil tap (8-byte words: conflict-free both for 32 banks serving 16 lanes
// and 64 banks serving 32 lanes per cycle). Points beyond ft/32 in a residue
// class take free lanes; lanes still free repeat a point of their own residue
// class (same address, so no conflict) and store nothing.
void lane_table(int ex, int ey, int ez, int ft, std::vector<uint16_t> &out) {
    const int n = ex * ey * ez, groups = ft / 32;
    std::vector<std::vector<int>> res(32);
    for (int p = 0; p < n; p++) {
        const int li = p % ex, lj = (p / ex) % ey, lk = p / (ex * ey);
        res[(li + FRX * lj + FRXY * lk) % 32].push_back(p);
    }
    std::vector<int> lane(ft, -1), over;
    for (int r = 0; r < 32; r++)
        for (int g = 0; g < (int)res[r].size(); g++) {
            if (g < groups) lane[32 * g + r] = res[r][g];
            else over.push_back(res[r][g]);
        }
    size_t o = 0;
    for (int l = 0; l < ft && o < over.size(); l++)
        if (lane[l] < 0) lane[l] = over[o++];
    for (int l = 0; l < ft; l++) {
        if (lane[l] >= 0) {
            out.push_back((uint16_t)lane[l]);
        } else {
            const std::vector<int> &rr = res[l % 32];
            out.push_back((uint16_t)((rr.empty() ? 0 : rr[0]) | LANE_SHADOW));
        }
    }
}

// Slot table of a tile configuration: slot (it, tid) -> the position
// w = i0 + 13 j + 169 k it evaluates and writes, first D_x for DXIT passes,
// then D_y for DYIT passes, grouped like lane_table; unused slots get w = 7
// (a column no stencil reads). D_x: i0 < 7 and either a tile plane
// (PW <= k < PW+FZ, any j < 13) or a tile row (PW <= j < PW+7, any k);
// D_y: i0 < 7, j < 7, any k. (k < FZ+6 throughout; the tile sits at j, k >= PW.)
template <class C>
bool dtable(std::vector<uint16_t> &out) {
    bool ok = true;
    auto fill = [&](int nj, int nit, bool dx) {
        std::vector<std::vector<int>> res(32);
        int n = 0;
        for (int k = 0; k < C::FZ + 2 * PW; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < TMAX; i++) {
                    const bool plane = k >= PW && k < PW + C::FZ, row = j >= PW && j < PW + TMAX;
                    if (dx && !plane && !row) continue;
                    const int w = i + FRX * j + FRXY * k;
                    res[w % 32].push_back(w);
                    n++;
                }
        ok = ok && n == (dx ? C::NDX : C::NDY) && n <= nit * C::FT;
        const int nslot = nit * C::FT, groups = nslot / 32;
        std::vector<int> slot(nslot, -1), over;
        for (int r = 0; r < 32; r++)
            for (int g = 0; g < (int)res[r].size(); g++) {
                if (g < groups) slot[32 * g + r] = res[r][g];
                else over.push_back(res[r][g]);
            }
        size_t o = 0;
        for (int q = 0; q < nslot && o < over.size(); q++)
            if (slot[q] < 0) slot[q] = over[o++];
        for (int q = 0; q < nslot; q++) out.push_back((uint16_t)(slot[q] < 0 ? TMAX : slot[q]));
    };
    fill(RMAX, C::DXIT, true);
    fill(TMAX, C::DYIT, false);
    return ok;
}

// m interior points -> ceil(m/emax) near-equal parts of at most emax points
void split_axis_max(int m, int emax, std::vector<std::pair<int, int>> &parts) {
    const int np = (m + emax - 1) / emax;
    int o = PW;
    for (int p = 0; p < np; p++) {
        const int e = m / np + (p < m % np ? 1 : 0);
        parts.push_back({o, e});
        o += e;
    }
}

int thread_count(int group) {
    return group == FG_THIN ? CfgThin::FT : group == FG_FULL ? CfgFull::FT : CfgWide::FT;
}

// Appends the tiles of block b under scheme `sch` to per[group][slot]. Blocks
// with flagged faces are cut along z into near-equal parts of at most 5 (each
// >= 4 points, so the one-sided closures, which reach 5 points along their
// axis, stay inside the staged region), whatever the scheme.
void block_tiles(const DBlock &d, uint32_t b, int sch, std::vector<std::vector<Tile>> *per, size_t slot) {
    std::vector<std::pair<int, int>> px, py, pz;
    split_axis_max(d.n[0] - 6, TMAX, px);
    split_axis_max(d.n[1] - 6, TMAX, py);
    if (d.bflag || sch == SCH_W) {
        split_axis_max(d.n[2] - 6, CfgWide::FZ, pz);
    } else if (sch == SCH_FULL) {
        split_axis_max(d.n[2] - 6, CfgFull::FZ, pz);
    } else {
        for (int o = PW, m = d.n[2] - 6; m > 0;) {
            const int e = std::min(m, CfgWide::FZ);
            pz.push_back({o, e});
            o += e;
            m -= e;
        }
    }
    for (auto &z : pz)
        for (auto &y : py)
            for (auto &x : px) {
                Tile t{};
                t.block = (int)b;
                t.q0 = 0;
                t.o[0] = x.first; t.o[1] = y.first; t.o[2] = z.first;
                t.e[0] = x.second; t.e[1] = y.second; t.e[2] = z.second;
                const int g = d.bflag                                    ? FG_FACE
                              : sch == SCH_FULL                          ? FG_FULL
                              : sch == SCH_WT && t.e[2] <= CfgThin::FZ ? FG_THIN
                                                                         : FG_WIDE;
                per[g][slot].push_back(t);
            }
}

// Concatenates the groups; within a group the blocks (slots) are taken eight
// at a time and their tiles interleaved, so that with workgroups dealt
// round-robin over the 8 XCDs the tiles of one block share an XCD and its L2
// for their overlapping halos.
void group_tiles(std::vector<std::vector<Tile>> *per, std::vector<Tile> &out, int *fgroup) {
    for (int g = 0; g < NFGROUP; g++) {
        fgroup[g] = (int)out.size();
        std::vector<size_t> bl;
        for (size_t b = 0; b < per[g].size(); b++)
            if (!per[g][b].empty()) bl.push_back(b);
        // blocks with the same number of tiles together, most tiles first: a
        // group of 8 then interleaves its tiles to the end, so every tile of
        // a block lands on the same XCD (workgroups are dealt to the 8 XCDs
        // round-robin) and its neighbours' halos are in that XCD's L2; in a
        // group of mixed sizes the large block's remaining tiles would be
        // dealt over all 8 XCDs
        std::stable_sort(bl.begin(), bl.end(),
                         [&](size_t a, size_t b) { return per[g][a].size() > per[g][b].size(); });
        for (size_t i0 = 0; i0 < bl.size(); i0 += 8) {
            const size_t i1 = std::min(bl.size(), i0 + 8);
            size_t mx = 0;
            for (size_t i = i0; i < i1; i++) mx = std::max(mx, per[g][bl[i]].size());
            for (size_t t = 0; t < mx; t++)
                for (size_t i = i0; i < i1; i++)
                    if (t < per[g][bl[i]].size()) out.push_back(per[g][bl[i]][t]);
        }
    }
    fgroup[NFGROUP] = (int)out.size();
}

// Tiling of an interior block for precision p (0 FP64, 1 mixed): the tuned
// scheme of its size, else SCH_WT in FP64 and, in FP32, SCH_FULL for blocks
// one 7-point tile deep (13^3) and SCH_WT for the others (measured defaults).
int tiling_of(const dgr_plan *plan, int p, const DBlock &d) {
    auto it = plan->tiling[p].find({d.n[0], d.n[1], d.n[2]});
    if (it != plan->tiling[p].end()) return it->second;
    if (p == 1 && d.n[2] - 6 <= CfgFull::FZ) return SCH_FULL;
    return SCH_WT;
}

// Sets lanetab of every tile from the plan's lane tables (all shapes exist).
bool assign_lane_tables(const dgr_plan *plan, std::vector<Tile> &tiles, const int *fgroup) {
    for (int g = 0; g < NFGROUP; g++)
        for (int i = fgroup[g]; i < fgroup[g + 1]; i++) {
            Tile &t = tiles[i];
            auto it = plan->lane_off.find({t.e[0], t.e[1], t.e[2], thread_count(g)});
            if (it == plan->lane_off.end()) return false;
            t.lanetab = it->second;
        }
    return true;
}

// (Re)builds and uploads the tile list of precision p.
hipError_t build_ftiles(dgr_plan *plan, int p) {
    const size_t nb = plan->hblocks.size();
    std::vector<std::vector<Tile>> per[NFGROUP];
    for (auto &v : per) v.resize(nb);
    for (size_t b = 0; b < nb; b++) {
        const DBlock &d = plan->hblocks[b];
        block_tiles(d, (uint32_t)b, d.bflag ? SCH_WT : tiling_of(plan, p, d), per, b);
    }
    std::vector<Tile> tiles;
    int fg[NFGROUP + 1];
    group_tiles(per, tiles, fg);
    if (!assign_lane_tables(plan, tiles, fg)) return hipErrorInvalidValue;
    Tile *d = nullptr;
    hipError_t e = hipSuccess;
    if (!tiles.empty()) {
        e = hipMalloc(&d, sizeof(Tile) * tiles.size());
        if (e == hipSuccess) e = hipMemcpy(d, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            return e;
        }
    }
    if (plan->d_ftiles[p]) (void)hipFree(plan->d_ftiles[p]);
    plan->d_ftiles[p] = d;
    std::copy(fg, fg + NFGROUP + 1, plan->fgroup[p]);
    return hipSuccess;
}

int link_neighbors(dgr_plan *plan);
int build_fill(dgr_plan *plan);

}  // namespace

extern "C" {

void dgr_default_params(dgr_bssn_params *p) {
    std::memset(p, 0, sizeof(*p));
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
    p->eta_func = 0;  // the reference's default build (BSSN_ETA_FUNCTION off)
    p->eta_r0 = 1.31;  // parameters.cpp:59-60
    p->eta_power[0] = 2.0;
    p->eta_power[1] = 2.0;
    p->rochester = 0;  // the reference's default build: standard gauge, no slow-start lapse
    p->ssl = 0;
    p->xi[0] = p->xi[1] = p->xi[2] = 0.0;  // parameters.cpp:157
    p->time = 0.0;
}

const char *dgr_last_error(void) { return g_err.c_str(); }


static void split_axis(int m, std::vector<std::pair<int, int>> &parts) {
    // m interior points -> ceil(m/7) near-equal parts (each >= 4 when m >= 7)
    const int np = (m + TMAX - 1) / TMAX;
    int o = PW;
    for (int p = 0; p < np; p++) {
        const int e = m / np + (p < m % np ? 1 : 0);
        parts.push_back({o, e});
        o += e;
    }
}

int dgr_plan_create(int device, const dgr_block *blocks, uint32_t nblocks, uint64_t szpdof, dgr_plan **out) {
    g_err.clear();
    if (!out || (!blocks && nblocks)) return fail(DGR_E_INVALID, "null argument");
    *out = nullptr;
    // the fused kernel's in-kernel padding refresh indexes a variable with 32 bits
    if (szpdof >= (1ull << 32)) return fail(DGR_E_INVALID, "szpdof must be below 2^32 points per variable");
    std::vector<DBlock> hb(nblocks);
    for (uint32_t b = 0; b < nblocks; b++) {
        const dgr_block &s = blocks[b];
        uint64_t n3 = 1;
        for (int a = 0; a < 3; a++) {
            if (s.sz[a] < 13 || s.sz[a] > 4096)
                return fail(DGR_E_INVALID, "block " + std::to_string(b) + ": size must be in [13, 4096] per axis");
            n3 *= s.sz[a];
        }
        if (s.offset + n3 > szpdof || s.offset + n3 < s.offset)
            return fail(DGR_E_INVALID, "block " + std::to_string(b) + " extends past szpdof");
        if (s.bflag >> 6) return fail(DGR_E_INVALID, "block " + std::to_string(b) + ": unknown face bits");
        DBlock &d = hb[b];
        d.offset = s.offset;
        d.bflag = s.bflag;
        for (int a = 0; a < 3; a++) {
            d.n[a] = (int)s.sz[a];
            d.pmin[a] = s.pmin[a];
            d.h[a] = (s.pmax[a] - s.pmin[a]) / (s.sz[a] - 1);  // rhs.cpp:144-146
            if (!(d.h[a] > 0.0) || !std::isfinite(d.h[a]))
                return fail(DGR_E_INVALID, "block " + std::to_string(b) + ": pmax must exceed pmin");
        }
    }

    int64_t max_chunk = 131072;
    if (const char *s = std::getenv("DGR_CHUNK_POINTS")) max_chunk = std::max<int64_t>(4096, std::atoll(s));

    auto *plan = new dgr_plan();
    plan->device = device;
    plan->szpdof = szpdof;
    std::vector<Tile> tiles;
    std::vector<int> wg_tile, face_tiles;
    int64_t chunk_pts = 0;
    Chunk cur{0, 0, 0, 0, 0, 0, 0};
    auto close_chunk = [&]() {
        if (cur.tile_end == cur.tile_begin) return;
        cur.npoints = chunk_pts;
        cur.wg_begin = (int)wg_tile.size();
        cur.nwg = (int)((chunk_pts + NTHREADS - 1) / NTHREADS);
        int t = cur.tile_begin;
        for (int w = 0; w < cur.nwg; w++) {
            const int64_t q = (int64_t)w * NTHREADS;
            while (t + 1 < cur.tile_end && q >= tiles[t + 1].q0) t++;
            wg_tile.push_back(t);
        }
        cur.face_begin = (int)face_tiles.size();
        for (int tt = cur.tile_begin; tt < cur.tile_end; tt++)
            if (hb[tiles[tt].block].bflag) face_tiles.push_back(tt);
        cur.nface = (int)face_tiles.size() - cur.face_begin;
        plan->max_chunk_points = std::max(plan->max_chunk_points, chunk_pts);
        plan->chunks.push_back(cur);
        cur = Chunk{(int)tiles.size(), (int)tiles.size(), 0, 0, 0, 0, 0};
        chunk_pts = 0;
    };
    for (uint32_t b = 0; b < nblocks; b++) {
        const DBlock &d = hb[b];
        const int64_t pts = (int64_t)(d.n[0] - 6) * (d.n[1] - 6) * (d.n[2] - 6);
        if (pts > (int64_t(1) << 24)) {  // keeps chunk-local point indices in 32 bits
            delete plan;
            return fail(DGR_E_INVALID, "block " + std::to_string(b) + ": more than 2^24 interior points");
        }
        if (chunk_pts > 0 && chunk_pts + pts > max_chunk) close_chunk();
        std::vector<std::pair<int, int>> px, py, pz;
        split_axis(d.n[0] - 6, px);
        split_axis(d.n[1] - 6, py);
        split_axis(d.n[2] - 6, pz);
        for (auto &z : pz)
            for (auto &y : py)
                for (auto &x : px) {
                    Tile t{};
                    t.block = (int)b;
                    t.q0 = (int)chunk_pts;
                    t.o[0] = x.first; t.o[1] = y.first; t.o[2] = z.first;
                    t.e[0] = x.second; t.e[1] = y.second; t.e[2] = z.second;
                    tiles.push_back(t);
                    chunk_pts += (int64_t)t.e[0] * t.e[1] * t.e[2];
                }
        cur.tile_end = (int)tiles.size();
        hb[b].q0 = plan->interior;
        plan->interior += pts;
    }
    close_chunk();
    plan->nwg_total = (int)wg_tile.size();

    // k_fused: slot tables per configuration, lane tables for every tile
    // shape any tiling can produce (one block per size and kind suffices);
    // the tile lists themselves are built per precision after the uploads
    plan->hblocks = hb;
    std::vector<uint16_t> tabs;
    plan->dtab_wide = (int)tabs.size();
    bool tabs_ok = dtable<CfgWide>(tabs);
    plan->dtab_thin = (int)tabs.size();
    tabs_ok = dtable<CfgThin>(tabs) && tabs_ok;
    plan->dtab_full = (int)tabs.size();
    tabs_ok = dtable<CfgFull>(tabs) && tabs_ok;
    if (!tabs_ok) {
        dgr_plan_destroy(plan);
        return fail(DGR_E_INVALID, "internal error: D_x / D_y slot table does not match its configuration");
    }
    {
        std::map<std::array<int, 4>, bool> seen;
        for (uint32_t b = 0; b < nblocks; b++) {
            const DBlock &d = hb[b];
            if (!seen.emplace(std::array<int, 4>{d.n[0], d.n[1], d.n[2], d.bflag ? 1 : 0}, true).second)
                continue;
            for (int sch = 0; sch < (d.bflag ? 1 : NSCHEME); sch++) {
                std::vector<std::vector<Tile>> per[NFGROUP];
                for (auto &v : per) v.resize(1);
                block_tiles(d, b, sch, per, 0);
                for (int g = 0; g < NFGROUP; g++)
                    for (const Tile &t : per[g][0]) {
                        const std::array<int, 4> key{t.e[0], t.e[1], t.e[2], thread_count(g)};
                        if (plan->lane_off.count(key)) continue;
                        plan->lane_off[key] = (int)tabs.size();
                        lane_table(t.e[0], t.e[1], t.e[2], thread_count(g), tabs);
                    }
            }
        }
    }
    std::vector<dgr::AxisConsts> bconst(3 * (size_t)nblocks);
    for (uint32_t b = 0; b < nblocks; b++)
        for (int a = 0; a < 3; a++) bconst[3 * b + a] = dgr::axis_consts(hb[b].h[a]);
    if (const char *k = std::getenv("DGR_KERNEL")) plan->fused = std::string(k) != "split";

    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = scratch_address(device, &plan->d_scratch);
    if (e == hipSuccess && nblocks) e = hipMalloc(&plan->d_blocks, sizeof(DBlock) * hb.size());
    if (e == hipSuccess && !tiles.empty()) e = hipMalloc(&plan->d_tiles, sizeof(Tile) * tiles.size());
    if (e == hipSuccess && !wg_tile.empty()) e = hipMalloc(&plan->d_wg_tile, sizeof(int) * wg_tile.size());
    if (e == hipSuccess && !face_tiles.empty()) e = hipMalloc(&plan->d_face_tiles, sizeof(int) * face_tiles.size());
    if (e == hipSuccess) e = hipMalloc(&plan->d_tabs, sizeof(uint16_t) * tabs.size());
    if (e == hipSuccess) e = hipMemcpy(plan->d_tabs, tabs.data(), sizeof(uint16_t) * tabs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && nblocks) e = hipMalloc(&plan->d_bconst, sizeof(dgr::AxisConsts) * bconst.size());
    if (e == hipSuccess && nblocks)
        e = hipMemcpy(plan->d_bconst, bconst.data(), sizeof(dgr::AxisConsts) * bconst.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&plan->d_err, sizeof(int));
    if (e == hipSuccess && nblocks) e = hipMemcpy(plan->d_blocks, hb.data(), sizeof(DBlock) * hb.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !tiles.empty()) e = hipMemcpy(plan->d_tiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !wg_tile.empty()) e = hipMemcpy(plan->d_wg_tile, wg_tile.data(), sizeof(int) * wg_tile.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !face_tiles.empty())
        e = hipMemcpy(plan->d_face_tiles, face_tiles.data(), sizeof(int) * face_tiles.size(), hipMemcpyHostToDevice);
    for (int p = 0; p < 2 && e == hipSuccess; p++) e = build_ftiles(plan, p);
    if (e == hipSuccess) e = hipMemset(plan->d_err, 0, sizeof(int));
    if (e != hipSuccess) {
        dgr_plan_destroy(plan);
        return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, std::string("plan setup: ") + hipGetErrorString(e));
    }
    *out = plan;
    return DGR_OK;
}

int dgr_plan_destroy(dgr_plan *plan) {
    if (!plan) return DGR_OK;
    // best-effort release: errors here have no one to report to
    (void)hipSetDevice(plan->device);
    for (auto &v : plan->pending)
        for (auto &pr : v) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (hipEvent_t e : plan->event_pool) (void)hipEventDestroy(e);
    (void)hipFree(plan->d_blocks);
    (void)hipFree(plan->d_tiles);
    (void)hipFree(plan->d_wg_tile);
    (void)hipFree(plan->d_face_tiles);
    (void)hipFree(plan->d_W);
    (void)hipFree(plan->d_stage);
    (void)hipFree(plan->d_ftiles[0]);
    (void)hipFree(plan->d_ftiles[1]);
    (void)hipFree(plan->d_bconst);
    (void)hipFree(plan->d_tabs);
    (void)hipFree(plan->d_istage);
    (void)hipFree(plan->d_rk45);
    (void)hipFree(plan->d_rk45_emax);
    (void)hipFree(plan->d_ssl);
    (void)hipFree(plan->d_range);
    (void)hipFree(plan->d_wstate);
    (void)hipFree(plan->d_nbr);
    (void)hipFree(plan->d_nbo);
    (void)hipFree(plan->d_fill);
    (void)hipFree(plan->d_fill_groups);
    (void)hipFree(plan->d_halo_wg);
    (void)hipFree(plan->d_err);
    (void)hipFree(plan->d_norm);
    if (plan->rk4_exec) (void)hipGraphExecDestroy(plan->rk4_exec);
    if (plan->cap_stream) (void)hipStreamDestroy(plan->cap_stream);
    delete plan;
    return DGR_OK;
}

int64_t dgr_plan_interior_points(const dgr_plan *plan) { return plan ? plan->interior : -1; }

int dgr_plan_move_block(dgr_plan *plan, uint32_t b, uint64_t offset, const double *pmin, const double *pmax) {
    g_err.clear();
    if (!plan || !pmin || !pmax) return fail(DGR_E_INVALID, "null argument");
    if (b >= plan->hblocks.size()) return fail(DGR_E_INVALID, "block index out of range");
    DBlock d = plan->hblocks[b];
    const uint64_t n3 = (uint64_t)d.n[0] * d.n[1] * d.n[2];
    if (offset + n3 > plan->szpdof || offset + n3 < offset)
        return fail(DGR_E_INVALID, "block " + std::to_string(b) + " extends past szpdof");
    d.offset = offset;
    for (int a = 0; a < 3; a++) {
        d.pmin[a] = pmin[a];
        d.h[a] = (pmax[a] - pmin[a]) / (d.n[a] - 1);  // rhs.cpp:144-146
        if (!(d.h[a] > 0.0) || !std::isfinite(d.h[a]))
            return fail(DGR_E_INVALID, "block " + std::to_string(b) + ": pmax must exceed pmin");
    }
    dgr::AxisConsts k[3];
    for (int a = 0; a < 3; a++) k[a] = dgr::axis_consts(d.h[a]);
    HIP_OK(hipSetDevice(plan->device));
    // earlier launches may still read the tables
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(plan->d_blocks + b, &d, sizeof(DBlock), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(plan->d_bconst + 3 * (size_t)b, k, sizeof(k), hipMemcpyHostToDevice));
    plan->hblocks[b] = d;
    if (plan->nlinks >= 0) {  // the neighbour table follows the geometry
        int rc = link_neighbors(plan);
        if (rc) return rc;
        if (plan->fill_general && (rc = build_fill(plan))) return rc;
    }
    return DGR_OK;
}

int dgr_plan_set_precision(dgr_plan *plan, int precision) {
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    if (precision != DGR_PRECISION_FP64 && precision != DGR_PRECISION_MIXED)
        return fail(DGR_E_INVALID, "unknown precision");
    if (precision != DGR_PRECISION_FP64 && !plan->fused)
        return fail(DGR_E_INVALID, "mixed precision needs the fused kernel (DGR_KERNEL=split is set)");
    if (precision != plan->precision && plan->rk4_exec) {  // the captured step used the other kernels
        HIP_OK(hipSetDevice(plan->device));
        HIP_OK(hipGraphExecDestroy(plan->rk4_exec));
        plan->rk4_exec = nullptr;
    }
    plan->precision = precision;
    return DGR_OK;
}

int dgr_plan_set_timing(dgr_plan *plan, int enable) {
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    plan->timing = enable != 0;
    return DGR_OK;
}

int dgr_plan_kernel_times(dgr_plan *plan, double *ms, int64_t *launches) {
    if (!plan || !ms || !launches) return fail(DGR_E_INVALID, "null argument");
    HIP_OK(hipSetDevice(plan->device));
    for (int k = 0; k < DGR_TIMING_KINDS; k++) {
        ms[k] = 0.0;
        launches[k] = (int64_t)plan->pending[k].size();
        for (auto &pr : plan->pending[k]) {
            HIP_OK(hipEventSynchronize(pr.second));
            float t = 0.f;
            HIP_OK(hipEventElapsedTime(&t, pr.first, pr.second));
            ms[k] += t;
            plan->event_pool.push_back(pr.first);
            plan->event_pool.push_back(pr.second);
        }
        plan->pending[k].clear();
    }
    return DGR_OK;
}

}  // extern "C"

namespace {

// t: the time the slow-start lapse term reads
RhsParams to_params(const dgr_bssn_params *p, double t) {
    RhsParams r;
    for (int i = 0; i < 4; i++) r.lam[i] = p->lambda[i];
    r.lamf[0] = p->lambda_f[0];
    r.lamf[1] = p->lambda_f[1];
    r.eta_central = p->eta_central;
    r.eta_outer = p->eta_outer;
    r.eta_width = p->eta_width;
    r.ko_sigma = p->ko_sigma;
    r.chi_floor = p->chi_floor;
    r.cako = p->cako_enabled;
    r.eps_gauge = p->eps_cako_gauge;
    r.eps_other = p->eps_cako_other;
    r.eta_func = p->eta_func != 0;
    r.eta_r0 = p->eta_r0;
    r.eta_p0 = p->eta_power[0];
    r.eta_p1 = p->eta_power[1];
    r.rochester = p->rochester != 0;
    r.ssl = p->ssl != 0;
    for (int i = 0; i < 3; i++) r.xi[i] = p->xi[i];
    r.ssl_e = r.ssl ? std::exp(-1.0 / 800.0 * (t * t)) : 0.0;
    r.ssl_dev = nullptr;
    return r;
}

// dgr_rk*_step(s): every stage reads the step's start time, params->time
RhsParams to_params(const dgr_bssn_params *p) { return to_params(p, p->time); }

int load_ptrs(VarPtrs &vp, const double *const *src, const char *what) {
    if (!src) return fail(DGR_E_INVALID, std::string(what) + " is null");
    for (int v = 0; v < NV; v++) {
        if (!src[v]) return fail(DGR_E_INVALID, std::string(what) + "[" + std::to_string(v) + "] is null");
        vp.p[v] = const_cast<double *>(src[v]);
    }
    return DGR_OK;
}

// Runs `launch` on `st`, bracketed by timing events when the plan records them.
template <typename F>
void timed(dgr_plan *plan, int kind, hipStream_t st, F launch) {
    if (!plan->timing) {
        launch();
        return;
    }
    hipEvent_t a = plan->take_event(), b = plan->take_event();
    const bool ra = a && hipEventRecord(a, st) == hipSuccess;
    launch();
    const bool rb = b && hipEventRecord(b, st) == hipSuccess;
    if (ra && rb) {
        plan->pending[kind].push_back({a, b});
    } else {  // an unrecorded pair would fail in dgr_plan_kernel_times: drop it
        if (a) plan->event_pool.push_back(a);
        if (b) plan->event_pool.push_back(b);
    }
}

template <int MODE>
int run_chunks(dgr_plan *plan, const VarPtrs &state, const VarPtrs &rhs, const RKArgs &rk,
               const RhsParams &prm, hipStream_t st) {
    if (!plan->d_W && plan->max_chunk_points) {
        hipError_t e = hipMalloc(&plan->d_W, sizeof(double) * NSLOT * (size_t)plan->max_chunk_points);
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "derivative workspace");
    }
    for (const Chunk &c : plan->chunks) {
        const int ntile = c.tile_end - c.tile_begin;
        timed(plan, 0, st, [&] {
            hipLaunchKernelGGL(k_derivs, dim3(ntile, NV), dim3(NTHREADS), 0, st, state, plan->d_blocks,
                               plan->d_tiles, c.tile_begin, plan->d_W, (int64_t)c.npoints);
        });
        timed(plan, 1, st, [&] {
            hipLaunchKernelGGL(k_rhs<MODE>, dim3(c.nwg), dim3(NTHREADS), 0, st, state, rhs, rk, prm,
                               plan->d_blocks, plan->d_tiles, plan->d_wg_tile + c.wg_begin, c.tile_end,
                               plan->d_W, (int64_t)c.npoints);
        });
        if (c.nface)
            timed(plan, 2, st, [&] {
                hipLaunchKernelGGL(k_face<MODE>, dim3(c.nface), dim3(NTHREADS), 0, st, state, rhs, rk, prm,
                                   plan->d_blocks, plan->d_tiles, plan->d_face_tiles + c.face_begin, plan->d_W,
                                   (int64_t)c.npoints);
            });
    }
    HIP_OK(hipGetLastError());
    return DGR_OK;
}

// Address of a fused kernel: the FP64 ones live in this unit, the FP32 ones
// in the unit compiled with DGR_F32_KERNELS (see there).
template <bool FACE, class C, typename T, bool HALO, bool GAUGE = false>
const void *fused_kernel() {
    if constexpr (sizeof(T) == sizeof(double)) {
        return reinterpret_cast<const void *>(&k_fused<FACE, C, T, HALO, GAUGE>);
    } else {
        static_assert(!GAUGE, "the gauge variants are FP64 only");
        static const void *const *tab = [] {
            static const void *k[8];
            dgr_internal_fused_f32_kernels(k);
            return k;
        }();
        constexpr int h = HALO ? 4 : 0;  // the in-kernel refresh variants follow the others
        if constexpr (FACE) return tab[h + 3];
        else if constexpr (C::FT == CfgFull::FT) return tab[h + 2];
        else if constexpr (C::FT == CfgThin::FT) return tab[h + 1];
        else return tab[h];
    }
}

template <bool FACE, class C, typename T, bool HALO = false, bool GAUGE = false>
void launch_fused(dgr_plan *plan, const Tile *tiles, const int *fg, int group, int dtab, const StageIO &io,
                  const RhsParams &prm, hipStream_t st) {
    const int n = fg[group + 1] - fg[group];
    if (!n) return;
    // k_fused's parameters, in order
    StageIO a_io = io;
    RhsParams a_prm = prm;
    const DBlock *a_blocks = plan->d_blocks;
    const dgr::AxisConsts *a_bconst = plan->d_bconst;
    const Tile *a_tiles = tiles + fg[group];
    const uint16_t *a_tabs = plan->d_tabs;
    int a_dtab = dtab;
    void *args[] = {&a_io, &a_prm, &a_blocks, &a_bconst, &a_tiles, &a_tabs, &a_dtab};
    timed(plan, 3 + group, st, [&] {
        (void)hipLaunchKernel(fused_kernel<FACE, C, T, HALO, GAUGE>(), dim3(n), dim3(C::FT), args, 0, st);
    });
}

// The launches of one tile list (groups as in fg).
template <typename T>
void launch_fused_all(dgr_plan *plan, const Tile *tiles, const int *fg, const StageIO &io, const RhsParams &prm,
                      hipStream_t st) {
    if constexpr (sizeof(T) == sizeof(double)) {
        if (gauge_variant(prm)) {  // the reference's other builds (FP64 only)
            if (io.nbo) {
                launch_fused<false, CfgWide, T, true, true>(plan, tiles, fg, FG_WIDE, plan->dtab_wide, io, prm, st);
                launch_fused<false, CfgThin, T, true, true>(plan, tiles, fg, FG_THIN, plan->dtab_thin, io, prm, st);
                launch_fused<true, CfgWide, T, true, true>(plan, tiles, fg, FG_FACE, plan->dtab_wide, io, prm, st);
            } else {
                launch_fused<false, CfgWide, T, false, true>(plan, tiles, fg, FG_WIDE, plan->dtab_wide, io, prm, st);
                launch_fused<false, CfgThin, T, false, true>(plan, tiles, fg, FG_THIN, plan->dtab_thin, io, prm, st);
                launch_fused<true, CfgWide, T, false, true>(plan, tiles, fg, FG_FACE, plan->dtab_wide, io, prm, st);
            }
            return;
        }
    }
    if (io.nbo) {  // the padding refresh inside the stage kernel
        launch_fused<false, typename CfgOf<T>::Wide, T, true>(plan, tiles, fg, FG_WIDE, plan->dtab_wide, io, prm, st);
        launch_fused<false, typename CfgOf<T>::Thin, T, true>(plan, tiles, fg, FG_THIN, plan->dtab_thin, io, prm, st);
        if constexpr (sizeof(T) < sizeof(double))
            launch_fused<false, CfgFull, T, true>(plan, tiles, fg, FG_FULL, plan->dtab_full, io, prm, st);
        launch_fused<true, typename CfgOf<T>::Wide, T, true>(plan, tiles, fg, FG_FACE, plan->dtab_wide, io, prm, st);
        return;
    }
    launch_fused<false, typename CfgOf<T>::Wide, T>(plan, tiles, fg, FG_WIDE, plan->dtab_wide, io, prm, st);
    launch_fused<false, typename CfgOf<T>::Thin, T>(plan, tiles, fg, FG_THIN, plan->dtab_thin, io, prm, st);
    if constexpr (sizeof(T) < sizeof(double))
        launch_fused<false, CfgFull, T>(plan, tiles, fg, FG_FULL, plan->dtab_full, io, prm, st);
    launch_fused<true, typename CfgOf<T>::Wide, T>(plan, tiles, fg, FG_FACE, plan->dtab_wide, io, prm, st);
}

int run_fused(dgr_plan *plan, const StageIO &io, const RhsParams &prm, hipStream_t st) {
    if (plan->precision == DGR_PRECISION_MIXED && gauge_variant(prm))
        return fail(DGR_E_INVALID, "the eta function, the Rochester gauge and the slow-start lapse "
                                   "(eta_func, rochester, ssl) need FP64 precision");
    if (plan->precision == DGR_PRECISION_MIXED)
        launch_fused_all<float>(plan, plan->d_ftiles[1], plan->fgroup[1], io, prm, st);
    else
        launch_fused_all<double>(plan, plan->d_ftiles[0], plan->fgroup[0], io, prm, st);
    HIP_OK(hipGetLastError());
    return DGR_OK;
}

}  // namespace

extern "C" {

int dgr_bssn_rhs(dgr_plan *plan, double *const *d_rhs, const double *const *d_u, double t,
                 const dgr_bssn_params *params, void *stream) {
    g_err.clear();
    if (!plan || !params) return fail(DGR_E_INVALID, "null plan or params");
    VarPtrs u, r;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u")) || (rc = load_ptrs(r, d_rhs, "d_rhs"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    if (plan->fused) {
        StageIO io{};
        io.src = u;
        io.rhs = r;
        io.stage = -1;
        io.err = plan->d_err;
        return run_fused(plan, io, to_params(params, t), (hipStream_t)stream);
    }
    RKArgs rk{};
    return run_chunks<MODE_RHS>(plan, u, r, rk, to_params(params, t), (hipStream_t)stream);
}

int dgr_plan_autotune(dgr_plan *plan, double *const *d_rhs, const double *const *d_u,
                      const dgr_bssn_params *params, void *stream) {
    g_err.clear();
    if (!plan || !params) return fail(DGR_E_INVALID, "null plan or params");
    if (!plan->fused) return fail(DGR_E_INVALID, "autotuning needs the fused kernel (DGR_KERNEL=split is set)");
    VarPtrs u, r;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u")) || (rc = load_ptrs(r, d_rhs, "d_rhs"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    const hipStream_t st = (hipStream_t)stream;
    const int p = plan->precision == DGR_PRECISION_MIXED ? 1 : 0;
    StageIO io{};
    io.src = u;
    io.rhs = r;
    io.stage = -1;
    io.err = plan->d_err;
    const RhsParams prm = to_params(params);
    std::map<std::array<int, 3>, std::vector<uint32_t>> sizes;  // blocks without flagged faces by size
    for (uint32_t b = 0; b < (uint32_t)plan->hblocks.size(); b++) {
        const DBlock &d = plan->hblocks[b];
        if (!d.bflag) sizes[{d.n[0], d.n[1], d.n[2]}].push_back(b);
    }
    const bool timing = plan->timing;
    plan->timing = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Tile *d_t = nullptr;
    size_t cap = 0;
    auto tune = [&]() -> int {
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        for (auto &kv : sizes) {
            // a sample of at most 2048 blocks of this size
            const size_t ns = std::min<size_t>(kv.second.size(), 2048);
            int best = -1;
            float best_ms = 0.f;
            for (int sch = 0; sch < NSCHEME; sch++) {
                if (sch == SCH_FULL && p == 0) continue;
                std::vector<std::vector<Tile>> per[NFGROUP];
                for (auto &v : per) v.resize(ns);
                for (size_t i = 0; i < ns; i++) block_tiles(plan->hblocks[kv.second[i]], kv.second[i], sch, per, i);
                std::vector<Tile> tiles;
                int fg[NFGROUP + 1];
                group_tiles(per, tiles, fg);
                if (!assign_lane_tables(plan, tiles, fg)) return fail(DGR_E_INVALID, "internal error: no lane table");
                HIP_OK(hipStreamSynchronize(st));  // earlier candidates' launches read d_t
                if (tiles.size() > cap) {
                    HIP_OK(hipFree(d_t));
                    d_t = nullptr;
                    HIP_OK(hipMalloc(&d_t, sizeof(Tile) * tiles.size()));
                    cap = tiles.size();
                }
                HIP_OK(hipMemcpy(d_t, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice));
                float ms = 0.f;
                for (int rep = 0; rep < 4; rep++) {  // a warm-up, then the best of three
                    HIP_OK(hipEventRecord(e0, st));
                    if (p) launch_fused_all<float>(plan, d_t, fg, io, prm, st);
                    else launch_fused_all<double>(plan, d_t, fg, io, prm, st);
                    HIP_OK(hipGetLastError());
                    HIP_OK(hipEventRecord(e1, st));
                    HIP_OK(hipEventSynchronize(e1));
                    float t = 0.f;
                    HIP_OK(hipEventElapsedTime(&t, e0, e1));
                    if (rep == 1 || (rep > 1 && t < ms)) ms = t;
                }
                if (best < 0 || ms < best_ms) {
                    best = sch;
                    best_ms = ms;
                }
            }
            plan->tiling[p][kv.first] = best;
        }
        HIP_OK(build_ftiles(plan, p));
        if (plan->rk4_exec) {  // the captured step used the previous tile list
            HIP_OK(hipGraphExecDestroy(plan->rk4_exec));
            plan->rk4_exec = nullptr;
        }
        return DGR_OK;
    };
    rc = tune();
    plan->timing = timing;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (d_t) (void)hipFree(d_t);
    return rc;
}

int dgr_plan_set_tiling(dgr_plan *plan, int precision, uint32_t nx, uint32_t ny, uint32_t nz, int scheme) {
    g_err.clear();
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    if (precision != DGR_PRECISION_FP64 && precision != DGR_PRECISION_MIXED)
        return fail(DGR_E_INVALID, "unknown precision");
    if (scheme < 0 || scheme >= NSCHEME || (scheme == SCH_FULL && precision == DGR_PRECISION_FP64))
        return fail(DGR_E_INVALID, "unknown tiling for this precision");
    const int p = precision == DGR_PRECISION_MIXED ? 1 : 0;
    plan->tiling[p][{(int)nx, (int)ny, (int)nz}] = scheme;
    HIP_OK(hipSetDevice(plan->device));
    HIP_OK(build_ftiles(plan, p));
    if (plan->rk4_exec) {  // the captured step used the previous tile list
        HIP_OK(hipGraphExecDestroy(plan->rk4_exec));
        plan->rk4_exec = nullptr;
    }
    return DGR_OK;
}

int dgr_plan_tiling(const dgr_plan *plan, int precision, int max_sizes, uint32_t *sizes, int *scheme) {
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    if (precision != DGR_PRECISION_FP64 && precision != DGR_PRECISION_MIXED)
        return fail(DGR_E_INVALID, "unknown precision");
    if (max_sizes > 0 && (!sizes || !scheme)) return fail(DGR_E_INVALID, "null output");
    const int p = precision == DGR_PRECISION_MIXED ? 1 : 0;
    std::map<std::array<int, 3>, int> seen;
    for (const DBlock &d : plan->hblocks)
        if (!d.bflag) seen.emplace(std::array<int, 3>{d.n[0], d.n[1], d.n[2]}, tiling_of(plan, p, d));
    int i = 0;
    for (auto &kv : seen) {
        if (i < max_sizes) {
            for (int a = 0; a < 3; a++) sizes[3 * i + a] = (uint32_t)kv.first[a];
            scheme[i] = kv.second;
        }
        i++;
    }
    return i;
}

}  // extern "C"

namespace {

// RK4 table: include/rkBSSN.h:50-53
const double RKC[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
const double RKU[4] = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};

int ensure_istage(dgr_plan *plan) {
    const size_t ni = (size_t)plan->interior;
    if (!plan->d_istage && ni) {
        hipError_t e = hipMalloc(&plan->d_istage, sizeof(double) * 2 * NV * ni);
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "stage buffers");
    }
    if (!plan->d_wstate && ni) {
        hipError_t e = hipMalloc(&plan->d_wstate, sizeof(double) * NV * plan->szpdof);
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "work state");
    }
    return DGR_OK;
}

// The RK4 work buffers of the fused path: the padded work state W (24 x
// szpdof doubles) and the interior-only u0 and accumulator (2 x 24 x interior).
struct RkBuffers {
    double *wstate;
    double *istage;
};
RkBuffers plan_rk_buffers(const dgr_plan *plan) { return {plan->d_wstate, plan->d_istage}; }

// W's padding <- u's (before the steps of one RK4 entry-point call)
int launch_pad_copy(dgr_plan *plan, const RkBuffers &buf, const VarPtrs &u, hipStream_t st) {
    const uint32_t nb = (uint32_t)plan->hblocks.size();
    if (!nb) return DGR_OK;
    VarPtrs w;
    for (int v = 0; v < NV; v++) w.p[v] = buf.wstate + (size_t)v * plan->szpdof;
    hipLaunchKernelGGL(k_pad_copy, dim3(nb, NV), dim3(256), 0, st, w, u, plan->d_blocks);
    HIP_OK(hipGetLastError());
    return DGR_OK;
}

// Same-level neighbour table from the block geometry: nb is b's neighbour in
// direction d when both have the same sizes and spacing (within 1e-9
// relative) and pmin_nb = pmin_b + d (n-2*PW-1) h within 1e-6 h. Blocks are
// grouped by (sizes, spacing) and indexed by their lattice position
// relative to the group's first block.
int link_neighbors(dgr_plan *plan) {
    const size_t nb = plan->hblocks.size();
    plan->hnbr.assign(27 * nb, -1);
    std::vector<int> group(nb, -1);
    std::vector<int> rep;  // first block of each group
    for (size_t b = 0; b < nb; b++) {
        const DBlock &B = plan->hblocks[b];
        for (size_t g = 0; g < rep.size() && group[b] < 0; g++) {
            const DBlock &R = plan->hblocks[rep[g]];
            bool same = true;
            for (int a = 0; a < 3; a++)
                same = same && B.n[a] == R.n[a] && std::fabs(B.h[a] - R.h[a]) <= 1e-9 * R.h[a];
            if (same) group[b] = (int)g;
        }
        if (group[b] < 0) {
            group[b] = (int)rep.size();
            rep.push_back((int)b);
        }
    }
    std::map<std::array<int64_t, 4>, int> at;
    std::vector<std::array<int64_t, 3>> cell(nb);
    for (size_t b = 0; b < nb; b++) {
        const DBlock &B = plan->hblocks[b], &R = plan->hblocks[rep[group[b]]];
        for (int a = 0; a < 3; a++) cell[b][a] = std::llround((B.pmin[a] - R.pmin[a]) / ((B.n[a] - 2 * PW - 1) * R.h[a]));
        at[{group[b], cell[b][0], cell[b][1], cell[b][2]}] = (int)b;
    }
    int64_t links = 0;
    for (size_t b = 0; b < nb; b++) {
        const DBlock &B = plan->hblocks[b];
        for (int d = 0; d < 27; d++) {
            const int di = d % 3 - 1, dj = (d / 3) % 3 - 1, dk = d / 9 - 1;
            if (!(di | dj | dk)) continue;
            auto it = at.find({group[b], cell[b][0] + di, cell[b][1] + dj, cell[b][2] + dk});
            if (it == at.end()) continue;
            const DBlock &N = plan->hblocks[it->second];
            const int dd[3] = {di, dj, dk};
            bool ok = true;
            for (int a = 0; a < 3; a++)
                ok = ok && std::fabs(N.pmin[a] - (B.pmin[a] + dd[a] * (B.n[a] - 2 * PW - 1) * B.h[a])) <= 1e-6 * B.h[a];
            if (!ok) continue;
            plan->hnbr[27 * b + d] = it->second;
            links++;
        }
    }
    if (!plan->d_nbr && nb) {
        hipError_t e = hipMalloc(&plan->d_nbr, sizeof(int) * 27 * nb);
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "neighbour table");
    }
    if (nb) HIP_OK(hipMemcpy(plan->d_nbr, plan->hnbr.data(), sizeof(int) * 27 * nb, hipMemcpyHostToDevice));
    std::vector<int64_t> off(27 * nb);
    for (size_t i = 0; i < off.size(); i++)
        off[i] = plan->hnbr[i] >= 0 ? (int64_t)plan->hblocks[plan->hnbr[i]].offset : -1;
    if (!plan->d_nbo && nb) {
        hipError_t e = hipMalloc(&plan->d_nbo, sizeof(int64_t) * 27 * nb);
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "neighbour table");
    }
    if (nb) HIP_OK(hipMemcpy(plan->d_nbo, off.data(), sizeof(int64_t) * 27 * nb, hipMemcpyHostToDevice));
    plan->nlinks = links;
    plan->nfill = -1;  // the general fill records follow the links
    // k_halo_fill's workgroups: (block, chunk of 256 points) for the blocks
    // with at least one link, chunk-major (consecutive workgroups take
    // different blocks)
    std::vector<int2> wgs;
    int maxch = 0;
    std::vector<int> nch(nb, 0);
    for (size_t b = 0; b < nb; b++) {
        bool linked = false;
        for (int d = 0; d < 27 && !linked; d++) linked = plan->hnbr[27 * b + d] >= 0;
        const DBlock &B = plan->hblocks[b];
        if (linked) nch[b] = (int)(((int64_t)B.n[0] * B.n[1] * B.n[2] + 255) / 256);
        maxch = std::max(maxch, nch[b]);
    }
    for (int c = 0; c < maxch; c++)
        for (size_t b = 0; b < nb; b++)
            if (c < nch[b]) wgs.push_back(make_int2((int)b, c));
    if (wgs.size() > (size_t)INT32_MAX) return fail(DGR_E_INVALID, "padding fill: too many blocks x points");
    if (plan->rk4_exec && plan->halo) {  // a captured step with the refresh holds the old table
        HIP_OK(hipGraphExecDestroy(plan->rk4_exec));
        plan->rk4_exec = nullptr;
    }
    if (plan->d_halo_wg) {
        HIP_OK(hipFree(plan->d_halo_wg));
        plan->d_halo_wg = nullptr;
    }
    if (!wgs.empty()) {
        hipError_t e = hipMalloc(&plan->d_halo_wg, sizeof(int2) * wgs.size());
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "padding fill table");
        HIP_OK(hipMemcpy(plan->d_halo_wg, wgs.data(), sizeof(int2) * wgs.size(), hipMemcpyHostToDevice));
    }
    plan->nhalo_wg = (int64_t)wgs.size();
    return DGR_OK;
}

// Where a padding point at coordinate X along axis a lies on block C's grid,
// within C's interior: mirrors the oracle's padding_source. rel 0 (same
// spacing) and 1 (C finer) need a node of C: q is that node; rel 2 (C coarser)
// gives the first node q of C's element (6 intervals) holding X and cd = 2t,
// t the element-local coordinate.
bool fill_axis(const DBlock &C, int a, double X, int rel, int &q, int &cd) {
    const double qa = (X - C.pmin[a]) / C.h[a];
    const double q2 = 2.0 * qa;
    const double r2 = std::floor(q2 + 0.5);
    if (std::fabs(q2 - r2) > 2e-6) return false;
    const long i2 = (long)r2;
    if (i2 < 2 * PW || i2 > 2 * ((long)C.n[a] - PW - 1)) return false;
    if (rel < 2 && (i2 & 1)) return false;
    if (rel < 2) {
        q = (int)(i2 / 2);
        cd = 0;
    } else {
        const long t2 = i2 - 2 * PW, ne = ((long)C.n[a] - 2 * PW - 1) / 6;
        long e = t2 / 12;
        if (e > ne - 1) e = ne - 1;
        q = (int)(PW + 6 * e);
        cd = (int)(t2 - 12 * e);
    }
    return true;
}

// Spacing of C relative to B along x: 0 same, 1 half (C finer), 2 twice (C
// coarser), -1 other (the oracle's rule, relative tolerance 1e-9).
int fill_rel(const DBlock &B, const DBlock &C) {
    const double hb = B.h[0], hc = C.h[0];
    if (std::fabs(hc - hb) <= 1e-9 * hb) return 0;
    if (std::fabs(2.0 * hc - hb) <= 1e-9 * hb) return 1;
    if (std::fabs(hc - 2.0 * hb) <= 1e-9 * hb) return 2;
    return -1;
}

// The general fill records: for every padding point of every block whose
// direction has no same-size neighbour (the link table must be built), the
// block whose interior holds it, preferring the same spacing, then half, then
// twice the spacing, ties to the lower block index (as the oracle's
// orc_padding_fill). Candidates come from a hash of the blocks' interior boxes
// per spacing; the point set of a direction's padding region is a tensor
// product, so each candidate is matched one axis at a time.
int build_fill(dgr_plan *plan) {
    const size_t nb = plan->hblocks.size();
    const auto &H = plan->hblocks;
    std::vector<double> lvh;  // distinct x spacings
    std::vector<int> lv(nb);
    for (size_t b = 0; b < nb; b++) {
        int l = -1;
        for (size_t k = 0; k < lvh.size() && l < 0; k++)
            if (std::fabs(H[b].h[0] - lvh[k]) <= 1e-12 * lvh[k]) l = (int)k;
        if (l < 0) {
            l = (int)lvh.size();
            lvh.push_back(H[b].h[0]);
        }
        lv[b] = l;
    }
    std::vector<double> cell(lvh.size(), 0.0);  // hash cell edge per spacing: its largest interior extent
    for (size_t b = 0; b < nb; b++)
        for (int a = 0; a < 3; a++) cell[lv[b]] = std::max(cell[lv[b]], (H[b].n[a] - 2 * PW - 1) * H[b].h[a]);
    std::map<std::array<int64_t, 4>, std::vector<int>> grid;
    auto cell_of = [&](int l, double x) { return (int64_t)std::floor(x / cell[l]); };
    for (size_t b = 0; b < nb; b++) {
        const int l = lv[b];
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            const double e = 1e-6 * cell[l];
            lo[a] = cell_of(l, H[b].pmin[a] + PW * H[b].h[a] - e);
            hi[a] = cell_of(l, H[b].pmin[a] + (H[b].n[a] - PW - 1) * H[b].h[a] + e);
        }
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++) grid[{l, x, y, z}].push_back((int)b);
    }
    struct Best {
        int rel, c, q[3], cd[3];
    };
    std::vector<FillRec> recs;
    int64_t ninterp = 0;
    std::vector<Best> best;
    std::vector<int> cand;
    for (size_t b = 0; b < nb; b++) {
        const DBlock &B = H[b];
        for (int d = 0; d < 27; d++) {
            const int dd[3] = {d % 3 - 1, (d / 3) % 3 - 1, d / 9 - 1};
            if (!(dd[0] | dd[1] | dd[2]) || plan->hnbr[27 * b + d] >= 0) continue;
            int r0[3], r1[3];  // the region's index range per axis
            double xlo[3], xhi[3];
            for (int a = 0; a < 3; a++) {
                r0[a] = dd[a] < 0 ? 0 : (dd[a] == 0 ? PW : B.n[a] - PW);
                r1[a] = dd[a] < 0 ? PW : (dd[a] == 0 ? B.n[a] - PW : B.n[a]);
                xlo[a] = B.pmin[a] + r0[a] * B.h[a];
                xhi[a] = B.pmin[a] + (r1[a] - 1) * B.h[a];
            }
            cand.clear();
            for (int l = 0; l < (int)lvh.size(); l++) {
                const double h = lvh[l];
                if (!(std::fabs(h - B.h[0]) <= 1e-8 * B.h[0] || std::fabs(2.0 * h - B.h[0]) <= 1e-8 * B.h[0] ||
                      std::fabs(h - 2.0 * B.h[0]) <= 1e-8 * B.h[0]))
                    continue;
                int64_t lo[3], hi[3];
                for (int a = 0; a < 3; a++) {
                    const double e = 1e-6 * cell[l];
                    lo[a] = cell_of(l, xlo[a] - e);
                    hi[a] = cell_of(l, xhi[a] + e);
                }
                for (int64_t z = lo[2]; z <= hi[2]; z++)
                    for (int64_t y = lo[1]; y <= hi[1]; y++)
                        for (int64_t x = lo[0]; x <= hi[0]; x++) {
                            auto it = grid.find({l, x, y, z});
                            if (it == grid.end()) continue;
                            for (int c : it->second)
                                if (c != (int)b) cand.push_back(c);
                        }
            }
            if (cand.empty()) continue;
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            const int ex = r1[0] - r0[0], ey = r1[1] - r0[1], ez = r1[2] - r0[2];
            best.assign((size_t)ex * ey * ez, Best{3, -1, {0, 0, 0}, {0, 0, 0}});
            bool any = false;
            for (int c : cand) {
                const DBlock &C = H[c];
                const int rel = fill_rel(B, C);
                if (rel < 0) continue;
                std::vector<std::array<int, 3>> ax[3];  // (local index, q, cd) per axis
                bool ok = true;
                for (int a = 0; a < 3 && ok; a++) {
                    for (int i = r0[a]; i < r1[a]; i++) {
                        int q, cd;
                        if (fill_axis(C, a, B.pmin[a] + i * B.h[a], rel, q, cd)) ax[a].push_back({i - r0[a], q, cd});
                    }
                    ok = !ax[a].empty();
                }
                if (!ok) continue;
                for (const auto &z : ax[2])
                    for (const auto &y : ax[1])
                        for (const auto &x : ax[0]) {
                            Best &bb = best[(size_t)x[0] + (size_t)ex * (y[0] + (size_t)ey * z[0])];
                            if (rel < bb.rel || (rel == bb.rel && c < bb.c)) {
                                bb = Best{rel, c, {x[1], y[1], z[1]}, {x[2], y[2], z[2]}};
                                any = true;
                            }
                        }
            }
            if (!any) continue;
            for (int k = 0; k < ez; k++)
                for (int j = 0; j < ey; j++)
                    for (int i = 0; i < ex; i++) {
                        const Best &bb = best[(size_t)i + (size_t)ex * (j + (size_t)ey * k)];
                        if (bb.rel == 3) continue;
                        const DBlock &C = H[bb.c];
                        FillRec r{};
                        r.dst = B.offset + (uint64_t)(r0[0] + i) +
                                (uint64_t)B.n[0] * ((uint64_t)(r0[1] + j) + (uint64_t)B.n[1] * (r0[2] + k));
                        r.src = C.offset + (uint64_t)bb.q[0] +
                                (uint64_t)C.n[0] * ((uint64_t)bb.q[1] + (uint64_t)C.n[1] * bb.q[2]);
                        r.mx = (uint32_t)C.n[0];
                        r.mxy = (uint32_t)C.n[0] * (uint32_t)C.n[1];
                        r.code = bb.rel == 2 ? (uint32_t)bb.cd[0] | ((uint32_t)bb.cd[1] << 4) |
                                                   ((uint32_t)bb.cd[2] << 8) | FILL_INTERP
                                             : 0u;
                        ninterp += bb.rel == 2;
                        recs.push_back(r);
                    }
        }
    }
    // copies first; the interpolations by coarse element (first node), then by
    // the point they write (neighbouring threads write neighbouring points),
    // cut into groups of at most FILL_GROUP_MAX that share the element
    auto interp = [](const FillRec &r) { return (r.code & FILL_INTERP) != 0; };
    std::stable_sort(recs.begin(), recs.end(), [&](const FillRec &a, const FillRec &b) {
        const bool ia = interp(a), ib = interp(b);
        if (ia != ib) return ib;
        if (!ia) return false;
        if (a.src != b.src) return a.src < b.src;
        return a.dst < b.dst;
    });
    std::vector<FillGroup> groups;
    for (size_t k = 0; k < recs.size(); k++) {
        if (!interp(recs[k])) continue;
        const uint32_t cd = recs[k].code;
        if ((cd & 15u) > 12 || ((cd >> 4) & 15u) > 12 || ((cd >> 8) & 15u) > 12)
            return fail(DGR_E_INVALID, "padding fill: element coordinate out of range");
        if (groups.empty() || groups.back().src != recs[k].src || groups.back().count == FILL_GROUP_MAX) {
            if (k > UINT32_MAX) return fail(DGR_E_INVALID, "padding fill: too many records");
            groups.push_back(FillGroup{recs[k].src, recs[k].mx, recs[k].mxy, (uint32_t)k, 0});
        }
        groups.back().count++;
    }
    if (plan->rk4_exec && plan->fill_general) {  // a captured step with the full refresh holds the old records
        HIP_OK(hipGraphExecDestroy(plan->rk4_exec));
        plan->rk4_exec = nullptr;
    }
    if (plan->d_fill) {
        HIP_OK(hipFree(plan->d_fill));
        plan->d_fill = nullptr;
    }
    if (!recs.empty()) {
        hipError_t e = hipMalloc(&plan->d_fill, sizeof(FillRec) * recs.size());
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "padding fill records");
        HIP_OK(hipMemcpy(plan->d_fill, recs.data(), sizeof(FillRec) * recs.size(), hipMemcpyHostToDevice));
    }
    if (plan->d_fill_groups) {
        HIP_OK(hipFree(plan->d_fill_groups));
        plan->d_fill_groups = nullptr;
    }
    if (!groups.empty()) {
        hipError_t e = hipMalloc(&plan->d_fill_groups, sizeof(FillGroup) * groups.size());
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "padding fill groups");
        HIP_OK(hipMemcpy(plan->d_fill_groups, groups.data(), sizeof(FillGroup) * groups.size(), hipMemcpyHostToDevice));
    }
    plan->nfill_groups = (int64_t)groups.size();
    plan->nfill = (int64_t)recs.size();
    plan->nfill_interp = ninterp;
    return DGR_OK;
}

// Lagrange weights of the order-6 element at t = code/2, in the operation
// order of the oracle's orc_lagrange6_half (so the bits are the oracle's).
Lag6Table lag6_table() {
#pragma clang fp contract(off)
    Lag6Table tab;
    for (int code = 0; code < 13; code++) {
        const double t = 0.5 * (double)code;
        for (int m = 0; m < 7; m++) {
            double num = 1.0, den = 1.0;
            for (int l = 0; l < 7; l++) {
                if (l == m) continue;
                num *= (t - (double)l);
                den *= (double)(m - l);
            }
            tab.w[code][m] = num / den;
        }
    }
    return tab;
}

int launch_fill_general(dgr_plan *plan, const VarPtrs &u, hipStream_t st) {
    if (plan->nfill <= 0) return DGR_OK;
    static const Lag6Table tab = lag6_table();
    const int64_t ncopy = plan->nfill - plan->nfill_interp;  // the copies come first
    const int64_t nwg = (ncopy + 255) / 256;
    if (nwg > INT32_MAX || plan->nfill_groups > INT32_MAX) return fail(DGR_E_INVALID, "padding fill: too many records");
    if (ncopy)
        hipLaunchKernelGGL(k_fill_copy, dim3((uint32_t)nwg, NV), dim3(256), 0, st, u, plan->d_fill, ncopy);
    if (plan->nfill_groups)
        hipLaunchKernelGGL(k_fill_interp, dim3((uint32_t)plan->nfill_groups), dim3(FILL_WG), 0, st, u, plan->d_fill,
                           plan->d_fill_groups, tab);
    HIP_OK(hipGetLastError());
    return DGR_OK;
}

int launch_halo_fill(dgr_plan *plan, const VarPtrs &u, hipStream_t st) {
    if (plan->nhalo_wg <= 0) return DGR_OK;
    hipLaunchKernelGGL(k_halo_fill, dim3((uint32_t)plan->nhalo_wg), dim3(256), 0, st, u, plan->d_blocks, plan->d_nbr,
                       plan->d_halo_wg);
    HIP_OK(hipGetLastError());
    return DGR_OK;
}

// The whole refresh of u's padding: from same-size neighbours and, when the
// plan's refresh covers them, the general records.
int launch_padding_refresh(dgr_plan *plan, const VarPtrs &u, hipStream_t st) {
    int rc = launch_halo_fill(plan, u, st);
    if (!rc && plan->fill_general) rc = launch_fill_general(plan, u, st);
    return rc;
}

// With the padding refresh on, stages read linked padding straight from the
// neighbours (k_fused<..., HALO>); plans whose variables hold 2^29 points or
// more get the padding copied in by k_halo_fill before each stage instead.
bool halo_in_kernel(const dgr_plan *plan) {
    // the HALO kernels address the whole variable with 32-bit byte offsets;
    // DGR_HALO_IN_KERNEL=0 selects the copy for any plan (the path of larger plans)
    static const bool allowed = [] {
        const char *e = std::getenv("DGR_HALO_IN_KERNEL");
        return !(e && std::string(e) == "0");
    }();
    return plan->halo && plan->szpdof < (1ull << 29) && allowed;
}

// The four stage launches of one RK4 step on the fused path (rkBSSN.cpp:764-860):
// stages read u, W, u, W and write the next state into W, u, W, u (see
// StageIO). Needs ensure_istage() first.
int rk4_fused_launches(dgr_plan *plan, const RkBuffers &buf, const VarPtrs &u, double dt, const RhsParams &prm,
                       int enforce_per_stage, hipStream_t st) {
    const size_t ni = (size_t)plan->interior;
    VarPtrs w, u0, acc;
    for (int v = 0; v < NV; v++) {
        w.p[v] = buf.wstate + (size_t)v * plan->szpdof;
        u0.p[v] = buf.istage + (size_t)v * ni;
        acc.p[v] = buf.istage + (size_t)(NV + v) * ni;
    }
    for (int s = 0; s < 4; s++) {
        StageIO io{};
        io.src = (s & 1) ? w : u;
        io.dst = (s & 1) ? u : w;
        io.u0 = u0;
        io.acc = acc;
        io.cacc = RKC[s] * dt;
        io.cnext = s < 3 ? RKU[s + 1] * dt : 0.0;
        io.stage = s;
        io.enforce_stage = enforce_per_stage ? 1 : 0;
        io.err = plan->d_err;
        int rc;
        if (halo_in_kernel(plan)) {
            io.nbo = plan->d_nbo;  // the stage kernel reads the neighbours itself
        } else if (plan->halo && s > 0 && (rc = launch_halo_fill(plan, io.src, st))) {
            // copy first; stage 0 reads u, whose padding the entry point
            // (before the first step) or the previous step has refreshed
            return rc;
        }
        // padding across level jumps and from blocks of other sizes: written
        // into the stage input before every stage
        if (plan->halo && plan->fill_general && (rc = launch_fill_general(plan, io.src, st))) return rc;
        if ((rc = run_fused(plan, io, prm, st))) return rc;
    }
    if (plan->halo && !halo_in_kernel(plan)) return launch_padding_refresh(plan, u, st);
    return DGR_OK;
}

}  // namespace

extern "C" {

int dgr_rk4_step(dgr_plan *plan, double *const *d_u, double dt, const dgr_bssn_params *params,
                 int enforce_per_stage, void *stream) {
    g_err.clear();
    if (!plan || !params) return fail(DGR_E_INVALID, "null plan or params");
    VarPtrs u;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    hipStream_t st = (hipStream_t)stream;
    if (plan->fused) {
        if ((rc = ensure_istage(plan)) || (plan->halo && !halo_in_kernel(plan) && (rc = launch_halo_fill(plan, u, st))) ||
            (rc = launch_pad_copy(plan, plan_rk_buffers(plan), u, st)) ||
            (rc = rk4_fused_launches(plan, plan_rk_buffers(plan), u, dt, to_params(params), enforce_per_stage, st)))
            return rc;
        // in-kernel refresh: the caller's padding, refreshed from the new interiors
        return halo_in_kernel(plan) ? launch_padding_refresh(plan, u, st) : DGR_OK;
    }
    if (plan->halo) return fail(DGR_E_INVALID, "the padding refresh needs the fused path (DGR_KERNEL=split is set)");
    const size_t n = plan->szpdof;
    if (!plan->d_stage && n) {
        hipError_t e = hipMalloc(&plan->d_stage, sizeof(double) * 2 * NV * n);
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "stage buffers");
    }
    VarPtrs stage, acc;
    for (int v = 0; v < NV; v++) {
        stage.p[v] = plan->d_stage + (size_t)v * n;
        acc.p[v] = plan->d_stage + (size_t)(NV + v) * n;
        HIP_OK(hipMemcpyAsync(stage.p[v], u.p[v], sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    }
    const RhsParams prm = to_params(params);
    for (int s = 0; s < 4; s++) {
        RKArgs rk;
        rk.u0 = u;
        rk.acc = acc;
        rk.snext = stage;
        rk.cacc = RKC[s] * dt;
        rk.cnext = s < 3 ? RKU[s + 1] * dt : 0.0;
        rk.stage = s;
        rk.enforce_stage = enforce_per_stage ? 1 : 0;
        rk.err = plan->d_err;
        if ((rc = run_chunks<MODE_RK>(plan, stage, stage, rk, prm, st))) return rc;
    }
    return DGR_OK;
}

int dgr_rk4_steps(dgr_plan *plan, double *const *d_u, double dt, const dgr_bssn_params *params,
                  int enforce_per_stage, int nsteps, void *stream) {
    g_err.clear();
    if (!plan || !params) return fail(DGR_E_INVALID, "null plan or params");
    if (nsteps < 0) return fail(DGR_E_INVALID, "nsteps must be >= 0");
    VarPtrs u;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u"))) return rc;
    // per-launch timing and the split path run the steps eagerly (the
    // slow-start lapse's time moving on by dt per step, rkBSSN.cpp:1087)
    if (!plan->fused || plan->timing) {
        dgr_bssn_params p = *params;
        for (int i = 0; i < nsteps; i++, p.time += dt)
            if ((rc = dgr_rk4_step(plan, d_u, dt, &p, enforce_per_stage, stream))) return rc;
        return DGR_OK;
    }
    // the captured step reads the slow-start lapse's time from device memory,
    // where the graph moves it on: a new time keeps the graph
    dgr_bssn_params kp = *params;
    kp.time = 0.0;
    const bool ssl = params->ssl != 0;
    if (ssl && !plan->d_ssl) {
        hipError_t e = hipMalloc(&plan->d_ssl, 2 * sizeof(double));
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "lapse time");
    }
    HIP_OK(hipSetDevice(plan->device));
    if ((rc = ensure_istage(plan))) return rc;
    auto &key = plan->rk4_key;
    const int enf = enforce_per_stage ? 1 : 0;
    bool same = plan->rk4_exec && key.dt == dt && key.enforce == enf && key.halo == plan->halo &&
                key.fill_general == plan->fill_general &&
                std::memcmp(&key.prm, &kp, sizeof(kp)) == 0;
    for (int v = 0; same && v < NV; v++) same = key.u[v] == u.p[v];
    if (!same) {
        if (plan->rk4_exec) {
            HIP_OK(hipGraphExecDestroy(plan->rk4_exec));
            plan->rk4_exec = nullptr;
        }
        if (!plan->cap_stream) HIP_OK(hipStreamCreateWithFlags(&plan->cap_stream, hipStreamNonBlocking));
        HIP_OK(hipStreamBeginCapture(plan->cap_stream, hipStreamCaptureModeThreadLocal));
        RhsParams prm = to_params(params);
        if (ssl) {
            prm.ssl_dev = plan->d_ssl;
            hipLaunchKernelGGL(k_ssl_begin, dim3(1), dim3(1), 0, plan->cap_stream, plan->d_ssl);
        }
        rc = rk4_fused_launches(plan, plan_rk_buffers(plan), u, dt, prm, enf, plan->cap_stream);
        if (ssl) hipLaunchKernelGGL(k_ssl_advance, dim3(1), dim3(1), 0, plan->cap_stream, plan->d_ssl, dt);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(plan->cap_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return fail(DGR_E_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&plan->rk4_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            plan->rk4_exec = nullptr;
            return fail(DGR_E_HIP, std::string("graph instantiation: ") + hipGetErrorString(ei));
        }
        for (int v = 0; v < NV; v++) key.u[v] = u.p[v];
        key.dt = dt;
        key.prm = kp;
        key.enforce = enf;
        key.halo = plan->halo;
        key.fill_general = plan->fill_general;
    }
    if (nsteps > 0 && ssl) {
        unsigned long long bits;
        std::memcpy(&bits, &params->time, sizeof(bits));
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, (hipStream_t)stream,
                           reinterpret_cast<unsigned long long *>(plan->d_ssl), bits);
    }
    if (nsteps > 0 && ((plan->halo && !halo_in_kernel(plan) && (rc = launch_halo_fill(plan, u, (hipStream_t)stream))) ||
                       (rc = launch_pad_copy(plan, plan_rk_buffers(plan), u, (hipStream_t)stream))))
        return rc;
    for (int i = 0; i < nsteps; i++) HIP_OK(hipGraphLaunch(plan->rk4_exec, (hipStream_t)stream));
    // in-kernel refresh: the caller's padding, refreshed from the new interiors
    if (nsteps > 0 && halo_in_kernel(plan)) return launch_padding_refresh(plan, u, (hipStream_t)stream);
    return DGR_OK;
}

int dgr_variable_range(dgr_plan *plan, const double *d_var, double *vmin, double *vmax, int64_t *nonfinite,
                       void *stream) {
    g_err.clear();
    if (!plan || !d_var) return fail(DGR_E_INVALID, "null plan or variable");
    HIP_OK(hipSetDevice(plan->device));
    if (!plan->d_range) {
        hipError_t e = hipMalloc(&plan->d_range, 3 * sizeof(unsigned long long));
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "range scratch");
    }
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_var_range_init, dim3(1), dim3(1), 0, st, plan->d_range);
    for (const Chunk &c : plan->chunks)
        hipLaunchKernelGGL(k_var_range, dim3(c.nwg), dim3(NTHREADS), 0, st, d_var, plan->d_blocks, plan->d_tiles,
                           plan->d_wg_tile + c.wg_begin, c.tile_end, (int64_t)c.npoints, plan->d_range);
    HIP_OK(hipGetLastError());
    unsigned long long r[3];
    HIP_OK(hipMemcpyAsync(r, plan->d_range, sizeof(r), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    auto value = [](unsigned long long k) {
        const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
        double x;
        std::memcpy(&x, &b, sizeof(x));
        return x;
    };
    const bool any = r[0] <= r[1];  // no finite value leaves the initial keys
    if (vmin) *vmin = any ? value(r[0]) : std::nan("");
    if (vmax) *vmax = any ? value(r[1]) : std::nan("");
    if (nonfinite) *nonfinite = (int64_t)r[2];
    return DGR_OK;
}

int dgr_rk3_step(dgr_plan *plan, double *const *d_u, double dt, const dgr_bssn_params *params, void *stream) {
    g_err.clear();
    if (!plan || !params) return fail(DGR_E_INVALID, "null plan or params");
    VarPtrs u;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    if ((rc = ensure_istage(plan))) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t ni = (size_t)plan->interior;
    VarPtrs k, u0;
    for (int v = 0; v < NV; v++) {
        k.p[v] = plan->d_wstate + (size_t)v * plan->szpdof;  // the right-hand side, padded layout
        u0.p[v] = plan->d_istage + (size_t)v * ni;
    }
    const RhsParams prm = to_params(params);
    // rkBSSN.cpp:691-762: (a, b, c) of u_next = (a u0 + b u) + c k per stage
    const double A[3] = {1.0, 0.75, 1.0 / 3.0};
    const double Bw[3] = {0.0, 0.25, 2.0 / 3.0};
    const double Cw[3] = {dt, dt * 0.25, dt * (2.0 / 3.0)};
    for (int s = 0; s < 3; s++) {
        if (plan->halo && (rc = launch_padding_refresh(plan, u, st))) return rc;
        if (plan->fused) {
            StageIO io{};
            io.src = u;
            io.rhs = k;
            io.stage = -1;
            io.err = plan->d_err;
            if ((rc = run_fused(plan, io, prm, st))) return rc;
        } else {
            RKArgs rk{};
            if ((rc = run_chunks<MODE_RHS>(plan, u, k, rk, prm, st))) return rc;
        }
        for (const Chunk &c : plan->chunks)
            hipLaunchKernelGGL(k_rk3_combine, dim3(c.nwg), dim3(NTHREADS), 0, st, u, k, u0, plan->d_blocks,
                               plan->d_tiles, plan->d_wg_tile + c.wg_begin, c.tile_end, (int64_t)c.npoints, s, A[s],
                               Bw[s], Cw[s], prm.chi_floor, plan->d_err);
        HIP_OK(hipGetLastError());
    }
    return plan->halo ? launch_padding_refresh(plan, u, st) : DGR_OK;
}

int dgr_rk45_step(dgr_plan *plan, double *const *d_u, double *dt, double tol, const dgr_bssn_params *params,
                  int max_attempts, double *dt_used, double *err_norm, int *attempts, void *stream) {
    g_err.clear();
    if (!plan || !params || !dt) return fail(DGR_E_INVALID, "null plan, params or dt");
    if (!(*dt > 0.0) || !(tol > 0.0) || max_attempts < 1)
        return fail(DGR_E_INVALID, "dgr_rk45_step needs dt > 0, tol > 0 and max_attempts >= 1");
    VarPtrs u;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    if ((rc = ensure_istage(plan))) return rc;
    const size_t ni = (size_t)plan->interior;
    if (!plan->d_rk45 && ni) {
        hipError_t e = hipMalloc(&plan->d_rk45, sizeof(double) * 6 * NV * ni);
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "RK45 stage buffers");
    }
    if (!plan->d_rk45_emax) {
        hipError_t e = hipMalloc(&plan->d_rk45_emax, sizeof(unsigned long long));
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "RK45 error estimate");
    }
    hipStream_t st = (hipStream_t)stream;
    VarPtrs r, u0;
    for (int v = 0; v < NV; v++) {
        r.p[v] = plan->d_wstate + (size_t)v * plan->szpdof;  // the right-hand side, padded layout
        u0.p[v] = plan->d_istage + (size_t)v * ni;
    }
    const RhsParams prm = to_params(params);
    // rkBSSN.h:57-71 as the reference has them (its fourth-order weight of
    // stage 3 is 2197/4101, Fehlberg's 2197/4104)
    static const double U[6][6] = {
        {0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {1.0 / 4.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {3.0 / 32.0, 9.0 / 32.0, 0.0, 0.0, 0.0, 0.0},
        {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0.0, 0.0, 0.0},
        {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0, 0.0, 0.0},
        {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0, 0.0}};
    static const double C5[6] = {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0};
    static const double C4[5] = {25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4101.0, -1.0 / 5.0};
    double h = *dt, en = 0.0;
    int tries = 0;
    for (;;) {
        tries++;
        for (int s = 0; s < 6; s++) {
            if (plan->halo && (rc = launch_padding_refresh(plan, u, st))) return rc;
            if (plan->fused) {
                StageIO io{};
                io.src = u;
                io.rhs = r;
                io.stage = -1;
                io.err = plan->d_err;
                if ((rc = run_fused(plan, io, prm, st))) return rc;
            } else {
                RKArgs rk{};
                if ((rc = run_chunks<MODE_RHS>(plan, u, r, rk, prm, st))) return rc;
            }
            Rk45Coef cf{};
            for (int m = 0; m < 6; m++) cf.a[m] = s < 5 ? U[s + 1][m] * h : 0.0;
            for (int m = 0; m < 5; m++) cf.c4[m] = C4[m] * h;
            for (int m = 0; m < 6; m++) cf.c5[m] = C5[m] * h;
            if (s == 5) hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, plan->d_rk45_emax, 0ull);
            for (const Chunk &c : plan->chunks)
                hipLaunchKernelGGL(k_rk45_stage, dim3(c.nwg), dim3(NTHREADS), 0, st, u, r, u0, plan->d_rk45,
                                   (int64_t)ni, plan->d_blocks, plan->d_tiles, plan->d_wg_tile + c.wg_begin,
                                   c.tile_end, (int64_t)c.npoints, s, cf, prm.chi_floor, plan->d_err,
                                   plan->d_rk45_emax);
            HIP_OK(hipGetLastError());
        }
        unsigned long long bits = 0;
        HIP_OK(hipMemcpyAsync(&bits, plan->d_rk45_emax, sizeof(bits), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        std::memcpy(&en, &bits, sizeof(en));
        // rkBSSN.cpp:1062-1080: retry with a smaller step above the tolerance
        if (en > tol && tries < max_attempts) {
            h = 0.8 * h * std::pow(std::fabs(tol / en), 0.25);
            for (const Chunk &c : plan->chunks)
                hipLaunchKernelGGL(k_rk45_restore, dim3(c.nwg), dim3(NTHREADS), 0, st, u, u0, plan->d_blocks,
                                   plan->d_tiles, plan->d_wg_tile + c.wg_begin, c.tile_end, (int64_t)c.npoints);
            HIP_OK(hipGetLastError());
            continue;
        }
        break;
    }
    if (plan->halo && (rc = launch_padding_refresh(plan, u, st))) return rc;
    if (dt_used) *dt_used = h;
    if (err_norm) *err_norm = en;
    if (attempts) *attempts = tries;
    *dt = en > tol ? 0.8 * h * std::pow(std::fabs(tol / en), 0.25) : 0.8 * h * std::pow(std::fabs(tol / en), 0.20);
    if (en > tol) return fail(DGR_E_STEP, "dgr_rk45_step: the error estimate exceeds tol after max_attempts tries");
    return DGR_OK;
}

int dgr_plan_link_neighbors(dgr_plan *plan, int32_t *nbr, int64_t *nlinks) {
    g_err.clear();
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    HIP_OK(hipSetDevice(plan->device));
    int rc;
    if ((rc = link_neighbors(plan))) return rc;
    if (nbr) std::memcpy(nbr, plan->hnbr.data(), sizeof(int32_t) * plan->hnbr.size());
    if (nlinks) *nlinks = plan->nlinks;
    return DGR_OK;
}

int dgr_halo_fill(dgr_plan *plan, double *const *d_u, void *stream) {
    g_err.clear();
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    VarPtrs u;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    if (plan->nlinks < 0 && (rc = link_neighbors(plan))) return rc;
    return launch_halo_fill(plan, u, (hipStream_t)stream);
}

int dgr_plan_set_halo_refresh(dgr_plan *plan, int mode) {
    g_err.clear();
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    if (mode < DGR_REFRESH_OFF || mode > DGR_REFRESH_FULL) return fail(DGR_E_INVALID, "unknown refresh mode");
    if (mode && !plan->fused) return fail(DGR_E_INVALID, "the padding refresh needs the fused path (DGR_KERNEL=split is set)");
    HIP_OK(hipSetDevice(plan->device));
    int rc;
    if (mode && plan->nlinks < 0 && (rc = link_neighbors(plan))) return rc;
    if (mode == DGR_REFRESH_FULL && plan->nfill < 0 && (rc = build_fill(plan))) return rc;
    plan->halo = mode != DGR_REFRESH_OFF;
    plan->fill_general = mode == DGR_REFRESH_FULL;
    return DGR_OK;
}

int dgr_padding_fill(dgr_plan *plan, double *const *d_u, void *stream) {
    g_err.clear();
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    VarPtrs u;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    if (plan->nlinks < 0 && (rc = link_neighbors(plan))) return rc;
    if (plan->nfill < 0 && (rc = build_fill(plan))) return rc;
    if ((rc = launch_halo_fill(plan, u, (hipStream_t)stream))) return rc;
    return launch_fill_general(plan, u, (hipStream_t)stream);
}

int dgr_plan_fill_sources(dgr_plan *plan, int64_t *ncopy, int64_t *ninterp) {
    g_err.clear();
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    HIP_OK(hipSetDevice(plan->device));
    int rc;
    if (plan->nlinks < 0 && (rc = link_neighbors(plan))) return rc;
    if (plan->nfill < 0 && (rc = build_fill(plan))) return rc;
    if (ncopy) *ncopy = plan->nfill - plan->nfill_interp;
    if (ninterp) *ninterp = plan->nfill_interp;
    return DGR_OK;
}

int dgr_enforce_constraints(dgr_plan *plan, double *const *d_u, double chi_floor, void *stream) {
    g_err.clear();
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    VarPtrs u;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    for (const Chunk &c : plan->chunks)
        hipLaunchKernelGGL(k_enforce, dim3(c.nwg), dim3(NTHREADS), 0, (hipStream_t)stream, u, plan->d_blocks,
                           plan->d_tiles, plan->d_wg_tile + c.wg_begin, c.tile_end, (int64_t)c.npoints, chi_floor,
                           plan->d_err);
    HIP_OK(hipGetLastError());
    return DGR_OK;
}

int dgr_constraints(dgr_plan *plan, double *d_ham, double *const *d_mom, double *const *d_psi4,
                    const double *const *d_u, void *stream) {
    g_err.clear();
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    ConsPtrs o{{d_ham, nullptr, nullptr, nullptr, nullptr, nullptr}};
    int mask = d_ham ? CONS_HAM : 0;
    if (d_mom)
        for (int i = 0; i < 3; i++)
            if ((o.c[1 + i] = d_mom[i])) mask |= CONS_MOM;
    if (d_psi4)
        for (int i = 0; i < 2; i++)
            if ((o.c[4 + i] = d_psi4[i])) mask |= CONS_PSI4;
    if (!mask) return fail(DGR_E_INVALID, "no output: d_ham, every d_mom[i] and every d_psi4[i] are null");
    VarPtrs u;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    hipStream_t st = (hipStream_t)stream;
    if (plan->fused) {
        // one fused kernel per tile kind over the FP64 tile list; no workspace
        StageIO io{};
        io.src = u;
        io.stage = -1;
        const Tile *tiles = plan->d_ftiles[0];
        const int *fg = plan->fgroup[0];
        int n = fg[FG_WIDE + 1] - fg[FG_WIDE];
        if (n)
            hipLaunchKernelGGL((k_cons_fused<false, CfgWide>), dim3(n), dim3(CfgWide::FT), 0, st, io, o,
                               plan->d_blocks, plan->d_bconst, tiles + fg[FG_WIDE], plan->d_tabs, plan->dtab_wide);
        n = fg[FG_THIN + 1] - fg[FG_THIN];
        if (n)
            hipLaunchKernelGGL((k_cons_fused<false, CfgThin>), dim3(n), dim3(CfgThin::FT), 0, st, io, o,
                               plan->d_blocks, plan->d_bconst, tiles + fg[FG_THIN], plan->d_tabs, plan->dtab_thin);
        n = fg[FG_FACE + 1] - fg[FG_FACE];
        if (n)
            hipLaunchKernelGGL((k_cons_fused<true, CfgWide>), dim3(n), dim3(CfgWide::FT), 0, st, io, o,
                               plan->d_blocks, plan->d_bconst, tiles + fg[FG_FACE], plan->d_tabs, plan->dtab_wide);
        HIP_OK(hipGetLastError());
        return DGR_OK;
    }
    if (!plan->d_W && plan->max_chunk_points) {
        hipError_t e = hipMalloc(&plan->d_W, sizeof(double) * NSLOT * (size_t)plan->max_chunk_points);
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "derivative workspace");
    }
    // the smallest instantiated superset of the requested groups
    void (*kern)(VarPtrs, ConsPtrs, const DBlock *, const Tile *, const int *, int, const double *, int64_t) =
        (mask & CONS_PSI4) ? k_constraints<CONS_HAM | CONS_MOM | CONS_PSI4>
        : (mask & CONS_MOM) ? k_constraints<CONS_HAM | CONS_MOM>
                            : k_constraints<CONS_HAM>;
    for (const Chunk &c : plan->chunks) {
        const int ntile = c.tile_end - c.tile_begin;
        hipLaunchKernelGGL(k_derivs, dim3(ntile, NV), dim3(NTHREADS), 0, st, u, plan->d_blocks, plan->d_tiles,
                           c.tile_begin, plan->d_W, (int64_t)c.npoints);
        hipLaunchKernelGGL(kern, dim3(c.nwg), dim3(NTHREADS), 0, st, u, o, plan->d_blocks, plan->d_tiles,
                           plan->d_wg_tile + c.wg_begin, c.tile_end, plan->d_W, (int64_t)c.npoints);
    }
    HIP_OK(hipGetLastError());
    return DGR_OK;
}

int dgr_ham_constraint(dgr_plan *plan, double *d_ham, const double *const *d_u, void *stream) {
    if (!d_ham) {
        g_err.clear();
        return fail(DGR_E_INVALID, "null plan or d_ham");
    }
    return dgr_constraints(plan, d_ham, nullptr, nullptr, d_u, stream);
}

}  // extern "C"

namespace {
// The two norm kernels over every chunk: partials into psum/pcnt (one per
// k_rhs-style workgroup, nwg_total of each), then the sum and count into
// res[0..1] and, with d_norm, the norm into d_norm[0]. Linked neighbours
// (when the table holds links) count each shared node once.
int launch_norm(dgr_plan *plan, const double *d_c, const double *bh, double r1, double r2, double *psum,
                unsigned long long *pcnt, double *res, double *d_norm, hipStream_t st) {
    NormMask m;
    for (int i = 0; i < 6; i++) m.bh[i] = bh[i];
    m.r1 = r1;
    m.r2 = r2;
    const int *nbr = plan->nlinks > 0 ? plan->d_nbr : nullptr;
    for (const Chunk &c : plan->chunks)
        hipLaunchKernelGGL(k_norm_partial, dim3(c.nwg), dim3(NTHREADS), 0, st, d_c, m, plan->d_blocks, plan->d_tiles,
                           plan->d_wg_tile + c.wg_begin, c.tile_end, (int64_t)c.npoints, nbr, psum + c.wg_begin,
                           pcnt + c.wg_begin);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(NTHREADS), 0, st, psum, pcnt, plan->nwg_total, res, d_norm);
    HIP_OK(hipGetLastError());
    return DGR_OK;
}
}  // namespace

extern "C" {

int dgr_constraint_l2_norm(dgr_plan *plan, const double *d_c, const double *bh, double r1, double r2,
                           double *norm, void *stream) {
    g_err.clear();
    if (!plan || !d_c || !bh || !norm) return fail(DGR_E_INVALID, "null argument");
    HIP_OK(hipSetDevice(plan->device));
    if (plan->nlinks < 0) {
        const int rc = link_neighbors(plan);
        if (rc) return rc;
    }
    const int nwg = plan->nwg_total;
    if (!plan->d_norm) {
        hipError_t e = hipMalloc(&plan->d_norm, sizeof(double) * (2 * (size_t)nwg + 2));
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "norm partials");
    }
    double *psum = plan->d_norm;
    auto *pcnt = reinterpret_cast<unsigned long long *>(plan->d_norm + nwg);
    double *res = plan->d_norm + 2 * (size_t)nwg;
    hipStream_t st = (hipStream_t)stream;
    const int rc = launch_norm(plan, d_c, bh, r1, r2, psum, pcnt, res, nullptr, st);
    if (rc) return rc;
    double h[2] = {0.0, 0.0};
    HIP_OK(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    *norm = h[1] > 0.0 ? std::sqrt(h[0] / h[1]) : -1.0;
    return DGR_OK;
}

int dgr_constraint_l2_norm_noalloc(dgr_plan *plan, const double *d_c, const double *bh, double r1, double r2,
                                   double *d_norm, void *stream) {
    g_err.clear();
    if (!plan || !d_c || !bh || !d_norm) return fail(DGR_E_INVALID, "null argument");
    if (plan->nwg_total > NORM_SCRATCH_WG)
        return fail(DGR_E_INVALID, "dgr_constraint_l2_norm_noalloc: the plan needs " + std::to_string(plan->nwg_total) +
                                       " norm workgroups (256 interior points each); the scratch holds " +
                                       std::to_string(NORM_SCRATCH_WG));
    if (plan->nlinks < 0)
        return fail(DGR_E_INVALID, "dgr_constraint_l2_norm_noalloc: the neighbour table is not built; call "
                                   "dgr_plan_link_neighbors once at set-up");
    if (!plan->d_scratch) return fail(DGR_E_INVALID, "dgr_constraint_l2_norm_noalloc: no scratch address");
    HIP_OK(hipSetDevice(plan->device));
    double *psum = plan->d_scratch;
    auto *pcnt = reinterpret_cast<unsigned long long *>(plan->d_scratch + NORM_SCRATCH_WG);
    double *res = plan->d_scratch + 2 * (size_t)NORM_SCRATCH_WG;
    return launch_norm(plan, d_c, bh, r1, r2, psum, pcnt, res, d_norm, (hipStream_t)stream);
}

// Synchronizes `stream` and reports constraint violations met since the last
// call (det(gt) < 0 -> DGR_E_CONSTRAINT; the reference exits there,
// bssn_constraints.h:42-47), then clears the flag.
int dgr_plan_status(dgr_plan *plan, void *stream) {
    g_err.clear();
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    HIP_OK(hipSetDevice(plan->device));
    int h = 0;
    HIP_OK(hipMemcpyAsync(&h, plan->d_err, sizeof(int), hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIP_OK(hipStreamSynchronize((hipStream_t)stream));
    HIP_OK(hipMemsetAsync(plan->d_err, 0, sizeof(int), (hipStream_t)stream));
    HIP_OK(hipStreamSynchronize((hipStream_t)stream));
    if (h & 1) return fail(DGR_E_CONSTRAINT, "negative det(gt) met while enforcing constraints");
    if (h & 6) return fail(DGR_E_CONSTRAINT, "constraint projection exceeded 1e-6 (det(gt)-1 or tr(At))");
    return DGR_OK;
}

}  // extern "C"
#endif  // DGR_F32_KERNELS
