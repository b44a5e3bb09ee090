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
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
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
    int q0;     // first workspace point of the tile within its chunk
    int o[3];   // absolute index of the tile's first point along each axis
    int e[3];   // tile extents
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
};

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
    v14[0] = fmax(v14[0], chi_floor);
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
    const double eta = rit_eta(x, y, z, prm);

    PointIn in{&u, g.idx, W, P, q};
    PointOut<MODE> out;
    out.rhs = &rhs;
    out.rk = &rk;
    out.idx = g.idx;
    out.W = W;
    out.P = P;
    out.q = q;
    ko_sigmas(prm, u, g.idx, out.sig_gauge, out.sig_other);
    dgr::bssn_rhs_point<double>(in, eta, prm.lam, prm.lamf, out);
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
// One workgroup per tile; the 24 variables stream through LDS in the phase
// order of bssn_rhs_phased.h: while the equations of phase v run, the next
// variable's region is already in flight into registers. Each thread owns one
// point (threads beyond the tile's points shadow the last point and store
// nothing), computes its derivatives of the resident variable, and leaves that
// variable's Kreiss-Oliger term (on flagged faces: boundary value plus KO term)
// in LDS for the final stores.
// -------------------------------------------------------------------------
// One wave per SIMD: the equations need ~430 registers per lane (256 VGPRs
// plus AGPRs used as spill space), which only a single wave per SIMD has.
constexpr int FT = 256;                                   // threads per workgroup
constexpr int FZMAX = 5;                                  // max tile extent along z
constexpr int TPTS = FT;                                  // max points per tile (7 x 7 x 5 = 245)
constexpr int RPTS = RMAX * RMAX * (FZMAX + 2 * PW);      // staged region points
constexpr int DXPTS = TMAX * RMAX * (FZMAX + 2 * PW);     // D_x region
constexpr int DYPTS = TMAX * TMAX * (FZMAX + 2 * PW);     // D_y region
constexpr int PREF = (RPTS + FT - 1) / FT;                // prefetch registers per thread

// Where a stage reads its state and writes its results. Stage buffers hold
// interior points only, indexed by DBlock::q0 + (i-3) + mx*((j-3) + my*(k-3));
// padding always comes from the caller's state `u`.
struct StageIO {
    VarPtrs u;     // caller's padded state (padding; interior at stage 0 and in RHS mode)
    VarPtrs sin;   // interior-only stage input (stages 1..3)
    VarPtrs sout;  // interior-only next-stage state (stages 0..2)
    VarPtrs acc;   // interior-only accumulator
    VarPtrs rhs;   // RHS mode: padded output
    double cacc, cnext;
    int stage;     // -1 in RHS mode
    int enforce_stage;
    int *err;
};

struct FusedCtx {
    // tile geometry
    int nx, ny, nz, rx, ry, nreg, ex, ey;
    int gx0, gy0, gz0;
    uint64_t boff;
    int64_t bq0;
    int mx, my;
    bool lox, hix, loy, hiy, loz, hiz;
    double h[3];
    // this thread's point
    int p, li, lj, lk, ax, ay, az;
    bool active, face;
    uint64_t idx;
    int64_t q;
    double x, y, z, inv_r, sig_gauge, sig_other;
    // sources and LDS
    const StageIO *io;
    double *s_tile, *s_dx, *s_dy, *s_ko;
    double pre[PREF];
    double d1v[3], d2v[6];

    __device__ __forceinline__ void fetch(int v) {
        const bool from_stage = io->stage >= 1;
        const double *__restrict__ up = io->u.p[v];
        const double *__restrict__ sp_ = from_stage ? io->sin.p[v] : nullptr;
#pragma unroll
        for (int c = 0; c < PREF; c++) {
            const int e = threadIdx.x + c * FT;
            if (e < nreg) {
                const int ii = e % rx, jj = (e / rx) % ry, kk = e / (rx * ry);
                const int gi = gx0 + ii, gj = gy0 + jj, gk = gz0 + kk;
                const bool inner = gi >= PW && gi < nx - PW && gj >= PW && gj < ny - PW && gk >= PW && gk < nz - PW;
                if (from_stage && inner)
                    pre[c] = sp_[bq0 + (gi - PW) + (int64_t)mx * ((gj - PW) + (int64_t)my * (gk - PW))];
                else
                    pre[c] = up[boff + (uint64_t)gi + (uint64_t)nx * ((uint64_t)gj + (uint64_t)ny * gk)];
            }
        }
    }

    __device__ __forceinline__ void stash() {
#pragma unroll
        for (int c = 0; c < PREF; c++) {
            const int e = threadIdx.x + c * FT;
            if (e < nreg) s_tile[e] = pre[c];
        }
    }

    template <int V, int NEXT>
    __device__ __forceinline__ void phase() {
        __syncthreads();  // everyone is done with the previous variable's LDS data
        stash();
        if constexpr (NEXT >= 0) fetch(NEXT);
        __syncthreads();
        constexpr int m2 = d2slot(V);
        const dgr::Axis1<double> c1x(h[0]), c1y(h[1]), c1z(h[2]);
        const int ez = nreg / (rx * ry) - 2 * PW;
        if constexpr (m2 >= 0) {
            const int ndx = ex * ry * (ez + 2 * PW);
            for (int e = threadIdx.x; e < ndx; e += FT) {
                const int i0 = e % ex, jj = (e / ex) % ry, kk = e / (ex * ry);
                s_dx[e] = dgr::d1(&s_tile[(PW + i0) + rx * (jj + ry * kk)], 1, gx0 + PW + i0, nx, lox, hix, c1x);
            }
            const int ndy = ex * ey * (ez + 2 * PW);
            for (int e = threadIdx.x; e < ndy; e += FT) {
                const int i0 = e % ex, j0 = (e / ex) % ey, kk = e / (ex * ey);
                s_dy[e] = dgr::d1(&s_tile[(PW + i0) + rx * ((PW + j0) + ry * kk)], rx, gy0 + PW + j0, ny, loy, hiy, c1y);
            }
            __syncthreads();
        }
        const double *c = &s_tile[(PW + li) + rx * ((PW + lj) + ry * (PW + lk))];
        if constexpr (m2 >= 0) {
            const double *cx = &s_dx[li + ex * ((PW + lj) + ry * (PW + lk))];
            const double *cy = &s_dy[li + ex * (lj + ey * (PW + lk))];
            d1v[0] = cx[0];
            d1v[1] = cy[0];
            const dgr::Axis2<double> c2x(h[0]), c2y(h[1]), c2z(h[2]);
            d2v[0] = dgr::d2(c, 1, ax, nx, lox, hix, c2x);
            d2v[1] = dgr::d1(cx, ex, ay, ny, loy, hiy, c1y);        // D_y(D_x u)
            d2v[2] = dgr::d1(cx, ex * ry, az, nz, loz, hiz, c1z);   // D_z(D_x u)
            d2v[3] = dgr::d2(c, rx, ay, ny, loy, hiy, c2y);
            d2v[4] = dgr::d1(cy, ex * ey, az, nz, loz, hiz, c1z);   // D_z(D_y u)
            d2v[5] = dgr::d2(c, rx * ry, az, nz, loz, hiz, c2z);
        } else {
            d1v[0] = dgr::d1(c, 1, ax, nx, lox, hix, c1x);
            d1v[1] = dgr::d1(c, rx, ay, ny, loy, hiy, c1y);
        }
        d1v[2] = dgr::d1(c, rx * ry, az, nz, loz, hiz, c1z);
        // computed here, not at their last use (that would keep the LDS taps live)
        DGR_PIN(d1v[0]);
        DGR_PIN(d1v[1]);
        DGR_PIN(d1v[2]);
        if constexpr (m2 >= 0) {
#pragma unroll
            for (int k = 0; k < 6; k++) DGR_PIN(d2v[k]);
        }
        ko_term<V>(c);
    }

    // KO term sigma*(ko_x+ko_y+ko_z) (rhs.cpp:434-489); on a flagged face the
    // boundary value replaces the equations (bssn_bcs, rhs.cpp:511-631) and the
    // KO term is added to it, as the reference adds KO after the boundary pass.
    template <int V>
    __device__ __forceinline__ void ko_term(const double *c) {
#pragma clang fp contract(off)
        const dgr::AxisKO<double> kx(h[0]), ky(h[1]), kz(h[2]);
        const double kox = dgr::ko(c, 1, ax, nx, lox, hix, kx);
        const double koy = dgr::ko(c, rx, ay, ny, loy, hiy, ky);
        const double koz = dgr::ko(c, rx * ry, az, nz, loz, hiz, kz);
        const double kt = (is_gauge(V) ? sig_gauge : sig_other) * (kox + koy + koz);
        double v = kt;
        if (face) {
            const double bc = -inv_r * (x * d1v[0] + y * d1v[1] + z * d1v[2] + BC_FALLOFF[V] * (c[0] - BC_ASYM[V]));
            v = bc + kt;
        }
        if (active) s_ko[V * TPTS + p] = v;
    }

    template <int V, int D>
    __device__ __forceinline__ double d1() const { return d1v[D]; }
    template <int M, int C>
    __device__ __forceinline__ double d2() const { return d2v[C]; }
    template <int V>
    __device__ __forceinline__ double fld() const {
        return io->stage >= 1 ? io->sin.p[V][q] : io->u.p[V][idx];
    }
    __device__ __forceinline__ void finish() {}
};

struct FusedOut {
    const FusedCtx *ctx;
    const StageIO *io;
    double v14[14];

    template <int V>
    __device__ __forceinline__ void put(double val) {
#pragma clang fp contract(off)
        const double kb = ctx->s_ko[V * TPTS + ctx->p];
        val = ctx->face ? kb : val + kb;
        if (!ctx->active) return;
        const uint64_t idx = ctx->idx;
        const int64_t q = ctx->q;
        const int s = io->stage;
        if (s < 0) {
            io->rhs.p[V][idx] = val;
        } else if (s < 3) {
            const double u0 = io->u.p[V][idx];
            const double nxt = u0 + io->cnext * val;
            io->acc.p[V][q] = (s == 0 ? u0 : io->acc.p[V][q]) + io->cacc * val;
            if (slot14(V) >= 0 && io->enforce_stage) v14[slot14(V)] = nxt;
            else io->sout.p[V][q] = nxt;
        } else {
            const double fin = io->acc.p[V][q] + io->cacc * val;
            if (slot14(V) >= 0) v14[slot14(V)] = fin;
            else io->u.p[V][idx] = fin;
        }
    }

    __device__ __forceinline__ void finish(double chi_floor) {
        const int s = io->stage;
        if (!ctx->active || s < 0 || (s < 3 && !io->enforce_stage)) return;
        const int e = enforce_point(v14, chi_floor);
        if (e) atomicOr(io->err, e);
        if (s == 3) {
            const uint64_t idx = ctx->idx;
            io->u.p[0][idx] = v14[0];
            io->u.p[1][idx] = v14[1];
#pragma unroll
            for (int c = 0; c < 12; c++) io->u.p[12 + c][idx] = v14[2 + c];
        } else {
            const int64_t q = ctx->q;
            io->sout.p[0][q] = v14[0];
            io->sout.p[1][q] = v14[1];
#pragma unroll
            for (int c = 0; c < 12; c++) io->sout.p[12 + c][q] = v14[2 + c];
        }
    }
};

__global__ __launch_bounds__(FT) void k_fused(StageIO io, RhsParams prm, const DBlock *__restrict__ blocks,
                                             const Tile *__restrict__ tiles) {
    __shared__ double s_tile[RPTS];
    __shared__ double s_dx[DXPTS];
    __shared__ double s_dy[DYPTS];
    __shared__ double s_ko[NV * TPTS];

    const Tile T = tiles[blockIdx.x];
    const DBlock B = blocks[T.block];
    FusedCtx ctx;
    ctx.io = &io;
    ctx.s_tile = s_tile;
    ctx.s_dx = s_dx;
    ctx.s_dy = s_dy;
    ctx.s_ko = s_ko;
    ctx.nx = B.n[0];
    ctx.ny = B.n[1];
    ctx.nz = B.n[2];
    ctx.ex = T.e[0];
    ctx.ey = T.e[1];
    ctx.rx = T.e[0] + 2 * PW;
    ctx.ry = T.e[1] + 2 * PW;
    ctx.nreg = ctx.rx * ctx.ry * (T.e[2] + 2 * PW);
    ctx.gx0 = T.o[0] - PW;
    ctx.gy0 = T.o[1] - PW;
    ctx.gz0 = T.o[2] - PW;
    ctx.boff = B.offset;
    ctx.bq0 = B.q0;
    ctx.mx = B.n[0] - 2 * PW;
    ctx.my = B.n[1] - 2 * PW;
    const uint32_t bf = B.bflag;
    ctx.lox = bf & 1u; ctx.hix = bf & 2u; ctx.loy = bf & 4u; ctx.hiy = bf & 8u; ctx.loz = bf & 16u; ctx.hiz = bf & 32u;
    ctx.h[0] = B.h[0];
    ctx.h[1] = B.h[1];
    ctx.h[2] = B.h[2];
    const int npts = T.e[0] * T.e[1] * T.e[2];
    ctx.active = (int)threadIdx.x < npts;
    ctx.p = ctx.active ? (int)threadIdx.x : npts - 1;
    ctx.li = ctx.p % T.e[0];
    ctx.lj = (ctx.p / T.e[0]) % T.e[1];
    ctx.lk = ctx.p / (T.e[0] * T.e[1]);
    ctx.ax = T.o[0] + ctx.li;
    ctx.ay = T.o[1] + ctx.lj;
    ctx.az = T.o[2] + ctx.lk;
    ctx.idx = B.offset + (uint64_t)ctx.ax + (uint64_t)ctx.nx * ((uint64_t)ctx.ay + (uint64_t)ctx.ny * ctx.az);
    ctx.q = B.q0 + (ctx.ax - PW) + (int64_t)ctx.mx * ((ctx.ay - PW) + (int64_t)ctx.my * (ctx.az - PW));
    ctx.face = (ctx.lox && ctx.ax == PW) || (ctx.hix && ctx.ax == ctx.nx - PW - 1) || (ctx.loy && ctx.ay == PW) ||
               (ctx.hiy && ctx.ay == ctx.ny - PW - 1) || (ctx.loz && ctx.az == PW) || (ctx.hiz && ctx.az == ctx.nz - PW - 1);
    ctx.x = coord(B.pmin[0], ctx.ax, B.h[0]);
    ctx.y = coord(B.pmin[1], ctx.ay, B.h[1]);
    ctx.z = coord(B.pmin[2], ctx.az, B.h[2]);
    ctx.inv_r = inv_radius(ctx.x, ctx.y, ctx.z);
    if (prm.cako) {
        const double sg = sqrt(ctx.fld<1>());
        ctx.sig_gauge = sg * prm.eps_gauge;
        ctx.sig_other = sg * prm.eps_other;
    } else {
        ctx.sig_gauge = ctx.sig_other = prm.ko_sigma;
    }
    const double eta = rit_eta(ctx.x, ctx.y, ctx.z, prm);
    ctx.fetch(dgr::kBssnFirstPhaseVar);
    FusedOut out;
    out.ctx = &ctx;
    out.io = &io;
    dgr::bssn_rhs_phased<double>(ctx, eta, prm.lam, prm.lamf, out);
    out.finish(prm.chi_floor);
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
    bool fused = false;         // k_fused (DGR_KERNEL=fused) or the split k_derivs/k_rhs path (default)
    Tile *d_ftiles = nullptr;   // k_fused tiles (<= 7 x 7 x 5 points), XCD-grouped order
    int nftiles = 0;
    double *d_istage = nullptr; // k_fused RK4: 3 * 24 * interior (two stage states, accumulator)
    int *d_err = nullptr;
    // optional per-kernel timing: event pairs recorded around each launch
    bool timing = false;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[4];

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
}

const char *dgr_last_error(void) { return g_err.c_str(); }

// m interior points -> ceil(m/emax) near-equal parts of at most emax points
static void split_axis_max(int m, int emax, std::vector<std::pair<int, int>> &parts) {
    const int np = (m + emax - 1) / emax;
    int o = PW;
    for (int p = 0; p < np; p++) {
        const int e = m / np + (p < m % np ? 1 : 0);
        parts.push_back({o, e});
        o += e;
    }
}

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
                    Tile t;
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

    // k_fused tiles: at most 7 x 7 x 5 points (one point per thread of a 256-thread
    // workgroup). Blocks are taken eight at a time and their tiles interleaved, so
    // that (with workgroups dealt round-robin over the 8 XCDs) the tiles of one
    // block share an XCD and its L2 for their overlapping halos.
    std::vector<Tile> ftiles;
    {
        std::vector<std::vector<Tile>> per(nblocks);
        for (uint32_t b = 0; b < nblocks; b++) {
            const DBlock &d = hb[b];
            std::vector<std::pair<int, int>> px, py, pz;
            split_axis_max(d.n[0] - 6, TMAX, px);
            split_axis_max(d.n[1] - 6, TMAX, py);
            split_axis_max(d.n[2] - 6, FZMAX, pz);
            for (auto &z : pz)
                for (auto &y : py)
                    for (auto &x : px) {
                        Tile t;
                        t.block = (int)b;
                        t.q0 = 0;
                        t.o[0] = x.first; t.o[1] = y.first; t.o[2] = z.first;
                        t.e[0] = x.second; t.e[1] = y.second; t.e[2] = z.second;
                        per[b].push_back(t);
                    }
        }
        for (uint32_t g = 0; g < nblocks; g += 8) {
            const uint32_t ge = std::min<uint32_t>(nblocks, g + 8);
            size_t mx = 0;
            for (uint32_t b = g; b < ge; b++) mx = std::max(mx, per[b].size());
            for (size_t t = 0; t < mx; t++)
                for (uint32_t b = g; b < ge; b++)
                    if (t < per[b].size()) ftiles.push_back(per[b][t]);
        }
    }
    plan->nftiles = (int)ftiles.size();
    if (const char *k = std::getenv("DGR_KERNEL")) plan->fused = std::string(k) == "fused";

    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess && nblocks) e = hipMalloc(&plan->d_blocks, sizeof(DBlock) * hb.size());
    if (e == hipSuccess && !tiles.empty()) e = hipMalloc(&plan->d_tiles, sizeof(Tile) * tiles.size());
    if (e == hipSuccess && !wg_tile.empty()) e = hipMalloc(&plan->d_wg_tile, sizeof(int) * wg_tile.size());
    if (e == hipSuccess && !face_tiles.empty()) e = hipMalloc(&plan->d_face_tiles, sizeof(int) * face_tiles.size());
    if (e == hipSuccess && !ftiles.empty()) e = hipMalloc(&plan->d_ftiles, sizeof(Tile) * ftiles.size());
    if (e == hipSuccess) e = hipMalloc(&plan->d_err, sizeof(int));
    if (e == hipSuccess && nblocks) e = hipMemcpy(plan->d_blocks, hb.data(), sizeof(DBlock) * hb.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !tiles.empty()) e = hipMemcpy(plan->d_tiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !wg_tile.empty()) e = hipMemcpy(plan->d_wg_tile, wg_tile.data(), sizeof(int) * wg_tile.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !face_tiles.empty())
        e = hipMemcpy(plan->d_face_tiles, face_tiles.data(), sizeof(int) * face_tiles.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !ftiles.empty())
        e = hipMemcpy(plan->d_ftiles, ftiles.data(), sizeof(Tile) * ftiles.size(), hipMemcpyHostToDevice);
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
    hipSetDevice(plan->device);
    for (auto &v : plan->pending)
        for (auto &pr : v) {
            hipEventDestroy(pr.first);
            hipEventDestroy(pr.second);
        }
    for (hipEvent_t e : plan->event_pool) hipEventDestroy(e);
    hipFree(plan->d_blocks);
    hipFree(plan->d_tiles);
    hipFree(plan->d_wg_tile);
    hipFree(plan->d_face_tiles);
    hipFree(plan->d_W);
    hipFree(plan->d_stage);
    hipFree(plan->d_ftiles);
    hipFree(plan->d_istage);
    hipFree(plan->d_err);
    delete plan;
    return DGR_OK;
}

int64_t dgr_plan_interior_points(const dgr_plan *plan) { return plan ? plan->interior : -1; }

int dgr_plan_set_timing(dgr_plan *plan, int enable) {
    if (!plan) return fail(DGR_E_INVALID, "null plan");
    plan->timing = enable != 0;
    return DGR_OK;
}

int dgr_plan_kernel_times(dgr_plan *plan, double *ms, int64_t *launches) {
    if (!plan || !ms || !launches) return fail(DGR_E_INVALID, "null argument");
    HIP_OK(hipSetDevice(plan->device));
    for (int k = 0; k < 4; k++) {
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

RhsParams to_params(const dgr_bssn_params *p) {
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
    return r;
}

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
    if (a) hipEventRecord(a, st);
    launch();
    if (b) hipEventRecord(b, st);
    if (a && b) plan->pending[kind].push_back({a, b});
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

int run_fused(dgr_plan *plan, const StageIO &io, const RhsParams &prm, hipStream_t st) {
    if (plan->nftiles == 0) return DGR_OK;
    timed(plan, 3, st, [&] {
        hipLaunchKernelGGL(k_fused, dim3(plan->nftiles), dim3(FT), 0, st, io, prm, plan->d_blocks, plan->d_ftiles);
    });
    HIP_OK(hipGetLastError());
    return DGR_OK;
}

}  // namespace

extern "C" {

int dgr_bssn_rhs(dgr_plan *plan, double *const *d_rhs, const double *const *d_u, double t,
                 const dgr_bssn_params *params, void *stream) {
    (void)t;
    g_err.clear();
    if (!plan || !params) return fail(DGR_E_INVALID, "null plan or params");
    VarPtrs u, r;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u")) || (rc = load_ptrs(r, d_rhs, "d_rhs"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    if (plan->fused) {
        StageIO io{};
        io.u = u;
        io.rhs = r;
        io.stage = -1;
        io.err = plan->d_err;
        return run_fused(plan, io, to_params(params), (hipStream_t)stream);
    }
    RKArgs rk{};
    return run_chunks<MODE_RHS>(plan, u, r, rk, to_params(params), (hipStream_t)stream);
}

int dgr_rk4_step(dgr_plan *plan, double *const *d_u, double dt, const dgr_bssn_params *params,
                 int enforce_per_stage, void *stream) {
    g_err.clear();
    if (!plan || !params) return fail(DGR_E_INVALID, "null plan or params");
    VarPtrs u;
    int rc;
    if ((rc = load_ptrs(u, d_u, "d_u"))) return rc;
    HIP_OK(hipSetDevice(plan->device));
    hipStream_t st = (hipStream_t)stream;
    // RK4 table: include/rkBSSN.h:50-53
    static const double RKC[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    static const double RKU[4] = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};
    if (plan->fused) {
        const size_t ni = (size_t)plan->interior;
        if (!plan->d_istage && ni) {
            hipError_t e = hipMalloc(&plan->d_istage, sizeof(double) * 3 * NV * ni);
            if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? DGR_E_NOMEM : DGR_E_HIP, "stage buffers");
        }
        VarPtrs sa, sb, acc;
        for (int v = 0; v < NV; v++) {
            sa.p[v] = plan->d_istage + (size_t)v * ni;
            sb.p[v] = plan->d_istage + (size_t)(NV + v) * ni;
            acc.p[v] = plan->d_istage + (size_t)(2 * NV + v) * ni;
        }
        const RhsParams prm = to_params(params);
        // stage s reads u (s = 0) or the previous stage buffer, writes the other one;
        // the last stage writes the result into u (rkBSSN.cpp:764-860)
        const VarPtrs *ins[4] = {&sa, &sa, &sb, &sa};
        const VarPtrs *outs[4] = {&sa, &sb, &sa, &sb};
        for (int s = 0; s < 4; s++) {
            StageIO io{};
            io.u = u;
            io.sin = *ins[s];
            io.sout = *outs[s];
            io.acc = acc;
            io.cacc = RKC[s] * dt;
            io.cnext = s < 3 ? RKU[s + 1] * dt : 0.0;
            io.stage = s;
            io.enforce_stage = enforce_per_stage ? 1 : 0;
            io.err = plan->d_err;
            if ((rc = run_fused(plan, io, prm, st))) return rc;
        }
        return DGR_OK;
    }
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
