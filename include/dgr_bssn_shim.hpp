// dgr_bssn_shim.hpp — C++ host entry points with the reference's signatures,
// implemented over the C-ABI in dgr_bssn.h, so a Dendro-GR solver can link
// them in place of BSSN_GR/src/rhs.cpp.
//
//   bssnRHS(...)  exactly the argument list of BSSN_GR/include/rhs.h:28-30,
//                 templated on the block type (ot::Block lives in dendrolib).
//                 Per block it derives the offset, sizes, face flags and the
//                 padded corner coordinates as rhs.cpp:49-67 does
//                 (GRIDX_TO_X of the block node's corners -/+ PW dx, dx from
//                 Block::computeDx), from the domain set once with
//                 dgr_shim_set_domain. All blocks then run in one plan; the
//                 plan is kept and reused while the block list is unchanged
//                 (one plan per remesh, not per call).
//   bssnrhs(...)  exactly the signature of BSSN_GR/include/rhs.h:32-36, one
//                 block per call; one plan per block shape (sizes, face flags)
//                 is kept and moved to each call's block (dgr_plan_move_block).
//
// Variable pointers are DEVICE pointers (the unzipped arrays live in HBM); the
// 24-entry pointer arrays themselves are host arrays. Both functions run on
// the null stream and return when the result is ready, as the reference GPU
// launcher does (cuda/include/bssn_kernels.cuh:1533-1534). Parameters are
// read from dgr::bssn_params(), the counterpart of the bssn:: globals the
// reference reads inside bssnrhs (rhs.cpp:144-168). Where the reference has no
// error path (it returns void), these throw std::runtime_error carrying
// dgr_last_error(). The cached plans are process-wide and not thread-safe,
// like the reference's global derivative workspace (grUtils.cpp:2374-2396).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "dgr_bssn.h"

extern "C" {
/* Process-wide parameters used by the C++ shims (defaults: dgr_default_params). */
dgr_bssn_params *dgr_shim_params(void);

/* The computational domain and octree extent bssnRHS maps block nodes with
 * (the reference's bssn::BSSN_COMPD_MIN/MAX and BSSN_OCTREE_MIN/MAX,
 * parameters.cpp:106-110, used by GRIDX_TO_X, grDef.h:16-21), and the length
 * of each unzipped variable array (the reference's mesh->getDegOfFreedomUnZip(),
 * which its bssnRHS does not need because it never checks bounds). Set once
 * per run and again after a remesh changes szpdof. Drops the cached plans. */
void dgr_shim_set_domain(const double *compd_min, const double *compd_max, const double *octree_min,
                         const double *octree_max, uint64_t szpdof);

/* The plan-cached calls behind bssnRHS and bssnrhs: right-hand side of the
 * given blocks, synchronous on the null stream. dgr_shim_rhs keeps one plan for
 * the most recent block table (rebuilt when the table or szpdof changes);
 * dgr_shim_rhs_block keeps one plan per block shape and moves it to `blk`
 * (szpdof: the domain's, or offset + block size when none was set). */
int dgr_shim_rhs(const dgr_block *blocks, uint32_t nblocks, double *const *d_rhs, const double *const *d_u, double t);
int dgr_shim_rhs_block(const dgr_block *blk, double *const *d_rhs, const double *const *d_u, double t);

/* Plans the shims have built and calls they have served since the process
 * started (a steady state builds none). */
void dgr_shim_cache_stats(int64_t *plans_built, int64_t *calls);

/* Destroys the cached plans (e.g. before a remesh frees the arrays). */
void dgr_shim_release(void);

/* The domain set by dgr_shim_set_domain (set = 0 before the first call). */
typedef struct dgr_shim_domain_t {
    double compd_min[3], compd_max[3], octree_min[3], octree_max[3];
    uint64_t szpdof;
    int32_t set;
} dgr_shim_domain_t;
const dgr_shim_domain_t *dgr_shim_domain(void);
}

void bssnrhs(double **unzipVarsRHS, const double **uZipVars, const unsigned int &offset, const double *pmin,
             const double *pmax, const unsigned int *sz, const unsigned int &bflag, const double t,
             const double **uZipConstVars);

namespace dgr {

inline dgr_bssn_params &bssn_params() { return *dgr_shim_params(); }

namespace detail {
// The point type of Block::computeDx(const Point&, const Point&) const.
template <class B, class P, class R>
P point_arg(R (B::*)(const P &, const P &) const);
template <class Block>
using PointOf = decltype(point_arg(&Block::computeDx));
}  // namespace detail

}  // namespace dgr

// bssnRHS with the argument list of rhs.h:28-30. uZipConstVars is accepted for
// that list only: this gauge's equations do not read constraint variables.
template <typename Block>
void bssnRHS(double **uzipVarsRHS, const double **uZipVars, const Block *blkList, unsigned int numBlocks,
             const double curr_time, const double **uZipConstVars) {
    (void)uZipConstVars;
    const dgr_shim_domain_t &D = *dgr_shim_domain();
    if (!D.set) throw std::runtime_error("bssnRHS: call dgr_shim_set_domain first");
    using Point = dgr::detail::PointOf<Block>;
    const Point pt_min(D.compd_min[0], D.compd_min[1], D.compd_min[2]);
    const Point pt_max(D.compd_max[0], D.compd_max[1], D.compd_max[2]);
    const unsigned PW = DGR_PADDING_WIDTH;
    // GRIDX_TO_X of grDef.h:16-21: (R / Rg) * (xg - octree_min) + compd_min
    auto grid_to_phys = [&D](int a, double g) {
        return ((D.compd_max[a] - D.compd_min[a]) / (D.octree_max[a] - D.octree_min[a])) * (g - D.octree_min[a]) +
               D.compd_min[a];
    };
    std::vector<dgr_block> blocks(numBlocks);
    for (unsigned b = 0; b < numBlocks; b++) {  // rhs.cpp:49-67
        const Block &blk = blkList[b];
        dgr_block &d = blocks[b];
        d.offset = blk.getOffset();
        d.sz[0] = blk.getAllocationSzX();
        d.sz[1] = blk.getAllocationSzY();
        d.sz[2] = blk.getAllocationSzZ();
        d.bflag = blk.getBlkNodeFlag();
        const double dx[3] = {blk.computeDx(pt_min, pt_max), blk.computeDy(pt_min, pt_max),
                              blk.computeDz(pt_min, pt_max)};
        const auto &node = blk.getBlockNode();
        const double lo[3] = {(double)node.minX(), (double)node.minY(), (double)node.minZ()};
        const double hi[3] = {(double)node.maxX(), (double)node.maxY(), (double)node.maxZ()};
        for (int a = 0; a < 3; a++) {
            d.pmin[a] = grid_to_phys(a, lo[a]) - PW * dx[a];
            d.pmax[a] = grid_to_phys(a, hi[a]) + PW * dx[a];
        }
    }
    if (dgr_shim_rhs(blocks.data(), numBlocks, uzipVarsRHS, uZipVars, curr_time) != DGR_OK)
        throw std::runtime_error(std::string("bssnRHS: ") + dgr_last_error());
}
