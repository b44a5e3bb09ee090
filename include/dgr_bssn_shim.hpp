// dgr_bssn_shim.hpp — C++ host entry points with the reference's signatures,
// implemented over the C-ABI in dgr_bssn.h, so a Dendro-GR solver can link
// them in place of BSSN_GR/src/rhs.cpp.
//
//   bssnrhs(...)  same signature as BSSN_GR/include/rhs.h:32-36. The variable
//                 pointers are DEVICE pointers (the unzipped arrays live in
//                 HBM); the 24-entry pointer arrays themselves are host arrays.
//                 Runs on the null stream and returns when the result is ready,
//                 as the reference GPU launcher does (cuda/include/
//                 bssn_kernels.cuh:1533-1534).
//   bssnRHS<Block>(...) same argument list as rhs.h:28-30, templated on the
//                 block type (ot::Block lives in dendrolib); `to_phys` maps the
//                 block's octree corner to physical coordinates as
//                 GRIDX_TO_X/GRIDY_TO_Y/GRIDZ_TO_Z do (rhs.cpp:57-67).
//
// Parameters are read from dgr::bssn_params(), the counterpart of the bssn::
// globals the reference reads inside bssnrhs (rhs.cpp:144-168). Where the
// reference has no error path (it returns void), these throw
// std::runtime_error carrying dgr_last_error().
#pragma once

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "dgr_bssn.h"

extern "C" {
/* Process-wide parameters used by the C++ shims (defaults: dgr_default_params). */
dgr_bssn_params *dgr_shim_params(void);
}

void bssnrhs(double **unzipVarsRHS, const double **uZipVars, const unsigned int &offset, const double *pmin,
             const double *pmax, const unsigned int *sz, const unsigned int &bflag, const double t,
             const double **uZipConstVars);

namespace dgr {

inline dgr_bssn_params &bssn_params() { return *dgr_shim_params(); }

// Runs every block through one plan (one launch sequence instead of the
// reference's per-block loop, rhs.cpp:49-71). `szpdof` is the length of each
// unzipped variable array. ToPhys: (const Block&, double pmin[3], double pmax[3],
// double dx[3]) filling the block's physical corner coordinates and spacing.
template <typename Block, typename ToPhys>
void bssnRHS(double **uzipVarsRHS, const double **uZipVars, const Block *blkList, unsigned int numBlocks,
             double curr_time, uint64_t szpdof, ToPhys to_phys) {
    const unsigned PW = DGR_PADDING_WIDTH;
    std::vector<dgr_block> blocks(numBlocks);
    for (unsigned b = 0; b < numBlocks; b++) {
        const Block &blk = blkList[b];
        dgr_block &d = blocks[b];
        d.offset = blk.getOffset();
        d.sz[0] = blk.getAllocationSzX();
        d.sz[1] = blk.getAllocationSzY();
        d.sz[2] = blk.getAllocationSzZ();
        d.bflag = blk.getBlkNodeFlag();
        double lo[3], hi[3], dx[3];
        to_phys(blk, lo, hi, dx);
        for (int a = 0; a < 3; a++) {  // rhs.cpp:61-67
            d.pmin[a] = lo[a] - PW * dx[a];
            d.pmax[a] = hi[a] + PW * dx[a];
        }
    }
    dgr_plan *plan = nullptr;
    int rc = dgr_plan_create(0, blocks.data(), numBlocks, szpdof, &plan);
    if (rc == DGR_OK) rc = dgr_bssn_rhs(plan, uzipVarsRHS, uZipVars, curr_time, &bssn_params(), nullptr);
    if (rc == DGR_OK) rc = dgr_plan_status(plan, nullptr);
    std::string err = rc == DGR_OK ? "" : dgr_last_error();
    dgr_plan_destroy(plan);
    if (rc != DGR_OK) throw std::runtime_error("bssnRHS: " + err);
}

}  // namespace dgr
