// Host C++ entry point with the reference's bssnrhs() signature
// (BSSN_GR/include/rhs.h:32-36) over the C-ABI. See dgr_bssn_shim.hpp.
#include <cstdint>
#include <stdexcept>
#include <string>

#include "dgr_bssn.h"
#include "dgr_bssn_shim.hpp"

namespace {
struct ShimParams {
    dgr_bssn_params p;
    ShimParams() { dgr_default_params(&p); }
};
ShimParams g_params;
}  // namespace

extern "C" dgr_bssn_params *dgr_shim_params(void) { return &g_params.p; }

// One block at `offset` of the unzipped arrays: builds a one-block plan whose
// szpdof ends at the block (the arrays are at least that long), runs it on the
// null stream and waits, like a per-block call of the reference.
void bssnrhs(double **unzipVarsRHS, const double **uZipVars, const unsigned int &offset, const double *pmin,
             const double *pmax, const unsigned int *sz, const unsigned int &bflag, const double t,
             const double **uZipConstVars) {
    (void)uZipConstVars;  // constraint arrays: read only by CAHD variants, not this gauge
    dgr_block b;
    b.offset = offset;
    for (int a = 0; a < 3; a++) {
        b.sz[a] = sz[a];
        b.pmin[a] = pmin[a];
        b.pmax[a] = pmax[a];
    }
    b.bflag = bflag;
    const uint64_t szpdof = (uint64_t)offset + (uint64_t)sz[0] * sz[1] * sz[2];
    dgr_plan *plan = nullptr;
    int rc = dgr_plan_create(0, &b, 1, szpdof, &plan);
    if (rc == DGR_OK) rc = dgr_bssn_rhs(plan, unzipVarsRHS, uZipVars, t, &g_params.p, nullptr);
    if (rc == DGR_OK) rc = dgr_plan_status(plan, nullptr);
    std::string err = rc == DGR_OK ? "" : dgr_last_error();
    dgr_plan_destroy(plan);
    if (rc != DGR_OK) throw std::runtime_error("bssnrhs: " + err);
}
