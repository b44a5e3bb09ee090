// Host C++ entry points with the reference's signatures over the C-ABI
// (BSSN_GR/include/rhs.h:28-36). See dgr_bssn_shim.hpp.
//
// Plan reuse: the reference calls bssnRHS once per RK stage with the same
// block list until a remesh (BSSN_GR/src/bssnCtx.cpp:96-97), and its bssnrhs
// once per block (rhs.cpp:49-71). Building a plan costs device allocations and
// table uploads, so neither shim builds one per call: bssnRHS keeps the plan of
// the most recent block table, bssnrhs one plan per block shape, moved to each
// call's block.
#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "dgr_bssn.h"
#include "dgr_bssn_shim.hpp"

namespace {

struct ShimState {
    dgr_bssn_params p;
    dgr_shim_domain_t dom{};
    // bssnRHS: the plan of the last block table
    std::vector<dgr_block> table;
    uint64_t table_szpdof = 0;
    dgr_plan *plan = nullptr;
    // bssnrhs: one plan per (sizes, face flags)
    std::map<std::array<uint32_t, 4>, dgr_plan *> shape_plans;
    int64_t plans_built = 0, calls = 0;

    // no destructor: the plans live until dgr_shim_release or the process
    // ends (freeing device memory from a static destructor could run after
    // the HIP runtime has shut down)
    ShimState() { dgr_default_params(&p); }
    void release() {
        if (plan) dgr_plan_destroy(plan);
        plan = nullptr;
        table.clear();
        for (auto &kv : shape_plans) dgr_plan_destroy(kv.second);
        shape_plans.clear();
    }
};
ShimState g_shim;

bool same_table(const std::vector<dgr_block> &a, const dgr_block *b, uint32_t n) {
    return a.size() == n && (n == 0 || std::memcmp(a.data(), b, sizeof(dgr_block) * n) == 0);
}

// Runs the plan's right-hand side on the null stream and waits.
int run(dgr_plan *plan, double *const *d_rhs, const double *const *d_u, double t) {
    int rc = dgr_bssn_rhs(plan, d_rhs, d_u, t, &g_shim.p, nullptr);
    if (rc == DGR_OK) rc = dgr_plan_status(plan, nullptr);
    return rc;
}

}  // namespace

extern "C" {

dgr_bssn_params *dgr_shim_params(void) { return &g_shim.p; }

const dgr_shim_domain_t *dgr_shim_domain(void) { return &g_shim.dom; }

void dgr_shim_set_domain(const double *compd_min, const double *compd_max, const double *octree_min,
                         const double *octree_max, uint64_t szpdof) {
    g_shim.release();
    dgr_shim_domain_t &d = g_shim.dom;
    for (int a = 0; a < 3; a++) {
        d.compd_min[a] = compd_min[a];
        d.compd_max[a] = compd_max[a];
        d.octree_min[a] = octree_min[a];
        d.octree_max[a] = octree_max[a];
    }
    d.szpdof = szpdof;
    d.set = 1;
}

int dgr_shim_rhs(const dgr_block *blocks, uint32_t nblocks, double *const *d_rhs, const double *const *d_u,
                 double t) {
    const uint64_t szpdof = g_shim.dom.szpdof;
    if (!g_shim.plan || g_shim.table_szpdof != szpdof || !same_table(g_shim.table, blocks, nblocks)) {
        if (g_shim.plan) dgr_plan_destroy(g_shim.plan);
        g_shim.plan = nullptr;
        g_shim.table.clear();
        int rc = dgr_plan_create(0, blocks, nblocks, szpdof, &g_shim.plan);
        if (rc != DGR_OK) return rc;
        g_shim.table.assign(blocks, blocks + nblocks);
        g_shim.table_szpdof = szpdof;
        g_shim.plans_built++;
    }
    g_shim.calls++;
    return run(g_shim.plan, d_rhs, d_u, t);
}

int dgr_shim_rhs_block(const dgr_block *blk, double *const *d_rhs, const double *const *d_u, double t) {
    const std::array<uint32_t, 4> key{blk->sz[0], blk->sz[1], blk->sz[2], blk->bflag};
    // without a domain the arrays are known to reach at least to the block's end
    const uint64_t n3 = (uint64_t)blk->sz[0] * blk->sz[1] * blk->sz[2];
    const uint64_t szpdof = g_shim.dom.set ? g_shim.dom.szpdof : (uint64_t)0xffffffffu;
    if (blk->offset + n3 > szpdof) {
        dgr_block b = *blk;  // let the library report it
        dgr_plan *p = nullptr;
        return dgr_plan_create(0, &b, 1, szpdof, &p);
    }
    auto it = g_shim.shape_plans.find(key);
    if (it == g_shim.shape_plans.end()) {
        dgr_plan *p = nullptr;
        int rc = dgr_plan_create(0, blk, 1, szpdof, &p);
        if (rc != DGR_OK) return rc;
        it = g_shim.shape_plans.emplace(key, p).first;
        g_shim.plans_built++;
    } else {
        int rc = dgr_plan_move_block(it->second, 0, blk->offset, blk->pmin, blk->pmax);
        if (rc != DGR_OK) return rc;
    }
    g_shim.calls++;
    return run(it->second, d_rhs, d_u, t);
}

void dgr_shim_cache_stats(int64_t *plans_built, int64_t *calls) {
    if (plans_built) *plans_built = g_shim.plans_built;
    if (calls) *calls = g_shim.calls;
}

void dgr_shim_release(void) { g_shim.release(); }

}  // extern "C"

// rhs.h:32-36. uZipConstVars: read only by the CAHD variants, not this gauge.
void bssnrhs(double **unzipVarsRHS, const double **uZipVars, const unsigned int &offset, const double *pmin,
             const double *pmax, const unsigned int *sz, const unsigned int &bflag, const double t,
             const double **uZipConstVars) {
    (void)uZipConstVars;
    dgr_block b;
    b.offset = offset;
    for (int a = 0; a < 3; a++) {
        b.sz[a] = sz[a];
        b.pmin[a] = pmin[a];
        b.pmax[a] = pmax[a];
    }
    b.bflag = bflag;
    if (dgr_shim_rhs_block(&b, unzipVarsRHS, uZipVars, t) != DGR_OK)
        throw std::runtime_error(std::string("bssnrhs: ") + dgr_last_error());
}
