// Calls the reference-signature entry points of libdgr_bssn.so exactly as a
// Dendro-GR solver would: bssnRHS(double**, const double**, const Block*,
// unsigned, double, const double**) (BSSN_GR/include/rhs.h:28-30) over a block
// type with ot::Block's accessors, and the per-block loop of
// BSSN_GR/src/rhs.cpp:49-71 through bssnrhs (rhs.h:32-36).
//
// usage: shim_driver <input> <output>
// input (little-endian): int64 nblocks, szpdof, maxdepth, repeats;
//   double compd_min[3], compd_max[3], ko_sigma, eta_central, eta_outer,
//   eta_width;
//   per block int64 offset, node min x/y/z, element level, elements per
//   axis, bflag; then the state, 24 x szpdof doubles.
// output: the state after bssnRHS and after the bssnrhs loop (2 x 24 x
//   szpdof doubles, padding left as the input's), and int64 plans built /
//   calls served after the bssnRHS calls and after the bssnrhs loops.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dgr_bssn_shim.hpp"

namespace {

constexpr unsigned ELE_ORDER = 6;

struct TestPoint {  // Dendro's Point: x(), y(), z()
    double v[3];
    TestPoint(double x, double y, double z) : v{x, y, z} {}
    double x() const { return v[0]; }
    double y() const { return v[1]; }
    double z() const { return v[2]; }
};

struct TestNode {  // ot::TreeNode corners in octree coordinates
    unsigned mn[3], mx[3];
    unsigned minX() const { return mn[0]; }
    unsigned minY() const { return mn[1]; }
    unsigned minZ() const { return mn[2]; }
    unsigned maxX() const { return mx[0]; }
    unsigned maxY() const { return mx[1]; }
    unsigned maxZ() const { return mx[2]; }
};

// ot::Block's accessors used by bssnRHS (rhs.cpp:49-67): a block of ne^3
// elements of order 6 at octree level `lev`, padded by 3 points per side.
struct TestBlock {
    unsigned offset, sz, bflag, maxdepth, lev;
    TestNode node;
    unsigned getOffset() const { return offset; }
    unsigned getAllocationSzX() const { return sz; }
    unsigned getAllocationSzY() const { return sz; }
    unsigned getAllocationSzZ() const { return sz; }
    unsigned getBlkNodeFlag() const { return bflag; }
    TestNode getBlockNode() const { return node; }
    double spacing(double lo, double hi) const {
        return ((hi - lo) / ((double)(1u << maxdepth))) * ((1u << (maxdepth - lev)) / ((double)ELE_ORDER));
    }
    double computeDx(const TestPoint &a, const TestPoint &b) const { return spacing(a.x(), b.x()); }
    double computeDy(const TestPoint &a, const TestPoint &b) const { return spacing(a.y(), b.y()); }
    double computeDz(const TestPoint &a, const TestPoint &b) const { return spacing(a.z(), b.z()); }
};

#define CHECK_HIP(x)                                                         \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) {                                              \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));     \
            std::exit(3);                                                    \
        }                                                                    \
    } while (0)

template <typename T>
void rd(FILE *f, T *p, size_t n) {
    if (std::fread(p, sizeof(T), n, f) != n) {
        std::fprintf(stderr, "short input\n");
        std::exit(2);
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
        return 2;
    }
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int64_t hdr[4];
    rd(in, hdr, 4);
    const int64_t nb = hdr[0], szpdof = hdr[1], maxdepth = hdr[2], repeats = hdr[3];
    double dom[10];
    rd(in, dom, 10);
    std::vector<TestBlock> blocks(nb);
    for (int64_t b = 0; b < nb; b++) {
        int64_t r[7];
        rd(in, r, 7);
        TestBlock &t = blocks[b];
        t.offset = (unsigned)r[0];
        t.maxdepth = (unsigned)maxdepth;
        t.lev = (unsigned)r[4];
        const unsigned ne = (unsigned)r[5];
        t.sz = ne * ELE_ORDER + 1 + 6;
        t.bflag = (unsigned)r[6];
        const unsigned ext = ne << (maxdepth - t.lev);
        for (int a = 0; a < 3; a++) {
            t.node.mn[a] = (unsigned)r[1 + a];
            t.node.mx[a] = (unsigned)r[1 + a] + ext;
        }
    }
    const size_t n = (size_t)szpdof;
    std::vector<double> state(24 * n);
    rd(in, state.data(), state.size());
    std::fclose(in);

    const double octree_min[3] = {0.0, 0.0, 0.0};
    const double octree_max[3] = {(double)(1u << maxdepth), (double)(1u << maxdepth), (double)(1u << maxdepth)};
    dgr_shim_set_domain(dom, dom + 3, octree_min, octree_max, (uint64_t)szpdof);
    dgr::bssn_params().ko_sigma = dom[6];
    dgr::bssn_params().eta_central = dom[7];
    dgr::bssn_params().eta_outer = dom[8];
    dgr::bssn_params().eta_width = dom[9];

    double *d_u = nullptr, *d_r1 = nullptr, *d_r2 = nullptr;
    CHECK_HIP(hipMalloc(&d_u, sizeof(double) * 24 * n));
    CHECK_HIP(hipMalloc(&d_r1, sizeof(double) * 24 * n));
    CHECK_HIP(hipMalloc(&d_r2, sizeof(double) * 24 * n));
    CHECK_HIP(hipMemcpy(d_u, state.data(), sizeof(double) * 24 * n, hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d_r1, d_u, sizeof(double) * 24 * n, hipMemcpyDeviceToDevice));
    CHECK_HIP(hipMemcpy(d_r2, d_u, sizeof(double) * 24 * n, hipMemcpyDeviceToDevice));
    double *u[24], *r1[24], *r2[24];
    for (int v = 0; v < 24; v++) {
        u[v] = d_u + v * n;
        r1[v] = d_r1 + v * n;
        r2[v] = d_r2 + v * n;
    }

    int64_t stats[4];
    try {
        // once per RK stage with the same block list (bssnCtx.cpp:96-97)
        for (int64_t rep = 0; rep < repeats; rep++)
            bssnRHS(r1, (const double **)u, blocks.data(), (unsigned)nb, 0.0, nullptr);
        dgr_shim_cache_stats(&stats[0], &stats[1]);
        // rhs.cpp:49-71 with the same corners, block by block
        const dgr_shim_domain_t &D = *dgr_shim_domain();
        const TestPoint pt_min(D.compd_min[0], D.compd_min[1], D.compd_min[2]);
        const TestPoint pt_max(D.compd_max[0], D.compd_max[1], D.compd_max[2]);
        for (int64_t rep = 0; rep < repeats; rep++)
            for (int64_t b = 0; b < nb; b++) {
                const TestBlock &blk = blocks[b];
                const unsigned offset = blk.getOffset();
                const unsigned sz[3] = {blk.getAllocationSzX(), blk.getAllocationSzY(), blk.getAllocationSzZ()};
                const unsigned bflag = blk.getBlkNodeFlag();
                const double dx[3] = {blk.computeDx(pt_min, pt_max), blk.computeDy(pt_min, pt_max),
                                      blk.computeDz(pt_min, pt_max)};
                const unsigned lo[3] = {blk.getBlockNode().minX(), blk.getBlockNode().minY(), blk.getBlockNode().minZ()};
                const unsigned hi[3] = {blk.getBlockNode().maxX(), blk.getBlockNode().maxY(), blk.getBlockNode().maxZ()};
                double ptmin[3], ptmax[3];
                for (int a = 0; a < 3; a++) {
                    const double R = D.compd_max[a] - D.compd_min[a], Rg = D.octree_max[a] - D.octree_min[a];
                    ptmin[a] = ((R / Rg) * (lo[a] - D.octree_min[a])) + D.compd_min[a] - 3 * dx[a];
                    ptmax[a] = ((R / Rg) * (hi[a] - D.octree_min[a])) + D.compd_min[a] + 3 * dx[a];
                }
                bssnrhs(r2, (const double **)u, offset, ptmin, ptmax, sz, bflag, 0.0, nullptr);
            }
        dgr_shim_cache_stats(&stats[2], &stats[3]);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 4;
    }

    FILE *out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    CHECK_HIP(hipMemcpy(state.data(), d_r1, sizeof(double) * 24 * n, hipMemcpyDeviceToHost));
    std::fwrite(state.data(), sizeof(double), state.size(), out);
    CHECK_HIP(hipMemcpy(state.data(), d_r2, sizeof(double) * 24 * n, hipMemcpyDeviceToHost));
    std::fwrite(state.data(), sizeof(double), state.size(), out);
    std::fwrite(stats, sizeof(int64_t), 4, out);
    std::fclose(out);
    dgr_shim_release();
    CHECK_HIP(hipFree(d_u));
    CHECK_HIP(hipFree(d_r1));
    CHECK_HIP(hipFree(d_r2));
    std::printf("shim_driver: %lld blocks, plans built %lld / %lld, calls %lld / %lld\n", (long long)nb,
                (long long)stats[0], (long long)stats[2], (long long)stats[1], (long long)stats[3]);
    return 0;
}
