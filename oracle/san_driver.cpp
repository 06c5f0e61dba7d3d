// Sanitizer driver (test infrastructure): runs the oracle and the library's host C++ (terrain
// generator, host table build, image validation) on a small scene under AddressSanitizer +
// UndefinedBehaviorSanitizer or ThreadSanitizer. Built by `make -C oracle sanitize`; run by
// tests/test_sanitizers.py. Prints "OK" and exits 0 when every step succeeds.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../hashdag_amd/csrc/hd_internal.h"
#include "../hashdag_amd/csrc/hd_scene.h"
#include "../include/hashdag_hip.h"
#include "oracle_capi.h"

#define REQUIRE(c)                                                      \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                   \
        }                                                               \
    } while (0)

int main(int argc, char** argv) {
    const int threads = argc > 1 ? std::atoi(argv[1]) : 2;
    // scene: 2^8 coloured terrain, colour tree down to level 5 (the library's generator)
    hd::TerrainParams tp;
    tp.levels = 8;
    tp.seed = 7;
    tp.color_tree_levels = 5;
    tp.colors = true;
    tp.color_seed = 3;
    tp.threads = (uint32_t)threads;
    const hd::Scene sc = hd::generate_terrain(tp);
    const std::vector<uint32_t>& dag = sc.dag;
    std::vector<uint32_t> nodes = sc.color_nodes, weights = sc.weights;
    std::vector<uint64_t> offsets = sc.color_offsets, blocks = sc.blocks, macro = sc.macro;
    REQUIRE(!dag.empty() && !nodes.empty() && sc.num_voxels > 0);

    // the oracle's table and colours
    const orc_geometry g{8, 512, 5, 6, 8, 1024, 1024, 5};
    const uint32_t poolPages = 1u << 14;
    void* od = orc_dag_from_basic(&g, poolPages, dag.data(), dag.size());
    REQUIRE(od);
    void* oc = orc_colors_new(nodes.data(), nodes.size(), offsets.data(), offsets.size(), blocks.data(), blocks.size(),
                              weights.data(), weights.size(), macro.data(), macro.size());
    REQUIRE(oc);

    // image validation (host C++ of the library) on the oracle's image, then on damaged copies
    hd_config cfg{};
    cfg.levels = g.levels;
    cfg.page_size = g.page_size;
    cfg.num_top_levels = g.num_top_levels;
    cfg.bucket_bits_top = g.bucket_bits_top;
    cfg.bucket_bits_low = g.bucket_bits_low;
    cfg.bucket_size_top = g.bucket_size_top;
    cfg.bucket_size_low = g.bucket_size_low;
    cfg.color_tree_levels = g.color_tree_levels;
    cfg.pool_pages = poolPages;
    const uint64_t totalPages = (uint64_t)5 * 64 * 2 + (uint64_t)3 * 256 * 2;
    const uint64_t totalBuckets = (uint64_t)5 * 64 + (uint64_t)3 * 256;
    std::vector<uint32_t> pool((uint64_t)poolPages * 512), pt(totalPages), bs(totalBuckets), full(g.levels);
    REQUIRE(orc_dag_export(od, pool.data(), pt.data(), bs.data(), full.data()) == 0);
    const uint32_t top = orc_dag_pool_top(od), root = orc_dag_root(od);
    char why[256];
    REQUIRE(hd_validate_image(&cfg, pool.data(), pt.data(), bs.data(), full.data(), top, root, why, sizeof(why)) ==
            HD_OK);
    std::vector<uint32_t> bad = pt;
    for (uint32_t& v : bad)
        if (v) {
            v = top + 5;  // a page beyond the pool
            break;
        }
    REQUIRE(hd_validate_image(&cfg, pool.data(), bad.data(), bs.data(), full.data(), top, root, why, sizeof(why)) !=
            HD_OK);
    REQUIRE(hd_validate_image(&cfg, pool.data(), pt.data(), bs.data(), full.data(), top, 0xFFFFFFF0u, why,
                              sizeof(why)) != HD_OK);
    // the library's host table build (load_from_DAG) on the same scene reaches the same pool top
    hd::HostImage img;
    img.g = hd::Geom{g.levels, g.page_size, g.num_top_levels, g.bucket_bits_top, g.bucket_bits_low, g.bucket_size_top,
                     g.bucket_size_low, g.color_tree_levels};
    img.pool_pages = poolPages;
    hd::load_from_basic_dag(img, dag.data(), dag.size());
    REQUIRE(img.pool_top == top);
    REQUIRE(hd_validate_image(&cfg, img.pool.data(), img.page_table.data(), img.bucket_sizes.data(),
                              img.full_nodes.data(), img.pool_top, img.root, why, sizeof(why)) == HD_OK);

    // every editor kind
    uint64_t st[3];
    const float c0[3] = {100.f, 60.f, 120.f}, c1[3] = {140.f, 50.f, 90.f}, c2[3] = {60.f, 70.f, 60.f};
    REQUIRE(orc_edit_threads(od, oc, 0, c0, 9.f, threads, st) == 0);  // sphere add
    REQUIRE(orc_edit_threads(od, oc, 1, c1, 7.f, threads, st) == 0);  // sphere carve
    REQUIRE(orc_edit(od, oc, 2, c2, 5.f, st) == 0);                    // box add
    REQUIRE(orc_edit(od, oc, 3, c0, 4.f, st) == 0);                    // box carve
    REQUIRE(orc_edit(od, oc, 4, c1, 12.f, st) == 0);                   // sphere paint
    const float src[3] = {100.f, 60.f, 120.f}, dst[3] = {180.f, 60.f, 180.f};
    REQUIRE(orc_edit_copy_kind(od, oc, 5, c0, 10.f, src, dst, st) == 0);  // copy
    REQUIRE(orc_edit_copy_kind(od, oc, 7, c0, 6.5f, src, dst, st) == 0);  // copy with decompression
    const double tf[9] = {0, 0, 1, 0, 1, 0, -1, 0, 0};
    REQUIRE(orc_edit_copy_tf(od, oc, c0, 6.f, src, dst, tf, 1, 40.f, st) == 0);  // transform + swirl
    REQUIRE(orc_edit(od, oc, 6, c2, 6.f, st) == 0);                             // fill
    REQUIRE(orc_edit(od, oc, 8, c1, 6.f, st) == 0);                             // noise add
    REQUIRE(orc_edit(od, oc, 9, c0, 5.f, st) == 0);                             // noise carve

    // a small frame: paths, colours, shadows
    const double pos[3] = {30.0, 140.0, 30.0}, rot[9] = {-0.7071, 0, 0.7071, -0.3, 0.9, -0.3, 0.64, 0.42, 0.64};
    const double bmin[3] = {0, 0, 0}, bmax[3] = {256, 256, 256};
    double prm[12];
    orc_ray_params(pos, rot, bmin, bmax, g.levels, 48, 32, prm);
    std::vector<uint32_t> paths(48 * 32 * 4), cols(48 * 32);
    uint64_t bytes = 0;
    REQUIRE(orc_trace_paths(od, prm, 48, 32, 0, 32, paths.data(), &bytes, threads) == 0);
    REQUIRE(orc_trace_colors(od, oc, 48, 32, 0, 32, paths.data(), cols.data(), threads) == 0);
    REQUIRE(orc_trace_shadows(od, prm, 48, 32, 0, 32, paths.data(), cols.data(), 1.f, 0.f, &bytes, threads) == 0);

    // GC, then the canonical forms
    REQUIRE(orc_gc(od, g.levels - 2) == 0);
    const uint64_t nCanon = orc_canonical(od, nullptr, 0);
    std::vector<uint32_t> canon(nCanon);
    REQUIRE(orc_canonical(od, canon.data(), nCanon) == nCanon);
    const uint64_t nColors = orc_color_canonical(od, oc, nullptr, 0);
    std::vector<uint32_t> ccanon(nColors);
    REQUIRE(orc_color_canonical(od, oc, ccanon.data(), nColors) == nColors);

    // colour codec round trip
    const uint64_t n = 40000;
    std::vector<uint32_t> cb(n);
    std::vector<uint8_t> w(n), bpw(n);
    for (uint64_t i = 0; i < n; ++i) {
        cb[i] = (uint32_t)(i / 300) * 2654435761u;
        bpw[i] = (uint8_t)((i / 300) % 4);
        w[i] = bpw[i] ? (uint8_t)(i % (1u << bpw[i])) : 0;
    }
    std::vector<uint64_t> eb(n), em(64);
    std::vector<uint32_t> ew(n);
    uint64_t es[3];
    REQUIRE(orc_encode_colors(cb.data(), w.data(), bpw.data(), n, eb.data(), ew.data(), em.data(), es, eb.size(),
                              ew.size(), em.size()) == 0);
    for (uint64_t i = 0; i < n; i += 997) {
        uint32_t out[2];
        REQUIRE(orc_decode_color(eb.data(), es[0], ew.data(), es[1], em.data(), es[2], 0, i, out) == 0);
        REQUIRE(out[0] == cb[i] && (out[1] & 0xFFu) == w[i] && (out[1] >> 8) == bpw[i]);
    }
    orc_colors_free(oc);
    orc_dag_free(od);
    std::printf("OK\n");
    return 0;
}
