// Context, device arena, reference-layout image upload/download, colour state and camera
// setup behind the C ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "hd_ctx.h"
#include "hd_scene.h"

namespace hd {

static thread_local std::string t_err;

hd_status fail(hd_ctx* ctx, hd_status code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    t_err = msg;
    return code;
}

hd_status hip_fail(hd_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, e == hipErrorOutOfMemory ? HD_ERR_OOM : HD_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

DevDag dev_dag(const hd_ctx* ctx) {
    DevDag d;
    d.vpool = ctx->vpool;
    d.root = ctx->root;
    d.levels = ctx->g.levels;
    d.leaf_level = ctx->g.leaf_level();
    d.color_tree_levels = ctx->g.color_tree_levels;
    d.root_ptr = nullptr;
    return d;
}

DevColors dev_colors(const hd_ctx* ctx) {
    DevColors c;
    c.nodes = ctx->cnodes;
    c.offsets = ctx->coffsets;
    c.leaves = ctx->cleaves;
    c.arena = ctx->arena;
    c.main = ctx->main;
    c.valid = ctx->colors_loaded ? 1u : 0u;
    return c;
}

// Copies every allocated virtual page of a physical pool into the virtual-address pool.
__global__ void k_scatter_pages(uint32_t* __restrict__ vpool, const uint32_t* __restrict__ phys,
                                const uint32_t* __restrict__ page_table, uint64_t total_pages, uint32_t page_size) {
    for (uint64_t v = blockIdx.x; v < total_pages; v += gridDim.x) {
        const uint32_t p = page_table[v];
        if (p == 0) continue;
        for (uint32_t i = threadIdx.x; i < page_size; i += blockDim.x)
            vpool[v * page_size + i] = phys[(uint64_t)p * page_size + i];
    }
}

// Inverse of k_scatter_pages.
__global__ void k_gather_pages(uint32_t* __restrict__ phys, const uint32_t* __restrict__ vpool,
                               const uint32_t* __restrict__ page_table, uint64_t total_pages, uint32_t page_size) {
    for (uint64_t v = blockIdx.x; v < total_pages; v += gridDim.x) {
        const uint32_t p = page_table[v];
        if (p == 0) continue;
        for (uint32_t i = threadIdx.x; i < page_size; i += blockDim.x)
            phys[(uint64_t)p * page_size + i] = vpool[v * page_size + i];
    }
}

static uint32_t grid_for(uint64_t items) { return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(items, 1), 65536); }

static hd_status upload_image_impl(hd_ctx* ctx, const uint32_t* pool, const uint32_t* page_table,
                                   const uint32_t* bucket_sizes, const uint32_t* full_nodes, uint32_t pool_top,
                                   uint32_t root) {
    const Geom& g = ctx->g;
    if (pool_top > ctx->cfg.pool_pages) return fail(ctx, HD_ERR_POOL_FULL, "image uses more pages than the pool budget");
    const uint64_t physWords = (uint64_t)pool_top * g.page_size;
    uint32_t* dphys = nullptr;
    HD_HIP(ctx, hipMalloc(&dphys, std::max<uint64_t>(physWords, 1) * 4));
    hipError_t e = hipMemcpyAsync(dphys, pool, physWords * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->page_table, page_table, g.total_pages() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->bucket_sizes, bucket_sizes, g.total_buckets() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ctx->full_nodes, full_nodes, g.levels * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        k_scatter_pages<<<grid_for(g.total_pages()), 256, 0, ctx->stream>>>(ctx->vpool, dphys, ctx->page_table,
                                                                           g.total_pages(), g.page_size);
        e = hipGetLastError();
    }
    const uint32_t counters[2] = {pool_top, 0u};
    if (e == hipSuccess) e = hipMemcpyAsync(ctx->counters, counters, sizeof(counters), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ctx->d_root, &root, 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(dphys);
    if (e != hipSuccess) return hip_fail(ctx, e, "upload image");
    const hd_status st = leaf_tags_rebuild(ctx, ctx->stream);
    if (st != HD_OK) return st;
    ctx->root = root;
    ctx->loaded = true;
    return HD_OK;
}

}  // namespace hd

using namespace hd;

extern "C" {

void hd_default_config(uint32_t levels, hd_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->levels = levels;
    c->page_size = 512;
    c->num_top_levels = 9;
    c->bucket_bits_top = 10;
    c->bucket_bits_low = 16;
    c->bucket_size_top = 1024;
    c->bucket_size_low = 4096;
    c->color_tree_levels = 10;
    c->pool_pages = 0x8FFFFFFFu / 512u / 4u;  // main.cpp:46
    c->device = 0;
    c->color_arena_words = 0;  // sized at hd_load_colors
    c->color_nodes_cap = 0;
    c->color_leaves_cap = 1u << 20;
}

hd_status hd_create(const hd_config* cfg, hd_ctx** out) {
    if (!cfg || !out) return fail(nullptr, HD_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->levels < 4 || cfg->levels > 17) return fail(nullptr, HD_ERR_INVALID, "levels must be in [4, 17]");
    if (cfg->page_size == 0 || (cfg->page_size & (cfg->page_size - 1)) || cfg->bucket_size_top % cfg->page_size ||
        cfg->bucket_size_low % cfg->page_size)
        return fail(nullptr, HD_ERR_INVALID, "bucket sizes must be multiples of a power-of-two page size");
    // a full node's header counts 8^(levels - level) voxels in 24 bits from color_tree_levels
    // down (hash_table.h:749-789; the reference aborts beyond that, hash_table.h:772); a colour
    // tree at or below the leaf level means no colour leaves
    if (cfg->color_tree_levels + 7 < cfg->levels || cfg->color_tree_levels == 0)
        return fail(nullptr, HD_ERR_INVALID, "color_tree_levels must be at least levels - 7");
    auto* ctx = new hd_ctx();
    ctx->cfg = *cfg;
    ctx->g = Geom{cfg->levels,          cfg->page_size,       cfg->num_top_levels, cfg->bucket_bits_top,
                  cfg->bucket_bits_low, cfg->bucket_size_top, cfg->bucket_size_low, cfg->color_tree_levels};
    if (ctx->g.total_virtual() >= 0xFFFFFFFFull) {
        delete ctx;
        return fail(nullptr, HD_ERR_INVALID, "virtual address space exceeds 32 bits");
    }
    ctx->device = (int)cfg->device;
    auto cleanup = [&](hd_status s) {
        hd_destroy(ctx);
        return s;
    };
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return cleanup(hip_fail(nullptr, e, "hipSetDevice"));
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(hip_fail(nullptr, e, "hipStreamCreate"));
    const Geom& g = ctx->g;
    const uint64_t vwords = g.total_virtual() + kVpoolPad;
    if ((e = hipMalloc(&ctx->vpool, vwords * 4)) != hipSuccess) return cleanup(hip_fail(nullptr, e, "vpool"));
    if ((e = hipMalloc(&ctx->page_table, g.total_pages() * 4)) != hipSuccess) return cleanup(hip_fail(nullptr, e, "page table"));
    if ((e = hipMalloc(&ctx->bucket_sizes, g.total_buckets() * 4)) != hipSuccess) return cleanup(hip_fail(nullptr, e, "buckets"));
    if ((e = hipMalloc(&ctx->bucket_snapshot, g.total_buckets() * 4)) != hipSuccess)
        return cleanup(hip_fail(nullptr, e, "bucket snapshot"));
    if ((e = hipMalloc(&ctx->full_nodes, g.levels * 4)) != hipSuccess) return cleanup(hip_fail(nullptr, e, "full nodes"));
    if ((e = hipMalloc(&ctx->counters, 64 * 4)) != hipSuccess) return cleanup(hip_fail(nullptr, e, "counters"));
    if ((e = hipMalloc(&ctx->byte_counters, 8 * sizeof(unsigned long long))) != hipSuccess)
        return cleanup(hip_fail(nullptr, e, "byte counters"));
    if ((e = hipMalloc(&ctx->d_root, 64)) != hipSuccess) return cleanup(hip_fail(nullptr, e, "root"));
    if ((e = hipMalloc(&ctx->leaf_tags, std::max<uint64_t>(leaf_tag_bytes(g), 4))) != hipSuccess)
        return cleanup(hip_fail(nullptr, e, "leaf tags"));
    if ((e = hipMalloc(&ctx->edit_bytes, 8 * sizeof(unsigned long long))) != hipSuccess)
        return cleanup(hip_fail(nullptr, e, "edit byte counters"));
    hipMemsetAsync(ctx->vpool, 0, vwords * 4, ctx->stream);
    hipMemsetAsync(ctx->page_table, 0, g.total_pages() * 4, ctx->stream);
    hipMemsetAsync(ctx->bucket_sizes, 0, g.total_buckets() * 4, ctx->stream);
    hipMemsetAsync(ctx->full_nodes, 0, g.levels * 4, ctx->stream);
    hipMemsetAsync(ctx->counters, 0, 64 * 4, ctx->stream);
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return cleanup(hip_fail(nullptr, e, "init"));
    // the static edit scratch of this device (hd_edit_batch_static), looked up once per device
    const hd_status st = static_scratch_setup(ctx);
    if (st != HD_OK) return cleanup(fail(nullptr, st, ctx->err));
    *out = ctx;
    return HD_OK;
}

hd_status hd_destroy(hd_ctx* ctx) {
    if (!ctx) return HD_OK;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    hipDeviceSynchronize();
    edit_free(ctx);
    gc_scratch_free(ctx);
    hipFree(ctx->d_root);
    hipFree(ctx->vpool);
    hipFree(ctx->page_table);
    hipFree(ctx->bucket_sizes);
    hipFree(ctx->bucket_snapshot);
    hipFree(ctx->full_nodes);
    hipFree(ctx->counters);
    hipFree(ctx->byte_counters);
    trace_scratch_free(ctx);
    hipFree(ctx->leaf_tags);
    hipFree(ctx->edit_bytes);
    hipFree(ctx->cnodes);
    hipFree(ctx->coffsets);
    hipFree(ctx->cleaves);
    hipFree(ctx->arena);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return HD_OK;
}

const char* hd_last_error(hd_ctx* ctx) { return ctx ? ctx->err.c_str() : t_err.c_str(); }

hd_status hd_load_basic_dag(hd_ctx* ctx, const uint32_t* data, uint64_t n) {
    if (!ctx || !data || n < 2) return fail(ctx, HD_ERR_INVALID, "bad BasicDAG");
    HostImage img;
    img.g = ctx->g;
    img.pool_pages = ctx->cfg.pool_pages;
    try {
        load_from_basic_dag(img, data, n);
    } catch (const std::exception& ex) {
        const std::string m = ex.what();
        return fail(ctx, m.find("pool") != std::string::npos ? HD_ERR_POOL_FULL : HD_ERR_BUCKET_FULL, m);
    }
    return upload_image_impl(ctx, img.pool.data(), img.page_table.data(), img.bucket_sizes.data(),
                             img.full_nodes.data(), img.pool_top, img.root);
}

hd_status hd_image_info_get(hd_ctx* ctx, hd_image_info* out) {
    if (!ctx || !out) return fail(ctx, HD_ERR_INVALID, "null argument");
    uint32_t top = 0;
    HD_HIP(ctx, hipMemcpyAsync(&top, ctx->counters, 4, hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    out->root = ctx->root;
    out->pool_top = top;
    out->pool_pages = ctx->cfg.pool_pages;
    out->levels = ctx->g.levels;
    out->total_pages = ctx->g.total_pages();
    out->total_buckets = ctx->g.total_buckets();
    out->total_virtual = ctx->g.total_virtual();
    return HD_OK;
}

hd_status hd_download_image(hd_ctx* ctx, uint32_t* pool, uint32_t* page_table, uint32_t* bucket_sizes,
                            uint32_t* full_nodes) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    const Geom& g = ctx->g;
    if (pool) {
        const hd_status st = download_pool(ctx, pool, (uint64_t)ctx->cfg.pool_pages * g.page_size);
        if (st != HD_OK) return st;
    }
    if (page_table) HD_HIP(ctx, hipMemcpyAsync(page_table, ctx->page_table, g.total_pages() * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (bucket_sizes)
        HD_HIP(ctx, hipMemcpyAsync(bucket_sizes, ctx->bucket_sizes, g.total_buckets() * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (full_nodes) HD_HIP(ctx, hipMemcpyAsync(full_nodes, ctx->full_nodes, g.levels * 4, hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

}  // extern "C"

namespace hd {
// The physical pool pages in use (page 0 .. pool_top - 1), then zeros up to cap_words.
hd_status download_pool(hd_ctx* ctx, uint32_t* pool, uint64_t cap_words) {
    const Geom& g = ctx->g;
    {
        uint32_t top = 0;
        HD_HIP(ctx, hipMemcpyAsync(&top, ctx->counters, 4, hipMemcpyDeviceToHost, ctx->stream));
        HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const uint64_t words = (uint64_t)std::max<uint32_t>(top, 1) * g.page_size;
        uint32_t* dphys = nullptr;
        HD_HIP(ctx, hipMalloc(&dphys, words * 4));
        hipError_t e = hipMemsetAsync(dphys, 0, words * 4, ctx->stream);
        if (e == hipSuccess) {
            k_gather_pages<<<grid_for(g.total_pages()), 256, 0, ctx->stream>>>(dphys, ctx->vpool, ctx->page_table,
                                                                              g.total_pages(), g.page_size);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(pool, dphys, std::min(words, cap_words) * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        hipFree(dphys);
        if (e != hipSuccess) return hip_fail(ctx, e, "download pool");
        // pages at and above pool_top are zero
        if (cap_words > words) std::memset(pool + words, 0, (cap_words - words) * 4);
    }
    return HD_OK;
}
}  // namespace hd

extern "C" {

hd_status hd_upload_image(hd_ctx* ctx, const uint32_t* pool, const uint32_t* page_table, const uint32_t* bucket_sizes,
                          const uint32_t* full_nodes, uint32_t pool_top, uint32_t root) {
    if (!ctx || !pool || !page_table || !bucket_sizes || !full_nodes) return fail(ctx, HD_ERR_INVALID, "null argument");
    const std::string why =
        validate_image(ctx->g, ctx->cfg.pool_pages, pool, page_table, bucket_sizes, full_nodes, pool_top, root);
    if (!why.empty()) return fail(ctx, HD_ERR_INVALID, "invalid image: " + why);
    return upload_image_impl(ctx, pool, page_table, bucket_sizes, full_nodes, pool_top, root);
}

hd_status hd_get_root(hd_ctx* ctx, uint32_t* root) {
    if (!ctx || !root) return fail(ctx, HD_ERR_INVALID, "null argument");
    *root = ctx->root;
    return HD_OK;
}

hd_status hd_set_root(hd_ctx* ctx, uint32_t root) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    if ((uint64_t)root >= ctx->g.total_virtual()) return fail(ctx, HD_ERR_INVALID, "root outside the virtual space");
    ctx->root = root;
    HD_HIP(ctx, hipMemcpy(ctx->d_root, &root, 4, hipMemcpyHostToDevice));
    return HD_OK;
}

hd_status hd_load_colors(hd_ctx* ctx, const uint32_t* nodes, uint64_t n_nodes, const uint64_t* offsets,
                         uint64_t n_offsets, const uint64_t* blocks, uint64_t n_blocks, const uint32_t* weights,
                         uint64_t n_weights, const uint64_t* macro, uint64_t n_macro) {
    if (!ctx || !nodes || n_nodes < 8) return fail(ctx, HD_ERR_INVALID, "colour tree needs a root node");
    if (n_macro % 2) return fail(ctx, HD_ERR_INVALID, "macro array holds pairs");
    const uint64_t weights64 = (n_weights + 1) / 2 + 1;  // + 1 word: bit-window reads past the end
    const uint64_t mainWords = n_blocks + n_macro + weights64;
    const uint64_t arenaWords = std::max<uint64_t>(ctx->cfg.color_arena_words, mainWords + mainWords / 2 + (1u << 22));
    const uint64_t nodesCap = std::max<uint64_t>(ctx->cfg.color_nodes_cap, n_nodes * 2 + (1u << 16));
    const uint64_t leavesCap = std::max<uint64_t>(ctx->cfg.color_leaves_cap, 1024);
    hipFree(ctx->cnodes);
    hipFree(ctx->coffsets);
    hipFree(ctx->cleaves);
    hipFree(ctx->arena);
    ctx->cnodes = nullptr;
    ctx->coffsets = nullptr;
    ctx->cleaves = nullptr;
    ctx->arena = nullptr;
    ctx->colors_loaded = false;
    HD_HIP(ctx, hipMalloc(&ctx->cnodes, nodesCap * 4));
    HD_HIP(ctx, hipMalloc(&ctx->coffsets, std::max<uint64_t>(n_offsets, 1) * 8));
    HD_HIP(ctx, hipMalloc(&ctx->cleaves, leavesCap * sizeof(LeafDesc)));
    HD_HIP(ctx, hipMalloc(&ctx->arena, arenaWords * 8));
    ctx->cfg.color_arena_words = arenaWords;
    ctx->cfg.color_nodes_cap = (uint32_t)nodesCap;
    ctx->cfg.color_leaves_cap = (uint32_t)leavesCap;
    HD_HIP(ctx, hipMemsetAsync(ctx->cnodes, 0, nodesCap * 4, ctx->stream));
    HD_HIP(ctx, hipMemsetAsync(ctx->arena, 0, arenaWords * 8, ctx->stream));
    HD_HIP(ctx, hipMemcpyAsync(ctx->cnodes, nodes, n_nodes * 4, hipMemcpyHostToDevice, ctx->stream));
    if (n_offsets) HD_HIP(ctx, hipMemcpyAsync(ctx->coffsets, offsets, n_offsets * 8, hipMemcpyHostToDevice, ctx->stream));
    LeafDesc m{};
    m.blocks = 0;
    m.n_blocks = n_blocks;
    m.macro = n_blocks;
    m.n_macro = n_macro;
    m.weights32 = (n_blocks + n_macro) * 2;
    m.n_weights = n_weights;
    m.offset = 0;
    if (n_blocks) HD_HIP(ctx, hipMemcpyAsync(ctx->arena, blocks, n_blocks * 8, hipMemcpyHostToDevice, ctx->stream));
    if (n_macro) HD_HIP(ctx, hipMemcpyAsync(ctx->arena + n_blocks, macro, n_macro * 8, hipMemcpyHostToDevice, ctx->stream));
    if (n_weights)
        HD_HIP(ctx, hipMemcpyAsync(reinterpret_cast<uint32_t*>(ctx->arena) + m.weights32, weights, n_weights * 4,
                                   hipMemcpyHostToDevice, ctx->stream));
    // colour-tree allocation tops used by edits: entries in use, unique leaves in use
    const uint32_t tops[2] = {(uint32_t)n_nodes, 0u};
    HD_HIP(ctx, hipMemcpyAsync(ctx->counters + 2, tops, sizeof(tops), hipMemcpyHostToDevice, ctx->stream));
    const uint64_t top64 = mainWords;  // committed colour arena top (device copy)
    HD_HIP(ctx, hipMemcpyAsync(ctx->counters + kArenaTopWord, &top64, 8, hipMemcpyHostToDevice, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->main = m;
    ctx->arena_top = mainWords;
    ctx->n_cnodes = n_nodes;
    ctx->n_coffsets = n_offsets;
    ctx->n_cleaves = 0;
    ctx->h_leaves.clear();
    ctx->colors_loaded = true;
    ctx->ptr_gen++;
    return HD_OK;
}

hd_status hd_colors_info(hd_ctx* ctx, uint64_t* n_nodes, uint64_t* n_leaves, uint64_t* n_offsets) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    if (n_nodes) *n_nodes = ctx->n_cnodes;
    if (n_leaves) *n_leaves = ctx->n_cleaves;
    if (n_offsets) *n_offsets = ctx->n_coffsets;
    return HD_OK;
}

hd_status hd_download_color_nodes(hd_ctx* ctx, uint32_t* nodes) {
    if (!ctx || !nodes || !ctx->colors_loaded) return fail(ctx, HD_ERR_INVALID, "no colours");
    HD_HIP(ctx, hipMemcpyAsync(nodes, ctx->cnodes, ctx->n_cnodes * 4, hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

static hd_status resolve_leaf(hd_ctx* ctx, uint32_t handle, LeafDesc& d) {
    if (!ctx || !ctx->colors_loaded) return fail(ctx, HD_ERR_INVALID, "no colours");
    if (handle & 0x80000000u) {
        const uint32_t i = handle & 0x7FFFFFFFu;
        if (i >= ctx->n_cleaves) return fail(ctx, HD_ERR_INVALID, "leaf handle out of range");
        HD_HIP(ctx, hipMemcpyAsync(&d, ctx->cleaves + i, sizeof(d), hipMemcpyDeviceToHost, ctx->stream));
        HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return HD_OK;
    }
    if (handle >= ctx->n_coffsets) return fail(ctx, HD_ERR_INVALID, "shared handle out of range");
    uint64_t off = 0;
    HD_HIP(ctx, hipMemcpyAsync(&off, ctx->coffsets + handle, 8, hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    d = ctx->main;
    d.offset = off;
    return HD_OK;
}

hd_status hd_color_leaf_sizes(hd_ctx* ctx, uint32_t handle, uint64_t sizes[4]) {
    LeafDesc d;
    hd_status s = resolve_leaf(ctx, handle, d);
    if (s != HD_OK) return s;
    sizes[0] = d.n_blocks;
    sizes[1] = d.n_weights;
    sizes[2] = d.n_macro;
    sizes[3] = d.offset;
    return HD_OK;
}

hd_status hd_download_color_leaf(hd_ctx* ctx, uint32_t handle, uint64_t* blocks, uint32_t* weights, uint64_t* macro) {
    LeafDesc d;
    hd_status s = resolve_leaf(ctx, handle, d);
    if (s != HD_OK) return s;
    if (blocks && d.n_blocks)
        HD_HIP(ctx, hipMemcpyAsync(blocks, ctx->arena + d.blocks, d.n_blocks * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (macro && d.n_macro)
        HD_HIP(ctx, hipMemcpyAsync(macro, ctx->arena + d.macro, d.n_macro * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (weights && d.n_weights)
        HD_HIP(ctx, hipMemcpyAsync(weights, reinterpret_cast<uint32_t*>(ctx->arena) + d.weights32, d.n_weights * 4,
                                   hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

// dag_tracer.cu:71-113, FP64 in the reference's operation order.
hd_status hd_ray_params_from_camera(const double position[3], const double rot[9], const double bmin[3],
                                    const double bmax[3], uint32_t levels, uint32_t W, uint32_t H, hd_ray_params* out) {
    if (!position || !rot || !bmin || !bmax || !out || W == 0 || H == 0 || levels > 30)
        return fail(nullptr, HD_ERR_INVALID, "bad camera arguments");
    struct V { double x, y, z; };
    auto add = [](V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; };
    auto sub = [](V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; };
    auto mul = [](V a, V b) { return V{a.x * b.x, a.y * b.y, a.z * b.z}; };
    auto muls = [](V a, double s) { return V{a.x * s, a.y * s, a.z * s}; };
    auto div = [](V a, V b) { return V{a.x / b.x, a.y / b.y, a.z / b.z}; };
    const V pos{position[0], position[1], position[2]};
    const V right{-rot[0], -rot[1], -rot[2]};
    const V up{rot[3], rot[4], rot[5]};
    const V fwd{rot[6], rot[7], rot[8]};
    const V lo{bmin[0], bmin[1], bmin[2]};
    const V hi{bmax[0], bmax[1], bmax[2]};
    const double fov = 60.0 / 2.0 * (double(M_PI) / 180.);
    const double aspect = double(W) / double(H);
    const V X = muls(muls(right, std::sin(fov)), aspect);
    const V Y = muls(up, std::sin(fov));
    const V Z = muls(fwd, std::cos(fov));
    const V bl = sub(sub(add(pos, Z), Y), X);
    const V br = add(sub(add(pos, Z), Y), X);
    const V tl = sub(add(add(pos, Z), Y), X);
    const V tr{-lo.x, -lo.y, -lo.z};
    const double side = double(1u << levels);
    const V scale = div(V{side, side, side}, sub(hi, lo));
    const V fPos = mul(add(pos, tr), scale);
    const V fBl = mul(add(bl, tr), scale);
    const V fTl = mul(add(tl, tr), scale);
    const V fBr = mul(add(br, tr), scale);
    const V dx = muls(sub(fBr, fBl), 1.0 / W);
    const V dy = muls(sub(fTl, fBl), 1.0 / H);
    const double v[12] = {fPos.x, fPos.y, fPos.z, fBl.x, fBl.y, fBl.z, dx.x, dx.y, dx.z, dy.x, dy.y, dy.z};
    std::memcpy(out, v, sizeof(v));
    return HD_OK;
}

// ---------------------------------------------------------------------------------------------
// Synthetic scenes
// ---------------------------------------------------------------------------------------------
struct hd_scene {
    hd::Scene s;
};

static TerrainParams to_terrain(const hd_terrain_params* p) {
    TerrainParams t;
    t.levels = p->levels;
    t.footprint = p->footprint;
    t.seed = p->seed;
    t.octaves = p->octaves ? p->octaves : 4;
    t.amplitude = p->amplitude;
    t.frequency = p->frequency;
    t.color_tree_levels = p->color_tree_levels ? p->color_tree_levels : 10;
    t.colors = p->colors != 0;
    t.color_seed = p->color_seed;
    t.threads = p->threads;
    return t;
}

hd_status hd_scene_generate(const hd_terrain_params* p, hd_scene** out) {
    if (!p || !out) return fail(nullptr, HD_ERR_INVALID, "null argument");
    try {
        auto* s = new hd_scene();
        s->s = generate_terrain(to_terrain(p));
        *out = s;
        return HD_OK;
    } catch (const std::exception& ex) {
        return fail(nullptr, HD_ERR_INVALID, ex.what());
    }
}

void hd_scene_free(hd_scene* s) { delete s; }

hd_status hd_scene_sizes(const hd_scene* s, uint64_t sizes[7]) {
    if (!s) return fail(nullptr, HD_ERR_INVALID, "null scene");
    sizes[0] = s->s.dag.size();
    sizes[1] = s->s.color_nodes.size();
    sizes[2] = s->s.color_offsets.size();
    sizes[3] = s->s.blocks.size();
    sizes[4] = s->s.weights.size();
    sizes[5] = s->s.macro.size();
    sizes[6] = s->s.num_voxels;
    return HD_OK;
}

hd_status hd_scene_copy(const hd_scene* s, uint32_t* dag, uint32_t* color_nodes, uint64_t* color_offsets,
                        uint64_t* blocks, uint32_t* weights, uint64_t* macro) {
    if (!s) return fail(nullptr, HD_ERR_INVALID, "null scene");
    if (dag) std::copy(s->s.dag.begin(), s->s.dag.end(), dag);
    if (color_nodes) std::copy(s->s.color_nodes.begin(), s->s.color_nodes.end(), color_nodes);
    if (color_offsets) std::copy(s->s.color_offsets.begin(), s->s.color_offsets.end(), color_offsets);
    if (blocks) std::copy(s->s.blocks.begin(), s->s.blocks.end(), blocks);
    if (weights) std::copy(s->s.weights.begin(), s->s.weights.end(), weights);
    if (macro) std::copy(s->s.macro.begin(), s->s.macro.end(), macro);
    return HD_OK;
}

hd_status hd_scene_level_nodes(const hd_scene* s, uint64_t* out) {
    if (!s) return fail(nullptr, HD_ERR_INVALID, "null scene");
    std::copy(s->s.level_nodes.begin(), s->s.level_nodes.end(), out);
    std::copy(s->s.level_words.begin(), s->s.level_words.end(), out + s->s.level_nodes.size());
    return HD_OK;
}

int32_t hd_terrain_height(const hd_terrain_params* p, uint32_t x, uint32_t z) {
    return terrain_height(to_terrain(p), x, z);
}

}  // extern "C"
