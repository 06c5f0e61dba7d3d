// Request serving over one device-resident HashDAG: a host scheduler that admits trace and
// edit requests as they arrive and keeps the GPU busy with continuous batching.
//
//   order     requests take effect in arrival order: a trace sees exactly the edits that
//             arrived before it, so its frame is the same alone and in any mix
//   batching  consecutive edit requests at the head of the queue are fused into one edit
//             pass; consecutive trace requests run concurrently, one per trace slot (its own
//             stream and frame buffers); whatever arrives while work is on the GPU joins the
//             next batch
//   overlap   an edit pass waits (on the device, through events) only for the traces that
//             precede it; traces behind a pass start as soon as it has committed
//   admission an edit batch is admitted when the free page budget covers an estimate of what
//             it appends; otherwise stale nodes are collected first (GC), and a batch that still
//             does not fit, or overflows a bucket, is rejected with its status
//
// There is no counterpart in the reference (its engine traces one frame and applies edits
// from the UI thread, engine.cpp:575-718); the trace and edit calls are the ones that replace
// DAGTracer::resolve_* and HashDAG::edit_threads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <thread>
#include <vector>

#include "hd_ctx.h"

namespace hd {
namespace {

// Order-independent 64-bit frame checksum: sum over pixels of fmix64(pixel word, index).
__global__ __launch_bounds__(256) void k_checksum(const uint32_t* __restrict__ px, uint64_t n,
                                                  unsigned long long* __restrict__ out) {
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        acc += fmix64(((uint64_t)px[i] << 32) ^ (i * 0x9E3779B97F4A7C15ull));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor((unsigned long long)acc, off, 64);
    if ((threadIdx.x & 63u) == 0) atomicAdd(out, (unsigned long long)acc);
}

__global__ void k_zero64(unsigned long long* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// Per bucket: used words, and mapped pages past the used ones (pages are mapped in order as a
// bucket grows, so the mapped ones form a prefix). out[0] += used words, out[1] += reusable.
__global__ __launch_bounds__(256) void k_mem_info(Geom g, const uint32_t* __restrict__ bucket_sizes,
                                                  const uint32_t* __restrict__ page_table,
                                                  unsigned long long* __restrict__ out) {
    const uint64_t nb = g.total_buckets();
    unsigned long long used = 0, reusable = 0;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t top = g.top();
        const uint64_t topBuckets = (uint64_t)top << g.bucket_bits_top;
        uint32_t l, bucket;
        if (b < topBuckets) {
            l = (uint32_t)(b >> g.bucket_bits_top);
            bucket = (uint32_t)(b & ((1u << g.bucket_bits_top) - 1u));
        } else {
            l = top + (uint32_t)((b - topBuckets) >> g.bucket_bits_low);
            bucket = (uint32_t)((b - topBuckets) & ((1u << g.bucket_bits_low) - 1u));
        }
        const uint32_t bs = bucket_sizes[b];
        used += bs;
        const uint32_t pages = g.bucket_size(l) / g.page_size;
        const uint32_t usedPages = (bs + g.page_size - 1) / g.page_size;
        const uint64_t first = g.make_ptr(l, bucket, 0) / g.page_size;
        for (uint32_t k = usedPages; k < pages && page_table[first + k] != 0; ++k) reusable++;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        used += __shfl_xor(used, off, 64);
        reusable += __shfl_xor(reusable, off, 64);
    }
    if ((threadIdx.x & 63u) == 0) {
        atomicAdd(out + 0, used);
        atomicAdd(out + 1, reusable);
    }
}

using Clock = std::chrono::steady_clock;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint32_t* paths = nullptr;
    uint32_t* colors = nullptr;
    unsigned long long* sum = nullptr;  // device checksum
    unsigned long long* hsum = nullptr;  // pinned host copy
    int64_t req = -1;                    // request in flight, -1 = free
};

// Pages an edit batch may append: boundary leaves of each editor (surface / 16 voxels per leaf
// face) at 2 words plus about as many words of interior nodes above them.
uint64_t estimate_pages(const hd_edit* e, uint32_t n, uint32_t pageSize) {
    double words = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const double r = std::max(1.0f, e[i].radius);
        words += 4.0 * 3.14159265358979 * r * r / 16.0 * 4.0 + 64.0;
    }
    return (uint64_t)(words / pageSize) + 1;
}

}  // namespace
}  // namespace hd

using namespace hd;

extern "C" {

hd_status hd_frame_checksum(hd_ctx* ctx, const uint32_t* d_colors, uint32_t width, uint32_t height, uint64_t* out,
                            void* stream) {
    if (!ctx || !d_colors || !out) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    unsigned long long* d = reinterpret_cast<unsigned long long*>(ctx->byte_counters + 3);
    k_zero64<<<1, 64, 0, s>>>(d);
    const uint64_t n = (uint64_t)width * height;
    k_checksum<<<(uint32_t)std::min<uint64_t>((n + 255) / 256, 2048), 256, 0, s>>>(d_colors, n, d);
    HD_HIP(ctx, hipGetLastError());
    unsigned long long v = 0;
    HD_HIP(ctx, hipMemcpyAsync(&v, d, 8, hipMemcpyDeviceToHost, s));
    HD_HIP(ctx, hipStreamSynchronize(s));
    *out = v;
    return HD_OK;
}

hd_status hd_mem_info(hd_ctx* ctx, hd_mem_info_t* out) {
    if (!ctx || !out) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded) return fail(ctx, HD_ERR_INVALID, "no DAG loaded");
    unsigned long long* d = reinterpret_cast<unsigned long long*>(ctx->byte_counters + 4);
    HD_HIP(ctx, hipMemsetAsync(d, 0, 16, ctx->stream));
    k_mem_info<<<1024, 256, 0, ctx->stream>>>(ctx->g, ctx->bucket_sizes, ctx->page_table, d);
    HD_HIP(ctx, hipGetLastError());
    unsigned long long v[2];
    uint32_t top = 0;
    HD_HIP(ctx, hipMemcpyAsync(v, d, 16, hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipMemcpyAsync(&top, ctx->counters, 4, hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    out->pool_pages = ctx->cfg.pool_pages;
    out->pool_top = top;
    out->used_words = v[0];
    out->reusable_pages = v[1];
    out->free_pages = (top < ctx->cfg.pool_pages ? ctx->cfg.pool_pages - top : 0) + v[1];
    out->arena_words = ctx->cfg.color_arena_words;
    out->arena_top = ctx->arena_top;
    return HD_OK;
}

hd_status hd_serve_run(hd_ctx* ctx, const hd_request* reqs, uint32_t n, const hd_edit* edits, uint32_t n_edits,
                       const hd_serve_config* cfg, hd_request_result* out) {
    if (!ctx || (!reqs && n) || !cfg || (!out && n)) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded || !ctx->colors_loaded) return fail(ctx, HD_ERR_INVALID, "no DAG or colours loaded");
    const uint32_t nSlots = std::max<uint32_t>(1, std::min<uint32_t>(cfg->trace_slots, 16));
    uint64_t maxPixels = 1;
    for (uint32_t i = 0; i < n; ++i) {
        const hd_request& r = reqs[i];
        if (r.kind == HD_REQ_TRACE) {
            if (r.width == 0 || r.height == 0) return fail(ctx, HD_ERR_INVALID, "trace request without a frame size");
            maxPixels = std::max<uint64_t>(maxPixels, (uint64_t)r.width * r.height);
        } else if (r.kind == HD_REQ_EDIT) {
            if ((uint64_t)r.first_edit + r.n_edits > n_edits) return fail(ctx, HD_ERR_INVALID, "edit range out of bounds");
        } else if (r.kind != HD_REQ_UNDO && r.kind != HD_REQ_REDO) {
            return fail(ctx, HD_ERR_INVALID, "unknown request kind");
        }
        if (i && r.arrival_s < reqs[i - 1].arrival_s) return fail(ctx, HD_ERR_INVALID, "requests not in arrival order");
    }
    // trace slots
    std::vector<Slot> slots(nSlots);
    hipStream_t editStream = nullptr;
    auto cleanup = [&]() {
        for (Slot& s : slots) {
            if (s.stream) hipStreamSynchronize(s.stream);
            hipFree(s.paths);
            hipFree(s.colors);
            hipFree(s.sum);
            hipHostFree(s.hsum);
            if (s.done) hipEventDestroy(s.done);
            if (s.stream) hipStreamDestroy(s.stream);
        }
        if (editStream) {
            hipStreamSynchronize(editStream);
            hipStreamDestroy(editStream);
        }
    };
    hipError_t e = hipStreamCreateWithFlags(&editStream, hipStreamNonBlocking);
    for (Slot& s : slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(&s.paths, maxPixels * 16);
        if (e == hipSuccess) e = hipMalloc(&s.colors, maxPixels * 4);
        if (e == hipSuccess) e = hipMalloc(&s.sum, 8);
        if (e == hipSuccess) e = hipHostMalloc(&s.hsum, 8);
    }
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(ctx, e, "serving slots");
    }
    for (uint32_t i = 0; i < n; ++i) out[i] = hd_request_result{-1.0, -1.0, 0, 0, 0, HD_OK, 0, 0};

    const auto t0 = Clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    uint32_t next = 0;           // next request to arrive
    std::deque<uint32_t> queue;  // arrived, not started (arrival order)
    uint32_t finished = 0, batch = 0;
    hd_status result = HD_OK;
    // a HIP error ends the run with its status; cleanup() below still runs
    auto hip_ok = [&](hipError_t err, const char* what) {
        if (err == hipSuccess) return true;
        if (result == HD_OK) result = hip_fail(ctx, err, what);
        return false;
    };

    auto poll_slots = [&](bool wait_one) {
        for (;;) {
            bool any = false, busy = false;
            for (Slot& s : slots) {
                if (s.req < 0) continue;
                busy = true;
                if (hipEventQuery(s.done) == hipSuccess) {
                    hd_request_result& r = out[s.req];
                    r.done_s = now_s();
                    r.checksum = *s.hsum;
                    s.req = -1;
                    ++finished;
                    any = true;
                }
            }
            if (!wait_one || any || !busy) return;
            std::this_thread::yield();
        }
    };

    while (finished < n && result == HD_OK) {
        // admit arrivals
        const double t = now_s();
        while (next < n && (!cfg->realtime || reqs[next].arrival_s <= t)) queue.push_back(next++);
        poll_slots(false);
        if (queue.empty()) {
            if (next < n) {
                // sleep until the next arrival, but keep noticing finished traces while any runs
                bool busy = false;
                for (const Slot& sl : slots) busy |= sl.req >= 0;
                const double wait = reqs[next].arrival_s - now_s();
                const double nap = busy ? std::min(wait - 100e-6, 20e-6) : wait - 100e-6;
                if (nap > 0) std::this_thread::sleep_for(std::chrono::duration<double>(nap));
                else std::this_thread::yield();
            } else {
                poll_slots(true);
            }
            continue;
        }
        const uint32_t head = queue.front();
        if (reqs[head].kind == HD_REQ_UNDO || reqs[head].kind == HD_REQ_REDO) {
            // undo / redo swap the root and colour leaves in place: everything queued before
            // finishes first, everything after starts after
            queue.pop_front();
            const double start = now_s();
            bool ok = hip_ok(hipStreamSynchronize(editStream), "serving: edit stream");
            for (Slot& sl : slots)
                if (ok && sl.req >= 0) ok = hip_ok(hipStreamSynchronize(sl.stream), "serving: trace slot");
            if (!ok) break;
            poll_slots(false);
            const hd_status st = reqs[head].kind == HD_REQ_UNDO ? hd_undo(ctx) : hd_redo(ctx);
            out[head].start_s = start;
            out[head].done_s = now_s();
            out[head].status = st;
            out[head].batch = batch++;
            out[head].checksum = ctx->root;
            ++finished;
            if (st != HD_OK) result = st;
            continue;
        }
        if (reqs[head].kind == HD_REQ_EDIT) {
            // fuse the run of edit requests at the head of the queue
            std::vector<hd_edit> fused;
            std::vector<uint32_t> members;
            while (!queue.empty() && reqs[queue.front()].kind == HD_REQ_EDIT) {
                const hd_request& r = reqs[queue.front()];
                if (!members.empty() && fused.size() + r.n_edits > cfg->max_batch_edits) break;
                fused.insert(fused.end(), edits + r.first_edit, edits + r.first_edit + r.n_edits);
                members.push_back(queue.front());
                queue.pop_front();
            }
            // the pass waits on the device for the traces queued before it
            bool ok = true;
            for (Slot& s : slots)
                if (ok && s.req >= 0) ok = hip_ok(hipStreamWaitEvent(editStream, s.done, 0), "serving: wait for traces");
            if (!ok) break;
            const double start = now_s();
            for (uint32_t m : members) {
                out[m].start_s = start;
                out[m].batch = batch;
            }
            // admission by free pages: collect stale nodes below the watermark; reject a batch
            // whose estimate does not fit even after that
            if (!hip_ok(hipStreamSynchronize(editStream), "serving: edit stream")) break;
            hd_mem_info_t mem;
            hd_status st = hd_mem_info(ctx, &mem);
            const uint64_t need = estimate_pages(fused.data(), (uint32_t)fused.size(), ctx->g.page_size);
            uint32_t gcRan = 0;
            // a GC that fails leaves the table unusable (the context refuses further calls), so
            // it ends the run with its status rather than rejecting one batch
            hd_status gcSt = HD_OK;
            if (st == HD_OK && mem.free_pages < need + cfg->reserve_pages) {
                gcSt = hd_gc_roots(ctx, ctx->g.leaf_level(), cfg->keep_undo, nullptr, 0, nullptr, editStream);
                gcRan = 1;
                st = gcSt;
                if (st == HD_OK) st = hd_mem_info(ctx, &mem);
                if (st == HD_OK && mem.free_pages < need) st = HD_ERR_POOL_FULL;
            }
            hd_edit_stats es{};
            if (st == HD_OK) st = hd_edit_batch(ctx, fused.data(), (uint32_t)fused.size(), 1, &es, editStream);
            if ((st == HD_ERR_BUCKET_FULL || st == HD_ERR_POOL_FULL) && !gcRan) {
                // a full bucket or page budget: collect stale nodes once and retry
                gcRan = 1;
                gcSt = hd_gc_roots(ctx, ctx->g.leaf_level(), cfg->keep_undo, nullptr, 0, nullptr, editStream);
                if (gcSt == HD_OK) st = hd_edit_batch(ctx, fused.data(), (uint32_t)fused.size(), 1, &es, editStream);
                else st = gcSt;
            }
            if (st == HD_OK && !cfg->keep_undo) hd_undo_clear(ctx);
            const double done = now_s();
            for (uint32_t m : members) {
                out[m].done_s = done;
                out[m].status = st;
                out[m].voxels = es.num_voxels;
                out[m].checksum = es.root;
                out[m].gc = gcRan;
                ++finished;
            }
            if (gcSt != HD_OK) result = gcSt;
            else if (st != HD_OK && st != HD_ERR_BUCKET_FULL && st != HD_ERR_POOL_FULL) result = st;
            ++batch;
        } else {
            // run the traces at the head of the queue, one per free slot
            bool launched = false;
            while (!queue.empty() && reqs[queue.front()].kind == HD_REQ_TRACE) {
                Slot* free = nullptr;
                for (Slot& s : slots)
                    if (s.req < 0) {
                        free = &s;
                        break;
                    }
                if (!free) break;
                const uint32_t q = queue.front();
                queue.pop_front();
                const hd_request& r = reqs[q];
                out[q].start_s = now_s();
                out[q].batch = batch;
                hd_status st = hd_trace_paths(ctx, &r.params, r.width, r.height, free->paths, free->stream);
                if (st == HD_OK) st = hd_trace_colors(ctx, r.width, r.height, free->paths, free->colors, free->stream);
                if (st == HD_OK)
                    st = hd_trace_shadows(ctx, &r.params, r.width, r.height, free->paths, free->colors, cfg->shadow_bias,
                                          cfg->fog_density, free->stream);
                if (st != HD_OK) {
                    result = st;
                    break;
                }
                const uint64_t px = (uint64_t)r.width * r.height;
                k_zero64<<<1, 64, 0, free->stream>>>(free->sum);
                k_checksum<<<(uint32_t)std::min<uint64_t>((px + 255) / 256, 2048), 256, 0, free->stream>>>(free->colors, px,
                                                                                                           free->sum);
                if (!hip_ok(hipMemcpyAsync(free->hsum, free->sum, 8, hipMemcpyDeviceToHost, free->stream),
                            "serving: checksum readback") ||
                    !hip_ok(hipEventRecord(free->done, free->stream), "serving: trace event"))
                    break;
                out[q].voxels = px;
                free->req = q;
                launched = true;
            }
            if (launched) ++batch;
            else poll_slots(true);  // every slot busy: wait for one
        }
    }
    // drain
    while (finished < n && result == HD_OK) poll_slots(true);
    cleanup();
    return result;
}

}  // extern "C"
