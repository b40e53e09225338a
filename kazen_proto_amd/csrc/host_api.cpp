// Host-only part of the C ABI: BVH8 builds and queries that need no HIP
// device (kp_bvh8_*), and the per-thread error text that kp_last_error(NULL)
// returns.  This file includes no HIP header, so it builds with any host
// compiler (and with the sanitizers).
#include <string>

#include "../../include/kp_hip.h"
#include "host_trace.h"
#include "kp_bvh.h"

namespace kp {
namespace {
thread_local std::string g_thread_error;
}

void set_thread_error(const std::string& msg) { g_thread_error = msg; }
const char* thread_error() { return g_thread_error.c_str(); }

}  // namespace kp

struct kp_bvh8_blob {
    kp::HostBvh h;
    uint32_t max_depth = 0;
};

extern "C" {

int kp_bvh8_build_host(const float* tris9, uint32_t n, const kp_build_opts* o, kp_bvh8_blob** out) {
    if (!out || !tris9) {
        kp::set_thread_error("kp_bvh8_build_host: NULL argument");
        return KP_ERR_INVALID;
    }
    *out = nullptr;
    kp::BuildOptions bo;
    if (o) {
        if (o->leaf_size) bo.leaf_size = o->leaf_size;
        if (o->sah_bins) bo.sah_bins = o->sah_bins;
        if (o->node_width > 8 || o->node_width == 1) {   // no device to tune on: AUTO is not accepted here
            kp::set_thread_error("kp_bvh8_build_host: node_width must be 2..8 or 0");
            return KP_ERR_INVALID;
        }
        if (o->node_width) bo.max_children = o->node_width;
    }
    kp::Bvh8Host h;
    std::string err;
    if (!kp::build_bvh8(tris9, n, bo, h, err)) {
        kp::set_thread_error("kp_bvh8_build_host: " + err);
        return KP_ERR_INVALID;
    }
    kp_bvh8_blob* b = new kp_bvh8_blob();
    b->h.nodes = std::move(h.nodes);
    b->h.tris = std::move(h.tris);
    b->max_depth = h.max_depth;
    *out = b;
    return KP_OK;
}

int kp_bvh8_blob_info(const kp_bvh8_blob* b, uint64_t* nn, uint64_t* nt, uint32_t* md) {
    if (!b) return KP_ERR_INVALID;
    if (nn) *nn = b->h.nodes.size();
    if (nt) *nt = b->h.tris.size() / 12;
    if (md) *md = b->max_depth;
    return KP_OK;
}

const void* kp_bvh8_blob_nodes(const kp_bvh8_blob* b) { return b ? b->h.nodes.data() : nullptr; }
const void* kp_bvh8_blob_tris(const kp_bvh8_blob* b) { return b ? b->h.tris.data() : nullptr; }
void kp_bvh8_blob_free(kp_bvh8_blob* b) { delete b; }

int kp_bvh8_blob_closest(const kp_bvh8_blob* b, const kp_ray* rays, uint32_t n, kp_hit* hits) {
    if (!b || (n && (!rays || !hits))) {
        kp::set_thread_error("kp_bvh8_blob_closest: NULL argument");
        return KP_ERR_INVALID;
    }
    kp::host_closest(b->h, (const float*)rays, n, (float*)hits);
    return KP_OK;
}

int kp_bvh8_blob_occluded(const kp_bvh8_blob* b, const kp_ray* rays, uint32_t n, uint32_t* out) {
    if (!b || (n && (!rays || !out))) {
        kp::set_thread_error("kp_bvh8_blob_occluded: NULL argument");
        return KP_ERR_INVALID;
    }
    kp::host_occluded(b->h, (const float*)rays, n, out);
    return KP_OK;
}

}  // extern "C"
