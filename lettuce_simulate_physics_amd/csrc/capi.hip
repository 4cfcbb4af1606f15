// C ABI of include/lettuce_hip.h over lhip::Context (context.hpp).
#include <vector>

#include "context.hpp"

// ---------------------------------------------------------------- C ABI

namespace {

using lhip::guarded;

void copy_str(const std::string& s, char* buf, size_t len) {
    if (!buf || len == 0) return;
    const size_t k = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = 0;
}

}  // namespace

extern "C" {

int lhip_device_count(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return count;
}

int lhip_compiled(void) { return 1; }

int lhip_create(int device, lhip_ctx** out, char* why, size_t why_len) {
    if (out) *out = nullptr;
    if (!out) return LHIP_INVALID;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        copy_str(std::string("no usable HIP device: ") + hipGetErrorString(e), why, why_len);
        return LHIP_UNAVAILABLE;
    }
    if (count == 0) {
        copy_str("no HIP device found", why, why_len);
        return LHIP_UNAVAILABLE;
    }
    if (device < 0 || device >= count) {
        copy_str("HIP device " + std::to_string(device) + " out of range (" + std::to_string(count) + " devices)", why,
                 why_len);
        return LHIP_UNAVAILABLE;
    }
    try {
        auto* c = new lhip::Context(device);
        *out = new lhip_ctx{c};
        return LHIP_OK;
    } catch (const lhip::Error& ex) {
        copy_str(ex.what(), why, why_len);
        return ex.code == LHIP_OUT_OF_MEMORY ? LHIP_OUT_OF_MEMORY : LHIP_UNAVAILABLE;
    } catch (const std::exception& ex) {
        copy_str(ex.what(), why, why_len);
        return LHIP_UNAVAILABLE;
    }
}

void lhip_destroy(lhip_ctx* ctx) {
    if (!ctx) return;
    delete ctx->c;
    delete ctx;
}

int lhip_name(const lhip_ctx* ctx, char* buf, size_t len) {
    if (!ctx || !ctx->c) return LHIP_INVALID;
    copy_str(ctx->c->name, buf, len);
    return LHIP_OK;
}

const char* lhip_last_error(const lhip_ctx* ctx) { return ctx && ctx->c ? ctx->c->err.c_str() : "null context"; }

int lhip_upload(lhip_ctx* ctx, uint64_t n, const float* x, const float* y, const float* z, const float* vx,
                const float* vy, const float* vz, const float* m, const uint32_t* id) {
    return guarded(ctx, [&](lhip::Context& c) { c.upload(n, x, y, z, vx, vy, vz, m, id); });
}

uint64_t lhip_size(const lhip_ctx* ctx) { return ctx && ctx->c ? uint64_t(ctx->c->n) : 0; }

int lhip_download(lhip_ctx* ctx, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* m,
                  uint32_t* id) {
    return guarded(ctx, [&](lhip::Context& c) { c.download(x, y, z, vx, vy, vz, m, id); });
}

int lhip_compute_forces(lhip_ctx* ctx, const lhip_force_params* p, int potential) {
    if (!p) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { c.compute_forces(*p, potential != 0); });
}

int lhip_kick(lhip_ctx* ctx, double dt) {
    return guarded(ctx, [&](lhip::Context& c) { c.kick(dt); });
}

int lhip_drift(lhip_ctx* ctx, double dt) {
    return guarded(ctx, [&](lhip::Context& c) { c.drift(dt); });
}

int lhip_diagnostics(lhip_ctx* ctx, lhip_diag* out) {
    if (!out) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { *out = c.diagnostics(); });
}

int lhip_exact_potential_energy(lhip_ctx* ctx, double G, double eps, double* out) {
    if (!out) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { *out = c.exact_pe(G, eps); });
}

int lhip_times(lhip_ctx* ctx, lhip_phase_times* out) {
    if (!out) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { *out = c.times(); });
}

int lhip_local_bounds(lhip_ctx* ctx, float lo[3], float hi[3]) {
    return guarded(ctx, [&](lhip::Context& c) { c.local_bounds(lo, hi); });
}

int lhip_build_tree(lhip_ctx* ctx, const lhip_force_params* p) {
    if (!p) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { c.build_tree(*p); });
}

int lhip_compute_forces_with(lhip_ctx* ctx, const lhip_force_params* p, int potential) {
    if (!p) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { c.compute_forces_with(*p, potential != 0); });
}

int lhip_compute_forces_remote(lhip_ctx* ctx, const lhip_force_params* p, int potential, const lhip_let* remote,
                               int n_remote) {
    if (!p || n_remote < 0 || (n_remote > 0 && !remote)) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { c.compute_forces_with(*p, potential != 0, remote, n_remote); });
}

int lhip_export_let(lhip_ctx* ctx, const float lo[3], const float hi[3], const lhip_force_params* p,
                    uint64_t* n_nodes, uint64_t* n_bodies) {
    if (!lo || !hi || !p || !n_nodes || !n_bodies) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { c.export_let(lo, hi, *p, n_nodes, n_bodies); });
}

int lhip_export_let_copy(lhip_ctx* ctx, lhip_let_node* nodes, float* bx, float* by, float* bz, float* bm) {
    return guarded(ctx, [&](lhip::Context& c) { c.export_let_copy(nodes, bx, by, bz, bm); });
}

int lhip_snapshot_write(lhip_ctx* ctx, const char* path, const lhip_snapshot_header* h) {
    if (!path || !h) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { c.snapshot_write(path, *h); });
}

int lhip_snapshot_restore(lhip_ctx* ctx, const char* path, lhip_snapshot_header* h) {
    if (!path || !h) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { c.snapshot_restore(path, h); });
}

int lhip_memory_need(uint64_t n, int32_t method, uint64_t* device_bytes, uint64_t* host_bytes) {
    return lhip_memory_need_mode(n, method, LHIP_MODE_PLAIN, device_bytes, host_bytes);
}

int lhip_memory_need_mode(uint64_t n, int32_t method, int32_t mode, uint64_t* device_bytes, uint64_t* host_bytes) {
    if (!device_bytes || !host_bytes || (mode != LHIP_MODE_PLAIN && mode != LHIP_MODE_MIXED)) return LHIP_INVALID;
    lhip::memory_need(n, method, mode == LHIP_MODE_MIXED, device_bytes, host_bytes);
    return LHIP_OK;
}

int lhip_device_bytes_held(lhip_ctx* ctx, uint64_t* held) {
    if (!held) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { *held = c.held_bytes(); });
}

int lhip_device_memory(lhip_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!free_bytes || !total_bytes) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { c.device_memory(free_bytes, total_bytes); });
}

int lhip_body_costs(lhip_ctx* ctx, float* cost) {
    if (!cost) return LHIP_INVALID;
    bool ok = false;
    const int rc = guarded(ctx, [&](lhip::Context& c) { ok = c.body_costs(cost); });
    return rc != LHIP_OK ? rc : (ok ? LHIP_OK : LHIP_NOT_PROVIDED);
}

int lhip_accelerations(lhip_ctx* ctx, float* ax, float* ay, float* az, float* pot) {
    return guarded(ctx, [&](lhip::Context& c) { c.accelerations(ax, ay, az, pot); });
}
int lhip_accelerations_f64(lhip_ctx* ctx, double* ax, double* ay, double* az, double* pot) {
    return guarded(ctx, [&](lhip::Context& c) { c.accelerations_f64(ax, ay, az, pot); });
}
int lhip_direct_f64_targets(lhip_ctx* ctx, uint64_t nt, const uint32_t* targets, double G, double eps, double* ax,
                            double* ay, double* az, double* pot) {
    if (nt > 0 && !targets) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { c.direct_f64_targets(nt, targets, G, eps, ax, ay, az, pot); });
}
int lhip_set_precision(lhip_ctx* ctx, int32_t mode, const lhip_walk_options* options) {
    return guarded(ctx, [&](lhip::Context& c) { c.set_precision(mode, options); });
}
int lhip_get_precision(const lhip_ctx* ctx, int32_t* mode, lhip_walk_options* options) {
    if (!ctx || !ctx->c) return LHIP_INVALID;
    if (mode) *mode = ctx->c->mixed ? LHIP_MODE_MIXED : LHIP_MODE_PLAIN;
    if (options) *options = ctx->c->wopt;
    return LHIP_OK;
}
int lhip_upload_f64(lhip_ctx* ctx, uint64_t n, const double* x, const double* y, const double* z, const double* vx,
                    const double* vy, const double* vz, const float* m, const uint32_t* id) {
    if (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !m || !id)) return LHIP_INVALID;
    return guarded(ctx, [&](lhip::Context& c) { c.upload_f64(n, x, y, z, vx, vy, vz, m, id); });
}
int lhip_download_f64(lhip_ctx* ctx, double* x, double* y, double* z, double* vx, double* vy, double* vz, float* m,
                      uint32_t* id) {
    return guarded(ctx, [&](lhip::Context& c) { c.download_f64(x, y, z, vx, vy, vz, m, id); });
}
int lhip_tree_keys(lhip_ctx* ctx, uint64_t* keys) {
    return guarded(ctx, [&](lhip::Context& c) {
        if (!c.tree_ok) throw lhip::Error(LHIP_LOGIC, "no current tree");
        const std::size_t n = std::size_t(c.n);
        LHIP_CHECK(hipMemcpyAsync(keys, c.key.as<void>(), n * 8, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        if (c.frames > 1) {
            // A body of a restarted cell carries the key of its deepest frame, as the reference's
            // key array does after its restarts (src/cpu_backend.cpp:501-529).
            std::vector<uint64_t> fk(std::size_t(c.frames - 1) * n);
            std::vector<uint32_t> cap(n);
            LHIP_CHECK(hipMemcpyAsync(fk.data(), c.fkey.as<void>(), fk.size() * 8, hipMemcpyDeviceToHost, c.stream));
            LHIP_CHECK(hipMemcpyAsync(cap.data(), c.capb.as<void>(), n * 4, hipMemcpyDeviceToHost, c.stream));
            c.sync();
            for (std::size_t i = 0; i < n; ++i) {
                const int r = int(cap[i]) / lhip::kMaxLevel - 1;  // restarts of body i
                if (r > 0) keys[i] = fk[std::size_t(r - 1) * n + i];
            }
        }
    });
}

int lhip_tree_counts(lhip_ctx* ctx, uint64_t* nodes, uint64_t* groups, uint64_t* overfull) {
    return guarded(ctx, [&](lhip::Context& c) {
        if (!c.tree_ok) throw lhip::Error(LHIP_LOGIC, "no current tree");
        uint32_t M = 0, G = 0;
        lhip::WalkCounters wc{};
        LHIP_CHECK(hipMemcpyAsync(&M, c.cnt.as<uint32_t>() + c.n, 4, hipMemcpyDeviceToHost, c.stream));
        LHIP_CHECK(hipMemcpyAsync(&G, c.gflag.as<uint32_t>() + c.n, 4, hipMemcpyDeviceToHost, c.stream));
        LHIP_CHECK(hipMemcpyAsync(&wc, c.counters.as<void>(), sizeof wc, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        if (nodes) *nodes = M;
        if (groups) *groups = G;
        if (overfull) *overfull = uint64_t(wc.overfull);
    });
}

int lhip_tree_nodes(lhip_ctx* ctx, int32_t* link, uint64_t* prefix, float* mom) {
    return guarded(ctx, [&](lhip::Context& c) {
        if (!c.tree_ok) throw lhip::Error(LHIP_LOGIC, "no current tree");
        uint32_t M = 0;
        LHIP_CHECK(hipMemcpyAsync(&M, c.cnt.as<uint32_t>() + c.n, 4, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        if (link) LHIP_CHECK(hipMemcpy2DAsync(link, 16, c.node.as<char>() + 16, sizeof(lhip::WalkNode), 16, M,
                                              hipMemcpyDeviceToHost, c.stream));
        if (prefix) LHIP_CHECK(hipMemcpyAsync(prefix, c.prefix.as<void>(), std::size_t(M) * 8, hipMemcpyDeviceToHost, c.stream));
        if (mom) LHIP_CHECK(hipMemcpyAsync(mom, c.mom.as<void>(), std::size_t(M) * 64, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

int lhip_tree_groups(lhip_ctx* ctx, uint32_t* gstart) {
    return guarded(ctx, [&](lhip::Context& c) {
        if (!c.tree_ok) throw lhip::Error(LHIP_LOGIC, "no current tree");
        uint32_t G = 0;
        LHIP_CHECK(hipMemcpyAsync(&G, c.gflag.as<uint32_t>() + c.n, 4, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        LHIP_CHECK(hipMemcpyAsync(gstart, c.gstart.as<void>(), (std::size_t(G) + 1) * 4, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

}  // extern "C"
