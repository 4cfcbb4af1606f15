// HipBackend: the reference Backend interface (src/backend.hpp:89-141) implemented over the C ABI
// of liblettuce_hip.so (include/lettuce_hip.h). C++ side only: no HIP headers, no device code.
// Errors keep the reference's exception types: device/runtime failures std::runtime_error,
// protocol misuse std::logic_error (src/cuda_backend.cu:45-51, 1399-1400, 1570-1571); the factory
// never throws and reports `why` (src/cuda_backend.cu:2026-2048).
#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <string>

#include "backend.hpp"
#include "lettuce_hip.h"

namespace {

static_assert(sizeof(LetNode) == sizeof(lhip_let_node), "LetNode and lhip_let_node are one record");
static_assert(offsetof(LetNode, rcrit2) == offsetof(lhip_let_node, rcrit2) &&
                  offsetof(LetNode, first_child) == offsetof(lhip_let_node, first_child) &&
                  offsetof(LetNode, body_count) == offsetof(lhip_let_node, body_count),
              "LetNode and lhip_let_node share their layout");

lhip_force_params to_c(const ForceParams& p) {
    lhip_force_params c{};
    c.G = p.G;
    c.eps = p.eps;
    c.theta = p.theta;
    c.method = p.method == ForceMethod::tree ? LHIP_METHOD_TREE : LHIP_METHOD_DIRECT;
    return c;
}

class HipBackend final : public Backend {
public:
    explicit HipBackend(lhip_ctx* c) : c_(c) {}
    lhip_ctx* ctx() const { return c_; }
    ~HipBackend() override { lhip_destroy(c_); }
    HipBackend(const HipBackend&) = delete;
    HipBackend& operator=(const HipBackend&) = delete;

    std::string name() const override {
        char buf[512];
        lhip_name(c_, buf, sizeof buf);
        return buf;
    }

    void upload(const Bodies& b) override {
        ck(lhip_upload(c_, b.size(), b.x.data(), b.y.data(), b.z.data(), b.vx.data(), b.vy.data(), b.vz.data(),
                       b.m.data(), b.id.data()));
    }

    void download(Bodies& b) override {
        b.resize(lhip_size(c_));
        ck(lhip_download(c_, b.x.data(), b.y.data(), b.z.data(), b.vx.data(), b.vy.data(), b.vz.data(), b.m.data(),
                         b.id.data()));
    }

    void compute_forces(const ForceParams& p, bool potential) override {
        const lhip_force_params c = to_c(p);
        ck(lhip_compute_forces(c_, &c, potential ? 1 : 0));
    }

    void kick(double dt) override { ck(lhip_kick(c_, dt)); }
    void drift(double dt) override { ck(lhip_drift(c_, dt)); }

    Diag diagnostics() override {
        lhip_diag d{};
        ck(lhip_diagnostics(c_, &d));
        Diag o;
        o.mass = d.mass; o.ke = d.ke; o.pe = d.pe; o.has_pe = d.has_pe != 0;
        o.px = d.px; o.py = d.py; o.pz = d.pz;
        o.lx = d.lx; o.ly = d.ly; o.lz = d.lz;
        o.cx = d.cx; o.cy = d.cy; o.cz = d.cz;
        return o;
    }

    double exact_potential_energy(double G, double eps) override {
        double v = 0;
        ck(lhip_exact_potential_energy(c_, G, eps, &v));
        return v;
    }

    PhaseTimes times() const override {
        lhip_phase_times t{};
        ck(lhip_times(c_, &t));
        PhaseTimes o;
        o.force_ms = t.force_ms; o.sort_ms = t.sort_ms; o.build_ms = t.build_ms; o.walk_ms = t.walk_ms;
        o.step_ms = t.step_ms; o.interactions = t.interactions; o.comm_ms = t.comm_ms; o.balance_ms = t.balance_ms;
        return o;
    }

    Box3 local_bounds() override {
        Box3 b;
        ck(lhip_local_bounds(c_, b.lo, b.hi));
        return b;
    }

    void build_tree(const ForceParams& p) override {
        const lhip_force_params c = to_c(p);
        ck(lhip_build_tree(c_, &c));
    }

    // LETs travel as the reference's LetNode records; lhip_let_node has the same layout, so the
    // records pass through the ABI without conversion.
    LetTree export_let(const Box3& target, const ForceParams& p) override {
        const lhip_force_params c = to_c(p);
        uint64_t nn = 0, nb = 0;
        ck(lhip_export_let(c_, target.lo, target.hi, &c, &nn, &nb));
        LetTree out;
        out.nodes.resize(nn);
        out.bx.resize(nb); out.by.resize(nb); out.bz.resize(nb); out.bm.resize(nb);
        ck(lhip_export_let_copy(c_, reinterpret_cast<lhip_let_node*>(out.nodes.data()), out.bx.data(), out.by.data(),
                                out.bz.data(), out.bm.data()));
        return out;
    }

    void compute_forces_with(const ForceParams& p, bool potential, const std::vector<const LetTree*>& remote) override {
        const lhip_force_params c = to_c(p);
        if (remote.empty()) {
            ck(lhip_compute_forces_with(c_, &c, potential ? 1 : 0));
            return;
        }
        std::vector<lhip_let> lets;
        lets.reserve(remote.size());
        for (const LetTree* t : remote) {
            if (!t) continue;
            const std::size_t nb = t->bx.size();
            if (t->by.size() != nb || t->bz.size() != nb || t->bm.size() != nb)
                throw std::runtime_error("hip compute_forces_with: LET body arrays differ in length");
            lets.push_back(lhip_let{reinterpret_cast<const lhip_let_node*>(t->nodes.data()), t->nodes.size(),
                                    t->bx.data(), t->by.data(), t->bz.data(), t->bm.data(), nb});
        }
        ck(lhip_compute_forces_remote(c_, &c, potential ? 1 : 0, lets.data(), int(lets.size())));
    }

    // For this backend's mode: the mixed-precision form holds the FP64 state besides the f32 one.
    MemoryNeed memory_need(uint64_t n, ForceMethod m) const override {
        MemoryNeed r;
        int32_t mode = LHIP_MODE_PLAIN;
        lhip_get_precision(c_, &mode, nullptr);
        lhip_memory_need_mode(n, m == ForceMethod::tree ? LHIP_METHOD_TREE : LHIP_METHOD_DIRECT, mode, &r.device,
                              &r.host);
        return r;
    }

    bool device_memory(uint64_t& free_bytes, uint64_t& total_bytes) override {
        ck(lhip_device_memory(c_, &free_bytes, &total_bytes));
        return true;
    }

    bool body_costs(std::vector<float>& cost) override {
        std::vector<float> c(lhip_size(c_));
        if (c.empty()) return false;
        const int rc = lhip_body_costs(c_, c.data());
        if (rc == LHIP_NOT_PROVIDED) return false;
        ck(rc);
        cost.swap(c);
        return true;
    }

private:
    void ck(int rc) const {
        if (rc == LHIP_OK) return;
        const std::string msg = lhip_last_error(c_);
        if (rc == LHIP_LOGIC) throw std::logic_error(msg);
        throw std::runtime_error(msg.empty() ? "hip backend error " + std::to_string(rc) : msg);
    }
    lhip_ctx* c_;
};

}  // namespace

std::unique_ptr<Backend> make_hip_backend(int device, std::string& why) {
    return make_hip_backend(device, why, HipOptions{});
}

std::unique_ptr<Backend> make_hip_backend(int device, std::string& why, const HipOptions& opt) {
    lhip_ctx* c = nullptr;
    char buf[512] = {0};
    const int rc = lhip_create(device, &c, buf, sizeof buf);
    if (rc != LHIP_OK || !c) {
        why = buf[0] ? buf : "hip backend unavailable";
        return nullptr;
    }
    lhip_walk_options w{};
    const int mode = opt.mixed ? LHIP_MODE_MIXED : LHIP_MODE_PLAIN;
    lhip_set_precision(c, mode, nullptr);  // the mode's defaults, then any explicit option
    lhip_get_precision(c, nullptr, &w);
    if (opt.group_max >= 0) w.group_max = opt.group_max;
    if (opt.small_leaf >= 0) w.small_leaf = opt.small_leaf;
    if (opt.quad_tau >= 0) w.quad_tau = opt.quad_tau;
    if (lhip_set_precision(c, mode, &w) != LHIP_OK) {
        why = lhip_last_error(c);
        lhip_destroy(c);
        return nullptr;
    }
    return std::make_unique<HipBackend>(c);
}

lhip_ctx* hip_context(Backend& be) {
    auto* h = dynamic_cast<HipBackend*>(&be);
    return h ? h->ctx() : nullptr;
}

bool hip_compiled() { return lhip_compiled() != 0; }
int hip_device_count() { return lhip_device_count(); }
