// The reference's force-backend interface, restated so this repository builds standalone. Names,
// members, argument meaning and error contract are kept identical to the reference's
// src/backend.hpp:12-148 and src/bodies.hpp:6-35 on purpose: HipBackend must be a drop-in Backend.
// In the reference tree this header is not needed; hip_backend.cpp compiles against theirs.
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

struct Bodies {
    std::vector<float> x, y, z;
    std::vector<float> vx, vy, vz;
    std::vector<float> m;
    std::vector<uint32_t> id;
    std::size_t size() const { return x.size(); }
    void resize(std::size_t n) {
        for (auto* v : {&x, &y, &z, &vx, &vy, &vz, &m}) v->resize(n);
        id.resize(n);
    }
    void reserve(std::size_t n) {
        for (auto* v : {&x, &y, &z, &vx, &vy, &vz, &m}) v->reserve(n);
        id.reserve(n);
    }
    void push(double px, double py, double pz, double pvx, double pvy, double pvz, double pm) {
        id.push_back(static_cast<uint32_t>(x.size()));
        x.push_back(float(px)); y.push_back(float(py)); z.push_back(float(pz));
        vx.push_back(float(pvx)); vy.push_back(float(pvy)); vz.push_back(float(pvz));
        m.push_back(float(pm));
    }
};

enum class ForceMethod { direct, tree };

struct ForceParams {
    double G = 1;
    double eps = 0.01;
    double theta = 0.6;
    ForceMethod method = ForceMethod::direct;
};

struct MemoryNeed {
    uint64_t device = 0;
    uint64_t host = 0;
};

struct Diag {
    double mass = 0, ke = 0, pe = 0;
    bool has_pe = false;
    double px = 0, py = 0, pz = 0;
    double lx = 0, ly = 0, lz = 0;
    double cx = 0, cy = 0, cz = 0;
};

struct PhaseTimes {
    double force_ms = 0, sort_ms = 0, build_ms = 0, walk_ms = 0, step_ms = 0;
    uint64_t interactions = 0;
    double comm_ms = 0, balance_ms = 0;
};

struct Box3 {
    float lo[3] = {0, 0, 0};
    float hi[3] = {0, 0, 0};
};

struct LetNode {
    float cx, cy, cz;
    float mass;
    float qxx, qxy, qxz, qyy, qyz, qzz;
    float rcrit2;
    int32_t first_child, child_count, body_first, body_count;
};
static_assert(sizeof(LetNode) == 60, "LetNode is a 60-byte wire record");

struct LetTree {
    std::vector<LetNode> nodes;
    std::vector<float> bx, by, bz, bm;
};

class Backend {
public:
    virtual ~Backend() = default;
    virtual std::string name() const = 0;
    virtual void upload(const Bodies& b) = 0;
    virtual void download(Bodies& b) = 0;
    virtual void compute_forces(const ForceParams& p, bool potential) = 0;
    virtual void kick(double dt) = 0;
    virtual void drift(double dt) = 0;
    virtual Diag diagnostics() = 0;
    virtual double exact_potential_energy(double G, double eps) = 0;
    virtual PhaseTimes times() const = 0;
    virtual Box3 local_bounds() = 0;
    virtual void build_tree(const ForceParams& p) = 0;
    virtual LetTree export_let(const Box3& target, const ForceParams& p) = 0;
    virtual void compute_forces_with(const ForceParams& p, bool potential,
                                     const std::vector<const LetTree*>& remote) = 0;
    virtual MemoryNeed memory_need(uint64_t n, ForceMethod m) const { (void)n; (void)m; return {}; }
    virtual bool device_memory(uint64_t& free_bytes, uint64_t& total_bytes) {
        (void)free_bytes; (void)total_bytes;
        return false;
    }
    virtual bool body_costs(std::vector<float>& cost) { (void)cost; return false; }
};

// The HIP factory family, beside the reference's make_cpu_backend/make_cuda_backend.
std::unique_ptr<Backend> make_hip_backend(int device, std::string& why);
// Options of the HIP backend with no reference counterpart: the mixed-precision form (FP64 state,
// FP32 relative-coordinate pair arithmetic) and the walk options; negative values keep the mode's
// defaults (plain: the reference walk, group 128, small leaf 4, every quadrupole; mixed: group 128,
// small leaf 1, quadrupoles dropped below 0.05 of the monopole term).
struct HipOptions {
    bool mixed = false;
    int group_max = -1;
    int small_leaf = -1;
    double quad_tau = -1;
};
std::unique_ptr<Backend> make_hip_backend(int device, std::string& why, const HipOptions& opt);
bool hip_compiled();
int hip_device_count();
// The C ABI context behind a backend that make_hip_backend made (nullptr for any other backend), for
// the calls that have no Backend member: the FP64 state and its snapshots (lhip_upload_f64,
// lhip_snapshot_write, lhip_snapshot_restore).
struct lhip_ctx;
lhip_ctx* hip_context(Backend& be);
