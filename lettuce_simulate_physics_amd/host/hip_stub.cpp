// The HIP factory family for builds without the HIP backend: linked instead of hip_backend.cpp and
// liblettuce_hip.so, as the reference links src/cuda_stub.cpp instead of src/cuda_backend.cu when it
// has no CUDA compiler (CMakeLists.txt:94-108). backend=hip then fails at selection time with `why`,
// and nothing else in the build changes.
#include "backend.hpp"

std::unique_ptr<Backend> make_hip_backend(int, std::string& why) {
    why = "this build has no HIP support (configure with LETTUCE_HIP_DIR pointing at a built liblettuce_hip.so)";
    return nullptr;
}

std::unique_ptr<Backend> make_hip_backend(int device, std::string& why, const HipOptions&) {
    return make_hip_backend(device, why);
}

lhip_ctx* hip_context(Backend&) { return nullptr; }

bool hip_compiled() { return false; }

int hip_device_count() { return 0; }
