// The stub build: backend.hpp + hip_stub.cpp only (no liblettuce_hip.so, no HIP runtime). The factory
// returns nullptr and says why, as the reference's cuda_stub.cpp does (src/cuda_stub.cpp:4-11).
#include <cstdio>
#include <string>

#include "backend.hpp"

int main() {
    int fails = 0;
    std::string why;
    auto be = make_hip_backend(0, why);
    if (be) ++fails, std::fprintf(stderr, "make_hip_backend returned a backend in the stub build\n");
    if (why.find("no HIP support") == std::string::npos) ++fails, std::fprintf(stderr, "why: '%s'\n", why.c_str());
    std::string why2;
    if (make_hip_backend(0, why2, HipOptions{true, 64, 1, 0.05}) || why2 != why) ++fails;
    if (hip_compiled()) ++fails;
    if (hip_device_count() != 0) ++fails;
    std::fprintf(stderr, "stub: %s; %d failed\n", why.c_str(), fails);
    return fails == 0 ? 0 : 1;
}
