// Snapshot files (host code): the reference's LTCSNAP3 layout (src/snapshot.hpp:1-2,
// src/snapshot.cpp), extended with the FP64 state of the mixed-precision form.
//
// Layout (little-endian). A 256-byte header, then planar arrays, each starting on a 4096-byte
// boundary: f32 x, y, z, vx, vy, vz, m, then u32 id (the reference's eight arrays). Header fields
// at byte offsets: magic "LTCSNAP3" 0, u32 version 3 @8, u32 header size 256 @12, u64 n @16,
// u64 step @24, f64 time @32, G @40, eps @48, dt @56, u64 seed @64, scenario (31 bytes + NUL) @72,
// u64 array offsets[8] @104, u32 pieces @168 (bytes after it are zero in the reference's files).
//
// FP64 extension. u32 ext @172 = 1 marks six more arrays, f64 x, y, z, vx, vy, vz, at the u64
// offsets @176..224, after the eight others and 4096-aligned likewise. The f32 arrays then hold the
// FP64 values rounded to nearest, so a reader of LTCSNAP3 (the reference) reads the file as an
// ordinary snapshot, and this reader restores the FP64 state bit for bit.
//
// Writes go to <path>.tmp, are flushed and fsync'ed, then renamed into place, so `path` is never
// left half-written. Reads validate the magic, version, header size, body count and every array's
// extent against the file size before reading.
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/lettuce_hip.h"

namespace {

constexpr char kMagic[8] = {'L', 'T', 'C', 'S', 'N', 'A', 'P', '3'};
constexpr uint32_t kVersion = 3;
constexpr uint32_t kHeader = 256;
constexpr uint64_t kPage = 4096;
constexpr int kF32Arrays = 8;  // x y z vx vy vz m (f32), id (u32)
constexpr int kF64Arrays = 6;  // x y z vx vy vz (f64)
constexpr uint32_t kExtF64 = 1;
enum : std::size_t {
    kAtN = 16, kAtStep = 24, kAtTime = 32, kAtG = 40, kAtEps = 48, kAtDt = 56, kAtSeed = 64, kAtScenario = 72,
    kAtOffsets = 104, kAtPieces = 168, kAtExt = 172, kAtOffsets64 = 176
};

thread_local std::string g_error;

struct Closer {
    void operator()(std::FILE* f) const {
        if (f) std::fclose(f);
    }
};
using FilePtr = std::unique_ptr<std::FILE, Closer>;

struct SnapError : std::runtime_error {
    SnapError(const std::string& path, const std::string& what) : std::runtime_error("snapshot '" + path + "': " + what) {}
};

uint64_t page_up(uint64_t v) { return (v + kPage - 1) / kPage * kPage; }

template <class T>
void store(unsigned char* h, std::size_t at, T v) {
    std::memcpy(h + at, &v, sizeof v);
}
template <class T>
T load(const unsigned char* h, std::size_t at) {
    T v;
    std::memcpy(&v, h + at, sizeof v);
    return v;
}

struct Parsed {
    lhip_snapshot_header h{};
    uint64_t off[kF32Arrays]{};
    uint64_t off64[kF64Arrays]{};
};

// Opens and validates; every array offset is checked against the file size.
Parsed parse(const std::string& path, FilePtr& f) {
    f.reset(std::fopen(path.c_str(), "rb"));
    if (!f) throw SnapError(path, std::string("cannot open: ") + std::strerror(errno));
    if (std::fseek(f.get(), 0, SEEK_END) != 0) throw SnapError(path, "cannot seek");
    const long end = std::ftell(f.get());
    if (end < 0) throw SnapError(path, "cannot determine the size");
    const uint64_t size = uint64_t(end);
    std::rewind(f.get());
    unsigned char h[kHeader];
    if (size < kHeader || std::fread(h, 1, kHeader, f.get()) != kHeader) throw SnapError(path, "too short for a header");
    if (std::memcmp(h, kMagic, 8) != 0) throw SnapError(path, "bad magic (not an LTCSNAP3 snapshot)");
    if (load<uint32_t>(h, 8) != kVersion) throw SnapError(path, "unsupported version");
    if (load<uint32_t>(h, 12) != kHeader) throw SnapError(path, "unexpected header size");
    Parsed p;
    p.h.n = load<uint64_t>(h, kAtN);
    p.h.step = load<uint64_t>(h, kAtStep);
    p.h.time = load<double>(h, kAtTime);
    p.h.G = load<double>(h, kAtG);
    p.h.eps = load<double>(h, kAtEps);
    p.h.dt = load<double>(h, kAtDt);
    p.h.seed = load<uint64_t>(h, kAtSeed);
    std::memcpy(p.h.scenario, h + kAtScenario, 31);
    p.h.scenario[31] = 0;
    p.h.pieces = load<uint32_t>(h, kAtPieces);
    // The count is bounded by the file size before any product with it can overflow.
    if (p.h.n > size / 4) throw SnapError(path, "body count exceeds the file size");
    std::memcpy(p.off, h + kAtOffsets, sizeof p.off);
    for (int i = 0; i < kF32Arrays; ++i)
        if (p.off[i] < kHeader || p.off[i] > size || size - p.off[i] < p.h.n * 4)
            throw SnapError(path, "array " + std::to_string(i) + " extends past the end of the file");
    const uint32_t ext = load<uint32_t>(h, kAtExt);
    if (ext == kExtF64) {
        if (p.h.n > size / 8) throw SnapError(path, "body count exceeds the file size");
        std::memcpy(p.off64, h + kAtOffsets64, sizeof p.off64);
        for (int i = 0; i < kF64Arrays; ++i)
            if (p.off64[i] < kHeader || p.off64[i] > size || size - p.off64[i] < p.h.n * 8)
                throw SnapError(path, "FP64 array " + std::to_string(i) + " extends past the end of the file");
        p.h.fp64 = 1;
    } else if (ext != 0) {
        throw SnapError(path, "unknown extension " + std::to_string(ext));
    }
    return p;
}

void read_array(const std::string& path, std::FILE* f, uint64_t off, void* dst, std::size_t elem, uint64_t n,
                const char* what) {
    if (n == 0 || !dst) return;
    if (std::fseek(f, long(off), SEEK_SET) != 0 || std::fread(dst, elem, n, f) != n)
        throw SnapError(path, std::string("short read in ") + what);
}

template <class F>
int guard(F&& f) {
    try {
        f();
        return LHIP_OK;
    } catch (const std::bad_alloc&) {
        g_error = "snapshot: out of host memory";
        return LHIP_OUT_OF_MEMORY;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return LHIP_INVALID;
    } catch (const std::exception& e) {
        g_error = e.what();
        return LHIP_RUNTIME;
    }
}

}  // namespace

extern "C" {

const char* lhip_snapshot_error(void) { return g_error.c_str(); }

int lhip_snapshot_read_header(const char* path, lhip_snapshot_header* h) {
    if (!path || !h) return LHIP_INVALID;
    return guard([&] {
        FilePtr f;
        *h = parse(path, f).h;
    });
}

int lhip_snapshot_save(const char* path, const lhip_snapshot_header* h, uint64_t n, const float* x, const float* y,
                       const float* z, const float* vx, const float* vy, const float* vz, const float* m,
                       const uint32_t* id, const double* xd, const double* yd, const double* zd, const double* vxd,
                       const double* vyd, const double* vzd) {
    if (!path || !h) return LHIP_INVALID;
    const void* a32[kF32Arrays] = {x, y, z, vx, vy, vz, m, id};
    const double* a64[kF64Arrays] = {xd, yd, zd, vxd, vyd, vzd};
    const bool f64 = xd != nullptr;
    if (n > 0)
        for (const void* p : a32)
            if (!p) return LHIP_INVALID;
    if (f64 && n > 0)
        for (const double* p : a64)
            if (!p) return LHIP_INVALID;
    return guard([&] {
        const std::string file(path);
        unsigned char hd[kHeader] = {};
        uint64_t off[kF32Arrays], off64[kF64Arrays] = {};
        uint64_t at = kPage;
        for (int i = 0; i < kF32Arrays; ++i) {
            off[i] = at;
            at = page_up(at + n * 4);
        }
        if (f64)
            for (int i = 0; i < kF64Arrays; ++i) {
                off64[i] = at;
                at = page_up(at + n * 8);
            }
        std::memcpy(hd, kMagic, 8);
        store<uint32_t>(hd, 8, kVersion);
        store<uint32_t>(hd, 12, kHeader);
        store<uint64_t>(hd, kAtN, n);  // the arrays are the truth, not h->n
        store<uint64_t>(hd, kAtStep, h->step);
        store<double>(hd, kAtTime, h->time);
        store<double>(hd, kAtG, h->G);
        store<double>(hd, kAtEps, h->eps);
        store<double>(hd, kAtDt, h->dt);
        store<uint64_t>(hd, kAtSeed, h->seed);
        std::memcpy(hd + kAtScenario, h->scenario, strnlen(h->scenario, 31));
        std::memcpy(hd + kAtOffsets, off, sizeof off);
        store<uint32_t>(hd, kAtPieces, h->pieces);
        if (f64) {
            store<uint32_t>(hd, kAtExt, kExtF64);
            std::memcpy(hd + kAtOffsets64, off64, sizeof off64);
        }
        const std::string tmp = file + ".tmp";
        bool ok;
        {
            FilePtr f(std::fopen(tmp.c_str(), "wb"));
            if (!f) throw SnapError(file, "cannot create '" + tmp + "': " + std::strerror(errno));
            static const unsigned char zero[kPage] = {};
            uint64_t pos = 0;
            auto pad_to = [&](uint64_t to) {
                bool good = true;
                while (good && pos < to) {
                    const std::size_t k = std::size_t(std::min<uint64_t>(to - pos, kPage));
                    good = std::fwrite(zero, 1, k, f.get()) == k;
                    pos += k;
                }
                return good;
            };
            ok = std::fwrite(hd, 1, kHeader, f.get()) == kHeader;
            pos = kHeader;
            for (int i = 0; ok && i < kF32Arrays; ++i) {
                ok = pad_to(off[i]) && (n == 0 || std::fwrite(a32[i], 4, n, f.get()) == n);
                pos = off[i] + n * 4;
            }
            for (int i = 0; f64 && ok && i < kF64Arrays; ++i) {
                ok = pad_to(off64[i]) && (n == 0 || std::fwrite(a64[i], 8, n, f.get()) == n);
                pos = off64[i] + n * 8;
            }
            ok = std::fflush(f.get()) == 0 && ok && !std::ferror(f.get());
            ok = ok && ::fsync(fileno(f.get())) == 0;
            ok = std::fclose(f.release()) == 0 && ok;
        }
        if (!ok) {
            std::remove(tmp.c_str());
            throw SnapError(file, "write failed (disk full or I/O error)");
        }
        if (std::rename(tmp.c_str(), file.c_str()) != 0) {
            std::remove(tmp.c_str());
            throw SnapError(file, "cannot rename '" + tmp + "' into place");
        }
    });
}

int lhip_snapshot_load(const char* path, lhip_snapshot_header* h, uint64_t capacity, float* x, float* y, float* z,
                       float* vx, float* vy, float* vz, float* m, uint32_t* id, double* xd, double* yd, double* zd,
                       double* vxd, double* vyd, double* vzd) {
    if (!path || !h) return LHIP_INVALID;
    return guard([&] {
        const std::string file(path);
        FilePtr f;
        const Parsed p = parse(file, f);
        if (p.h.n > capacity) throw std::invalid_argument("snapshot '" + file + "': " + std::to_string(p.h.n) +
                                                          " bodies exceed the capacity of " + std::to_string(capacity));
        void* a32[kF32Arrays] = {x, y, z, vx, vy, vz, m, id};
        double* a64[kF64Arrays] = {xd, yd, zd, vxd, vyd, vzd};
        for (int i = 0; i < kF32Arrays; ++i) read_array(file, f.get(), p.off[i], a32[i], 4, p.h.n, "an f32 array");
        if (p.h.fp64)
            for (int i = 0; i < kF64Arrays; ++i) read_array(file, f.get(), p.off64[i], a64[i], 8, p.h.n, "an f64 array");
        *h = p.h;
    });
}

}  // extern "C"
