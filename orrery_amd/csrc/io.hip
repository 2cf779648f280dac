// Checkpoint and trajectory files written from (and, for checkpoints, read back
// into) device-resident state.
//
// The byte layouts are the reference's (docs/formats/checkpoint.md,
// docs/formats/trajectory.md; written by sim/checkpoint.cpp:86-131 and
// sim/trajectory.cpp:38-101): unsigned little-endian integers, IEEE 754 FP64
// bit patterns, 64-bit FNV-1a checksums (offset basis 0xcbf29ce484222325, prime
// 0x100000001b3). A checkpoint goes to "<path>.partial" and is renamed over
// the target, so a reader sees the old file or the new one, never half of one.
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"

namespace {

constexpr char kCheckpointMagic[8] = {'O', 'R', 'R', 'E', 'R', 'Y', 'C', 'K'};
constexpr char kTrajectoryMagic[8] = {'O', 'R', 'R', 'E', 'R', 'Y', 'T', 'J'};
constexpr std::uint32_t kVersion = 1;
constexpr std::uint32_t kFlagSingle = 1u << 0;     // never set: state is FP64
constexpr std::uint32_t kFlagVelocities = 1u << 1;  // trajectory frames carry velocities
constexpr std::uint64_t kParticleLimit = 1ull << 40;
constexpr std::size_t kConfigurationLimit = 1u << 20;

// Appends little-endian fields to a byte buffer, hashing as it goes.
class Bytes {
public:
    void raw(const void* p, std::size_t n) {
        const auto* b = static_cast<const unsigned char*>(p);
        for (std::size_t k = 0; k < n; ++k) hash_ = (hash_ ^ b[k]) * 0x100000001b3ull;
        data_.insert(data_.end(), b, b + n);
    }
    void u32(std::uint32_t v) { put_le(v, 4); }
    void u64(std::uint64_t v) { put_le(v, 8); }
    void real(double v) {
        std::uint64_t bits;
        std::memcpy(&bits, &v, 8);
        put_le(bits, 8);
    }
    void reals(const double* v, std::size_t n) {
        for (std::size_t k = 0; k < n; ++k) real(v[k]);
    }
    void checksum() { u64(hash_); }
    const std::vector<unsigned char>& data() const { return data_; }

private:
    void put_le(std::uint64_t v, int width) {
        unsigned char b[8];
        for (int k = 0; k < width; ++k) b[k] = static_cast<unsigned char>(v >> (8 * k));
        raw(b, static_cast<std::size_t>(width));
    }
    std::vector<unsigned char> data_;
    std::uint64_t hash_ = 0xcbf29ce484222325ull;
};

// Reads little-endian fields from a file image, hashing what it consumes.
class Reader {
public:
    explicit Reader(std::vector<unsigned char> data) : data_(std::move(data)) {}
    bool ok() const { return ok_; }
    std::uint64_t hash() const { return hash_; }
    bool take(void* out, std::size_t n) {
        if (!ok_ || data_.size() - pos_ < n) {
            ok_ = false;
            return false;
        }
        std::memcpy(out, data_.data() + pos_, n);
        for (std::size_t k = 0; k < n; ++k) hash_ = (hash_ ^ data_[pos_ + k]) * 0x100000001b3ull;
        pos_ += n;
        return true;
    }
    std::uint64_t le(int width) {
        unsigned char b[8] = {};
        if (!take(b, static_cast<std::size_t>(width))) return 0;
        std::uint64_t v = 0;
        for (int k = 0; k < width; ++k) v |= static_cast<std::uint64_t>(b[k]) << (8 * k);
        return v;
    }
    double real() {
        const std::uint64_t bits = le(8);
        double v;
        std::memcpy(&v, &bits, 8);
        return v;
    }

private:
    std::vector<unsigned char> data_;
    std::size_t pos_ = 0;
    std::uint64_t hash_ = 0xcbf29ce484222325ull;
    bool ok_ = true;
};

[[noreturn]] void refuse(const std::string& path, const std::string& why) {
    orr::fail(ORR_ERR_INVALID_ARGUMENT, path + ": " + why);
}

void write_file(const std::filesystem::path& path, const std::vector<unsigned char>& bytes,
                bool append) {
    std::ofstream f(path, std::ios::binary | (append ? std::ios::app : std::ios::trunc));
    if (!f) orr::fail(ORR_ERR_RUNTIME, path.string() + ": could not be opened for writing");
    f.write(reinterpret_cast<const char*>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
    f.flush();
    if (!f) orr::fail(ORR_ERR_RUNTIME, path.string() + ": could not be written");
}

// The system's arrays on the host: x y z vx vy vz ax ay az m.
std::vector<double> download(orr_system* s, int first, int count) {
    std::vector<double> host(static_cast<std::size_t>(count) * orr::system_count(s));
    orr::system_download(s, first, count, host.data());
    return host;
}

}  // namespace

struct orr_trajectory {
    std::string path;
    std::uint64_t count = 0;
    double timestep = 0.0;
    bool velocities = false;
};

using orr::guarded;
using orr::require;

extern "C" {

int orr_checkpoint_write(orr_system* sys, const char* path, uint64_t step, const char* configuration,
                         uint64_t configuration_length) {
    return guarded([&] {
        require(sys != nullptr && path != nullptr, "null argument");
        require(configuration != nullptr || configuration_length == 0, "configuration is null");
        require(configuration_length <= kConfigurationLimit, "configuration longer than 1 MiB");
        const std::uint64_t n = orr::system_count(sys);
        const std::vector<double> state = download(sys, 0, 10);
        Bytes b;
        b.raw(kCheckpointMagic, 8);
        b.u32(kVersion);
        b.u32(0);
        b.u64(step);
        b.u64(n);
        b.u32(static_cast<std::uint32_t>(configuration_length));
        b.raw(configuration, configuration_length);
        // Masses, then positions, velocities and accelerations by component
        // (checkpoint.cpp:43-61).
        b.reals(state.data() + 9 * n, n);
        for (int k = 0; k < 9; ++k) b.reals(state.data() + static_cast<std::size_t>(k) * n, n);
        b.checksum();
        std::filesystem::path target(path), partial(path);
        partial += ".partial";
        write_file(partial, b.data(), false);
        std::error_code e;
        std::filesystem::rename(partial, target, e);
        if (e) orr::fail(ORR_ERR_RUNTIME, std::string(path) + ": could not be replaced: " + e.message());
    });
}

int orr_checkpoint_read(orr_ctx* ctx, const char* path, orr_system** out, uint64_t* step,
                        char* configuration, uint64_t capacity, uint64_t* configuration_length) {
    return guarded([&] {
        require(ctx != nullptr && path != nullptr && out != nullptr && step != nullptr, "null argument");
        *out = nullptr;
        std::ifstream f(path, std::ios::binary);
        if (!f) orr::fail(ORR_ERR_INVALID_ARGUMENT, std::string(path) + ": could not be opened for reading");
        std::vector<unsigned char> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        Reader r(std::move(data));
        char magic[8] = {};
        if (!r.take(magic, 8) || std::memcmp(magic, kCheckpointMagic, 8) != 0) {
            refuse(path, "is not an Orrery checkpoint");
        }
        const std::uint32_t version = static_cast<std::uint32_t>(r.le(4));
        if (version != kVersion) refuse(path, "is version " + std::to_string(version) + ", not 1");
        const std::uint32_t flags = static_cast<std::uint32_t>(r.le(4));
        if ((flags & ~kFlagSingle) != 0) refuse(path, "sets header flags this build does not know about");
        if ((flags & kFlagSingle) != 0) refuse(path, "was written in single precision; the engine's state is FP64");
        const std::uint64_t at = r.le(8);
        const std::uint64_t n = r.le(8);
        if (!r.ok() || n > kParticleLimit) refuse(path, "has an unreadable header");
        if (n >= (std::uint64_t{1} << 31)) orr::fail(ORR_ERR_LENGTH, "checkpoint of 2^31 or more particles");
        const std::uint32_t length = static_cast<std::uint32_t>(r.le(4));
        if (!r.ok() || length > kConfigurationLimit) refuse(path, "has an unreadable configuration in it");
        std::string text(length, '\0');
        if (!r.take(text.data(), length)) refuse(path, "has an unreadable configuration in it");
        std::vector<double> state(10 * n);
        // File order: masses, then x y z vx vy vz ax ay az.
        for (std::uint64_t i = 0; i < n; ++i) state[9 * n + i] = r.real();
        for (int k = 0; k < 9; ++k) {
            for (std::uint64_t i = 0; i < n; ++i) state[static_cast<std::size_t>(k) * n + i] = r.real();
        }
        const std::uint64_t expected = r.hash();
        const std::uint64_t stored = r.le(8);
        if (!r.ok()) refuse(path, "ends before its state is complete");
        if (stored != expected) refuse(path, "does not match its checksum");
        orr_system* s = nullptr;
        const int status = orr_system_create(ctx, n, &s);
        if (status != ORR_OK) orr::fail(status, orr_last_error());
        try {
            orr::system_upload(s, 0, 10, state.data());
        } catch (...) {
            orr_system_destroy(s);
            throw;
        }
        *step = at;
        if (configuration_length != nullptr) *configuration_length = length;
        if (configuration != nullptr && capacity > 0) {
            const std::size_t k = std::min<std::size_t>(length, capacity - 1);
            std::memcpy(configuration, text.data(), k);
            configuration[k] = '\0';
        }
        *out = s;
    });
}

int orr_trajectory_open(orr_system* sys, const char* path, double timestep, int32_t with_velocities,
                        orr_trajectory** out) {
    return guarded([&] {
        require(sys != nullptr && path != nullptr && out != nullptr, "null argument");
        *out = nullptr;
        auto t = std::make_unique<orr_trajectory>();
        t->path = path;
        t->count = orr::system_count(sys);
        t->timestep = timestep;
        t->velocities = with_velocities != 0;
        const std::vector<double> masses = download(sys, 9, 1);
        Bytes b;
        b.raw(kTrajectoryMagic, 8);
        b.u32(kVersion);
        b.u32(t->velocities ? kFlagVelocities : 0u);
        b.u64(t->count);
        b.real(timestep);
        b.reals(masses.data(), t->count);
        b.checksum();
        write_file(t->path, b.data(), false);
        *out = t.release();
    });
}

int orr_trajectory_append(orr_trajectory* t, orr_system* sys, uint64_t step) {
    return guarded([&] {
        require(t != nullptr && sys != nullptr, "null argument");
        require(orr::system_count(sys) == t->count, "system size differs from the trajectory's");
        const std::uint64_t n = t->count;
        const std::vector<double> state = download(sys, 0, t->velocities ? 6 : 3);
        Bytes b;  // each frame hashes from its own step number (trajectory.cpp:71-101)
        b.u64(step);
        b.real(static_cast<double>(step) * t->timestep);
        b.reals(state.data(), (t->velocities ? 6 : 3) * n);
        b.checksum();
        write_file(t->path, b.data(), true);
    });
}

int orr_trajectory_close(orr_trajectory* t) {
    return guarded([&] { delete t; });
}

}  // extern "C"
