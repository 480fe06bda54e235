// Minimal named-array container ("SWFX") used for golden fixtures.
// Layout, little-endian: magic "SWFX" u32 version=1, then records of
//   u32 name_length, name bytes, u8 dtype (0 f64, 1 i64, 2 u64, 3 u8),
//   u64 element_count, raw elements.
// tests/golden/swfx.py reads the same layout with numpy.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace swfx {

class Writer {
public:
    explicit Writer(const std::string& path) : file_(std::fopen(path.c_str(), "wb")) {
        if (!file_) throw std::runtime_error("cannot open " + path);
        std::fwrite("SWFX", 1, 4, file_);
        const std::uint32_t version = 1;
        std::fwrite(&version, 4, 1, file_);
    }
    ~Writer() {
        if (file_) std::fclose(file_);
    }
    Writer(const Writer&) = delete;
    Writer& operator=(const Writer&) = delete;

    void f64(const std::string& name, const std::vector<double>& v) { put(name, 0, v.data(), v.size(), 8); }
    void i64(const std::string& name, const std::vector<std::int64_t>& v) { put(name, 1, v.data(), v.size(), 8); }
    void u64(const std::string& name, const std::vector<std::uint64_t>& v) { put(name, 2, v.data(), v.size(), 8); }
    void u8(const std::string& name, const std::vector<std::uint8_t>& v) { put(name, 3, v.data(), v.size(), 1); }

private:
    void put(const std::string& name, std::uint8_t dtype, const void* data, std::uint64_t count,
             std::size_t width) {
        const std::uint32_t length = static_cast<std::uint32_t>(name.size());
        std::fwrite(&length, 4, 1, file_);
        std::fwrite(name.data(), 1, name.size(), file_);
        std::fwrite(&dtype, 1, 1, file_);
        std::fwrite(&count, 8, 1, file_);
        if (count) std::fwrite(data, width, count, file_);
    }
    std::FILE* file_;
};

inline std::uint64_t fnv1a(const void* data, std::size_t bytes,
                           std::uint64_t hash = 14695981039346656037ULL) {
    const auto* p = static_cast<const unsigned char*>(data);
    for (std::size_t i = 0; i < bytes; ++i) {
        hash ^= p[i];
        hash *= 1099511628211ULL;
    }
    return hash;
}

}  // namespace swfx
