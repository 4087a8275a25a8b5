// tm_io.h -- snapshot file helpers shared by tm_api.cpp and tm_io.cpp (internal).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace tmio {

// What a snapshot records about one instance besides its fields.
struct SnapInfo {
    int32_t model, precision, nf, L;
    uint32_t seed, instance_id;
    int64_t step;
    int32_t params_index;
};

// Write / read a snapshot: a 64-byte header, then nf*L*L values in the stored
// precision, species-major.  Return 0 on success, -1 on an I/O or format error.
int write_snapshot(const char* path, const SnapInfo& info, const void* data, size_t bytes);
int read_snapshot(const char* path, SnapInfo* info, std::vector<char>* data);

}  // namespace tmio
