// common.hpp — shared host-side definitions for the native library.
#pragma once

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

namespace hspz {

constexpr uint32_t INVALID_KMER = 1u << 31;  // seed.rs:25-27

struct Shape {
    uint32_t size = 0;       // span of the pattern
    uint32_t kmer_size = 0;  // care positions
    uint32_t pos[32] = {0};
};

int parse_shape(const char *spec, Shape *s);

// io.cpp: the scoring alphabet of every byte (sequence.rs:66-79) and the
// IUPAC complement of every byte (sequence.rs:80-86), 256 entries each.
const uint8_t *encoding_table();
const uint8_t *complement_table();

// Thread-local last-error text behind hspz_last_error().
inline char *error_buffer() {
    static thread_local char buf[512];
    return buf;
}

inline void clear_error() { error_buffer()[0] = 0; }

inline int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(error_buffer(), 512, fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace hspz
