// Host plane copies of the filter layer (engine.cpp): into pinned staging memory before the H2D
// copy, and out of the download staging into the caller's frames (src/bm3d.zig:295-373, 427-478
// copy the same planes with memcpy).
#pragma once

#include <immintrin.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace vsz_host {

// Plane copies into pinned staging memory (read next by the DMA engine) and out of it into the
// caller's frames (read next by another filter): 32-byte non-temporal stores skip this core's
// cache, so no destination line is read for ownership first. The fence orders the weakly ordered
// stores before whatever the caller does next (an H2D copy of the staging buffer, returning the
// frame).
__attribute__((target("avx2"))) inline void copy_rows_nt(char *dst, ptrdiff_t dst_stride, const char *src,
                                                  ptrdiff_t src_stride, size_t row_bytes, int rows) {
    for (int y = 0; y < rows; ++y) {
        char *d = dst + y * dst_stride;
        const char *s = src + y * src_stride;
        size_t n = row_bytes;
        const size_t head = std::min<size_t>(n, (32 - (reinterpret_cast<uintptr_t>(d) & 31)) & 31);
        std::memcpy(d, s, head);
        d += head;
        s += head;
        n -= head;
        size_t i = 0;
        for (; i + 128 <= n; i += 128) {
            const __m256i a0 = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(s + i));
            const __m256i a1 = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(s + i + 32));
            const __m256i a2 = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(s + i + 64));
            const __m256i a3 = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(s + i + 96));
            _mm256_stream_si256(reinterpret_cast<__m256i *>(d + i), a0);
            _mm256_stream_si256(reinterpret_cast<__m256i *>(d + i + 32), a1);
            _mm256_stream_si256(reinterpret_cast<__m256i *>(d + i + 64), a2);
            _mm256_stream_si256(reinterpret_cast<__m256i *>(d + i + 96), a3);
        }
        for (; i + 32 <= n; i += 32)
            _mm256_stream_si256(reinterpret_cast<__m256i *>(d + i),
                                _mm256_loadu_si256(reinterpret_cast<const __m256i *>(s + i)));
        std::memcpy(d + i, s + i, n - i);
    }
    _mm_sfence();
}

inline void copy_plane(void *dst, ptrdiff_t dst_stride, const void *src, ptrdiff_t src_stride, size_t row_bytes, int rows) {
    static const bool avx2 = __builtin_cpu_supports("avx2");
    if (avx2 && row_bytes >= 256) {
        copy_rows_nt(static_cast<char *>(dst), dst_stride, static_cast<const char *>(src), src_stride, row_bytes, rows);
        return;
    }
    for (int y = 0; y < rows; ++y)
        std::memcpy(static_cast<char *>(dst) + y * dst_stride, static_cast<const char *>(src) + y * src_stride,
                    row_bytes);
}

}  // namespace vsz_host
