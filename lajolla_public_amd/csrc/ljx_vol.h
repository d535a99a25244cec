// Volumetric path tracing hooks of the wavefront loop (homogeneous media).
#pragma once

#include "ljx_kernels.h"

#include <vector>

namespace ljx {

constexpr bool kVolpathEnabled = false;

__global__ void k_reset_stats(unsigned long long *stat) {
    if (threadIdx.x < S_NUM) stat[threadIdx.x] = 0;
}

template <class Alloc, class Bufs>
void vol_alloc(bool has_media, size_t capacity, PathBufs &P, Bufs &bufs, Alloc alloc) {
    (void)has_media;
    (void)capacity;
    (void)P;
    (void)bufs;
    (void)alloc;
}

inline void vol_shade_all(const DeviceScene &, PathBufs &, int, int, hipStream_t) {}
inline void vol_shadow(const DeviceScene &, PathBufs &, int, hipStream_t, bool) {}

}  // namespace ljx
