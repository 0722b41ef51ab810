// Host-side entry points shared between the kernel translation units of libnvq.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nvq_device.h"

namespace nvq {

// Large-batch point sampling through spatial bins and LDS bricks (nvq_binned.hip).
bool sampleBinnedSupported(const GridView& g, uint64_t n);
int sampleBinned(const GridView& g, const float* xyz, uint64_t n, int world, float* nn, float* val, float* grad,
                 int mode, hipStream_t s);

} // namespace nvq
