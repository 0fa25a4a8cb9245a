// Host helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/egg_sync.h"

namespace egg {

// Largest float s with sqrtf(s) <= e: for correctly rounded sqrtf (monotone),
// sqrtf(s) <= e  <=>  s <= sqrt_threshold(e), so membership tests compare
// squared distances without a square root.
inline float sqrt_threshold(float e) {
    float s = e * e;
    if (std::isinf(s)) s = 3.4028234663852886e38f;
    while (s > 0.0f && std::sqrt(s) > e) s = std::nextafter(s, 0.0f);
    for (;;) {
        float nx = std::nextafter(s, INFINITY);
        if (std::isinf(nx) || std::sqrt(nx) > e) break;
        s = nx;
    }
    return s;
}

inline int hip_err(hipError_t e) {
    if (e == hipSuccess) return 0;
    if (e == hipErrorOutOfMemory) return EGG_SYNC_E_ALLOC;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return EGG_SYNC_E_NODEV;
    return EGG_SYNC_E_HIP;
}

}  // namespace egg
