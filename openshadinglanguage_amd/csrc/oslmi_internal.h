// Host-side helpers shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace oslmi {

// Records msg as this thread's last error and returns code.
int set_error(int code, const std::string& msg);
// Converts a pending launch error (hipGetLastError) into an error code.
int check_launch(const char* what);
// Converts a HIP API status into an error code.
int check_hip(hipError_t e, const char* what);

// Grid size for a grid-stride loop over n items with 256-thread blocks:
// enough blocks to fill 256 CUs many times over, capped so the index
// arithmetic stays in range.
inline unsigned grid_for(int64_t n, int per_thread = 1)
{
    int64_t items  = (n + per_thread - 1) / per_thread;
    int64_t blocks = (items + 255) / 256;
    if (blocks < 1)
        blocks = 1;
    if (blocks > (int64_t)1 << 20)
        blocks = (int64_t)1 << 20;
    return (unsigned)blocks;
}

}  // namespace oslmi
