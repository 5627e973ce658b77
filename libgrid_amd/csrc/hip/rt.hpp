// Runtime state shared by the HIP translation units: the library stream and error text.
#pragma once
#include <hip/hip_runtime.h>

namespace lg {
hipStream_t stream();
// Record a HIP error (if any) from the last call / launch; returns 0 or -1.
int check(hipError_t e, const char *what);
int check_launch(const char *what);
// Record an argument error (nothing was launched); returns -1.
int fail(const char *what);
// Look up this device's allocation-free reduction scratch (pointwise.hip); 0 or -1.
int noalloc_setup(int device);
}  // namespace lg
