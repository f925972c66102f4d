// Plane-pass kernels (see kw_plane.inc).
#include <hip/hip_runtime.h>

#include "kw_plane.inc"
