// Basic types of the source-compatible field API (the names the reference's
// apps use: Int, Scalar, FORCEINLINE; src/tensor/types.h).  Everything here
// is usable in host and device code.
#ifndef NSEM_FIELD_TYPES_H
#define NSEM_FIELD_TYPES_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

typedef unsigned int Int;
typedef double Scalar;

#define FORCEINLINE inline
// host and device
#define NSEM_HD __host__ __device__ inline

#endif
