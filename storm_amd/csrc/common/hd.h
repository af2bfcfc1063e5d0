// Host/device qualifier shim for headers that are compiled twice: once by
// hipcc for gfx950 kernels and once by g++ for the CPU replay of the same
// arithmetic.  Nothing else lives here.
#pragma once

#if defined(__HIPCC__)
#define STORM_HD __host__ __device__ __forceinline__
#else
#define STORM_HD inline
#endif
