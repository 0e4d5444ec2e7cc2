// Test-only stand-in for <hip/hip_runtime.h>: lets g++ compile the device
// headers of 3d_aero_lab_amd/csrc as host code so that a CPU run (one OS thread
// per GPU thread, a barrier per __syncthreads) can check indexing under
// AddressSanitizer. Never used by the product build.
#pragma once
#include <barrier>
#include <cmath>
#include <cstddef>
#include <cstdint>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__

struct emul_dim3 { int x = 0, y = 0, z = 0; };
extern thread_local emul_dim3 threadIdx;
extern thread_local emul_dim3 blockIdx;
inline emul_dim3 emul_grid_dim() { emul_dim3 d; d.x = 1; d.y = 1; d.z = 1; return d; }
#define gridDim emul_grid_dim()
// one OS thread per GPU thread: a wavefront shuffle has no CPU meaning here; the
// kernels that use it are not run by the CPU harness
template <class T> inline T __shfl_down(T v, int, int) { return v; }
// a wave vote used only as a guard around per-lane selections: the lane's own vote gives the same result
inline uint64_t __builtin_amdgcn_ballot_w64(bool b) { return b ? 1u : 0u; }
extern thread_local std::barrier<> *emul_block_barrier;
inline void __syncthreads() { emul_block_barrier->arrive_and_wait(); }
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
inline float __builtin_amdgcn_sqrtf(float x) { return std::sqrt(x); }
inline float __builtin_amdgcn_rsqf(float x) { return 1.0f / std::sqrt(x); }
inline float __builtin_amdgcn_fmed3f(float a, float b, float c) { return std::max(std::min(a, b), std::min(std::max(a, b), c)); }
// wave issue priority: scheduling only, no effect on results
inline void __builtin_amdgcn_s_setprio(int) {}
#include <math.h>
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
