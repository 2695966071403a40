// Workgroup barrier for data that waves exchange through LDS.
//
// __syncthreads() on gfx950 compiles to a workgroup fence on LDS plus s_barrier. The
// compiler's memory model takes LDS operations of all waves to be observed in one global
// order, so it may leave a wave's last ds_write in flight when the wave reaches the barrier.
// Measured on MI355X: with kernels of other streams sharing the CUs, a wave that reads the
// word after the barrier now and then saw the old value (rare wrong Merkle nodes at the
// in-LDS levels of the lane-parallel tree kernel; gone once every wave waits for its own LDS
// operations before the barrier). mam_barrier() makes that wait explicit.
#pragma once

#include <hip/hip_runtime.h>

__device__ __forceinline__ void mam_barrier() {
    // s_waitcnt lgkmcnt(0): vmcnt and expcnt fields at their maxima (no wait on them)
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __syncthreads();
}
