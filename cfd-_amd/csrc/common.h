// Shared device helpers for the cfd2d MI355X kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CFD_WAVE 64

// Index of entry (i, j) in a row-major array of row pitch `pitch` (j contiguous).
__host__ __device__ __forceinline__ long ix(int i, int j, int pitch) {
    return (long)i * pitch + j;
}

// ---------------------------------------------------------------------------
// Wave / block reductions (deterministic: fixed shuffle tree, fixed wave order)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, CFD_WAVE);
    return v;
}
// max that propagates NaN (as numpy's max does), so a blown-up field is seen
__device__ __forceinline__ double nanmax(double a, double b) {
    return (a != a) ? a : ((b != b) ? b : fmax(a, b));
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, CFD_WAVE));
    return v;
}

// Sums NV values over the block.  `lds` needs NV * (blockDim/64) doubles.
// Every thread returns the block totals in out[].
template <int NV>
__device__ __forceinline__ void block_sum(double (&val)[NV], double* lds) {
    const int lane = threadIdx.x & (CFD_WAVE - 1);
    const int wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = wave_sum(val[q]);
        if (lane == 0) lds[q * nw + wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += lds[q * nw + w];
        val[q] = s;
    }
    __syncthreads();
}

template <int NV>
__device__ __forceinline__ void block_max(double (&val)[NV], double* lds) {
    const int lane = threadIdx.x & (CFD_WAVE - 1);
    const int wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = wave_max(val[q]);
        if (lane == 0) lds[q * nw + wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = lds[q * nw];
        for (int w = 1; w < nw; ++w) s = nanmax(s, lds[q * nw + w]);
        val[q] = s;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Grid-wide barrier for persistent (cooperatively launched) kernels.
// `counter` is zeroed by a separate kernel before the launch; barrier number k
// (1, 2, ...) completes when the counter reaches k * gridDim.x.  Release and
// acquire are at agent scope so stores made before the barrier by any block
// (any XCD) are visible to loads made after it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void grid_barrier(unsigned* counter, unsigned& epoch) {
    __syncthreads();
    epoch += 1;
    if (threadIdx.x == 0) {
        const unsigned target = epoch * gridDim.x;
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        while (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target)
            __builtin_amdgcn_s_sleep(1);
    }
    __syncthreads();
}

// Grid-wide deterministic sum of NV partials: each block stores its partials
// in red[NV*block + q], then after the barrier every block adds all partials
// in block order with one wave, so all blocks obtain bit-identical totals.
// `red` holds two slots of NV*gridDim.x doubles used by barrier parity: a block
// writes slot (e&1) for barrier e only after passing barrier e-1, by which time
// every block has finished reading that slot's previous use (barrier e-2).
template <int NV>
__device__ __forceinline__ void grid_sum(double (&val)[NV], double* red, double* lds,
                                         unsigned* counter, unsigned& epoch) {
    block_sum<NV>(val, lds);
    red += ((epoch + 1) & 1) * (long)NV * gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) red[NV * blockIdx.x + q] = val[q];
    }
    grid_barrier(counter, epoch);
    if (threadIdx.x < CFD_WAVE) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int b = threadIdx.x; b < (int)gridDim.x; b += CFD_WAVE)
                s += __hip_atomic_load(&red[NV * b + q], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            s = wave_sum(s);
            if (threadIdx.x == 0) lds[q] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) val[q] = lds[q];
    __syncthreads();
}

template <int NV>
__device__ __forceinline__ void grid_max(double (&val)[NV], double* red, double* lds,
                                         unsigned* counter, unsigned& epoch) {
    block_max<NV>(val, lds);
    red += ((epoch + 1) & 1) * (long)NV * gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) red[NV * blockIdx.x + q] = val[q];
    }
    grid_barrier(counter, epoch);
    if (threadIdx.x < CFD_WAVE) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int b = threadIdx.x; b < (int)gridDim.x; b += CFD_WAVE)
                s = nanmax(s, __hip_atomic_load(&red[NV * b + q], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT));
            s = wave_max(s);
            if (threadIdx.x == 0) lds[q] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) val[q] = lds[q];
    __syncthreads();
}
