// Shared device helpers for the cfd2d MI355X kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CFD_WAVE 64

// Index of entry (i, j) in a row-major array of row pitch `pitch` (j contiguous).
__host__ __device__ __forceinline__ long ix(int i, int j, int pitch) {
    return (long)i * pitch + j;
}

// Work-groups are dispatched round-robin over the 8 XCDs by their flattened
// index (x fastest), each XCD with its own L2.  xcd_block maps the hardware
// index to a logical one such that the work-groups of one XCD hold
// consecutive logical indices: neighbouring rows then share an L2 instead of
// being fetched again by another XCD.
__device__ __forceinline__ int xcd_block(int hw, int total) {
    const int x = hw & 7;
    int off = 0;
    for (int k = 0; k < x; ++k) off += (total - k + 7) >> 3;
    return off + (hw >> 3);
}

// (column chunk, row) of a work-group of a (chunks x rows) grid, the rows
// handled by one XCD contiguous (xcd_block)
__device__ __forceinline__ int2 xcd_rowgrid() {
    const int gx = gridDim.x;
    const int b = xcd_block(blockIdx.x + blockIdx.y * gx, gx * gridDim.y);
    const int row = b / gx;
    return make_int2(b - row * gx, row);
}

// One-pass reduction over the w-wide block of n (< 2^31) entries starting at
// (i0, j0), spread over the whole grid, U entries per thread at a time:
// load(i, j) reads an entry's values (in-range indices: past the end the last
// entry is read again), then acc(values, ok) folds them in, ok false past the
// end.  The empty asm statement between the phases keeps every load of a
// batch ahead of the first use, so the U entries' loads are in flight
// together instead of one HBM round trip after the other (the compiler turns
// the guarded accumulations into branches and does not move loads across
// them).
template <int U, typename L, typename A>
__device__ __forceinline__ void flat_reduce(int n, int w, int i0, int j0, L load, A acc) {
    using T = decltype(load(0, 0));
    const int stride = gridDim.x * blockDim.x;
    for (int k0 = blockIdx.x * blockDim.x + threadIdx.x; k0 < n; k0 += U * stride) {
        T val[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * stride;
            ok[u] = k < n;
            const int kk = ok[u] ? k : n - 1;
            const int i = kk / w;
            val[u] = load(i0 + i, j0 + (kk - i * w));
        }
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int u = 0; u < U; ++u) acc(val[u], ok[u]);
    }
}

// Bound of the spin-waits on other work-groups (block 0 of a one-pass
// reduction waiting for the others' flags, every block of a persistent kernel
// waiting for block 0's release; see grid_reduce for block 0's own wait).  A
// wait longer than `limit` ticks of the 100 MHz real-time counter gives up,
// stores 1 in *err -- the host reads it whenever it reads the device state
// and reports the error -- and goes on, so the kernel still ends (with
// meaningless values).  limit <= 0 gives up at once without waiting (the
// host's test of this path).  The start time and the limit are uniform
// (scalar registers).
struct SpinGuard {
    unsigned* err;
    long long limit;
};

__device__ __forceinline__ long long spin_now() {
    return (long long)__builtin_amdgcn_s_memrealtime();   // a scalar (uniform) read
}

__device__ __forceinline__ void spin_fail(const SpinGuard& g) {
    __hip_atomic_store(g.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// a wait that is forced to give up at once (limit <= 0)
__device__ __forceinline__ bool spin_forced(const SpinGuard& g) {
    if (g.limit > 0) return false;
    spin_fail(g);
    return true;
}

// Hands the per-block partials of a one-pass grid reduction to block 0 with no
// read-modify-write on a shared address (those serialise across the chip):
// every other block publishes with one release store of its own flag, block 0
// waits for all the flags, clears them for the next launch and returns true
// once all its threads see the partials.  Thread 0 of each block writes its
// partials before the call.  Only block 0 waits, holding one work-group slot,
// so the other blocks can always be scheduled.
__device__ __forceinline__ bool gather_block0(unsigned* flags, const SpinGuard& spin) {
    if (blockIdx.x != 0) {
        if (threadIdx.x == 0)
            __hip_atomic_store(&flags[blockIdx.x], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return false;
    }
    const bool forced = spin_forced(spin);
    const long long t0 = spin_now();
    for (int b = 1 + threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
        while (!forced &&
               __hip_atomic_load(&flags[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
            if (spin_now() - t0 >= spin.limit) {
                spin_fail(spin);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __hip_atomic_store(&flags[b], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __syncthreads();
    return true;
}

// ---------------------------------------------------------------------------
// Wave / block reductions (deterministic: fixed shuffle tree, fixed wave order)
// ---------------------------------------------------------------------------
// the value of lane (lane ^ M): M = 1, 2 by DPP quad permutes, 4, 8, 16 by
// ds_swizzle's bit-mask mode (no LDS access), 32 by gfx950's swap of the
// wave's halves (v_permlane32_swap: the first result holds the lower half in
// both halves, the second the upper half)
template <int M>
__device__ __forceinline__ double lane_xor(double v) {
    const long long b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
    if constexpr (M == 32) {
        const bool low = __lane_id() < 32;   // the hardware lane
        const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        lo = (int)(low ? rl[1] : rl[0]);
        hi = (int)(low ? rh[1] : rh[0]);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    } else {
        if constexpr (M == 1 || M == 2) {
            constexpr int ctrl = M == 1 ? 0xB1 : 0x4E;   // quad_perm [1,0,3,2] / [2,3,0,1]
            lo = __builtin_amdgcn_mov_dpp(lo, ctrl, 0xF, 0xF, true);
            hi = __builtin_amdgcn_mov_dpp(hi, ctrl, 0xF, 0xF, true);
        } else {
            constexpr int pat = 0x1F | (M << 10);   // and 0x1f, or 0, xor M
            lo = __builtin_amdgcn_ds_swizzle(lo, pat);
            hi = __builtin_amdgcn_ds_swizzle(hi, pat);
        }
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
}

// the butterfly over lane distances 32, 16, .., 1: every lane ends with the
// same sum (each step adds two equal-shaped partial sums, so the pairing, not
// the operand order, fixes the result)
__device__ __forceinline__ double wave_sum(double v) {
    v += lane_xor<32>(v);
    v += lane_xor<16>(v);
    v += lane_xor<8>(v);
    v += lane_xor<4>(v);
    v += lane_xor<2>(v);
    v += lane_xor<1>(v);
    return v;
}
// max that propagates NaN (as numpy's max does), so a blown-up field is seen
__device__ __forceinline__ double nanmax(double a, double b) {
    return (a != a) ? a : ((b != b) ? b : fmax(a, b));
}
__device__ __forceinline__ double wave_max(double v) {
    v = nanmax(v, lane_xor<32>(v));
    v = nanmax(v, lane_xor<16>(v));
    v = nanmax(v, lane_xor<8>(v));
    v = nanmax(v, lane_xor<4>(v));
    v = nanmax(v, lane_xor<2>(v));
    v = nanmax(v, lane_xor<1>(v));
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
    if (NV > 2) {
        // thread q adds value q's wave sums (the same order as below) and
        // leaves the total in the first slot: with many values, every thread
        // reading every wave's sums cost more LDS bandwidth than a barrier
        if ((int)threadIdx.x < NV) {
            const int q = threadIdx.x;
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += lds[q * nw + w];
            lds[q * nw] = s;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NV; ++q) val[q] = lds[q * nw];
    } else {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += lds[q * nw + w];
            val[q] = s;
        }
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
    if (NV > 2) {
        if ((int)threadIdx.x < NV) {   // as in block_sum
            const int q = threadIdx.x;
            double s = lds[q * nw];
            for (int w = 1; w < nw; ++w) s = nanmax(s, lds[q * nw + w]);
            lds[q * nw] = s;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NV; ++q) val[q] = lds[q * nw];
    } else {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = lds[q * nw];
            for (int w = 1; w < nw; ++w) s = nanmax(s, lds[q * nw + w]);
            val[q] = s;
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Grid-wide synchronisation for persistent (cooperatively launched) kernels.
//
// No atomic read-modify-write: every block publishes its partial values and an
// epoch tag on its own 128-byte line (`slots`, 16 doubles per block; tag in the
// last word), block 0 polls all tags, adds the partials in block order with a
// fixed tree, and publishes the totals and the tag on a release line (`rel`,
// two lines used alternately).  Every block then reads the same totals, so the
// reduction is deterministic and identical in all blocks.  Tags are
// release-stored and acquire-loaded at agent scope, which makes every store a
// block made before arriving visible to every block after the release (across
// XCDs).  The tags are zeroed by a small kernel before each launch.
//
// A block writes its slot for epoch e+1 only after it has seen the release of
// epoch e, which block 0 publishes only after reading every slot for epoch e.
// The release line of epoch e is reused at epoch e+2, which block 0 can only
// reach after every block has arrived at e+1, i.e. has finished reading e.
// ---------------------------------------------------------------------------
#define CFD_SLOT 16   // doubles per block slot (128 bytes); values in [0, 15), tag in [15]
#define CFD_MAXV 14

struct GridSync {
    double* slots;    // gridDim.x * CFD_SLOT doubles
    double* rel;      // 2 * CFD_SLOT doubles
    SpinGuard spin;   // bound of the waits
};

__device__ __forceinline__ unsigned* slot_tag(double* line) {
    return reinterpret_cast<unsigned*>(line + CFD_SLOT - 1);
}

// Bounds (SpinGuard): every block's wait for block 0's release gives up after
// the limit (storing the error word) and passes; tags compare with >= so a
// stale or advanced tag never blocks.  Block 0's wait for the other blocks'
// tags has no clock: any bound inside that loop makes the persistent kernels
// spill registers (measured: the FP64 Jacobi-PCG 15-28 % slower).  It waits
// only for blocks that are resident -- the cooperative launch guarantees it --
// and whose own waits are bounded, so it ends once they arrive.  A forced
// limit (<= 0) skips every wait.
template <int NV, bool IS_MAX>
__device__ __forceinline__ void grid_reduce(double (&val)[NV], const GridSync& gs, double* lds,
                                            unsigned& epoch) {
    static_assert(NV >= 1 && NV <= CFD_MAXV, "1..14 reduction values");
    if (IS_MAX) block_max<NV>(val, lds);
    else block_sum<NV>(val, lds);
    const bool forced = spin_forced(gs.spin);
    epoch += 1;
    const unsigned e = epoch;
    if (threadIdx.x == 0) {
        double* my = gs.slots + (long)blockIdx.x * CFD_SLOT;
#pragma unroll
        for (int q = 0; q < NV; ++q) my[q] = val[q];
        __hip_atomic_store(slot_tag(my), e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    double* rel = gs.rel + (e & 1) * CFD_SLOT;
    if (blockIdx.x == 0) {
        double acc[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] = 0.0;
        for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
            double* line = gs.slots + (long)b * CFD_SLOT;
            // poll without acquire (an acquire load invalidates the caches on
            // every try); one acquire fence once the tag has been seen
            while (!forced && __hip_atomic_load(slot_tag(line), __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT) < e) {
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                const double x = __hip_atomic_load(&line[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                acc[q] = IS_MAX ? nanmax(acc[q], x) : acc[q] + x;
            }
        }
        if (IS_MAX) block_max<NV>(acc, lds);
        else block_sum<NV>(acc, lds);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < NV; ++q) rel[q] = acc[q];
            __hip_atomic_store(slot_tag(rel), e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (threadIdx.x == 0) {
        const long long t0 = spin_now();
        while (!forced && __hip_atomic_load(slot_tag(rel), __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT) < e) {
            if (spin_now() - t0 >= gs.spin.limit) {
                spin_fail(gs.spin);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
        for (int q = 0; q < NV; ++q)
            lds[q] = __hip_atomic_load(&rel[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) val[q] = lds[q];
    __syncthreads();
}

__device__ __forceinline__ void grid_sync(const GridSync& gs, double* lds, unsigned& epoch) {
    double none[1] = {0.0};
    grid_reduce<1, false>(none, gs, lds, epoch);
}

template <int NV>
__device__ __forceinline__ void grid_sum(double (&val)[NV], const GridSync& gs, double* lds,
                                         unsigned& epoch) {
    grid_reduce<NV, false>(val, gs, lds, epoch);
}

template <int NV>
__device__ __forceinline__ void grid_max(double (&val)[NV], const GridSync& gs, double* lds,
                                         unsigned& epoch) {
    grid_reduce<NV, true>(val, gs, lds, epoch);
}
