// Device fill of the benchmark harness's deterministic random field: the arithmetic of
// splitmix64, unit_random_from_state<T> and fill_random_real_{3d,4d}_functor
// (source/tests/detail/fft_benchmark_common.h:201-257).  Element (x, y, z[, w]) gets
// unit_random<T>(seed ^ gx K0 ^ gy K1 ^ gz K2 [^ gw K3]) with g = start + index, the same
// value whatever the memory arrangement; the kernel walks the destination in memory
// order so that its stores are coalesced for any axis permutation.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace fftm_mi
{

struct fill_params
{
    int32_t  nd;       // 3 or 4 logical axes
    int64_t  ext[4];   // extents in memory order (slowest first)
    int64_t  str[4];   // element strides in memory order
    int32_t  ax[4];    // logical axis of each memory-order axis
    uint64_t start[4]; // logical origin per logical axis (a slab's offset)
    uint64_t seed;
    int64_t  total;    // product of ext
};

__device__ __forceinline__ uint64_t splitmix64( uint64_t x )
{
    x += 0x9E3779B97F4A7C15ull;
    x = ( x ^ ( x >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
    x = ( x ^ ( x >> 27 ) ) * 0x94D049BB133111EBull;
    return x ^ ( x >> 31 );
}

// 2 u - 1 in T's arithmetic, u = (bits >> 11) 2^-53 converted to T first
template <typename T>
__device__ __forceinline__ T unit_random( uint64_t state )
{
    const double unit = (double)( splitmix64( state ) >> 11 ) * ( 1.0 / 9007199254740992.0 );
    const T      u    = (T)unit;
    return (T)( T( 2 ) * u - T( 1 ) );
}

template <typename T>
__global__ void __launch_bounds__( 256 ) k_fill_random( const fill_params p, T *dst )
{
    for ( int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < p.total; i += (int64_t)gridDim.x * 256 )
    {
        int64_t  r     = i;
        int64_t  off   = 0;
        uint64_t state = p.seed;
        for ( int m = p.nd - 1; m >= 0; --m )
        {
            const int64_t c = r % p.ext[m];
            r /= p.ext[m];
            off += c * p.str[m];
            const int a = p.ax[m];
            const uint64_t k = a == 0 ? 0xD6E8FEB86659FD93ull
                               : a == 1 ? 0xA5A3564E27F886A5ull
                               : a == 2 ? 0x9E3779B97F4A7C15ull
                                        : 0x94D049BB133111EBull;
            state ^= ( p.start[a] + (uint64_t)c ) * k;
        }
        dst[off] = unit_random<T>( state );
    }
}

} // namespace fftm_mi
