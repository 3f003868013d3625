// 4D axis permutation of a memory array: dst = transpose(src, perm), i.e. dst axis k
// is src axis perm[k] (memory axes listed slowest to fastest).
//
// Replaces the ffts stage transposes (source/detail/direct_transpose_4d.h:42-120,
// an element-per-thread functor with permuted, uncoalesced stores) and the
// hipMemcpy3DAsync slab loops of source/detail/runtime_memcpy_4d_slab_transposer.h.
// Two kernels:
//   * the fastest axis is kept  -> contiguous row copies (every ffts 4D stage
//     transpose is of this kind or the identity);
//   * the fastest axis changes -> 32 x 32 LDS tiles between the source's fastest
//     axis and the destination's fastest axis, 128-byte coalesced on both sides
//     (the one-rank native_xzwy reorder [x][y][z][wh] -> [y][z][wh][x]).
#pragma once

#include "mi_complex.hpp"

namespace fftm_mi
{

struct tr_params
{
    const void *src;
    void       *dst;
    int64_t     d[4];  // source dims, slowest..fastest
    int64_t     ds[4]; // destination stride (elements) of each source axis
};

// one row of the source (fastest axis kept) per blockIdx.y chunk
template <typename E>
__global__ void __launch_bounds__( 256 ) k_tr_rows( const tr_params p )
{
    const int64_t row = (int64_t)blockIdx.y + (int64_t)blockIdx.z * gridDim.y;
    const int64_t nrows = p.d[0] * p.d[1] * p.d[2];
    if ( row >= nrows )
        return;
    const int64_t i2 = row % p.d[2];
    const int64_t r1 = row / p.d[2];
    const int64_t i1 = r1 % p.d[1];
    const int64_t i0 = r1 / p.d[1];
    const E      *s  = static_cast<const E *>( p.src ) + row * p.d[3];
    E            *d  = static_cast<E *>( p.dst ) + i0 * p.ds[0] + i1 * p.ds[1] + i2 * p.ds[2];
    for ( int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x; x < p.d[3]; x += (int64_t)gridDim.x * 256 )
        d[x] = s[x];
}

// LDS tile transpose: tile over (source axis 3 = i, source axis A = j); the other two
// source axes (u, v) are flattened into blockIdx.z.
template <typename E>
__global__ void __launch_bounds__( 256 ) k_tr_tile( const tr_params p, int A, int U, int V )
{
    __shared__ E  tile[32][33];
    const int     tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    const int64_t i0 = (int64_t)blockIdx.x * 32;
    const int64_t j0 = (int64_t)blockIdx.y * 32;
    const int64_t uv = blockIdx.z;
    const int64_t iu = uv % p.d[U];
    const int64_t iv = uv / p.d[U];
    int64_t       sst[4];
    sst[3] = 1;
    sst[2] = p.d[3];
    sst[1] = p.d[3] * p.d[2];
    sst[0] = sst[1] * p.d[1];
    const E *s = static_cast<const E *>( p.src ) + iu * sst[U] + iv * sst[V];
    E       *d = static_cast<E *>( p.dst ) + iu * p.ds[U] + iv * p.ds[V];
    for ( int k = ty; k < 32; k += 8 )
    {
        const int64_t j = j0 + k, i = i0 + tx;
        if ( j < p.d[A] && i < p.d[3] )
            tile[k][tx] = s[j * sst[A] + i];
    }
    __syncthreads();
    for ( int k = ty; k < 32; k += 8 )
    {
        const int64_t i = i0 + k, j = j0 + tx;
        if ( j < p.d[A] && i < p.d[3] )
            d[i * p.ds[3] + j * p.ds[A]] = tile[tx][k];
    }
}

} // namespace fftm_mi
