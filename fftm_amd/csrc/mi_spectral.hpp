// Diagonal operators on an ffts half spectrum (no reference counterpart as a library
// entry point; the reference applies such operators in caller functors, e.g.
// solve_poisson_3d_functor, versioned_ffts_3d_test_driver.h:128-154, and the
// spectral kernels of examples/turbulence/source/spacetime_kernels.h:66-145).
//
// Wavenumbers follow the reference drivers: k = i <= n/2 ? i : i - n on full axes
// (Nyquist positive), k = i on the halved real axis.  A radial shell filter uses
// S shells of equal width over [0, kmax], kmax = sqrt(sum_a (n_a/2)^2):
//   shell(k) = min(S - 1, floor(S * |k| / kmax)), evaluated exactly (shell_of).
#pragma once

#include "mi_complex.hpp"
#include "mi_pass.hpp"

namespace fftm_mi
{

struct spec_geom
{
    int32_t nd;        // 3 or 4
    int32_t half_axis; // logical index of the halved (real) axis
    int64_t ext[4];    // memory extents, slowest..fastest
    int32_t ax[4];     // logical axis of each memory axis
    int64_t n[4];      // full logical lengths per logical axis
    int64_t total;     // number of complex bins
    int64_t kmax2;     // sum over axes of (n_a / 2)^2
    int32_t nshells;
    uint64_t mg[4];    // division by ext[m] as a multiply and a shift (mg = 0: divide)
    int32_t  sh[4];
};

__device__ __forceinline__ int spec_shell( const spec_geom &g, int64_t i )
{
    int64_t k2 = 0;
    int64_t r  = i;
    for ( int m = g.nd - 1; m >= 0; --m )
    {
        const int64_t q   = div_col( r, g.ext[m], g.mg[m], g.sh[m] );
        const int64_t idx = r - q * g.ext[m];
        r                 = q;
        const int     a   = g.ax[m];
        const int64_t na  = g.n[a];
        const int64_t k   = ( a == g.half_axis || idx <= na / 2 ) ? idx : idx - na;
        k2 += k * k;
    }
    return shell_of( k2, g.nshells, g.kmax2 );
}

// out[i] = H_i * in[i];  H_i = filter[i] (per bin, real) or coeffs[shell(i)]
template <typename T>
__global__ void __launch_bounds__( 256 )
    k_spectral_multiply( const spec_geom g, const cpx<T> *in, cpx<T> *out, const T *filter, const T *coeffs )
{
    for ( int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < g.total; i += (int64_t)gridDim.x * 256 )
    {
        const T h = filter ? filter[i] : coeffs[spec_shell( g, i )];
        out[i]    = scl( in[i], h );
    }
}

// out[i] = H(k_i) in[i] for an analytic operator (Poisson, Gaussian, radial shells) on
// the dense half-spectrum layout described by f
template <typename T>
__global__ void __launch_bounds__( 256 )
    k_spectral_op( const filter_params<T> f, const cpx<T> *in, cpx<T> *out, int64_t total )
{
    for ( int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256 )
    {
        float       kd = 0.f;
        const float k2 = filter_k2( f, i, kd );
        out[i]         = apply_h( in[i], filter_h( f, k2, kd ), f.kind );
    }
}

// partial[b * S + s] = sum over this block's bins in shell s of Re(conj(G_i) X_i), in FP64
template <typename T>
__global__ void __launch_bounds__( 256 )
    k_shell_partials( const spec_geom g, const cpx<T> *G, const cpx<T> *X, double *partial )
{
    extern __shared__ double acc[];
    for ( int s = threadIdx.x; s < g.nshells; s += 256 )
        acc[s] = 0.0;
    __syncthreads();
    for ( int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < g.total; i += (int64_t)gridDim.x * 256 )
    {
        const cpx<T> a = G[i], b = X[i];
        const double v = (double)a.x * (double)b.x + (double)a.y * (double)b.y;
        atomicAdd( &acc[spec_shell( g, i )], v );
    }
    __syncthreads();
    for ( int s = threadIdx.x; s < g.nshells; s += 256 )
        partial[(int64_t)blockIdx.x * g.nshells + s] = acc[s];
}

// grad[s] = sum_b partial[b * S + s], in a fixed order
template <typename T>
__global__ void __launch_bounds__( 256 ) k_shell_reduce( const double *partial, int nblocks, int nshells, T *grad )
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if ( s >= nshells )
        return;
    double sum = 0.0;
    for ( int b = 0; b < nblocks; ++b )
        sum += partial[(int64_t)b * nshells + s];
    grad[s] = (T)sum;
}

} // namespace fftm_mi
