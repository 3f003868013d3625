// Plane pass: the real axis A and one complex axis B of a 2D plane in one HBM pass.
//
// When a whole (B x A) plane of the half spectrum fits in LDS (gfx950: 160 KB per
// workgroup), the R2C along A and the C2C along B run on-chip: the plane is read once
// and written once.  Forward:
//   1. load rows b = 2p, 2p+1 of the real plane into Z[p][n] = (x_2p[n], x_2p+1[n]);
//   2. FFT along n of the NB/2 lines of Z (two real rows per complex line);
//   3. separate the pairs into S[b][k], k <= NA/2 (S aliases Z);
//   4. FFT along b of the NA/2+1 columns of S;
//   5. store S.
// Backward runs the mirror image (C2C inverse along b, then the paired C2R along n).
// The 1D transforms are the Stockham stages of the pass kernel, entirely in LDS with
// register staging.  The load/store thread mappings follow whichever plane axis is
// contiguous in HBM, so every layout of ffts (3D z-y planes, 4D slab z-w planes,
// 4D pencil / native xzwy w-x planes) reads and writes 128-byte segments.
#pragma once

#include "mi_pass.hpp"

namespace fftm_mi
{

template <typename T>
struct plane_params
{
    const void   *in;
    void         *out;
    const cpx<T> *twa; // length-NA table
    const cpx<T> *twb; // length-NB table
    int64_t       ra, rb;           // real-array element strides along A and B
    int64_t       ca, cb;           // complex-array element strides along A (half axis) and B
    int64_t       nplanes, d0;      // planes; batch dim 0 extent (dim 1 = rest)
    int64_t       ri0, ri1, ci0, ci1; // real / complex strides of the two batch dims
    int32_t       wmode;            // real-axis weights, as pass_params::wmode
    int32_t       pad_;
    T             scale;
};

template <int NA, int NB, int CS>
struct plane_geom
{
    static constexpr int NAH   = NA / 2 + 1;
    static constexpr int PA    = NAH;                   // pitch of S rows (complex)
    static constexpr int ZSH   = 4;                     // Z rows: one pad element every 16
    static constexpr int PZ    = NA + ( NA >> ZSH );    // pitch of Z rows (complex)
    static constexpr int ZELEM = ( NB / 2 ) * PZ;
    static constexpr int SELEM = NB * PA;
    static constexpr int ELEMS = ZELEM > SELEM ? ZELEM : SELEM;
    static constexpr int TWOFF = ELEMS;                 // twiddle tables (NA, then NB entries)
    static constexpr int TOTAL = ELEMS + NA + NB;
    static constexpr int BYTES = TOTAL * CS;
    static constexpr int NT    = ELEMS * CS > 32768 ? 512 : 256;
};

constexpr int plane_lds_limit() { return 163840; }

// In-LDS Stockham transform of NL lines of length N.  ROWS: the line elements are
// contiguous (element stride 1, line stride LS) and lanes walk along a line; otherwise
// lanes walk across lines (line stride 1, element stride ES).
template <typename T, int N, int NL, bool ROWS, bool INV, int NT, int PSH, int S = 0>
__device__ __forceinline__ void lds_lines( cpx<T> *buf, int es, int ls, const cpx<T> *tw )
{
    // element e of a line sits at e + (e >> PSH) (PSH = 0: unpadded)
    auto at = [&]( int line, int e ) { return line * ls + ( PSH ? e + ( e >> PSH ) : e ) * es; };
    constexpr rplan RP = make_rplan( N );
    if constexpr ( S < RP.n )
    {
        constexpr int R     = RP.r[S];
        constexpr int NS    = rplan_span( RP, S );
        constexpr int NBF   = N / R;
        constexpr int ITEMS = NL * NBF;
        constexpr int PER   = ( ITEMS + NT - 1 ) / NT;
        constexpr int TST   = N / ( NS * R );
        const int     tid   = threadIdx.x;
        cpx<T>        v[PER][R];
#pragma unroll
        for ( int i = 0; i < PER; ++i )
        {
            const int item = tid + i * NT;
            if ( ITEMS % NT == 0 || item < ITEMS )
            {
                const int line = ROWS ? item / NBF : item % NL;
                const int j    = ROWS ? item % NBF : item / NL;
#pragma unroll
                for ( int r = 0; r < R; ++r )
                    v[i][r] = buf[at( line, j + r * NBF )];
                const int k = j % NS;
                if constexpr ( NS > 1 )
                {
#pragma unroll
                    for ( int r = 1; r < R; ++r )
                        v[i][r] = twmul<INV>( v[i][r], tw[r * k * TST] );
                }
                dft_fixed<R, INV>( v[i], tw, N );
            }
        }
        __syncthreads();
#pragma unroll
        for ( int i = 0; i < PER; ++i )
        {
            const int item = tid + i * NT;
            if ( ITEMS % NT == 0 || item < ITEMS )
            {
                const int line = ROWS ? item / NBF : item % NL;
                const int j    = ROWS ? item % NBF : item / NL;
                const int d    = ( j / NS ) * NS * R + ( j % NS );
#pragma unroll
                for ( int m = 0; m < R; ++m )
                    buf[at( line, d + m * NS )] = v[i][m];
            }
        }
        __syncthreads();
        lds_lines<T, N, NL, ROWS, INV, NT, PSH, S + 1>( buf, es, ls, tw );
    }
}

// plane index -> real / complex base offsets
template <typename T>
__device__ __forceinline__ void plane_offsets( const plane_params<T> &p, int64_t plane, int64_t &ro, int64_t &co )
{
    const int64_t i0 = plane % p.d0, i1 = plane / p.d0;
    ro               = i0 * p.ri0 + i1 * p.ri1;
    co               = i0 * p.ci0 + i1 * p.ci1;
}

// FWD: real -> spectrum (R2C along A, C2C forward along B); else the inverse.
// Every HBM phase is unrolled so that a thread issues all of its loads before the
// first LDS write; the twiddle tables are copied to LDS with the first loads.
template <typename T, int NA, int NB, bool FWD>
__global__ void __launch_bounds__( ( plane_geom<NA, NB, (int)sizeof( cpx<T> )>::NT ) )
    k_plane( const plane_params<T> p )
{
    using C                 = cpx<T>;
    using G                 = plane_geom<NA, NB, (int)sizeof( C )>;
    constexpr int NT        = G::NT;
    constexpr int NAH       = G::NAH;
    constexpr int PA        = G::PA;
    constexpr int PZ        = G::PZ;
    constexpr int ZSH       = G::ZSH;
    constexpr int NP        = NB / 2;
    static_assert( NB % 2 == 0, "plane pass pairs the rows of B" );
    static_assert( G::BYTES <= plane_lds_limit(), "plane does not fit in LDS" );
    __shared__ C  lds[G::TOTAL];
    C            *twa   = lds + G::TWOFF;
    C            *twb   = twa + NA;
    const int     tid   = threadIdx.x;
    const int64_t plane = (int64_t)xcd_tile( blockIdx.x, gridDim.x );
    int64_t       ro, co;
    plane_offsets( p, plane, ro, co );
    auto zi = []( int q, int n ) { return q * PZ + n + ( n >> ZSH ); }; // Z[q][n]

    // twiddle tables into LDS (read once per workgroup, from L2)
    for ( int e = tid; e < NA + NB; e += NT )
        twa[e] = e < NA ? p.twa[e] : p.twb[e - NA];

    constexpr int ITZ = NP * NA;  // Z elements (pairs of reals)
    constexpr int PZT = ( ITZ + NT - 1 ) / NT;
    constexpr int ITS = NB * NAH; // S elements (half-spectrum values)
    constexpr int PST = ( ITS + NT - 1 ) / NT;

    if constexpr ( FWD )
    {
        const T *src = static_cast<const T *>( p.in ) + ro;
        // 1. Z[q][n] = (x[2q][n], x[2q+1][n])
        T ua[PZT], ub[PZT];
#pragma unroll
        for ( int i = 0; i < PZT; ++i )
        {
            const int e = tid + i * NT;
            if ( ITZ % NT == 0 || e < ITZ )
            {
                const int     q   = p.ra == 1 ? e / NA : e % NP;
                const int     n   = p.ra == 1 ? e % NA : e / NP;
                const int64_t off = (int64_t)n * p.ra + (int64_t)( 2 * q ) * p.rb;
                ua[i]             = src[off];
                ub[i]             = src[off + p.rb];
            }
        }
#pragma unroll
        for ( int i = 0; i < PZT; ++i )
        {
            const int e = tid + i * NT;
            if ( ITZ % NT == 0 || e < ITZ )
            {
                const int q = p.ra == 1 ? e / NA : e % NP;
                const int n = p.ra == 1 ? e % NA : e / NP;
                lds[zi( q, n )] = mk<T>( ua[i], ub[i] );
            }
        }
        __syncthreads();
        // 2. rows of Z
        lds_lines<T, NA, NP, true, false, NT, ZSH>( lds, 1, PZ, twa );
        // 3. separate pairs: S[2q][k] = A, S[2q+1][k] = B   (S aliases Z)
        constexpr int ITEMS = NP * NAH;
        constexpr int PER   = ( ITEMS + NT - 1 ) / NT;
        C             a[PER], b[PER];
#pragma unroll
        for ( int i = 0; i < PER; ++i )
        {
            const int e = tid + i * NT;
            if ( ITEMS % NT == 0 || e < ITEMS )
            {
                const int q = e / NAH, k = e % NAH;
                const C   zk = lds[zi( q, k )];
                const C   zm = lds[zi( q, k == 0 ? 0 : NA - k )];
                const T   h  = ( p.wmode == 1 && k != 0 && 2 * k != NA ) ? (T)1 : (T)0.5;
                a[i]         = mk<T>( h * ( zk.x + zm.x ), h * ( zk.y - zm.y ) );
                b[i]         = mk<T>( h * ( zk.y + zm.y ), h * ( zm.x - zk.x ) );
            }
        }
        __syncthreads();
#pragma unroll
        for ( int i = 0; i < PER; ++i )
        {
            const int e = tid + i * NT;
            if ( ITEMS % NT == 0 || e < ITEMS )
            {
                const int q = e / NAH, k = e % NAH;
                lds[( 2 * q ) * PA + k]     = a[i];
                lds[( 2 * q + 1 ) * PA + k] = b[i];
            }
        }
        __syncthreads();
        // 4. columns of S
        lds_lines<T, NB, NAH, false, false, NT, 0>( lds, PA, 1, twb );
        // 5. store
        C *dst = static_cast<C *>( p.out ) + co;
#pragma unroll
        for ( int i = 0; i < PST; ++i )
        {
            const int e = tid + i * NT;
            if ( ITS % NT == 0 || e < ITS )
            {
                const int bb = p.ca == 1 ? e / NAH : e % NB;
                const int k  = p.ca == 1 ? e % NAH : e / NB;
                dst[(int64_t)k * p.ca + (int64_t)bb * p.cb] = scl( lds[bb * PA + k], p.scale );
            }
        }
    }
    else
    {
        const C *src = static_cast<const C *>( p.in ) + co;
        // 1. load S
        C u[PST];
#pragma unroll
        for ( int i = 0; i < PST; ++i )
        {
            const int e = tid + i * NT;
            if ( ITS % NT == 0 || e < ITS )
            {
                const int bb = p.ca == 1 ? e / NAH : e % NB;
                const int k  = p.ca == 1 ? e % NAH : e / NB;
                u[i]         = src[(int64_t)k * p.ca + (int64_t)bb * p.cb];
            }
        }
#pragma unroll
        for ( int i = 0; i < PST; ++i )
        {
            const int e = tid + i * NT;
            if ( ITS % NT == 0 || e < ITS )
            {
                const int bb = p.ca == 1 ? e / NAH : e % NB;
                const int k  = p.ca == 1 ? e % NAH : e / NB;
                lds[bb * PA + k] = u[i];
            }
        }
        __syncthreads();
        // 2. columns of S, inverse
        lds_lines<T, NB, NAH, false, true, NT, 0>( lds, PA, 1, twb );
        // 3. Z[q][n] = A~[n] + i B~[n] from rows 2q, 2q+1 (Hermitian extension)
        C z[PZT];
#pragma unroll
        for ( int i = 0; i < PZT; ++i )
        {
            const int e = tid + i * NT;
            if ( ITZ % NT == 0 || e < ITZ )
            {
                const int q = e / NA, n = e % NA;
                const int nn = n <= NA / 2 ? n : NA - n;
                C         av = lds[( 2 * q ) * PA + nn];
                C         bv = lds[( 2 * q + 1 ) * PA + nn];
                if ( n > NA / 2 )
                {
                    av.y = -av.y;
                    bv.y = -bv.y;
                }
                if ( n == 0 || 2 * n == NA )
                {
                    av.y = 0;
                    bv.y = 0;
                }
                else if ( p.wmode == 2 )
                {
                    av = scl( av, (T)0.5 );
                    bv = scl( bv, (T)0.5 );
                }
                z[i] = mk<T>( av.x - bv.y, av.y + bv.x );
            }
        }
        __syncthreads();
#pragma unroll
        for ( int i = 0; i < PZT; ++i )
        {
            const int e = tid + i * NT;
            if ( ITZ % NT == 0 || e < ITZ )
                lds[zi( e / NA, e % NA )] = z[i];
        }
        __syncthreads();
        // 4. rows of Z, inverse
        lds_lines<T, NA, NP, true, true, NT, ZSH>( lds, 1, PZ, twa );
        // 5. store the real rows
        T *dst = static_cast<T *>( p.out ) + ro;
#pragma unroll
        for ( int i = 0; i < PZT; ++i )
        {
            const int e = tid + i * NT;
            if ( ITZ % NT == 0 || e < ITZ )
            {
                const int     q   = p.ra == 1 ? e / NA : e % NP;
                const int     n   = p.ra == 1 ? e % NA : e / NP;
                const C       v   = lds[zi( q, n )];
                const int64_t off = (int64_t)n * p.ra + (int64_t)( 2 * q ) * p.rb;
                dst[off]          = v.x * p.scale;
                dst[off + p.rb]   = v.y * p.scale;
            }
        }
    }
}

} // namespace fftm_mi
