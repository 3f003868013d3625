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
    int32_t       order;            // pass_params::vec order bits (VEC_RR, VEC_CHUNK8, VEC_CHUNK32)
    int32_t       pair;             // k_plane_rx: real rows 2q, 2q+1 move as one aligned pair
    int32_t       pad_;
    T             scale;
};

// PPW planes per workgroup share every barrier (small planes do little work between
// barriers); each plane owns an LDS region of ELEMS values.
template <int NA, int NB, int CS, int PPW = 1>
struct plane_geom
{
    static constexpr int NAH   = NA / 2 + 1;
    static constexpr int PA    = NAH;                   // pitch of S rows (complex)
    static constexpr int ZSH   = 4;                     // Z rows: one pad element every 16
    static constexpr int PZ    = NA + ( NA >> ZSH );    // pitch of Z rows (complex)
    static constexpr int ZELEM = ( NB / 2 ) * PZ;
    static constexpr int SELEM = NB * PA;
    static constexpr int ELEMS = ZELEM > SELEM ? ZELEM : SELEM;
    static constexpr int TWOFF = PPW * ELEMS;           // twiddle tables (NA, then NB entries)
    static constexpr int TOTAL = PPW * ELEMS + NA + NB;
    static constexpr int BYTES = TOTAL * CS;
    // threads: more waves per workgroup for the large planes (the LDS tile caps the
    // workgroups per CU at two, so memory-level parallelism comes from the waves)
    static constexpr int NT    = PPW * ELEMS * CS > 65536 ? 1024 : ( PPW * ELEMS * CS > 32768 ? 512 : 256 );
    // waves per SIMD for two such workgroups per CU when the tile allows two
    static constexpr int WPE   = ( 2 * BYTES <= 163840 && NT == 1024 ) ? 8 : 1;
};

constexpr int plane_lds_limit() { return 163840; }

// In-LDS Stockham transform of NL lines of length N.  ROWS: the line elements are
// contiguous (element stride 1, line stride LS) and lanes walk along a line; otherwise
// lanes walk across lines (line stride 1, element stride ES).
template <typename T, int N, int NL, bool ROWS, bool INV, int NT, int PSH, int LPG, int S = 0>
__device__ __forceinline__ void lds_lines( cpx<T> *buf, int es, int ls, int gs, const cpx<T> *tw )
{
    // element e of a line sits at e + (e >> PSH) (PSH = 0: unpadded); lines come in
    // groups of LPG (one group per plane), groups gs apart
    auto at = [&]( int line, int e ) {
        return ( line / LPG ) * gs + ( line % LPG ) * ls + ( PSH ? e + ( e >> PSH ) : e ) * es;
    };
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
        lds_lines<T, N, NL, ROWS, INV, NT, PSH, LPG, S + 1>( buf, es, ls, gs, tw );
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
template <typename T, int NA, int NB, bool FWD, int PPW>
__global__ void __launch_bounds__( ( plane_geom<NA, NB, (int)sizeof( cpx<T> ), PPW>::NT ) )
    __attribute__( ( amdgpu_waves_per_eu( plane_geom<NA, NB, (int)sizeof( cpx<T> ), PPW>::WPE ) ) )
    k_plane( const plane_params<T> p )
{
    using C                 = cpx<T>;
    using G                 = plane_geom<NA, NB, (int)sizeof( C ), PPW>;
    constexpr int NT        = G::NT;
    constexpr int NAH       = G::NAH;
    constexpr int PA        = G::PA;
    constexpr int PZ        = G::PZ;
    constexpr int ZSH       = G::ZSH;
    constexpr int EL        = G::ELEMS;
    constexpr int NP        = NB / 2;
    static_assert( NB % 2 == 0, "plane pass pairs the rows of B" );
    static_assert( G::BYTES <= plane_lds_limit(), "plane does not fit in LDS" );
    static_assert( G::ZELEM <= EL && G::SELEM <= EL, "plane region" );
    __shared__ C  lds[G::TOTAL];
    C            *twa   = lds + G::TWOFF;
    C            *twb   = twa + NA;
    const int     tid   = threadIdx.x;
    const int64_t first = ordered_tile( p.order ) * PPW;
    int64_t       ro[PPW], co[PPW];
    bool          pv[PPW];
#pragma unroll
    for ( int g = 0; g < PPW; ++g )
    {
        pv[g] = first + g < p.nplanes;
        plane_offsets( p, pv[g] ? first + g : 0, ro[g], co[g] );
    }
    auto zi = []( int g, int q, int n ) { return g * EL + q * PZ + n + ( n >> ZSH ); }; // Z[q][n] of plane g
    // twiddle tables into LDS (read once per workgroup, from L2), after the plane's loads
    // are issued so that their latency overlaps
    auto copy_twiddles = [&]() {
        for ( int e = tid; e < NA + NB; e += NT )
            twa[e] = e < NA ? p.twa[e] : p.twb[e - NA];
    };

    constexpr int ITZ = NP * NA;  // Z elements (pairs of reals) per plane
    constexpr int PZT = ( PPW * ITZ + NT - 1 ) / NT;
    constexpr int ITS = NB * NAH; // S elements (half-spectrum values) per plane
    constexpr int PST = ( PPW * ITS + NT - 1 ) / NT;

    if constexpr ( FWD )
    {
        // 1. Z[q][n] = (x[2q][n], x[2q+1][n])
        T ua[PZT], ub[PZT];
#pragma unroll
        for ( int i = 0; i < PZT; ++i )
        {
            const int e = tid + i * NT, g = e / ITZ, r = e % ITZ;
            if ( ( ( PPW * ITZ ) % NT == 0 || e < PPW * ITZ ) && pv[g] )
            {
                const T      *src = static_cast<const T *>( p.in ) + ro[g];
                const int     q   = p.ra == 1 ? r / NA : r % NP;
                const int     n   = p.ra == 1 ? r % NA : r / NP;
                const int64_t off = (int64_t)n * p.ra + (int64_t)( 2 * q ) * p.rb;
                ua[i]             = src[off];
                ub[i]             = src[off + p.rb];
            }
            else
            {
                ua[i] = ub[i] = (T)0;
            }
        }
        copy_twiddles();
#pragma unroll
        for ( int i = 0; i < PZT; ++i )
        {
            const int e = tid + i * NT, g = e / ITZ, r = e % ITZ;
            if ( ( PPW * ITZ ) % NT == 0 || e < PPW * ITZ )
            {
                const int q = p.ra == 1 ? r / NA : r % NP;
                const int n = p.ra == 1 ? r % NA : r / NP;
                lds[zi( g, q, n )] = mk<T>( ua[i], ub[i] );
            }
        }
        __syncthreads();
        // 2. rows of Z
        lds_lines<T, NA, PPW * NP, true, false, NT, ZSH, NP>( lds, 1, PZ, EL, twa );
        // 3. separate pairs: S[2q][k] = A, S[2q+1][k] = B   (S aliases Z)
        constexpr int ITEMS = NP * NAH;
        constexpr int PER   = ( PPW * ITEMS + NT - 1 ) / NT;
        C             a[PER], b[PER];
#pragma unroll
        for ( int i = 0; i < PER; ++i )
        {
            const int e = tid + i * NT, g = e / ITEMS, r = e % ITEMS;
            if ( ( PPW * ITEMS ) % NT == 0 || e < PPW * ITEMS )
            {
                const int q = r / NAH, k = r % NAH;
                const C   zk = lds[zi( g, q, k )];
                const C   zm = lds[zi( g, q, k == 0 ? 0 : NA - k )];
                const T   h  = ( p.wmode == 1 && k != 0 && 2 * k != NA ) ? (T)1 : (T)0.5;
                a[i]         = mk<T>( h * ( zk.x + zm.x ), h * ( zk.y - zm.y ) );
                b[i]         = mk<T>( h * ( zk.y + zm.y ), h * ( zm.x - zk.x ) );
            }
        }
        __syncthreads();
#pragma unroll
        for ( int i = 0; i < PER; ++i )
        {
            const int e = tid + i * NT, g = e / ITEMS, r = e % ITEMS;
            if ( ( PPW * ITEMS ) % NT == 0 || e < PPW * ITEMS )
            {
                const int q = r / NAH, k = r % NAH;
                lds[g * EL + ( 2 * q ) * PA + k]     = a[i];
                lds[g * EL + ( 2 * q + 1 ) * PA + k] = b[i];
            }
        }
        __syncthreads();
        // 4. columns of S
        lds_lines<T, NB, PPW * NAH, false, false, NT, 0, NAH>( lds, PA, 1, EL, twb );
        // 5. store
#pragma unroll
        for ( int i = 0; i < PST; ++i )
        {
            const int e = tid + i * NT, g = e / ITS, r = e % ITS;
            if ( ( ( PPW * ITS ) % NT == 0 || e < PPW * ITS ) && pv[g] )
            {
                const int bb = p.ca == 1 ? r / NAH : r % NB;
                const int k  = p.ca == 1 ? r % NAH : r / NB;
                C        *dst = static_cast<C *>( p.out ) + co[g];
                dst[(int64_t)k * p.ca + (int64_t)bb * p.cb] = scl( lds[g * EL + bb * PA + k], p.scale );
            }
        }
    }
    else
    {
        // 1. load S
        C u[PST];
#pragma unroll
        for ( int i = 0; i < PST; ++i )
        {
            const int e = tid + i * NT, g = e / ITS, r = e % ITS;
            u[i]        = mk<T>( 0, 0 );
            if ( ( ( PPW * ITS ) % NT == 0 || e < PPW * ITS ) && pv[g] )
            {
                const int bb  = p.ca == 1 ? r / NAH : r % NB;
                const int k   = p.ca == 1 ? r % NAH : r / NB;
                const C  *src = static_cast<const C *>( p.in ) + co[g];
                u[i]          = src[(int64_t)k * p.ca + (int64_t)bb * p.cb];
            }
        }
        copy_twiddles();
#pragma unroll
        for ( int i = 0; i < PST; ++i )
        {
            const int e = tid + i * NT, g = e / ITS, r = e % ITS;
            if ( ( PPW * ITS ) % NT == 0 || e < PPW * ITS )
            {
                const int bb = p.ca == 1 ? r / NAH : r % NB;
                const int k  = p.ca == 1 ? r % NAH : r / NB;
                lds[g * EL + bb * PA + k] = u[i];
            }
        }
        __syncthreads();
        // 2. columns of S, inverse
        lds_lines<T, NB, PPW * NAH, false, true, NT, 0, NAH>( lds, PA, 1, EL, twb );
        // 3. Z[q][n] = A~[n] + i B~[n] from rows 2q, 2q+1 (Hermitian extension)
        C z[PZT];
#pragma unroll
        for ( int i = 0; i < PZT; ++i )
        {
            const int e = tid + i * NT, g = e / ITZ, r = e % ITZ;
            if ( ( PPW * ITZ ) % NT == 0 || e < PPW * ITZ )
            {
                const int q = r / NA, n = r % NA;
                const int nn = n <= NA / 2 ? n : NA - n;
                C         av = lds[g * EL + ( 2 * q ) * PA + nn];
                C         bv = lds[g * EL + ( 2 * q + 1 ) * PA + nn];
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
            const int e = tid + i * NT, g = e / ITZ, r = e % ITZ;
            if ( ( PPW * ITZ ) % NT == 0 || e < PPW * ITZ )
                lds[zi( g, r / NA, r % NA )] = z[i];
        }
        __syncthreads();
        // 4. rows of Z, inverse
        lds_lines<T, NA, PPW * NP, true, true, NT, ZSH, NP>( lds, 1, PZ, EL, twa );
        // 5. store the real rows
#pragma unroll
        for ( int i = 0; i < PZT; ++i )
        {
            const int e = tid + i * NT, g = e / ITZ, r = e % ITZ;
            if ( ( ( PPW * ITZ ) % NT == 0 || e < PPW * ITZ ) && pv[g] )
            {
                const int     q   = p.ra == 1 ? r / NA : r % NP;
                const int     n   = p.ra == 1 ? r % NA : r / NP;
                const C       v   = lds[zi( g, q, n )];
                T            *dst = static_cast<T *>( p.out ) + ro[g];
                const int64_t off = (int64_t)n * p.ra + (int64_t)( 2 * q ) * p.rb;
                dst[off]          = v.x * p.scale;
                dst[off + p.rb]   = v.y * p.scale;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Plane pass, register variant (square planes).  Each 1D transform keeps its
// butterflies in registers: stage 0 takes its inputs straight from HBM (or, for the
// second axis, from the first axis's result in LDS), the last stage hands its outputs
// straight to HBM (or to LDS), and consecutive stages meet in one LDS exchange.  Per
// plane that is about six LDS sweeps instead of k_plane's fourteen.  Consecutive
// threads walk whichever axis is contiguous in HBM (runtime choice per phase).  The
// arithmetic (stages, twiddles, pair separation, Hermitian extension, scaling) is
// k_plane's, so the two kernels give bitwise equal results and the tuner can choose
// either.
// ---------------------------------------------------------------------------
template <int N, int CS>
struct plane_rx_geom
{
    static constexpr int NAH = N / 2 + 1;
    static constexpr int NP  = N / 2;
    // planes per workgroup and threads: small planes share a workgroup.  Fewer planes per
    // workgroup (smaller LDS tiles) keep more workgroups in flight per CU: at 64 points,
    // one plane per workgroup runs the 64^4 plane passes 13-15 % faster than two
    static constexpr int PPW = N <= 16 ? 8 : ( N <= 32 ? 2 : 1 );
    static constexpr int NT  = N >= 128 ? 512 : 256;
    static constexpr int XSH = 3;                    // one pad element every 8
    static constexpr int PX  = N + ( N >> XSH ) + 1; // row lines (A axis): odd pitch
    // column lines (B axis): an even pitch when consecutive threads take the butterflies
    // of one column, an odd one when they take consecutive columns
    static constexpr int PC  = N + ( N >> XSH );
    static constexpr int PCO = PC + 1;
    static constexpr int ROWE  = PPW * NP * PX;
    static constexpr int COLE  = PPW * NAH * PCO;
    static constexpr int ELEMS = ROWE > COLE ? ROWE : COLE;
    static constexpr int TOTAL = ELEMS + 2 * N; // + twiddle tables (A, then B)
    static constexpr int BYTES = TOTAL * CS;
};

// Stockham stages of NL lines of length N: stage 0 reads ld(line, e), the last stage
// calls st(line, e, value), and stages in between exchange through LDS at
// xb[line * xp + e + (e >> XSH)].  lf: consecutive threads take consecutive lines, else
// consecutive butterflies of one line.  st_lds: st writes LDS that the stages may have
// read, so a barrier precedes it.  hook runs once stage 0 has issued its loads.
template <typename T, int N, int NL, bool INV, int NT, int XSH>
struct rx_lines
{
    using C                   = cpx<T>;
    static constexpr rplan RP = make_rplan( N );
    __device__ static __forceinline__ int at( int line, int e, int xp ) { return line * xp + ( XSH ? e + ( e >> XSH ) : e ); }

    template <int S, typename LD, typename ST, typename HK>
    __device__ static __forceinline__ void run( C *xb, int xp, const C *tw, const LD &ld, const ST &st, bool lf,
                                                bool st_lds, const HK &hook )
    {
        constexpr int R     = RP.r[S];
        constexpr int NS    = rplan_span( RP, S );
        constexpr int NBF   = N / R;
        constexpr int ITEMS = NL * NBF;
        constexpr int PER   = ( ITEMS + NT - 1 ) / NT;
        constexpr int TST   = N / ( NS * R );
        const int     tid   = threadIdx.x;
        C             v[PER][R];
#pragma unroll
        for ( int i = 0; i < PER; ++i )
        {
            const int item = tid + i * NT;
            if ( ITEMS % NT == 0 || item < ITEMS )
            {
                const int line = lf ? item % NL : item / NBF;
                const int j    = lf ? item / NL : item % NBF;
#pragma unroll
                for ( int r = 0; r < R; ++r )
                {
                    if constexpr ( S == 0 )
                        v[i][r] = ld( line, j + r * NBF );
                    else
                        v[i][r] = xb[at( line, j + r * NBF, xp )];
                }
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
        if constexpr ( S == 0 )
            hook(); // after stage 0 has issued its loads, before the first barrier
        if constexpr ( S + 1 < RP.n )
        {
            __syncthreads(); // every read of xb (this stage, or the previous phase) is done
#pragma unroll
            for ( int i = 0; i < PER; ++i )
            {
                const int item = tid + i * NT;
                if ( ITEMS % NT == 0 || item < ITEMS )
                {
                    const int line = lf ? item % NL : item / NBF;
                    const int j    = lf ? item / NL : item % NBF;
                    const int d    = ( j / NS ) * NS * R + ( j % NS );
#pragma unroll
                    for ( int m = 0; m < R; ++m )
                        xb[at( line, d + m * NS, xp )] = v[i][m];
                }
            }
            __syncthreads();
            run<S + 1>( xb, xp, tw, ld, st, lf, st_lds, hook );
        }
        else
        {
            if ( st_lds )
                __syncthreads();
#pragma unroll
            for ( int i = 0; i < PER; ++i )
            {
                const int item = tid + i * NT;
                if ( ITEMS % NT == 0 || item < ITEMS )
                {
                    const int line = lf ? item % NL : item / NBF;
                    const int j    = lf ? item / NL : item % NBF;
                    const int d    = ( j / NS ) * NS * R + ( j % NS );
#pragma unroll
                    for ( int m = 0; m < R; ++m )
                        st( line, d + m * NS, v[i][m] );
                }
            }
        }
    }
};

// FWD: real -> spectrum (R2C along A, C2C forward along B); else the inverse.  Square
// planes (NA = NB = N).  STAGED (for planes whose B axis is not contiguous in the real
// array): the real rows go through LDS with consecutive threads along A.  Otherwise the
// threads walk the rows q, and p.pair (set by the host for rb = 1, even real-array
// offsets and an aligned real array) moves rows 2q, 2q+1 as one pair of values.
template <typename T, int N, bool FWD, bool STAGED>
__global__ void __launch_bounds__( ( plane_rx_geom<N, (int)sizeof( cpx<T> )>::NT ) )
    k_plane_rx( const plane_params<T> p )
{
    using C               = cpx<T>;
    using G               = plane_rx_geom<N, (int)sizeof( C )>;
    using V2              = typename pair_of<T>::type;
    constexpr int NT      = G::NT;
    constexpr int PPW     = G::PPW;
    constexpr int NAH     = G::NAH;
    constexpr int NP      = G::NP;
    static_assert( G::BYTES <= plane_lds_limit(), "plane does not fit in LDS" );
    using ROWS = rx_lines<T, N, PPW * NP, !FWD, NT, G::XSH>;
    using COLS = rx_lines<T, N, PPW * NAH, !FWD, NT, G::XSH>;
    constexpr int PX = G::PX;
    // consecutive threads along the axis that is contiguous in HBM: the rows (q) when B
    // is contiguous in the real array; else the real rows go through LDS with coalesced
    // accesses along A (`staged`).  The columns (k) when the half axis is contiguous in
    // the spectrum, else the butterflies of a column (b)
    constexpr bool staged = STAGED;
    const bool     lf_cols = p.ca == 1;
    const int      pc      = lf_cols ? G::PCO : G::PC;
    constexpr int ITZ  = PPW * NP * N; // staged real-row elements (pairs of reals)
    constexpr int PZT  = ( ITZ + NT - 1 ) / NT;
    __shared__ __align__( 16 ) C lds[G::TOTAL]; // 16-byte pair reads of S below
    C            *twa   = lds + G::ELEMS;
    C            *twb   = twa + N;
    const int     tid   = threadIdx.x;
    const int64_t first = ordered_tile( p.order ) * PPW;
    int64_t       ro[PPW], co[PPW];
    bool          pv[PPW];
#pragma unroll
    for ( int g = 0; g < PPW; ++g )
    {
        pv[g] = first + g < p.nplanes;
        plane_offsets( p, pv[g] ? first + g : 0, ro[g], co[g] );
    }
    // twiddle tables into LDS once the first HBM loads are issued; stage 0 of every
    // transform (radix 2, 4, 8 or 16) needs none, and later stages read them after an
    // exchange barrier
    auto copy_twiddles = [&]() {
        for ( int e = tid; e < 2 * N; e += NT )
            twa[e] = e < N ? p.twa[e] : p.twb[e - N];
    };
    auto no_hook = []() {};
    if constexpr ( FWD )
    {
        // rows: Z[q][n] = (x[2q][n], x[2q+1][n]) straight from HBM, Z to LDS in natural order
        if constexpr ( staged )
        {
            // Z[q][n] into the row layout, consecutive threads along n; loads in chunks
            // of 8 per thread, all issued before the chunk's LDS writes
            constexpr int CH = 8;
#pragma unroll
            for ( int i0 = 0; i0 < PZT; i0 += CH )
            {
                T ua[CH], ub[CH];
#pragma unroll
                for ( int c = 0; c < CH; ++c )
                {
                    // unconditional loads (an element past the end reads the last one, an
                    // empty plane slot reads plane 0) so that none waits on a branch
                    const int i = i0 + c < PZT ? i0 + c : PZT - 1, e0 = tid + i * NT;
                    const int e = ( ITZ % NT == 0 || e0 < ITZ ) ? e0 : ITZ - 1;
                    const int g = e / ( NP * N ), r = e % ( NP * N );
                    const T  *src = static_cast<const T *>( p.in ) + ro[g] + (int64_t)( r % N ) * p.ra +
                                   (int64_t)( 2 * ( r / N ) ) * p.rb;
                    const T   a   = src[0];
                    const T   b   = src[p.rb];
                    ua[c]         = pv[g] ? a : (T)0;
                    ub[c]         = pv[g] ? b : (T)0;
                }
                if ( i0 == 0 )
                    copy_twiddles();
#pragma unroll
                for ( int c = 0; c < CH; ++c )
                {
                    const int i = i0 + c, e = tid + i * NT, g = e / ( NP * N ), r = e % ( NP * N );
                    if ( i < PZT && ( ITZ % NT == 0 || e < ITZ ) )
                        lds[ROWS::at( g * NP + r / N, r % N, PX )] = mk<T>( ua[c], ub[c] );
                }
            }
            __syncthreads();
        }
        auto ld_row = [&]( int line, int n ) -> C {
            if constexpr ( staged )
                return lds[ROWS::at( line, n, PX )];
            const int g = line / NP, q = line % NP;
            if ( !pv[g] )
                return mk<T>( 0, 0 );
            const T *src = static_cast<const T *>( p.in ) + ro[g] + (int64_t)n * p.ra + (int64_t)( 2 * q ) * p.rb;
            if ( p.pair )
            {
                const V2 w = *reinterpret_cast<const V2 *>( src );
                return mk<T>( w.x, w.y );
            }
            return mk<T>( src[0], src[p.rb] );
        };
        auto st_row = [&]( int line, int n, const C &v ) { lds[ROWS::at( line, n, PX )] = v; };
        if constexpr ( staged )
            ROWS::template run<0>( lds, PX, twa, ld_row, st_row, true, true, no_hook );
        else
            ROWS::template run<0>( lds, PX, twa, ld_row, st_row, true, true, copy_twiddles );
        __syncthreads();
        // columns: S[b][k] separated from Z rows b / 2 (k and its mirror) on the way in
        auto ld_col = [&]( int col, int b ) -> C {
            const int g = col / NAH, k = col % NAH, line = g * NP + b / 2;
            const C   zk = lds[ROWS::at( line, k, PX )];
            const C   zm = lds[ROWS::at( line, k == 0 ? 0 : N - k, PX )];
            const T   h  = ( p.wmode == 1 && k != 0 && 2 * k != N ) ? (T)1 : (T)0.5;
            return ( b & 1 ) ? mk<T>( h * ( zk.y + zm.y ), h * ( zm.x - zk.x ) )
                             : mk<T>( h * ( zk.x + zm.x ), h * ( zk.y - zm.y ) );
        };
        auto st_col = [&]( int col, int b, const C &v ) {
            const int g = col / NAH, k = col % NAH;
            if ( pv[g] )
                static_cast<C *>( p.out )[co[g] + (int64_t)k * p.ca + (int64_t)b * p.cb] = scl( v, p.scale );
        };
        COLS::template run<0>( lds, pc, twb, ld_col, st_col, lf_cols, false, no_hook );
    }
    else
    {
        // columns, inverse: S straight from HBM, S to LDS in natural order
        auto ld_col = [&]( int col, int b ) -> C {
            const int g = col / NAH, k = col % NAH;
            if ( !pv[g] )
                return mk<T>( 0, 0 );
            return static_cast<const C *>( p.in )[co[g] + (int64_t)k * p.ca + (int64_t)b * p.cb];
        };
        // S in natural order.  Where threads walk the butterflies of a column (pc even) and
        // the rows take consecutive q, S is stored unpadded (pitch pc): the pair S[2q], S[2q+1]
        // of a column is then one aligned 16-byte LDS read (FP32), and 16 threads reading
        // consecutive pairs cover all banks once
        const bool pair_s = !lf_cols && !staged && sizeof( T ) == 4;
        auto       s_at   = [&]( int col, int b ) { return pair_s ? col * pc + b : COLS::at( col, b, pc ); };
        auto       st_col = [&]( int col, int b, const C &v ) { lds[s_at( col, b )] = v; };
        COLS::template run<0>( lds, pc, twb, ld_col, st_col, lf_cols, true, copy_twiddles );
        __syncthreads();
        // rows, inverse: Z[q][n] = A~[n] + i B~[n] (Hermitian extension) on the way in
        auto ld_row = [&]( int line, int n ) -> C {
            const int g = line / NP, q = line % NP;
            const int nn = n <= N / 2 ? n : N - n;
            C         av, bv;
            if ( pair_s )
            {
                const f32x4 w = *reinterpret_cast<const f32x4 *>( &lds[s_at( g * NAH + nn, 2 * q )] );
                av            = mk<T>( (T)w.x, (T)w.y );
                bv            = mk<T>( (T)w.z, (T)w.w );
            }
            else
            {
                av = lds[s_at( g * NAH + nn, 2 * q )];
                bv = lds[s_at( g * NAH + nn, 2 * q + 1 )];
            }
            if ( n > N / 2 )
            {
                av.y = -av.y;
                bv.y = -bv.y;
            }
            if ( n == 0 || 2 * n == N )
            {
                av.y = 0;
                bv.y = 0;
            }
            else if ( p.wmode == 2 )
            {
                av = scl( av, (T)0.5 );
                bv = scl( bv, (T)0.5 );
            }
            return mk<T>( av.x - bv.y, av.y + bv.x );
        };
        auto st_row = [&]( int line, int n, const C &v ) {
            if constexpr ( staged )
            {
                lds[ROWS::at( line, n, PX )] = v;
                return;
            }
            const int g = line / NP, q = line % NP;
            if ( !pv[g] )
                return;
            T *dst = static_cast<T *>( p.out ) + ro[g] + (int64_t)n * p.ra + (int64_t)( 2 * q ) * p.rb;
            if ( p.pair )
                *reinterpret_cast<V2 *>( dst ) = V2{ v.x * p.scale, v.y * p.scale };
            else
            {
                dst[0]    = v.x * p.scale;
                dst[p.rb] = v.y * p.scale;
            }
        };
        ROWS::template run<0>( lds, PX, twa, ld_row, st_row, true, staged, no_hook );
        if constexpr ( staged )
        {
            // the real rows out of the row layout, consecutive threads along n
            __syncthreads();
#pragma unroll
            for ( int i = 0; i < PZT; ++i )
            {
                const int e = tid + i * NT, g = e / ( NP * N ), r = e % ( NP * N );
                if ( ( ITZ % NT == 0 || e < ITZ ) && pv[g] )
                {
                    const C v   = lds[ROWS::at( g * NP + r / N, r % N, PX )];
                    T      *dst = static_cast<T *>( p.out ) + ro[g] + (int64_t)( r % N ) * p.ra +
                             (int64_t)( 2 * ( r / N ) ) * p.rb;
                    dst[0]    = v.x * p.scale;
                    dst[p.rb] = v.y * p.scale;
                }
            }
        }
    }
}

} // namespace fftm_mi
