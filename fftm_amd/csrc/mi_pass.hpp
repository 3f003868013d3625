// One FFT "pass": a batch of length-N transforms along one axis of an array in HBM.
//
// A pass reads every element once and writes every element once.  The batch is a
// set of "columns" (one column = one 1D line to transform), enumerated by a
// flattened index col in [0, ncols) decomposed over up to three batch dimensions
// (col = c0 + d0 * (c1 + d1 * c2)) with independent input and output strides, so
// a pass can also permute the batch layout (the 4D stage transposes of ffts are
// folded into the passes this way).
//
// Two thread mappings, chosen by the input geometry:
//   A  (element stride 1: the line is contiguous)       -> lanes walk along the line
//   B  (line strided, adjacent columns contiguous)       -> lanes walk across columns
// so that first-stage loads and last-stage stores are 128-byte coalesced either way.
// When the output geometry wants the other mapping, the last stage is re-staged
// through LDS before the store.
//
// Real axes use the two-for-one trick: real columns 2q and 2q+1 are packed into one
// complex line z = a + i b, transformed once, and separated with Hermitian symmetry
// (R2C: A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = (Z[k] - conj Z[N-k]) / 2i; C2R: the
// inverse, built at load time from the half spectra, ignoring the imaginary parts
// of the DC and Nyquist bins).  This works for every N and for both mappings.
//
// The 1D transform is a Stockham autosort FFT: stage s with radix R and span Ns
// reads x[j + r N/R], twiddles by W_{Ns R}^{r (j mod Ns)}, runs a radix-R DFT and
// writes y[(j / Ns) Ns R + j mod Ns + m Ns].  Stage 0 reads HBM, the last stage
// writes HBM (or LDS when post-processing / re-staging), the stages in between
// exchange through LDS.  Twiddles come from a length-N table e^{-2 pi i k / N}
// generated on the host in FP64 (long double) and rounded once.
#pragma once

#include "mi_complex.hpp"

namespace fftm_mi
{

enum : int
{
    KIND_C2C = 0,
    KIND_R2C = 1,
    KIND_C2R = 2
};

template <typename T>
struct pass_params
{
    const void   *in;
    void         *out;
    const cpx<T> *tw;
    int64_t       ies, oes; // element strides along the transformed axis
    int64_t       ncols;    // columns (real columns for R2C / C2R)
    int64_t       d0, d1;   // batch dims 0 and 1 (dim 2 is the rest)
    int64_t       is0, is1, is2;
    int64_t       os0, os1, os2;
    int32_t       store_b; // 1: store with the column-fastest mapping, 0: line-fastest
    int32_t       wmode;   // real-axis bin weights: 0 none, 1 R2C output x w_k, 2 C2R input / w_k
    T             scale;   // multiplies every output (1 for the plain transforms)
};

// Runtime description for the generic (any-radix) kernel.
struct radix_list
{
    int32_t n;      // number of stages
    int32_t r[20];  // radices (<= 16 each)
};

// ---------------------------------------------------------------------------
// compile-time radix schedule for a given N
// ---------------------------------------------------------------------------
struct rplan
{
    int n;
    int r[24];
};

constexpr rplan make_rplan( int N )
{
    rplan p{ 0, {} };
    int   m = N, k = 0;
    while ( m % 2 == 0 )
    {
        m /= 2;
        ++k;
    }
    if ( k % 3 == 1 && k >= 4 )
    {
        p.r[p.n++] = 16;
        k -= 4;
    }
    else if ( k % 3 == 2 )
    {
        p.r[p.n++] = 4;
        k -= 2;
    }
    else if ( k % 3 == 1 )
    {
        p.r[p.n++] = 2;
        k -= 1;
    }
    while ( k >= 3 )
    {
        p.r[p.n++] = 8;
        k -= 3;
    }
    for ( int f = 3; f <= m; f += 2 )
        while ( m % f == 0 )
        {
            p.r[p.n++] = f;
            m /= f;
        }
    return p;
}

constexpr int rplan_max( const rplan &p )
{
    int r = 1;
    for ( int i = 0; i < p.n; ++i )
        r = p.r[i] > r ? p.r[i] : r;
    return r;
}

constexpr int rplan_span( const rplan &p, int s ) // product of radices before stage s
{
    int v = 1;
    for ( int i = 0; i < s; ++i )
        v *= p.r[i];
    return v;
}

// ---------------------------------------------------------------------------
// Launch geometry (shared by the kernel instantiations and the host launcher)
// ---------------------------------------------------------------------------
struct geom
{
    int tw;  // columns (complex lines) per workgroup
    int tpc; // threads per column
    int pad; // mode A: one pad element every 2^pad elements (0 = no padding)
};

constexpr int lds_limit_bytes() { return 65536; }
constexpr int lds_limit_wide_bytes() { return 163840; }

constexpr int geom_pa( int N, int pad ) { return pad ? N + ( N >> pad ) : N; }

// var selects a tile-shape variant (0 = default); the tuner measures them per pass:
//   1: half the columns per workgroup (smaller LDS tile, more workgroups per CU)
//   2: twice the threads per column (fewer elements per thread, larger workgroup)
//   3: both (same workgroup size, half the tile)
//   4: twice the columns (256-byte row segments; LDS up to 160 KB on gfx950)
//   5: twice the columns and twice the threads per column (1024-thread workgroups)
constexpr int NUM_VARIANTS = 6;

constexpr geom pick_geom( int N, int csize, bool loadb, int var = 0 )
{
    const rplan p  = make_rplan( N );
    const int   rm = rplan_max( p );
    int         lo = N / 32 > 1 ? N / 32 : 1;
    int         hi = N / rm > 1 ? N / rm : 1;
    if ( lo > hi )
        lo = hi;
    geom g{ 1, 1, 3 };
    if ( loadb )
    {
        const int tw0 = 128 / csize; // one 128-byte row segment
        int       tpc = 256 / tw0;
        tpc           = tpc < lo ? lo : ( tpc > hi ? hi : tpc );
        int tw        = 256 / tpc > 0 ? 256 / tpc : 1;
        while ( tw > 1 && N * tw * csize > lds_limit_bytes() )
        {
            tw /= 2;
            if ( tpc * 2 <= hi )
                tpc *= 2;
        }
        g.tw  = tw;
        g.tpc = tpc;
        g.pad = 0;
    }
    else
    {
        int tpc = hi > 256 ? 256 : hi;
        if ( tpc < lo )
            tpc = lo;
        int tw  = 256 / tpc > 0 ? 256 / tpc : 1;
        int pad = 3;
        while ( tw > 1 && tw * geom_pa( N, pad ) * csize > lds_limit_bytes() )
            tw /= 2;
        if ( tw * geom_pa( N, pad ) * csize > lds_limit_bytes() )
            pad = 0;
        g.tw  = tw;
        g.tpc = tpc;
        g.pad = pad;
    }
    if ( ( var == 1 || var == 3 ) && g.tw > 1 )
        g.tw /= 2;
    if ( ( var == 2 || var == 3 ) && g.tpc * 2 <= hi && g.tpc * 2 * g.tw <= 1024 )
        g.tpc *= 2;
    if ( var == 4 || var == 5 )
    {
        const int pa    = loadb ? N : geom_pa( N, g.pad );
        const int tw2   = g.tw * 2;
        const int tpc2  = var == 5 && g.tpc * 2 <= hi ? g.tpc * 2 : g.tpc;
        if ( pa * tw2 * csize <= lds_limit_wide_bytes() && tw2 * tpc2 <= 1024 )
        {
            g.tw  = tw2;
            g.tpc = tpc2;
        }
    }
    return g;
}

// ---------------------------------------------------------------------------
// column address helpers
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void col_offsets( const pass_params<T> &p, int64_t col, int64_t &io, int64_t &oo )
{
    const int64_t c0 = col % p.d0;
    const int64_t r  = col / p.d0;
    const int64_t c1 = r % p.d1;
    const int64_t c2 = r / p.d1;
    io               = c0 * p.is0 + c1 * p.is1 + c2 * p.is2;
    oo               = c0 * p.os0 + c1 * p.os1 + c2 * p.os2;
}

// element n of the (possibly paired) input line, as one complex value
template <int KIND, typename T>
__device__ __forceinline__ cpx<T> load_line( const pass_params<T> &p, int64_t ia, int64_t ib, bool va, bool vb,
                                             int n, int N )
{
    if constexpr ( KIND == KIND_C2C )
    {
        if ( !va )
            return mk<T>( 0, 0 );
        return static_cast<const cpx<T> *>( p.in )[ia + (int64_t)n * p.ies];
    }
    else if constexpr ( KIND == KIND_R2C )
    {
        const T *src = static_cast<const T *>( p.in );
        const T  a   = va ? src[ia + (int64_t)n * p.ies] : (T)0;
        const T  b   = vb ? src[ib + (int64_t)n * p.ies] : (T)0;
        return mk<T>( a, b );
    }
    else
    {
        const cpx<T> *src = static_cast<const cpx<T> *>( p.in );
        const int     h   = N / 2;
        const int     nn  = n <= h ? n : N - n;
        cpx<T>        a   = va ? src[ia + (int64_t)nn * p.ies] : mk<T>( 0, 0 );
        cpx<T>        b   = vb ? src[ib + (int64_t)nn * p.ies] : mk<T>( 0, 0 );
        if ( n > h )
        {
            a.y = -a.y;
            b.y = -b.y;
        }
        if ( n == 0 || 2 * n == N )
        {
            a.y = 0;
            b.y = 0;
        }
        else if ( p.wmode == 2 )
        {
            // adjoint of R2C: interior bins of the half spectrum carry weight 1/2
            a = scl( a, (T)0.5 );
            b = scl( b, (T)0.5 );
        }
        // z = A~ + i B~
        return mk<T>( a.x - b.y, a.y + b.x );
    }
}

// store element n of a transformed line (C2C / C2R; R2C goes through r2c_store)
template <int KIND, typename T>
__device__ __forceinline__ void store_line( const pass_params<T> &p, int64_t oa, int64_t ob, bool va, bool vb, int n,
                                            cpx<T> z )
{
    if constexpr ( KIND == KIND_C2C )
    {
        if ( va )
            static_cast<cpx<T> *>( p.out )[oa + (int64_t)n * p.oes] = scl( z, p.scale );
    }
    else
    {
        T *dst = static_cast<T *>( p.out );
        if ( va )
            dst[oa + (int64_t)n * p.oes] = z.x * p.scale;
        if ( vb )
            dst[ob + (int64_t)n * p.oes] = z.y * p.scale;
    }
}

// ---------------------------------------------------------------------------
// the compile-time-N pass kernel
// ---------------------------------------------------------------------------
template <typename T, int N, int KIND, bool INV, bool LOADB, int VAR = 0>
struct pass_kernel
{
    using C                    = cpx<T>;
    static constexpr rplan RP  = make_rplan( N );
    static constexpr geom  G   = pick_geom( N, (int)sizeof( C ), LOADB, VAR );
    static constexpr int   TW  = G.tw;
    static constexpr int   TPC = G.tpc;
    static constexpr int   NT  = TW * TPC;
    static constexpr int   PA  = geom_pa( N, G.pad );
    static constexpr int   PB  = TW;
    static constexpr int   LDS = LOADB ? N * PB : TW * PA;
    static constexpr int   CPT = KIND == KIND_C2C ? 1 : 2;
    static_assert( LDS * (int)sizeof( C ) <= ( VAR >= 4 ? lds_limit_wide_bytes() : lds_limit_bytes() ),
                   "LDS tile too large" );

    __device__ static __forceinline__ int L( int n, int c )
    {
        if constexpr ( LOADB )
            return n * PB + c;
        else
            return c * PA + ( G.pad ? n + ( n >> G.pad ) : n );
    }

    struct ctx
    {
        const pass_params<T> *p;
        C                    *lds;
        int                   c, t;
        int64_t               tile;
        int64_t               ia, ib, oa, ob;
        bool                  va, vb;
    };

    // ---- final store of the last stage -------------------------------------------
    template <int R, int NS, int NB>
    __device__ static __forceinline__ void final_store( C ( &v )[NB][R], ctx &x )
    {
        const pass_params<T> &p      = *x.p;
        constexpr int         NBF    = N / R;
        const bool            direct = ( KIND != KIND_R2C ) && ( p.store_b == (int)LOADB );
        if ( direct )
        {
#pragma unroll
            for ( int b = 0; b < NB; ++b )
            {
                const int j = x.t + TPC * b;
                if ( NBF % TPC == 0 || j < NBF )
                {
                    const int d = ( j / NS ) * NS * R + ( j % NS );
#pragma unroll
                    for ( int m = 0; m < R; ++m )
                        store_line<KIND>( p, x.oa, x.ob, x.va, x.vb, d + m * NS, v[b][m] );
                }
            }
            return;
        }
        // stage the transformed lines through LDS, then store with the output mapping
#pragma unroll
        for ( int b = 0; b < NB; ++b )
        {
            const int j = x.t + TPC * b;
            if ( NBF % TPC == 0 || j < NBF )
            {
                const int d = ( j / NS ) * NS * R + ( j % NS );
#pragma unroll
                for ( int m = 0; m < R; ++m )
                    x.lds[L( d + m * NS, x.c )] = v[b][m];
            }
        }
        __syncthreads();
        const int tid = threadIdx.x;
        const int c2  = p.store_b ? tid % TW : tid / TPC;
        const int t2  = p.store_b ? tid / TW : tid % TPC;
        const int64_t col = ( x.tile * TW + c2 ) * CPT;
        const bool    va  = col < p.ncols;
        const bool    vb  = CPT == 2 && col + 1 < p.ncols;
        int64_t       ia, oa, ib = 0, ob = 0;
        col_offsets( p, col, ia, oa );
        if ( CPT == 2 )
            col_offsets( p, col + 1, ib, ob );
        if constexpr ( KIND == KIND_R2C )
        {
            // A[k] = (Z[k] + conj Z[N-k]) / 2 ;  B[k] = (Z[k] - conj Z[N-k]) / (2i)
            T *dst = nullptr;
            (void)dst;
            cpx<T> *out = static_cast<cpx<T> *>( p.out );
            for ( int k = t2; k <= N / 2; k += TPC )
            {
                // adjoint of C2R: interior bins of the half spectrum carry weight 2
                const T h  = ( p.wmode == 1 && k != 0 && 2 * k != N ) ? p.scale : (T)0.5 * p.scale;
                const C zk = x.lds[L( k, c2 )];
                const C zm = x.lds[L( k == 0 ? 0 : N - k, c2 )];
                if ( va )
                    out[oa + (int64_t)k * p.oes] = mk<T>( h * ( zk.x + zm.x ), h * ( zk.y - zm.y ) );
                if ( vb )
                    out[ob + (int64_t)k * p.oes] = mk<T>( h * ( zk.y + zm.y ), h * ( zm.x - zk.x ) );
            }
        }
        else
        {
#pragma unroll 4
            for ( int n = t2; n < N; n += TPC )
                store_line<KIND>( p, oa, ob, va, vb, n, x.lds[L( n, c2 )] );
        }
    }

    // ---- stage chain ----------------------------------------------------------------
    template <int S, int NB, int R>
    __device__ static __forceinline__ void chain( C ( &v )[NB][R], ctx &x )
    {
        constexpr int NS = rplan_span( RP, S );
        if constexpr ( S == RP.n - 1 )
        {
            final_store<R, NS, NB>( v, x );
        }
        else
        {
            constexpr int NBF = N / R;
#pragma unroll
            for ( int b = 0; b < NB; ++b )
            {
                const int j = x.t + TPC * b;
                if ( NBF % TPC == 0 || j < NBF )
                {
                    const int d = ( j / NS ) * NS * R + ( j % NS );
#pragma unroll
                    for ( int m = 0; m < R; ++m )
                        x.lds[L( d + m * NS, x.c )] = v[b][m];
                }
            }
            __syncthreads();
            constexpr int R2   = RP.r[S + 1];
            constexpr int NS2  = NS * R;
            constexpr int NBF2 = N / R2;
            constexpr int NB2  = ( NBF2 + TPC - 1 ) / TPC;
            constexpr int TST  = N / ( NS2 * R2 );
            C             w[NB2][R2];
#pragma unroll
            for ( int b = 0; b < NB2; ++b )
            {
                const int j = x.t + TPC * b;
                if ( NBF2 % TPC == 0 || j < NBF2 )
                {
#pragma unroll
                    for ( int r = 0; r < R2; ++r )
                        w[b][r] = x.lds[L( j + r * NBF2, x.c )];
                    const int k = j % NS2;
#pragma unroll
                    for ( int r = 1; r < R2; ++r )
                        w[b][r] = twmul<INV>( w[b][r], x.p->tw[r * k * TST] );
                    dft_fixed<R2, INV>( w[b], x.p->tw, N );
                }
            }
            __syncthreads();
            chain<S + 1, NB2, R2>( w, x );
        }
    }

    __device__ static __forceinline__ void run( const pass_params<T> &p )
    {
        __shared__ C lds[LDS];
        ctx          x;
        x.p          = &p;
        x.lds        = lds;
        const int tid = threadIdx.x;
        x.c          = LOADB ? tid % TW : tid / TPC;
        x.t          = LOADB ? tid / TW : tid % TPC;
        x.tile       = xcd_tile( blockIdx.x, gridDim.x );
        const int64_t col = ( x.tile * TW + x.c ) * CPT;
        x.va         = col < p.ncols;
        x.vb         = CPT == 2 && col + 1 < p.ncols;
        x.ib = x.ob = 0;
        col_offsets( p, col, x.ia, x.oa );
        if ( CPT == 2 )
            col_offsets( p, col + 1, x.ib, x.ob );

        constexpr int R0   = RP.r[0];
        constexpr int NBF0 = N / R0;
        constexpr int NB0  = ( NBF0 + TPC - 1 ) / TPC;
        C             v[NB0][R0];
#pragma unroll
        for ( int b = 0; b < NB0; ++b )
        {
            const int j = x.t + TPC * b;
            if ( NBF0 % TPC == 0 || j < NBF0 )
            {
#pragma unroll
                for ( int r = 0; r < R0; ++r )
                    v[b][r] = load_line<KIND>( p, x.ia, x.ib, x.va, x.vb, j + r * NBF0, N );
                dft_fixed<R0, INV>( v[b], p.tw, N );
            }
        }
        chain<0, NB0, R0>( v, x );
    }
};

template <typename T, int N, int KIND, bool INV, bool LOADB, int VAR>
__global__ void __launch_bounds__( ( pass_kernel<T, N, KIND, INV, LOADB, VAR>::NT ) )
    k_pass( const pass_params<T> p )
{
    pass_kernel<T, N, KIND, INV, LOADB, VAR>::run( p );
}

// ---------------------------------------------------------------------------
// Generic kernel: runtime N and radix list (radices <= 16), ping-pong LDS,
// column-fastest mapping.  Used for lengths without a compiled specialisation.
// ---------------------------------------------------------------------------
constexpr int GEN_THREADS = 256;

template <typename T, int R, bool INV>
__device__ __forceinline__ void gen_stage( const cpx<T> *src, cpx<T> *dst, int N, int Ns, int tw_cols, int c,
                                           int t, int tpc, const cpx<T> *__restrict__ tw )
{
    const int nbf = N / R;
    for ( int j = t; j < nbf; j += tpc )
    {
        cpx<T> v[R];
        for ( int r = 0; r < R; ++r )
            v[r] = src[( j + r * nbf ) * tw_cols + c];
        const int k    = j % Ns;
        const int tstp = N / ( Ns * R );
        for ( int r = 1; r < R; ++r )
            v[r] = twmul<INV>( v[r], tw[r * k * tstp] );
        dft_fixed<R, INV>( v, tw, N );
        const int d = ( j / Ns ) * Ns * R + k;
        for ( int m = 0; m < R; ++m )
            dst[( d + m * Ns ) * tw_cols + c] = v[m];
    }
}

template <typename T, bool INV>
__device__ __forceinline__ void gen_stage_rt( int R, const cpx<T> *src, cpx<T> *dst, int N, int Ns, int tw_cols,
                                              int c, int t, int tpc, const cpx<T> *__restrict__ tw )
{
    switch ( R )
    {
    case 2: gen_stage<T, 2, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 3: gen_stage<T, 3, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 4: gen_stage<T, 4, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 5: gen_stage<T, 5, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 7: gen_stage<T, 7, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 8: gen_stage<T, 8, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 11: gen_stage<T, 11, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 13: gen_stage<T, 13, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 16: gen_stage<T, 16, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    default: break; // host never plans other radices
    }
}

template <typename T, int KIND, bool INV>
__global__ void __launch_bounds__( GEN_THREADS )
    k_pass_generic( const pass_params<T> p, const radix_list rl, int N, int tw_cols )
{
    extern __shared__ unsigned char gen_lds_raw[];
    cpx<T>       *buf0 = reinterpret_cast<cpx<T> *>( gen_lds_raw );
    cpx<T>       *buf1 = buf0 + (int64_t)N * tw_cols;
    constexpr int CPT  = KIND == KIND_C2C ? 1 : 2;
    const int     tid  = threadIdx.x;
    const int     tpc  = GEN_THREADS / tw_cols;
    const int     c    = tid % tw_cols;
    const int     t    = tid / tw_cols;
    const int64_t tile = xcd_tile( blockIdx.x, gridDim.x );
    const int64_t col  = ( tile * tw_cols + c ) * CPT;
    const bool    va   = c < tw_cols && col < p.ncols;
    const bool    vb   = CPT == 2 && c < tw_cols && col + 1 < p.ncols;
    int64_t       ia = 0, oa = 0, ib = 0, ob = 0;
    col_offsets( p, col, ia, oa );
    if ( CPT == 2 )
        col_offsets( p, col + 1, ib, ob );
    if ( t < tpc )
        for ( int n = t; n < N; n += tpc )
            buf0[n * tw_cols + c] = load_line<KIND>( p, ia, ib, va, vb, n, N );
    __syncthreads();
    cpx<T> *src = buf0, *dst = buf1;
    int     Ns  = 1;
    for ( int s = 0; s < rl.n; ++s )
    {
        if ( t < tpc )
            gen_stage_rt<T, INV>( rl.r[s], src, dst, N, Ns, tw_cols, c, t, tpc, p.tw );
        __syncthreads();
        Ns *= rl.r[s];
        cpx<T> *tmp = src;
        src         = dst;
        dst         = tmp;
    }
    if ( t >= tpc )
        return;
    if constexpr ( KIND == KIND_R2C )
    {
        cpx<T> *out = static_cast<cpx<T> *>( p.out );
        for ( int k = t; k <= N / 2; k += tpc )
        {
            const T      h  = ( p.wmode == 1 && k != 0 && 2 * k != N ) ? p.scale : (T)0.5 * p.scale;
            const cpx<T> zk = src[k * tw_cols + c];
            const cpx<T> zm = src[( k == 0 ? 0 : N - k ) * tw_cols + c];
            if ( va )
                out[oa + (int64_t)k * p.oes] = mk<T>( h * ( zk.x + zm.x ), h * ( zk.y - zm.y ) );
            if ( vb )
                out[ob + (int64_t)k * p.oes] = mk<T>( h * ( zk.y + zm.y ), h * ( zm.x - zk.x ) );
        }
    }
    else
    {
        for ( int n = t; n < N; n += tpc )
            store_line<KIND>( p, oa, ob, va, vb, n, src[n * tw_cols + c] );
    }
}

} // namespace fftm_mi
