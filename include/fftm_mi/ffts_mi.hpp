// C++ ffts frontend on the MI355X backend: the interface of fftm::ffts
// (source/ffts.hpp:159-1167) -- init once, forward / backward unnormalised -- over the
// fused C-ABI transforms.  SCFD's tensor arrays are replaced by device_array<E, Dim>,
// an owning device buffer with the raw_ptr() / size_nd() / init(...) calls ffts uses.
//
// Memory layouts (slowest..fastest), identical to the reference's arrangers
// (ffts.hpp:104-155):
//   3D            real [x][y][z]       spectrum [x][y][z/2+1]
//   4D pencil     real [w][z][y][x]    spectrum [y][z][w/2+1][x]
//   4D slab       real [x][y][z][w]    spectrum [x][y][z][w/2+1]
//   4D slab+xzwy  real [x][y][z][w]    spectrum [y][z][w/2+1][x]   (fftm native_xzwy)
// 2D uses the named plans of ffts.hpp:659-675 through fft_wrap_many.
#ifndef FFTM_MI_FFTS_HPP
#define FFTM_MI_FFTS_HPP

#include <array>
#include <cstddef>
#include <stdexcept>
#include <string>

#include "mi_fft_wrap_many.h"

namespace fftm
{
namespace mi
{

using ::fftm::direction;

template <class Strategy>
struct strategy_traits;

struct strategy_4d_pencil_pencil
{
};
struct strategy_4d_slab_slab
{
};
struct strategy_4d_slab_native_xzwy
{
};

template <>
struct strategy_traits<strategy_4d_pencil_pencil>
{
    static constexpr int layout = FFTM_MI_LAYOUT_4D_PENCIL;
    static const char   *name() { return "pencil-pencil-fused"; }
};
template <>
struct strategy_traits<strategy_4d_slab_slab>
{
    static constexpr int layout = FFTM_MI_LAYOUT_4D_SLAB;
    static const char   *name() { return "slab-slab-fused"; }
};
template <>
struct strategy_traits<strategy_4d_slab_native_xzwy>
{
    static constexpr int layout = FFTM_MI_LAYOUT_4D_XZWY;
    static const char   *name() { return "slab-native-xzwy-fused"; }
};

// Owning device array; dimensions listed slowest..fastest (memory order).
template <class E, std::size_t Dim>
class device_array
{
public:
    template <class... S>
    void init( S... sizes )
    {
        static_assert( sizeof...( S ) == Dim, "device_array::init: wrong number of sizes" );
        const std::size_t s[Dim] = { static_cast<std::size_t>( sizes )... };
        std::size_t       n      = 1;
        for ( std::size_t k = 0; k < Dim; ++k )
        {
            dims_[k] = s[k];
            n *= s[k];
        }
        buf_.init( n );
    }
    E                            *raw_ptr() const { return buf_.raw_ptr(); }
    std::size_t                   size() const { return buf_.size(); }
    std::size_t                   total_size() const { return buf_.size(); }
    std::array<std::size_t, Dim> size_nd() const { return dims_; }

private:
    ::fftm::wrap::mi::device_buffer<E> buf_;
    std::array<std::size_t, Dim>       dims_{};
};

struct ffts_init_options
{
    bool sync                 = true;  // device-synchronise each call (reference safe-call behaviour)
    bool destructive_backward = false; // 4D/3D backward may use its input as scratch (no work area)
    bool fp16_spectrum_storage = false; // 3D float: FP16 half spectra between passes (FFTM_MI_F32_H16)
};

// Diagonal spectral operators of ffts::filter / ffts::spectral_apply (fftm_mi_spectral_op).
template <class T>
struct spectral_op
{
    int      kind    = FFTM_MI_OP_POISSON; // FFTM_MI_OP_POISSON / _GAUSSIAN / _SHELL
    double   scale   = 1.0;                // e.g. 1/N to normalise a round trip
    double   sigma   = 0.0;                // Gaussian width (wavenumber units)
    const T *coeffs  = nullptr;            // shell coefficients (device)
    int      nshells = 0;
    int      axis    = -1;                 // derivative axis (logical)
    double   klo = 0.0, khi = 0.0;         // band-pass edges

    fftm_mi_spectral_op c() const
    {
        fftm_mi_spectral_op o;
        o.kind     = kind;
        o.nshells  = nshells;
        o.scale    = scale;
        o.sigma    = sigma;
        o.coeffs   = coeffs;
        o.axis     = axis;
        o.reserved = 0;
        o.klo      = klo;
        o.khi      = khi;
        return o;
    }
};

template <class T, class Strategy4D = strategy_4d_pencil_pencil>
class ffts
{
public:
    using real    = T;
    using complex = ::fftm::wrap::mi::complex_value<T>;
    template <std::size_t Dim>
    using real_array_t = device_array<T, Dim>;
    template <std::size_t Dim>
    using complex_array_t = device_array<complex, Dim>;

    static const char *strategy_name() { return strategy_traits<Strategy4D>::name(); }

    ffts() = default;
    ffts( const ffts & )            = delete;
    ffts &operator=( const ffts & ) = delete;
    ~ffts()
    {
        // the handle keeps the work-area pointer: destroy it before the buffer
        if ( h_ != nullptr )
            fftm_mi_ffts_destroy( h_ );
    }

    void init( std::size_t nx, std::size_t ny, const ffts_init_options &o = ffts_init_options() )
    {
        ensure_can_init_();
        opt_ = o;
        nx_  = nx;
        ny_  = ny;
        dim_ = 2;
        const long long nyh = static_cast<long long>( ny / 2 + 1 );
        const long long X = (long long)nx, Y = (long long)ny;
        base_.template add_plan_2D<direction::R2C>( "forward_2d", X, Y, X, Y, 1, X * Y, X, nyh, 1, X * nyh, 1 );
        base_.template add_plan_2D<direction::C2R>( "inverse_2d", X, Y, X, nyh, 1, X * nyh, X, Y, 1, X * Y, 1 );
        work_.init( ( base_.activate_work_size() + 255 ) / 256 * 256 );
        base_.set_external_work_area( work_.raw_ptr() );
        if ( !o.sync )
            base_.set_hot_exec_no_sync_enabled( true );
        init_done_ = true;
    }

    void init( std::size_t nx, std::size_t ny, std::size_t nz, const ffts_init_options &o = ffts_init_options() )
    {
        ensure_can_init_();
        opt_ = o;
        nx_  = nx;
        ny_  = ny;
        nz_  = nz;
        dim_ = 3;
        const int64_t n[3] = { (int64_t)nx, (int64_t)ny, (int64_t)nz };
        create_( FFTM_MI_LAYOUT_3D, n );
        init_done_ = true;
    }

    void init( std::size_t nx, std::size_t ny, std::size_t nz, std::size_t nw,
               const ffts_init_options &o = ffts_init_options() )
    {
        ensure_can_init_();
        opt_ = o;
        nx_  = nx;
        ny_  = ny;
        nz_  = nz;
        nw_  = nw;
        dim_ = 4;
        const int64_t n[4] = { (int64_t)nx, (int64_t)ny, (int64_t)nz, (int64_t)nw };
        create_( strategy_traits<Strategy4D>::layout, n );
        init_done_ = true;
    }

    template <std::size_t Dim>
    void init( const std::array<std::size_t, Dim> &g, const ffts_init_options &o = ffts_init_options() )
    {
        init_array_( g, o );
    }

    bool        is_initialized() const { return init_done_; }
    std::size_t dimension() const
    {
        ensure_initialized_();
        return dim_;
    }

    void set_stream( hipStream_t s )
    {
        if ( h_ != nullptr )
            FFTM_MI_SAFE_CALL( fftm_mi_ffts_set_stream( h_, static_cast<void *>( s ) ) );
        if ( dim_ == 2 )
        {
            base_.set_stream( "forward_2d", s );
            base_.set_stream( "inverse_2d", s );
        }
    }

    void forward( const real_array_t<2> &in, complex_array_t<2> &out )
    {
        ensure_dimension_( 2 );
        base_.exec( "forward_2d", in, out );
    }
    void backward( const complex_array_t<2> &in, real_array_t<2> &out )
    {
        ensure_dimension_( 2 );
        base_.exec( "inverse_2d", in, out );
    }
    void forward( const real_array_t<3> &in, complex_array_t<3> &out )
    {
        ensure_dimension_( 3 );
        run_( true, in.raw_ptr(), out.raw_ptr() );
    }
    void backward( const complex_array_t<3> &in, real_array_t<3> &out )
    {
        ensure_dimension_( 3 );
        run_( false, in.raw_ptr(), out.raw_ptr() );
    }
    void forward( const real_array_t<4> &in, complex_array_t<4> &out )
    {
        ensure_dimension_( 4 );
        run_( true, in.raw_ptr(), out.raw_ptr() );
    }
    void backward( const complex_array_t<4> &in, real_array_t<4> &out )
    {
        ensure_dimension_( 4 );
        run_( false, in.raw_ptr(), out.raw_ptr() );
    }

    // real -> real backward(H forward(in)) with the spectrum kept in the work area; the last
    // forward axis, the multiply and the first backward axis run as one pass
    template <std::size_t Dim>
    void filter( const real_array_t<Dim> &in, real_array_t<Dim> &out, const spectral_op<T> &op )
    {
        ensure_dimension_( Dim );
        if ( h_ == nullptr )
            throw std::logic_error( "ffts::filter: 3D/4D transforms only" );
        const fftm_mi_spectral_op o = op.c();
        ::fftm::wrap::mi::safe_call( fftm_mi_ffts_filter( h_, in.raw_ptr(), out.raw_ptr(), &o, opt_.sync ? 1 : 0 ),
                                     "fftm_mi_ffts_filter", __FILE__, __LINE__, opt_.sync );
    }
    template <std::size_t Dim>
    void spectral_apply( const complex_array_t<Dim> &in, complex_array_t<Dim> &out, const spectral_op<T> &op )
    {
        ensure_dimension_( Dim );
        if ( h_ == nullptr )
            throw std::logic_error( "ffts::spectral_apply: 3D/4D transforms only" );
        const fftm_mi_spectral_op o = op.c();
        ::fftm::wrap::mi::safe_call( fftm_mi_ffts_spectral_apply( h_, in.raw_ptr(), out.raw_ptr(), &o, opt_.sync ? 1 : 0 ),
                                     "fftm_mi_ffts_spectral_apply", __FILE__, __LINE__, opt_.sync );
    }

    // memory shapes of the arrays (slowest..fastest)
    std::array<std::size_t, 3> real_shape_3d() const { return { nx_, ny_, nz_ }; }
    std::array<std::size_t, 3> complex_shape_3d() const { return { nx_, ny_, nz_ / 2 + 1 }; }
    std::array<std::size_t, 4> real_shape_4d() const
    {
        if ( strategy_traits<Strategy4D>::layout == FFTM_MI_LAYOUT_4D_PENCIL )
            return { nw_, nz_, ny_, nx_ };
        return { nx_, ny_, nz_, nw_ };
    }
    std::array<std::size_t, 4> complex_shape_4d() const
    {
        if ( strategy_traits<Strategy4D>::layout == FFTM_MI_LAYOUT_4D_SLAB )
            return { nx_, ny_, nz_, nw_ / 2 + 1 };
        return { ny_, nz_, nw_ / 2 + 1, nx_ };
    }

    std::size_t work_size_bytes() const { return work_.bytes(); }

private:
    void ensure_can_init_() const
    {
        if ( init_done_ )
            throw std::logic_error( "ffts::init can only be called once per instance." );
    }
    void ensure_initialized_() const
    {
        if ( !init_done_ )
            throw std::logic_error( "ffts: call init() before using the transform." );
    }
    void ensure_dimension_( std::size_t d ) const
    {
        ensure_initialized_();
        if ( dim_ != d )
            throw std::logic_error( "ffts: initialized for " + std::to_string( dim_ ) + "D but used as " +
                                    std::to_string( d ) + "D." );
    }

    void create_( int layout, const int64_t *n )
    {
        const int flags = opt_.destructive_backward ? FFTM_MI_FLAG_OVERWRITE_INPUT : 0;
        int       prec  = ::fftm::wrap::mi::precision_of<T>::value;
        if ( opt_.fp16_spectrum_storage )
        {
            if ( prec != FFTM_MI_F32 )
                throw std::logic_error( "ffts: FP16 spectrum storage needs float transforms" );
            prec = FFTM_MI_F32_H16;
        }
        FFTM_MI_SAFE_CALL( fftm_mi_ffts_create( layout, n, prec, flags, &h_ ) );
        std::size_t ws = 0;
        FFTM_MI_SAFE_CALL( fftm_mi_ffts_get_work_size( h_, &ws ) );
        if ( ws > 0 )
        {
            work_.init( ( ws + 255 ) / 256 * 256 );
            FFTM_MI_SAFE_CALL( fftm_mi_ffts_set_work_area( h_, work_.raw_ptr() ) );
        }
    }

    void run_( bool fwd, const void *in, void *out )
    {
        const int st = fwd ? fftm_mi_ffts_forward( h_, in, out, opt_.sync ? 1 : 0 )
                           : fftm_mi_ffts_backward( h_, in, out, opt_.sync ? 1 : 0 );
        ::fftm::wrap::mi::safe_call( st, fwd ? "fftm_mi_ffts_forward" : "fftm_mi_ffts_backward", __FILE__, __LINE__,
                                     opt_.sync );
    }

    void init_array_( const std::array<std::size_t, 2> &g, const ffts_init_options &o ) { init( g[0], g[1], o ); }
    void init_array_( const std::array<std::size_t, 3> &g, const ffts_init_options &o ) { init( g[0], g[1], g[2], o ); }
    void init_array_( const std::array<std::size_t, 4> &g, const ffts_init_options &o )
    {
        init( g[0], g[1], g[2], g[3], o );
    }

    ::fftm::wrap::mi::mi_wrap_many<T>   base_;
    ::fftm::wrap::mi::device_buffer<char> work_;
    fftm_mi_ffts                        h_ = nullptr;
    ffts_init_options                   opt_;
    bool                                init_done_ = false;
    std::size_t                         dim_       = 0;
    std::size_t                         nx_ = 0, ny_ = 0, nz_ = 0, nw_ = 0;
};

} // namespace mi
} // namespace fftm

#endif // FFTM_MI_FFTS_HPP
