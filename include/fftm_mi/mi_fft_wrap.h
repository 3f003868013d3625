// C++ plan wrapper over the C ABI: the drop-in for fftm::wrap::hipfft::fft<T,D>.
//
// Provides fft_base<T> (the virtual surface of source/external_wrap/hipfft_wrap.h:472-496)
// and fft<T,D> with the same constructors (:520-605), exec / exec_no_sync /
// exec_direction(_no_sync) (:665-734), work-area and stream binding, descriptor, and the
// opaque-plan statics that fft_wrap_many.h calls for fftm's plan arrays (:741-1216).
// Errors follow the reference: a failed status throws std::runtime_error after a device
// synchronisation (FFTM_HIPFFT_SAFE_CALL, hipfft_safe_call.h:58-75); misuse throws
// std::logic_error.
#ifndef FFTM_MI_FFT_WRAP_H
#define FFTM_MI_FFT_WRAP_H

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../fftm_mi.h"
#include "mi_runtime_api.h"

#ifndef __FFTM_FFT_DIRECTION_H__
#define __FFTM_FFT_DIRECTION_H__
namespace fftm
{
// same enumerators and values as source/fft_direction.h:9-15
enum class direction : std::uint8_t
{
    R2C = 1,
    C2R,
    C2CF,
    C2CB
};
} // namespace fftm
#endif

namespace fftm
{
namespace wrap
{

#ifndef FFTM_FFT_PLAN_DESCRIPTOR_H
#define FFTM_FFT_PLAN_DESCRIPTOR_H
// field-for-field the descriptor of source/external_wrap/fft_plan_descriptor.h:15-32
struct fft_plan_descriptor
{
    direction                    dir            = direction::R2C;
    int                          backend_type   = 0;
    int                          rank           = 0;
    bool                         default_layout = false;
    std::array<long long int, 3> n              = { { 0, 0, 0 } };
    std::array<long long int, 3> inembed        = { { 0, 0, 0 } };
    long long int                istride        = 0;
    long long int                idist          = 0;
    std::array<long long int, 3> onembed        = { { 0, 0, 0 } };
    long long int                ostride        = 0;
    long long int                odist          = 0;
    long long int                batch          = 0;
    std::size_t                  work_size      = 0;
    std::uintptr_t               work_area_token = 0;
    std::uintptr_t               stream_token    = 0;
};
#endif

namespace mi
{

inline const char *status_string( int s )
{
    switch ( s )
    {
    case FFTM_MI_SUCCESS: return "success";
    case FFTM_MI_INVALID_PLAN: return "invalid plan";
    case FFTM_MI_ALLOC_FAILED: return "allocation failed";
    case FFTM_MI_INVALID_TYPE: return "invalid type";
    case FFTM_MI_INVALID_VALUE: return "invalid value";
    case FFTM_MI_INTERNAL_ERROR: return "internal error";
    case FFTM_MI_EXEC_FAILED: return "execution failed";
    case FFTM_MI_SETUP_FAILED: return "setup failed";
    case FFTM_MI_INVALID_SIZE: return "invalid size";
    case FFTM_MI_UNALIGNED_DATA: return "unaligned data";
    case FFTM_MI_INVALID_DEVICE: return "invalid device";
    case FFTM_MI_NO_WORKSPACE: return "missing workspace";
    case FFTM_MI_NOT_IMPLEMENTED: return "not implemented";
    case FFTM_MI_NOT_SUPPORTED: return "not supported";
    }
    return "unknown fftm_mi error";
}

// throw on a failed status, after synchronising the device (safe-call semantics)
inline void safe_call( int status, const char *expr, const char *file, int line, bool sync = true )
{
    const hipError_t se = sync ? hipDeviceSynchronize() : hipSuccess;
    if ( status != FFTM_MI_SUCCESS )
        throw std::runtime_error( std::string( "fftm_mi call failed at " ) + file + ":" + std::to_string( line ) +
                                  ": " + expr + ": " + status_string( status ) + ": " + fftm_mi_last_error() );
    if ( se != hipSuccess )
        throw std::runtime_error( std::string( "HIP synchronization failed after fftm_mi call at " ) + file + ":" +
                                  std::to_string( line ) + ": " + expr + ": " + hipGetErrorString( se ) );
}

#define FFTM_MI_SAFE_CALL( expr ) ::fftm::wrap::mi::safe_call( ( expr ), #expr, __FILE__, __LINE__, true )
#define FFTM_MI_CHECK_NO_SYNC( expr ) ::fftm::wrap::mi::safe_call( ( expr ), #expr, __FILE__, __LINE__, false )

template <typename T>
struct precision_of;
template <>
struct precision_of<float>
{
    static constexpr int value = FFTM_MI_F32;
};
template <>
struct precision_of<double>
{
    static constexpr int value = FFTM_MI_F64;
};

// interleaved complex of the backend (source/external_wrap/runtime_api_types.h:50-55)
template <class T>
struct alignas( 2 * sizeof( T ) ) complex_value
{
    T x;
    T y;
};

template <typename T>
class fft_base
{
public:
    using plan_descriptor = fft_plan_descriptor;
    using complex         = complex_value<T>;
    using real            = T;
    using runtime_api     = mi_runtime_api;
    using memory_type     = typename runtime_api::memory_type;

    virtual ~fft_base() {}
    virtual void            exec( void *in, void *out )                                        = 0;
    virtual void            exec_no_sync( void *in, void *out )                                = 0;
    virtual void            exec_direction( direction exec_dir, void *in, void *out )          = 0;
    virtual void            exec_direction_no_sync( direction exec_dir, void *in, void *out )  = 0;
    virtual direction       get_direction() const                                              = 0;
    virtual plan_descriptor descriptor() const                                                 = 0;
    virtual std::size_t     get_work_size() const                                              = 0;
    virtual void            set_work_area( void *work_area )                                   = 0;
    virtual void            set_stream( typename runtime_api::stream_t stream )                = 0;
    virtual void recreate_with_stream_and_work_area( typename runtime_api::stream_t stream, void *work_area ) = 0;
};

// An opaque plan for fftm's plan arrays: a created plan plus its bindings.
struct opaque_plan
{
    fftm_mi_plan plan   = nullptr;
    hipStream_t  stream = nullptr;
    void        *work   = nullptr;
};
using opaque_plan_handle_t = opaque_plan *;

template <typename T, direction D>
class fft : public fft_base<T>
{
public:
    using base_t               = fft_base<T>;
    using opaque_plan_handle_t = ::fftm::wrap::mi::opaque_plan_handle_t;
    struct minimal_reference_c2c_plan_array_t
    {
        std::vector<opaque_plan_handle_t> handles;
        std::vector<hipStream_t>          streams;
        std::vector<std::size_t>          offsets;
        fft_plan_descriptor               descriptor;
        std::size_t                       work_stride_bytes = 0;
    };
    using minimal_reference_c2c_plan_array_handle_t = minimal_reference_c2c_plan_array_t *;

    explicit fft( long long int n, long long int inembed, long long int istride, long long int idist,
                  long long int onembed, long long int ostride, long long int odist, long long int batch )
        : fft( std::array<long long int, 1>{ n }, std::array<long long int, 1>{ inembed }, istride, idist,
               std::array<long long int, 1>{ onembed }, ostride, odist, batch )
    {
    }

    explicit fft( long long int n0, long long int n1, long long int inembed0, long long int inembed1,
                  long long int istride, long long int idist, long long int onembed0, long long int onembed1,
                  long long int ostride, long long int odist, long long int batch )
        : fft( std::array<long long int, 2>{ n0, n1 }, std::array<long long int, 2>{ inembed0, inembed1 }, istride,
               idist, std::array<long long int, 2>{ onembed0, onembed1 }, ostride, odist, batch )
    {
    }

    explicit fft( long long int n0, long long int n1, long long int n2, long long int inembed0, long long int inembed1,
                  long long int inembed2, long long int istride, long long int idist, long long int onembed0,
                  long long int onembed1, long long int onembed2, long long int ostride, long long int odist,
                  long long int batch )
        : fft( std::array<long long int, 3>{ n0, n1, n2 }, std::array<long long int, 3>{ inembed0, inembed1, inembed2 },
               istride, idist, std::array<long long int, 3>{ onembed0, onembed1, onembed2 }, ostride, odist, batch )
    {
    }

    template <std::size_t Rank>
    explicit fft( const std::array<long long int, Rank> &n, const std::array<long long int, Rank> &inembed,
                  long long int istride, long long int idist, const std::array<long long int, Rank> &onembed,
                  long long int ostride, long long int odist, long long int batch )
    {
        static_assert( Rank >= 1 && Rank <= 3, "fft supports only 1D, 2D, and 3D plans" );
        descriptor_.dir          = D;
        descriptor_.backend_type = static_cast<int>( D );
        descriptor_.rank         = static_cast<int>( Rank );
        descriptor_.istride      = istride;
        descriptor_.idist        = idist;
        descriptor_.ostride      = ostride;
        descriptor_.odist        = odist;
        descriptor_.batch        = batch;
        for ( std::size_t i = 0; i < Rank; ++i )
        {
            descriptor_.n[i]       = n[i];
            descriptor_.inembed[i] = inembed[i];
            descriptor_.onembed[i] = onembed[i];
        }
        create_plan_();
    }

    template <std::size_t Rank>
    explicit fft( const std::array<long long int, Rank> &n, long long int batch )
    {
        static_assert( Rank >= 1 && Rank <= 3, "fft supports only 1D, 2D, and 3D plans" );
        descriptor_.dir            = D;
        descriptor_.backend_type   = static_cast<int>( D );
        descriptor_.rank           = static_cast<int>( Rank );
        descriptor_.default_layout = true;
        descriptor_.batch          = batch;
        for ( std::size_t i = 0; i < Rank; ++i )
            descriptor_.n[i] = n[i];
        create_plan_();
    }

    fft( const fft & )            = delete;
    fft &operator=( const fft & ) = delete;

    ~fft() override
    {
        if ( handle_ != nullptr )
            fftm_mi_plan_destroy( handle_ );
    }

    std::size_t get_work_size() const override { return work_size_; }

    void set_work_area( void *work_area ) override
    {
        descriptor_.work_area_token = reinterpret_cast<std::uintptr_t>( work_area );
        FFTM_MI_SAFE_CALL( fftm_mi_plan_set_work_area( handle_, work_area ) );
    }

    void set_stream( typename base_t::runtime_api::stream_t stream ) override
    {
        descriptor_.stream_token = reinterpret_cast<std::uintptr_t>( stream );
        FFTM_MI_SAFE_CALL( fftm_mi_plan_set_stream( handle_, static_cast<void *>( stream ) ) );
    }

    void recreate_with_stream_and_work_area( typename base_t::runtime_api::stream_t stream, void *work_area ) override
    {
        if ( handle_ != nullptr )
        {
            fftm_mi_plan_destroy( handle_ );
            handle_ = nullptr;
        }
        create_plan_();
        set_stream( stream );
        set_work_area( work_area );
    }

    void exec_typed( void *in, void *out ) { FFTM_MI_SAFE_CALL( fftm_mi_plan_exec( handle_, (int32_t)D, in, out, 1 ) ); }

    void exec_typed_no_sync( void *in, void *out )
    {
        FFTM_MI_CHECK_NO_SYNC( fftm_mi_plan_exec( handle_, (int32_t)D, in, out, 0 ) );
    }

    void exec_typed_direction( direction exec_dir, void *in, void *out )
    {
        check_direction_( exec_dir );
        FFTM_MI_SAFE_CALL( fftm_mi_plan_exec( handle_, (int32_t)exec_dir, in, out, 1 ) );
    }

    void exec_typed_direction_no_sync( direction exec_dir, void *in, void *out )
    {
        check_direction_( exec_dir );
        FFTM_MI_CHECK_NO_SYNC( fftm_mi_plan_exec( handle_, (int32_t)exec_dir, in, out, 0 ) );
    }

    void exec( void *in, void *out ) override { exec_typed( in, out ); }
    void exec_no_sync( void *in, void *out ) override { exec_typed_no_sync( in, out ); }
    void exec_direction( direction d, void *in, void *out ) override { exec_typed_direction( d, in, out ); }
    void exec_direction_no_sync( direction d, void *in, void *out ) override
    {
        exec_typed_direction_no_sync( d, in, out );
    }
    direction           get_direction() const override { return D; }
    fft_plan_descriptor descriptor() const override { return descriptor_; }
    fftm_mi_plan        native_plan() const { return handle_; }
    opaque_plan_handle_t opaque_plan_handle() const { return nullptr; }

    // --- opaque plans for fftm's C2C plan arrays (hipfft_wrap.h:741-1216) ------------
    static opaque_plan_handle_t create_empty_opaque_plan() { return new opaque_plan(); }

    static std::size_t make_opaque_1d_default( opaque_plan_handle_t plan, long long int n, long long int batch )
    {
        fftm_mi_plan_desc d = desc_1d_( n, 0, 0, 0, 0, 0, 0, batch, true );
        return make_opaque_( plan, d );
    }

    static std::size_t make_opaque_1d( opaque_plan_handle_t plan, long long int n, long long int inembed,
                                       long long int istride, long long int idist, long long int onembed,
                                       long long int ostride, long long int odist, long long int batch )
    {
        fftm_mi_plan_desc d = desc_1d_( n, inembed, istride, idist, onembed, ostride, odist, batch, false );
        return make_opaque_( plan, d );
    }

    static opaque_plan_handle_t create_opaque_c2c_1d( long long int n, long long int inembed, long long int istride,
                                                      long long int idist, long long int onembed,
                                                      long long int ostride, long long int odist,
                                                      long long int batch, std::size_t &work_size_out )
    {
        return create_opaque_c2c_1d_with_stream( n, inembed, istride, idist, onembed, ostride, odist, batch,
                                                 work_size_out, nullptr );
    }

    static opaque_plan_handle_t create_opaque_c2c_1d_with_stream( long long int n, long long int inembed,
                                                                  long long int istride, long long int idist,
                                                                  long long int onembed, long long int ostride,
                                                                  long long int odist, long long int batch,
                                                                  std::size_t &work_size_out, hipStream_t stream )
    {
        opaque_plan_handle_t h = new opaque_plan();
        try
        {
            fftm_mi_plan_desc d = desc_1d_( n, inembed, istride, idist, onembed, ostride, odist, batch, false );
            d.dir               = FFTM_MI_C2CF;
            work_size_out       = make_opaque_( h, d );
            set_opaque_stream( h, stream );
        }
        catch ( ... )
        {
            destroy_opaque_plan_noexcept( h );
            throw;
        }
        return h;
    }

    static opaque_plan_handle_t create_direct_raw_reference_c2c_1d_with_stream(
        long long int n, long long int inembed, long long int istride, long long int idist, long long int onembed,
        long long int ostride, long long int odist, long long int batch, std::size_t &work_size_out,
        hipStream_t stream )
    {
        return create_opaque_c2c_1d_with_stream( n, inembed, istride, idist, onembed, ostride, odist, batch,
                                                 work_size_out, stream );
    }

    static minimal_reference_c2c_plan_array_handle_t create_minimal_reference_c2c_plan_array_1d(
        long long int n, long long int inembed, long long int istride, long long int idist, long long int onembed,
        long long int ostride, long long int odist, long long int batch, const std::vector<std::size_t> &offsets )
    {
        auto *b    = new minimal_reference_c2c_plan_array_t();
        b->offsets = offsets;
        try
        {
            for ( std::size_t i = 0; i < offsets.size(); ++i )
            {
                hipStream_t s = nullptr;
                if ( hipStreamCreate( &s ) != hipSuccess )
                    throw std::runtime_error( "fftm_mi: hipStreamCreate failed" );
                b->streams.push_back( s );
                std::size_t ws = 0;
                b->handles.push_back(
                    create_opaque_c2c_1d_with_stream( n, inembed, istride, idist, onembed, ostride, odist, batch, ws, s ) );
                b->work_stride_bytes = std::max( b->work_stride_bytes, ws );
            }
            b->descriptor.dir        = direction::C2CF;
            b->descriptor.rank       = 1;
            b->descriptor.n[0]       = n;
            b->descriptor.inembed[0] = inembed;
            b->descriptor.istride    = istride;
            b->descriptor.idist      = idist;
            b->descriptor.onembed[0] = onembed;
            b->descriptor.ostride    = ostride;
            b->descriptor.odist      = odist;
            b->descriptor.batch      = batch;
            b->descriptor.work_size  = b->work_stride_bytes;
        }
        catch ( ... )
        {
            destroy_minimal_reference_c2c_plan_array_noexcept( b );
            throw;
        }
        return b;
    }

    static void destroy_opaque_plan_noexcept( opaque_plan_handle_t plan )
    {
        if ( plan == nullptr )
            return;
        if ( plan->plan != nullptr )
            fftm_mi_plan_destroy( plan->plan );
        delete plan;
    }

    static void destroy_minimal_reference_c2c_plan_array_noexcept( minimal_reference_c2c_plan_array_handle_t b )
    {
        if ( b == nullptr )
            return;
        for ( auto h : b->handles )
            destroy_opaque_plan_noexcept( h );
        for ( auto s : b->streams )
            if ( s != nullptr )
                (void)hipStreamDestroy( s );
        delete b;
    }

    static std::size_t minimal_reference_c2c_plan_array_size( minimal_reference_c2c_plan_array_handle_t b )
    {
        return b == nullptr ? 0 : b->handles.size();
    }

    static std::size_t minimal_reference_c2c_plan_array_work_size( minimal_reference_c2c_plan_array_handle_t b )
    {
        return b == nullptr ? 0 : b->work_stride_bytes;
    }

    static fft_plan_descriptor minimal_reference_c2c_plan_array_descriptor( minimal_reference_c2c_plan_array_handle_t b )
    {
        if ( b == nullptr )
            throw std::logic_error( "mi::fft::minimal_reference_c2c_plan_array_descriptor: null bundle" );
        return b->descriptor;
    }

    static std::uintptr_t minimal_reference_c2c_plan_array_first_handle_token( minimal_reference_c2c_plan_array_handle_t b )
    {
        return b == nullptr || b->handles.empty() ? 0 : reinterpret_cast<std::uintptr_t>( b->handles.front() );
    }

    static std::uintptr_t minimal_reference_c2c_plan_array_last_handle_token( minimal_reference_c2c_plan_array_handle_t b )
    {
        return b == nullptr || b->handles.empty() ? 0 : reinterpret_cast<std::uintptr_t>( b->handles.back() );
    }

    static std::uintptr_t minimal_reference_c2c_plan_array_first_stream_token( minimal_reference_c2c_plan_array_handle_t b )
    {
        return b == nullptr || b->streams.empty() ? 0 : reinterpret_cast<std::uintptr_t>( b->streams.front() );
    }

    static std::uintptr_t minimal_reference_c2c_plan_array_last_stream_token( minimal_reference_c2c_plan_array_handle_t b )
    {
        return b == nullptr || b->streams.empty() ? 0 : reinterpret_cast<std::uintptr_t>( b->streams.back() );
    }

    static void bind_minimal_reference_c2c_plan_array_work_areas( minimal_reference_c2c_plan_array_handle_t b,
                                                                  void *base, std::size_t stride_bytes )
    {
        if ( b == nullptr )
            return;
        char *raw = static_cast<char *>( base );
        for ( std::size_t i = 0; i < b->handles.size(); ++i )
            set_opaque_work_area( b->handles[i], raw == nullptr ? nullptr : raw + i * stride_bytes );
    }

    static void synchronize_minimal_reference_c2c_plan_array_streams( minimal_reference_c2c_plan_array_handle_t b )
    {
        if ( b == nullptr )
            return;
        for ( auto s : b->streams )
            mi_runtime_api::stream_synchronize( s );
    }

    static void set_opaque_stream( opaque_plan_handle_t plan, hipStream_t stream )
    {
        plan->stream = stream;
        if ( plan->plan != nullptr )
            FFTM_MI_SAFE_CALL( fftm_mi_plan_set_stream( plan->plan, static_cast<void *>( stream ) ) );
    }

    static void set_opaque_work_area( opaque_plan_handle_t plan, void *work_area )
    {
        plan->work = work_area;
        if ( plan->plan != nullptr )
            FFTM_MI_SAFE_CALL( fftm_mi_plan_set_work_area( plan->plan, work_area ) );
    }

    static void set_direct_raw_work_area( opaque_plan_handle_t plan, void *work_area )
    {
        set_opaque_work_area( plan, work_area );
    }

    static void exec_opaque_c2c( opaque_plan_handle_t plan, direction exec_dir, void *in, void *out )
    {
        FFTM_MI_SAFE_CALL( fftm_mi_plan_exec( plan->plan, (int32_t)exec_dir, in, out, 1 ) );
    }

    static void exec_opaque_c2c_no_sync( opaque_plan_handle_t plan, direction exec_dir, void *in, void *out )
    {
        FFTM_MI_CHECK_NO_SYNC( fftm_mi_plan_exec( plan->plan, (int32_t)exec_dir, in, out, 0 ) );
    }

    static void exec_direct_raw_c2c( opaque_plan_handle_t plan, direction exec_dir, void *in, void *out )
    {
        exec_opaque_c2c( plan, exec_dir, in, out );
    }

    static void exec_direct_raw_c2c_no_sync( opaque_plan_handle_t plan, direction exec_dir, void *in, void *out )
    {
        exec_opaque_c2c_no_sync( plan, exec_dir, in, out );
    }

    static void exec_minimal_reference_c2c_plan_array_no_sync( minimal_reference_c2c_plan_array_handle_t b,
                                                               direction exec_dir, void *in, void *out )
    {
        using C = complex_value<T>;
        for ( std::size_t i = 0; i < b->handles.size(); ++i )
            exec_opaque_c2c_no_sync( b->handles[i], exec_dir, static_cast<C *>( in ) + b->offsets[i],
                                     static_cast<C *>( out ) + b->offsets[i] );
    }

    static void exec_minimal_reference_c2c_plan_array( minimal_reference_c2c_plan_array_handle_t b,
                                                       direction exec_dir, void *in, void *out )
    {
        exec_minimal_reference_c2c_plan_array_no_sync( b, exec_dir, in, out );
        synchronize_minimal_reference_c2c_plan_array_streams( b );
    }

    static void exec_minimal_reference_c2c_plan_array_repeated( minimal_reference_c2c_plan_array_handle_t b,
                                                                direction exec_dir, void *in, void *out,
                                                                std::size_t repeats )
    {
        for ( std::size_t r = 0; r < repeats; ++r )
            exec_minimal_reference_c2c_plan_array( b, exec_dir, in, out );
    }

    static void exec_minimal_reference_c2c_plan_array_no_sync_repeated( minimal_reference_c2c_plan_array_handle_t b,
                                                                        direction exec_dir, void *in, void *out,
                                                                        std::size_t repeats )
    {
        for ( std::size_t r = 0; r < repeats; ++r )
            exec_minimal_reference_c2c_plan_array_no_sync( b, exec_dir, in, out );
    }

private:
    fftm_mi_plan        handle_    = nullptr;
    std::size_t         work_size_ = 0;
    fft_plan_descriptor descriptor_;

    static fftm_mi_plan_desc desc_1d_( long long int n, long long int inembed, long long int istride,
                                       long long int idist, long long int onembed, long long int ostride,
                                       long long int odist, long long int batch, bool default_layout )
    {
        fftm_mi_plan_desc d{};
        d.dir            = (int32_t)D;
        d.precision      = precision_of<T>::value;
        d.rank           = 1;
        d.default_layout = default_layout ? 1 : 0;
        d.n[0]           = n;
        d.inembed[0]     = inembed;
        d.istride        = istride;
        d.idist          = idist;
        d.onembed[0]     = onembed;
        d.ostride        = ostride;
        d.odist          = odist;
        d.batch          = batch;
        return d;
    }

    static std::size_t make_opaque_( opaque_plan_handle_t plan, const fftm_mi_plan_desc &d )
    {
        if ( plan->plan != nullptr )
            fftm_mi_plan_destroy( plan->plan );
        plan->plan = nullptr;
        FFTM_MI_SAFE_CALL( fftm_mi_plan_create( &d, &plan->plan ) );
        if ( plan->stream != nullptr )
            FFTM_MI_SAFE_CALL( fftm_mi_plan_set_stream( plan->plan, static_cast<void *>( plan->stream ) ) );
        if ( plan->work != nullptr )
            FFTM_MI_SAFE_CALL( fftm_mi_plan_set_work_area( plan->plan, plan->work ) );
        std::size_t ws = 0;
        FFTM_MI_SAFE_CALL( fftm_mi_plan_get_work_size( plan->plan, &ws ) );
        return ws;
    }

    void check_direction_( direction exec_dir ) const
    {
        if ( exec_dir == D )
            return;
        const bool c2c = ( D == direction::C2CF || D == direction::C2CB ) &&
                         ( exec_dir == direction::C2CF || exec_dir == direction::C2CB );
        if ( !c2c )
            throw std::logic_error( "mi::fft::exec_typed_direction: incompatible execution direction" );
    }

    void create_plan_()
    {
        fftm_mi_plan_desc d{};
        d.dir            = (int32_t)D;
        d.precision      = precision_of<T>::value;
        d.rank           = descriptor_.rank;
        d.default_layout = descriptor_.default_layout ? 1 : 0;
        for ( int i = 0; i < 3; ++i )
        {
            d.n[i]       = descriptor_.n[i];
            d.inembed[i] = descriptor_.inembed[i];
            d.onembed[i] = descriptor_.onembed[i];
        }
        d.istride = descriptor_.istride;
        d.idist   = descriptor_.idist;
        d.ostride = descriptor_.ostride;
        d.odist   = descriptor_.odist;
        d.batch   = descriptor_.batch;
        FFTM_MI_SAFE_CALL( fftm_mi_plan_create( &d, &handle_ ) );
        FFTM_MI_SAFE_CALL( fftm_mi_plan_get_work_size( handle_, &work_size_ ) );
        descriptor_.work_size = work_size_;
    }
};

} // namespace mi
} // namespace wrap
} // namespace fftm

#endif // FFTM_MI_FFT_WRAP_H
