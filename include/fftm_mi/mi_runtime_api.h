// Host runtime helpers around the backend: the members of hip_runtime_api that ffts and
// fft_wrap_many use (source/external_wrap/hipfft_wrap.h:42-257), plus the owning device
// buffer this frontend uses in place of SCFD arrays (SCFD is not available here).
#ifndef FFTM_MI_RUNTIME_API_H
#define FFTM_MI_RUNTIME_API_H

#include <cstddef>
#include <stdexcept>
#include <string>

#include <hip/hip_runtime.h>

namespace fftm
{
namespace wrap
{
namespace mi
{

inline void hip_check( hipError_t e, const char *what )
{
    if ( e != hipSuccess )
        throw std::runtime_error( std::string( what ) + ": " + hipGetErrorString( e ) );
}

struct device_memory_tag
{
};

struct mi_runtime_api
{
    using memory_type = device_memory_tag;
    using stream_t    = hipStream_t;
    using event_t     = hipEvent_t;

    static void device_synchronize() { hip_check( hipDeviceSynchronize(), "hipDeviceSynchronize" ); }

    static stream_t create_nonblocking_stream()
    {
        // default priority (no stream priorities)
        stream_t s = nullptr;
        hip_check( hipStreamCreateWithFlags( &s, hipStreamNonBlocking ), "hipStreamCreateWithFlags" );
        return s;
    }

    static void destroy_stream( stream_t s )
    {
        if ( s != nullptr )
            (void)hipStreamDestroy( s );
    }

    static void stream_synchronize( stream_t s ) { hip_check( hipStreamSynchronize( s ), "hipStreamSynchronize" ); }

    static bool stream_ready( stream_t s ) { return hipStreamQuery( s ) == hipSuccess; }

    static event_t create_event()
    {
        event_t e = nullptr;
        hip_check( hipEventCreate( &e ), "hipEventCreate" );
        return e;
    }

    static void destroy_event( event_t e )
    {
        if ( e != nullptr )
            (void)hipEventDestroy( e );
    }

    static void record_event( event_t e, stream_t s ) { hip_check( hipEventRecord( e, s ), "hipEventRecord" ); }

    static float elapsed_ms( event_t a, event_t b )
    {
        float ms = 0.f;
        hip_check( hipEventElapsedTime( &ms, a, b ), "hipEventElapsedTime" );
        return ms;
    }

    static void memcpy_async( void *dst, const void *src, std::size_t bytes, stream_t s )
    {
        hip_check( hipMemcpyAsync( dst, src, bytes, hipMemcpyDefault, s ), "hipMemcpyAsync" );
    }

    static void memcpy( void *dst, const void *src, std::size_t bytes )
    {
        hip_check( hipMemcpy( dst, src, bytes, hipMemcpyDefault ), "hipMemcpy" );
    }

    static void *malloc_device( std::size_t bytes )
    {
        void *p = nullptr;
        if ( bytes > 0 )
            hip_check( hipMalloc( &p, bytes ), "hipMalloc" );
        return p;
    }

    static void free_device( void *p )
    {
        if ( p != nullptr )
            (void)hipFree( p );
    }
};

// Owning device buffer of n elements of E.
template <class E>
class device_buffer
{
public:
    device_buffer() = default;
    explicit device_buffer( std::size_t n ) { init( n ); }
    device_buffer( const device_buffer & )            = delete;
    device_buffer &operator=( const device_buffer & ) = delete;
    device_buffer( device_buffer &&o ) noexcept : ptr_( o.ptr_ ), n_( o.n_ )
    {
        o.ptr_ = nullptr;
        o.n_   = 0;
    }
    ~device_buffer() { free(); }

    void init( std::size_t n )
    {
        free();
        ptr_ = static_cast<E *>( mi_runtime_api::malloc_device( n * sizeof( E ) ) );
        n_   = n;
    }

    void free()
    {
        mi_runtime_api::free_device( ptr_ );
        ptr_ = nullptr;
        n_   = 0;
    }

    E          *raw_ptr() const { return ptr_; }
    std::size_t size() const { return n_; }
    std::size_t bytes() const { return n_ * sizeof( E ); }
    bool        is_free() const { return ptr_ == nullptr; }

private:
    E          *ptr_ = nullptr;
    std::size_t n_   = 0;
};

} // namespace mi
} // namespace wrap
} // namespace fftm

#endif // FFTM_MI_RUNTIME_API_H
