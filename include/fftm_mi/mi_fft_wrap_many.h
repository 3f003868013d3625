// Named-plan container over mi::fft<T,D>: the ffts subset of
// fftm::wrap::fft_wrap_many<BaseFFTWrap,T> (source/external_wrap/fft_wrap_many.h:120-292,
// 375-395, 1606-1680, 1753-1802) without the SCFD dependency.  mi::fft offers the whole
// BaseFFTWrap surface the reference container uses, but the reference container itself
// includes SCFD's arrays (fft_wrap_many.h:15), which are absent here, so compiling it
// against mi::fft is untested; this container exists for that reason.
#ifndef FFTM_MI_FFT_WRAP_MANY_H
#define FFTM_MI_FFT_WRAP_MANY_H

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mi_fft_wrap.h"

namespace fftm
{
namespace wrap
{
namespace mi
{

template <template <typename, ::fftm::direction> class BaseFFTWrap, class T>
class fft_wrap_many
{
    using wrap_t = typename BaseFFTWrap<T, ::fftm::direction::C2CF>::base_t;

public:
    using real              = T;
    using complex           = typename wrap_t::complex;
    using runtime_api       = typename wrap_t::runtime_api;
    using plan_descriptor_t = typename wrap_t::plan_descriptor;
    using memory_type       = typename wrap_t::memory_type;

    fft_wrap_many() = default;
    ~fft_wrap_many()
    {
        try
        {
            release();
        }
        catch ( ... )
        {
        }
    }

    void release()
    {
        container_.clear();
        work_area_.free();
        work_area_size_     = 0;
        activated_          = false;
        external_activated_ = false;
        hot_exec_no_sync_   = false;
    }

    template <::fftm::direction D, std::size_t Rank>
    void add_plan( const std::string &name, const std::array<long long int, Rank> &n,
                   const std::array<long long int, Rank> &inembed, long long int istride, long long int idist,
                   const std::array<long long int, Rank> &onembed, long long int ostride, long long int odist,
                   long long int batch )
    {
        if ( activated_ )
            throw std::logic_error( "fft_wrap_many::add_plan: cannot add new plans after activation." );
        container_.emplace( name, std::unique_ptr<wrap_t>( new BaseFFTWrap<T, D>( n, inembed, istride, idist, onembed,
                                                                                  ostride, odist, batch ) ) );
    }

    template <::fftm::direction D>
    void add_plan_1D( const std::string &name, long long int n, long long int inembed, long long int istride,
                      long long int idist, long long int onembed, long long int ostride, long long int odist,
                      long long int batch )
    {
        add_plan<D>( name, std::array<long long int, 1>{ n }, std::array<long long int, 1>{ inembed }, istride, idist,
                     std::array<long long int, 1>{ onembed }, ostride, odist, batch );
    }

    template <::fftm::direction D>
    void add_plan_1D_default( const std::string &name, long long int n, long long int batch )
    {
        if ( activated_ )
            throw std::logic_error( "fft_wrap_many::add_plan_1D_default: cannot add new plans after activation." );
        container_.emplace( name,
                            std::unique_ptr<wrap_t>( new BaseFFTWrap<T, D>( std::array<long long int, 1>{ n }, batch ) ) );
    }

    template <::fftm::direction D>
    void add_plan_2D( const std::string &name, long long int n0, long long int n1, long long int inembed0,
                      long long int inembed1, long long int istride, long long int idist, long long int onembed0,
                      long long int onembed1, long long int ostride, long long int odist, long long int batch )
    {
        add_plan<D>( name, std::array<long long int, 2>{ n0, n1 }, std::array<long long int, 2>{ inembed0, inembed1 },
                     istride, idist, std::array<long long int, 2>{ onembed0, onembed1 }, ostride, odist, batch );
    }

    template <::fftm::direction D>
    void add_plan_3D( const std::string &name, long long int n0, long long int n1, long long int n2,
                      long long int inembed0, long long int inembed1, long long int inembed2, long long int istride,
                      long long int idist, long long int onembed0, long long int onembed1, long long int onembed2,
                      long long int ostride, long long int odist, long long int batch )
    {
        add_plan<D>( name, std::array<long long int, 3>{ n0, n1, n2 },
                     std::array<long long int, 3>{ inembed0, inembed1, inembed2 }, istride, idist,
                     std::array<long long int, 3>{ onembed0, onembed1, onembed2 }, ostride, odist, batch );
    }

    void activate()
    {
        if ( activated_ )
            throw std::logic_error( "fft_wrap_many::activate: cannot activate again." );
        if ( external_activated_ )
            throw std::logic_error( "fft_wrap_many::activate: cannot use external and internal activation, the class "
                                    "is externally activated." );
        activate_work_size();
        if ( work_area_size_ > 0 )
        {
            work_area_.init( work_area_size_ );
            for ( auto &el : container_ )
                el.second->set_work_area( static_cast<void *>( work_area_.raw_ptr() ) );
        }
        activated_ = true;
    }

    std::size_t activate_work_size()
    {
        for ( const auto &el : container_ )
            work_area_size_ = std::max( work_area_size_, el.second->get_work_size() );
        return work_area_size_;
    }

    void set_external_work_area( void *external_work_area )
    {
        if ( activated_ )
            throw std::logic_error( "fft_wrap_many::set_external_work_area: cannot set external work_area with "
                                    "activated local work_area." );
        for ( auto &el : container_ )
            el.second->set_work_area( external_work_area );
        external_activated_ = true;
    }

    std::size_t get_work_size() const { return work_area_size_; }

    void set_hot_exec_no_sync_enabled( bool enabled ) { hot_exec_no_sync_ = enabled; }

    template <class ArrayIn, class ArrayOut>
    void exec( const std::string &name, const ArrayIn &in, ArrayOut &out )
    {
        wrap_t *plan = container_.at( name ).get();
        void   *i    = const_cast<void *>( static_cast<const void *>( in.raw_ptr() ) );
        void   *o    = static_cast<void *>( out.raw_ptr() );
        if ( hot_exec_no_sync_ )
            plan->exec_no_sync( i, o );
        else
            plan->exec( i, o );
    }

    template <class ArrayIn, class ArrayOut>
    void exec_named_no_sync( const std::string &name, const ArrayIn &in, ArrayOut &out )
    {
        container_.at( name )->exec_no_sync( const_cast<void *>( static_cast<const void *>( in.raw_ptr() ) ),
                                             static_cast<void *>( out.raw_ptr() ) );
    }

    void        synchronize_device_execution() const { runtime_api::device_synchronize(); }
    std::size_t work_size( const std::string &name ) const { return container_.at( name )->get_work_size(); }
    void set_work_area( const std::string &name, void *work_area ) { container_.at( name )->set_work_area( work_area ); }
    void set_stream( const std::string &name, typename runtime_api::stream_t stream )
    {
        container_.at( name )->set_stream( stream );
    }
    void recreate_plan_with_stream_and_work_area( const std::string &name, typename runtime_api::stream_t stream,
                                                  void *work_area )
    {
        container_.at( name )->recreate_with_stream_and_work_area( stream, work_area );
    }
    plan_descriptor_t plan_descriptor( const std::string &name ) const { return container_.at( name )->descriptor(); }

    std::vector<std::pair<std::string, plan_descriptor_t>> plan_descriptors() const
    {
        std::vector<std::pair<std::string, plan_descriptor_t>> r;
        for ( const auto &el : container_ )
            r.emplace_back( el.first, el.second->descriptor() );
        return r;
    }

private:
    device_buffer<char>                            work_area_;
    std::size_t                                    work_area_size_     = 0;
    bool                                           activated_          = false;
    bool                                           external_activated_ = false;
    bool                                           hot_exec_no_sync_   = false;
    std::map<std::string, std::unique_ptr<wrap_t>> container_;
};

// the backend alias the reference's selector would name (cf. hipfft_wrap_many.h:12-13)
template <class T>
using mi_wrap_many = fft_wrap_many<::fftm::wrap::mi::fft, T>;

} // namespace mi
} // namespace wrap
} // namespace fftm

#endif // FFTM_MI_FFT_WRAP_MANY_H
