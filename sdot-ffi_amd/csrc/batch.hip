// Many independent solves at once. A solve of a few thousand to a few hundred thousand seeds leaves most of the chip
// idle between its launches ( its diagrams take a fraction of a millisecond, its multigrid levels are small ); several
// solves on separate streams fill it. Each stream is driven by its own host thread ( the solve waits on its stream at
// its decision points, so one thread per stream keeps them independent ) with its own arena of device buffers and
// pinned staging. Jobs are taken largest first from a shared counter ( longest processing time first ).

#include "batch.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <numeric>
#include <thread>

namespace sdot_amd {

void solve_2d_batch( std::vector<BatchJob> &jobs, int nb_streams ) {
    SDOT_CHECK( nb_streams >= 1 && nb_streams <= 32, "batch: 1 to 32 streams" );
    if ( jobs.empty() ) return;
    int dev = 0;
    SDOT_HIP( hipGetDevice( &dev ) );
    std::vector<size_t> order( jobs.size() );
    std::iota( order.begin(), order.end(), size_t( 0 ) );
    std::stable_sort( order.begin(), order.end(), [&]( size_t a, size_t b ) { return jobs[ a ].in.n > jobs[ b ].in.n; } );
    std::atomic<size_t> next{ 0 };
    const auto t0 = std::chrono::steady_clock::now();
    auto secs = [&]() { return std::chrono::duration<double>( std::chrono::steady_clock::now() - t0 ).count(); };
    const int nw = int( std::min<size_t>( size_t( nb_streams ), jobs.size() ) );
    std::vector<std::string> setup_err( ( size_t( nw ) ) );
    auto worker = [&]( int wid ) {
        hipStream_t st = nullptr;
        auto arena_own = std::make_unique<Arena>();
        try {
            SDOT_HIP( hipSetDevice( dev ) );
            SDOT_HIP( hipStreamCreateWithFlags( &st, hipStreamNonBlocking ) );
        } catch ( const std::exception &e ) {
            setup_err[ size_t( wid ) ] = e.what();
            return;
        }
        set_thread_arena( arena_own.get() );
        for ( ;; ) {
            const size_t q = next.fetch_add( 1 );
            if ( q >= order.size() ) break;
            BatchJob &J = jobs[ order[ q ] ];
            J.t_start = secs();
            try {
                solve_2d( J.in, J.out, st );
                SDOT_HIP( hipStreamSynchronize( st ) );
            } catch ( const Failure &e ) {
                J.kind = e.kind; J.code = e.code; J.value = e.value; J.msg = e.what();
                (void) hipStreamSynchronize( st );
            } catch ( const std::exception &e ) {
                J.kind = 1; J.msg = e.what();
                (void) hipStreamSynchronize( st );
            }
            J.t_end = secs();
        }
        set_thread_arena( nullptr );
        arena_own->release();
        (void) hipStreamDestroy( st );
    };
    std::vector<std::thread> threads;
    for ( int w = 0; w < nw; ++w ) threads.emplace_back( worker, w );
    for ( auto &t : threads ) t.join();
    for ( int w = 0; w < nw; ++w )
        SDOT_CHECK( setup_err[ size_t( w ) ].empty(), "batch: a stream could not be set up: " + setup_err[ size_t( w ) ] );
}

} // namespace sdot_amd
