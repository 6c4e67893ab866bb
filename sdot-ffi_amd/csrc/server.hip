// The solve server ( see server.h ). Admission is by device memory: each slot's arena is grow-only, so what a slot holds
// is known after each request ( `Arena::bytes_held` ); a request is admitted on a slot when the other slots' holdings
// plus the larger of this slot's holding and the request's estimate stay within the budget.

#include "server.h"

#include <algorithm>
#include <chrono>

namespace sdot_amd {

Server::Server( int nb_slots, double budget_bytes ) : budget_( budget_bytes ), t0_( std::chrono::steady_clock::now() ) {
    SDOT_CHECK( nb_slots >= 1 && nb_slots <= 32, "server: 1 to 32 slots" );
    SDOT_CHECK( budget_bytes > 0, "server: a positive memory budget" );
    SDOT_HIP( hipGetDevice( &dev_ ) );
    held_.assign( size_t( nb_slots ), 0.0 );
    busy_.assign( size_t( nb_slots ), 0 );
    drop_.assign( size_t( nb_slots ), 0 );
    arenas_.assign( size_t( nb_slots ), nullptr );
    for ( int s = 0; s < nb_slots; ++s ) threads_.emplace_back( [ this, s ] { run( s ); } );
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk( mtx_ );
        stop_ = true;
    }
    cv_work_.notify_all();
    for ( auto &t : threads_ ) t.join();
}

void Server::sample() {
    double s = 0;
    for ( Arena *a : arenas_ ) if ( a ) s += double( a->bytes_held() );
    peak_ = std::max( peak_, s );
}

void Server::info( double &peak, I64 &done ) {
    std::lock_guard<std::mutex> lk( mtx_ );
    peak = peak_;
    done = nb_done_;
}

double Server::now() const { return std::chrono::duration<double>( std::chrono::steady_clock::now() - t0_ ).count(); }

I64 Server::submit( std::unique_ptr<Request> r ) {
    r->t_arrival = now();
    I64 t;
    {
        std::lock_guard<std::mutex> lk( mtx_ );
        SDOT_CHECK( ! stop_, "server: stopped" );
        t = next_ticket_++;
        r->job.in.options = r->opts.data();
        reqs_[ t ] = std::move( r );
        queue_.push_back( t );
    }
    cv_work_.notify_all();
    return t;
}

std::unique_ptr<Request> Server::wait( I64 ticket ) {
    std::unique_lock<std::mutex> lk( mtx_ );
    auto it = reqs_.find( ticket );
    SDOT_CHECK( it != reqs_.end(), "server: unknown ticket" );
    cv_done_.wait( lk, [ & ] { return it->second->done; } );
    std::unique_ptr<Request> r = std::move( it->second );
    reqs_.erase( it );
    return r;
}

void Server::run( int slot ) {
    hipStream_t st = nullptr;
    auto arena_own = std::make_unique<Arena>();
    bool ok = hipSetDevice( dev_ ) == hipSuccess && hipStreamCreateWithFlags( &st, hipStreamNonBlocking ) == hipSuccess;
    set_thread_arena( arena_own.get() );
    {
        std::lock_guard<std::mutex> lk( mtx_ );
        arenas_[ size_t( slot ) ] = arena_own.get();
    }
    for ( ;; ) {
        Request *r = nullptr;
        bool release = false;
        {
            std::unique_lock<std::mutex> lk( mtx_ );
            auto others_of = [&]( size_t s ) {
                double o = 0;
                for ( size_t u = 0; u < held_.size(); ++u ) if ( u != s ) o += held_[ u ];
                return o;
            };
            // the oldest waiting request this slot can take within the budget. THE FALLBACK, for a request that fits on no
            // slot: when every other slot idles and no idle slot could take the oldest request within the budget, this slot
            // takes it, and every other slot that holds memory is told to free it; the request starts once they have
            cv_work_.wait( lk, [ & ] {
                if ( stop_ || drop_[ size_t( slot ) ] ) return true;
                if ( queue_.empty() ) return false;
                const double others = others_of( size_t( slot ) );
                for ( auto q = queue_.begin(); q != queue_.end(); ++q ) {
                    Request *c = reqs_[ *q ].get();
                    if ( others + std::max( held_[ size_t( slot ) ], c->bytes ) <= budget_ ) {
                        r = c;
                        queue_.erase( q );
                        held_[ size_t( slot ) ] = std::max( held_[ size_t( slot ) ], c->bytes );   // ( reserved while it runs )
                        return true;
                    }
                }
                for ( size_t s = 0; s < busy_.size(); ++s ) if ( int( s ) != slot && busy_[ s ] ) return false;
                Request *c = reqs_[ queue_.front() ].get();
                for ( size_t s = 0; s < held_.size(); ++s )
                    if ( others_of( s ) + std::max( held_[ s ], c->bytes ) <= budget_ ) return false;   // ( that slot takes it )
                r = c;
                queue_.pop_front();
                release = true;
                held_[ size_t( slot ) ] = c->bytes;
                for ( size_t s = 0; s < drop_.size(); ++s ) if ( int( s ) != slot && held_[ s ] > 0 ) drop_[ s ] = 1;
                return true;
            } );
            if ( ! r && drop_[ size_t( slot ) ] && ! stop_ ) {   // another slot needs the memory this idle one holds
                lk.unlock();
                arena_own->release();
                lk.lock();
                drop_[ size_t( slot ) ] = 0;
                held_[ size_t( slot ) ] = 0;                 // ( only once the memory is freed )
                lk.unlock();
                cv_work_.notify_all();
                continue;
            }
            if ( ! r ) break;                           // ( stopping, nothing taken )
            busy_[ size_t( slot ) ] = 1;
            if ( release ) {                             // the other slots free their memory first
                cv_work_.notify_all();
                cv_work_.wait( lk, [ & ] {
                    if ( stop_ ) return true;
                    for ( size_t s = 0; s < drop_.size(); ++s ) if ( drop_[ s ] ) return false;
                    return true;
                } );
            }
        }
        BatchJob &J = r->job;
        J.t_start = now();
        if ( ! ok ) { J.kind = 1; J.msg = "server: the slot's stream could not be set up"; }
        else
            try {
                if ( release ) arena_own->release();
                solve_2d( J.in, J.out, st );
                SDOT_HIP( hipStreamSynchronize( st ) );
            } catch ( const Failure &e ) {
                J.kind = e.kind; J.code = e.code; J.value = e.value; J.msg = e.what();
                (void) hipStreamSynchronize( st );
            } catch ( const std::exception &e ) {
                J.kind = 1; J.msg = e.what();
                (void) hipStreamSynchronize( st );
            }
        J.t_end = now();
        {
            std::lock_guard<std::mutex> lk( mtx_ );
            held_[ size_t( slot ) ] = double( arena_own->bytes_held() );
            busy_[ size_t( slot ) ] = 0;
            r->done = true;
            ++nb_done_;
            sample();
        }
        cv_done_.notify_all();
        cv_work_.notify_all();                           // ( the memory this slot holds may have changed )
    }
    {
        std::lock_guard<std::mutex> lk( mtx_ );
        arenas_[ size_t( slot ) ] = nullptr;
    }
    set_thread_arena( nullptr );
    arena_own->release();
    if ( st ) (void) hipStreamDestroy( st );
}

} // namespace sdot_amd
