// A solve server: requests admitted by device memory as slots free up, each slot a stream with its own host thread.
#pragma once

#include "batch.h"

#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>

namespace sdot_amd {

/// one request: a solve's inputs and outputs ( the caller keeps them alive until the request is done ), its estimated
/// device footprint, what became of it
struct Request {
    BatchJob job;
    std::vector<double> opts;          ///< the request's own option vector
    double bytes = 0;                  ///< the admission's estimate of its device footprint
    double t_arrival = 0;              ///< seconds from the server's start: submitted, started, done
    bool done = false;
};

/// CONTINUOUS BATCHING of solves: `nb_slots` slots, each a HIP stream driven by its own thread with its own arena of
/// device buffers ( kept between its requests: the next solve of a similar size allocates nothing ). A slot takes the
/// OLDEST waiting request whose footprint fits the budget with what the other slots hold; a request that fits nowhere
/// while every slot idles gets a slot whose arena is released first, and the other idle slots free theirs. Requests
/// that arrive while others run are admitted as soon as a slot and the memory allow, and their work overlaps.
class Server {
public:
    Server( int nb_slots, double budget_bytes );
    ~Server();
    /// queues a request; returns its ticket
    I64 submit( std::unique_ptr<Request> r );
    /// blocks until the request is done, then hands it back ( and forgets it )
    std::unique_ptr<Request> wait( I64 ticket );
    double now() const;
    /// the largest sum of the slots' device bytes seen at the start and the end of a request, and the requests done
    void info( double &peak, I64 &done );

private:
    void run( int slot );
    int dev_ = 0;
    double budget_;
    std::vector<std::thread> threads_;
    std::vector<double> held_;         ///< per slot: the device bytes its arena holds ( or reserves )
    std::vector<Arena *> arenas_;      ///< per slot: its arena ( set by its thread )
    double peak_ = 0;                  ///< the largest sum of the arenas' bytes seen
    I64 nb_done_ = 0;
    void sample();                     ///< updates `peak_` ( the caller holds `mtx_` )
    std::vector<char> busy_, drop_;    ///< per slot: running a request; to free its buffers ( a request fits nowhere )
    std::deque<I64> queue_;            ///< waiting tickets, oldest first
    std::map<I64, std::unique_ptr<Request>> reqs_;
    I64 next_ticket_ = 1;
    bool stop_ = false;
    std::mutex mtx_;
    std::condition_variable cv_work_, cv_done_;
    std::chrono::steady_clock::time_point t0_;
};

} // namespace sdot_amd
