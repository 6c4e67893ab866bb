// Many independent solves at once: each on its own stream, driven by its own host thread with its own arena.
#pragma once

#include "solver.h"

namespace sdot_amd {

/// one instance of a batch and what became of it
struct BatchJob {
    SolveInputs  in;
    SolveOutputs out;
    int          kind = 0;        ///< 0, or the error kind ( `Failure::kind` ) that stopped this instance
    int          code = 0;
    I64          value = 0;
    std::string  msg;
    double       t_start = 0, t_end = 0;   ///< seconds from the batch's start
};

/// runs every job, the largest first, on `nb_streams` streams ( at most 32 ), each with its own host thread and arena;
/// the jobs' results do not depend on the mix ( every kernel is deterministic and nothing is shared between two jobs
/// but the device ). Returns when all are done; each job records its own failure.
void solve_2d_batch( std::vector<BatchJob> &jobs, int nb_streams );

} // namespace sdot_amd
