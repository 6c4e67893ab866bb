// The multigrid preconditioned CG of the Newton direction: `L d = b` for the Laguerre-graph Laplacian.
#pragma once

#include "solver.h"

namespace sdot_amd {

struct MgOptions {
    int    pack = 8;          ///< seeds per aggregate ( consecutive tree ranks )
    int    nu = 2;            ///< Jacobi sweeps before and after the coarse correction
    double omega = 0.7;       ///< their damping when `chebyshev` is off
    /// the `nu` sweeps as one Chebyshev smoother: sweep `s` damped by `1 / lambda_s`, the Chebyshev nodes of
    /// `[ 2 / ratio, 2 ]` ( 2 bounds the spectrum of `D^-1 A` for a Laplacian whose diagonal is its row sum, by
    /// Gershgorin ), so that the `nu` sweeps together damp that interval by the Chebyshev polynomial's bound
    bool   chebyshev = true;
    double ratio = 8;
    int    kcycle = 2;        ///< levels 1 .. kcycle take their coarse correction by two flexible CG steps ( a K-cycle ); below, one cycle
    I64    bottom = 1024;     ///< the levels from the first of at most this many rows run as one single-workgroup kernel ( 0: none )
    bool   graphs = true;     ///< the batches of iterations replayed from HIP graphs
    /// THE FP32 PRECONDITIONER: the flexible CG, its SpMV and its vectors in FP64, its multigrid cycle ( levels, smoothing,
    /// coarse solves ) in FP32 from an FP32 copy of the rows ( the reference card's float levels, `gpu/Linear2D.cuh:993-1000` );
    /// the FP64 cycle when that solve does not converge
    bool   fp32 = false;
};

/// `L d = b` ( `b` of zero sum ) by flexible CG preconditioned with one multigrid cycle whose aggregates are `pack`
/// consecutive tree ranks ( the cycle in FP32 when `o.fp32` ); the solution shifted so that `d[ gauge ] = 0`.
/// Returns the iteration count, or -1 when it did not reach `tol` ( relative residual ) in `maxit`, or
/// -2 when the hierarchy could not be built ( a coarse row with too many neighbours ).
int mg_solve( const Csr &L, int gauge, const double *b, double *d, double tol, int maxit, const MgOptions &o, hipStream_t st, CgStats &cs,
              int &nb_levels );

} // namespace sdot_amd
