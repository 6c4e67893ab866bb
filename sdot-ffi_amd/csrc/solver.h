// The damped-Newton solve of a 2D or 3D semi-discrete transport on the device, and its pieces.
#pragma once

#include "cells.h"
#include "cells3.h"

namespace sdot_amd {

/// the reference card's option vector ( `src/sdot/SdotPlanNd.py:197-201`, 41 doubles, same order )
enum Opt {
    O_TOL_ABS = 0, O_TOL_REL, O_T_MIN, O_MULT_OK, O_FACTOR, O_MAXIT, O_MAX_BACKTRACKS, O_STEP, O_RESIDUAL, O_POWER, O_SWITCH,
    O_LIN, O_HOST_METHOD, O_LIN_TOL, O_AMG_VARIANT, O_MG_SHIFT, O_MG_RECYCLE, O_MG_REBUILD, O_MG_STOP, O_MG_NU, O_MG_KCYCLE,
    O_LIN_MAXIT, O_TRACE, O_MG_FLOAT, O_MG_SMOOTHED,
    O_CONTINUATION, O_CONV_THRESHOLD, O_CONV_S0, O_CONV_RATIO, O_CONV_MIN, O_CONV_POSSIBLE, O_MIN_SCALE,
    O_IMG_X0, O_IMG_Y0, O_IMG_HX, O_IMG_HY, O_IMG_NX, O_IMG_NY, O_AGG_MARGIN, O_AGG_GAP, O_AGG_NB_DUPS,
    NB_OPTS
};

/// `stats` ( `src/sdot/include/sdot/sdotplan/Report.h:15-21` )
enum Stat {
    STATUS = 0, RESIDUAL, RESIDUAL0, NB_ITER, NB_DIAG, NB_BACKTRACKS, T_MAJORANT, T_DIAG, T_ASM, T_LIN, T_LIM, EPS,
    DOMAIN_MASS, NB_OVERFLOWED, NB_CELL_LIM, NB_LIMIT_ROUNDS, LIN_NB_HIERARCHIES, LIN_NB_ITER, LIN_WORST, START, T_TOTAL,
    NB_CONTINUATION_STEPS, MIN_START_MASS, IT_SWITCH, IT_DOUBLE, SCRATCH_BYTES,
    NB_CLUSTERS, NB_AGGREGATED, NB_DUPLICATES, RESIDUAL_FULL, NB_POLISH, NB_STATS
};
enum Hist { H_STEP = 0, H_T, H_RESIDUAL_L2, H_MIN_MASS, H_MAX_RESIDUAL, H_NB_DIAG, H_NB_EVALS, H_S, NB_HIST };
enum Status { S_RUNNING = 0, S_CONVERGED = 1, S_MAX_ITERATIONS = 2, S_STAGNATION = 3, S_LINEAR_FAILURE = 4, S_CAPACITY = 5, S_FAILURE = 6,
              S_CONVERGED_AGGREGATED = 7 };
enum Residual { R_LIN = 0, R_LOG = 1, R_POWER = 2 };

/// the sparse Laplacian of one diagram, rank order ( rows of `DiagramOut` )
struct Csr {
    I64 n = 0, nnz = 0, step = 1;       ///< entry `q` of row `i` at `row_off[ i ] + q * step`
    I64 cap = 0;                        ///< the entries `col` / `val` hold ( every `row_off[ i ] + q * step` is below it )
    const int *row_off = nullptr, *row_cnt = nullptr, *col = nullptr;
    const double *val = nullptr, *dia = nullptr;
};

struct CgStats {
    int nb_iter = 0;
    double worst = 0;
    // the last solve's end: its residual by the recurrence and the true one ( relative ), a breakdown ( `p . q <= 0` )
    double last_err = 0, last_true = 0;
    bool last_breakdown = false;
    int nb_fp64_fallbacks = 0;          ///< systems the FP32 preconditioner did not bring to the tolerance, redone with the FP64 one
};

/// Jacobi-preconditioned CG for `L d = b`, the unknown `gauge` held at 0 ( reference
/// `sdotplan/Linear.cpp:44-88` ); `b`, `d` device vectors of `n`. Returns the iteration count or -1.
int cg_solve( const Csr &L, int gauge, const double *b, double *d, double tol, int maxit, hipStream_t st, CgStats &cs );

/// one solve; every pointer as documented in `include/sdot_amd.h`
struct SolveInputs {
    I64 n = 0;
    const double *positions = nullptr, *nu = nullptr, *w0 = nullptr;   ///< device, user order
    const double *options = nullptr;                                     ///< host [ 41 ]
    const long long *dups = nullptr;                                     ///< device [ 2 nb_dups ], ( dup, rep ) user ids
    I64 nb_dups = 0;
    double box[ 4 ] = { 0, 0, 1, 1 };
    Density dens;
    // 3D ( `dim == 3` ): positions `[ n ][ 3 ]`, the box `lo x, y, z, hi x, y, z`, a constant or a sum of Gaussians; the
    // step is by trials, and there are no duplicates
    int dim = 2;
    int precision = 0;                                                   ///< 2D cells: 0 FP64, 1 FP32, 2 mixed
    double box3[ 6 ] = { 0, 0, 0, 1, 1, 1 };
    Density3 dens3;
};
struct SolveOutputs {
    double *weights = nullptr, *cell_masses = nullptr, *barycenters = nullptr;   ///< device, user order
    long long *clusters = nullptr;                                              ///< device, user order ( may be null )
    double *history = nullptr;                                                  ///< host [ ( maxit + 1 ) * 8 ]
    long long *nb_steps = nullptr;                                              ///< host
    double *stats = nullptr;                                                    ///< host [ 31 ]
    double *cost = nullptr;                                                     ///< host [ 1 ]
};

/// one solve, 2D or 3D after `in.dim`
void solve_2d( const SolveInputs &in, const SolveOutputs &out, hipStream_t st );

/// the masses ( and facets ) of one diagram, user order ( see `sdot_amd_measures_2d` )
void measures_2d( hipStream_t st, I64 n, const double *positions, const double *weights, const double *box, const Density &dn,
                  double *masses, long long *fi, long long *fj, double *fc, I64 facet_cap, I64 *nb_facets, unsigned long long *events,
                  bool fp32 = false );

/// the masses ( and facets ) of one 3D diagram, user order ( see `sdot_amd_measures_3d` )
void measures_3d( hipStream_t st, I64 n, const double *positions, const double *weights, const double *box, const Density3 &dn,
                  double *masses, long long *fi, long long *fj, double *fc, I64 facet_cap, I64 *nb_facets, unsigned long long *events );

/// `( masses, first, second )` of each cell ( `int rho`, `int x rho`, `int |x|^2 rho` ), user order ( see `sdot_amd_moments_2d` )
void moments_2d( hipStream_t st, I64 n, const double *positions, const double *weights, const double *box, const Density &dn, double *masses,
                 double *first, double *second );
void moments_3d( hipStream_t st, I64 n, const double *positions, const double *weights, const double *box, const Density3 &dn, double *masses,
                 double *first, double *second );

/// the cells' vertices ( see `sdot_amd_cells_2d` ); returns the vertex count, and writes nothing past the offsets' count when
/// `xy` is null or the count passes `cap`
I64 cells_2d( hipStream_t st, I64 n, const double *positions, const double *weights, const double *box, const Density &dn, long long *offsets,
              double *xy, long long *cut, I64 cap );

/// the adjoint of the masses ( see `sdot_amd_measures_vjp_2d` )
void measures_vjp_2d( hipStream_t st, I64 n, const double *positions, const double *weights, const double *box, const Density &dn,
                      const double *grad_masses, double *grad_weights, double *grad_positions );

} // namespace sdot_amd
