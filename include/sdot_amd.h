/* The C ABI of the MI355X semi-discrete optimal transport library ( libsdot_amd.so ).
 *
 * It replaces the FFI calls that the reference `sdot` makes on its GPU path for one damped-Newton
 * solve and the diagrams it is made of. Each entry point names the reference call it stands for
 * ( paths relative to the reference's repository root ). The calling convention is the same for all:
 *
 *   - `stream` is a `hipStream_t` ( null: the default stream ); work is queued on it, and the call
 *     returns once its results are in place;
 *   - arrays marked DEVICE are device pointers ( `hipMalloc`, or a torch tensor's `data_ptr()` ),
 *     arrays marked HOST are host pointers;
 *   - 2D positions are `[ n ][ 2 ]` float64, C order, in the USER's order ( the tree's permutation is
 *     internal ); weights and masses are `[ n ]` float64 in the user's order;
 *   - the return value is 0 on success, otherwise the error kind: 1 invalid argument / HIP error,
 *     2 kernel failure ( `code` 1: a cell past the vertex limit, `value` its seed ), 3 capacity,
 *     4 device memory; `sdot_amd_last_error` gives the message and the details of the calling thread's
 *     last failure.
 */
#ifndef SDOT_AMD_H
#define SDOT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define SDOT_AMD_ABI_VERSION 1

/* the density: SDOT_AMD_CONSTANT ( `rho` on the domain box ), SDOT_AMD_IMAGE ( piecewise constant on
 * `nx * ny` pixels, `img[ j * nx + i ]` the value of pixel `( i, j )` covering
 * `[ x0 + i hx, x0 + ( i + 1 ) hx ] x [ y0 + j hy, y0 + ( j + 1 ) hy ]` ) or SDOT_AMD_GAUSSIANS ( a sum of
 * `nx` isotropic Gaussians truncated to the domain box, `img[ 4 q .. 4 q + 3 ]` = centre x, centre y, standard
 * deviation and mass of Gaussian `q`: the reference card's `g_pos`, `g_sigma`, `g_mass`, SdotPlanNd.py:837-862 ) */
#define SDOT_AMD_CONSTANT  0
#define SDOT_AMD_IMAGE     1
#define SDOT_AMD_GAUSSIANS 2

int         sdot_amd_abi_version( void );

/* the message of the calling thread's last failure, and its kind / code / value */
const char *sdot_amd_last_error( int *kind, int *code, long long *value );

/* the bytes of device memory the library holds ( its grow-only arena ), and their release */
long long   sdot_amd_device_bytes( void );
void        sdot_amd_release( void );

/* LIVE KERNEL TIMING ( no counterpart in the reference, which times stages on the host: `Newton2D.cuh`'s
 * `StageTimer` ). When on, HIP events are recorded on the launch stream around each launch of the cell
 * kernels ( index 0: first pass, 1: overflow pass ), one SpMV launch in sixteen of the CG ( index 2 ) and
 * the majorant refresh ( index 3 ).
 * `sdot_amd_kernel_times` returns, per index, the summed milliseconds and the launches timed, the
 * algorithmic bytes of the timed SpMV launches, and the cell kernels' event counters summed over every
 * diagram since the last reset ( HOST [ 5 ]: side tests, pruning vertex tests, new vertices, node
 * visits, cuts ); `reset` clears them afterwards. */
int sdot_amd_set_timing( int on );
int sdot_amd_kernel_times( double *ms, long long *launches, double *spmv_bytes, unsigned long long *events, int reset );

/* THE MULTIGRID'S HIP GRAPHS ( no counterpart in the reference, whose card launches every kernel ): the batches of flexible-CG
 * iterations captured as graphs and the replays of them, since the last `sdot_amd_kernel_times( ..., reset = 1 )`, counted
 * whether timing is on or not; a solve that replays far more than it captures reuses its graphs. HOST [ 1 ] each. */
int sdot_amd_graph_counts( long long *captures, long long *replays );

/* THE FP32 KERNEL'S TOPOLOGY REPAIRS ( no counterpart in the reference, whose float kernel keeps what it found ): over the
 * diagrams solves made with the FP32 kernel since the last `sdot_amd_kernel_times( ..., reset = 1 )`, the diagrams, the cells
 * redone by the FP64 kernel because a facet was found by one of its two cells only, and the facets still one-sided after that
 * ( dropped from the Laplacian ). HOST [ 1 ] each, may be null. */
int sdot_amd_fp32_counts( long long *diagrams, long long *redone, long long *dropped );

/* THE MULTIGRID'S FP32 PRECONDITIONER ( the reference's card keeps its levels in float, `gpu/Linear2D.cuh:993-1000` ): since
 * the last `sdot_amd_kernel_times( ..., reset = 1 )`, the Newton systems solved by the FP64 flexible CG with the FP32 multigrid
 * cycle as its preconditioner, their iterations, and the systems redone with the FP64 cycle because that solve did not converge.
 * HOST [ 1 ] each, may be null. */
int sdot_amd_mg32_counts( long long *solves, long long *iterations, long long *fallbacks );

/* ONE DIAGRAM: the masses of the power cells of `( positions, weights )` in the box
 * `box = { lo_x, lo_y, hi_x, hi_y }` ( HOST ) against a density, and optionally the facets
 * `c_ij = int_facet rho / ( 2 |p_i - p_j| )` as triplets ( rows in user order, each row in its cell's
 * edge order ).
 * Replaces `power_diagram_measures_card_2d` -> `sdot::gpu2d::measures`
 * ( src/sdot/include/sdot/gpu/Cell2D.cuh:1801-1811, called from src/sdot/PowerDiagram_Bsp.py:271-311 ).
 *   positions    DEVICE [ n ][ 2 ]
 *   weights      DEVICE [ n ]       ( null: all zero, a Voronoi diagram )
 *   img          DEVICE [ ny * nx ] ( SDOT_AMD_IMAGE; img_geo HOST { x0, y0, hx, hy } ) or [ nx ][ 4 ]
 *                                   ( SDOT_AMD_GAUSSIANS, `nx` Gaussians; `ny`, `img_geo` unused )
 *   masses       DEVICE [ n ]       ( output )
 *   facet_i/j/c  DEVICE [ facet_cap ] ( outputs; null: no facets )
 *   nb_facets    HOST   [ 1 ]       ( output, may be null )
 *   events       HOST   [ 5 ]       ( may be null: side tests, pruning vertex tests, new vertices, node visits, cuts )
 */
int sdot_amd_measures_2d( void *stream, long long n, const double *positions, const double *weights, const double *box,
                          int density_kind, double rho, const double *img, int nx, int ny, const double *img_geo,
                          double *masses, long long *facet_i, long long *facet_j, double *facet_c, long long facet_cap,
                          long long *nb_facets, unsigned long long *events );

/* ONE DIAGRAM with the kernel's float chosen: `sdot_amd_measures_2d`'s arguments, then `precision` SDOT_AMD_FP64 ( that call )
 * or SDOT_AMD_FP32 ( the walk and the cuts in float on 32-byte float node records, every vertex re-solved in double from its two
 * planes; a cell that fails the re-solve's conditioning test, and both cells of a facet only one of them found, are redone by
 * the FP64 kernel; the rows made symmetric ). SDOT_AMD_FP64 .. SDOT_AMD_MIXED are defined below, with the solve.
 * Replaces `power_diagram_measures_card_2d` with the kernel variant of `kernel_dtype` ( `card_variant_for( 32 | 64, ... )`,
 * src/sdot/PowerDiagram_Bsp.py:421-443, src/sdot/PowerDiagram.py:375-391 ). */
int sdot_amd_measures_2d_precision( void *stream, long long n, const double *positions, const double *weights, const double *box,
                                    int density_kind, double rho, const double *img, int nx, int ny, const double *img_geo,
                                    double *masses, long long *facet_i, long long *facet_j, double *facet_c, long long facet_cap,
                                    long long *nb_facets, unsigned long long *events, int precision );

/* ONE DIAGRAM IN A POLYGON DOMAIN: `sdot_amd_measures_2d_precision`'s arguments, then the domain, a convex polygon of
 * `nb_domain` ( 3 to 12 ) vertices `domain_xy` HOST [ nb_domain ][ 2 ], counterclockwise. Each cell starts from the polygon
 * instead of the box ( `box` stays the polygon's bounding box ); edge `i` ( vertex `i` to `i + 1` ) is the cut `-5 - i`. The
 * density is zero outside the polygon. Replaces the card's start cell ( `Cuts`, `start_vertices`: src/sdot/include/sdot/gpu/
 * Cell2D.cuh:295-333, 465-478, laid down from src/sdot/StartCell.py and src/sdot/PowerDiagram.py:166-195 ). */
int sdot_amd_measures_2d_domain( void *stream, long long n, const double *positions, const double *weights, const double *box,
                                 int density_kind, double rho, const double *img, int nx, int ny, const double *img_geo,
                                 double *masses, long long *facet_i, long long *facet_j, double *facet_c, long long facet_cap,
                                 long long *nb_facets, unsigned long long *events, int precision, int nb_domain, const double *domain_xy );

/* THE MOMENTS OF ONE DIAGRAM: for each cell `int rho` ( `masses` ), `int x rho` ( `first`, DEVICE [ n ][ 2 ] ) and `int |x|^2 rho`
 * ( `second` ), user order: enough for a transport cost ( `second - 2 p . first + |p|^2 mass` ) and the barycentres
 * ( `first / mass` ). The arguments of `sdot_amd_measures_2d` otherwise ( any of its densities ).
 * Replaces `power_diagram_moments` ( src/sdot/PowerDiagram.py:433-459, `diagram/Ops.h:187-228` ). */
int sdot_amd_moments_2d( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                         double rho, const double *img, int nx, int ny, const double *img_geo, double *masses, double *first, double *second );

/* The same in 3D ( `first` DEVICE [ n ][ 3 ] ), the density as `sdot_amd_measures_3d`'s. */
int sdot_amd_moments_3d( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                         double rho, const double *gauss, int ng, double *masses, double *first, double *second );

/* THE CELLS THEMSELVES: the vertices of every cell, counter-clockwise, in user order. Cell `i` owns the vertices
 * `offsets[ i ] .. offsets[ i + 1 ]` ( DEVICE [ n + 1 ] int64 ) of `vertices` ( DEVICE [ vertex_cap ][ 2 ], absolute coordinates ), and
 * `cuts[ v ]` ( DEVICE [ vertex_cap ] int64 ) names the cut of the edge that leaves vertex `v`: a seed's user index, or `-1 - f` for
 * side `f` of the box ( 0 bottom, 1 right, 2 top, 3 left ). `nb_vertices` HOST [ 1 ]: the total. With `vertices` null, only the
 * total is computed ( a first call to size the buffers ); past `vertex_cap`, kind 3 with the total in `value`.
 * Replaces `power_diagram_cells` ( src/sdot/PowerDiagram.py:490-515, `diagram/Ops.h` `build_cell` ). */
int sdot_amd_cells_2d( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                       double rho, const double *img, int nx, int ny, const double *img_geo, long long *offsets, double *vertices,
                       long long *cuts, long long vertex_cap, long long *nb_vertices );

/* THE ADJOINT OF THE MASSES: the cotangent `grad_masses` ( DEVICE [ n ], user order ) -> `grad_weights` ( DEVICE [ n ],
 * `sum_i g_i d m_i / d w` = `L g` ) and `grad_positions` ( DEVICE [ n ][ 2 ], `sum_i g_i d m_i / d p` ); either may be null. A
 * constant or an image density ( the facets' first moments of a sum of Gaussians are not written: kind 1 ).
 * Replaces `sdot::gpu2d::measures_vjp` ( src/sdot/include/sdot/gpu/Cell2D.cuh:1816-1839; the CPU's `measures_bwd`,
 * `diagram/Ops.h:437-459` ). */
int sdot_amd_measures_vjp_2d( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                              double rho, const double *img, int nx, int ny, const double *img_geo, const double *grad_masses,
                              double *grad_weights, double *grad_positions );

/* ONE DIAGRAM'S MASSES WITHOUT ALLOCATION. No counterpart in the reference: this entry point is for callers that cannot allocate
 * device memory ( code under HIP graph capture, tight serving loops, harnesses that forbid `hipMalloc` ). It gives the masses of
 * `sdot_amd_measures_2d` ( the same arguments without the facets; the same values, to the bit ) and allocates nothing, device
 * or pinned host: its scratch ( the tree, the majorants, the rank-order masses, the overflow lists and slots ) is a fixed device
 * array of 40 MiB that each GPU the library runs on holds until the process ends. At most
 * `sdot_amd_measures_2d_noalloc_max_seeds()` seeds ( 131072 ); more is an invalid argument ( kind 1 ). Calls on one GPU share
 * that GPU's array: the library serializes them ( a lock held for the whole call, which returns once the masses are in place ),
 * and a caller must not rely on two calls on one GPU overlapping. */
int       sdot_amd_measures_2d_noalloc( void *stream, long long n, const double *positions, const double *weights, const double *box,
                                        int density_kind, double rho, const double *img, int nx, int ny, const double *img_geo,
                                        double *masses, unsigned long long *events );
long long sdot_amd_measures_2d_noalloc_max_seeds( void );

/* THE SOLVE: weights `w` such that every cell's mass matches its target `nu_i`, by the damped Newton of
 * Kitagawa-Merigot-Thibert, FP64 cells; the step by trials or by the limits of the cells ( option `step` ),
 * directions by Jacobi-PCG or multigrid-PCG ( option `lin` ), the width continuation of an image or of
 * Gaussians ( options `continuation`, `conv_*` ).
 * Replaces `sdotplan_solve_card_2d` / `_image` / `_gauss` -> `sdot::gpu2d::solve`
 * ( src/sdot/include/sdot/gpu/Newton2D.cuh:846-850, called from src/sdot/SdotPlanNd.py:864-890 ).
 *   positions    DEVICE [ n ][ 2 ]
 *   nu           DEVICE [ n ]       target masses ( rescaled to the mass of the domain )
 *   w0           DEVICE [ n ]       starting weights ( null: zero, the Voronoi diagram )
 *   dups         DEVICE [ 2 k ]     exact duplicates, ( dup, rep ) user-id pairs, 0 <= rep < dup, each dup once and
 *                                   no rep a dup; the first `agg_nb_dups` pairs are read ( `nb_dup_values` >= 2 x that ).
 *                                   Each duplicate's weight is tied `j agg_gap` under its representative's ( its j-th ),
 *                                   its cell left empty, its target carried by the representative; may be null.
 *                                   With `agg_margin` > 0 the solve also merges the near-coincident pairs it finds
 *                                   on its diagrams ( src/sdot/include/sdot/sdotplan/Aggregation.h ); `clusters` and
 *                                   `stats` report both kinds of merge
 *   options      HOST   [ 41 ]      the reference card's option vector, same order
 *                                   ( src/sdot/SdotPlanNd.py:197-201 ); the image geometry is in
 *                                   img_x0 .. img_ny as there, the Gaussians' count in img_nx; `lin_tol` 0 is the
 *                                   reference's 1e-6, a negative `lin_tol` a forcing term ( inexact Newton, no
 *                                   counterpart in the reference ) floored at its absolute value
 *   box          HOST   [ 4 ]       the domain { lo_x, lo_y, hi_x, hi_y }
 *   img          DEVICE [ ny * nx ] ( SDOT_AMD_IMAGE ) or [ img_nx ][ 4 ] ( SDOT_AMD_GAUSSIANS )
 *   weights      DEVICE [ n ]       output, `weights[ 0 ] == 0` ( the gauge )
 *   history      HOST   [ ( maxit + 1 ) * 8 ] output rows ( src/sdot/include/sdot/sdotplan/Report.h:25 )
 *   nb_steps     HOST   [ 1 ]       output, rows written
 *   stats        HOST   [ 31 ]      output ( src/sdot/include/sdot/sdotplan/Report.h:15-21 order )
 *   cell_masses  DEVICE [ n ]       output
 *   barycenters  DEVICE [ n ][ 2 ]  output ( the seed for an empty cell )
 *   cost         HOST   [ 1 ]       output, sum_i int_{cell_i} |x - p_i|^2 rho
 *   clusters     DEVICE [ n ] int64 output ( each seed its own index, a duplicate its representative's; may be null )
 */
int sdot_amd_solve_2d( void *stream, long long n, const double *positions, const double *nu, const double *w0,
                       const long long *dups, long long nb_dup_values, const double *options, const double *box,
                       int density_kind, double rho, const double *img,
                       double *weights, double *history, long long *nb_steps, double *stats,
                       double *cell_masses, double *barycenters, double *cost, long long *clusters );

/* THE SOLVE with the kernel's float chosen: `sdot_amd_solve_2d`'s arguments, then `precision`
 *   SDOT_AMD_FP64   the FP64 cells ( `sdot_amd_solve_2d` )
 *   SDOT_AMD_FP32   the walk and the cuts in float ( seed frame, planes rounded from double ), every vertex then re-solved
 *                   in double from its two planes ( kept within what the float vertex can be off by, else the float vertex
 *                   projected on one of them ), measures and facets in double; cells past the first pass by the FP64 kernel
 *   SDOT_AMD_MIXED  the FP32 kernel until one of its steps stagnates, then the FP64 kernel, the targets rescaled to the
 *                   domain's mass as the FP64 kernel measures it; `stats[ 24 ]` ( it_double ) the iteration of the switch,
 *                   -1 if none
 * The barycenters and the cost come from an FP64 diagram whatever `precision`.
 * Replaces the kernel variants of `sdotplan_solve_card_2d*`: `card_variant_for( 32 | 64, ... )`, the pair of them for
 * 'mixed' ( src/sdot/SdotPlanNd.py:772-783 ), `sdot::gpu2d::solve<V,VD>` ( src/sdot/include/sdot/gpu/Newton2D.cuh:860-866,
 * the switch 1701-1721 ).
 */
#define SDOT_AMD_FP64  0
#define SDOT_AMD_FP32  1
#define SDOT_AMD_MIXED 2
int sdot_amd_solve_2d_precision( void *stream, long long n, const double *positions, const double *nu, const double *w0,
                                 const long long *dups, long long nb_dup_values, const double *options, const double *box,
                                 int density_kind, double rho, const double *img,
                                 double *weights, double *history, long long *nb_steps, double *stats,
                                 double *cell_masses, double *barycenters, double *cost, long long *clusters, int precision );

/* THE SOLVE IN A POLYGON DOMAIN: `sdot_amd_solve_2d_precision`'s arguments, then the domain as in
 * `sdot_amd_measures_2d_domain` ( a convex counterclockwise polygon of 3 to 12 vertices; `box` its bounding box ). The targets
 * are rescaled to the polygon's mass of the density. Replaces the card's solve with `start_vertices` / `bnd_*` set
 * ( src/sdot/SdotPlanNd.py:864-890 with a `Polytope` target, src/sdot/distributions/Polytope.py ). */
int sdot_amd_solve_2d_domain( void *stream, long long n, const double *positions, const double *nu, const double *w0,
                              const long long *dups, long long nb_dup_values, const double *options, const double *box,
                              int density_kind, double rho, const double *img,
                              double *weights, double *history, long long *nb_steps, double *stats,
                              double *cell_masses, double *barycenters, double *cost, long long *clusters, int precision,
                              int nb_domain, const double *domain_xy );

/* ONE 3D DIAGRAM: the masses of the power cells of `( positions, weights )` in the box
 * `box = { lo_x, lo_y, lo_z, hi_x, hi_y, hi_z }` ( HOST ) against a constant density or a sum of Gaussians, and optionally
 * the facets `c_ij = int_facet rho / ( 2 |p_i - p_j| )` as triplets ( rows in user order, each row in its cell's face order ).
 * Replaces `power_diagram_measures_card_3d` -> `sdot::gpu3d::measures` ( src/sdot/include/sdot/gpu/Cell3D.cuh:1941-1952,
 * called from src/sdot/PowerDiagram_Bsp.py:271-311 ) and, for a sum of Gaussians, the CPU `power_diagram_measures`
 * ( src/sdot/PowerDiagram.py:375-391 ).
 *   positions    DEVICE [ n ][ 3 ]
 *   weights      DEVICE [ n ]       ( null: all zero, a Voronoi diagram )
 *   density_kind SDOT_AMD_CONSTANT ( `rho` ) or SDOT_AMD_GAUSSIANS ( `gauss` DEVICE [ ng ][ 5 ]: centre x, y, z, width, mass )
 *   masses       DEVICE [ n ]       ( output )
 *   facet_i/j/c  DEVICE [ facet_cap ] ( outputs; null: no facets )
 *   nb_facets    HOST   [ 1 ]       ( output, may be null )
 *   events       HOST   [ 5 ]       ( may be null: side tests, pruning vertex tests, new vertices, node visits, cuts )
 * A cell of more than 64 vertices or 64 faces is a failure ( kind 2, code 1, the seed in `value` ). */
int sdot_amd_measures_3d( void *stream, long long n, const double *positions, const double *weights, const double *box,
                          int density_kind, double rho, const double *gauss, int ng, double *masses, long long *facet_i,
                          long long *facet_j, double *facet_c, long long facet_cap, long long *nb_facets, unsigned long long *events );

/* ONE 3D DIAGRAM AGAINST A VOXEL IMAGE: `sdot_amd_measures_3d`'s arguments with the density a piecewise-constant image:
 * `img` DEVICE [ nz ][ ny ][ nx ] ( voxel `( i, j, k )` at `( k ny + j ) nx + i` ), `img_geo` HOST [ 6 ]: the origin x, y, z,
 * then the voxel's sides hx, hy, hz; the voxel `( i, j, k )` covers `[ x0 + i hx, x0 + ( i + 1 ) hx ] x ...`. Replaces the
 * reference's N-D `Image` against 3D cells ( src/sdot/distributions/Image.py, `Image.cxx:106-232`, run on the host there ). */
int sdot_amd_measures_3d_image( void *stream, long long n, const double *positions, const double *weights, const double *box,
                                const double *img, int nx, int ny, int nz, const double *img_geo, double *masses, long long *facet_i,
                                long long *facet_j, double *facet_c, long long facet_cap, long long *nb_facets );

/* THE MOMENTS OF 3D CELLS AGAINST A VOXEL IMAGE: `sdot_amd_moments_3d` with the image of `sdot_amd_measures_3d_image` as the
 * density ( the reference's `PowerDiagram.moments` with an `Image`, src/sdot/PowerDiagram.py:433-460, host code there ). */
int sdot_amd_moments_3d_image( void *stream, long long n, const double *positions, const double *weights, const double *box, const double *img,
                               int nx, int ny, int nz, const double *img_geo, double *masses, double *first, double *second );

/* THE 3D SOLVE AGAINST A VOXEL IMAGE: `sdot_amd_solve_3d`'s arguments with the image of `sdot_amd_measures_3d_image` as the
 * density ( no width continuation: it is the Gaussians' in 3D ). Replaces the reference's 3D solve with an `Image` target
 * ( src/sdot/SdotPlanNd.py:442-641, CPU there ). */
int sdot_amd_solve_3d_image( void *stream, long long n, const double *positions, const double *nu, const double *w0, const double *options,
                             const double *box, const double *img, int nx, int ny, int nz, const double *img_geo, double *weights,
                             double *history, long long *nb_steps, double *stats, double *cell_masses, double *barycenters, double *cost );

/* THE 3D SOLVE: the same damped Newton as `sdot_amd_solve_2d` on 3D cells, the step by trials ( the limits are a 2D step ),
 * a constant density or a sum of Gaussians ( the width continuation widens them ), no duplicates.
 * Replaces the 3D card solve `sdot::gpu2d::solve` with `SolveCard<3,...>` ( src/sdot/include/sdot/gpu/Newton2D.cuh:819,
 * 846-850 ) and, for a sum of Gaussians, the CPU `sdotplan_solve` ( src/sdot/include/sdot/sdotplan/Solve.h:111-114, called
 * from src/sdot/SdotPlanNd.py:568-626 ).
 *   positions    DEVICE [ n ][ 3 ]
 *   nu, w0       as `sdot_amd_solve_2d`
 *   options      HOST   [ 41 ]      the reference card's option vector ( the step, the image and duplicate entries unused )
 *   box          HOST   [ 6 ]       the domain { lo_x, lo_y, lo_z, hi_x, hi_y, hi_z }
 *   gauss        DEVICE [ ng ][ 5 ] ( SDOT_AMD_GAUSSIANS ); may be null for a constant density
 *   weights, history, nb_steps, stats, cell_masses, cost   as `sdot_amd_solve_2d`
 *   barycenters  DEVICE [ n ][ 3 ]  output ( the seed for an empty cell ) */
int sdot_amd_solve_3d( void *stream, long long n, const double *positions, const double *nu, const double *w0,
                       const double *options, const double *box, int density_kind, double rho, const double *gauss, int ng,
                       double *weights, double *history, long long *nb_steps, double *stats,
                       double *cell_masses, double *barycenters, double *cost );

/* THE BATCH: `k` independent solves at once ( the reference solves one problem per call; this entry point has no
 * counterpart there and serves the batched configuration: many instances of uneven size ). Instance q owns the seeds
 * `offsets[ q ] .. offsets[ q + 1 ]` of the concatenated arrays and has its own box and density; all share `options`
 * ( the 41 doubles of `sdot_amd_solve_2d`, minus the image geometry, which is per instance ). The instances run on
 * `nb_streams` HIP streams ( 1 .. 32 ), largest first, each stream driven by its own host thread with its own device
 * buffers: an instance's results do not depend on what it ran beside. `stream` is waited on before the start; every
 * output is complete when the call returns. No duplicates are merged ( `aggregate` is per-instance work done by
 * `sdot_amd_solve_2d` ).
 *   offsets      HOST   [ k + 1 ]  int64, offsets[ 0 ] == 0, increasing by at least 2
 *   positions    DEVICE [ N ][ 2 ], nu DEVICE [ N ], w0 DEVICE [ N ] or null   ( N = offsets[ k ] )
 *   boxes        HOST   [ k ][ 4 ]  each instance's domain
 *   rho          HOST   [ k ]       the constant density ( SDOT_AMD_CONSTANT ); may be null otherwise
 *   imgs         HOST   [ k ]       device pointers: each instance's image [ ny * nx ] ( SDOT_AMD_IMAGE ) or Gaussians
 *                                   [ ng ][ 4 ] ( SDOT_AMD_GAUSSIANS ); may be null for a constant density
 *   img_geo      HOST   [ k ][ 6 ]  x0, y0, hx, hy, nx, ny ( an image ) or ng in nx ( Gaussians ); may be null for a constant
 *   weights, cell_masses DEVICE [ N ], barycenters DEVICE [ N ][ 2 ]  outputs, per instance as `sdot_amd_solve_2d`'s
 *   stats        HOST   [ k ][ 31 ], costs HOST [ k ]  outputs
 *   errors       HOST   [ k ][ 3 ]  output: per instance 0, or the error kind, code and value that stopped it
 *   times        HOST   [ k ][ 2 ]  output ( may be null ): each instance's start and end, seconds from the call's start
 * Returns 0, or an error of the batch itself ( an instance's failure is only in `errors` ). */
int sdot_amd_solve_2d_batch( void *stream, int nb_streams, long long k, const long long *offsets, const double *positions,
                             const double *nu, const double *w0, const double *options, const double *boxes, int density_kind,
                             const double *rho, const double *const *imgs, const double *img_geo,
                             double *weights, double *cell_masses, double *barycenters, double *stats, double *costs,
                             long long *errors, double *times );

/* THE BATCH with the kernel's float chosen: `sdot_amd_solve_2d_batch`'s arguments, then `precision` as `sdot_amd_solve_2d_precision`'s
 * ( every instance solved as `sdot_amd_solve_2d_precision` would solve it alone ). */
int sdot_amd_solve_2d_batch_precision( void *stream, int nb_streams, long long k, const long long *offsets, const double *positions,
                                       const double *nu, const double *w0, const double *options, const double *boxes, int density_kind,
                                       const double *rho, const double *const *imgs, const double *img_geo,
                                       double *weights, double *cell_masses, double *barycenters, double *stats, double *costs,
                                       long long *errors, double *times, int precision );

/* THE SERVER: continuous batching of solves ( no counterpart in the reference, which solves one problem per call; for
 * callers that serve a stream of requests ). `nb_slots` slots ( 1 .. 32 ), each a HIP stream driven by its own host
 * thread with its own device buffers, kept between its requests. A free slot takes the oldest waiting request whose
 * footprint ( `sdot_amd_solve_bytes` ) fits `memory_budget` bytes beside what the other slots hold; requests that arrive
 * while others run are admitted as slots and memory free up, and their work overlaps on the device. A request that fits on
 * no slot runs once every other slot idles and has freed its buffers. A request's result does not depend on what ran beside
 * it. Returns a handle, or 0 on error ( `sdot_amd_last_error` ). */
long long sdot_amd_server_start( int nb_slots, double memory_budget );

/* A request ( 2D, FP64 cells ): the arguments of `sdot_amd_solve_2d` minus the stream, the duplicates, the history and the clusters. Every
 * array must stay alive, and the outputs untouched, until `sdot_amd_server_wait` returns for its ticket. */
int sdot_amd_server_submit( long long server, long long n, const double *positions, const double *nu, const double *w0,
                            const double *options, const double *box, int density_kind, double rho, const double *img,
                            double *weights, double *cell_masses, double *barycenters, double *stats, double *cost,
                            long long *ticket );

/* A request of either dimension and kernel float: `dim` 2 or 3, `precision` as `sdot_amd_solve_2d_precision` ( 3D: FP64 or
 * MIXED, both FP64 cells ), then the arguments of `sdot_amd_server_submit` with the density as `sdot_amd_solve_2d`'s ( 2D ) or
 * `sdot_amd_solve_3d`'s ( 3D: a constant, or `nb_density` Gaussians of 5 doubles in `density` ); a 3D request steps by
 * trials. Its admission footprint is `sdot_amd_solve_bytes` of the request. */
int sdot_amd_server_submit_nd( long long server, int dim, int precision, long long n, const double *positions, const double *nu,
                               const double *w0, const double *options, const double *box, int density_kind, double rho,
                               const double *density, int nb_density, double *weights, double *cell_masses, double *barycenters,
                               double *stats, double *cost, long long *ticket );

/* A 3D REQUEST AGAINST A VOXEL IMAGE: `sdot_amd_server_submit_nd`'s 3D request with the image of `sdot_amd_measures_3d_image`
 * ( `img` DEVICE [ nz ][ ny ][ nx ], `img_geo` HOST [ 6 ] ) as the density; FP64 cells. No counterpart in the reference ( it has
 * no server ): the served 3D share's density. */
int sdot_amd_server_submit_3d_image( long long server, long long n, const double *positions, const double *nu, const double *w0,
                                     const double *options, const double *box, const double *img, int nx, int ny, int nz,
                                     const double *img_geo, double *weights, double *cell_masses, double *barycenters, double *stats,
                                     double *cost, long long *ticket );

/* What the server's slots held: `info` HOST [ 2 ]: the largest sum of the slots' device bytes seen at the end of a request,
 * and the requests done. */
int sdot_amd_server_info( long long server, double *info );

/* Blocks until the request is done. `error` HOST [ 3 ]: 0, or the kind, code and value that stopped it; `times` HOST [ 3 ]:
 * its arrival, start and end, seconds from the server's start ( may be null ). */
int sdot_amd_server_wait( long long server, long long ticket, long long *error, double *times );

/* Stops the server once its running requests are done ( waiting requests are dropped ) and frees its slots. */
int sdot_amd_server_stop( long long server );

/* WHAT A SOLVE TAKES FROM THE DEVICE, part by part, before it starts ( the model the server admits requests by ).
 * Replaces `card_solve_bytes` ( src/sdot/CardMemory.py:47-103, the reference's model of its card's takes ) for this library's
 * own buffers. `parts` HOST [ SDOT_AMD_NB_MEMORY_PARTS ] ( output, bytes ): 0 the tree and its build, 1 the two diagram
 * slots ( masses, Laplacian rows, overflow lists ), 2 the cells' edges for the step by the limits, 3 the FP32 kernel's
 * polygons ( 2D ) or the polytopes of a 3D sum of Gaussians, 4 the third pass's vertex slots, 5 Newton's vectors, 6 the
 * moments, 7 the linear solver ( multigrid levels and vectors ), 8 the density's copies ( blurred images, the pyramid of
 * minima ), 9 reductions and small state. `options` the 41 doubles of `sdot_amd_solve_2d` ( may be null: defaults ),
 * `precision` an SDOT_AMD_FP64 / FP32 / MIXED. Held to what a solve really takes by the tests ( within 10 % above it on
 * generic clouds ); a diagram whose rows overflow their first room takes more. */
#define SDOT_AMD_NB_MEMORY_PARTS 10
int sdot_amd_solve_bytes( long long n, int dim, int density_kind, const double *options, int precision, double *parts );

/* THE TREE: the BSP tree of a 2D cloud ( median cut on the longest axis, at most `leaf` seeds per leaf,
 * perfect binary tree of depth ceil( log2( n / leaf ) ) + 1 in preorder ).
 * Replaces `bsp_build_card` -> `sdot::gpu2d::build_tree`
 * ( src/sdot/include/sdot/gpu/Bsp2D.cuh:367, called from src/sdot/AaBsp.py:733-765 ).
 *   positions     DEVICE [ n ][ 2 ]
 *   seed_indices  DEVICE [ n ] int64   rank -> user index
 *   rank_of       DEVICE [ n ] int64   user index -> rank
 *   node_begin / node_end / node_left / node_right  DEVICE [ N ] int64 ( -1: no child )
 *   node_box      DEVICE [ N ][ 2 ][ 2 ] ( lo then hi )
 * N = 2^depth - 1, see `sdot_amd_tree_nodes`.
 */
long long sdot_amd_tree_nodes( long long n, int leaf );
int sdot_amd_build_tree_2d( void *stream, long long n, const double *positions, int leaf,
                            long long *seed_indices, long long *rank_of, long long *node_begin, long long *node_end,
                            long long *node_left, long long *node_right, double *node_box );

/* THE MAJORANTS: for every node of a tree built as above, `( wa, wb )` with `w_k <= wa . p_k + wb` on
 * its seeds ( affine least-squares fit kept when it tightens the spread, else constant ).
 * Replaces `bsp_refresh_majorants_card_2d` -> `sdot::gpu2d::refresh_majorants`
 * ( src/sdot/include/sdot/gpu/Majorant2D.cuh:276, called from src/sdot/AaBsp.py:356-372 ).
 *   node_box / node_begin / node_end  DEVICE, as built
 *   sorted_positions DEVICE [ n ][ 2 ], sorted_weights DEVICE [ n ] ( in rank order )
 *   node_wa DEVICE [ N ][ 2 ], node_wb DEVICE [ N ] ( outputs )
 */
int sdot_amd_refresh_majorants_2d( void *stream, long long n, long long nb_nodes, const double *node_box,
                                   const long long *node_begin, const long long *node_end,
                                   const double *sorted_positions, const double *sorted_weights,
                                   double *node_wa, double *node_wb );

/* THE LINEAR SOLVER ALONE: `L d_j = rhs_j` for `k` right-hand sides of zero sum, `L` the Laplacian
 * given as CSR of its off-diagonal entries ( `val > 0` are `c_ij`, the matrix holds `-c_ij` ) plus its
 * diagonal; the solution in the gauge `d[ 0 ] = 0`.
 * Replaces `sdotplan_card_linear_2d` -> `sdot::gpu2d::linear_solves`
 * ( src/sdot/include/sdot/gpu/Newton2D.cuh:1916, called from src/sdot/SdotPlanNd.py:204-236 ).
 *   row DEVICE [ n + 1 ] int64, col DEVICE [ nnz ] int64, val DEVICE [ nnz ], dia DEVICE [ n ]
 *   rhs DEVICE [ k ][ n ], options HOST [ 10 ] ( n, k, method, tol, ... as the reference; tol used )
 *   solution DEVICE [ k ][ n ] ( output ), iterations HOST [ k ] ( output, -1: failed )
 */
int sdot_amd_linear_solves_2d( void *stream, long long n, const long long *row, const long long *col, const double *val,
                               const double *dia, const double *rhs, long long k, const double *options,
                               double *solution, double *iterations );

#ifdef __cplusplus
}
#endif

#endif
