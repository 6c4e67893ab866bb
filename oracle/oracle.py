"""Python access to the oracle ( `oracle/sdot_oracle.cpp` ) -- TEST INFRASTRUCTURE.

Only the tests, `__graft_entry__.smoke()` and the CPU-baseline leg of `bench.py` load this module;
the product ( `sdot-ffi_amd/` ) never does. Every function takes and returns numpy arrays, in the
reference's user order ( `positions[ n, 2 ]`, `weights[ n ]` ), and image values as
`img[ ny, nx ]` ( row `j` = pixel row along y ).
"""

import contextlib
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname( os.path.abspath( __file__ ) )
_LIB_PATH = os.path.join( _HERE, "_build", "liboracle.so" )
_lib = None

_D = ctypes.POINTER( ctypes.c_double )
_L = ctypes.POINTER( ctypes.c_longlong )


def build():
    """compiles the oracle ( `make -C oracle` )"""
    subprocess.run( [ "make", "-s", "-C", _HERE ], check = True )


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists( _LIB_PATH ):
            build()
        _lib = ctypes.CDLL( _LIB_PATH )
        _lib.oracle_build_tree.restype = ctypes.c_int
        _lib.oracle_measures.restype = ctypes.c_longlong
        _lib.oracle_cell.restype = ctypes.c_int
        _lib.oracle_cg.restype = ctypes.c_int
        _lib.oracle_mg.restype = ctypes.c_int
        _lib.oracle_solve.restype = ctypes.c_int
        _lib.oracle_max_depth_for.restype = ctypes.c_int
        _lib.oracle_get_threads.restype = ctypes.c_int
        _lib.oracle_build_tree_3d.restype = ctypes.c_int
        _lib.oracle_measures_3d.restype = ctypes.c_longlong
        _lib.oracle_cell_3d.restype = ctypes.c_int
        _lib.oracle_solve_3d.restype = ctypes.c_int
    return _lib


def _d( a ):
    return None if a is None else a.ctypes.data_as( _D )


def _l( a ):
    return None if a is None else a.ctypes.data_as( _L )


def _f64( a ):
    return np.ascontiguousarray( np.asarray( a, dtype = np.float64 ) )


def set_threads( nt ):
    lib().oracle_set_threads( int( nt ) )


def get_threads():
    return int( lib().oracle_get_threads() )


def max_depth_for( n, leaf = 10 ):
    return int( lib().oracle_max_depth_for( ctypes.c_longlong( n ), ctypes.c_longlong( leaf ) ) )


def build_tree( pos, w = None, leaf = 10 ):
    """the tree: `dict( order, beg, end, box[ N, 2, 2 ], wa[ N, 2 ], wb[ N ], depth )`, preorder"""
    pos = _f64( pos )
    n = pos.shape[ 0 ]
    depth = max_depth_for( n, leaf )
    N = 2 ** depth - 1
    order = np.zeros( n, np.int64 )
    beg = np.zeros( N, np.int64 )
    end = np.zeros( N, np.int64 )
    box = np.zeros( ( N, 2, 2 ) )
    wa = np.zeros( ( N, 2 ) )
    wb = np.zeros( N )
    ww = None if w is None else _f64( w )
    lib().oracle_build_tree( ctypes.c_longlong( n ), _d( pos ), _d( ww ), ctypes.c_longlong( leaf ), _l( order ), _l( beg ), _l( end ),
                             _d( box ), _d( wa ), _d( wb ) )
    return dict( order = order, beg = beg, end = end, box = box, wa = wa, wb = wb, depth = depth )


def weight_majorant( pos, w ):
    pos = _f64( pos )
    w = _f64( w )
    wa = np.zeros( 2 )
    wb = np.zeros( 1 )
    lib().oracle_weight_majorant( ctypes.c_longlong( len( w ) ), _d( pos ), _d( w ), _d( wa ), _d( wb ) )
    return wa, float( wb[ 0 ] )


def _density_args( density ):
    """`density`: `None` / a float ( constant ), `dict( values = img[ ny, nx ], x0, y0, hx, hy )` ( an image ), or
    `dict( gaussians = [ ng, 4 ] )` ( centre x, y, sigma, mass per row: a sum of isotropic Gaussians )"""
    if density is None or np.isscalar( density ):
        return 0, 1.0 if density is None else float( density ), 0, 0, None, None, None
    if "gaussians" in density:
        g = _f64( density[ "gaussians" ] ).reshape( -1, 4 )
        return 2, 0.0, g.shape[ 0 ], 0, None, g, ( g, )
    img = _f64( density[ "values" ] )
    geo = _f64( [ density[ "x0" ], density[ "y0" ], density[ "hx" ], density[ "hy" ] ] )
    return 1, 0.0, img.shape[ 1 ], img.shape[ 0 ], geo, img, ( img, geo )


def _support( lo, hi, density ):
    """the domain of the calls: the box, cut for an image density by the block of its grid, as the reference cuts the box by
    the distribution's bounding half-spaces ( src/sdot/PowerDiagram.py:158 ). The image is zero off its grid ( its cells
    are clipped by the pixels ): what the cut changes is where the limits' frozen cells start from"""
    lo, hi = _f64( lo ).copy(), _f64( hi ).copy()
    if isinstance( density, dict ) and ( "values" in density or "voxels" in density ):
        v = _f64( density[ "values" ] if "values" in density else density[ "voxels" ] )
        if "values" in density:                          # ( `values[ ny, nx ]` )
            o, h, nb = [ density[ "x0" ], density[ "y0" ] ], [ density[ "hx" ], density[ "hy" ] ], [ v.shape[ 1 ], v.shape[ 0 ] ]
        else:                                            # ( `voxels[ nx, ny, nz ]` )
            o = [ density[ "x0" ], density[ "y0" ], density[ "z0" ] ]
            h, nb = [ density[ "hx" ], density[ "hy" ], density[ "hz" ] ], list( v.shape )
        for a in range( len( lo ) ):
            lo[ a ] = max( lo[ a ], o[ a ] )
            hi[ a ] = min( hi[ a ], o[ a ] + nb[ a ] * h[ a ] )
        if np.any( hi <= lo ):
            raise ValueError( "oracle: the image's grid and the box do not meet" )
    return lo, hi


@contextlib.contextmanager
def polygon_domain( xy ):
    """the calls inside take the convex polygon `xy` ( `[ nb, 2 ]`, counterclockwise ) as their domain: each cell starts from
    it, edge `i` the cut `-5 - i` ( `None`: the box )"""
    if xy is None:
        yield
        return
    v = np.ascontiguousarray( _f64( xy ).reshape( -1, 2 ) )
    lib().oracle_set_domain( ctypes.c_int( len( v ) ), _d( v ) )
    try:
        yield
    finally:
        lib().oracle_set_domain( ctypes.c_int( 0 ), None )


def measures( pos, w = None, lo = ( 0, 0 ), hi = ( 1, 1 ), density = None, facets = False, brute_force = False, events = False, domain = None ):
    """masses `[ n ]` ( and if asked the facets `( fi, fj, fc )` in row order, and the event counts ); `domain`: a convex polygon
    ( counterclockwise vertices ) in place of the box"""
    with polygon_domain( domain ):
        return _measures( pos, w, lo, hi, density, facets, brute_force, events )


def _measures( pos, w, lo, hi, density, facets, brute_force, events ):
    pos = _f64( pos )
    n = pos.shape[ 0 ]
    ww = None if w is None else _f64( w )
    lo, hi = _support( lo, hi, density )
    kind, value, nx, ny, geo, img, keep = _density_args( density )
    masses = np.zeros( n )
    cap = 8 * n + 64 if facets else 0
    fi = np.zeros( cap, np.int64 ) if facets else None
    fj = np.zeros( cap, np.int64 ) if facets else None
    fc = np.zeros( cap ) if facets else None
    ev = np.zeros( 5, np.int64 )
    nf = lib().oracle_measures( ctypes.c_longlong( n ), _d( pos ), _d( ww ), _d( lo ), _d( hi ), ctypes.c_int( kind ), ctypes.c_double( value ),
                                ctypes.c_int( nx ), ctypes.c_int( ny ), _d( geo ), _d( img ), _d( masses ), _l( fi ), _l( fj ), _d( fc ),
                                ctypes.c_longlong( cap ), _l( ev ), ctypes.c_int( int( brute_force ) ) )
    res = [ masses ]
    if facets:
        if nf < 0:
            raise RuntimeError( "oracle: facet capacity exceeded" )
        res.append( ( fi[ :nf ], fj[ :nf ], fc[ :nf ] ) )
    if events:
        res.append( dict( zip( ( "side_tests", "prune_vertex_tests", "new_vertices", "node_visits", "cuts" ), ev.tolist() ) ) )
    return res[ 0 ] if len( res ) == 1 else tuple( res )


def cell( pos, i, w = None, lo = ( 0, 0 ), hi = ( 1, 1 ), brute_force = False, cap = 4096, domain = None ):
    """the vertices `[ nv, 2 ]` of the cell of seed `i` and the ids of its edges ( -1 - f: side f of the box, -5 - e: edge e
    of the polygon `domain` )"""
    with polygon_domain( domain ):
        return _cell( pos, i, w, lo, hi, brute_force, cap )


def _cell( pos, i, w, lo, hi, brute_force, cap ):
    pos = _f64( pos )
    n = pos.shape[ 0 ]
    ww = None if w is None else _f64( w )
    xy = np.zeros( ( cap, 2 ) )
    ids = np.zeros( cap, np.int64 )
    nv = lib().oracle_cell( ctypes.c_longlong( n ), _d( pos ), _d( ww ), _d( _f64( lo ) ), _d( _f64( hi ) ), ctypes.c_longlong( i ),
                            _d( xy ), _l( ids ), ctypes.c_int( cap ), ctypes.c_int( int( brute_force ) ) )
    if nv < 0:
        raise RuntimeError( "oracle: cell capacity exceeded" )
    return xy[ :nv ].copy(), ids[ :nv ].copy()


def along( pos, w, d, t = 0.0, eps = 0.0, lo = ( 0, 0 ), hi = ( 1, 1 ), density = None ):
    """the step's frozen cells of the diagram of ( pos, w ) along d: dict( poly[ n, 3 ], ok[ n ], frozen_mass[ n ] at t,
    forward[ n ] at floor eps, backward[ n ] from the trial t )"""
    pos = _f64( pos )
    n = pos.shape[ 0 ]
    lo, hi = _support( lo, hi, density )
    kind, value, nx, ny, geo, img, keep = _density_args( density )
    poly = np.zeros( ( n, 3 ) )
    ok = np.zeros( n, np.int32 )
    fm, fw, bw = np.zeros( n ), np.zeros( n ), np.zeros( n )
    lib().oracle_along( ctypes.c_longlong( n ), _d( pos ), _d( _f64( w ) ), _d( _f64( d ) ), _d( _f64( lo ) ), _d( _f64( hi ) ), ctypes.c_int( kind ),
                        ctypes.c_double( value ), ctypes.c_int( nx ), ctypes.c_int( ny ), _d( geo ), _d( img ), ctypes.c_double( t ),
                        ctypes.c_double( eps ), _d( poly ), ok.ctypes.data_as( ctypes.c_void_p ), _d( fm ), _d( fw ), _d( bw ) )
    return dict( poly = poly, ok = ok.astype( bool ), frozen_mass = fm, forward = fw, backward = bw )


def blur( values, s, hx, hy ):
    """the image `values[ ny, nx ]` at width s ( the continuation's filter )"""
    v = _f64( values )
    ny, nx = v.shape
    out = np.zeros_like( v )
    lib().oracle_blur( ctypes.c_int( nx ), ctypes.c_int( ny ), ctypes.c_double( hx ), ctypes.c_double( hy ), _d( v ), ctypes.c_double( s ), _d( out ) )
    return out


def cg( n, fi, fj, fc, b, tol = 1e-12, maxit = 20000 ):
    """Jacobi-PCG on the Laplacian of the facets; returns `( d, iterations )`"""
    fi = np.ascontiguousarray( fi, dtype = np.int64 )
    fj = np.ascontiguousarray( fj, dtype = np.int64 )
    fc = _f64( fc )
    b = _f64( b )
    d = np.zeros( n )
    its = lib().oracle_cg( ctypes.c_longlong( n ), ctypes.c_longlong( len( fi ) ), _l( fi ), _l( fj ), _d( fc ), _d( b ), ctypes.c_double( tol ),
                           ctypes.c_int( maxit ), _d( d ) )
    return d, its


def mg( n, fi, fj, fc, b, rank_of = None, nu = 3, tol = 1e-12, maxit = 20000 ):
    """the multigrid-preconditioned CG on the Laplacian of the facets ( first level aggregated by `rank_of >> 3` );
    returns `( d, iterations, levels )`"""
    fi = np.ascontiguousarray( fi, dtype = np.int64 )
    fj = np.ascontiguousarray( fj, dtype = np.int64 )
    fc = _f64( fc )
    b = _f64( b )
    ro = None if rank_of is None else np.ascontiguousarray( rank_of, dtype = np.int64 )
    d = np.zeros( n )
    lv = np.zeros( 1, dtype = np.int32 )
    its = lib().oracle_mg( ctypes.c_longlong( n ), ctypes.c_longlong( len( fi ) ), _l( fi ), _l( fj ), _d( fc ), _d( b ),
                           None if ro is None else _l( ro ), ctypes.c_int( nu ), ctypes.c_double( tol ), ctypes.c_int( maxit ), _d( d ),
                           lv.ctypes.data_as( ctypes.POINTER( ctypes.c_int ) ) )
    return d, its, int( lv[ 0 ] )


STATS = [ "status", "residual", "residual0", "nb_iter", "nb_diag", "nb_backtracks", "t_majorant", "t_diag", "t_asm", "t_lin", "t_lim", "eps",
          "domain_mass", "nb_overflowed", "nb_cell_lim", "nb_limit_rounds", "lin_nb_hierarchies", "lin_nb_iter", "lin_worst", "start", "t_total",
          "nb_continuation_steps", "min_start_mass", "it_switch", "it_double", "scratch_bytes",
          "nb_clusters", "nb_aggregated", "nb_duplicates", "residual_full", "nb_polish" ]
HISTORY = [ "step", "t", "residual_l2", "min_measure", "max_abs_residual", "nb_diag", "nb_evals", "s" ]


def solve( pos, nu = None, w0 = None, lo = ( 0, 0 ), hi = ( 1, 1 ), density = None, tol = 1e-9, tol_rel = 0.0, maxit = 100, max_backtracks = 60,
           t_min = 1e-10, mult_ok = 4.0, residual = 1, power = 0.5, switch = 2.0, lin_tol = 1e-12, lin_maxit = 20000, keep_start = False,
           start_box = None, step = "trials", factor = 0.9, continuation = "never", conv_threshold = 1e-2, conv_s0 = 0.0,
           conv_ratio = 2 ** 0.5, conv_min = 0.0, dups = None, gap = 0.0, margin = 0.0, linear = "cg", domain = None ):
    """the damped-Newton solve ( step "trials" or "limits"; continuation "never", "auto" or "always" ); `dups`: the exact
    duplicates, `( dup, rep )` rows of user ids, whose weights are tied at `w_rep - k gap`; `margin`: the aggregation of
    near-coincident pairs ( 0: off ); `linear`: the Newton systems by Jacobi-PCG ( "cg" ) or the multigrid ( "mg" ); `domain`: a convex
    polygon ( counterclockwise vertices; `lo`, `hi` its bounding box ) in place of the box; returns a dict ( weights, masses, barycenters, cost, stats, history, clusters )"""
    pos = _f64( pos )
    n = pos.shape[ 0 ]
    nu = np.full( n, 1.0 / n ) if nu is None else _f64( nu )
    w0 = None if w0 is None else _f64( w0 )
    lo, hi = _support( lo, hi, density )
    kind, value, nx, ny, geo, img, keep = _density_args( density )
    opts = _f64( [ tol, tol_rel, maxit, max_backtracks, t_min, mult_ok, residual, power, switch, lin_tol, lin_maxit, int( keep_start ),
                   { "trials": 0, "limits": 1 }[ step ], factor, { "never": 0, "auto": 1, "always": 2 }[ continuation ], conv_threshold,
                   conv_s0, conv_ratio, conv_min, gap, 0 if dups is None else len( dups ), margin, { "cg": 0, "mg": 1 }[ linear ] ] )
    dp = None if dups is None or len( dups ) == 0 else np.ascontiguousarray( np.asarray( dups, dtype = np.int64 ).reshape( -1, 2 ) )
    sb = _f64( [ 0, 0, 0, 0, 0 ] if start_box is None else [ 1, *start_box[ 0 ], *start_box[ 1 ] ] )
    weights = np.zeros( n )
    stats = np.zeros( len( STATS ) )
    hist = np.zeros( ( maxit + 1, len( HISTORY ) ) )
    masses = np.zeros( n )
    bary = np.zeros( ( n, 2 ) )
    cost = np.zeros( 1 )
    clusters = np.zeros( n, dtype = np.int64 )
    with polygon_domain( domain ):
        nb = lib().oracle_solve( ctypes.c_longlong( n ), _d( pos ), _d( nu ), _d( w0 ), _d( _f64( lo ) ), _d( _f64( hi ) ), ctypes.c_int( kind ),
                                 ctypes.c_double( value ), ctypes.c_int( nx ), ctypes.c_int( ny ), _d( geo ), _d( img ), _d( opts ), _d( sb ),
                                 None if dp is None else dp.ctypes.data_as( ctypes.POINTER( ctypes.c_longlong ) ), _d( weights ), _d( stats ), _d( hist ), _d( masses ), _d( bary ), _d( cost ),
                                 clusters.ctypes.data_as( ctypes.POINTER( ctypes.c_longlong ) ) )
    return dict( weights = weights, masses = masses, barycenters = bary, cost = float( cost[ 0 ] ),
                 stats = dict( zip( STATS, stats.tolist() ) ), history = hist[ :nb ].copy(), clusters = clusters )


# ---- 3D ------------------------------------------------------------------------------------------------------------

def _density3_args( density ):
    """`density`: `None` / a float ( constant ), `dict( gaussians = [ ng, 5 ] )` ( centre x, y, z, sigma, mass per row ), or
    `dict( voxels = v[ nx, ny, nz ], x0, y0, z0, hx, hy, hz )` ( a voxel image, first index along x )"""
    if density is None or np.isscalar( density ):
        return 0, 1.0 if density is None else float( density ), 0, None
    if "voxels" in density:
        return 1, 0.0, 0, None
    g = _f64( density[ "gaussians" ] ).reshape( -1, 5 )
    return 2, 0.0, g.shape[ 0 ], g


@contextlib.contextmanager
def _image3( density ):
    """the voxel image of `density` ( if it is one ) set for the calls inside"""
    if density is None or np.isscalar( density ) or "voxels" not in density:
        yield
        return
    v = _f64( density[ "voxels" ] )
    flat = np.ascontiguousarray( v.transpose( 2, 1, 0 ) ).reshape( -1 )        # ( `( k ny + j ) nx + i` )
    geo = _f64( [ density.get( "x0", 0.0 ), density.get( "y0", 0.0 ), density.get( "z0", 0.0 ), density[ "hx" ], density[ "hy" ], density[ "hz" ] ] )
    lib().oracle_set_image3( ctypes.c_int( v.shape[ 0 ] ), ctypes.c_int( v.shape[ 1 ] ), ctypes.c_int( v.shape[ 2 ] ), _d( geo ), _d( flat ) )
    try:
        yield
    finally:
        lib().oracle_set_image3( ctypes.c_int( 0 ), ctypes.c_int( 0 ), ctypes.c_int( 0 ), None, None )


def rules():
    """the quadrature rules of the 3D integrals: `tet[ 14, 5 ]` ( barycentric coordinates, weight ), `tri[ 7, 4 ]`"""
    tet = np.zeros( ( 14, 5 ) )
    tri = np.zeros( ( 7, 4 ) )
    lib().oracle_rules( _d( tet ), _d( tri ) )
    return tet, tri


def build_tree_3d( pos, w = None, leaf = 10 ):
    """dict( order, beg, end, box[ N, 6 ], wa[ N, 3 ], wb[ N ], depth ) of the 3D tree"""
    pos = _f64( pos )
    n = pos.shape[ 0 ]
    depth = max_depth_for( n, leaf )
    N = ( 1 << depth ) - 1
    order = np.zeros( n, np.int64 )
    beg = np.zeros( N, np.int64 )
    end = np.zeros( N, np.int64 )
    box = np.zeros( ( N, 6 ) )
    wa = np.zeros( ( N, 3 ) )
    wb = np.zeros( N )
    ww = None if w is None else _f64( w )
    d = lib().oracle_build_tree_3d( ctypes.c_longlong( n ), _d( pos ), _d( ww ), ctypes.c_longlong( leaf ), _l( order ), _l( beg ), _l( end ),
                                    _d( box ), _d( wa ), _d( wb ) )
    return dict( order = order, beg = beg, end = end, box = box, wa = wa, wb = wb, depth = d )


def measures_3d( pos, w = None, lo = ( 0, 0, 0 ), hi = ( 1, 1, 1 ), density = None, facets = False, events = False ):
    """masses `[ n ]` of the 3D diagram ( and if asked the facets `( fi, fj, fc )` in row order, and the event counts with
    `broken`, the cells whose cut could not be closed )"""
    pos = _f64( pos )
    n = pos.shape[ 0 ]
    ww = np.zeros( n ) if w is None else _f64( w )
    lo, hi = _support( lo, hi, density )
    kind, value, ng, g = _density3_args( density )
    masses = np.zeros( n )
    cap = 24 * n + 64 if facets else 0
    fi = np.zeros( cap, np.int64 )
    fj = np.zeros( cap, np.int64 )
    fc = np.zeros( cap )
    info = np.zeros( 6, np.int64 )
    with _image3( density ):
        nf = lib().oracle_measures_3d( ctypes.c_longlong( n ), _d( pos ), _d( ww ), _d( _f64( lo ) ), _d( _f64( hi ) ), ctypes.c_int( kind ),
                                       ctypes.c_double( value ), ctypes.c_int( ng ), _d( g ), _d( masses ), _l( fi ), _l( fj ), _d( fc ),
                                       ctypes.c_longlong( cap ), _l( info ) )
    res = [ masses ]
    if facets:
        if nf > cap:
            raise RuntimeError( "oracle: facet capacity exceeded" )
        res.append( ( fi[ :nf ], fj[ :nf ], fc[ :nf ] ) )
    if events:
        res.append( dict( zip( ( "broken", "side_tests", "prune_vertex_tests", "new_vertices", "node_visits", "cuts" ), info.tolist() ) ) )
    return res[ 0 ] if len( res ) == 1 else tuple( res )


def cell_3d( pos, i, w = None, lo = ( 0, 0, 0 ), hi = ( 1, 1, 1 ), brute_force = False, cap = 512 ):
    """the vertices `[ nv, 3 ]` of the 3D cell of seed `i`, the ids of its faces ( user index, or `-1 - f` for face `f` of
    the box: x lo, x hi, y lo, y hi, z lo, z hi ) and their areas"""
    pos = _f64( pos )
    n = pos.shape[ 0 ]
    ww = np.zeros( n ) if w is None else _f64( w )
    xyz = np.zeros( ( cap, 3 ) )
    ids = np.zeros( cap, np.int64 )
    areas = np.zeros( cap )
    nf = ctypes.c_int( 0 )
    nv = lib().oracle_cell_3d( ctypes.c_longlong( n ), _d( pos ), _d( ww ), _d( _f64( lo ) ), _d( _f64( hi ) ), ctypes.c_longlong( i ),
                               ctypes.c_int( int( brute_force ) ), _d( xyz ), ctypes.c_int( cap ), _l( ids ), _d( areas ), ctypes.byref( nf ) )
    if nv < 0:
        raise RuntimeError( "oracle: the cell's topology broke" )
    if nv > cap or nf.value > cap:
        raise RuntimeError( "oracle: cell capacity exceeded" )
    return xyz[ :nv ].copy(), ids[ :nf.value ].copy(), areas[ :nf.value ].copy()


def solve_3d( pos, nu = None, w0 = None, lo = ( 0, 0, 0 ), hi = ( 1, 1, 1 ), density = None, tol = 1e-9, tol_rel = 0.0, maxit = 100,
              max_backtracks = 60, t_min = 1e-10, mult_ok = 4.0, residual = 1, power = 0.5, switch = 2.0, lin_tol = 1e-12, lin_maxit = 20000,
              keep_start = False, continuation = "never", conv_threshold = 1e-2, conv_s0 = 0.0, conv_ratio = 2 ** 0.5, conv_min = 0.0,
              linear = "cg" ):
    """the 3D damped-Newton solve ( step by trials ); `linear` as `solve`; returns a dict ( weights, masses, barycenters, cost, stats, history )"""
    pos = _f64( pos )
    n = pos.shape[ 0 ]
    nu = np.full( n, 1.0 / n ) if nu is None else _f64( nu )
    w0 = None if w0 is None else _f64( w0 )
    lo, hi = _support( lo, hi, density )
    kind, value, ng, g = _density3_args( density )
    opts = _f64( [ tol, tol_rel, maxit, max_backtracks, t_min, mult_ok, residual, power, switch, lin_tol, lin_maxit, int( keep_start ),
                   0, 0.9, { "never": 0, "auto": 1, "always": 2 }[ continuation ], conv_threshold, conv_s0, conv_ratio, conv_min, 0, 0, 0,
                   { "cg": 0, "mg": 1 }[ linear ] ] )
    weights = np.zeros( n )
    stats = np.zeros( len( STATS ) )
    hist = np.zeros( ( maxit + 1, len( HISTORY ) ) )
    masses = np.zeros( n )
    bary = np.zeros( ( n, 3 ) )
    cost = np.zeros( 1 )
    with _image3( density ):
        nb = lib().oracle_solve_3d( ctypes.c_longlong( n ), _d( pos ), _d( nu ), _d( w0 ), _d( _f64( lo ) ), _d( _f64( hi ) ), ctypes.c_int( kind ),
                                    ctypes.c_double( value ), ctypes.c_int( ng ), _d( g ), _d( opts ), _d( weights ), _d( stats ), _d( hist ),
                                    _d( masses ), _d( bary ), _d( cost ) )
    return dict( weights = weights, masses = masses, barycenters = bary, cost = float( cost[ 0 ] ),
                 stats = dict( zip( STATS, stats.tolist() ) ), history = hist[ :nb ].copy() )
