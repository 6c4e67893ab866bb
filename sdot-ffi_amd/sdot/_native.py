"""The C ABI of `libsdot_amd.so` ( `include/sdot_amd.h` ), loaded with ctypes.

Device memory is torch's ( ROCm ): every array crosses as a `data_ptr()` with its size, on torch's
current stream. There is no fallback: without the library, or without a GPU, the calls raise.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname( os.path.abspath( __file__ ) )
LIB_PATH = os.path.join( _HERE, "_lib", "libsdot_amd.so" )
_lib = None

_P = ctypes.c_void_p
_LL = ctypes.c_longlong
_D = ctypes.c_double
_I = ctypes.c_int

#: the option vector of the solve, in the reference card's order ( `src/sdot/SdotPlanNd.py:197-201` )
CARD_OPTIONS = [ "tol_abs", "tol_rel", "t_min", "mult_ok", "factor", "maxit", "max_backtracks", "step", "residual", "power", "switch",
                 "lin", "host_method", "lin_tol", "amg_variant", "mg_shift", "mg_recycle", "mg_rebuild", "mg_stop", "mg_nu", "mg_kcycle",
                 "lin_maxit", "trace", "mg_float", "mg_smoothed",
                 "continuation", "conv_threshold", "conv_s0", "conv_ratio", "conv_min", "conv_possible", "min_scale",
                 "img_x0", "img_y0", "img_hx", "img_hy", "img_nx", "img_ny", "agg_margin", "agg_gap", "agg_nb_dups" ]

#: what the library exports ( every symbol of `include/sdot_amd.h` )
EXPORTS = [ "sdot_amd_abi_version", "sdot_amd_last_error", "sdot_amd_device_bytes", "sdot_amd_release", "sdot_amd_set_timing",
            "sdot_amd_kernel_times", "sdot_amd_measures_2d",
            "sdot_amd_solve_2d", "sdot_amd_tree_nodes", "sdot_amd_build_tree_2d", "sdot_amd_refresh_majorants_2d",
            "sdot_amd_linear_solves_2d", "sdot_amd_solve_2d_batch", "sdot_amd_server_start", "sdot_amd_server_submit",
            "sdot_amd_server_wait", "sdot_amd_server_stop", "sdot_amd_measures_3d", "sdot_amd_solve_3d",
            "sdot_amd_solve_2d_precision", "sdot_amd_graph_counts", "sdot_amd_solve_bytes", "sdot_amd_server_submit_nd", "sdot_amd_server_info",
            "sdot_amd_measures_2d_noalloc", "sdot_amd_measures_2d_noalloc_max_seeds", "sdot_amd_fp32_counts", "sdot_amd_mg32_counts",
            "sdot_amd_measures_2d_precision", "sdot_amd_moments_2d", "sdot_amd_moments_3d", "sdot_amd_cells_2d",
            "sdot_amd_measures_vjp_2d", "sdot_amd_solve_2d_batch_precision", "sdot_amd_measures_2d_domain", "sdot_amd_solve_2d_domain",
            "sdot_amd_measures_3d_image", "sdot_amd_solve_3d_image", "sdot_amd_moments_3d_image", "sdot_amd_server_submit_3d_image" ]

CONSTANT, IMAGE, GAUSSIANS = 0, 1, 2


class SdotError( RuntimeError ):
    """a failure reported by the library: `kind` ( 1 argument / HIP, 2 kernel failure, 3 capacity, 4 memory ), `code`, `value`"""

    def __init__( self, msg, kind, code, value ):
        super().__init__( msg )
        self.kind, self.code, self.value = kind, code, value


class KernelFailure( SdotError ):
    """a cell the kernels could not build ( `code` 1: more vertices than the limit, `value` the seed )"""


def lib():
    """the loaded library ( raises when it is not built )"""
    global _lib
    if _lib is None:
        if not os.path.exists( LIB_PATH ):
            raise RuntimeError( f"libsdot_amd.so is not built ( expected at { LIB_PATH } ): `make -C sdot-ffi_amd/csrc` "
                                "or `python -c 'import __graft_entry__ as g; g.build()'`" )
        # ONE HIP runtime per process: torch's is loaded first, so that the library's `libamdhip64.so.7` resolves to
        # it ( the same soname ) instead of loading the system's beside it -- two runtimes do not share devices
        torch().cuda.is_available()
        L = ctypes.CDLL( LIB_PATH )
        L.sdot_amd_abi_version.restype = _I
        L.sdot_amd_last_error.restype = ctypes.c_char_p
        L.sdot_amd_last_error.argtypes = [ ctypes.POINTER( _I ), ctypes.POINTER( _I ), ctypes.POINTER( _LL ) ]
        L.sdot_amd_device_bytes.restype = _LL
        L.sdot_amd_tree_nodes.restype = _LL
        L.sdot_amd_tree_nodes.argtypes = [ _LL, _I ]
        L.sdot_amd_measures_2d.argtypes = [ _P, _LL, _P, _P, _P, _I, _D, _P, _I, _I, _P, _P, _P, _P, _P, _LL, _P, _P ]
        L.sdot_amd_measures_2d_precision.argtypes = [ _P, _LL, _P, _P, _P, _I, _D, _P, _I, _I, _P, _P, _P, _P, _P, _LL, _P, _P, _I ]
        L.sdot_amd_measures_2d_precision.restype = _I
        L.sdot_amd_measures_2d_domain.argtypes = [ _P, _LL, _P, _P, _P, _I, _D, _P, _I, _I, _P, _P, _P, _P, _P, _LL, _P, _P, _I, _I, _P ]
        L.sdot_amd_measures_2d_domain.restype = _I
        L.sdot_amd_measures_3d_image.argtypes = [ _P, _LL, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _LL, _P ]
        L.sdot_amd_measures_3d_image.restype = _I
        L.sdot_amd_solve_3d_image.argtypes = [ _P, _LL, _P, _P, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P ]
        L.sdot_amd_solve_3d_image.restype = _I
        L.sdot_amd_moments_3d_image.argtypes = [ _P, _LL, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P ]
        L.sdot_amd_moments_3d_image.restype = _I
        L.sdot_amd_server_submit_3d_image.argtypes = [ _LL, _LL, _P, _P, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P ]
        L.sdot_amd_server_submit_3d_image.restype = _I
        L.sdot_amd_moments_2d.argtypes = [ _P, _LL, _P, _P, _P, _I, _D, _P, _I, _I, _P, _P, _P, _P ]
        L.sdot_amd_moments_2d.restype = _I
        L.sdot_amd_moments_3d.argtypes = [ _P, _LL, _P, _P, _P, _I, _D, _P, _I, _P, _P, _P ]
        L.sdot_amd_moments_3d.restype = _I
        L.sdot_amd_cells_2d.argtypes = [ _P, _LL, _P, _P, _P, _I, _D, _P, _I, _I, _P, _P, _P, _P, _LL, _P ]
        L.sdot_amd_cells_2d.restype = _I
        L.sdot_amd_measures_vjp_2d.argtypes = [ _P, _LL, _P, _P, _P, _I, _D, _P, _I, _I, _P, _P, _P, _P ]
        L.sdot_amd_measures_vjp_2d.restype = _I
        L.sdot_amd_measures_2d_noalloc.argtypes = [ _P, _LL, _P, _P, _P, _I, _D, _P, _I, _I, _P, _P, _P ]
        L.sdot_amd_measures_2d_noalloc_max_seeds.restype = _LL
        L.sdot_amd_solve_2d.argtypes = [ _P, _LL, _P, _P, _P, _P, _LL, _P, _P, _I, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P ]
        L.sdot_amd_solve_2d_precision.argtypes = [ _P, _LL, _P, _P, _P, _P, _LL, _P, _P, _I, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I ]
        L.sdot_amd_solve_2d_domain.argtypes = [ _P, _LL, _P, _P, _P, _P, _LL, _P, _P, _I, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _P ]
        L.sdot_amd_solve_2d_domain.restype = _I
        L.sdot_amd_build_tree_2d.argtypes = [ _P, _LL, _P, _I, _P, _P, _P, _P, _P, _P, _P ]
        L.sdot_amd_refresh_majorants_2d.argtypes = [ _P, _LL, _LL, _P, _P, _P, _P, _P, _P, _P ]
        L.sdot_amd_linear_solves_2d.argtypes = [ _P, _LL, _P, _P, _P, _P, _P, _LL, _P, _P, _P ]
        L.sdot_amd_solve_2d_batch.argtypes = [ _P, _I, _LL, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P ]
        L.sdot_amd_solve_2d_batch_precision.argtypes = L.sdot_amd_solve_2d_batch.argtypes + [ _I ]
        L.sdot_amd_solve_2d_batch_precision.restype = _I
        L.sdot_amd_measures_3d.argtypes = [ _P, _LL, _P, _P, _P, _I, _D, _P, _I, _P, _P, _P, _P, _LL, _P, _P ]
        L.sdot_amd_solve_3d.argtypes = [ _P, _LL, _P, _P, _P, _P, _P, _I, _D, _P, _I, _P, _P, _P, _P, _P, _P, _P ]
        L.sdot_amd_server_start.argtypes = [ _I, _D ]
        L.sdot_amd_server_start.restype = _LL
        L.sdot_amd_server_submit.argtypes = [ _LL, _LL, _P, _P, _P, _P, _P, _I, _D, _P, _P, _P, _P, _P, _P, _P ]
        L.sdot_amd_server_wait.argtypes = [ _LL, _LL, _P, _P ]
        L.sdot_amd_server_submit_nd.argtypes = [ _LL, _I, _I, _LL, _P, _P, _P, _P, _P, _I, _D, _P, _I, _P, _P, _P, _P, _P, _P ]
        L.sdot_amd_server_submit_nd.restype = _I
        L.sdot_amd_server_info.argtypes = [ _LL, _P ]
        L.sdot_amd_server_info.restype = _I
        L.sdot_amd_server_stop.argtypes = [ _LL ]
        L.sdot_amd_set_timing.argtypes = [ _I ]
        L.sdot_amd_set_timing.restype = _I
        L.sdot_amd_kernel_times.argtypes = [ _P, _P, _P, _P, _I ]
        L.sdot_amd_kernel_times.restype = _I
        L.sdot_amd_fp32_counts.argtypes = [ _P, _P, _P ]
        L.sdot_amd_fp32_counts.restype = _I
        L.sdot_amd_mg32_counts.argtypes = [ _P, _P, _P ]
        L.sdot_amd_mg32_counts.restype = _I
        L.sdot_amd_graph_counts.argtypes = [ _P, _P ]
        L.sdot_amd_graph_counts.restype = _I
        L.sdot_amd_solve_bytes.argtypes = [ _LL, _I, _I, _P, _I, _P ]
        L.sdot_amd_solve_bytes.restype = _I
        for name in ( "sdot_amd_measures_2d", "sdot_amd_solve_2d", "sdot_amd_build_tree_2d", "sdot_amd_refresh_majorants_2d",
                      "sdot_amd_linear_solves_2d", "sdot_amd_solve_2d_batch", "sdot_amd_server_submit", "sdot_amd_server_wait",
                      "sdot_amd_server_stop", "sdot_amd_measures_3d", "sdot_amd_solve_3d", "sdot_amd_solve_2d_precision",
                      "sdot_amd_measures_2d_noalloc" ):
            getattr( L, name ).restype = _I
        if L.sdot_amd_abi_version() != 1:
            raise RuntimeError( "libsdot_amd.so: unexpected ABI version" )
        _lib = L
    return _lib


def _check( rc ):
    if rc == 0:
        return
    k, c, v = _I(), _I(), _LL()
    msg = lib().sdot_amd_last_error( ctypes.byref( k ), ctypes.byref( c ), ctypes.byref( v ) ).decode()
    cls = KernelFailure if k.value == 2 else SdotError
    if k.value == 4:
        raise MemoryError( msg )
    raise cls( msg, k.value, c.value, v.value )


# ---- torch plumbing ------------------------------------------------------------------------------------


def torch():
    import torch as T
    return T


def device():
    """the GPU the library runs on ( raises without one: there is no CPU path )"""
    T = torch()
    if not T.cuda.is_available():
        raise RuntimeError( "sdot (MI355X): no GPU visible -- this implementation runs on the device only" )
    return T.device( "cuda", T.cuda.current_device() )


def on_device( a, dtype = None ):
    """`a` as a contiguous device tensor ( float64 unless told otherwise )"""
    T = torch()
    dtype = dtype or T.float64
    if isinstance( a, T.Tensor ):
        return a.detach().to( device = device(), dtype = dtype ).contiguous()
    return T.as_tensor( np.ascontiguousarray( np.asarray( a ) ), dtype = dtype, device = device() ).contiguous()


def _ptr( t ):
    return None if t is None else _P( t.data_ptr() )


def _stream():
    return _P( torch().cuda.current_stream().cuda_stream )


def _host( a ):
    a = np.ascontiguousarray( np.asarray( a, dtype = np.float64 ) )
    return a, a.ctypes.data_as( _P )


# ---- the calls -------------------------------------------------------------------------------------------


def _domain_args( domain ):
    """`( nb, pointer, keep )` of a polygon domain ( counterclockwise vertices ), or `( 0, None, None )`"""
    if domain is None:
        return 0, None, None
    v = np.ascontiguousarray( np.asarray( domain, dtype = np.float64 ).reshape( -1, 2 ) )
    return len( v ), v.ctypes.data_as( _P ), v


def measures_2d( positions, weights, box, density, facets = False, events = False, precision = 0, domain = None ):
    """masses ( device tensor, user order ); with `facets` also `( fi, fj, fc )` ( device tensors ); with `events` the counts;
    `precision` FP64 or FP32 ( the kernel's float ); `domain`: a convex polygon ( counterclockwise vertices, `box` its bounding
    box ) the cells start from instead of the box"""
    T = torch()
    pos = on_device( positions ).reshape( -1, 2 ).contiguous()
    n = pos.shape[ 0 ]
    w = None if weights is None else on_device( weights ).reshape( -1 )
    boxh, boxp = _host( box )
    kind, rho, img, nx, ny, geo = density.native_args( device_tensor = on_device )
    geoh, geop = _host( geo )
    masses = T.empty( n, dtype = T.float64, device = device() )
    cap = 8 * n + 1024 if facets else 0
    fi = T.empty( cap, dtype = T.int64, device = device() ) if facets else None
    fj = T.empty( cap, dtype = T.int64, device = device() ) if facets else None
    fc = T.empty( cap, dtype = T.float64, device = device() ) if facets else None
    nbf = np.zeros( 1, np.int64 )
    ev = np.zeros( 5, np.uint64 )
    evp = ev.ctypes.data_as( _P ) if events else None
    if domain is None:
        _check( lib().sdot_amd_measures_2d_precision( _stream(), n, _ptr( pos ), _ptr( w ), boxp, kind, rho, _ptr( img ), nx, ny, geop,
                                                      _ptr( masses ), _ptr( fi ), _ptr( fj ), _ptr( fc ), cap, nbf.ctypes.data_as( _P ), evp,
                                                      int( precision ) ) )
    else:
        nd, dp, keep = _domain_args( domain )
        _check( lib().sdot_amd_measures_2d_domain( _stream(), n, _ptr( pos ), _ptr( w ), boxp, kind, rho, _ptr( img ), nx, ny, geop,
                                                   _ptr( masses ), _ptr( fi ), _ptr( fj ), _ptr( fc ), cap, nbf.ctypes.data_as( _P ), evp,
                                                   int( precision ), nd, dp ) )
    res = [ masses ]
    if facets:
        k = int( nbf[ 0 ] )
        res.append( ( fi[ :k ], fj[ :k ], fc[ :k ] ) )
    if events:
        res.append( dict( zip( ( "side_tests", "prune_vertex_tests", "new_vertices", "node_visits", "cuts" ), [ int( x ) for x in ev ] ) ) )
    return res[ 0 ] if len( res ) == 1 else tuple( res )


def _diagram_args( positions, weights, box, density, dim = 2 ):
    pos = on_device( positions ).reshape( -1, dim ).contiguous()
    w = None if weights is None else on_device( weights ).reshape( -1 )
    return pos, w, _host( box )


def moments( positions, weights, box, density ):
    """`( masses, first, second )` of each cell ( device tensors, user order; `first` `[ n, d ]` )"""
    T = torch()
    dim = len( box ) // 2
    pos, w, ( boxh, boxp ) = _diagram_args( positions, weights, box, density, dim )
    n = pos.shape[ 0 ]
    m = T.empty( n, dtype = T.float64, device = device() )
    f = T.empty( ( n, dim ), dtype = T.float64, device = device() )
    s2 = T.empty( n, dtype = T.float64, device = device() )
    if dim == 3:
        kind, rho, g, ng = _density3_args( density )
        if kind == IMAGE:
            nx, ny, nz, geo = ng
            geoh, geop = _host( geo )
            _check( lib().sdot_amd_moments_3d_image( _stream(), n, _ptr( pos ), _ptr( w ), boxp, _ptr( g ), nx, ny, nz, geop, _ptr( m ),
                                                     _ptr( f ), _ptr( s2 ) ) )
        else:
            _check( lib().sdot_amd_moments_3d( _stream(), n, _ptr( pos ), _ptr( w ), boxp, kind, rho, _ptr( g ), ng, _ptr( m ), _ptr( f ),
                                               _ptr( s2 ) ) )
    else:
        kind, rho, img, nx, ny, geo = density.native_args( device_tensor = on_device )
        geoh, geop = _host( geo )
        _check( lib().sdot_amd_moments_2d( _stream(), n, _ptr( pos ), _ptr( w ), boxp, kind, rho, _ptr( img ), nx, ny, geop, _ptr( m ),
                                           _ptr( f ), _ptr( s2 ) ) )
    return m, f, s2


def cells_2d( positions, weights, box, density ):
    """every cell's vertices: `( offsets [ n + 1 ], vertices [ V, 2 ], cuts [ V ] )` ( device tensors, user order; a cut is a seed's
    user index, or `-1 - f` for side `f` of the box )"""
    T = torch()
    pos, w, ( boxh, boxp ) = _diagram_args( positions, weights, box, density )
    n = pos.shape[ 0 ]
    kind, rho, img, nx, ny, geo = density.native_args( device_tensor = on_device )
    geoh, geop = _host( geo )
    nv = np.zeros( 1, np.int64 )
    _check( lib().sdot_amd_cells_2d( _stream(), n, _ptr( pos ), _ptr( w ), boxp, kind, rho, _ptr( img ), nx, ny, geop, None, None, None, 0,
                                     nv.ctypes.data_as( _P ) ) )
    cap = int( nv[ 0 ] )
    off = T.empty( n + 1, dtype = T.int64, device = device() )
    xy = T.empty( ( max( cap, 1 ), 2 ), dtype = T.float64, device = device() )
    cut = T.empty( max( cap, 1 ), dtype = T.int64, device = device() )
    _check( lib().sdot_amd_cells_2d( _stream(), n, _ptr( pos ), _ptr( w ), boxp, kind, rho, _ptr( img ), nx, ny, geop, _ptr( off ), _ptr( xy ),
                                     _ptr( cut ), cap, nv.ctypes.data_as( _P ) ) )
    return off, xy[ :cap ], cut[ :cap ]


def measures_vjp_2d( positions, weights, box, density, grad_masses ):
    """`( grad_weights [ n ], grad_positions [ n, 2 ] )`: the cotangent of the masses pulled back ( device tensors, user order )"""
    T = torch()
    pos, w, ( boxh, boxp ) = _diagram_args( positions, weights, box, density )
    n = pos.shape[ 0 ]
    g = on_device( grad_masses ).reshape( -1 )
    kind, rho, img, nx, ny, geo = density.native_args( device_tensor = on_device )
    geoh, geop = _host( geo )
    gw = T.empty( n, dtype = T.float64, device = device() )
    gp = T.empty( ( n, 2 ), dtype = T.float64, device = device() )
    _check( lib().sdot_amd_measures_vjp_2d( _stream(), n, _ptr( pos ), _ptr( w ), boxp, kind, rho, _ptr( img ), nx, ny, geop, _ptr( g ),
                                            _ptr( gw ), _ptr( gp ) ) )
    return gw, gp


def measures_2d_noalloc( positions, weights, box, density, masses = None ):
    """the masses of `measures_2d` by the entry point that allocates nothing ( its scratch a fixed device array ): `masses`
    ( a device tensor of `n` float64, given or made here: the caller's allocation, not the library's )"""
    T = torch()
    pos = on_device( positions ).reshape( -1, 2 ).contiguous()
    n = pos.shape[ 0 ]
    w = None if weights is None else on_device( weights ).reshape( -1 )
    boxh, boxp = _host( box )
    kind, rho, img, nx, ny, geo = density.native_args( device_tensor = on_device )
    geoh, geop = _host( geo )
    if masses is None:
        masses = T.empty( n, dtype = T.float64, device = device() )
    _check( lib().sdot_amd_measures_2d_noalloc( _stream(), n, _ptr( pos ), _ptr( w ), boxp, kind, rho, _ptr( img ), nx, ny, geop,
                                                _ptr( masses ), None ) )
    return masses


def _density3_args( density ):
    """`( kind, rho, gaussians device tensor [ ng, 5 ] or None, ng )` of a 3D density: `None` / a float ( a constant ), an
    object with `native_args3`, or `[ ng, 5 ]` rows ( centre x, y, z, width, mass ). A voxel image gives `( IMAGE, 0, voxels,
    ( nx, ny, nz, geo ) )` ( `geo`: origin x, y, z and the voxel's sides )"""
    if density is None or np.isscalar( density ):
        return CONSTANT, 1.0 if density is None else float( density ), None, 0
    if hasattr( density, "native_args3" ):
        return density.native_args3( device_tensor = on_device )
    g = on_device( np.asarray( density, dtype = np.float64 ).reshape( -1, 5 ) )
    return GAUSSIANS, 0.0, g, int( g.shape[ 0 ] )


def measures_3d( positions, weights, box, density = None, facets = False, events = False ):
    """masses of a 3D diagram ( device tensor, user order ); with `facets` also `( fi, fj, fc )`; with `events` the counts"""
    T = torch()
    pos = on_device( positions ).reshape( -1, 3 ).contiguous()
    n = pos.shape[ 0 ]
    w = None if weights is None else on_device( weights ).reshape( -1 )
    boxh, boxp = _host( box )
    kind, rho, g, ng = _density3_args( density )
    masses = T.empty( n, dtype = T.float64, device = device() )
    cap = 24 * n + 1024 if facets else 0
    fi = T.empty( cap, dtype = T.int64, device = device() ) if facets else None
    fj = T.empty( cap, dtype = T.int64, device = device() ) if facets else None
    fc = T.empty( cap, dtype = T.float64, device = device() ) if facets else None
    nbf = np.zeros( 1, np.int64 )
    ev = np.zeros( 5, np.uint64 )
    if kind == IMAGE:
        nx, ny, nz, geo = ng
        geoh, geop = _host( geo )
        _check( lib().sdot_amd_measures_3d_image( _stream(), n, _ptr( pos ), _ptr( w ), boxp, _ptr( g ), nx, ny, nz, geop, _ptr( masses ),
                                                  _ptr( fi ), _ptr( fj ), _ptr( fc ), cap, nbf.ctypes.data_as( _P ) ) )
    else:
        _check( lib().sdot_amd_measures_3d( _stream(), n, _ptr( pos ), _ptr( w ), boxp, kind, rho, _ptr( g ), ng, _ptr( masses ),
                                            _ptr( fi ), _ptr( fj ), _ptr( fc ), cap, nbf.ctypes.data_as( _P ),
                                            ev.ctypes.data_as( _P ) if events else None ) )
    res = [ masses ]
    if facets:
        k = int( nbf[ 0 ] )
        res.append( ( fi[ :k ], fj[ :k ], fc[ :k ] ) )
    if events:
        res.append( dict( zip( ( "side_tests", "prune_vertex_tests", "new_vertices", "node_visits", "cuts" ), [ int( x ) for x in ev ] ) ) )
    return res[ 0 ] if len( res ) == 1 else tuple( res )


def solve_3d( positions, nu, w0, options, box, density = None ):
    """one 3D solve; returns a dict of device tensors and host arrays ( see `sdot_amd_solve_3d` )"""
    T = torch()
    pos = on_device( positions ).reshape( -1, 3 ).contiguous()
    n = pos.shape[ 0 ]
    nu = on_device( nu ).reshape( -1 )
    w0 = None if w0 is None else on_device( w0 ).reshape( -1 )
    opts, optp = _host( options )
    boxh, boxp = _host( box )
    kind, rho, g, ng = _density3_args( density )
    maxit = int( opts[ CARD_OPTIONS.index( "maxit" ) ] )
    dev = device()
    weights = T.empty( n, dtype = T.float64, device = dev )
    masses = T.empty( n, dtype = T.float64, device = dev )
    bary = T.empty( ( n, 3 ), dtype = T.float64, device = dev )
    hist = np.zeros( ( maxit + 1, 8 ) )
    nb_steps = np.zeros( 1, np.int64 )
    stats = np.zeros( 31 )
    cost = np.zeros( 1 )
    if kind == IMAGE:
        nx, ny, nz, geo = ng
        geoh, geop = _host( geo )
        _check( lib().sdot_amd_solve_3d_image( _stream(), n, _ptr( pos ), _ptr( nu ), _ptr( w0 ), optp, boxp, _ptr( g ), nx, ny, nz, geop,
                                               _ptr( weights ), hist.ctypes.data_as( _P ), nb_steps.ctypes.data_as( _P ), stats.ctypes.data_as( _P ),
                                               _ptr( masses ), _ptr( bary ), cost.ctypes.data_as( _P ) ) )
    else:
        _check( lib().sdot_amd_solve_3d( _stream(), n, _ptr( pos ), _ptr( nu ), _ptr( w0 ), optp, boxp, kind, rho, _ptr( g ), ng,
                                         _ptr( weights ), hist.ctypes.data_as( _P ), nb_steps.ctypes.data_as( _P ), stats.ctypes.data_as( _P ),
                                         _ptr( masses ), _ptr( bary ), cost.ctypes.data_as( _P ) ) )
    return dict( weights = weights, cell_masses = masses, barycenters = bary, clusters = None, cost = float( cost[ 0 ] ),
                 stats = stats, history = hist[ :int( nb_steps[ 0 ] ) ] )


#: the kernel's float ( `sdot_amd_solve_2d_precision` )
FP64, FP32, MIXED = 0, 1, 2


def solve_2d( positions, nu, w0, options, box, density, dups = None, precision = FP64, domain = None ):
    """one solve; returns a dict of device tensors and host arrays ( see `sdot_amd_solve_2d` ); `dups`: the exact
    duplicates, `[ k, 2 ]` rows `( dup, rep )` of user ids; `precision`: FP64, FP32 or MIXED cells; `domain`: a convex polygon
    ( counterclockwise vertices, `box` its bounding box ) the cells start from"""
    T = torch()
    pos = on_device( positions ).reshape( -1, 2 ).contiguous()
    n = pos.shape[ 0 ]
    nu = on_device( nu ).reshape( -1 )
    w0 = None if w0 is None else on_device( w0 ).reshape( -1 )
    opts, optp = _host( options )
    boxh, boxp = _host( box )
    kind, rho, img, nx, ny, geo = density.native_args( device_tensor = on_device )
    maxit = int( opts[ CARD_OPTIONS.index( "maxit" ) ] )
    dev = device()
    weights = T.empty( n, dtype = T.float64, device = dev )
    masses = T.empty( n, dtype = T.float64, device = dev )
    bary = T.empty( ( n, 2 ), dtype = T.float64, device = dev )
    clusters = T.empty( n, dtype = T.int64, device = dev )
    hist = np.zeros( ( maxit + 1, 8 ) )
    nb_steps = np.zeros( 1, np.int64 )
    stats = np.zeros( 31 )
    cost = np.zeros( 1 )
    dp = None if dups is None or len( dups ) == 0 else on_device( np.asarray( dups, dtype = np.int64 ).reshape( -1 ), dtype = T.int64 )
    args = [ _stream(), n, _ptr( pos ), _ptr( nu ), _ptr( w0 ), _ptr( dp ), 0 if dp is None else dp.numel(), optp, boxp, kind, rho, _ptr( img ),
             _ptr( weights ), hist.ctypes.data_as( _P ), nb_steps.ctypes.data_as( _P ), stats.ctypes.data_as( _P ), _ptr( masses ), _ptr( bary ),
             cost.ctypes.data_as( _P ), _ptr( clusters ), int( precision ) ]
    if domain is None:
        _check( lib().sdot_amd_solve_2d_precision( *args ) )
    else:
        nd, ddp, keep = _domain_args( domain )
        _check( lib().sdot_amd_solve_2d_domain( *args, nd, ddp ) )
    return dict( weights = weights, cell_masses = masses, barycenters = bary, clusters = clusters, cost = float( cost[ 0 ] ),
                 stats = stats, history = hist[ :int( nb_steps[ 0 ] ) ] )


def solve_2d_batch( instances, options, streams = 8, precision = 0 ):
    """`k` independent solves at once ( see `sdot_amd_solve_2d_batch` ): `instances` is a list of `( positions, nu, w0,
    box, density )`, every density of the same kind; `options` the 41-double vector they share. Returns one dict per
    instance: device tensors ( views of the concatenated outputs ), `stats`, `cost`, `error` ( kind, code, value: 0 when
    it converged or stopped on its own ), `times` ( start, end in seconds from the call's start )."""
    T = torch()
    k = len( instances )
    if k == 0:
        return []
    dev = device()
    pos = [ on_device( p ).reshape( -1, 2 ) for p, _, _, _, _ in instances ]
    sizes = [ int( p.shape[ 0 ] ) for p in pos ]
    offsets = np.zeros( k + 1, np.int64 )
    offsets[ 1: ] = np.cumsum( sizes )
    P = T.cat( pos ).contiguous()
    nu = T.cat( [ on_device( q[ 1 ] ).reshape( -1 ) for q in instances ] ).contiguous()
    has_w0 = [ q[ 2 ] is not None for q in instances ]
    w0 = None
    if any( has_w0 ):
        w0 = T.cat( [ on_device( q[ 2 ] ).reshape( -1 ) if q[ 2 ] is not None else T.zeros( sizes[ i ], dtype = T.float64, device = dev )
                      for i, q in enumerate( instances ) ] ).contiguous()
    boxes = np.ascontiguousarray( np.asarray( [ q[ 3 ] for q in instances ], dtype = np.float64 ).reshape( k, 4 ) )
    args = [ q[ 4 ].native_args( device_tensor = on_device ) for q in instances ]
    kinds = { a[ 0 ] for a in args }
    if len( kinds ) != 1:
        raise ValueError( "solve_2d_batch: every instance's density must be of the same kind" )
    kind = kinds.pop()
    rho = np.asarray( [ a[ 1 ] for a in args ], dtype = np.float64 )
    imgs = [ a[ 2 ] for a in args ]                       # ( kept alive through the call )
    img_ptrs = ( ctypes.c_void_p * k )( *[ ( None if im is None else im.data_ptr() ) for im in imgs ] )
    geo = np.zeros( ( k, 6 ) )
    for i, a in enumerate( args ):
        g = a[ 5 ]
        if g is not None:
            geo[ i, :4 ] = np.asarray( g, dtype = np.float64 ).reshape( -1 )[ :4 ]
        geo[ i, 4 ], geo[ i, 5 ] = a[ 3 ], a[ 4 ]
    opts, optp = _host( options )
    N = int( offsets[ -1 ] )
    weights = T.empty( N, dtype = T.float64, device = dev )
    masses = T.empty( N, dtype = T.float64, device = dev )
    bary = T.empty( ( N, 2 ), dtype = T.float64, device = dev )
    stats = np.zeros( ( k, 31 ) )
    costs = np.zeros( k )
    errors = np.zeros( ( k, 3 ), np.int64 )
    times = np.zeros( ( k, 2 ) )
    _check( lib().sdot_amd_solve_2d_batch_precision( _stream(), int( streams ), k, offsets.ctypes.data_as( _P ), _ptr( P ), _ptr( nu ), _ptr( w0 ), optp,
                                           boxes.ctypes.data_as( _P ), kind, rho.ctypes.data_as( _P ), ctypes.cast( img_ptrs, _P ),
                                           geo.ctypes.data_as( _P ), _ptr( weights ), _ptr( masses ), _ptr( bary ),
                                           stats.ctypes.data_as( _P ), costs.ctypes.data_as( _P ), errors.ctypes.data_as( _P ),
                                           times.ctypes.data_as( _P ), int( precision ) ) )
    out = []
    for i in range( k ):
        a, b = int( offsets[ i ] ), int( offsets[ i + 1 ] )
        out.append( dict( weights = weights[ a:b ], cell_masses = masses[ a:b ], barycenters = bary[ a:b ], stats = stats[ i ],
                          cost = float( costs[ i ] ), error = tuple( int( v ) for v in errors[ i ] ), times = tuple( times[ i ] ) ) )
    return out


def build_tree_2d( positions, leaf = 10 ):
    T = torch()
    pos = on_device( positions ).reshape( -1, 2 ).contiguous()
    n = pos.shape[ 0 ]
    N = int( lib().sdot_amd_tree_nodes( n, leaf ) )
    dev = device()
    out = dict( seed_indices = T.empty( n, dtype = T.int64, device = dev ), rank_of = T.empty( n, dtype = T.int64, device = dev ),
                node_begin = T.empty( N, dtype = T.int64, device = dev ), node_end = T.empty( N, dtype = T.int64, device = dev ),
                node_left = T.empty( N, dtype = T.int64, device = dev ), node_right = T.empty( N, dtype = T.int64, device = dev ),
                node_box = T.empty( ( N, 2, 2 ), dtype = T.float64, device = dev ) )
    _check( lib().sdot_amd_build_tree_2d( _stream(), n, _ptr( pos ), leaf, *[ _ptr( out[ k ] ) for k in (
        "seed_indices", "rank_of", "node_begin", "node_end", "node_left", "node_right", "node_box" ) ] ) )
    return out


def refresh_majorants_2d( tree, sorted_positions, sorted_weights ):
    T = torch()
    sp = on_device( sorted_positions ).reshape( -1, 2 ).contiguous()
    sw = on_device( sorted_weights ).reshape( -1 )
    N = tree[ "node_begin" ].shape[ 0 ]
    wa = T.empty( ( N, 2 ), dtype = T.float64, device = device() )
    wb = T.empty( N, dtype = T.float64, device = device() )
    _check( lib().sdot_amd_refresh_majorants_2d( _stream(), sp.shape[ 0 ], N, _ptr( tree[ "node_box" ] ), _ptr( tree[ "node_begin" ] ),
                                                 _ptr( tree[ "node_end" ] ), _ptr( sp ), _ptr( sw ), _ptr( wa ), _ptr( wb ) ) )
    return wa, wb


def linear_solves_2d( row, col, val, dia, rhs, tol = 1e-6, method = "cg", kcycle = -1, nu = 0 ):
    """`L d_j = rhs[ j ]` ( `SdotPlanNd._card_linear_solves` of the reference: `method` "cg" or "mg" )"""
    T = torch()
    row = on_device( row, T.int64 )
    col = on_device( col, T.int64 )
    val = on_device( val )
    dia = on_device( dia )
    rhs = on_device( rhs )
    if rhs.dim() == 1:
        rhs = rhs[ None ]
    k, n = rhs.shape
    sol = T.empty( ( k, n ), dtype = T.float64, device = device() )
    its = np.zeros( k )
    if method not in ( "cg", "mg" ):
        raise ValueError( f"linear_solves_2d: method 'cg' or 'mg' ( not { method!r} )" )
    opts, optp = _host( [ n, k, 0 if method == "cg" else 1, tol, -1, 0, 0, 0, kcycle, nu ] )
    _check( lib().sdot_amd_linear_solves_2d( _stream(), n, _ptr( row ), _ptr( col ), _ptr( val ), _ptr( dia ), _ptr( rhs.contiguous() ), k,
                                             optp, _ptr( sol ), its.ctypes.data_as( _P ) ) )
    return sol, its.astype( int )


TIMED_KERNELS = [ "cells", "cells_overflow", "spmv", "majorants" ]


def set_timing( on ):
    _check( lib().sdot_amd_set_timing( int( bool( on ) ) ) )


def kernel_times( reset = False ):
    """per timed kernel: `( ms, launches )`, the timed SpMV bytes, and the cell kernels' event counts"""
    ms = np.zeros( 4 )
    la = np.zeros( 4, np.int64 )
    sb = np.zeros( 1 )
    ev = np.zeros( 5, np.uint64 )
    _check( lib().sdot_amd_kernel_times( ms.ctypes.data_as( _P ), la.ctypes.data_as( _P ), sb.ctypes.data_as( _P ), ev.ctypes.data_as( _P ),
                                         int( bool( reset ) ) ) )
    return ( { k: ( float( ms[ i ] ), int( la[ i ] ) ) for i, k in enumerate( TIMED_KERNELS ) }, float( sb[ 0 ] ),
             dict( zip( ( "side_tests", "prune_vertex_tests", "new_vertices", "node_visits", "cuts" ), [ int( x ) for x in ev ] ) ) )


def fp32_counts():
    """the FP32 kernel's diagrams since the last reset of the kernel times, the cells it had redone in FP64 for a one-sided
    facet, and the facets dropped as still one-sided: `{ diagrams, redone, dropped }`"""
    d, r, x = _LL(), _LL(), _LL()
    _check( lib().sdot_amd_fp32_counts( ctypes.byref( d ), ctypes.byref( r ), ctypes.byref( x ) ) )
    return dict( diagrams = int( d.value ), redone = int( r.value ), dropped = int( x.value ) )


def mg32_counts():
    """the multigrid's FP32 preconditioner since the last reset of the kernel times: the systems solved with it, their
    flexible-CG iterations, and the systems redone with the FP64 preconditioner: `{ solves, iterations, fallbacks }`"""
    s, r, f = _LL(), _LL(), _LL()
    _check( lib().sdot_amd_mg32_counts( ctypes.byref( s ), ctypes.byref( r ), ctypes.byref( f ) ) )
    return dict( solves = int( s.value ), iterations = int( r.value ), fallbacks = int( f.value ) )


def graph_counts():
    """the multigrid's graphs since the last reset of the kernel times: `( captured, replayed )`"""
    c, r = _LL(), _LL()
    _check( lib().sdot_amd_graph_counts( ctypes.byref( c ), ctypes.byref( r ) ) )
    return int( c.value ), int( r.value )


#: the parts of `solve_bytes`, in the ABI's order
MEMORY_PARTS = [ "tree", "diagram slots", "step edges", "fp32 polygons / 3D polytopes", "overflow slots", "newton vectors", "moments",
                 "linear solver", "density copies", "reductions" ]


def solve_bytes( n, dim, density_kind, options = None, precision = 0 ):
    """what a solve of `n` seeds takes from the device, per part ( `{ part: bytes }`, `MEMORY_PARTS` )"""
    parts = np.zeros( len( MEMORY_PARTS ) )
    op = None if options is None else np.ascontiguousarray( np.asarray( options, dtype = np.float64 ) )
    _check( lib().sdot_amd_solve_bytes( int( n ), int( dim ), int( density_kind ), None if op is None else op.ctypes.data_as( _P ), int( precision ),
                                        parts.ctypes.data_as( _P ) ) )
    return dict( zip( MEMORY_PARTS, parts.tolist() ) )


def release():
    """frees the device memory the library's arena holds ( for the calling thread's device )"""
    lib().sdot_amd_release()


def device_bytes():
    return int( lib().sdot_amd_device_bytes() )
