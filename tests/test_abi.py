"""The C ABI on the CPU: the library builds, loads, and exports exactly what `include/sdot_amd.h`
declares. No compute call ( those need the GPU )."""

import os
import re
import subprocess

ROOT = os.path.dirname( os.path.dirname( os.path.abspath( __file__ ) ) )
HEADER = os.path.join( ROOT, "include", "sdot_amd.h" )


def _declared():
    text = open( HEADER ).read()
    return sorted( set( re.findall( r"\b(sdot_amd_\w+)\s*\(", text ) ) )


def test_library_exports_every_declared_symbol():
    from sdot import _native
    lib = _native.lib()
    declared = _declared()
    assert len( declared ) == len( set( _native.EXPORTS ) )
    for name in declared:
        assert hasattr( lib, name ), name
    assert sorted( _native.EXPORTS ) == declared
    # and nothing else with the prefix is exported
    out = subprocess.run( [ "nm", "-D", "--defined-only", _native.LIB_PATH ], capture_output = True, text = True, check = True ).stdout
    exported = sorted( set( re.findall( r"\b(sdot_amd_\w+)\b", out ) ) )
    assert exported == declared


def test_abi_version_and_tree_sizes():
    from sdot import _native
    lib = _native.lib()
    assert lib.sdot_amd_abi_version() == 1
    # the node count is a function of n alone ( AaBsp.max_nb_nodes_for )
    for n, depth in ( ( 1, 1 ), ( 10, 1 ), ( 11, 2 ), ( 21, 3 ), ( 10 ** 6, 18 ) ):
        assert lib.sdot_amd_tree_nodes( n, 10 ) == 2 ** depth - 1


def test_errors_cross_the_abi_without_a_gpu_call():
    # argument checks happen before any device work: a null array is kind 1 with a message
    import ctypes
    from sdot import _native
    lib = _native.lib()
    box = ( ctypes.c_double * 4 )( 0, 0, 1, 1 )
    rc = lib.sdot_amd_measures_2d( None, 0, None, None, box, 0, 1.0, None, 0, 0, None, None, None, None, None, 0, None, None )
    assert rc == 1
    k, c, v = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
    msg = lib.sdot_amd_last_error( ctypes.byref( k ), ctypes.byref( c ), ctypes.byref( v ) ).decode()
    assert k.value == 1 and "required" in msg


def test_options_vector_matches_the_reference_layout():
    # 41 doubles, the reference card's order ( src/sdot/SdotPlanNd.py:197-201 )
    from sdot import Iterative, Image, Tuning
    from sdot import _native
    from sdot.SdotPlanNd import card_options
    import numpy as np
    assert len( _native.CARD_OPTIONS ) == 41
    img = Image( np.arange( 12.0 ).reshape( 3, 4 ) + 1, origin = [ 1, 2 ], frame = np.diag( [ 0.5, 0.25 ] ) ).normalized_version()
    o = card_options( Iterative( tol = 3e-9, max_iter = 17, tuning = Tuning( residual = "lin" ) ), img )
    idx = { k: i for i, k in enumerate( _native.CARD_OPTIONS ) }
    assert o[ idx[ "tol_abs" ] ] == 3e-9 and o[ idx[ "maxit" ] ] == 17 and o[ idx[ "residual" ] ] == 0
    assert list( o[ idx[ "img_x0" ]: idx[ "img_ny" ] + 1 ] ) == [ 1, 2, 0.5, 0.25, 3, 4 ]
