"""`PowerDiagram`: one diagram of seeds and weights in a box, its cell masses and facets.

Mirrors the constructor and `measures` of `src/sdot/PowerDiagram.py` ( `:120`, `:325-391` ) for
the 2D box domain; `measures` is one call of `sdot_amd_measures_2d` ( standing for
`power_diagram_measures_card_2d`, `PowerDiagram_Bsp.py:271-311` ).
"""

import numpy as np

from . import _native
from .distributions import UniformBox


def box_half_spaces( lo, hi ):
    """the half-spaces `dir . x <= off` of the box `[ lo, hi ]` ( `PowerDiagram.py:706` )"""
    lo = np.asarray( lo, dtype = np.float64 ).reshape( -1 )
    hi = np.asarray( hi, dtype = np.float64 ).reshape( -1 )
    d = len( lo )
    return np.concatenate( [ np.eye( d ), -np.eye( d ) ] ), np.concatenate( [ hi, -lo ] )


def axis_aligned_box( dirs, offs ):
    """`( lo, hi )` when the half-spaces are exactly those of a box, else `None` ( `PowerDiagram.py:657` )"""
    dirs = np.asarray( dirs, dtype = np.float64 )
    offs = np.asarray( offs, dtype = np.float64 ).reshape( -1 )
    d = dirs.shape[ 1 ]
    lo, hi = np.full( d, -np.inf ), np.full( d, np.inf )
    for k in range( len( offs ) ):
        nz = np.nonzero( dirs[ k ] )[ 0 ]
        if len( nz ) != 1:
            return None
        a = nz[ 0 ]
        s = dirs[ k, a ]
        if s > 0:
            hi[ a ] = min( hi[ a ], offs[ k ] / s )
        else:
            lo[ a ] = max( lo[ a ], offs[ k ] / s )
    if np.any( ~np.isfinite( lo ) ) or np.any( ~np.isfinite( hi ) ):
        return None
    return lo, hi


class PowerDiagram:
    """the power diagram of `( positions, weights )` in the domain given by `boundaries` ( half-spaces )
    or by the support of `distribution`"""

    def __init__( self, positions, weights = None, boundaries = None, accelerator = None, distribution = None, kernel_dtype = None, **kwargs ):
        self.positions = positions
        self.weights = weights
        if distribution is not None:
            dist = distribution.normalized_version()
            lo, hi = dist.box()
        elif boundaries is not None:
            box = axis_aligned_box( *boundaries )
            if box is None:
                raise NotImplementedError( "PowerDiagram on the MI355X path: a box domain" )
            lo, hi = box
            dist = UniformBox( lo, hi )
        else:
            raise ValueError( "PowerDiagram: a domain is needed ( `boundaries` or `distribution` )" )
        if len( lo ) not in ( 2, 3 ):
            raise NotImplementedError( "PowerDiagram on the MI355X path: 2D or 3D" )
        self.distribution = dist
        self.box_min, self.box_max = lo, hi
        self.nb_dims = len( lo )

    @property
    def nb_points( self ):
        return int( self.positions.shape[ 0 ] )

    def _box( self ):
        return [ *self.box_min, *self.box_max ]

    def _measures( self, facets ):
        if self.nb_dims == 3:
            return _native.measures_3d( self.positions, self.weights, self._box(), self.distribution, facets = facets )
        return _native.measures_2d( self.positions, self.weights, self._box(), self.distribution, facets = facets )

    @property
    def measures( self ):
        """the mass of each cell against the density ( device tensor, user order )"""
        return self._measures( False )

    def facets( self ):
        """`( i, j, c_ij )`: each cell's own view of its facets, `c_ij = int_facet rho / ( 2 |p_i - p_j| )`"""
        _, f = self._measures( True )
        return f
