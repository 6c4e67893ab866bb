"""The distributions of the solve: the discrete source and the target densities.

Mirrors `src/sdot/distributions/{SumOfDiracs,Image,SumOfGaussians}.py` of the reference: the same
constructors, `normalized_version()` ( total mass 1 ) and the support the target declares
( `bounding_half_spaces`, the domain of the problem ).
"""

import numpy as np


def _host( a ):
    try:
        import torch
        if isinstance( a, torch.Tensor ):
            return a.detach().cpu().numpy()
    except ImportError:
        pass
    return np.asarray( a )


class SumOfDiracs:
    """weighted Dirac masses at `positions[ n, d ]` ( reference `distributions/SumOfDiracs.py:16` )"""

    _is_dirac_source = True

    def __init__( self, positions, weights = None, target_mass = 1.0 ):
        self.positions = positions
        self.weights = weights
        self.target_mass = target_mass

    @property
    def nb_diracs( self ):
        return int( self.positions.shape[ 0 ] )

    @property
    def nb_dims( self ):
        return int( self.positions.shape[ 1 ] )

    @property
    def mass( self ):
        return float( self.nb_diracs if self.weights is None else _host( self.weights ).sum() )

    def normalized_version( self, nb_dims = None ):
        """the weights scaled to `target_mass` ( `1 / n` each when none were given ), `SumOfDiracs.py:43`"""
        if self.weights is None:
            w = np.full( self.nb_diracs, self.target_mass / self.nb_diracs )
        else:
            w = _host( self.weights ).astype( np.float64 ).reshape( -1 )
            w = w * ( self.target_mass / w.sum() )
        res = SumOfDiracs( self.positions, w, self.target_mass )
        return res


def _affine( x, A, b ):
    """`x A^T + b` for rows `x`, numpy or torch ( a torch tensor stays one, on its device, differentiable )"""
    try:
        import torch
        if isinstance( x, torch.Tensor ):
            At = torch.as_tensor( np.ascontiguousarray( A.T ), dtype = x.dtype, device = x.device )
            return x @ At + torch.as_tensor( b, dtype = x.dtype, device = x.device )
    except ImportError:
        pass
    return np.asarray( x, dtype = np.float64 ) @ A.T + b


class Image:
    """piecewise-constant density on a grid: pixel `( i, j )` ( first index along x ) covers
    `origin + frame^T [ i, i + 1 ] x [ j, j + 1 ]` ( reference `distributions/Image.py:12`, `Image.cxx:97` ).
    Regular grids: a diagonal frame, or a TURNED one whose rows are orthogonal ( a rotated or reflected grid of rectangular
    pixels ). A turned grid is solved in its own frame: `y = R ( x - origin )` ( `R` the frame's rows made unit ) moves and
    turns the plane without changing a distance, so power diagrams, masses and weights are the same there, and the grid is
    axis-aligned ( `local_version` ); positions go there ( `to_local` ) and barycentres come back ( `from_local` ). A frame
    whose rows are not orthogonal ( sheared pixels ) is refused: the reference's card refuses every frame that is not
    diagonal ( `PowerDiagram_Bsp.py:247-248` )."""

    def __init__( self, values, origin = None, frame = None, target_mass = 1.0 ):
        self.values = values
        v = _host( values )
        self.nb_dims = v.ndim
        self.origin = np.zeros( v.ndim ) if origin is None else np.asarray( _host( origin ), dtype = np.float64 ).reshape( -1 )
        self.frame = np.eye( v.ndim ) if frame is None else np.asarray( _host( frame ), dtype = np.float64 ).reshape( v.ndim, v.ndim )
        self.target_mass = target_mass
        #: the turn of a turned grid ( `None`: a diagonal frame ), and its pixels' sides
        self.rotation = None
        if np.any( self.frame != np.diag( np.diag( self.frame ) ) ) or np.any( np.diag( self.frame ) <= 0 ):
            steps = np.sqrt( ( self.frame ** 2 ).sum( axis = 1 ) )
            gram = self.frame @ self.frame.T
            if not np.all( steps > 0 ) or np.abs( gram - np.diag( np.diag( gram ) ) ).max() > 1e-12 * float( steps.max() ) ** 2:
                raise NotImplementedError( "Image: a diagonal frame, or one whose rows are orthogonal ( a turned grid ); sheared pixels "
                                           "are not taken" )
            self.rotation = self.frame / steps[ :, None ]
            self.steps = steps

    def to_local( self, x ):
        """positions `[ n, d ]` in the frame of a turned grid: `R ( x - origin )`"""
        return _affine( x, self.rotation, -self.rotation @ self.origin )

    def from_local( self, y ):
        """positions `[ n, d ]` back from the frame of a turned grid: `origin + R^T y`"""
        return _affine( y, self.rotation.T, self.origin )

    def local_version( self ):
        """the same image in the frame of a turned grid: origin 0, the diagonal frame of its pixels' sides"""
        return Image( self.values, np.zeros( self.nb_dims ), np.diag( self.steps ), self.target_mass )

    @property
    def shape( self ):
        return tuple( _host( self.values ).shape )

    @property
    def mass( self ):
        return float( _host( self.values ).sum() * abs( np.linalg.det( self.frame ) ) )

    def normalized_version( self, nb_dims = None ):
        v = _host( self.values ).astype( np.float64 )
        m = self.mass
        if not ( m > 0 ):
            raise ValueError( "Image: the image has no mass" )
        return Image( v * ( self.target_mass / m ), self.origin, self.frame, self.target_mass )

    def box( self ):
        """`( lo, hi )` of the grid ( of its corners, for a turned grid )"""
        n = np.asarray( self.shape, dtype = np.float64 )
        if self.rotation is None:
            h = np.diag( self.frame )
            return self.origin.copy(), self.origin + h * n
        import itertools
        corners = np.array( [ self.origin + self.frame.T @ ( n * np.array( c ) ) for c in itertools.product( ( 0, 1 ), repeat = self.nb_dims ) ] )
        return corners.min( axis = 0 ), corners.max( axis = 0 )

    def bounding_half_spaces( self ):
        """the block of the grid as half-spaces `dir . x <= off` ( `Image.py:66-80` ); a turned grid's along its own axes"""
        d = self.nb_dims
        if self.rotation is None:
            lo, hi = self.box()
            return np.concatenate( [ np.eye( d ), -np.eye( d ) ] ), np.concatenate( [ hi, -lo ] )
        R = self.rotation
        ext = self.steps * np.asarray( self.shape, dtype = np.float64 )
        return np.concatenate( [ R, -R ] ), np.concatenate( [ R @ self.origin + ext, -( R @ self.origin ) ] )

    def native_args3( self, device_tensor ):
        """`( kind, rho, voxels, ( nx, ny, nz, geo ) )` for the 3D calls: voxel `( i, j, k )` at `( k ny + j ) nx + i`, `geo` the
        origin and the voxel's sides"""
        v = _host( self.values ).astype( np.float64 )
        if v.ndim != 3:
            raise NotImplementedError( "Image: a 3D diagram takes a 3D image" )
        if self.rotation is not None:
            raise ValueError( "Image: a turned grid is integrated in its own frame ( `local_version`, positions by `to_local` )" )
        nx, ny, nz = v.shape
        h = np.diag( self.frame )
        vox = device_tensor( np.ascontiguousarray( v.transpose( 2, 1, 0 ) ).reshape( -1 ) )
        return 1, 0.0, vox, ( nx, ny, nz, [ *self.origin[ :3 ], *h[ :3 ] ] )

    def native_args( self, device_tensor ):
        """`( kind, rho, img, nx, ny, geo )` for the C ABI: `img[ j * nx + i ]` ( `PowerDiagram_Bsp.py:266-268` )"""
        v = _host( self.values ).astype( np.float64 )
        if v.ndim != 2:
            raise NotImplementedError( "Image: 2D images on this path" )
        if self.rotation is not None:
            raise ValueError( "Image: a turned grid is integrated in its own frame ( `local_version`, positions by `to_local` )" )
        if v.size and np.all( v == v.flat[ 0 ] ):        # equal values: the constant density ( `PowerDiagram_Bsp.py:223-231` )
            return 0, float( v.flat[ 0 ] ), None, 0, 0, [ 0.0, 0.0, 1.0, 1.0 ]
        nx, ny = v.shape
        h = np.diag( self.frame )
        return 1, 0.0, device_tensor( np.ascontiguousarray( v.T ).reshape( -1 ) ), nx, ny, [ self.origin[ 0 ], self.origin[ 1 ], h[ 0 ], h[ 1 ] ]


class UniformBox:
    """the Lebesgue measure of a box, normalized: `rho = 1 / |box|` ( the constant density of the reference card )"""

    def __init__( self, lo, hi, target_mass = 1.0 ):
        self.lo = np.asarray( lo, dtype = np.float64 ).reshape( -1 )
        self.hi = np.asarray( hi, dtype = np.float64 ).reshape( -1 )
        self.nb_dims = len( self.lo )
        self.target_mass = target_mass

    def normalized_version( self, nb_dims = None ):
        return self

    def box( self ):
        return self.lo.copy(), self.hi.copy()

    def bounding_half_spaces( self ):
        d = self.nb_dims
        return np.concatenate( [ np.eye( d ), -np.eye( d ) ] ), np.concatenate( [ self.hi, -self.lo ] )

    @property
    def rho( self ):
        return self.target_mass / float( np.prod( self.hi - self.lo ) )

    def native_args( self, device_tensor ):
        return 0, self.rho, None, 0, 0, [ 0.0, 0.0, 1.0, 1.0 ]

    def native_args3( self, device_tensor ):
        """`( kind, rho, gauss, ng )` of the 3D calls"""
        return 0, self.rho, None, 0


class Polytope:
    """a CONSTANT density on a convex polygon, zero elsewhere ( reference `distributions/Polytope.py:36-80` ): the half-planes
    `directions . x <= offsets`, normalized to the total mass `target_mass`. Its support is the polygon, which makes it a
    domain: the cells of a power diagram start from it ( `StartCell.start_cell` ). 2D on this path, at most 12 vertices."""

    def __init__( self, directions, offsets, density = 1.0, target_mass = 1.0 ):
        from .StartCell import MAX_VERTICES, start_cell
        cell = start_cell( directions, offsets )
        if cell is None:
            raise ValueError( "Polytope: the half-planes do not bound a 2D polygon ( empty, flat, unbounded, or not 2D )" )
        if len( cell[ "vertices" ] ) > MAX_VERTICES:
            raise NotImplementedError( f"Polytope on the MI355X path: at most { MAX_VERTICES } vertices ( got { len( cell[ 'vertices' ] ) } )" )
        self.directions = np.asarray( directions, dtype = np.float64 )
        self.offsets = np.asarray( offsets, dtype = np.float64 ).reshape( -1 )
        self.vertices = cell[ "vertices" ]
        self.area = cell[ "area" ]
        self.density = float( density )
        self.nb_dims = 2
        self.target_mass = target_mass

    def normalized_version( self, nb_dims = None ):
        return self

    def box( self ):
        return self.vertices.min( axis = 0 ), self.vertices.max( axis = 0 )

    def bounding_half_spaces( self ):
        return self.directions.copy(), self.offsets.copy()

    def domain_polygon( self ):
        """the counterclockwise vertices of the support"""
        return self.vertices.copy()

    @property
    def rho( self ):
        return self.target_mass / self.area

    def native_args( self, device_tensor ):
        return 0, self.rho, None, 0, 0, [ 0.0, 0.0, 1.0, 1.0 ]


class SumOfGaussians:
    """isotropic Gaussians `weights_k N( positions_k, sigmas_k^2 I )` truncated to `support_box`
    ( reference `distributions/SumOfGaussians.py:11-101` ): `weights_k` is the MASS of Gaussian `k` ( all 1 by
    default ), the declared support is `support_box`, else the box of `support_sigmas` deviations around them."""

    def __init__( self, positions, sigmas, weights = None, target_mass = 1.0, support_sigmas = 6.0, support_box = None ):
        self.positions = np.asarray( _host( positions ), dtype = np.float64 )
        self.sigmas = np.asarray( _host( sigmas ), dtype = np.float64 ).reshape( -1 )
        k = self.positions.shape[ 0 ]
        self.weights = np.ones( k ) if weights is None else np.asarray( _host( weights ), dtype = np.float64 ).reshape( -1 )
        if np.any( self.sigmas <= 0 ):
            raise ValueError( "SumOfGaussians: the widths must be positive" )
        self.nb_dims = self.positions.shape[ 1 ]
        self.target_mass = target_mass
        if support_box is None:
            lo = ( self.positions - support_sigmas * self.sigmas[ :, None ] ).min( axis = 0 )
            hi = ( self.positions + support_sigmas * self.sigmas[ :, None ] ).max( axis = 0 )
            support_box = ( lo, hi )
        self.support_box = ( np.asarray( support_box[ 0 ], dtype = np.float64 ), np.asarray( support_box[ 1 ], dtype = np.float64 ) )

    @property
    def mass( self ):
        return float( self.weights.sum() )

    def normalized_version( self, nb_dims = None ):
        """the masses scaled to `target_mass` ( a division: the total is the sum of the masses )"""
        return SumOfGaussians( self.positions, self.sigmas, self.weights * ( self.target_mass / self.weights.sum() ),
                               target_mass = self.target_mass, support_box = self.support_box )

    def box( self ):
        return self.support_box[ 0 ].copy(), self.support_box[ 1 ].copy()

    def bounding_half_spaces( self ):
        d = self.nb_dims
        lo, hi = self.box()
        return np.concatenate( [ np.eye( d ), -np.eye( d ) ] ), np.concatenate( [ hi, -lo ] )

    def native_args( self, device_tensor ):
        """`( kind, rho, gauss, ng, 0, geo )`: the Gaussians as rows ( centre x, y, width, mass ) ( `SdotPlanNd.py:837-862` )"""
        if self.nb_dims != 2:
            raise NotImplementedError( "SumOfGaussians: 2D on this path" )
        g = np.concatenate( [ self.positions, self.sigmas[ :, None ], self.weights[ :, None ] ], axis = 1 )
        return 2, 0.0, device_tensor( np.ascontiguousarray( g ).reshape( -1 ) ), g.shape[ 0 ], 0, [ 0.0, 0.0, 1.0, 1.0 ]

    def native_args3( self, device_tensor ):
        """`( kind, rho, gauss, ng )` of the 3D calls: rows ( centre x, y, z, width, mass )"""
        if self.nb_dims != 3:
            raise ValueError( "SumOfGaussians: native_args3 is for 3D Gaussians" )
        g = np.concatenate( [ self.positions, self.sigmas[ :, None ], self.weights[ :, None ] ], axis = 1 )
        return 2, 0.0, device_tensor( np.ascontiguousarray( g ) ), g.shape[ 0 ]
