"""`PowerDiagram`: one diagram of seeds and weights in a box and its queries: the cell masses ( differentiable with respect
to the positions and the weights ), the facets and Hessian rows, the moments, the cells themselves, the Laplacian solve.

Mirrors `src/sdot/PowerDiagram.py` ( the constructor `:120`, `measures` `:325-391`, `laplacian_solve` `:395-430`, `moments`
`:433-459`, `hessian_rows` `:461-488`, `cells` `:490-515` ) for box domains; each query is one call of the C ABI
( `sdot_amd_measures_2d_precision`, `_moments_2d / _3d`, `_cells_2d`, `_measures_vjp_2d`, `_linear_solves_2d` ).
"""

import numpy as np

from . import _native
from .distributions import Image, UniformBox


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
        # THE DOMAIN ( `PowerDiagram.py:166-195` ): the half-planes given, intersected with the support the distribution declares;
        # a box, or a convex polygon the cells start from ( `StartCell.start_cell` )
        self.domain = None
        dist = distribution.normalized_version() if distribution is not None else None
        if dist is None and boundaries is None:
            raise ValueError( "PowerDiagram: a domain is needed ( `boundaries` or `distribution` )" )
        # a turned grid ( `Image.rotation` ): the diagram built in the grid's frame, where the grid is axis-aligned and the power
        # distances are the same; the positions ( differentiably ) and the half-planes taken there, the coordinates that come out
        # ( moments, cells ) brought back
        self._turn = None
        if isinstance( dist, Image ) and dist.rotation is not None:
            self._turn = dist
            self.positions = dist.to_local( positions )
            if boundaries is not None:
                bd = np.asarray( boundaries[ 0 ], dtype = np.float64 )
                boundaries = ( bd @ dist.rotation.T, np.asarray( boundaries[ 1 ], dtype = np.float64 ).reshape( -1 ) - bd @ dist.origin )
            dist = dist.local_version()
        halves = [ b for b in ( boundaries, dist.bounding_half_spaces() if dist is not None else None ) if b is not None ]
        dirs = np.concatenate( [ np.asarray( h[ 0 ], dtype = np.float64 ) for h in halves ] )
        offs = np.concatenate( [ np.asarray( h[ 1 ], dtype = np.float64 ).reshape( -1 ) for h in halves ] )
        box = axis_aligned_box( dirs, offs )
        if box is not None:
            lo, hi = box
        else:
            from .StartCell import MAX_VERTICES, start_cell
            cell = start_cell( dirs, offs ) if dirs.shape[ 1 ] == 2 else None
            if cell is None:
                raise NotImplementedError( "PowerDiagram on the MI355X path: a box domain, or a bounded 2D convex polygon" )
            if len( cell[ "vertices" ] ) > MAX_VERTICES:
                raise NotImplementedError( f"PowerDiagram on the MI355X path: a polygon domain of at most { MAX_VERTICES } vertices" )
            self.domain = cell[ "vertices" ]
            lo, hi = cell[ "bounding_box" ]
        if dist is None:                                 # ( no distribution: the volumes of the cells, `rho = 1` )
            vol = float( np.prod( np.asarray( hi ) - np.asarray( lo ) ) )
            dist = UniformBox( lo, hi, target_mass = vol )
        if len( lo ) not in ( 2, 3 ):
            raise NotImplementedError( "PowerDiagram on the MI355X path: 2D or 3D" )
        self.distribution = dist
        self.box_min, self.box_max = lo, hi
        self.nb_dims = len( lo )
        # the kernel's float ( `kernel_dtype`, `PowerDiagram.py:375-391` ): float64 ( default ) or float32 ( 2D: the FP32 kernel,
        # vertices re-solved in double, degenerate cells redone in FP64 )
        self.kernel_dtype = kernel_dtype
        name = "float64" if kernel_dtype is None else str( getattr( kernel_dtype, "name", kernel_dtype ) ).replace( "torch.", "" )
        if name not in ( "float64", "float32", "double", "float" ):
            raise ValueError( f"PowerDiagram: kernel_dtype is float32 or float64 ( not { kernel_dtype } )" )
        self._precision = _native.FP32 if name in ( "float32", "float" ) else _native.FP64
        if self._precision == _native.FP32 and self.nb_dims == 3:
            raise NotImplementedError( "PowerDiagram on the MI355X path: the 3D cells are FP64 ( kernel_dtype float64 )" )

    @property
    def nb_points( self ):
        return int( self.positions.shape[ 0 ] )

    def _box( self ):
        return [ *self.box_min, *self.box_max ]

    def _measures( self, facets ):
        if self.nb_dims == 3:
            return _native.measures_3d( self.positions, self.weights, self._box(), self.distribution, facets = facets )
        return _native.measures_2d( self.positions, self.weights, self._box(), self.distribution, facets = facets, precision = self._precision,
                                    domain = self.domain )

    def _box_only( self, what ):
        if self.domain is not None:
            raise NotImplementedError( f"PowerDiagram on the MI355X path: { what } in a box domain ( the masses and facets take a polygon )" )

    @property
    def measures( self ):
        """the mass of each cell against the density ( device tensor, user order ). Differentiable in 2D with respect to the
        positions and the weights when they are tensors that require a gradient ( a constant or image density ): the backward
        is one call of the adjoint kernel ( `sdot_amd_measures_vjp_2d` )"""
        import torch
        pos, w = self.positions, self.weights
        needs = any( isinstance( t, torch.Tensor ) and t.requires_grad for t in ( pos, w ) )
        if needs and self.nb_dims == 2:
            return _Measures.apply( self, pos, w )
        return self._measures( False )

    def facets( self ):
        """`( i, j, c_ij )`: each cell's own view of its facets, `c_ij = int_facet rho / ( 2 |p_i - p_j| )`"""
        _, f = self._measures( True )
        return f

    @property
    def moments( self ):
        """`( masses, first, second )`: for each cell `int rho`, `int x rho` ( `[ n, d ]` ) and `int |x|^2 rho` ( device tensors,
        user order ): a transport cost is `second - 2 p . first + |p|^2 masses`, the barycentres `first / masses`"""
        self._box_only( "moments" )
        m, f, s = _native.moments( self.positions, self.weights, self._box(), self.distribution )
        if self._turn is not None:                       # ( `x = o + R^T y`: `int x rho = o m + R^T int y rho`, `|x|^2 = |o|^2 + 2 R o . y + |y|^2` )
            R, o = self._turn.rotation, self._turn.origin
            T = _native.torch()
            Rt = T.as_tensor( R, dtype = f.dtype, device = f.device )
            ot = T.as_tensor( o, dtype = f.dtype, device = f.device )
            s = s + 2 * ( f @ ( Rt @ ot ) ) + float( o @ o ) * m
            f = f @ Rt + m[ :, None ] * ot
        return m, f, s

    def hessian_rows( self ):
        """`( nb_nbrs, ids, vals )` ( host arrays ): each cell's neighbours `ids[ i, :nb_nbrs[ i ] ]` ( user indices, in the cell's
        edge order; the box's sides are left out ) and `vals[ i, r ] = int_facet rho / ( 2 |p_i - p_j| )`: the Jacobian of the
        masses, `d m_i / d w_j = -vals`, `d m_i / d w_i = sum_j vals`"""
        fi, fj, fc = ( t.cpu().numpy() for t in self.facets() )
        n = self.nb_points
        counts = np.bincount( fi, minlength = n ).astype( int )
        width = max( int( counts.max() ) if n else 0, 1 )
        ids = np.full( ( n, width ), -1, dtype = np.int64 )
        vals = np.zeros( ( n, width ) )
        start = np.concatenate( [ [ 0 ], np.cumsum( counts )[ :-1 ] ] )
        rank = np.arange( len( fi ) ) - start[ fi ]     # ( facets come row by row, each row in its cell's edge order )
        ids[ fi, rank ] = fj
        vals[ fi, rank ] = fc
        return counts, ids, vals

    @property
    def cells( self ):
        """every cell, counter-clockwise, user order: `( offsets [ n + 1 ], vertices [ V, 2 ], cuts [ V ] )` ( device tensors );
        cell `i` is `vertices[ offsets[ i ] : offsets[ i + 1 ] ]`, and `cuts[ v ]` names what bounds the edge leaving vertex `v`
        ( a seed's index, or `-1 - f` for side `f` of the box: 0 bottom, 1 right, 2 top, 3 left )"""
        if self.nb_dims != 2:
            raise NotImplementedError( "PowerDiagram.cells on the MI355X path: 2D" )
        self._box_only( "cells" )
        off, xy, cuts = _native.cells_2d( self.positions, self.weights, self._box(), self.distribution )
        if self._turn is not None:
            xy = self._turn.from_local( xy )
        return off, xy, cuts

    def laplacian_solve( self, rhs, linear_solver = "auto" ):
        """`x = L^+ rhs` with `x[ 0 ] == 0`: the Laplacian of the diagram's facets ( `d m_i / d w_j` ), `rhs` summing to zero.
        On the device ( Jacobi-PCG, or the multigrid with `linear_solver` "mg" / "amg" / "auto" ), tolerance 1e-12"""
        import torch
        fi, fj, fc = self.facets()
        n = self.nb_points
        order = torch.argsort( fi * n + fj )
        fi, fj, fc = fi[ order ], fj[ order ], fc[ order ]
        row = torch.zeros( n + 1, dtype = torch.int64, device = fi.device )
        row[ 1: ] = torch.cumsum( torch.bincount( fi, minlength = n ), 0 )
        dia = torch.zeros( n, dtype = torch.float64, device = fi.device ).index_add_( 0, fi, fc )
        dia = torch.where( dia > 0, dia, torch.ones_like( dia ) )
        method = "cg" if linear_solver == "cg" else "mg"
        sol, its = _native.linear_solves_2d( row, fj, fc, dia, torch.as_tensor( rhs, dtype = torch.float64, device = fi.device ).reshape( -1 ),
                                              tol = 1e-12, method = method )
        if its[ 0 ] < 0:
            raise RuntimeError( "PowerDiagram.laplacian_solve: the linear solver did not converge" )
        return sol[ 0 ]


def _tensor_or_none( t ):
    import torch
    return t if isinstance( t, torch.Tensor ) else None


def _make_measures_fn():
    import torch

    class Measures( torch.autograd.Function ):
        """the masses of a 2D diagram with the adjoint kernel as backward ( gradients of the positions and the weights )"""

        @staticmethod
        def forward( ctx, pd, pos, w ):
            ctx.pd = pd
            ctx.save_for_backward( *( t for t in ( _tensor_or_none( pos ), _tensor_or_none( w ) ) if t is not None ) )
            ctx.has = ( _tensor_or_none( pos ) is not None, _tensor_or_none( w ) is not None )
            return pd._measures( False )

        @staticmethod
        def backward( ctx, g ):
            pd = ctx.pd
            pd._box_only( "the masses' adjoint" )
            gw, gp = _native.measures_vjp_2d( pd.positions.detach() if hasattr( pd.positions, "detach" ) else pd.positions,
                                              None if pd.weights is None else ( pd.weights.detach() if hasattr( pd.weights, "detach" ) else pd.weights ),
                                              pd._box(), pd.distribution, g.contiguous() )
            gpos = gp.reshape( pd.positions.shape ).to( pd.positions.dtype ) if ctx.has[ 0 ] and ctx.needs_input_grad[ 1 ] else None
            gwt = gw.to( pd.weights.dtype ) if ctx.has[ 1 ] and ctx.needs_input_grad[ 2 ] else None
            return None, gpos, gwt

    return Measures


class _Lazy:
    fn = None

    @staticmethod
    def apply( *a ):
        if _Lazy.fn is None:
            _Lazy.fn = _make_measures_fn()
        return _Lazy.fn.apply( *a )


_Measures = _Lazy
