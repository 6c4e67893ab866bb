"""Images on a turned grid ( a frame whose rows are orthogonal: a rotated grid of rectangular pixels ) on the MI355X: the
problem is solved in the grid's frame, where the grid is axis-aligned and the power distances are the same, so the plan is
the axis-aligned problem's to the bit, its barycentres brought back; the diagram's masses against the oracle in that frame,
its moments against the image's closed form in the plane's own coordinates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _turn( th ):
    return np.array( [ [ np.cos( th ), np.sin( th ) ], [ -np.sin( th ), np.cos( th ) ] ] )


def _turned_case( n, seed ):
    import torch
    from sdot import Image
    rng = np.random.default_rng( seed )
    v = rng.uniform( 0.3, 1.7, ( 64, 48 ) )
    F = np.diag( [ 1 / 64, 1 / 48 ] ) @ _turn( 0.4 )
    im = Image( v, origin = [ 0.3, -0.2 ], frame = F )
    x = im.from_local( rng.uniform( 0, 1, ( n, 2 ) ) )
    return im, torch.as_tensor( x, device = "cuda" )


def test_a_turned_grid_solve_is_the_axis_aligned_one( native, oracle ):
    import torch
    import sdot
    im, x = _turned_case( 20_000, 1 )
    it = sdot.Iterative( tol = 1e-10, precision = "fp64", continuation = "never" )
    print( "turned grid: solve", flush = True )
    sol = sdot.OtProblem( sdot.SumOfDiracs( x ), im ).solve( it )
    y = im.to_local( x )
    loc = sdot.OtProblem( sdot.SumOfDiracs( y ), im.local_version() ).solve( it )
    assert sol.converged and loc.converged
    assert torch.equal( sol.weights, loc.weights ) and torch.equal( sol.cell_masses, loc.cell_masses ) and sol.cost == loc.cost
    assert torch.allclose( sol.barycenters, im.from_local( loc.barycenters ), rtol = 0, atol = 1e-15 )
    # the oracle's cells of the plan, in the grid's frame
    vals = np.asarray( loc.problem.target.values )
    d = dict( values = np.ascontiguousarray( vals.T ), x0 = 0.0, y0 = 0.0, hx = 1 / 64, hy = 1 / 48 )
    mo = oracle.measures( y.cpu().numpy(), sol.weights.cpu().numpy(), lo = ( 0, 0 ), hi = ( 1, 1 ), density = d )
    assert np.abs( mo - sol.target_masses ).max() <= 1.01e-10


def test_a_turned_grid_diagram_and_its_moments( native, oracle ):
    import torch
    import sdot
    im, x = _turned_case( 30_000, 2 )
    w = torch.as_tensor( np.random.default_rng( 3 ).normal( 0, 1e-7, len( x ) ), device = "cuda" )
    print( "turned grid: diagram", flush = True )
    pd = sdot.PowerDiagram( x, w, distribution = im )
    m = pd.measures.cpu().numpy()
    y = im.to_local( x ).cpu().numpy()
    n = im.normalized_version()
    d = dict( values = np.ascontiguousarray( np.asarray( n.values ).T ), x0 = 0.0, y0 = 0.0, hx = 1 / 64, hy = 1 / 48 )
    mo = oracle.measures( y, w.cpu().numpy(), density = d )
    assert np.abs( m - mo ).max() <= 1e-12 * mo.max() and abs( m.sum() - 1 ) < 1e-12
    mm, first, second = ( t.cpu().numpy() for t in pd.moments )
    # the image's own integrals, pixel by pixel in the plane's coordinates: a pixel's mass at its centre, its second moment
    # |c|^2 + the parallelogram's own spread ( ( a^2 + b^2 ) / 12 for sides a, b of orthogonal pixels )
    v = np.asarray( n.values )
    i, j = np.meshgrid( np.arange( 64 ) + 0.5, np.arange( 48 ) + 0.5, indexing = "ij" )
    F = im.frame
    cx = 0.3 + F[ 0, 0 ] * i + F[ 1, 0 ] * j
    cy = -0.2 + F[ 0, 1 ] * i + F[ 1, 1 ] * j
    area = abs( np.linalg.det( F ) )
    assert np.allclose( first.sum( 0 ), [ ( v * cx ).sum() * area, ( v * cy ).sum() * area ], rtol = 1e-12, atol = 0 )
    spread = ( ( 1 / 64 ) ** 2 + ( 1 / 48 ) ** 2 ) / 12
    assert abs( second.sum() - ( v * ( cx ** 2 + cy ** 2 + spread ) ).sum() * area ) < 1e-12 * second.sum()
    # the cells' vertices lie on the turned grid
    off, xy, cuts = pd.cells
    dirs, offs = im.bounding_half_spaces()
    assert ( xy.cpu().numpy() @ dirs.T - offs ).max() < 1e-12
